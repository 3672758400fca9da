// Wavefront path-tracing kernels: camera ray generation and the diffuse
// shade/bounce step, in a plain form (one slot per pixel, rays rewritten in
// place, dead paths kept as disabled rays) and a compacting form (live paths
// appended to the next level's queue with one atomic per wavefront).
//
// Semantics follow the reference's pieces: camera rays of cam_ray
// (libgi/camera.cpp:44-61); hit point and geometric normal of diff_geom
// (libgi/rt.cpp:7-15); cosine-weighted bounce about the normal flipped to the
// ray side (rt/cpu/bounce.cpp:41-72); a disabled ray is o = 0, d = (1,1,1),
// t_max = -FLT_MAX (bounce.cpp:50-51).  The path loop is simple_pt's
// (gi/pt.cpp:57-119) with two fixed choices of this workload: emission is
// added and the path continues (emitters also reflect), and Russian roulette
// keeps a path with a fixed probability from level `rr_start` on.  Each
// path/level draws from its own PCG32 stream, so both forms produce the same
// paths.
#include "kernels.h"
#include "pt_kernels.h"

#include "../core/raygen.h"

namespace rt {
namespace dev {

namespace {

__device__ __forceinline__ void write_ray(const ray_out &o, uint32_t i, vec3f org, float tmin, vec3f dir,
                                          float tmax) {
    o.ox[i] = org.x;
    o.oy[i] = org.y;
    o.oz[i] = org.z;
    o.tmin[i] = tmin;
    o.dx[i] = dir.x;
    o.dy[i] = dir.y;
    o.dz[i] = dir.z;
    o.tmax[i] = tmax;
}

__device__ __forceinline__ void write_disabled(const ray_out &o, uint32_t i) {
    write_ray(o, i, v3(0.0f), kRayEps, v3(1.0f), -FLT_MAX);
}

__device__ __forceinline__ vec3f f3(float4 v) { return vec3f{v.x, v.y, v.z}; }

// Shading of one hit.  Returns true and the next ray when the path goes on.
struct shade_out {
    bool live;
    vec3f org, dir;
};

__device__ __forceinline__ shade_out shade_one(const pt_scene &ps, const pt_params &pp, uint32_t path,
                                               int level, uint64_t frame_seed, vec3f in_dir, float t, float beta,
                                               float gamma, uint32_t tri, float *thr, float *rad) {
    shade_out so{false, v3(0.0f), v3(0.0f)};
    if (t == FLT_MAX) return so;
    const float4 *g = ps.tri_geo + 3 * (size_t)tri;
    float4 g0 = g[0], g1 = g[1], g2 = g[2];
    vec3f A = f3(g0), B = f3(g1), C = f3(g2);
    uint32_t m = __float_as_uint(g0.w);
    float4 alb = ps.mats[2 * m], emi = ps.mats[2 * m + 1];
    vec3f T{thr[3 * path], thr[3 * path + 1], thr[3 * path + 2]};
    rad[3 * path + 0] += T.x * emi.x;
    rad[3 * path + 1] += T.y * emi.y;
    rad[3 * path + 2] += T.z * emi.z;
    if (level >= pp.max_depth) return so;
    pcg32 rng(pcg_stream_id(frame_seed, path, (uint64_t)level));
    float xi0 = rng.next_float(), xi1 = rng.next_float();
    float scale = 1.0f;
    if (level + 1 >= pp.rr_start) {
        float xi = rng.next_float();
        if (!(xi < pp.rr_survive)) return so;
        scale = 1.0f / pp.rr_survive;
    }
    vec3f n = geom_normal(A, B, C);
    so.dir = cosine_bounce(n, in_dir, xi0, xi1);
    so.org = hit_point(A, B, C, beta, gamma);
    thr[3 * path + 0] = T.x * alb.x * scale;
    thr[3 * path + 1] = T.y * alb.y * scale;
    thr[3 * path + 2] = T.z * alb.z * scale;
    so.live = true;
    return so;
}

// Shading blocks are 1024 threads (16 wavefronts) so that appends to the
// next queue cost one global atomic per block: atomics on one address
// serialize (about 11 ns apart on MI355X), and one per wavefront over a 2M-ray
// level costs more than the shading itself.
constexpr int kShadeBlock = 1024;
constexpr int kShadeWaves = kShadeBlock / 64;

// Block-aggregated append: every live thread gets a distinct slot in
// [*count, *count + live threads of the block).  Every thread of the block
// must call it.  `sh` is kShadeWaves + 1 words of LDS.
__device__ __forceinline__ uint32_t block_append(bool live, uint32_t *count, uint32_t *sh) {
    uint64_t mask = __ballot(live);
    uint32_t lane = threadIdx.x & 63u, wave = threadIdx.x >> 6;
    if (lane == 0) sh[wave] = (uint32_t)__popcll(mask);
    __syncthreads();
    if (threadIdx.x == 0) {
        uint32_t sum = 0;
        for (int w = 0; w < kShadeWaves; ++w) {
            uint32_t c = sh[w];
            sh[w] = sum;
            sum += c;
        }
        sh[kShadeWaves] = sum ? atomicAdd(count, sum) : 0u;
    }
    __syncthreads();
    return sh[kShadeWaves] + sh[wave] + (uint32_t)__popcll(mask & ((1ull << lane) - 1ull));
}

} // namespace

// Camera rays for every pixel; path i = pixel i.  Also resets throughput (when
// thr is non-null), the accumulator (when `clear_radiance`) and the per-path
// id table (when path_ids is non-null).
__global__ __launch_bounds__(256) void k_camera_rays(cam_frame f, ray_out out, uint32_t *path_ids, float *thr,
                                                    float *rad, int clear_radiance, uint32_t n) {
    uint32_t i = blockIdx.x * 256u + threadIdx.x;
    if (i >= n) return;
    int x = (int)(i % (uint32_t)f.w), y = (int)(i / (uint32_t)f.w);
    write_ray(out, i, f.pos, kRayEps, cam_dir(f, x, y), FLT_MAX);
    if (path_ids) path_ids[i] = i;
    if (thr) thr[3 * i] = thr[3 * i + 1] = thr[3 * i + 2] = 1.0f;
    if (clear_radiance) rad[3 * i] = rad[3 * i + 1] = rad[3 * i + 2] = 0.0f;
}

// Plain shading: slot i is path i; the next ray replaces this one in place.
// counts[level + 1] receives the number of live next-level rays.
__global__ __launch_bounds__(kShadeBlock) void k_shade_plain(pt_scene ps, pt_params pp, int level,
                                                            const uint64_t *frame_seed, ray_out rays, hit_soa hits,
                                                            float *thr, float *rad, uint32_t *counts, uint32_t n) {
    __shared__ uint32_t sh[kShadeWaves + 1];
    uint32_t i = blockIdx.x * (uint32_t)kShadeBlock + threadIdx.x;
    bool live = false;
    if (i < n && rays.tmax[i] >= 0.0f) {
        vec3f in_dir{rays.dx[i], rays.dy[i], rays.dz[i]};
        shade_out so = shade_one(ps, pp, i, level, *frame_seed, in_dir, hits.t[i], hits.u[i], hits.v[i],
                                 hits.id[i], thr, rad);
        live = so.live;
        if (live) write_ray(rays, i, so.org, kRayEps, so.dir, FLT_MAX);
        else write_disabled(rays, i);
    }
    if (level < pp.max_depth) block_append(live, counts + level + 1, sh);
}

// Compacting shading: reads queue slot i < counts[level] (path path_in[i]),
// appends live paths to the next queue.
__global__ __launch_bounds__(kShadeBlock) void k_shade_compact(pt_scene ps, pt_params pp, int level,
                                                              const uint64_t *frame_seed, ray_out rays_in,
                                                              hit_soa hits, const uint32_t *path_in,
                                                              ray_out rays_out, uint32_t *path_out, float *thr,
                                                              float *rad, uint32_t *counts) {
    __shared__ uint32_t sh[kShadeWaves + 1];
    uint32_t i = blockIdx.x * (uint32_t)kShadeBlock + threadIdx.x;
    uint32_t n = counts[level];
    if (blockIdx.x * (uint32_t)kShadeBlock >= n) return;   // whole block past the queue: uniform exit
    bool live = false;
    shade_out so{};
    uint32_t path = 0;
    if (i < n) {
        path = path_in[i];
        vec3f in_dir{rays_in.dx[i], rays_in.dy[i], rays_in.dz[i]};
        so = shade_one(ps, pp, path, level, *frame_seed, in_dir, hits.t[i], hits.u[i], hits.v[i], hits.id[i], thr,
                       rad);
        live = so.live;
    }
    if (level >= pp.max_depth) return;
    uint32_t slot = block_append(live, counts + level + 1, sh);
    if (live) {
        write_ray(rays_out, slot, so.org, kRayEps, so.dir, FLT_MAX);
        path_out[slot] = path;
    }
}

// The "sample cos weighted dir" step on its own (rt/cpu/bounce.cpp:41-72):
// slot i gets a cosine-weighted bounce ray from hit i of ray i, or the
// disabled ray when ray i missed.  Random numbers: PCG32 stream (seed, i, level).
__global__ __launch_bounds__(256) void k_cosine_bounce(pt_scene ps, ray_soa in, hit_soa hits, ray_out out,
                                                      uint64_t seed, int level, uint32_t n) {
    uint32_t i = blockIdx.x * 256u + threadIdx.x;
    if (i >= n) return;
    float t = hits.t[i];
    if (t == FLT_MAX || !(in.tmax[i] >= 0.0f)) {
        write_disabled(out, i);
        return;
    }
    const float4 *g = ps.tri_geo + 3 * (size_t)hits.id[i];
    float4 g0 = g[0], g1 = g[1], g2 = g[2];
    vec3f A = f3(g0), B = f3(g1), C = f3(g2);
    pcg32 rng(pcg_stream_id(seed, i, (uint64_t)level));
    float xi0 = rng.next_float(), xi1 = rng.next_float();
    vec3f dir = cosine_bounce(geom_normal(A, B, C), vec3f{in.dx[i], in.dy[i], in.dz[i]}, xi0, xi1);
    write_ray(out, i, hit_point(A, B, C, hits.u[i], hits.v[i]), kRayEps, dir, FLT_MAX);
}

hipError_t launch_cosine_bounce(const pt_scene &ps, const ray_soa &in, const hit_soa &hits, const ray_out &out,
                                uint64_t seed, int level, uint32_t n, hipStream_t s) {
    if (n == 0) return hipSuccess;
    hipLaunchKernelGGL(k_cosine_bounce, dim3((n + 255) / 256), dim3(256), 0, s, ps, in, hits, out, seed, level, n);
    return hipGetLastError();
}

__global__ void k_add_counts(unsigned long long *totals, const uint32_t *counts, int levels) {
    int i = threadIdx.x;
    if (blockIdx.x == 0 && i < levels) totals[i] += counts[i];
}

hipError_t launch_add_counts(unsigned long long *totals, const uint32_t *counts, int levels, hipStream_t s) {
    hipLaunchKernelGGL(k_add_counts, dim3(1), dim3(64 * ((levels + 63) / 64)), 0, s, totals, counts, levels);
    return hipGetLastError();
}

__global__ void k_set_u64(uint64_t *p, uint64_t v) {
    if (threadIdx.x == 0 && blockIdx.x == 0) *p = v;
}

hipError_t launch_camera_rays(const cam_frame &f, const ray_out &out, uint32_t *path_ids, float *thr, float *rad,
                              bool clear_radiance, hipStream_t s) {
    uint32_t n = (uint32_t)f.w * (uint32_t)f.h;
    hipLaunchKernelGGL(k_camera_rays, dim3((n + 255) / 256), dim3(256), 0, s, f, out, path_ids, thr, rad,
                       clear_radiance ? 1 : 0, n);
    return hipGetLastError();
}

hipError_t launch_shade_plain(const pt_scene &ps, const pt_params &pp, int level, const uint64_t *frame_seed,
                              const ray_out &rays, const hit_soa &hits, float *thr, float *rad, uint32_t *counts,
                              uint32_t n, hipStream_t s) {
    hipLaunchKernelGGL(k_shade_plain, dim3((n + kShadeBlock - 1) / kShadeBlock), dim3(kShadeBlock), 0, s, ps, pp,
                       level, frame_seed, rays, hits, thr, rad, counts, n);
    return hipGetLastError();
}

hipError_t launch_shade_compact(const pt_scene &ps, const pt_params &pp, int level, const uint64_t *frame_seed,
                                const ray_out &rays_in, const hit_soa &hits, const uint32_t *path_in,
                                const ray_out &rays_out, uint32_t *path_out, float *thr, float *rad,
                                uint32_t *counts, uint32_t n_max, hipStream_t s) {
    hipLaunchKernelGGL(k_shade_compact, dim3((n_max + kShadeBlock - 1) / kShadeBlock), dim3(kShadeBlock), 0, s, ps,
                       pp, level, frame_seed, rays_in, hits, path_in, rays_out, path_out, thr, rad, counts);
    return hipGetLastError();
}

hipError_t launch_set_u64(uint64_t *p, uint64_t v, hipStream_t s) {
    hipLaunchKernelGGL(k_set_u64, dim3(1), dim3(64), 0, s, p, v);
    return hipGetLastError();
}

} // namespace dev
} // namespace rt
