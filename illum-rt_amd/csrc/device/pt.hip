// Wavefront path-tracing kernels: camera ray generation and the diffuse
// shade/bounce step, in a plain form (one slot per pixel, rays rewritten in
// place, dead paths kept as disabled rays) and a compacting form (live paths
// appended to the next level's queue with one atomic per wavefront).
//
// Semantics follow the reference's pieces: camera rays of cam_ray
// (libgi/camera.cpp:44-61); hit point and geometric normal of diff_geom
// (libgi/rt.cpp:7-15); cosine-weighted bounce about the geometric normal
// flipped to the ray side by the shading normal (rt/cpu/bounce.cpp:41-72,
// libgi/util.h:226-232); a disabled ray is o = 0, d = (1,1,1),
// t_max = -FLT_MAX (bounce.cpp:50-51).  The path loop is simple_pt's
// (gi/pt.cpp:72-115): a miss ends the path (no sky light), an emitter adds
// throughput x emission and ends it (unless emitters_reflect is set), a
// zero throughput ends it; with two fixed choices of this workload: the
// bounce is the cosine-weighted wavefront step above (throughput x albedo),
// and Russian roulette keeps a path with a fixed probability from level
// `rr_start` on.  Each path/level draws from its own PCG32 stream, so every
// schedule produces the same paths.
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

// Shading of one hit.  Returns true and the next ray when the path goes on;
// T is the path's throughput, updated when it does.  Emission is added to
// the path's radiance only where the material emits, and ends the path there
// unless emitters reflect.
struct shade_out {
    bool live;
    vec3f org, dir;
};

// The path is pixel `pixel` of sample `sample`; its radiance accumulates in
// rad[0..2].
__device__ __forceinline__ shade_out shade_one(const pt_scene &ps, const pt_params &pp, uint32_t sample,
                                               uint32_t pixel, int level, const uint64_t *frame_seed, vec3f in_dir,
                                               float t, float beta, float gamma, uint32_t tri, vec3f &T, float *rad) {
    shade_out so{false, v3(0.0f), v3(0.0f)};
    if (!ps.valid(t, tri)) return so;
    const float4 *g = ps.tri_geo + 3 * (size_t)tri;
    float4 g0 = g[0], g1{}, g2{};
    if (level < pp.max_depth) {   // uniform; the last level only adds emission (material in g0.w)
        g1 = g[1];
        g2 = g[2];
    }
    vec3f A = f3(g0), B = f3(g1), C = f3(g2);
    uint32_t m = __float_as_uint(g0.w);
    float4 alb = ps.mats[2 * m], emi = ps.mats[2 * m + 1];
    if (emi.x != 0.0f || emi.y != 0.0f || emi.z != 0.0f) {
        rad[0] += T.x * emi.x;
        rad[1] += T.y * emi.y;
        rad[2] += T.z * emi.z;
        if (!pp.emitters_reflect) return so;
    }
    if (level >= pp.max_depth) return so;
    pcg32 rng(pcg_stream_id(frame_seed[sample], pixel, (uint64_t)level));
    float xi0 = rng.next_float(), xi1 = rng.next_float();
    float scale = 1.0f;
    if (level + 1 >= pp.rr_start) {
        float xi = rng.next_float();
        if (!(xi < pp.rr_survive)) return so;
        scale = 1.0f / pp.rr_survive;
    }
    vec3f n = geom_normal(A, B, C);
    so.dir = cosine_bounce(n, ps.shading_normal(tri, beta, gamma, n), in_dir, xi0, xi1);
    so.org = hit_point(A, B, C, beta, gamma);
    T = vec3f{T.x * alb.x * scale, T.y * alb.y * scale, T.z * alb.z * scale};
    // a path that can no longer carry light ends (gi/pt.cpp:110-111)
    so.live = T.x != 0.0f || T.y != 0.0f || T.z != 0.0f;
    return so;
}

// shade_one for path `path` (numbered as pt_params::decode says), radiance in
// rad[3 path ..].
__device__ __forceinline__ shade_out shade_path(const pt_scene &ps, const pt_params &pp, uint32_t path, int level,
                                                const uint64_t *frame_seed, vec3f in_dir, float t, float beta,
                                                float gamma, uint32_t tri, vec3f &T, float *rad) {
    uint32_t sample, pixel;
    pp.decode(path, sample, pixel);
    return shade_one(ps, pp, sample, pixel, level, frame_seed, in_dir, t, beta, gamma, tri, T, rad + 3ull * path);
}

__device__ __forceinline__ vec3f load3(const float *p, uint32_t i) { return vec3f{p[3 * i], p[3 * i + 1], p[3 * i + 2]}; }
__device__ __forceinline__ void store3(float *p, uint32_t i, vec3f v) {
    p[3 * i] = v.x;
    p[3 * i + 1] = v.y;
    p[3 * i + 2] = v.z;
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

// Camera rays for n paths, ray i for path i: with `px_samples` k > 1 path i
// is sample i % k of pixel i / k (pixel-major: the samples of a pixel share a
// wavefront and traverse the tree almost as one ray), else sample i / n_px of
// pixel i % n_px.  Also resets the path's throughput (when thr is non-null),
// its pixel's accumulator (when `clear_radiance`, once per pixel), the path's
// radiance (when rad_paths is non-null) and its id (when path_ids is
// non-null).
__global__ __launch_bounds__(256) void k_camera_rays(cam_frame f, ray_out out, uint32_t *path_ids, float *thr,
                                                    float *rad, int clear_radiance, uint32_t n, uint32_t n_px,
                                                    float *rad_paths, const uint64_t *jitter_seeds,
                                                    uint32_t px_samples) {
    uint32_t i = blockIdx.x * 256u + threadIdx.x;
    if (i >= n) return;
    uint32_t px, sample;
    if (px_samples > 1) {
        px = i / px_samples;
        sample = i - px * px_samples;
    } else {
        sample = i / n_px;
        px = i - sample * n_px;
    }
    int x = (int)(px % (uint32_t)f.w), y = (int)(px / (uint32_t)f.w);
    float ox = 0.0f, oy = 0.0f;
    if (jitter_seeds) pixel_jitter(jitter_seeds[sample], px, ox, oy);
    write_ray(out, i, f.pos, kRayEps, cam_dir(f, x, y, ox, oy), FLT_MAX);
    if (path_ids) path_ids[i] = i;
    if (thr) thr[3ull * i] = thr[3ull * i + 1] = thr[3ull * i + 2] = 1.0f;
    if (rad_paths) rad_paths[3ull * i] = rad_paths[3ull * i + 1] = rad_paths[3ull * i + 2] = 0.0f;
    if (clear_radiance && sample == 0) rad[3ull * px] = rad[3ull * px + 1] = rad[3ull * px + 2] = 0.0f;
}

__global__ __launch_bounds__(256) void k_add_samples(float *image, const float *rad_paths, uint32_t n_px,
                                                    uint32_t samples, int pixel_major) {
    uint32_t px = blockIdx.x * 256u + threadIdx.x;
    if (px >= n_px) return;
    float r = image[3ull * px], g = image[3ull * px + 1], b = image[3ull * px + 2];
    for (uint32_t k = 0; k < samples; ++k) {
        const uint64_t path = pixel_major ? (uint64_t)px * samples + k : (uint64_t)k * n_px + px;
        const float *q = rad_paths + 3ull * path;
        r += q[0];
        g += q[1];
        b += q[2];
    }
    image[3ull * px] = r;
    image[3ull * px + 1] = g;
    image[3ull * px + 2] = b;
}

__global__ void k_set_seeds(uint64_t *p, uint64_t base, uint32_t n) {
    uint32_t k = blockIdx.x * 64u + threadIdx.x;
    if (k < n) p[k] = base + k;
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
        vec3f T = load3(thr, i);
        shade_out so = shade_path(ps, pp, i, level, frame_seed, in_dir, hits.t[i], hits.u[i], hits.v[i],
                                 hits.id[i], T, rad);
        live = so.live;
        if (live) store3(thr, i, T);
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
        vec3f T = load3(thr, path);
        so = shade_path(ps, pp, path, level, frame_seed, in_dir, hits.t[i], hits.u[i], hits.v[i], hits.id[i], T,
                       rad);
        live = so.live;
        if (live) store3(thr, path, T);
    }
    if (level >= pp.max_depth) return;
    uint32_t slot = block_append(live, counts + level + 1, sh);
    if (live) {
        write_ray(rays_out, slot, so.org, kRayEps, so.dir, FLT_MAX);
        path_out[slot] = path;
    }
}

// ---- sorted bounce queues (the batched schedule's default) ----
//
// A queue written in shading order is, after a bounce or two, incoherent:
// neighbouring slots start anywhere in the scene and point anywhere, so the
// rays a wavefront traces together share few nodes.  In the sorting form the
// shade step writes each next-level ray as a 32-byte record, its (path id,
// throughput) as a 16-byte record and a key (the direction's cell of 1024 on
// its octahedral projection and the Morton code of the origin's cell in a
// 128^3 grid over the scene box, in three levels, coarse first: 31 bits),
// all in append order; the
// stable radix sort of the keys (device/ray_sort.hip) gives the permutation the trace kernel reads
// the ray records through, and the next shade step reads directions and
// path records through the same permutation.  A ray and its path share one
// 48-byte record (origin, t_min | direction, t_max | path id, throughput):
// the trace kernel reads its first 32 bytes, the shade step its last 32.  Sorted rays fill the
// persistent kernel's work chunks key range by key range, so each XCD
// traces rays of a few direction cells and one part of the scene.  The
// throughput travels with the ray because a sorted queue visits pixels in
// no useful order: a per-pixel throughput array would be read and written
// at random addresses (that alone cost 110 us per level at 2M rays).
// Paths keep their own random streams, so the image does not depend on the
// queue order.

__device__ __forceinline__ uint32_t spread3(uint32_t v) {   // bit b -> bit 3b
    uint32_t r = 0;
#pragma unroll
    for (int b = 0; b < kSortCellBits; ++b) r |= ((v >> b) & 1u) << (3 * b);
    return r;
}

__device__ __forceinline__ uint32_t cell_of(float x, float lo, float scale) {
    float q = (x - lo) * scale;
    q = fminf(fmaxf(q, 0.0f), (float)((1u << kSortCellBits) - 1));   // fmaxf maps NaN to 0
    return (uint32_t)q;
}

__device__ __forceinline__ uint32_t origin_cell(const sort_grid &g, vec3f o) {
    return (spread3(cell_of(o.x, g.lo[0], g.scale[0])) << 2) | (spread3(cell_of(o.y, g.lo[1], g.scale[1])) << 1) |
           spread3(cell_of(o.z, g.lo[2], g.scale[2]));
}

// Key of a bounce ray in the samples-in-flight queues, from the direction's
// cell in a 32 x 32 grid over its octahedral projection (|x| + |y| + |z| = 1,
// the lower half folded over the diagonals; kDirHalfBits bits per axis,
// Morton order) and the origin's cell in a 128^3 grid, in three levels: the
// direction's cell of a 4 x 4 grid and the origin's of a 16^3 grid, then
// those of 16 x 16 and 64^3, then the last direction and origin bits.  NaN
// coordinates land in cell 0.
__device__ __forceinline__ uint32_t ray_key(const sort_grid &g, vec3f o, vec3f d) {
    const float inv = __builtin_amdgcn_rcpf(fabsf(d.x) + fabsf(d.y) + fabsf(d.z));
    float u = d.x * inv, v = d.y * inv;
    if (d.z < 0.0f) {
        const float uu = copysignf(1.0f - fabsf(v), u), vv = copysignf(1.0f - fabsf(u), v);
        u = uu;
        v = vv;
    }
    constexpr float kCells = (float)(1 << kDirHalfBits);
    const uint32_t qu = (uint32_t)fminf(fmaxf((u * 0.5f + 0.5f) * kCells, 0.0f), kCells - 1.0f);
    const uint32_t qv = (uint32_t)fminf(fmaxf((v * 0.5f + 0.5f) * kCells, 0.0f), kCells - 1.0f);
    uint32_t dir = 0;
#pragma unroll
    for (int b = 0; b < kDirHalfBits; ++b) dir |= (((qu >> b) & 1u) << (2 * b + 1)) | (((qv >> b) & 1u) << (2 * b));
    static_assert(kDirHalfBits == 5 && kSortCellBits == 7, "levels: 4 + 4 + 2 direction bits, 12 + 6 + 3 origin bits");
    const uint32_t org = origin_cell(g, o);
    return ((dir >> 6) << 27) | ((org >> 9) << 15) | (((dir >> 2) & 15u) << 11) | (((org >> 3) & 63u) << 5) |
           ((dir & 3u) << 3) | (org & 7u);
}

// Key of a bounce ray in the regeneration queues and in one-sample queues:
// direction octant, then the origin's cell in a 16^3 grid (the Morton code
// without each axis's low kSortCellBits - kOctantCellBits bits;
// kOctantKeyBits).
__device__ __forceinline__ uint32_t octant_key(const sort_grid &g, vec3f o, vec3f d) {
    const uint32_t oct = (d.x < 0.0f ? 4u : 0u) | (d.y < 0.0f ? 2u : 0u) | (d.z < 0.0f ? 1u : 0u);
    return (oct << (3 * kOctantCellBits)) | (origin_cell(g, o) >> (3 * (kSortCellBits - kOctantCellBits)));
}

// Shade step of the sorting form.  Queue slot i < counts[level] holds the
// ray of record src = perm[i] (src = i without a permutation).  At the camera
// level (in_rec null) the direction comes from plane i of rays_in and the
// path and throughput from path_in[i] and thr (path i and throughput 1 when
// they are null); later, direction and (path id bits, throughput) from record
// src of in_rec.  The last level reads no direction: it only adds emission.
// The hit is slot i.  Live
// next-level rays are appended as records to out_rec with their keys: the
// throughput travels with the ray, so no step reads or writes per-pixel
// throughput at a random address.
__global__ __launch_bounds__(kShadeBlock) void k_shade_stage(pt_scene ps, pt_params pp, int level,
                                                            const uint64_t *frame_seed, ray_out rays_in,
                                                            const uint32_t *path_in, const float *thr,
                                                            const float4 *in_rec, const uint32_t *perm,
                                                            hit_soa hits, float4 *out_rec, uint32_t *keys,
                                                            float *rad, uint32_t *counts, sort_grid grid) {
    __shared__ uint32_t sh[kShadeWaves + 1];
    uint32_t i = blockIdx.x * (uint32_t)kShadeBlock + threadIdx.x;
    uint32_t n = counts[level];
    if (blockIdx.x * (uint32_t)kShadeBlock >= n) return;   // whole block past the queue: uniform exit
    bool live = false;
    shade_out so{};
    uint32_t path = 0;
    vec3f T{};
    if (i < n) {
        vec3f in_dir{};
        const bool bounce = level < pp.max_depth;   // uniform
        if (in_rec) {
            uint32_t src = perm ? perm[i] : i;
            float4 pr = in_rec[3ull * src + 2];
            if (bounce) {
                float4 d = in_rec[3ull * src + 1];
                in_dir = vec3f{d.x, d.y, d.z};
            }
            path = __float_as_uint(pr.x);
            T = vec3f{pr.y, pr.z, pr.w};
        } else {
            path = path_in ? path_in[i] : i;
            if (bounce) in_dir = vec3f{rays_in.dx[i], rays_in.dy[i], rays_in.dz[i]};
            T = thr ? load3(thr, path) : vec3f{1.0f, 1.0f, 1.0f};
        }
        so = shade_path(ps, pp, path, level, frame_seed, in_dir, hits.t[i], hits.u[i], hits.v[i], hits.id[i], T,
                       rad);
        live = so.live;
    }
    if (level >= pp.max_depth) return;
    uint32_t slot = block_append(live, counts + level + 1, sh);
    if (live) {
        out_rec[3ull * slot] = make_float4(so.org.x, so.org.y, so.org.z, kRayEps);
        out_rec[3ull * slot + 1] = make_float4(so.dir.x, so.dir.y, so.dir.z, FLT_MAX);
        out_rec[3ull * slot + 2] = make_float4(__uint_as_float(path), T.x, T.y, T.z);
        // 1024 direction cells over a 128^3 grid when a pixel's samples share
        // the queue (their rays leave one point in many directions), the
        // octant over a 16^3 grid for one sample per pixel: each traced its
        // queue fastest in that setting
        keys[slot] = pp.px_samples > 1 ? ray_key(grid, so.org, so.dir) : octant_key(grid, so.org, so.dir);
    }
}

// ---- path regeneration ----
//
// A fixed pool of lanes x n_px path slots.  Slot j renders pixel j % n_px,
// samples j / n_px, j / n_px + lanes, ... of the render call, one after
// another: when a path ends, the shade step that ended it appends the camera
// ray of the slot's next sample in its place, so every iteration traces a
// full pool until the slots run out of samples.  The queue holds records
// (origin, - | direction, sample << 8 | level | slot, throughput), traced
// with ray_gather::fixed_t (every ray of the path tracer spans
// [kRayEps, FLT_MAX]).  With one lane per pixel a pixel's paths end in
// sample order and add their radiance straight to the image, in the same
// order as the in-flight schedule adds its samples; with more lanes each
// slot accumulates its own samples in `acc` and the lanes are added at the
// end of the render call.

// Key of a camera ray: above every bounce ray's key (the top bit of
// regen_key_bits), so that camera rays and bounce rays do not share
// wavefronts; then its
// direction octant and the Morton code of its pixel's cell in a 64 x 64
// grid over the image (camera rays share one origin, so the origin cell
// would not tell them apart).
__device__ __forceinline__ uint32_t camera_key(const cam_frame &f, int x, int y, vec3f d, uint32_t lanes) {
    uint32_t oct = (d.x < 0.0f ? 4u : 0u) | (d.y < 0.0f ? 2u : 0u) | (d.z < 0.0f ? 1u : 0u);
    uint32_t cx = (uint32_t)((uint64_t)x * 64u / (uint32_t)f.w), cy = (uint32_t)((uint64_t)y * 64u / (uint32_t)f.h);
    uint32_t m = 0;
#pragma unroll
    for (int b = 0; b < 6; ++b) m |= (((cx >> b) & 1u) << (2 * b + 1)) | (((cy >> b) & 1u) << (2 * b));
    return (1u << (regen_key_bits(lanes) - 1)) | (oct << 12) | m;
}

// Camera-ray record of sample `sample` of pixel `px` for slot `slot`, at
// queue position `at`.  The ray is k_camera_rays' ray for that path.
__device__ __forceinline__ void write_camera_record(const regen_params &rp, uint32_t slot, uint32_t px,
                                                    uint32_t sample, float4 *out_rec, uint32_t *keys, uint32_t at) {
    int x = (int)(px % (uint32_t)rp.f.w), y = (int)(px / (uint32_t)rp.f.w);
    float ox = 0.0f, oy = 0.0f;
    if (rp.jitter) pixel_jitter(rp.seeds[sample], px, ox, oy);
    vec3f d = cam_dir(rp.f, x, y, ox, oy);
    out_rec[3ull * at] = make_float4(rp.f.pos.x, rp.f.pos.y, rp.f.pos.z, 0.0f);
    out_rec[3ull * at + 1] = make_float4(d.x, d.y, d.z, __uint_as_float(sample << 8));
    out_rec[3ull * at + 2] = make_float4(__uint_as_float(slot), 1.0f, 1.0f, 1.0f);
    if (keys) keys[at] = camera_key(rp.f, x, y, d, rp.lanes);
}

// The pool's first paths: slot j < lanes * n_px gets sample j / n_px of its
// pixel at queue position j; clears the slot's accumulator.
__global__ __launch_bounds__(256) void k_regen_init(regen_params rp, float4 *out_rec, uint32_t *keys) {
    uint32_t j = blockIdx.x * 256u + threadIdx.x;
    if (j >= rp.lanes * rp.n_px) return;
    uint32_t lane = j / rp.n_px, px = j - lane * rp.n_px;
    write_camera_record(rp, j, px, lane, out_rec, keys, j);
    if (rp.acc) rp.acc[3ull * j] = rp.acc[3ull * j + 1] = rp.acc[3ull * j + 2] = 0.0f;
}

// One regeneration iteration's shade step: queue slot i < *count_in is the
// ray of record perm[i] (i without a permutation), its hit is slot i.  A
// path that goes on is appended with level + 1; one that ends appends its
// slot's next sample's camera ray, if the slot has one left.  Rays shaded
// per level are added to rp.level_counts.
__global__ __launch_bounds__(kShadeBlock) void k_shade_regen(pt_scene ps, pt_params pp, regen_params rp,
                                                            const float4 *in_rec, const uint32_t *perm, hit_soa hits,
                                                            float4 *out_rec, uint32_t *keys, const uint32_t *count_in,
                                                            uint32_t *count_out, sort_grid grid) {
    __shared__ uint32_t sh[kShadeWaves + 1];
    __shared__ uint32_t per_level[kMaxPathLevels];
    const uint32_t i = blockIdx.x * (uint32_t)kShadeBlock + threadIdx.x;
    const uint32_t n = *count_in;
    if (blockIdx.x * (uint32_t)kShadeBlock >= n) return;   // whole block past the queue: uniform exit
    const int levels = pp.max_depth + 1;
    if ((int)threadIdx.x < levels) per_level[threadIdx.x] = 0u;
    __syncthreads();
    bool emit = false, live = false;
    shade_out so{};
    vec3f T{};
    uint32_t slot = 0, px = 0, sample = 0, next_word = 0;
    if (i < n) {
        uint32_t src = perm ? perm[i] : i;
        float4 d = in_rec[3ull * src + 1];
        float4 pr = in_rec[3ull * src + 2];
        uint32_t word = __float_as_uint(d.w);
        sample = word >> 8;
        int level = (int)(word & 255u);
        slot = __float_as_uint(pr.x);
        px = slot % rp.n_px;
        T = vec3f{pr.y, pr.z, pr.w};
        float *r3 = rp.acc ? rp.acc + 3ull * slot : rp.image + 3ull * px;
        so = shade_one(ps, pp, sample, px, level, rp.seeds, vec3f{d.x, d.y, d.z}, hits.t[i], hits.u[i], hits.v[i],
                       hits.id[i], T, r3);
        atomicAdd(&per_level[level < levels ? level : levels - 1], 1u);
        live = so.live;
        next_word = (sample << 8) | (uint32_t)(level + 1);
        emit = live || sample + rp.lanes < rp.samples;
    }
    __syncthreads();
    if ((int)threadIdx.x < levels && per_level[threadIdx.x])
        atomicAdd(rp.level_counts + threadIdx.x, per_level[threadIdx.x]);
    uint32_t at = block_append(emit, count_out, sh);
    if (!emit) return;
    if (live) {
        out_rec[3ull * at] = make_float4(so.org.x, so.org.y, so.org.z, 0.0f);
        out_rec[3ull * at + 1] = make_float4(so.dir.x, so.dir.y, so.dir.z, __uint_as_float(next_word));
        out_rec[3ull * at + 2] = make_float4(__uint_as_float(slot), T.x, T.y, T.z);
        if (keys) keys[at] = rp.lanes > 1 ? ray_key(grid, so.org, so.dir) : octant_key(grid, so.org, so.dir);
    } else {
        write_camera_record(rp, slot, px, sample + rp.lanes, out_rec, keys, at);
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
    if (!ps.valid(t, hits.id[i]) || !(in.tmax[i] >= 0.0f)) {
        write_disabled(out, i);
        return;
    }
    const float4 *g = ps.tri_geo + 3 * (size_t)hits.id[i];
    float4 g0 = g[0], g1 = g[1], g2 = g[2];
    vec3f A = f3(g0), B = f3(g1), C = f3(g2);
    pcg32 rng(pcg_stream_id(seed, i, (uint64_t)level));
    float xi0 = rng.next_float(), xi1 = rng.next_float();
    vec3f ng = geom_normal(A, B, C);
    vec3f dir = cosine_bounce(ng, ps.shading_normal(hits.id[i], hits.u[i], hits.v[i], ng),
                              vec3f{in.dx[i], in.dy[i], in.dz[i]}, xi0, xi1);
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

// Incoherent benchmark rays: setup_ray_incoherent of the reference's camera
// step under "incoherence r1 r2" (rt/cuda/kernels.cu:75-156).  Ray i starts
// at sphere 1's centre moved by a uniform point of the ball of radius
// r1 * r_max and runs to sphere 2's centre moved by a uniform point of the
// ball of radius r2 * r_max (a zero radius keeps the centre); the direction
// is that difference, unnormalized as the reference leaves it, over
// [0, FLT_MAX].  A ball point takes three numbers: phi = 2 pi xi1,
// cos theta = 2 xi2 - 1, r = R cbrt(xi3).  The numbers come from ray i's
// PCG32 stream (the reference draws from cuRAND's MTGP32), the origin's
// three first.
__device__ __forceinline__ vec3f ball_point(pcg32 &g, float radius) {
    const float pi = 3.14159265358979323846f;
    const float phi = g.next_float() * 2 * pi;
    const float costheta = g.next_float() * 2 - 1;
    const float u = g.next_float();
    const float theta = acosf(costheta);
    const float r = radius * cbrtf(u);
    return vec3f{r * sinf(theta) * cosf(phi), r * sinf(theta) * sinf(phi), r * cosf(theta)};
}

__global__ __launch_bounds__(256) void k_incoherent_rays(vec3f s1, vec3f s2, float r1, float r2, float r_max,
                                                        uint64_t seed, ray_out out, uint32_t n) {
    const uint32_t i = blockIdx.x * 256u + threadIdx.x;
    if (i >= n) return;
    pcg32 g(pcg_stream_id(seed, i, 0));
    vec3f o = s1, t = s2;
    if (r1 > 0) {
        const vec3f b = ball_point(g, r1 * r_max);
        o = vec3f{s1.x + b.x, s1.y + b.y, s1.z + b.z};
    }
    if (r2 > 0) {
        const vec3f b = ball_point(g, r2 * r_max);
        t = vec3f{s2.x + b.x, s2.y + b.y, s2.z + b.z};
    }
    write_ray(out, i, o, 0.0f, vec3f{t.x - o.x, t.y - o.y, t.z - o.z}, FLT_MAX);
}

hipError_t launch_incoherent_rays(vec3f s1, vec3f s2, float r1, float r2, float r_max, uint64_t seed,
                                  const ray_out &out, uint32_t n, hipStream_t s) {
    if (n == 0) return hipSuccess;
    hipLaunchKernelGGL(k_incoherent_rays, dim3((n + 255) / 256), dim3(256), 0, s, s1, s2, r1, r2, r_max, seed, out, n);
    return hipGetLastError();
}

hipError_t launch_camera_rays(const cam_frame &f, const ray_out &out, uint32_t *path_ids, float *thr, float *rad,
                              bool clear_radiance, hipStream_t s, uint32_t n_paths, float *rad_paths,
                              const uint64_t *jitter_seeds, uint32_t px_samples) {
    uint32_t n_px = (uint32_t)f.w * (uint32_t)f.h;
    uint32_t n = n_paths ? n_paths : n_px;
    if (px_samples > 1 && (uint64_t)px_samples * n_px != n) return hipErrorInvalidValue;   // whole pixels only
    hipLaunchKernelGGL(k_camera_rays, dim3((n + 255) / 256), dim3(256), 0, s, f, out, path_ids, thr, rad,
                       clear_radiance ? 1 : 0, n, n_px, rad_paths, jitter_seeds, px_samples);
    return hipGetLastError();
}

// Pixel-major per-path radiance: the samples of pixel px are the contiguous
// paths px * samples + k.  A block stages the radiance of `per_block` pixels
// through LDS with coalesced loads (each pixel's 3 * samples floats padded by
// one word, so the summing threads below fall in different banks), then one
// thread per (pixel, channel) adds that channel's samples to the image in
// sample order, as k_add_samples does.
constexpr uint32_t kAddLdsFloats = 12288;   // 48 KiB
__global__ __launch_bounds__(256) void k_add_samples_px(float *image, const float *rad_paths, uint32_t n_px,
                                                       uint32_t samples, uint32_t per_block) {
    extern __shared__ float lds[];
    const uint32_t px0 = blockIdx.x * per_block;
    if (px0 >= n_px) return;
    const uint32_t np = min(per_block, n_px - px0);
    const uint32_t row = 3 * samples, stride = row + 1;
    const float *src = rad_paths + (uint64_t)px0 * row;
    // element j of the block's rows is word e of row p; a step of 256 moves
    // (p, e) by (256 / row, 256 % row) and at most one carry
    // and eight loads are in flight before their LDS stores
    constexpr uint32_t kBatch = 8;
    const uint32_t dq = 256u / row, dr = 256u - dq * row, total = np * row;
    uint32_t p = threadIdx.x / row, e = threadIdx.x - p * row;
    for (uint32_t j0 = threadIdx.x; j0 < total; j0 += 256 * kBatch) {
        float v[kBatch];
#pragma unroll
        for (uint32_t u = 0; u < kBatch; ++u) v[u] = j0 + 256 * u < total ? src[j0 + 256 * u] : 0.0f;
#pragma unroll
        for (uint32_t u = 0; u < kBatch; ++u) {
            if (j0 + 256 * u < total) lds[p * stride + e] = v[u];
            p += dq;
            e += dr;
            if (e >= row) {
                e -= row;
                ++p;
            }
        }
    }
    __syncthreads();
    for (uint32_t c = threadIdx.x; c < np * 3; c += 256) {
        const uint32_t p = c / 3, ch = c - p * 3;
        float acc = image[3ull * (px0 + p) + ch];
        for (uint32_t k = 0; k < samples; ++k) acc += lds[p * stride + 3 * k + ch];
        image[3ull * (px0 + p) + ch] = acc;
    }
}

hipError_t launch_add_samples(float *image, const float *rad_paths, uint32_t n_px, uint32_t samples, hipStream_t s,
                              bool pixel_major) {
    const uint32_t per_block = pixel_major ? kAddLdsFloats / (3 * samples + 1) : 0u;
    if (per_block > 0) {
        const uint32_t blocks = (n_px + per_block - 1) / per_block;
        hipLaunchKernelGGL(k_add_samples_px, dim3(blocks), dim3(256), (size_t)per_block * (3 * samples + 1) * 4, s,
                           image, rad_paths, n_px, samples, per_block);
        return hipGetLastError();
    }
    hipLaunchKernelGGL(k_add_samples, dim3((n_px + 255) / 256), dim3(256), 0, s, image, rad_paths, n_px, samples,
                       pixel_major ? 1 : 0);
    return hipGetLastError();
}

hipError_t launch_set_seeds(uint64_t *p, uint64_t base, uint32_t n, hipStream_t s) {
    hipLaunchKernelGGL(k_set_seeds, dim3((n + 63) / 64), dim3(64), 0, s, p, base, n);
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

hipError_t launch_shade_stage(const pt_scene &ps, const pt_params &pp, int level, const uint64_t *frame_seed,
                              const ray_out &rays_in, const uint32_t *path_in, const float *thr,
                              const float4 *in_rec, const uint32_t *perm, const hit_soa &hits, float4 *out_rec,
                              uint32_t *keys, float *rad, uint32_t *counts, const sort_grid &grid, uint32_t n_max,
                              hipStream_t s) {
    hipLaunchKernelGGL(k_shade_stage, dim3((n_max + kShadeBlock - 1) / kShadeBlock), dim3(kShadeBlock), 0, s, ps, pp,
                       level, frame_seed, rays_in, path_in, thr, in_rec, perm, hits, out_rec, keys, rad, counts,
                       grid);
    return hipGetLastError();
}

hipError_t launch_regen_init(const regen_params &rp, float4 *out_rec, uint32_t *keys, hipStream_t s) {
    const uint32_t n = rp.lanes * rp.n_px;
    if (n == 0) return hipSuccess;
    hipLaunchKernelGGL(k_regen_init, dim3((n + 255) / 256), dim3(256), 0, s, rp, out_rec, keys);
    return hipGetLastError();
}

hipError_t launch_shade_regen(const pt_scene &ps, const pt_params &pp, const regen_params &rp,
                              const float4 *in_rec, const uint32_t *perm, const hit_soa &hits, float4 *out_rec,
                              uint32_t *keys, const uint32_t *count_in, uint32_t *count_out, const sort_grid &grid,
                              uint32_t n_max, hipStream_t s) {
    if (pp.max_depth + 1 > kMaxPathLevels) return hipErrorInvalidValue;
    hipLaunchKernelGGL(k_shade_regen, dim3((n_max + kShadeBlock - 1) / kShadeBlock), dim3(kShadeBlock), 0, s, ps, pp,
                       rp, in_rec, perm, hits, out_rec, keys, count_in, count_out, grid);
    return hipGetLastError();
}

hipError_t launch_set_u64(uint64_t *p, uint64_t v, hipStream_t s) {
    hipLaunchKernelGGL(k_set_u64, dim3(1), dim3(64), 0, s, p, v);
    return hipGetLastError();
}

} // namespace dev
} // namespace rt
