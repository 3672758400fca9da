// Types and launchers of the wavefront path-tracing kernels (device/pt.hip).
#pragma once

#include <hip/hip_runtime.h>

#include "../core/layout.h"
#include "../core/raygen.h"

#include <cstdint>

namespace rt {
namespace dev {

// Writable ray planes.
struct ray_out {
    float *ox, *oy, *oz, *tmin, *dx, *dy, *dz, *tmax;
    ray_soa as_in() const { return ray_soa{ox, oy, oz, tmin, dx, dy, dz, tmax}; }
};

// Shading data of a device scene: three float4 per triangle
// (A.xyz + material id bits, B.xyz, C.xyz) and two float4 per material
// (albedo, emission).
struct pt_scene {
    const float4 *tri_geo;
    const float4 *mats;
    uint32_t n_tris;
    // corner shading normals, 3 float4 per triangle; null when the scene has
    // none and the geometric normal shades
    const float4 *tri_ns = nullptr;
    // A hit the shading steps may use: finite t below FLT_MAX (misses are
    // FLT_MAX, a traversal that lost stack entries is NaN) on a triangle of
    // this scene.
    __device__ __forceinline__ bool valid(float t, uint32_t id) const { return t < FLT_MAX && id < n_tris; }
    // diff_geom::ns (libgi/rt.cpp:7-15): the corner normals interpolated at
    // (beta, gamma) and normalized; ng for a scene without normals.
    __device__ __forceinline__ vec3f shading_normal(uint32_t tri, float beta, float gamma, vec3f ng) const {
        if (!tri_ns) return ng;
        const float4 *q = tri_ns + 3 * (size_t)tri;
        float4 a = q[0], b = q[1], c = q[2];
        float w = 1.0f - beta - gamma;
        return normalize(w * vec3f{a.x, a.y, a.z} + beta * vec3f{b.x, b.y, b.z} + gamma * vec3f{c.x, c.y, c.z});
    }
};

struct pt_params {
    int max_depth;      // last trace level (primary = 0)
    int rr_start;       // first level whose ray is subject to Russian roulette
    float rr_survive;   // survival probability
    // Pixels per sample: path p is pixel p % n_px of sample p / n_px, whose
    // frame seed is frame_seed[p / n_px] (several samples in flight at once).
    uint32_t n_px;
    // 0: a path that meets an emitter adds its emission and ends there, as
    // simple_pt::path does (gi/pt.cpp:86-89); 1: it adds the emission and
    // bounces on like any other surface.
    int emitters_reflect;
    // > 1: paths are numbered pixel-major instead -- path p is sample
    // p % px_samples of pixel p / px_samples -- so the samples of a pixel sit
    // in adjacent queue slots and share wavefronts (samples in flight).
    uint32_t px_samples = 0;
    __host__ __device__ __forceinline__ void decode(uint32_t path, uint32_t &sample, uint32_t &pixel) const {
        if (px_samples > 1) {
            pixel = path / px_samples;
            sample = path - pixel * px_samples;
        } else {
            sample = path / n_px;
            pixel = path - sample * n_px;
        }
    }
};

// Sub-pixel offset (ox, oy) in [-0.5, 0.5)^2 of path `path`'s camera ray:
// the first two uniforms of its PCG32 stream (frame seed, pixel,
// kCameraStream), as simple_pt::sample_pixel offsets every sample
// (gi/pt.cpp:36).
RT_HD void pixel_jitter(uint64_t frame_seed, uint32_t pixel, float &ox, float &oy) {
    pcg32 rng(pcg_stream_id(frame_seed, pixel, kCameraStream));
    ox = rng.next_float() - 0.5f;
    oy = rng.next_float() - 0.5f;
}

// Camera rays for n_paths paths (0: one per pixel), path i looking through
// pixel i % (w*h) -- at the pixel centre, or, when jitter_seeds is non-null,
// at the offset pixel_jitter gives for frame seed jitter_seeds[i / (w*h)];
// resets path ids and throughput (when non-null), clears the image
// accumulator `rad` (per pixel, when clear_radiance) and the per-path
// radiance `rad_paths` (when non-null).
// px_samples k > 1: paths numbered pixel-major, as pt_params::px_samples
// (n_paths = k * n_px).
hipError_t launch_camera_rays(const cam_frame &f, const ray_out &out, uint32_t *path_ids, float *thr, float *rad,
                              bool clear_radiance, hipStream_t s, uint32_t n_paths = 0, float *rad_paths = nullptr,
                              const uint64_t *jitter_seeds = nullptr, uint32_t px_samples = 0);
// image[pixel] += the pixel's per-path radiance for samples k = 0 .. samples-1
// in order: rad_paths[k * n_px + pixel], or rad_paths[pixel * samples + k]
// when `pixel_major`.
hipError_t launch_add_samples(float *image, const float *rad_paths, uint32_t n_px, uint32_t samples, hipStream_t s,
                              bool pixel_major = false);
// p[k] = base + k for k < n (the frame seeds of consecutive samples).
hipError_t launch_set_seeds(uint64_t *p, uint64_t base, uint32_t n, hipStream_t s);
hipError_t launch_shade_plain(const pt_scene &ps, const pt_params &pp, int level, const uint64_t *frame_seed,
                              const ray_out &rays, const hit_soa &hits, float *thr, float *rad, uint32_t *counts,
                              uint32_t n, hipStream_t s);
hipError_t launch_shade_compact(const pt_scene &ps, const pt_params &pp, int level, const uint64_t *frame_seed,
                                const ray_out &rays_in, const hit_soa &hits, const uint32_t *path_in,
                                const ray_out &rays_out, uint32_t *path_out, float *thr, float *rad,
                                uint32_t *counts, uint32_t n_max, hipStream_t s);
// Sorted bounce queues with a pixel's samples in flight: ray keys combine the
// direction's cell in a 2^kDirHalfBits-square grid over its octahedral
// projection (1024 cells of similar solid angle, Morton order) and the Morton
// code of the origin's cell in a grid of 2^kSortCellBits cells per axis over
// the scene box (128^3), in three levels, coarse above fine: 31 bits, four
// radix passes.
constexpr int kDirHalfBits = 5;
constexpr int kSortCellBits = 7;
constexpr int kRayKeyBits = 2 * kDirHalfBits + 3 * kSortCellBits;
// One sample per pixel in flight sorts by the direction octant (3 bits) and
// the origin's cell in a 16^3 grid instead, and so does path regeneration,
// which puts camera rays above bounce rays with one more key bit: key and
// camera bit fit 16 bits, two passes.
constexpr int kOctantCellBits = 4;
constexpr int kOctantKeyBits = 3 + 3 * kOctantCellBits;
struct sort_grid {
    float lo[3];      // scene box minimum
    float scale[3];   // cells per axis / extent (0 for a flat axis)
};
// Ray records of the sorting form: kRayRecord float4 per ray,
// (ox, oy, oz, t_min | dx, dy, dz, t_max | path id bits, throughput).
constexpr uint32_t kRayRecord = 3;
// Shade step of the sorting form: slot i < counts[level] is the ray of
// record src = perm[i] (or i when perm is null).  Camera level (in_rec
// null): direction from the planes of rays_in, path from path_in[i],
// throughput from thr.  Later levels: direction, path and throughput from
// record src of in_rec.  Live next-level rays are appended to out_rec and
// keys and counted in counts[level + 1].
hipError_t launch_shade_stage(const pt_scene &ps, const pt_params &pp, int level, const uint64_t *frame_seed,
                              const ray_out &rays_in, const uint32_t *path_in, const float *thr,
                              const float4 *in_rec, const uint32_t *perm, const hit_soa &hits, float4 *out_rec,
                              uint32_t *keys, float *rad, uint32_t *counts, const sort_grid &grid, uint32_t n_max,
                              hipStream_t s);
// Path regeneration (a fixed pool of lanes * n_px path slots, each slot
// rendering its pixel's samples one after another; device/pt.hip).
constexpr int kMaxPathLevels = 65;   // max_depth <= 64
// Regeneration keys: bounce-ray keys as with samples in flight (the 31-bit
// keys with more than one slot per pixel, the 15-bit octant keys with one),
// camera-ray keys above them (one more bit).
constexpr int kRegenKeyBits = kOctantKeyBits + 1;
constexpr int kRegenFineKeyBits = kRayKeyBits + 1;
__host__ __device__ inline int regen_key_bits(uint32_t lanes) { return lanes > 1 ? kRegenFineKeyBits : kRegenKeyBits; }
struct regen_params {
    cam_frame f;
    const uint64_t *seeds;     // frame seed of each sample of the render call
    int jitter;                // camera rays through pixel_jitter offsets
    uint32_t n_px;
    uint32_t samples;          // samples per pixel in the render call (< 2^24)
    uint32_t lanes;            // slots per pixel (<= samples)
    float *acc;                // per-slot radiance when lanes > 1, else null
    float *image;              // per-pixel radiance the paths add to when acc is null
    uint32_t *level_counts;    // rays shaded per level, added to
};
// Writes the first sample of every slot to out_rec (and keys, when non-null)
// in slot order; the queue length is lanes * n_px.
hipError_t launch_regen_init(const regen_params &rp, float4 *out_rec, uint32_t *keys, hipStream_t s);
// One iteration's shade step over the *count_in rays of in_rec (through
// perm when non-null); appends continued paths and regenerated camera rays
// to out_rec / keys and counts them in *count_out.
hipError_t launch_shade_regen(const pt_scene &ps, const pt_params &pp, const regen_params &rp,
                              const float4 *in_rec, const uint32_t *perm, const hit_soa &hits, float4 *out_rec,
                              uint32_t *keys, const uint32_t *count_in, uint32_t *count_out, const sort_grid &grid,
                              uint32_t n_max, hipStream_t s);
hipError_t launch_set_u64(uint64_t *p, uint64_t v, hipStream_t s);
// totals[i] += counts[i] for i < levels.
hipError_t launch_add_counts(unsigned long long *totals, const uint32_t *counts, int levels, hipStream_t s);
// Incoherent benchmark rays between two balls (device/pt.hip, k_incoherent_rays).
hipError_t launch_incoherent_rays(vec3f s1, vec3f s2, float r1, float r2, float r_max, uint64_t seed,
                                  const ray_out &out, uint32_t n, hipStream_t s);
hipError_t launch_cosine_bounce(const pt_scene &ps, const ray_soa &in, const hit_soa &hits, const ray_out &out,
                                uint64_t seed, int level, uint32_t n, hipStream_t s);

} // namespace dev
} // namespace rt
