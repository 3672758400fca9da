// Launchers of the wavefront steps around the trace (device/steps.hip).
#pragma once

#include "pt_kernels.h"

namespace rt {
namespace dev {

hipError_t launch_add_albedo(const pt_scene &ps, const hit_soa &hits, float4 *fb, uint32_t n, hipStream_t s);
hipError_t launch_shadow_rays(const pt_scene &ps, const ray_soa &in, const hit_soa &hits, vec3f light,
                              const ray_out &out, uint32_t n, hipStream_t s);
hipError_t launch_integrate_light(const pt_scene &ps, const ray_soa &cam, const hit_soa &hits, const ray_soa &shadow,
                                  const uint8_t *occluded, vec3f light, vec3f col, float4 *fb, uint32_t n,
                                  hipStream_t s);

// A scene's light list on the device (core/lights.h): 4 float4 per light,
// the power of each light, the running sums divided by their total (n + 1
// entries) and that total.
struct light_table {
    const float4 *rec;
    const float *f;
    const float *cdf;
    uint32_t n;
    float integral;
};

enum { kDirectUniform = 0, kDirectCosine = 1, kDirectLight = 2 };

// The direct-lighting sample steps: sample_uniform_dir, sample_cos_weighted_dir
// and sample_light_dir (rt/cpu/bounce.cpp:11-106) for camera ray i and its hit.
// Random numbers: PCG32 stream (seed, i, 0) -- light mode draws the light
// index first, then the point on it.  light_col (3 floats per ray) only in
// light mode.
hipError_t launch_sample_direct(const pt_scene &ps, const light_table &lt, int mode, const ray_soa &cam,
                                const hit_soa &hits, uint64_t seed, const ray_out &out, float *pdf, float *light_col,
                                float4 *fb, uint32_t n, hipStream_t s);
// integrate_dir_sample (rt/cpu/bounce.cpp:108-130): the emission of the
// surface the sampled direction hit.
hipError_t launch_integrate_dir(const pt_scene &ps, const ray_soa &cam, const hit_soa &hits, const ray_soa &shadow,
                                const hit_soa &shadow_hits, const float *pdf, float4 *fb, uint32_t n, hipStream_t s);
// integrate_light_sample (rt/cpu/bounce.cpp:132-151): the sampled light's
// colour where the shadow ray is valid and unoccluded.
hipError_t launch_integrate_light_sample(const pt_scene &ps, const ray_soa &cam, const hit_soa &hits,
                                         const ray_soa &shadow, const uint8_t *occluded, const float *pdf,
                                         const float *light_col, float4 *fb, uint32_t n, hipStream_t s);

} // namespace dev
} // namespace rt
