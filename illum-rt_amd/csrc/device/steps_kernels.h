// Launchers of the wavefront steps around the trace (device/steps.hip).
#pragma once

#include "pt_kernels.h"

namespace rt {
namespace dev {

hipError_t launch_add_albedo(const pt_scene &ps, const hit_soa &hits, float4 *fb, uint32_t n, hipStream_t s);
hipError_t launch_shadow_rays(const pt_scene &ps, const ray_soa &in, const hit_soa &hits, vec3f light,
                              const ray_out &out, uint32_t n, hipStream_t s);
hipError_t launch_integrate_light(const pt_scene &ps, const ray_soa &cam, const hit_soa &hits, const ray_soa &shadow,
                                  const uint8_t *occluded, vec3f light, vec3f power, float4 *fb, uint32_t n,
                                  hipStream_t s);

} // namespace dev
} // namespace rt
