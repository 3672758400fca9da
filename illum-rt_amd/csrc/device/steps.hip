// Wavefront steps of the reference's primary and direct-lighting algorithms
// that sit around the trace, on the GPU:
//   add hitpoint albedo        rt/cpu/wavefront.cpp:105-119
//   sample light dir (point)   rt/cpu/bounce.cpp:74-106 with a point light;
//                              shadow ray = ray(hit.x, l - x) with
//                              length_exclusive(|l - x|) (libgi/rt.h:30)
//   integrate light sample     rt/cpu/bounce.cpp:132-151 for a diffuse BRDF
// The framebuffer is float4 per pixel (rgb, weight), as the reference's.
#include "pt_kernels.h"
#include "steps_kernels.h"

namespace rt {
namespace dev {

namespace {
__device__ __forceinline__ vec3f f3(float4 v) { return vec3f{v.x, v.y, v.z}; }
} // namespace

__global__ __launch_bounds__(256) void k_add_albedo(pt_scene ps, hit_soa hits, float4 *fb, uint32_t n) {
    uint32_t i = blockIdx.x * 256u + threadIdx.x;
    if (i >= n) return;
    float4 acc = fb[i];
    if (hits.t[i] != FLT_MAX) {
        uint32_t m = __float_as_uint(ps.tri_geo[3 * (size_t)hits.id[i]].w);
        float4 a = ps.mats[2 * m];
        acc.x += a.x;
        acc.y += a.y;
        acc.z += a.z;
    }
    acc.w += 1.0f;
    fb[i] = acc;
}

__global__ __launch_bounds__(256) void k_shadow_rays(pt_scene ps, ray_soa in, hit_soa hits, vec3f light,
                                                    ray_out out, uint32_t n) {
    uint32_t i = blockIdx.x * 256u + threadIdx.x;
    if (i >= n) return;
    vec3f o = v3(0.0f), d = v3(1.0f);
    float tmax = -FLT_MAX;
    if (hits.t[i] != FLT_MAX && in.tmax[i] >= 0.0f) {
        const float4 *g = ps.tri_geo + 3 * (size_t)hits.id[i];
        vec3f A = f3(g[0]), B = f3(g[1]), C = f3(g[2]);
        o = hit_point(A, B, C, hits.u[i], hits.v[i]);
        vec3f to = light - o;
        float dist = sqrtf(dot(to, to));
        d = to / dist;
        tmax = dist - kRayEps;
    }
    out.ox[i] = o.x;
    out.oy[i] = o.y;
    out.oz[i] = o.z;
    out.tmin[i] = kRayEps;
    out.dx[i] = d.x;
    out.dy[i] = d.y;
    out.dz[i] = d.z;
    out.tmax[i] = tmax;
}

// Diffuse direct light from a point light of power `power` (radiant intensity
// power / 4pi): L = albedo/pi * I * cos / dist^2 when the shadow ray is
// unoccluded; the geometric normal is flipped to the camera ray's side.
__global__ __launch_bounds__(256) void k_integrate_light(pt_scene ps, ray_soa cam, hit_soa hits, ray_soa shadow,
                                                        const uint8_t *occluded, vec3f light, vec3f power,
                                                        float4 *fb, uint32_t n) {
    uint32_t i = blockIdx.x * 256u + threadIdx.x;
    if (i >= n) return;
    float4 acc = fb[i];
    bool valid_shadow = shadow.tmax[i] > 0.0f;
    if (hits.t[i] != FLT_MAX && valid_shadow && !occluded[i]) {
        const float4 *g = ps.tri_geo + 3 * (size_t)hits.id[i];
        float4 g0 = g[0];
        vec3f A = f3(g0), B = f3(g[1]), C = f3(g[2]);
        vec3f ng = geom_normal(A, B, C);
        vec3f in_dir{cam.dx[i], cam.dy[i], cam.dz[i]};
        if (dot(in_dir, ng) > 0) ng = -ng;
        vec3f wi{shadow.dx[i], shadow.dy[i], shadow.dz[i]};
        float cosv = dot(wi, ng);
        if (cosv > 0) {
            float dist = shadow.tmax[i] + kRayEps;
            float4 a = ps.mats[2 * __float_as_uint(g0.w)];
            float s = cosv / (dist * dist) / (4.0f * kPi) / kPi;
            acc.x += a.x * power.x * s;
            acc.y += a.y * power.y * s;
            acc.z += a.z * power.z * s;
        }
    }
    acc.w += 1.0f;
    fb[i] = acc;
}

hipError_t launch_add_albedo(const pt_scene &ps, const hit_soa &hits, float4 *fb, uint32_t n, hipStream_t s) {
    if (n == 0) return hipSuccess;
    hipLaunchKernelGGL(k_add_albedo, dim3((n + 255) / 256), dim3(256), 0, s, ps, hits, fb, n);
    return hipGetLastError();
}

hipError_t launch_shadow_rays(const pt_scene &ps, const ray_soa &in, const hit_soa &hits, vec3f light,
                              const ray_out &out, uint32_t n, hipStream_t s) {
    if (n == 0) return hipSuccess;
    hipLaunchKernelGGL(k_shadow_rays, dim3((n + 255) / 256), dim3(256), 0, s, ps, in, hits, light, out, n);
    return hipGetLastError();
}

hipError_t launch_integrate_light(const pt_scene &ps, const ray_soa &cam, const hit_soa &hits, const ray_soa &shadow,
                                  const uint8_t *occluded, vec3f light, vec3f power, float4 *fb, uint32_t n,
                                  hipStream_t s) {
    if (n == 0) return hipSuccess;
    hipLaunchKernelGGL(k_integrate_light, dim3((n + 255) / 256), dim3(256), 0, s, ps, cam, hits, shadow, occluded,
                       light, power, fb, n);
    return hipGetLastError();
}

} // namespace dev
} // namespace rt
