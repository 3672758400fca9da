// Wavefront steps of the reference's primary and direct-lighting algorithms
// that sit around the trace, on the GPU:
//   add hitpoint albedo        rt/cpu/wavefront.cpp:105-119
//   sample uniform dir         rt/cpu/bounce.cpp:11-39
//   sample cos weighted dir    rt/cpu/bounce.cpp:41-72
//   sample light dir           rt/cpu/bounce.cpp:74-106 over the scene's
//                              light list (core/lights.h)
//   integrate dir sample       rt/cpu/bounce.cpp:108-130
//   integrate light sample     rt/cpu/bounce.cpp:132-151
// and the point-light pair of the first version of this path (shadow ray to
// one point light, its integration).  The framebuffer is float4 per pixel
// (rgb, weight), as the reference's.  Materials are Lambertian
// (lambertian_reflection::f, libgi/material.cpp:55-65: albedo / pi where the
// incoming direction lies on the side of the shading normal).  The shading
// normal is the corners' normals interpolated at the hit (pt_scene::
// shading_normal), or the geometric normal in a scene without normals (the
// face normals the reference's loader generates then); the sample steps flip
// the normals to the camera ray's side by it (flip_normals_to_ray) and align
// their samples to the geometric normal, the integration steps weigh by it
// unflipped, as the reference does.
#include "pt_kernels.h"
#include "steps_kernels.h"

#include "../core/lights.h"

namespace rt {
namespace dev {

namespace {
__device__ __forceinline__ vec3f f3(float4 v) { return vec3f{v.x, v.y, v.z}; }

// Lambertian f(hit, w_out, w_in) * cdot(w_in, ns) for the unflipped shading
// normal ns.
__device__ __forceinline__ vec3f lambert_cos(vec3f albedo, vec3f ns, vec3f w_in, float &cosv) {
    cosv = cdot(w_in, ns);
    return dot(w_in, ns) > 0 ? kOneOverPi * albedo : v3(0.0f);
}

__device__ __forceinline__ void add_fb(float4 *fb, uint32_t i, vec3f rgb, float w) {
    float4 acc = fb[i];
    acc.x += rgb.x;
    acc.y += rgb.y;
    acc.z += rgb.z;
    acc.w += w;
    fb[i] = acc;
}
} // namespace

__global__ __launch_bounds__(256) void k_add_albedo(pt_scene ps, hit_soa hits, float4 *fb, uint32_t n) {
    uint32_t i = blockIdx.x * 256u + threadIdx.x;
    if (i >= n) return;
    float4 acc = fb[i];
    if (ps.valid(hits.t[i], hits.id[i])) {
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
    if (ps.valid(hits.t[i], hits.id[i]) && in.tmax[i] >= 0.0f) {
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

// integrate_light_sample for one point light of colour col (pointlight:
// col / dist^2, pdf 1): col / dist^2 * albedo / pi * cdot(w_in, ns) where the
// shadow ray is valid and unoccluded; dist = |light - hit point|.
__global__ __launch_bounds__(256) void k_integrate_light(pt_scene ps, ray_soa cam, hit_soa hits, ray_soa shadow,
                                                        const uint8_t *occluded, vec3f light, vec3f col, float4 *fb,
                                                        uint32_t n) {
    uint32_t i = blockIdx.x * 256u + threadIdx.x;
    if (i >= n) return;
    vec3f rad = v3(0.0f);
    bool valid_shadow = shadow.tmax[i] > 0.0f;
    if (ps.valid(hits.t[i], hits.id[i]) && valid_shadow && !occluded[i]) {
        const float4 *g = ps.tri_geo + 3 * (size_t)hits.id[i];
        float4 g0 = g[0];
        vec3f A = f3(g0), B = f3(g[1]), C = f3(g[2]);
        vec3f x = hit_point(A, B, C, hits.u[i], hits.v[i]);
        vec3f to = light - x;
        float dist = sqrtf(dot(to, to));
        vec3f c = col / (dist * dist);
        vec3f wi{shadow.dx[i], shadow.dy[i], shadow.dz[i]};
        float cosv;
        vec3f ns = ps.shading_normal(hits.id[i], hits.u[i], hits.v[i], geom_normal(A, B, C));
        vec3f f = lambert_cos(f3(ps.mats[2 * __float_as_uint(g0.w)]), ns, wi, cosv);
        rad = c * f * cosv / 1.0f;
    }
    add_fb(fb, i, rad, 1.0f);
}

__global__ __launch_bounds__(256) void k_sample_direct(pt_scene ps, light_table lt, int mode, ray_soa cam,
                                                      hit_soa hits, uint64_t seed, ray_out out, float *pdf_out,
                                                      float *light_col, float4 *fb, uint32_t n) {
    uint32_t i = blockIdx.x * 256u + threadIdx.x;
    if (i >= n) return;
    float pdf = mode == kDirectUniform ? kOneOver2Pi : (mode == kDirectCosine ? kOneOverPi : 0.0f);
    vec3f o = v3(0.0f), d = v3(1.0f), lc = v3(0.0f);
    float tmax = -FLT_MAX;
    if (ps.valid(hits.t[i], hits.id[i])) {
        const float4 *g = ps.tri_geo + 3 * (size_t)hits.id[i];
        float4 g0 = g[0];
        vec3f A = f3(g0), B = f3(g[1]), C = f3(g[2]);
        uint32_t m = __float_as_uint(g0.w);
        vec3f emi = f3(ps.mats[2 * m + 1]);
        vec3f ng = geom_normal(A, B, C);
        vec3f ns = ps.shading_normal(hits.id[i], hits.u[i], hits.v[i], ng);
        vec3f in_dir{cam.dx[i], cam.dy[i], cam.dz[i]};
        if (dot(in_dir, ns) > 0) ng = -ng;   // flip_normals_to_ray (ns too, unused here)
        vec3f x = hit_point(A, B, C, hits.u[i], hits.v[i]);
        if (emi != v3(0.0f)) {
            add_fb(fb, i, emi, mode == kDirectLight ? 0.0f : 1.0f);
        } else {
            pcg32 rng(pcg_stream_id(seed, i, 0));
            if (mode == kDirectUniform) {
                float xi0 = rng.next_float(), xi1 = rng.next_float();
                o = x;
                d = align(uniform_hemisphere(xi0, xi1), ng);
                tmax = FLT_MAX;
            } else if (mode == kDirectCosine) {
                float xi0 = rng.next_float(), xi1 = rng.next_float();
                o = x;
                d = align(cosine_hemisphere(xi0, xi1), ng);
                pdf *= cdot(d, ng);
                tmax = FLT_MAX;
            } else {
                float l_pdf;
                uint32_t l = sample_index(lt.cdf, lt.f, lt.n, lt.integral, rng.next_float(), l_pdf);
                float xi0 = rng.next_float(), xi1 = rng.next_float();
                const float4 *r = lt.rec + 4 * (size_t)l;
                float4 r0 = r[0], r1 = r[1], r2 = r[2], r3 = r[3];
                vec3f col = f3(r3), w_i;
                float a_pdf = 1.0f, dist;
                if (__float_as_uint(r0.w) == kLightPoint) {   // pointlight::sample_Li
                    w_i = f3(r0) - x;
                    dist = sqrtf(dot(w_i, w_i));
                    w_i = w_i / dist;
                    col = col / (dist * dist);
                } else {   // trianglelight::sample_Li
                    vec3f la = f3(r0), lb = f3(r1), lcn = f3(r2);
                    float su0 = sqrtf(xi0);
                    float bx = 1.0f - su0, by = xi1 * su0;
                    vec3f target = (1.0f - bx - by) * la + bx * lb + by * lcn;
                    // the light's normal at the sampled point (its triangle id in r2.w)
                    vec3f ln = ps.shading_normal(__float_as_uint(r2.w), bx, by, geom_normal(la, lb, lcn));
                    w_i = target - x;
                    dist = sqrtf(dot(w_i, w_i));
                    w_i = w_i / dist;
                    float cos_l = dot(ln, -w_i);
                    if (cos_l <= 0.0f) {
                        col = v3(0.0f);
                        a_pdf = 0.0f;
                    } else {
                        a_pdf = dist * dist / (cos_l * r1.w);
                    }
                }
                if (col != v3(0.0f)) {
                    o = x;
                    d = w_i;
                    tmax = dist - kRayEps;   // ray::length_exclusive
                    lc = col;
                }
                pdf = a_pdf * l_pdf;
            }
        }
    }
    out.ox[i] = o.x;
    out.oy[i] = o.y;
    out.oz[i] = o.z;
    out.tmin[i] = kRayEps;
    out.dx[i] = d.x;
    out.dy[i] = d.y;
    out.dz[i] = d.z;
    out.tmax[i] = tmax;
    pdf_out[i] = pdf;
    if (light_col) {
        light_col[3 * (size_t)i] = lc.x;
        light_col[3 * (size_t)i + 1] = lc.y;
        light_col[3 * (size_t)i + 2] = lc.z;
    }
}

__global__ __launch_bounds__(256) void k_integrate_dir(pt_scene ps, ray_soa cam, hit_soa hits, ray_soa shadow,
                                                      hit_soa shadow_hits, const float *pdf, float4 *fb, uint32_t n) {
    uint32_t i = blockIdx.x * 256u + threadIdx.x;
    if (i >= n) return;
    vec3f rad = v3(0.0f);
    if (ps.valid(hits.t[i], hits.id[i]) && ps.valid(shadow_hits.t[i], shadow_hits.id[i])) {
        uint32_t lm = __float_as_uint(ps.tri_geo[3 * (size_t)shadow_hits.id[i]].w);
        vec3f emi = f3(ps.mats[2 * lm + 1]);
        if (emi != v3(0.0f)) {
            const float4 *g = ps.tri_geo + 3 * (size_t)hits.id[i];
            float4 g0 = g[0];
            vec3f ns = ps.shading_normal(hits.id[i], hits.u[i], hits.v[i], geom_normal(f3(g0), f3(g[1]), f3(g[2])));
            vec3f w_in{shadow.dx[i], shadow.dy[i], shadow.dz[i]};
            float cosv;
            vec3f f = lambert_cos(f3(ps.mats[2 * __float_as_uint(g0.w)]), ns, w_in, cosv);
            rad = emi * f * cosv / pdf[i];
        }
    }
    add_fb(fb, i, rad, 1.0f);
}

__global__ __launch_bounds__(256) void k_integrate_light_sample(pt_scene ps, ray_soa cam, hit_soa hits,
                                                               ray_soa shadow, const uint8_t *occluded,
                                                               const float *pdf, const float *light_col, float4 *fb,
                                                               uint32_t n) {
    uint32_t i = blockIdx.x * 256u + threadIdx.x;
    if (i >= n) return;
    vec3f rad = v3(0.0f);
    if (ps.valid(hits.t[i], hits.id[i]) && shadow.tmax[i] > 0.0f && !occluded[i]) {
        const float4 *g = ps.tri_geo + 3 * (size_t)hits.id[i];
        float4 g0 = g[0];
        vec3f ns = ps.shading_normal(hits.id[i], hits.u[i], hits.v[i], geom_normal(f3(g0), f3(g[1]), f3(g[2])));
        vec3f w_in{shadow.dx[i], shadow.dy[i], shadow.dz[i]};
        vec3f lc{light_col[3 * (size_t)i], light_col[3 * (size_t)i + 1], light_col[3 * (size_t)i + 2]};
        float cosv;
        vec3f f = lambert_cos(f3(ps.mats[2 * __float_as_uint(g0.w)]), ns, w_in, cosv);
        rad = lc * f * cosv / pdf[i];
    }
    add_fb(fb, i, rad, 1.0f);
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
                                  const uint8_t *occluded, vec3f light, vec3f col, float4 *fb, uint32_t n,
                                  hipStream_t s) {
    if (n == 0) return hipSuccess;
    hipLaunchKernelGGL(k_integrate_light, dim3((n + 255) / 256), dim3(256), 0, s, ps, cam, hits, shadow, occluded,
                       light, col, fb, n);
    return hipGetLastError();
}

hipError_t launch_sample_direct(const pt_scene &ps, const light_table &lt, int mode, const ray_soa &cam,
                                const hit_soa &hits, uint64_t seed, const ray_out &out, float *pdf, float *light_col,
                                float4 *fb, uint32_t n, hipStream_t s) {
    if (n == 0) return hipSuccess;
    hipLaunchKernelGGL(k_sample_direct, dim3((n + 255) / 256), dim3(256), 0, s, ps, lt, mode, cam, hits, seed, out,
                       pdf, light_col, fb, n);
    return hipGetLastError();
}

hipError_t launch_integrate_dir(const pt_scene &ps, const ray_soa &cam, const hit_soa &hits, const ray_soa &shadow,
                                const hit_soa &shadow_hits, const float *pdf, float4 *fb, uint32_t n, hipStream_t s) {
    if (n == 0) return hipSuccess;
    hipLaunchKernelGGL(k_integrate_dir, dim3((n + 255) / 256), dim3(256), 0, s, ps, cam, hits, shadow, shadow_hits,
                       pdf, fb, n);
    return hipGetLastError();
}

hipError_t launch_integrate_light_sample(const pt_scene &ps, const ray_soa &cam, const hit_soa &hits,
                                         const ray_soa &shadow, const uint8_t *occluded, const float *pdf,
                                         const float *light_col, float4 *fb, uint32_t n, hipStream_t s) {
    if (n == 0) return hipSuccess;
    hipLaunchKernelGGL(k_integrate_light_sample, dim3((n + 255) / 256), dim3(256), 0, s, ps, cam, hits, shadow,
                       occluded, pdf, light_col, fb, n);
    return hipGetLastError();
}

} // namespace dev
} // namespace rt
