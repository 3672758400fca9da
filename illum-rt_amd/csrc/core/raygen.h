// Per-ray generation math shared by the host generators and the device
// ray-generation / bounce kernels.
//
//   camera rays     libgi/camera.h:18-24, libgi/camera.cpp:44-61
//   cosine bounce   rt/cpu/bounce.cpp:41-72 with libgi/sampling.h:13-36 and
//                   libgi/util.h:212-233 (align, flip to the ray side)
//   hit point       libgi/rt.cpp:7-15 (x and the geometric normal)
//
// Camera rays use only +,-,*,/ and sqrtf, all correctly rounded on both sides,
// so host and device produce identical bits.  Bounce directions go through
// cosf/sinf, whose host (glibc) and device (ocml) results may differ by an ulp.
#pragma once

#include "rng.h"
#include "vec.h"

namespace rt {

constexpr float kPi = 3.14159265358979323846f;
constexpr float kRayEps = 1e-4f;   // ray::eps, libgi/rt.h:25

// Precomputed camera basis (camera::update_frustum + the basis of cam_ray).
struct cam_frame {
    vec3f pos, dir, U, V;
    float near_w, near_h;
    int w, h;
};

// The host computes near_h with tanf, which the device never needs.
inline cam_frame make_cam_frame(vec3f pos, vec3f dir, vec3f up, float fovy, int w, int h) {
    cam_frame f;
    f.pos = pos;
    f.dir = dir;
    f.w = w;
    f.h = h;
    float aspect = (float)w / (float)h;
    f.near_h = tanf(kPi * fovy * 0.5f / 180.0f);
    f.near_w = aspect * f.near_h;
    f.U = normalize(cross(dir, up));
    f.V = normalize(cross(f.U, dir));
    return f;
}

// Pixel (x, y) with sub-pixel offset (ox, oy); returns the direction.
RT_HD vec3f cam_dir(const cam_frame &f, int x, int y, float ox = 0.0f, float oy = 0.0f) {
    float u = ((float)x + 0.5f + ox) / (float)f.w * 2.0f - 1.0f;
    float v = ((float)y + 0.5f + oy) / (float)f.h * 2.0f - 1.0f;
    u = f.near_w * u;
    v = f.near_h * v;
    return normalize(f.dir + f.U * u + f.V * v);
}

// align(v, axis) of libgi/util.h:212-218.
RT_HD vec3f align(vec3f v, vec3f axis) {
    float s = copysignf(1.0f, axis.z);
    vec3f w{v.x, v.y, v.z * s};
    vec3f h{axis.x, axis.y, axis.z + s};
    float k = dot(w, h) / (1.0f + (axis.z < 0 ? -axis.z : axis.z));
    return k * h - w;
}

// cosine_sample_hemisphere(uniform_sample_disk(xi)), libgi/sampling.h:13-36.
RT_HD vec3f cosine_hemisphere(float xi0, float xi1) {
    float r = sqrtf(xi0);
    float theta = 2.0f * kPi * xi1;
    float dx = r * cosf(theta), dy = r * sinf(theta);
    float z = 1.0f - dx * dx - dy * dy;
    return vec3f{dx, dy, z > 0 ? sqrtf(z) : 0.0f};
}

// Hit point and geometric normal of libgi/rt.cpp:9-12.
RT_HD vec3f hit_point(vec3f A, vec3f B, vec3f C, float beta, float gamma) {
    float w = 1.0f - beta - gamma;
    return w * A + beta * B + gamma * C;
}
RT_HD vec3f geom_normal(vec3f A, vec3f B, vec3f C) { return normalize(cross(B - A, C - A)); }

// Cosine-weighted bounce direction about the geometric normal ng flipped to
// face away from the incoming direction; flip_normals_to_ray decides the flip
// by the shading normal ns (flip when dot(d, ns) > 0), and
// sample_cos_weighted_dir aligns the sample to ng (rt/cpu/bounce.cpp:41-72).
RT_HD vec3f cosine_bounce(vec3f ng, vec3f ns, vec3f in_dir, float xi0, float xi1) {
    if (dot(in_dir, ns) > 0) ng = -ng;
    return align(cosine_hemisphere(xi0, xi1), ng);
}

// Uniform direction on the sphere: z = 1 - 2 xi0, phi = 2 pi xi1.
RT_HD vec3f uniform_sphere(float xi0, float xi1) {
    float z = 1.0f - 2.0f * xi0;
    float r0 = 1.0f - z * z;
    float r = sqrtf(r0 < 0 ? 0.0f : r0);
    float phi = 2.0f * kPi * xi1;
    return vec3f{r * cosf(phi), r * sinf(phi), z};
}

} // namespace rt
