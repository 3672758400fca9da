// Light sources and direct-lighting math shared by the host (light table)
// and the device (the direct-lighting wavefront steps).
//
//   light list      every emissive triangle in triangle order
//                   (scene::find_light_geometry, libgi/scene.cpp:366-370),
//                   then the point lights (pointlight, driver/interaction.cpp:723-753)
//   light power     trianglelight::power = emissive * 0.5 * |e1 x e2| * pi,
//                   pointlight::power = 4 pi col (libgi/scene.cpp:494-527),
//                   weighed by luma (libgi/color.h)
//   distribution    distribution_1d over the powers: running sums, then
//                   divided by their total; sample_index = lower_bound - 1
//                   (libgi/discrete_distributions.cpp:25-52)
//   sample_Li       triangle: uniform point (uniform_sample_triangle,
//                   libgi/sampling.h:43-46), pdf = dist^2 / (cos_light area)
//                   with the light's normal interpolated at that point,
//                   nothing behind the light (libgi/scene.cpp:534-559);
//                   point: col / dist^2, pdf 1 (libgi/scene.cpp:500-508)
//   hemispheres     uniform_sample_hemisphere, cosine_sample_hemisphere
//                   (libgi/sampling.h:19-38)
//
// Float operations follow the reference's expressions in order (no FMA
// contraction on either side), so host and device agree bit for bit.
#pragma once

#include "vec.h"

namespace rt {

constexpr float kOneOverPi = (float)(1.0 / 3.14159265358979323846);          // one_over_pi, libgi/rt.h:16
constexpr float kOneOver2Pi = (float)(1.0 / (2.0 * 3.14159265358979323846)); // one_over_2pi, libgi/rt.h:17
constexpr float kPiF = 3.14159265358979323846f;                              // pi, libgi/rt.h:15

enum : uint32_t { kLightTriangle = 0, kLightPoint = 1 };

// One light, as 4 float4 on the device: (P0, type bits), (P1, area),
// (P2, triangle id bits), (colour, 0).  Triangle: P0..P2 = A, B, C and
// colour = emission; point: P0 = position and colour = col.
struct light_rec {
    float r[16];
};

RT_HD float luma(vec3f c) { return 0.212671f * c.x + 0.715160f * c.y + 0.072169f * c.z; }

RT_HD float tri_area(vec3f a, vec3f b, vec3f c) {
    vec3f x = cross(b - a, c - a);
    return 0.5f * sqrtf(dot(x, x));
}

RT_HD float triangle_light_power(vec3f a, vec3f b, vec3f c, vec3f emissive) {
    vec3f x = cross(b - a, c - a);
    return luma(emissive * 0.5f * sqrtf(dot(x, x)) * kPiF);
}

RT_HD float point_light_power(vec3f col) { return luma(4.0f * kPiF * col); }

// lower_bound(cdf, cdf + n + 1, xi) - 1 (0 for the first entry), clamped to
// a valid light; pdf = f[index] / integral (0 when the total is 0).
RT_HD uint32_t sample_index(const float *cdf, const float *f, uint32_t n, float integral, float xi, float &pdf) {
    uint32_t lo = 0, hi = n + 1;
    while (lo < hi) {
        uint32_t mid = (lo + hi) >> 1;
        if (cdf[mid] < xi) lo = mid + 1;
        else hi = mid;
    }
    uint32_t index = lo > 0 ? lo - 1 : lo;
    if (index >= n) index = n - 1;
    pdf = integral > 0.0f ? f[index] / integral : 0.0f;
    return index;
}

RT_HD vec3f uniform_hemisphere(float xi0, float xi1) {
    float z = xi0;
    float r0 = 1.0f - z * z;
    float r = sqrtf(r0 < 0 ? -r0 : r0);
    float phi = 2.0f * kPiF * xi1;
    return vec3f{r * cosf(phi), r * sinf(phi), z};
}

RT_HD float cdot(vec3f a, vec3f b) {
    float x = a.x * b.x + a.y * b.y + a.z * b.z;
    return x < 0.0f ? 0.0f : x;
}

} // namespace rt
