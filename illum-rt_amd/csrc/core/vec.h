// Minimal fp32 3-vector used by host and device code.
//
// Every operation is written out in the evaluation order that the reference's
// glm-based code produces, and nothing here may be contracted into FMAs: the
// host files are compiled with -ffp-contract=off and the device files with the
// same flag, so that a float computed here on the CPU and on gfx950 is the same
// bit pattern.  Semantics follow glm 0.9.9:
//   dot(a,b)       = (a.x*b.x + a.y*b.y) + a.z*b.z
//   cross(a,b)     = (a.y*b.z - b.y*a.z, a.z*b.x - b.z*a.x, a.x*b.y - b.x*a.y)
//   normalize(v)   = v * (1.0f / sqrtf(dot(v,v)))
//   min(x,y)       = (y < x) ? y : x      max(x,y) = (x < y) ? y : x
#pragma once

#include <cmath>
#include <cstdint>
#include <cfloat>

#if defined(__HIPCC__)
#include <hip/hip_runtime.h>
#define RT_HD __host__ __device__ __forceinline__
#else
#define RT_HD inline
#endif

namespace rt {

struct vec3f {
    float x, y, z;
};

RT_HD vec3f v3(float x, float y, float z) { return vec3f{x, y, z}; }
RT_HD vec3f v3(float s) { return vec3f{s, s, s}; }

RT_HD vec3f operator+(vec3f a, vec3f b) { return vec3f{a.x + b.x, a.y + b.y, a.z + b.z}; }
RT_HD vec3f operator-(vec3f a, vec3f b) { return vec3f{a.x - b.x, a.y - b.y, a.z - b.z}; }
RT_HD vec3f operator-(vec3f a) { return vec3f{-a.x, -a.y, -a.z}; }
RT_HD vec3f operator*(vec3f a, vec3f b) { return vec3f{a.x * b.x, a.y * b.y, a.z * b.z}; }
RT_HD vec3f operator*(vec3f a, float s) { return vec3f{a.x * s, a.y * s, a.z * s}; }
RT_HD vec3f operator*(float s, vec3f a) { return vec3f{s * a.x, s * a.y, s * a.z}; }
RT_HD vec3f operator/(vec3f a, float s) { return vec3f{a.x / s, a.y / s, a.z / s}; }
RT_HD bool operator==(vec3f a, vec3f b) { return a.x == b.x && a.y == b.y && a.z == b.z; }
RT_HD bool operator!=(vec3f a, vec3f b) { return !(a == b); }

RT_HD float dot(vec3f a, vec3f b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
RT_HD vec3f cross(vec3f a, vec3f b) {
    return vec3f{a.y * b.z - b.y * a.z, a.z * b.x - b.z * a.x, a.x * b.y - b.x * a.y};
}
RT_HD vec3f normalize(vec3f v) { return v * (1.0f / sqrtf(dot(v, v))); }
RT_HD float vmin(float x, float y) { return (y < x) ? y : x; }
RT_HD float vmax(float x, float y) { return (x < y) ? y : x; }
RT_HD vec3f vmin(vec3f a, vec3f b) { return vec3f{vmin(a.x, b.x), vmin(a.y, b.y), vmin(a.z, b.z)}; }
RT_HD vec3f vmax(vec3f a, vec3f b) { return vec3f{vmax(a.x, b.x), vmax(a.y, b.y), vmax(a.z, b.z)}; }
RT_HD float comp(vec3f v, int axis) { return axis == 0 ? v.x : (axis == 1 ? v.y : v.z); }

// Axis-aligned box, empty = (FLT_MAX, -FLT_MAX) as in libgi/intersect.h:5-16.
struct aabb {
    vec3f min, max;
};
RT_HD aabb empty_box() { return aabb{v3(FLT_MAX), v3(-FLT_MAX)}; }
// grow(v): min = glm::min(v, min) -> (min < v) ? min : v   (libgi/intersect.h:8-11)
RT_HD void grow(aabb &b, vec3f p) {
    b.min = vmin(p, b.min);
    b.max = vmax(p, b.max);
}
RT_HD void grow(aabb &b, const aabb &o) {
    b.min = vmin(o.min, b.min);
    b.max = vmax(o.max, b.max);
}
// 2*(ex*ey + ex*ez + ey*ez), the surface-area expression of rt/cpu/bvh-ctor.h:387-390.
RT_HD float box_area(const aabb &b) {
    vec3f e = b.max - b.min;
    return 2.0f * (e.x * e.y + e.x * e.z + e.y * e.z);
}

} // namespace rt
