// Device-side ray/box and ray/triangle tests and the per-ray traversal step
// shared by every trace kernel.
//
// Exactness: with -ffp-contract=off and IEEE division (hipcc's default
// -fhip-fp32-correctly-rounded-divide-sqrt), box_exact and tri_exact evaluate
// the same IEEE operations in the same order as the reference's CPU tests
// (intersect4, libgi/intersect.h:230-263; intersect, :19-70), including the
// select-based min/max whose NaN behaviour decides rays with zero direction
// components.  Traversal in the reference's order (rt/cpu/bvh2.cpp:72-137)
// therefore returns bit-identical hits.
#pragma once

#include <hip/hip_runtime.h>

#include "../core/layout.h"
#include "../core/raygen.h"

#include <cfloat>
#include <cstdint>

namespace rt {
namespace dev {

struct dray {
    float ox, oy, oz, tmin, dx, dy, dz, tmax;
    float ix, iy, iz;   // 1/d, computed as ray::ray does (libgi/rt.h:28)
};

__device__ __forceinline__ dray load_ray(const ray_soa &r, uint32_t i) {
    dray q;
    q.ox = r.ox[i];
    q.oy = r.oy[i];
    q.oz = r.oz[i];
    q.tmin = r.tmin[i];
    q.dx = r.dx[i];
    q.dy = r.dy[i];
    q.dz = r.dz[i];
    q.tmax = r.tmax[i];
    q.ix = 1.0f / q.dx;
    q.iy = 1.0f / q.dy;
    q.iz = 1.0f / q.dz;
    return q;
}

// The same ray from a record of `stride` float4 that begins
// (ox, oy, oz, tmin | dx, dy, dz, tmax); with `fixed_t` the record's .w words
// carry other data and the ray spans [kRayEps, FLT_MAX].
__device__ __forceinline__ dray load_ray_aos(const float4 *p, uint32_t i, uint32_t stride, bool fixed_t) {
    float4 a = p[(size_t)stride * i], b = p[(size_t)stride * i + 1];
    dray q;
    q.ox = a.x;
    q.oy = a.y;
    q.oz = a.z;
    q.tmin = fixed_t ? kRayEps : a.w;
    q.dx = b.x;
    q.dy = b.y;
    q.dz = b.z;
    q.tmax = fixed_t ? FLT_MAX : b.w;
    q.ix = 1.0f / q.dx;
    q.iy = 1.0f / q.dy;
    q.iz = 1.0f / q.dz;
    return q;
}

// One slab: the two ternaries of intersect4 (t1 = min, t2 = max as selects).
__device__ __forceinline__ void slab(float lo, float hi, float o, float id, float &a, float &b) {
    float t1 = (lo - o) * id;
    float t2 = (hi - o) * id;
    a = (t1 < t2) ? t1 : t2;
    b = (t2 < t1) ? t1 : t2;
}

// intersect4 plus, for closest hit, the tracer's "dist < closest.t" cull.
template <bool ANY>
__device__ __forceinline__ bool box_exact(float xlo, float xhi, float ylo, float yhi, float zlo, float zhi,
                                          const dray &r, float best, float &dist) {
    float t1x, t2x, t1y, t2y, t1z, t2z;
    slab(xlo, xhi, r.ox, r.ix, t1x, t2x);
    slab(ylo, yhi, r.oy, r.iy, t1y, t2y);
    slab(zlo, zhi, r.oz, r.iz, t1z, t2z);
    float t1 = (t1x < t1y) ? t1y : t1x;
    t1 = (t1z < t1) ? t1 : t1z;
    float t2 = (t2x < t2y) ? t2x : t2y;
    t2 = (t2z < t2) ? t2z : t2;
    dist = t1;
    bool hit = !(t1 > t2) && !(t2 < r.tmin) && !(t1 > r.tmax);
    if (!ANY) hit = hit && (t1 < best);
    return hit;
}

// intersect4 when no slab value can be NaN: the operands are the same
// separately rounded (box - o) * (1/d) products, and for NaN-free inputs the
// select chains of intersect4 equal IEEE min/max (min3/max3 on gfx950),
// differing at most in the sign of a zero, which no comparison sees.  NaN
// slab values need a non-finite ray component or box bound; kernels take
// this path only for wavefronts whose rays and scene are all finite.
template <bool ANY>
__device__ __forceinline__ bool box_fast(float xlo, float xhi, float ylo, float yhi, float zlo, float zhi,
                                         const dray &r, float best, float &dist) {
    float ax = (xlo - r.ox) * r.ix, bx = (xhi - r.ox) * r.ix;
    float ay = (ylo - r.oy) * r.iy, by = (yhi - r.oy) * r.iy;
    float az = (zlo - r.oz) * r.iz, bz = (zhi - r.oz) * r.iz;
    float t1 = fmaxf(fmaxf(fminf(ax, bx), fminf(ay, by)), fminf(az, bz));
    float t2 = fminf(fminf(fmaxf(ax, bx), fmaxf(ay, by)), fmaxf(az, bz));
    dist = t1;
    bool hit = (t1 <= t2) && (t2 >= r.tmin) && (t1 <= r.tmax);
    if (!ANY) hit = hit && (t1 < best);
    return hit;
}

// True when every component of the ray and of its inverse direction is
// finite, i.e. box_fast decides exactly like box_exact for this ray.
__device__ __forceinline__ bool ray_is_finite(const dray &r) {
    return isfinite(r.ox) && isfinite(r.oy) && isfinite(r.oz) && isfinite(r.dx) && isfinite(r.dy) &&
           isfinite(r.dz) && isfinite(r.ix) && isfinite(r.iy) && isfinite(r.iz) && !isnan(r.tmin) &&
           !isnan(r.tmax);
}

template <bool ANY>
__device__ __forceinline__ bool box_test(bool exact, float xlo, float xhi, float ylo, float yhi, float zlo,
                                         float zhi, const dray &r, float best, float &dist) {
    if (exact) return box_exact<ANY>(xlo, xhi, ylo, yhi, zlo, zhi, r, best, dist);
    return box_fast<ANY>(xlo, xhi, ylo, yhi, zlo, zhi, r, best, dist);
}

// Shirley/Cramer test on an exact record (A, A-B, A-C).  Accepts iff
// tmin < t < tmax, beta > 0, gamma > 0, beta + gamma <= 1, and t < best.
__device__ __forceinline__ bool tri_exact(const float4 &r0, const float4 &r1, const float4 &r2, const dray &r,
                                          float best, float &t_out, float &u_out, float &v_out) {
    const float a = r1.x, b = r1.y, c = r1.z;
    const float d = r2.x, e = r2.y, f = r2.z;
    const float g = r.dx, h = r.dy, i = r.dz;
    const float j = r0.x - r.ox, k = r0.y - r.oy, l = r0.z - r.oz;
    float c1 = e * i - h * f;
    float c2 = g * f - d * i;
    float c3 = d * h - e * g;
    const float M = a * c1 + b * c2 + c * c3;
    float beta = j * c1 + k * c2 + l * c3;
    c1 = a * k - j * b;
    c2 = j * c - a * l;
    c3 = b * l - k * c;
    float gamma = i * c1 + h * c2 + g * c3;
    float tt = -(f * c1 + e * c2 + d * c3);
    // Sign pre-filter: a quotient whose numerator and M have opposite signs
    // (or a zero numerator) is <= 0, so "beta > 0" / "gamma > 0" fail -- and
    // with t_min >= 0 so does "t > t_min" -- whatever the division would
    // round to.  Only the three divisions are skipped; accepted hits are the
    // quotients of the reference's expressions.
    bool neg_m = M < 0;
    if ((beta > 0) == neg_m || (gamma > 0) == neg_m || ((tt > 0) == neg_m && r.tmin >= 0)) {
        if (M == M && M != 0) return false;   // (a NaN or zero M falls through to the IEEE divisions)
    }
    tt /= M;
    if (!(tt > r.tmin && tt < r.tmax && tt < best)) return false;
    beta /= M;
    gamma /= M;
    if (!(beta > 0 && gamma > 0 && beta + gamma <= 1)) return false;
    t_out = tt;
    u_out = beta;
    v_out = gamma;
    return true;
}

// Woop test on a local-origin record (core/records.h): the ray origin
// relative to the corner A (rows' .w), t from the plane row, then the two
// barycentric rows; the same acceptance rule as the exact test.  FMAs and
// the hardware reciprocal (1 ulp) are explicit: this format is not
// bit-compatible with the CPU test in any case.
__device__ __forceinline__ bool tri_woop(const float4 &m0, const float4 &m1, const float4 &m2, const dray &r,
                                         float best, float &t_out, float &u_out, float &v_out) {
    const float px = r.ox - m0.w, py = r.oy - m1.w, pz = r.oz - m2.w;
    float oz = fmaf(m2.x, px, fmaf(m2.y, py, m2.z * pz));
    float dz = fmaf(m2.x, r.dx, fmaf(m2.y, r.dy, m2.z * r.dz));
    float tt = -oz * __builtin_amdgcn_rcpf(dz);
    if (!(tt > r.tmin && tt < r.tmax && tt < best)) return false;
    float ou = fmaf(m0.x, px, fmaf(m0.y, py, m0.z * pz));
    float du = fmaf(m0.x, r.dx, fmaf(m0.y, r.dy, m0.z * r.dz));
    float u = fmaf(tt, du, ou);
    if (!(u > 0)) return false;
    float ov = fmaf(m1.x, px, fmaf(m1.y, py, m1.z * pz));
    float dv = fmaf(m1.x, r.dx, fmaf(m1.y, r.dy, m1.z * r.dz));
    float v = fmaf(tt, dv, ov);
    if (!(v > 0 && u + v <= 1)) return false;
    t_out = tt;
    u_out = u;
    v_out = v;
    return true;
}

__device__ __forceinline__ uint32_t as_u32(float f) { return __float_as_uint(f); }

// Texel of texture2d::sample (libgi/scene.h:28-36): wrap by subtracting the
// floor, nearest texel by (int)(u * w + 0.5), index w (a coordinate that
// rounds up to 1) wraps to 0.  Non-finite coordinates, for which the
// reference indexes outside the texture, take texel 0.
__host__ __device__ __forceinline__ uint32_t texel_index(float u, float v, uint32_t w, uint32_t h) {
    u = u - floorf(u);
    v = v - floorf(v);
    int x = (int)(u * (float)w + 0.5f);
    int y = (int)(v * (float)h + 0.5f);
    if (x == (int)w) x = 0;
    if (y == (int)h) y = 0;
    if (!(x >= 0 && x < (int)w)) x = 0;
    if (!(y >= 0 && y < (int)h)) y = 0;
    return (uint32_t)y * w + (uint32_t)x;
}

// Opacity test of the reference's alpha-aware tracers (rt/cpu/bvh2.cpp:
// 109-119 and 171-176): a hit counts when diff_geom::opacity() > 0.5
// (libgi/rt.cpp:32-35) -- 1 without an albedo texture, else the alpha of the
// texel at tc = (1 - beta - gamma) tc_a + beta tc_b + gamma tc_c
// (libgi/rt.cpp:11), products rounded separately and summed left to right.
// The mask byte holds (alpha > 0.5) for every texel.
__device__ __forceinline__ bool alpha_opaque(const scene_desc &sc, uint32_t tri, float beta, float gamma) {
    const float4 *at = static_cast<const float4 *>(sc.alpha_tri) + 2ull * tri;
    float4 p = at[0], q = at[1];
    uint32_t tex = __float_as_uint(q.z);
    if (tex == 0xffffffffu) return true;
    uint4 t = static_cast<const uint4 *>(sc.alpha_tex)[tex];
    float w0 = 1.0f - beta - gamma;
    float u = w0 * p.x + beta * p.z + gamma * q.x;
    float v = w0 * p.y + beta * p.w + gamma * q.y;
    return sc.alpha_mask[t.x + texel_index(u, v, t.y, t.z)] != 0;
}

template <int TRI>
__device__ __forceinline__ bool test_slot(const scene_desc &sc, uint32_t slot, const dray &r, float best, float &t,
                                          float &u, float &v) {
    const float4 *p = reinterpret_cast<const float4 *>(sc.tris + slot);
    float4 r0 = p[0], r1 = p[1], r2 = p[2];
    if (TRI == TRI_EXACT) return tri_exact(r0, r1, r2, r, best, t, u, v);
    return tri_woop(r0, r1, r2, r, best, t, u, v);
}

template <int TRI>
__device__ __forceinline__ uint32_t slot_tri_id(const scene_desc &sc, uint32_t slot) {
    if (TRI == TRI_EXACT) return as_u32(reinterpret_cast<const float4 *>(sc.tris + slot)[0].w);
    return sc.slot_ids[slot];
}

// Result of one ray.
struct dhit {
    float t, u, v;
    uint32_t slot;
    bool found;
    bool over;   // the traversal dropped a stack entry: the result is not known
};

__device__ __forceinline__ void store_hit(const hit_soa &h, uint32_t i, float t, float u, float v, uint32_t id) {
    h.t[i] = t;
    h.u[i] = u;
    h.v[i] = v;
    h.id[i] = id;
}

// The result of a ray whose traversal lost a stack entry (the host sizes
// every stack for the tree, so this marks a sizing error, never a hit):
// t = NaN and id 0xFFFFFFFF, occlusion byte 0xFF.  An any-hit ray that found
// an occluder is unaffected (a lost entry can only hide occluders).
__device__ __forceinline__ void store_lost(const hit_soa &h, uint8_t *occluded, uint32_t i) {
    if (occluded) occluded[i] = 0xFF;
    if (h.t) store_hit(h, i, __uint_as_float(0x7FC00000u), 0.0f, 0.0f, 0xffffffffu);
}

} // namespace dev
} // namespace rt
