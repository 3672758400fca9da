// Construction of the 48-byte triangle records of core/layout.h, shared by
// the host flattener and the device refit kernel so that both produce the same
// bits for the same vertices (host and device objects are compiled with
// -ffp-contract=off; FP64 division is correctly rounded on both sides).
#pragma once

#include "layout.h"
#include "vec.h"

#include <cmath>

namespace rt {

struct dvec {
    double x, y, z;
};
RT_HD dvec dv(vec3f v) { return dvec{v.x, v.y, v.z}; }
RT_HD dvec dsub(dvec a, dvec b) { return dvec{a.x - b.x, a.y - b.y, a.z - b.z}; }
RT_HD dvec dcross(dvec a, dvec b) {
    return dvec{a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x};
}
RT_HD double ddot(dvec a, dvec b) { return a.x * b.x + a.y * b.y + a.z * b.z; }

RT_HD float u32_as_float(uint32_t u) { return __builtin_bit_cast(float, u); }
RT_HD uint32_t float_as_u32(float f) { return __builtin_bit_cast(uint32_t, f); }

// Exact record: A and the two differences the Shirley/Cramer test of
// libgi/intersect.h:19-70 forms (a..c = A-B, d..f = A-C), triangle id in
// row0.w.
RT_HD tri_rec make_exact_rec(vec3f A, vec3f B, vec3f C, uint32_t tri_id) {
    tri_rec r{};
    r.r0[0] = A.x;
    r.r0[1] = A.y;
    r.r0[2] = A.z;
    r.r0[3] = u32_as_float(tri_id);
    r.r1[0] = A.x - B.x;
    r.r1[1] = A.y - B.y;
    r.r1[2] = A.z - B.z;
    r.r2[0] = A.x - C.x;
    r.r2[1] = A.y - C.y;
    r.r2[2] = A.z - C.z;
    return r;
}

// Woop record: rows of inverse([B-A, C-A, N]) with N = (B-A)x(C-A), in FP64,
// rounded to fp32, with the triangle's corner A as the local origin: row k
// is (q_k.xyz, A[k]).  For a point p, (q0.(p-A), q1.(p-A), q2.(p-A)) =
// (beta, gamma, s): beta and gamma are the barycentric weights of B and C
// (the reference's beta/gamma) and s the signed offset along N.  The test
// forms p - A from the ray origin first, as the Shirley test forms A - o, so
// no row carries an offset term q.A whose rounding error grows with |A|
// relative to the triangle.  Degenerate triangles get q = 0: every t is then
// -0/0 = NaN, i.e. no hit.
RT_HD tri_rec make_woop_rec(vec3f A, vec3f B, vec3f C) {
    dvec a = dv(A), e1 = dsub(dv(B), a), e2 = dsub(dv(C), a);
    dvec n = dcross(e1, e2);
    double det = ddot(n, n);   // det([e1 e2 n]) = |e1 x e2|^2
    tri_rec r{};
    float *out[3] = {r.r0, r.r1, r.r2};
    const float corner[3] = {A.x, A.y, A.z};
    for (int k = 0; k < 3; ++k) out[k][3] = corner[k];
    if (!(det > 0.0 && det < HUGE_VAL)) return r;
    // inverse of a matrix with columns c0,c1,c2: rows (c1xc2, c2xc0, c0xc1)/det
    dvec rows[3] = {dcross(e2, n), dcross(n, e1), dcross(e1, e2)};
    for (int k = 0; k < 3; ++k) {
        out[k][0] = (float)(rows[k].x / det);
        out[k][1] = (float)(rows[k].y / det);
        out[k][2] = (float)(rows[k].z / det);
    }
    return r;
}

} // namespace rt
