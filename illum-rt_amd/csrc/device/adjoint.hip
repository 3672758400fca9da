// Adjoint of the closest-hit intersection.
//
// For a ray with a (frozen) hit triangle (A, B, C) the forward quantities
// x = (beta, gamma, t) solve  M x = q  with  M = [B-A, C-A, -d],  q = o - A.
// Given upstream gradients g = (dL/dbeta, dL/dgamma, dL/dt), the adjoint
// lambda = M^-T g yields, by differentiating M x = q,
//   dL/do = lambda            dL/dd = t lambda
//   dL/dA = -(1-beta-gamma) lambda,  dL/dB = -beta lambda,  dL/dC = -gamma lambda.
// The intersection is recomputed from the hit triangle's vertices instead of
// being stored by the forward pass; rays without a hit get zero gradients.
// Vertex gradients are accumulated into the scene's shared vertex array with
// float atomics (their summation order is not fixed).
#include "adjoint_kernels.h"

namespace rt {
namespace dev {

namespace {
__device__ __forceinline__ vec3f f3(float4 v) { return vec3f{v.x, v.y, v.z}; }
__device__ __forceinline__ void atomic_add3(float *p, vec3f v) {
    atomicAdd(p + 0, v.x);
    atomicAdd(p + 1, v.y);
    atomicAdd(p + 2, v.z);
}
} // namespace

__global__ __launch_bounds__(256) void k_adjoint_closest(adjoint_args a, uint32_t n) {
    uint32_t i = blockIdx.x * 256u + threadIdx.x;
    if (i >= n) return;
    vec3f go = v3(0.0f), gd = v3(0.0f), gA = v3(0.0f), gB = v3(0.0f), gC = v3(0.0f);
    if (a.hits.t[i] != FLT_MAX) {
        uint32_t tri = a.hits.id[i];
        const float4 *g = a.tri_geo + 3 * (size_t)tri;
        vec3f A = f3(g[0]), B = f3(g[1]), C = f3(g[2]);
        vec3f o{a.rays.ox[i], a.rays.oy[i], a.rays.oz[i]};
        vec3f d{a.rays.dx[i], a.rays.dy[i], a.rays.dz[i]};
        vec3f c0 = B - A, c1 = C - A, c2 = -d, q = o - A;
        vec3f r0 = cross(c1, c2), r1 = cross(c2, c0), r2 = cross(c0, c1);
        float inv = 1.0f / dot(c0, r0);
        float beta = dot(r0, q) * inv, gamma = dot(r1, q) * inv, t = dot(r2, q) * inv;
        float gb = a.grad_u[i], gg = a.grad_v[i], gt = a.grad_t[i];
        vec3f lam = (r0 * gb + r1 * gg + r2 * gt) * inv;
        go = lam;
        gd = t * lam;
        gA = -(1.0f - beta - gamma) * lam;
        gB = -beta * lam;
        gC = -gamma * lam;
        if (a.grad_vertices) {
            const uint32_t *vi = a.tri_vidx + 3 * (size_t)tri;
            atomic_add3(a.grad_vertices + 3 * (size_t)vi[0], gA);
            atomic_add3(a.grad_vertices + 3 * (size_t)vi[1], gB);
            atomic_add3(a.grad_vertices + 3 * (size_t)vi[2], gC);
        }
    }
    if (a.grad_origin) {
        a.grad_origin[3 * (size_t)i] = go.x;
        a.grad_origin[3 * (size_t)i + 1] = go.y;
        a.grad_origin[3 * (size_t)i + 2] = go.z;
    }
    if (a.grad_dir) {
        a.grad_dir[3 * (size_t)i] = gd.x;
        a.grad_dir[3 * (size_t)i + 1] = gd.y;
        a.grad_dir[3 * (size_t)i + 2] = gd.z;
    }
    if (a.grad_corners) {
        float *c = a.grad_corners + 9 * (size_t)i;
        c[0] = gA.x;
        c[1] = gA.y;
        c[2] = gA.z;
        c[3] = gB.x;
        c[4] = gB.y;
        c[5] = gB.z;
        c[6] = gC.x;
        c[7] = gC.y;
        c[8] = gC.z;
    }
}

hipError_t launch_adjoint_closest(const adjoint_args &a, uint32_t n, hipStream_t s) {
    if (n == 0) return hipSuccess;
    hipLaunchKernelGGL(k_adjoint_closest, dim3((n + 255) / 256), dim3(256), 0, s, a, n);
    return hipGetLastError();
}

} // namespace dev
} // namespace rt
