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
//
// Vertex gradients are sums over every ray that hit a triangle of the
// vertex.  One global float atomic per ray and component (9 per ray) is what
// limits a direct accumulation: device-scope float atomics on MI355X retire
// at about 10 G/s, and same-address ones serialize (about 11 ns apart), so
// 1M diffuse rays of the hall spent 1.26 ms there against 17 us for
// everything else.  Rays are therefore grouped by hit triangle first:
//   k_adjoint_closest   per ray: origin, direction and corner gradients (the
//                       caller's array or a 36-byte scratch record per ray),
//                       and a sort key (the hit triangle; n_tris for a miss);
//   radix sort          stable, of the keys with their ray index
//                       (device/ray_sort.hip);
//   k_adjoint_vertices  per sorted ray: its corner-gradient record, summed
//                       over runs of equal triangles inside each wavefront
//                       (a segmented scan over lane shuffles), the last lane
//                       of a run adding its 9 sums to the 3 vertices.
// About 9 atomics per distinct hit triangle and wavefront remain; their
// order, and so the rounding of the sums, is not fixed.
#include "adjoint_kernels.h"

#include "ray_sort.h"

namespace rt {
namespace dev {

namespace {
__device__ __forceinline__ vec3f f3(float4 v) { return vec3f{v.x, v.y, v.z}; }
__device__ __forceinline__ void atomic_add3(float *p, vec3f v) {
    atomicAdd(p + 0, v.x);
    atomicAdd(p + 1, v.y);
    atomicAdd(p + 2, v.z);
}

struct ray_adjoint {
    vec3f go, gd, gA, gB, gC;
};

// The adjoint of ray i with hit triangle tri (valid hit).
__device__ __forceinline__ ray_adjoint adjoint_of(const adjoint_args &a, uint32_t i, uint32_t tri) {
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
    ray_adjoint r;
    r.go = lam;
    r.gd = t * lam;
    r.gA = -(1.0f - beta - gamma) * lam;
    r.gB = -beta * lam;
    r.gC = -gamma * lam;
    return r;
}
} // namespace

__global__ __launch_bounds__(256) void k_adjoint_closest(adjoint_args a, uint32_t n) {
    uint32_t i = blockIdx.x * 256u + threadIdx.x;
    if (i >= n) return;
    ray_adjoint r{v3(0.0f), v3(0.0f), v3(0.0f), v3(0.0f), v3(0.0f)};
    const uint32_t tri = a.hits.id[i];
    const bool valid = a.hits.t[i] < FLT_MAX && tri < a.n_tris;   // a hit on a triangle of this scene
    if (valid) r = adjoint_of(a, i, tri);
    if (a.keys) {
        a.keys[i] = valid ? tri : a.n_tris;
        if (i == 0) *a.count = n;
    }
    float *corners = a.grad_corners ? a.grad_corners : a.corner_scratch;
    if (a.grad_origin) {
        a.grad_origin[3 * (size_t)i] = r.go.x;
        a.grad_origin[3 * (size_t)i + 1] = r.go.y;
        a.grad_origin[3 * (size_t)i + 2] = r.go.z;
    }
    if (a.grad_dir) {
        a.grad_dir[3 * (size_t)i] = r.gd.x;
        a.grad_dir[3 * (size_t)i + 1] = r.gd.y;
        a.grad_dir[3 * (size_t)i + 2] = r.gd.z;
    }
    if (corners) {
        float *c = corners + 9 * (size_t)i;
        c[0] = r.gA.x;
        c[1] = r.gA.y;
        c[2] = r.gA.z;
        c[3] = r.gB.x;
        c[4] = r.gB.y;
        c[5] = r.gB.z;
        c[6] = r.gC.x;
        c[7] = r.gC.y;
        c[8] = r.gC.z;
    }
}

// Sorted position j holds key keys[j] of ray perm[j]; rays of one triangle
// are contiguous, misses (key n_tris) last.  corners: 9 floats per ray.
__global__ __launch_bounds__(256) void k_adjoint_vertices(adjoint_args a, const uint32_t *keys, const uint32_t *perm,
                                                         const float *corners, uint32_t n) {
    const uint32_t j = blockIdx.x * 256u + threadIdx.x;
    const uint32_t lane = threadIdx.x & 63u;
    uint32_t key = 0xFFFFFFFFu;   // no triangle
    float v[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
    if (j < n && keys[j] < a.n_tris) {
        key = keys[j];
        const float *c = corners + 9 * (size_t)perm[j];
#pragma unroll
        for (int k = 0; k < 9; ++k) v[k] = c[k];
    }
    // inclusive segmented scan: keys are sorted, so an equal key `off` lanes
    // down means the whole stretch between belongs to the same triangle
#pragma unroll
    for (int off = 1; off < 64; off <<= 1) {
        uint32_t k_up = __shfl_up(key, off, 64);
        bool same = lane >= (uint32_t)off && k_up == key;
#pragma unroll
        for (int c = 0; c < 9; ++c) {
            float up = __shfl_up(v[c], off, 64);
            if (same) v[c] += up;
        }
    }
    uint32_t k_next = __shfl_down(key, 1, 64);
    bool tail = lane == 63u || k_next != key;
    if (tail && key != 0xFFFFFFFFu) {
        const uint32_t *vi = a.tri_vidx + 3 * (size_t)key;
        atomic_add3(a.grad_vertices + 3 * (size_t)vi[0], vec3f{v[0], v[1], v[2]});
        atomic_add3(a.grad_vertices + 3 * (size_t)vi[1], vec3f{v[3], v[4], v[5]});
        atomic_add3(a.grad_vertices + 3 * (size_t)vi[2], vec3f{v[6], v[7], v[8]});
    }
}

size_t adjoint_sort_bytes(uint32_t n) {
    return 4ull * (13ull * n + 256ull * radix_tiles(n) + 256ull + 64ull);
}

hipError_t launch_adjoint_closest(const adjoint_args &a_in, uint32_t n, void *sort_scratch, hipStream_t s) {
    if (n == 0) return hipSuccess;
    adjoint_args a = a_in;
    const bool vertices = a.grad_vertices != nullptr;
    radix_sort_buffers rb{};
    if (vertices) {
        if (!sort_scratch) return hipErrorInvalidValue;
        uint32_t *w = static_cast<uint32_t *>(sort_scratch);
        const uint32_t tiles = radix_tiles(n);
        rb.keys_in = w;
        rb.keys_tmp = w + n;
        rb.vals_tmp = w + 2ull * n;
        rb.vals_out = w + 3ull * n;
        a.corner_scratch = reinterpret_cast<float *>(w + 4ull * n);
        rb.counts = w + 13ull * n;
        rb.totals = rb.counts + 256ull * tiles;
        a.keys = rb.keys_in;
        a.count = rb.totals + 256;
    }
    hipLaunchKernelGGL(k_adjoint_closest, dim3((n + 255) / 256), dim3(256), 0, s, a, n);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess || !vertices) return e;
    int bits = 1;
    while (bits < 32 && (a.n_tris >> bits) != 0u) ++bits;   // keys 0 .. n_tris
    e = launch_radix_sort(rb, n, a.count, bits, s, true);   // the vertex pass reads the sorted keys
    if (e != hipSuccess) return e;
    // the last of the sort's passes leaves the keys in keys_tmp after an odd
    // number of passes, in keys_in after an even one
    const uint32_t *sorted = ((bits + 7) / 8) & 1 ? rb.keys_tmp : rb.keys_in;
    const float *corners = a.grad_corners ? a.grad_corners : a.corner_scratch;
    hipLaunchKernelGGL(k_adjoint_vertices, dim3((n + 255) / 256), dim3(256), 0, s, a, sorted, rb.vals_out, corners,
                       n);
    return hipGetLastError();
}

} // namespace dev
} // namespace rt
