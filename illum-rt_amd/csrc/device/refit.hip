// Refit: new vertex positions, unchanged triangles and topology.
//
// Records are rebuilt per leaf slot with the host flattener's own functions
// (core/records.h), so a refit to the committed positions reproduces the
// committed records bit for bit.  Boxes are recomputed bottom-up, one launch
// per tree level (deepest first): a leaf child's box is the bounding box of
// its triangles' vertices, an inner child's box the union of that child's own
// child boxes.  Leaf boxes therefore bound whole triangles even where the
// builder's early split clipping had bounded fragments (a fragment box lies
// inside the triangle's box, so every hit stays reachable).
//
// Min/max are the select forms of glm::min/max (aabb::grow,
// libgi/intersect.h:9-15) starting from an empty box (+inf, -inf); a NaN
// coordinate never enters a box that way, so refit boxes are NaN-free.
#include "refit_kernels.h"

#include "../core/records.h"

namespace rt {
namespace dev {

namespace {

struct dbox {
    float lo[3], hi[3];
};

__device__ __forceinline__ dbox empty_box() {
    return dbox{{INFINITY, INFINITY, INFINITY}, {-INFINITY, -INFINITY, -INFINITY}};
}

__device__ __forceinline__ void grow(dbox &b, float x, float y, float z) {
    b.lo[0] = (x < b.lo[0]) ? x : b.lo[0];
    b.lo[1] = (y < b.lo[1]) ? y : b.lo[1];
    b.lo[2] = (z < b.lo[2]) ? z : b.lo[2];
    b.hi[0] = (b.hi[0] < x) ? x : b.hi[0];
    b.hi[1] = (b.hi[1] < y) ? y : b.hi[1];
    b.hi[2] = (b.hi[2] < z) ? z : b.hi[2];
}

__device__ __forceinline__ void merge(dbox &b, const dbox &c) {
    for (int k = 0; k < 3; ++k) {
        b.lo[k] = (c.lo[k] < b.lo[k]) ? c.lo[k] : b.lo[k];
        b.hi[k] = (b.hi[k] < c.hi[k]) ? c.hi[k] : b.hi[k];
    }
}

__device__ __forceinline__ vec3f vertex(const float *pos, uint32_t v) {
    return vec3f{pos[3ull * v], pos[3ull * v + 1], pos[3ull * v + 2]};
}

__device__ __forceinline__ uint32_t slot_triangle(const refit_args &a, uint32_t slot) {
    return a.slot_ids[slot];
}

__device__ dbox leaf_box(const refit_args &a, int32_t ref) {
    dbox b = empty_box();
    uint32_t slot = leaf_slot(ref), cnt = leaf_count(ref);
    for (uint32_t k = 0; k < cnt; ++k) {
        uint32_t t = slot_triangle(a, slot + k);
        for (int c = 0; c < 3; ++c) {
            vec3f p = vertex(a.pos, a.vidx[3ull * t + c]);
            grow(b, p.x, p.y, p.z);
        }
    }
    return b;
}

// Union of the child boxes stored in inner node `i`.
__device__ dbox inner_box(const refit_args &a, int32_t i) {
    dbox b = empty_box();
    if (a.bvh == BVH2) {
        const node2 &n = static_cast<const node2 *>(a.nodes)[i];
        merge(b, dbox{{n.lxlo, n.lylo, n.lzlo}, {n.lxhi, n.lyhi, n.lzhi}});
        merge(b, dbox{{n.rxlo, n.rylo, n.rzlo}, {n.rxhi, n.ryhi, n.rzhi}});
    } else {
        const node4 &n = static_cast<const node4 *>(a.nodes)[i];
        for (int c = 0; c < 4; ++c)
            if (n.ref[c] != kInvalidRef) merge(b, dbox{{n.lox[c], n.loy[c], n.loz[c]}, {n.hix[c], n.hiy[c], n.hiz[c]}});
    }
    return b;
}

__device__ __forceinline__ dbox child_box(const refit_args &a, int32_t ref) {
    if (ref == kInvalidRef) return empty_box();
    return ref_is_leaf(ref) ? leaf_box(a, ref) : inner_box(a, ref);
}

__global__ __launch_bounds__(256) void k_refit_records(refit_args a) {
    uint32_t i = blockIdx.x * 256u + threadIdx.x;
    if (i < a.n_slots) {
        uint32_t t = a.slot_ids[i];
        vec3f A = vertex(a.pos, a.vidx[3ull * t]);
        vec3f B = vertex(a.pos, a.vidx[3ull * t + 1]);
        vec3f C = vertex(a.pos, a.vidx[3ull * t + 2]);
        a.tris[i] = a.tri == TRI_WOOP ? make_woop_rec(A, B, C) : make_exact_rec(A, B, C, t);
    }
    if (a.tri_geo && i < a.n_tris) {
        float *g = a.tri_geo + 12ull * i;
        for (int c = 0; c < 3; ++c) {
            vec3f p = vertex(a.pos, a.vidx[3ull * i + c]);
            g[4 * c] = p.x;
            g[4 * c + 1] = p.y;
            g[4 * c + 2] = p.z;
        }
    }
}

__global__ __launch_bounds__(256) void k_refit_level(refit_args a, const int32_t *list, uint32_t n) {
    uint32_t i = blockIdx.x * 256u + threadIdx.x;
    if (i >= n) return;
    int32_t id = list[i];
    if (a.bvh == BVH2) {
        node2 &nd = static_cast<node2 *>(a.nodes)[id];
        dbox l = child_box(a, nd.ref_l), r = child_box(a, nd.ref_r);
        nd.lxlo = l.lo[0];
        nd.lxhi = l.hi[0];
        nd.lylo = l.lo[1];
        nd.lyhi = l.hi[1];
        nd.lzlo = l.lo[2];
        nd.lzhi = l.hi[2];
        nd.rxlo = r.lo[0];
        nd.rxhi = r.hi[0];
        nd.rylo = r.lo[1];
        nd.ryhi = r.hi[1];
        nd.rzlo = r.lo[2];
        nd.rzhi = r.hi[2];
    } else {
        node4 &nd = static_cast<node4 *>(a.nodes)[id];
        for (int c = 0; c < 4; ++c) {
            dbox b = child_box(a, nd.ref[c]);
            nd.lox[c] = b.lo[0];
            nd.hix[c] = b.hi[0];
            nd.loy[c] = b.lo[1];
            nd.hiy[c] = b.hi[1];
            nd.loz[c] = b.lo[2];
            nd.hiz[c] = b.hi[2];
        }
    }
}

} // namespace

hipError_t launch_refit_records(const refit_args &a, hipStream_t stream) {
    uint32_t n = a.n_slots > a.n_tris ? a.n_slots : a.n_tris;
    if (n == 0) return hipSuccess;
    hipLaunchKernelGGL(k_refit_records, dim3((n + 255) / 256), dim3(256), 0, stream, a);
    return hipGetLastError();
}

hipError_t launch_refit_level(const refit_args &a, const int32_t *list, uint32_t n, hipStream_t stream) {
    if (n == 0) return hipSuccess;
    hipLaunchKernelGGL(k_refit_level, dim3((n + 255) / 256), dim3(256), 0, stream, a, list, n);
    return hipGetLastError();
}

} // namespace dev
} // namespace rt
