// Linear BVH built on the device (Karras, "Maximizing parallelism in the
// construction of BVHs, octrees, and k-d trees", HPG 2012), written straight
// into the flattened layout the trace kernels read.
//
//   1. per triangle: bounding box, and the centre of the box (the reduction of
//      the centres' bounds runs in two passes, one partial per block);
//   2. 63-bit Morton code of the centre, 21 bits per axis over those bounds;
//   3. radix sort of (code, triangle) pairs (rocPRIM);
//   4. the binary radix tree over the sorted codes, one thread per inner node;
//      equal codes are told apart by their sorted positions;
//   5. boxes bottom-up, one thread per leaf climbing while it is the second
//      child to arrive (acquire/release counters), which also counts every
//      subtree's emitted inner nodes; a subtree of at most `max_leaf`
//      triangles becomes one leaf of its parent (its triangles are
//      contiguous in Morton order, so it is one slot range);
//   6. pre-order numbering of the emitted nodes (each node climbs to the root
//      adding 1 per edge and the left sibling's size on right turns), so
//      children always follow their parent, as in the host builder's trees;
//   7. nodes and triangle records written in that order.
//
// Leaf boxes bound whole triangles.  Min/max are the select forms of
// glm::min/max starting from empty boxes, so a NaN coordinate never enters a
// box.  This tree is not the reference's SAH tree: closest hits agree with
// brute force up to equal-t ties and grazing hits, not bit for bit.
#include "lbvh_kernels.h"

#include "../core/records.h"

#include <cstring>

#include <rocprim/rocprim.hpp>

namespace rt {
namespace dev {

namespace {

constexpr int kBlock = 256;
constexpr uint32_t kMortonBits = 21;
constexpr int kMaxClimb = 4096;   // bound on any root-to-leaf walk (trees are far shallower)

struct fbox {
    float lo[3], hi[3];
};

__device__ __forceinline__ fbox empty_fbox() {
    return fbox{{INFINITY, INFINITY, INFINITY}, {-INFINITY, -INFINITY, -INFINITY}};
}
__device__ __forceinline__ void grow(fbox &b, float x, float y, float z) {
    b.lo[0] = (x < b.lo[0]) ? x : b.lo[0];
    b.lo[1] = (y < b.lo[1]) ? y : b.lo[1];
    b.lo[2] = (z < b.lo[2]) ? z : b.lo[2];
    b.hi[0] = (b.hi[0] < x) ? x : b.hi[0];
    b.hi[1] = (b.hi[1] < y) ? y : b.hi[1];
    b.hi[2] = (b.hi[2] < z) ? z : b.hi[2];
}
__device__ __forceinline__ void merge(fbox &b, const fbox &c) {
    for (int k = 0; k < 3; ++k) {
        b.lo[k] = (c.lo[k] < b.lo[k]) ? c.lo[k] : b.lo[k];
        b.hi[k] = (b.hi[k] < c.hi[k]) ? c.hi[k] : b.hi[k];
    }
}

// Scratch layout (offsets in bytes, 256-aligned), for n triangles, m = n - 1.
struct scratch_map {
    size_t tri_box, partial, bounds, keys0, keys1, vals0, vals1, left, right, par_inner, par_leaf, first, last,
        node_box, size, visits, pre, sort_tmp, sort_bytes, total;
};

size_t align256(size_t x) { return (x + 255) & ~size_t(255); }

scratch_map map_scratch(uint32_t n, size_t sort_bytes) {
    scratch_map s{};
    size_t m = n > 1 ? n - 1 : 1, nb = (n + kBlock - 1) / kBlock + 1, off = 0;
    auto take = [&](size_t bytes) {
        size_t at = off;
        off = align256(off + bytes);
        return at;
    };
    s.tri_box = take(sizeof(fbox) * (size_t)n);
    s.partial = take(sizeof(fbox) * nb);
    s.bounds = take(sizeof(fbox));
    s.keys0 = take(8ull * n);
    s.keys1 = take(8ull * n);
    s.vals0 = take(4ull * n);
    s.vals1 = take(4ull * n);
    s.left = take(4 * m);
    s.right = take(4 * m);
    s.par_inner = take(4 * m);
    s.par_leaf = take(4ull * n);
    s.first = take(4 * m);
    s.last = take(4 * m);
    s.node_box = take(sizeof(fbox) * m);
    s.size = take(4 * m);
    s.visits = take(4 * m);
    s.pre = take(4 * m);
    s.sort_bytes = sort_bytes;
    s.sort_tmp = take(sort_bytes);
    s.total = off;
    return s;
}

size_t sort_storage(uint32_t n) {
    size_t bytes = 0;
    (void)rocprim::radix_sort_pairs(nullptr, bytes, (uint64_t *)nullptr, (uint64_t *)nullptr, (uint32_t *)nullptr,
                                    (uint32_t *)nullptr, n, 0, 3 * kMortonBits);
    return bytes;
}

template <class T> T *at(void *base, size_t off) {
    return reinterpret_cast<T *>(static_cast<char *>(base) + off);
}

// ---- 1. triangle boxes, index check, partial bounds of the box centres ----
__global__ __launch_bounds__(kBlock) void k_tri_boxes(lbvh_in in, fbox *tri_box, fbox *partial, uint32_t *counts) {
    __shared__ float sh[6][kBlock];
    uint32_t i = blockIdx.x * kBlock + threadIdx.x;
    fbox c = empty_fbox();
    if (i < in.n_tris) {
        fbox b = empty_fbox();
        bool bad = false;
        for (int k = 0; k < 3; ++k) {
            uint32_t v = in.vidx[3ull * i + k];
            if (v >= in.n_vert) {
                bad = true;
                v = 0;
            }
            grow(b, in.pos[3ull * v], in.pos[3ull * v + 1], in.pos[3ull * v + 2]);
        }
        if (bad) atomicAdd(counts + 2, 1u);
        tri_box[i] = b;
        grow(c, 0.5f * (b.lo[0] + b.hi[0]), 0.5f * (b.lo[1] + b.hi[1]), 0.5f * (b.lo[2] + b.hi[2]));
    }
    for (int k = 0; k < 3; ++k) {
        sh[k][threadIdx.x] = c.lo[k];
        sh[3 + k][threadIdx.x] = c.hi[k];
    }
    __syncthreads();
    for (int w = kBlock / 2; w > 0; w >>= 1) {
        if (threadIdx.x < (unsigned)w)
            for (int k = 0; k < 3; ++k) {
                float a = sh[k][threadIdx.x + w], b = sh[3 + k][threadIdx.x + w];
                sh[k][threadIdx.x] = (a < sh[k][threadIdx.x]) ? a : sh[k][threadIdx.x];
                sh[3 + k][threadIdx.x] = (sh[3 + k][threadIdx.x] < b) ? b : sh[3 + k][threadIdx.x];
            }
        __syncthreads();
    }
    if (threadIdx.x == 0)
        partial[blockIdx.x] = fbox{{sh[0][0], sh[1][0], sh[2][0]}, {sh[3][0], sh[4][0], sh[5][0]}};
}

__global__ __launch_bounds__(kBlock) void k_reduce_bounds(const fbox *partial, uint32_t n_part, fbox *bounds) {
    __shared__ fbox sh[kBlock];
    fbox c = empty_fbox();
    for (uint32_t i = threadIdx.x; i < n_part; i += kBlock) merge(c, partial[i]);
    sh[threadIdx.x] = c;
    __syncthreads();
    for (int w = kBlock / 2; w > 0; w >>= 1) {
        if (threadIdx.x < (unsigned)w) merge(sh[threadIdx.x], sh[threadIdx.x + w]);
        __syncthreads();
    }
    if (threadIdx.x == 0) *bounds = sh[0];
}

// ---- 2. Morton codes ----
__device__ __forceinline__ uint64_t spread21(uint32_t x) {
    uint64_t v = x & 0x1fffffu;
    v = (v | (v << 32)) & 0x1f00000000ffffull;
    v = (v | (v << 16)) & 0x1f0000ff0000ffull;
    v = (v | (v << 8)) & 0x100f00f00f00f00full;
    v = (v | (v << 4)) & 0x10c30c30c30c30c3ull;
    v = (v | (v << 2)) & 0x1249249249249249ull;
    return v;
}

__global__ __launch_bounds__(kBlock) void k_morton(const fbox *tri_box, const fbox *bounds, uint32_t n,
                                                   uint64_t *keys, uint32_t *vals) {
    uint32_t i = blockIdx.x * kBlock + threadIdx.x;
    if (i >= n) return;
    const fbox b = tri_box[i];
    const fbox bb = *bounds;
    uint32_t q[3];
    for (int k = 0; k < 3; ++k) {
        float ext = bb.hi[k] - bb.lo[k];
        float scale = ext > 0.0f ? (float)((1u << kMortonBits) - 1) / ext : 0.0f;
        float x = (0.5f * (b.lo[k] + b.hi[k]) - bb.lo[k]) * scale;
        x = fminf(fmaxf(x, 0.0f), (float)((1u << kMortonBits) - 1));   // NaN -> 0
        q[k] = (uint32_t)x;
    }
    keys[i] = (spread21(q[0]) << 2) | (spread21(q[1]) << 1) | spread21(q[2]);
    vals[i] = i;
}

// ---- 4. binary radix tree ----
// Common-prefix length of sorted codes a and b; -1 outside [0, n).
__device__ __forceinline__ int delta(const uint64_t *keys, int n, int a, int b) {
    if (b < 0 || b >= n) return -1;
    uint64_t ka = keys[a], kb = keys[b];
    if (ka == kb) return 64 + __clz((uint32_t)a ^ (uint32_t)b);
    return __clzll(ka ^ kb);
}

// Child encoding: (index << 1) | is_leaf.
__global__ __launch_bounds__(kBlock) void k_karras(const uint64_t *keys, int n, int32_t *left, int32_t *right,
                                                   int32_t *par_inner, int32_t *par_leaf, int32_t *first,
                                                   int32_t *last) {
    int i = (int)(blockIdx.x * kBlock + threadIdx.x);
    if (i >= n - 1) return;
    int d = (delta(keys, n, i, i + 1) - delta(keys, n, i, i - 1)) >= 0 ? 1 : -1;
    int dmin = delta(keys, n, i, i - d);
    int lmax = 2;
    while (delta(keys, n, i, i + lmax * d) > dmin) lmax *= 2;
    int l = 0;
    for (int t = lmax / 2; t >= 1; t /= 2)
        if (delta(keys, n, i, i + (l + t) * d) > dmin) l += t;
    int j = i + l * d;
    int dnode = delta(keys, n, i, j);
    int s = 0, t = l;
    do {
        t = (t + 1) >> 1;
        if (delta(keys, n, i, i + (s + t) * d) > dnode) s += t;
    } while (t > 1);
    int gamma = i + s * d + (d < 0 ? -1 : 0);
    int lo = i < j ? i : j, hi = i < j ? j : i;
    first[i] = lo;
    last[i] = hi;
    bool left_leaf = lo == gamma, right_leaf = hi == gamma + 1;
    left[i] = (gamma << 1) | (left_leaf ? 1 : 0);
    right[i] = ((gamma + 1) << 1) | (right_leaf ? 1 : 0);
    if (left_leaf) par_leaf[gamma] = i;
    else par_inner[gamma] = i;
    if (right_leaf) par_leaf[gamma + 1] = i;
    else par_inner[gamma + 1] = i;
    if (i == 0) par_inner[0] = -1;
}

struct tree_view {
    int n, m;                // triangles, inner nodes
    const fbox *tri_box;
    const uint32_t *vals;
    const int32_t *left, *right, *par_inner, *par_leaf, *first, *last;
    fbox *node_box;
    int32_t *size;
    uint32_t *visits;
    int32_t *pre;
    int max_leaf;
};

__device__ __forceinline__ bool emitted(const tree_view &v, int i) { return v.last[i] - v.first[i] + 1 > v.max_leaf; }

// Every index read from the tree arrays is range-checked: a malformed tree
// ends up rejected by the host-side layout check, never as a stray access.
__device__ __forceinline__ bool child_ok(const tree_view &v, int32_t c) {
    return c >= 0 && ((c & 1) ? (c >> 1) < v.n : (c >> 1) < v.m);
}
__device__ __forceinline__ fbox child_box(const tree_view &v, int32_t c) {
    if (!child_ok(v, c)) return empty_fbox();
    return (c & 1) ? v.tri_box[v.vals[c >> 1]] : v.node_box[c >> 1];
}
__device__ __forceinline__ int32_t child_size(const tree_view &v, int32_t c) {
    return ((c & 1) || !child_ok(v, c)) ? 0 : v.size[c >> 1];
}

// ---- 5. boxes and emitted sizes, bottom-up ----
__global__ __launch_bounds__(kBlock) void k_bottom_up(tree_view v, int n) {
    int k = (int)(blockIdx.x * kBlock + threadIdx.x);
    if (k >= n) return;
    int p = v.par_leaf[k];
    for (int steps = 0; p >= 0 && p < v.m && steps < kMaxClimb; ++steps) {
        // the first child to arrive stops; the second finds both children final
        uint32_t before = __hip_atomic_fetch_add(v.visits + p, 1u, __ATOMIC_ACQ_REL, __HIP_MEMORY_SCOPE_AGENT);
        if (before == 0) return;
        int32_t l = v.left[p], r = v.right[p];
        fbox b = child_box(v, l);
        merge(b, child_box(v, r));
        v.node_box[p] = b;
        v.size[p] = emitted(v, p) ? 1 + child_size(v, l) + child_size(v, r) : 0;
        p = v.par_inner[p];
    }
}

// ---- 6. pre-order index of every emitted node ----
__global__ __launch_bounds__(kBlock) void k_preorder(tree_view v, int m, uint32_t *counts) {
    int i = (int)(blockIdx.x * kBlock + threadIdx.x);
    if (i >= m || !emitted(v, i)) return;
    int idx = 0, depth = 1, c = i;
    for (int steps = 0; c != 0; ++steps) {
        int p = v.par_inner[c];
        if (p < 0 || p >= m || steps >= kMaxClimb) {   // not a tree: reported, rejected by the host
            atomicAdd(counts + 3, 1u);
            v.pre[i] = -1;
            return;
        }
        bool is_right = v.right[p] == (c << 1);
        idx += 1 + (is_right ? child_size(v, v.left[p]) : 0);
        c = p;
        ++depth;
    }
    v.pre[i] = idx;
    atomicMax(counts + 1, (uint32_t)depth);
    if (i == 0) counts[0] = (uint32_t)v.size[0];
}

// ---- 7. flattened nodes and triangle records ----
__device__ __forceinline__ int32_t child_ref(const tree_view &v, int32_t c) {
    if (!child_ok(v, c)) return kInvalidRef;
    int idx = c >> 1;
    if (c & 1) return make_leaf_ref((uint32_t)idx, 1u);
    if (emitted(v, idx)) return v.pre[idx] >= 0 ? v.pre[idx] : kInvalidRef;
    return make_leaf_ref((uint32_t)v.first[idx], (uint32_t)(v.last[idx] - v.first[idx] + 1));
}

__global__ __launch_bounds__(kBlock) void k_emit(tree_view v, int m, node2 *nodes, int32_t *root_ref) {
    int i = (int)(blockIdx.x * kBlock + threadIdx.x);
    if (i >= m || !emitted(v, i)) return;
    int32_t l = v.left[i], r = v.right[i];
    fbox bl = child_box(v, l), br = child_box(v, r);
    node2 o;
    o.lxlo = bl.lo[0];
    o.lxhi = bl.hi[0];
    o.lylo = bl.lo[1];
    o.lyhi = bl.hi[1];
    o.rxlo = br.lo[0];
    o.rxhi = br.hi[0];
    o.rylo = br.lo[1];
    o.ryhi = br.hi[1];
    o.lzlo = bl.lo[2];
    o.lzhi = bl.hi[2];
    o.rzlo = br.lo[2];
    o.rzhi = br.hi[2];
    o.ref_l = child_ref(v, l);
    o.ref_r = child_ref(v, r);
    o.pad0 = 0;
    o.pad1 = 0;
    if (v.pre[i] < 0 || v.pre[i] >= m) return;   // reported by k_preorder
    nodes[v.pre[i]] = o;
    if (i == 0) *root_ref = 0;
}

__global__ __launch_bounds__(kBlock) void k_records(lbvh_in in, const uint32_t *vals, tri_rec *tris,
                                                    uint32_t *slot_ids) {
    uint32_t k = blockIdx.x * kBlock + threadIdx.x;
    if (k >= in.n_tris) return;
    uint32_t t = vals[k];
    vec3f P[3];
    for (int c = 0; c < 3; ++c) {
        uint32_t vi = in.vidx[3ull * t + c];
        if (vi >= in.n_vert) vi = 0;   // reported through counts[2]
        P[c] = vec3f{in.pos[3ull * vi], in.pos[3ull * vi + 1], in.pos[3ull * vi + 2]};
    }
    tris[k] = in.tri == TRI_WOOP ? make_woop_rec(P[0], P[1], P[2]) : make_exact_rec(P[0], P[1], P[2], t);
    slot_ids[k] = t;
}

// Trees of at most max_leaf triangles (including one triangle): the root is a leaf.
__global__ void k_root_leaf(uint32_t n, int32_t *root_ref, uint32_t *counts) {
    if (threadIdx.x == 0 && blockIdx.x == 0) {
        *root_ref = n ? make_leaf_ref(0u, n) : kInvalidRef;
        counts[0] = 0;
        counts[1] = 0;
    }
}

__global__ void k_fill_i32(int32_t *p, int32_t value, uint32_t n) {
    uint32_t i = blockIdx.x * kBlock + threadIdx.x;
    if (i < n) p[i] = value;
}

hipError_t fill_i32(int32_t *p, int32_t value, uint32_t n, hipStream_t s) {
    if (n == 0) return hipSuccess;
    hipLaunchKernelGGL(k_fill_i32, dim3((n + kBlock - 1) / kBlock), dim3(kBlock), 0, s, p, value, n);
    return hipGetLastError();
}

inline dim3 grid_for(uint32_t n) { return dim3((n + kBlock - 1) / kBlock); }

} // namespace

size_t lbvh_scratch_bytes(uint32_t n_tris) {
    if (n_tris == 0) return 256;
    return map_scratch(n_tris, sort_storage(n_tris)).total;
}

hipError_t launch_lbvh(const lbvh_in &in, const lbvh_out &out, void *scratch, hipStream_t s) {
    const uint32_t n = in.n_tris;
    hipError_t e = fill_i32(reinterpret_cast<int32_t *>(out.counts), 0, 4, s);
    if (e != hipSuccess || n == 0) {
        if (e == hipSuccess) {
            hipLaunchKernelGGL(k_root_leaf, dim3(1), dim3(64), 0, s, 0u, out.root_ref, out.counts);
            e = hipGetLastError();
        }
        return e;
    }
    const scratch_map sm = map_scratch(n, sort_storage(n));
    fbox *tri_box = at<fbox>(scratch, sm.tri_box), *partial = at<fbox>(scratch, sm.partial);
    fbox *bounds = at<fbox>(scratch, sm.bounds);
    uint64_t *keys0 = at<uint64_t>(scratch, sm.keys0), *keys1 = at<uint64_t>(scratch, sm.keys1);
    uint32_t *vals0 = at<uint32_t>(scratch, sm.vals0), *vals1 = at<uint32_t>(scratch, sm.vals1);
    const uint32_t nb = (n + kBlock - 1) / kBlock;

    hipLaunchKernelGGL(k_tri_boxes, grid_for(n), dim3(kBlock), 0, s, in, tri_box, partial, out.counts);
    if ((e = hipGetLastError()) != hipSuccess) return e;
    hipLaunchKernelGGL(k_reduce_bounds, dim3(1), dim3(kBlock), 0, s, partial, nb, bounds);
    if ((e = hipGetLastError()) != hipSuccess) return e;
    hipLaunchKernelGGL(k_morton, grid_for(n), dim3(kBlock), 0, s, tri_box, bounds, n, keys0, vals0);
    if ((e = hipGetLastError()) != hipSuccess) return e;
    size_t sort_bytes = sm.sort_bytes;
    e = rocprim::radix_sort_pairs(at<void>(scratch, sm.sort_tmp), sort_bytes, keys0, keys1, vals0, vals1, n, 0,
                                  3 * kMortonBits, s);
    if (e != hipSuccess) return e;

    // records in Morton order (every slot is one triangle)
    hipLaunchKernelGGL(k_records, grid_for(n), dim3(kBlock), 0, s, in, vals1, out.tris, out.slot_ids);
    if ((e = hipGetLastError()) != hipSuccess) return e;
    if (n <= (uint32_t)in.max_leaf) {
        hipLaunchKernelGGL(k_root_leaf, dim3(1), dim3(64), 0, s, n, out.root_ref, out.counts);
        return hipGetLastError();
    }

    const uint32_t m = n - 1;
    tree_view v{};
    v.tri_box = tri_box;
    v.vals = vals1;
    v.left = at<int32_t>(scratch, sm.left);
    v.right = at<int32_t>(scratch, sm.right);
    v.par_inner = at<int32_t>(scratch, sm.par_inner);
    v.par_leaf = at<int32_t>(scratch, sm.par_leaf);
    v.first = at<int32_t>(scratch, sm.first);
    v.last = at<int32_t>(scratch, sm.last);
    v.node_box = at<fbox>(scratch, sm.node_box);
    v.size = at<int32_t>(scratch, sm.size);
    v.visits = at<uint32_t>(scratch, sm.visits);
    v.pre = at<int32_t>(scratch, sm.pre);
    v.max_leaf = in.max_leaf;
    v.n = (int)n;
    v.m = (int)m;
    if ((e = fill_i32(reinterpret_cast<int32_t *>(v.visits), 0, m, s)) != hipSuccess) return e;
    if ((e = fill_i32(const_cast<int32_t *>(v.par_inner), -1, m, s)) != hipSuccess) return e;
    if ((e = fill_i32(const_cast<int32_t *>(v.par_leaf), -1, n, s)) != hipSuccess) return e;
    hipLaunchKernelGGL(k_karras, grid_for(m), dim3(kBlock), 0, s, keys1, (int)n, const_cast<int32_t *>(v.left),
                       const_cast<int32_t *>(v.right), const_cast<int32_t *>(v.par_inner),
                       const_cast<int32_t *>(v.par_leaf), const_cast<int32_t *>(v.first),
                       const_cast<int32_t *>(v.last));
    if ((e = hipGetLastError()) != hipSuccess) return e;
    hipLaunchKernelGGL(k_bottom_up, grid_for(n), dim3(kBlock), 0, s, v, (int)n);
    if ((e = hipGetLastError()) != hipSuccess) return e;
    hipLaunchKernelGGL(k_preorder, grid_for(m), dim3(kBlock), 0, s, v, (int)m, out.counts);
    if ((e = hipGetLastError()) != hipSuccess) return e;
    hipLaunchKernelGGL(k_emit, grid_for(m), dim3(kBlock), 0, s, v, (int)m, out.nodes, out.root_ref);
    return hipGetLastError();
}

} // namespace dev
} // namespace rt
