// Plain while-while trace kernels: one ray per lane, grid sized to the batch,
// stack in private memory.  This is the per-bounce launch form the batched
// persistent kernels are measured against, and the exact-order kernel the
// parity tests hold to the CPU tracer bit for bit.
//
// Traversal order follows binary_bvh_tracer::closest_hit/any_hit
// (rt/cpu/bvh2.cpp:72-189): at an inner node both child boxes are tested; if
// both are hit the nearer is entered next (left when dist_l < dist_r, else
// right) and the other pushed; leaf slots are tested in index order and a hit
// replaces the best one only when strictly closer.
#include "kernels.h"
#include "trace_core.h"

namespace rt {
namespace dev {

template <int TRI, bool ANY, bool COUNT>
__device__ __forceinline__ void traverse_plain(const scene_desc &sc, const dray &r, bool exact, dhit &h,
                                               uint32_t &n_nodes, uint32_t &n_tris) {
    int32_t stack[kPlainStackCap];
    int sp = 0;
    stack[0] = kInvalidRef;
    int32_t cur = sc.root_ref;
    const node2 *nodes = static_cast<const node2 *>(sc.nodes);
    h.found = false;
    h.over = false;
    h.t = ANY ? INFINITY : FLT_MAX;
    while (cur != kInvalidRef) {
        while (cur >= 0 && cur != kInvalidRef) {
            if (COUNT) ++n_nodes;
            const float4 *np = reinterpret_cast<const float4 *>(nodes + cur);
            float4 n0 = np[0], n1 = np[1], n2 = np[2], n3 = np[3];
            float dl, dr;
            bool hl = box_test<ANY>(exact, n0.x, n0.y, n0.z, n0.w, n2.x, n2.y, r, h.t, dl);
            bool hr = box_test<ANY>(exact, n1.x, n1.y, n1.z, n1.w, n2.z, n2.w, r, h.t, dr);
            int32_t rl = __float_as_int(n3.x), rr = __float_as_int(n3.y);
            if (hl && hr) {
                bool left_first = dl < dr;
                int32_t far = left_first ? rr : rl;
                cur = left_first ? rl : rr;
                if (sp < kPlainStackCap - 1) stack[++sp] = far;
                else h.over = true;
            } else if (hl) {
                cur = rl;
            } else if (hr) {
                cur = rr;
            } else {
                cur = stack[sp--];
            }
        }
        if (cur == kInvalidRef) break;
        uint32_t slot = leaf_slot(cur), cnt = leaf_count(cur);
        for (uint32_t k = 0; k < cnt; ++k) {
            float t, u, v;
            if (COUNT) ++n_tris;
            if (test_slot<TRI>(sc, slot + k, r, h.t, t, u, v) &&
                (!sc.alpha || alpha_opaque(sc, slot_tri_id<TRI>(sc, slot + k), u, v))) {
                h.found = true;
                h.slot = slot + k;
                if (ANY) return;
                h.t = t;
                h.u = u;
                h.v = v;
            }
        }
        cur = stack[sp--];
    }
}

template <int TRI, bool ANY, bool COUNT>
__global__ __launch_bounds__(256) void k_trace_plain(scene_desc sc, ray_soa rays, hit_soa hits, uint8_t *occluded,
                                                     uint32_t n, unsigned long long *stats) {
    uint32_t i = blockIdx.x * 256u + threadIdx.x;
    if (i >= n) return;
    dray r = load_ray(rays, i);
    // the select-based box test only where some ray of the wavefront (or the
    // scene) has a non-finite component; both decide identically otherwise
    const bool exact = !sc.nan_free || __ballot(!ray_is_finite(r)) != 0;
    dhit h;
    uint32_t c_nodes = 0, c_tris = 0;
    traverse_plain<TRI, ANY, COUNT>(sc, r, exact, h, c_nodes, c_tris);
    if (COUNT) {
        atomicAdd(stats, (unsigned long long)c_nodes);
        atomicAdd(stats + 1, (unsigned long long)c_tris);
    }
    if (h.over && !(ANY && h.found)) {
        store_lost(hits, ANY ? occluded : nullptr, i);
    } else if (ANY) {
        if (occluded) occluded[i] = h.found ? 1 : 0;
        // The reference's CPU adapter marks an occluded ray with
        // t = ref = (unsigned)-1 and resets the others (rt/cpu/wavefront.cpp:60-70).
        if (hits.t) store_hit(hits, i, h.found ? 4294967295.0f : FLT_MAX, 0.0f, 0.0f, h.found ? 0xffffffffu : 0u);
    } else if (h.found) {
        store_hit(hits, i, h.t, h.u, h.v, slot_tri_id<TRI>(sc, h.slot));
    } else {
        store_hit(hits, i, FLT_MAX, 0.0f, 0.0f, 0u);
    }
}

namespace {
template <int TRI, bool ANY>
void launch_tri_any(const scene_desc &sc, const ray_soa &rays, const hit_soa &hits, uint8_t *occluded, uint32_t n,
                    unsigned long long *stats, hipStream_t stream) {
    dim3 grid((n + 255u) / 256u), block(256);
    if (stats) hipLaunchKernelGGL((k_trace_plain<TRI, ANY, true>), grid, block, 0, stream, sc, rays, hits, occluded, n, stats);
    else hipLaunchKernelGGL((k_trace_plain<TRI, ANY, false>), grid, block, 0, stream, sc, rays, hits, occluded, n, stats);
}
} // namespace

hipError_t launch_trace_plain(const scene_desc &sc, const ray_soa &rays, const hit_soa &hits, uint8_t *occluded,
                              uint32_t n, bool any, hipStream_t stream, unsigned long long *stats) {
    if (n == 0) return hipSuccess;
    if (sc.tri == TRI_EXACT) {
        if (any) launch_tri_any<TRI_EXACT, true>(sc, rays, hits, occluded, n, stats, stream);
        else launch_tri_any<TRI_EXACT, false>(sc, rays, hits, occluded, n, stats, stream);
    } else {
        if (any) launch_tri_any<TRI_WOOP, true>(sc, rays, hits, occluded, n, stats, stream);
        else launch_tri_any<TRI_WOOP, false>(sc, rays, hits, occluded, n, stats, stream);
    }
    return hipGetLastError();
}

} // namespace dev
} // namespace rt
