// Persistent while-while trace kernels for wave64.
//
// A fixed grid of 256-thread blocks (four waves) stays resident; each wave
// takes work from one global counter with a single atomicAdd per fetch (the
// first lane adds the number of idle lanes and the wave splits the range by
// lane rank).  Per lane the traversal keeps the reference's decisions -- both
// child boxes tested, nearer child first, strict "<" on the best t -- and adds:
//
//   * a stack whose newest kRegStack entries live in VGPRs (a shift register
//     with static indices, so no scratch) and whose older entries spill to an
//     LDS column owned by the lane (level-major, lane-minor: conflict free);
//   * leaf postponement (SPEC): a lane that reaches a leaf parks it and keeps
//     descending until every lane still searching has a parked leaf, so the
//     wave runs triangle tests together instead of one lane at a time;
//   * dynamic refill (REFILL): when fewer than `refill_thr` lanes still have a
//     ray, the wave leaves the traversal loop and hands the idle lanes new
//     rays, so long rays do not hold 63 idle lanes hostage.
//
// Postponement and refill change the order in which boxes see the running
// best t, so the result can differ from the plain kernel only between
// triangles with exactly equal t.
#include "kernels.h"
#include "trace_core.h"

namespace rt {
namespace dev {

namespace {

template <int REG>
struct lane_stack {
    int32_t s[REG > 0 ? REG : 1];
    int rs;   // entries held in registers
    int ls;   // entries held in LDS
    int32_t *col;   // this lane's LDS column, entries 64 words apart

    __device__ __forceinline__ void reset() {
        rs = 0;
        ls = 0;
    }
    __device__ __forceinline__ void push(int32_t x) {
        if (REG == 0) {
            if (ls < kLdsStack) col[64 * ls++] = x;
            return;
        }
        if (rs == REG) {
            if (ls < kLdsStack) col[64 * ls++] = s[REG - 1];
        } else {
            ++rs;
        }
#pragma unroll
        for (int k = REG - 1; k > 0; --k) s[k] = s[k - 1];
        s[0] = x;
    }
    __device__ __forceinline__ int32_t pop() {
        if (REG > 0 && rs > 0) {
            int32_t x = s[0];
#pragma unroll
            for (int k = 0; k < REG - 1; ++k) s[k] = s[k + 1];
            --rs;
            return x;
        }
        if (ls > 0) return col[64 * --ls];
        return kInvalidRef;
    }
};

} // namespace

template <int TRI, bool ANY, bool SPEC, bool REFILL, int REG, bool COUNT>
__global__ __launch_bounds__(256) void k_trace_persistent(scene_desc sc, ray_soa rays, hit_soa hits,
                                                          uint8_t *occluded, uint32_t n_static,
                                                          const uint32_t *count_ptr, uint32_t *counter,
                                                          int refill_thr, unsigned long long *stats) {
    __shared__ int32_t lds_stack[kPersistentWavesPerBlock * kLdsStack * 64];
    const uint32_t n = count_ptr ? *count_ptr : n_static;
    const int lane = threadIdx.x & 63;
    const int wave = threadIdx.x >> 6;
    const node2 *nodes = static_cast<const node2 *>(sc.nodes);

    lane_stack<REG> st;
    st.col = lds_stack + wave * kLdsStack * 64 + lane;
    st.reset();

    dray r{};
    int32_t cur = kInvalidRef, leaf = kInvalidRef;
    uint32_t ray_id = 0;
    dhit h{};
    bool alive = false;
    bool exhausted = false;   // wave-uniform: the counter has run past n
    uint32_t c_nodes = 0, c_tris = 0;

    for (;;) {
        // ---- fetch: one atomic per wave for all idle lanes ----
        uint64_t idle = __ballot(!alive);
        int n_alive = 64 - __popcll(idle);
        bool fetch = !exhausted && (REFILL ? (n_alive < refill_thr) : (n_alive == 0));
        if (fetch) {
            uint32_t want = (uint32_t)__popcll(idle);
            uint32_t base = 0;
            if (lane == 0) base = atomicAdd(counter, want);
            base = __shfl(base, 0);
            if (base + want >= n) exhausted = true;
            if (!alive) {
                uint32_t rank = (uint32_t)__popcll(idle & ((1ull << lane) - 1ull));
                uint32_t idx = base + rank;
                if (idx < n) {
                    r = load_ray(rays, idx);
                    ray_id = idx;
                    cur = sc.root_ref;
                    leaf = kInvalidRef;
                    st.reset();
                    h.found = false;
                    h.t = ANY ? INFINITY : FLT_MAX;
                    alive = true;
                }
            }
        }
        if (__ballot(alive) == 0) break;

        if (alive) {
            // ---- traversal, while-while ----
            for (;;) {
                bool searching = true;
                while (cur >= 0 && cur != kInvalidRef) {
                    if (COUNT) ++c_nodes;
                    const float4 *np = reinterpret_cast<const float4 *>(nodes + cur);
                    float4 n0 = np[0], n1 = np[1], n2 = np[2], n3 = np[3];
                    float dl, dr;
                    bool hl = box_exact<ANY>(n0.x, n0.y, n0.z, n0.w, n2.x, n2.y, r, h.t, dl);
                    bool hr = box_exact<ANY>(n1.x, n1.y, n1.z, n1.w, n2.z, n2.w, r, h.t, dr);
                    int32_t rl = __float_as_int(n3.x), rr = __float_as_int(n3.y);
                    if (hl && hr) {
                        bool left_first = dl < dr;
                        st.push(left_first ? rr : rl);
                        cur = left_first ? rl : rr;
                    } else if (hl) {
                        cur = rl;
                    } else if (hr) {
                        cur = rr;
                    } else {
                        cur = st.pop();
                    }
                    if (SPEC) {
                        if (cur < 0 && leaf >= 0) {
                            leaf = cur;
                            cur = st.pop();
                            searching = false;
                        }
                        if (__ballot(searching) == 0) break;
                    }
                }
                // ---- leaves ----
                // Without postponement (or when the root itself is a leaf)
                // the leaf is still in cur.
                if (leaf >= 0 && cur < 0) {
                    leaf = cur;
                    cur = st.pop();
                }
                while (leaf < 0) {
                    uint32_t slot = leaf_slot(leaf), cnt = leaf_count(leaf);
                    for (uint32_t k = 0; k < cnt; ++k) {
                        float t, u, v;
                        if (COUNT) ++c_tris;
                        if (test_slot<TRI>(sc, slot + k, r, h.t, t, u, v)) {
                            h.found = true;
                            h.slot = slot + k;
                            if (!ANY) {
                                h.t = t;
                                h.u = u;
                                h.v = v;
                            }
                        }
                    }
                    if (ANY && h.found) {
                        cur = kInvalidRef;
                        leaf = kInvalidRef;
                        break;
                    }
                    leaf = cur;
                    if (cur < 0) cur = st.pop();
                }
                if (cur == kInvalidRef) break;
                if (REFILL && __popcll(__ballot(1)) < refill_thr) break;
            }
            if (cur == kInvalidRef) {
                // ---- terminated: write the result ----
                if (ANY) {
                    if (occluded) occluded[ray_id] = h.found ? 1 : 0;
                    if (hits.t)
                        store_hit(hits, ray_id, h.found ? 4294967295.0f : FLT_MAX, 0.0f, 0.0f,
                                  h.found ? 0xffffffffu : 0u);
                } else if (h.found) {
                    store_hit(hits, ray_id, h.t, h.u, h.v, slot_tri_id<TRI>(sc, h.slot));
                } else {
                    store_hit(hits, ray_id, FLT_MAX, 0.0f, 0.0f, 0u);
                }
                alive = false;
            }
        }
    }
    if (COUNT) {
        atomicAdd(stats, (unsigned long long)c_nodes);
        atomicAdd(stats + 1, (unsigned long long)c_tris);
    }
}

__global__ void k_fill_u32(uint32_t *p, uint32_t value, uint32_t n) {
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = value;
}

hipError_t launch_fill_u32(uint32_t *p, uint32_t value, uint32_t n, hipStream_t stream) {
    if (n == 0) return hipSuccess;
    hipLaunchKernelGGL(k_fill_u32, dim3((n + 255) / 256), dim3(256), 0, stream, p, value, n);
    return hipGetLastError();
}

namespace {

// Blocks of one instantiation that fit the device at once (occupancy x CUs),
// asked once per device.
template <int TRI, bool ANY, bool SPEC, bool REFILL>
hipError_t grid_of(int *blocks) {
    static int cached[64] = {0};
    int dev = 0;
    hipError_t e = hipGetDevice(&dev);
    if (e != hipSuccess) return e;
    if (dev < 0 || dev >= 64) return hipErrorInvalidDevice;
    if (cached[dev] == 0) {
        int cus = 0, per_cu = 0;
        e = hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev);
        if (e != hipSuccess) return e;
        e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, k_trace_persistent<TRI, ANY, SPEC, REFILL, kRegStack, false>,
                                                         256, 0);
        if (e != hipSuccess) return e;
        cached[dev] = cus * (per_cu > 0 ? per_cu : 1);
    }
    *blocks = cached[dev];
    return hipSuccess;
}

template <int TRI, bool ANY, bool SPEC, bool REFILL>
hipError_t launch_variant(const scene_desc &sc, const ray_soa &rays, const hit_soa &hits, uint8_t *occluded,
                          uint32_t n, const uint32_t *count_ptr, uint32_t *counter, const persistent_cfg &cfg,
                          hipStream_t stream) {
    int blocks = cfg.blocks;
    if (blocks <= 0) {
        hipError_t e = grid_of<TRI, ANY, SPEC, REFILL>(&blocks);
        if (e != hipSuccess) return e;
    }
    if (cfg.stats)
        hipLaunchKernelGGL((k_trace_persistent<TRI, ANY, SPEC, REFILL, kRegStack, true>), dim3(blocks), dim3(256), 0,
                           stream, sc, rays, hits, occluded, n, count_ptr, counter, cfg.refill_thr, cfg.stats);
    else
        hipLaunchKernelGGL((k_trace_persistent<TRI, ANY, SPEC, REFILL, kRegStack, false>), dim3(blocks), dim3(256), 0,
                           stream, sc, rays, hits, occluded, n, count_ptr, counter, cfg.refill_thr, cfg.stats);
    return hipGetLastError();
}

template <int TRI, bool ANY>
hipError_t grid_tri_any(const persistent_cfg &cfg, int *blocks) {
    if (cfg.speculative) {
        if (cfg.refill) return grid_of<TRI, ANY, true, true>(blocks);
        return grid_of<TRI, ANY, true, false>(blocks);
    }
    if (cfg.refill) return grid_of<TRI, ANY, false, true>(blocks);
    return grid_of<TRI, ANY, false, false>(blocks);
}

template <int TRI, bool ANY>
hipError_t launch_tri_any(const scene_desc &sc, const ray_soa &rays, const hit_soa &hits, uint8_t *occluded,
                          uint32_t n, const uint32_t *count_ptr, uint32_t *counter, const persistent_cfg &cfg,
                          hipStream_t stream) {
    if (cfg.speculative) {
        if (cfg.refill) return launch_variant<TRI, ANY, true, true>(sc, rays, hits, occluded, n, count_ptr, counter, cfg, stream);
        return launch_variant<TRI, ANY, true, false>(sc, rays, hits, occluded, n, count_ptr, counter, cfg, stream);
    }
    if (cfg.refill) return launch_variant<TRI, ANY, false, true>(sc, rays, hits, occluded, n, count_ptr, counter, cfg, stream);
    return launch_variant<TRI, ANY, false, false>(sc, rays, hits, occluded, n, count_ptr, counter, cfg, stream);
}

} // namespace

hipError_t persistent_grid(const scene_desc &sc, bool any, const persistent_cfg &cfg, int *blocks) {
    if (sc.tri == TRI_EXACT) return any ? grid_tri_any<TRI_EXACT, true>(cfg, blocks) : grid_tri_any<TRI_EXACT, false>(cfg, blocks);
    return any ? grid_tri_any<TRI_WOOP, true>(cfg, blocks) : grid_tri_any<TRI_WOOP, false>(cfg, blocks);
}

hipError_t launch_trace_persistent(const scene_desc &sc, const ray_soa &rays, const hit_soa &hits,
                                   uint8_t *occluded, uint32_t n, const uint32_t *count_ptr, uint32_t *counter,
                                   bool any, const persistent_cfg &cfg, hipStream_t stream) {
    hipError_t e = launch_fill_u32(counter, 0u, 1u, stream);
    if (e != hipSuccess) return e;
    if (sc.tri == TRI_EXACT) {
        if (any) return launch_tri_any<TRI_EXACT, true>(sc, rays, hits, occluded, n, count_ptr, counter, cfg, stream);
        return launch_tri_any<TRI_EXACT, false>(sc, rays, hits, occluded, n, count_ptr, counter, cfg, stream);
    }
    if (any) return launch_tri_any<TRI_WOOP, true>(sc, rays, hits, occluded, n, count_ptr, counter, cfg, stream);
    return launch_tri_any<TRI_WOOP, false>(sc, rays, hits, occluded, n, count_ptr, counter, cfg, stream);
}

} // namespace dev
} // namespace rt
