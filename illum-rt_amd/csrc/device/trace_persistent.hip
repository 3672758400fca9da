// Persistent while-while trace kernels for wave64.
//
// A fixed grid of 256-thread blocks (four waves) stays resident; each wave
// takes work from one global counter with a single atomicAdd per fetch (the
// first lane adds the number of idle lanes and the wave splits the range by
// lane rank).  Per lane the traversal keeps the reference's decisions -- both
// child boxes tested with intersect4's arithmetic, nearer child first, strict
// "<" on the best t -- and adds:
//
//   * a stack whose newest REG entries live in VGPRs (a shift register with
//     static indices, so no scratch), the next LDS entries in a ring in an
//     LDS column owned by the lane (level-major, lane-minor: conflict free),
//     and the oldest, when a deep traversal outgrows both, in a per-lane
//     column of a global spill buffer sized by the host for the tree;
//   * leaf postponement (SPEC): a lane that reaches a leaf parks it and keeps
//     descending until every lane still searching has a parked leaf, so the
//     wave runs triangle tests together instead of one lane at a time.  A
//     closest-hit lane parks up to `park` leaves (1..4, default
//     kParkClosest = 4; any-hit: one), so
//     lanes that find leaves early keep doing node work while the wave waits
//     for its slowest lane's first leaf (node-loop lane occupancy on diffuse
//     hall rays 33% with one parked leaf, 55% with four);
//   * dynamic refill (REFILL): when fewer than `refill_thr` lanes still have a
//     ray, the wave leaves the traversal loop and hands the idle lanes new
//     rays, so long rays do not hold 63 idle lanes hostage;
//   * 4-wide nodes (BVH == 4): four child boxes per 128-byte node, hit
//     children sorted by entry distance, the nearest entered, the others
//     pushed farthest first;
//   * a descriptor table (MULTI): the rays of many scenes form one index
//     space; each lane binds the scene of the ray it fetched, so one launch
//     traces uneven batches of many scenes;
//   * work chunks (n_chunks = 8): the ray range is split into one contiguous
//     chunk per XCD, each with its own counter; a block starts on the chunk
//     of the XCD it runs on (the dispatcher places workgroup b on XCD b % 8)
//     and moves on to the next chunk when its own runs dry.  Neighbouring
//     rays -- neighbouring pixels, neighbouring path origins -- then share
//     one XCD's L2 instead of being spread over all eight;
//   * a register cap for single-scene binary-tree kernels with a 20-entry
//     LDS stack and no VGPR entries (waves_hint): 8 waves per SIMD instead
//     of the 7 their SGPR count would allow.
//
// Postponement, refill and 4-wide nodes change the order in which boxes see
// the running best t, so results can differ from the plain kernel only
// between triangles whose t agree within rounding (ties); which of them a
// postponing kernel returns depends on when the lane tests its parked
// leaves, a wave-wide vote, and so can vary between runs.
#include "kernels.h"
#include "trace_core.h"

namespace rt {
namespace dev {

namespace {

// Spill column of one lane: `cap` stack entries `stride` words apart in a
// global buffer (entry k of every lane of the grid is contiguous).
struct spill_col {
    int32_t *p;
    uint32_t stride, cap;
};

// RING selects the LDS part: a ring that spills its oldest entries to the
// global column (any tree depth), or a plain LDS array (trees whose stack
// need fits REG + LDS; two fewer VALU operations per push and pop).
template <int REG, int LDS, bool RING>
struct lane_stack {
    static_assert(LDS > 0, "the LDS part needs at least one entry");
    int32_t s[REG > 0 ? REG : 1];
    int rs;         // entries held in registers
    int ls;         // entries held in the LDS ring
    int ts;         // ring slot the next LDS entry goes to
    uint32_t gs;    // entries spilled to the global column (the oldest ones)
    int32_t *col;   // this lane's LDS column, entries 64 words apart
    spill_col g;
    bool over;      // an entry was dropped: the host sized the stack too small

    __device__ __forceinline__ void reset() {
        rs = 0;
        ls = 0;
        ts = 0;
        gs = 0;
        over = false;
    }
    // Below the registers: an LDS ring of LDS entries.  When it is full, its
    // oldest entry (the one in slot ts) moves to the spill column, so the
    // order of the whole stack is kept: registers newest, then the ring, then
    // the spill column.  The host sizes the column for the tree's stack
    // need; entries beyond it are dropped, never written out of bounds, and
    // the ray's result is then marked lost (store_lost).
    __device__ __forceinline__ void push_lds(int32_t x) {
        if (!RING) {
            // the host picks this form only when the tree's need fits
            if (ls < LDS) col[64 * ls++] = x;
            else over = true;
            return;
        }
        if (ls == LDS) {
            if (gs < g.cap) g.p[(size_t)gs * g.stride] = col[64 * ts];
            else over = true;
            ++gs;
        } else {
            ++ls;
        }
        col[64 * ts] = x;
        if constexpr ((LDS & (LDS - 1)) == 0) ts = (ts + 1) & (LDS - 1);
        else ts = ts + 1 == LDS ? 0 : ts + 1;
    }
    __device__ __forceinline__ int32_t pop_lds() {
        if (!RING) return ls > 0 ? col[64 * --ls] : kInvalidRef;
        if (ls > 0) {
            --ls;
            if constexpr ((LDS & (LDS - 1)) == 0) ts = (ts - 1) & (LDS - 1);
            else ts = ts == 0 ? LDS - 1 : ts - 1;
            return col[64 * ts];
        }
        if (gs > 0) {
            --gs;
            return gs < g.cap ? g.p[(size_t)gs * g.stride] : kInvalidRef;
        }
        return kInvalidRef;
    }
    __device__ __forceinline__ void push(int32_t x) {
        if (REG == 0) {
            push_lds(x);
            return;
        }
        if (rs == REG) push_lds(s[REG - 1]);
        else ++rs;
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
        return pop_lds();
    }
};

// The scene a lane traces.  For single-scene launches these are copies of
// the kernel arguments and stay wave-uniform.
struct lane_scene {
    const void *nodes;
    const tri_rec *tris;
    const uint32_t *slot_ids;
    int32_t root;
};

__device__ __forceinline__ lane_scene scene_of(const scene_desc &sc) {
    return lane_scene{sc.nodes, sc.tris, sc.slot_ids, sc.root_ref};
}

template <int TRI>
__device__ __forceinline__ bool test_rec(const lane_scene &ls, uint32_t slot, const dray &r, float best, float &t,
                                         float &u, float &v) {
    const float4 *p = reinterpret_cast<const float4 *>(ls.tris + slot);
    float4 r0 = p[0], r1 = p[1], r2 = p[2];
    if (TRI == TRI_EXACT) return tri_exact(r0, r1, r2, r, best, t, u, v);
    return tri_woop(r0, r1, r2, r, best, t, u, v);
}

template <int TRI>
__device__ __forceinline__ uint32_t rec_tri_id(const lane_scene &ls, uint32_t slot) {
    if (TRI == TRI_EXACT) return __float_as_uint(reinterpret_cast<const float4 *>(ls.tris + slot)[0].w);
    return ls.slot_ids[slot];
}

__device__ __forceinline__ void cswap(float &ka, int32_t &ra, float &kb, int32_t &rb) {
    if (kb < ka) {
        float tk = ka;
        ka = kb;
        kb = tk;
        int32_t tr = ra;
        ra = rb;
        rb = tr;
    }
}

// Index of the descriptor whose ray range holds global ray `idx`
// (descriptors sorted by `first`, the first starting at 0).
__device__ __forceinline__ uint32_t find_desc(const batch_desc *table, uint32_t n_desc, uint32_t idx) {
    uint32_t lo = 0, hi = n_desc;
    while (hi - lo > 1) {
        uint32_t mid = (lo + hi) >> 1;
        if (table[mid].first <= idx) lo = mid;
        else hi = mid;
    }
    return lo;
}

} // namespace

// Waves per SIMD the register allocation must leave room for: 8 for
// single-scene binary-tree launches with the 20-entry LDS stack and no VGPR
// entries (20 KB per block lets 8 blocks share a CU's LDS; the register cap
// costs SGPR spills to VGPR lanes only), no bound otherwise (4-wide and
// descriptor-table kernels would spill VGPRs to scratch under it).
constexpr int waves_hint(int reg, int lds, int bvh, bool multi) {
    return reg == 0 && lds == 20 && bvh == 2 && !multi ? 8 : 1;
}

template <int TRI, bool ANY, bool SPEC, bool REFILL, int REG, int LDS, bool COUNT, int BVH, bool MULTI, bool RING>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(waves_hint(REG, LDS, BVH, MULTI))))
void k_trace_persistent(scene_desc sc, ray_soa rays, hit_soa hits, uint8_t *occluded, uint32_t n_static,
                        const uint32_t *count_ptr, uint32_t *counter, int lane_limits, unsigned long long *stats,
                        const batch_desc *table, uint32_t n_desc, uint32_t n_chunks, int32_t *spill,
                        uint32_t spill_cap, ray_gather gather) {
    __shared__ int32_t lds_stack[kPersistentWavesPerBlock * LDS * 64];
    const uint32_t n = count_ptr ? *count_ptr : n_static;
    uint32_t chunk = blockIdx.x % n_chunks;   // wave-uniform: the chunk this wave takes rays from
    uint32_t dry = 0;                         // chunks this wave has found exhausted
    const int lane = threadIdx.x & 63;
    const int wave = threadIdx.x >> 6;

    lane_stack<REG, LDS, RING> st;
    st.col = lds_stack + wave * LDS * 64 + lane;
    st.g = spill ? spill_col{spill + (size_t)blockIdx.x * 256u + threadIdx.x, gridDim.x * 256u, spill_cap}
                 : spill_col{nullptr, 0u, 0u};
    st.reset();

    lane_scene ls = scene_of(sc);
    uint32_t desc = 0;
    dray r{};
    int32_t cur = kInvalidRef;
    // parked leaves (SPEC): p0 is the oldest, pc how many are held
    // lane_limits: the refill threshold (low byte) and the parked-leaf limit
    // (next byte), one argument so the kernel takes no extra SGPR
    const int refill_thr = lane_limits & 0xFF;
    constexpr int PARK = ANY ? 1 : kParkClosest;   // registers held for parked leaves
    const int park = ANY ? 1 : ((lane_limits >> 8) & 0xFF);
    int32_t p0 = 0, p1 = 0, p2 = 0, p3 = 0;
    int pc = 0;
    uint32_t ray_id = 0;
    dhit h{};
    bool alive = false;
    bool exhausted = false;   // wave-uniform: every chunk's counter has run past its end
    bool finite = true;       // this lane's ray has only finite components
    bool exact = !sc.nan_free;   // wave-uniform: use the select-based box test
    bool alpha_on = !MULTI && sc.alpha;   // opacity test on hits (batches: the lane's descriptor)
    uint32_t c_nodes = 0, c_tris = 0;

    for (;;) {
        // ---- fetch: one atomic per wave for all idle lanes ----
        uint64_t idle = __ballot(!alive);
        int n_alive = 64 - __popcll(idle);
        bool fetch = !exhausted && (REFILL ? (n_alive < refill_thr) : (n_alive == 0));
        if (fetch) {
            uint32_t want = (uint32_t)__popcll(idle);
            const uint32_t lo = (uint32_t)((uint64_t)n * chunk / n_chunks);
            const uint32_t hi = (uint32_t)((uint64_t)n * (chunk + 1) / n_chunks);
            uint32_t base = 0;
            if (lane == 0) base = atomicAdd(counter + chunk * kCounterStride, want);
            base = __shfl(base, 0);
            // base counts rays taken from this chunk; it only grows
            const uint64_t first = (uint64_t)lo + base;
            if (first + want >= hi) {
                chunk = chunk + 1 == n_chunks ? 0 : chunk + 1;
                if (++dry >= n_chunks) exhausted = true;
            }
            if (!alive) {
                // idle lanes below this one (mbcnt: no 64-bit lane mask held in registers)
                uint32_t rank = __builtin_amdgcn_mbcnt_hi((uint32_t)(idle >> 32),
                                                          __builtin_amdgcn_mbcnt_lo((uint32_t)idle, 0u));
                uint64_t idx64 = first + rank;
                uint32_t idx = (uint32_t)idx64;
                if (idx64 < hi) {
                    if (MULTI) {
                        desc = find_desc(table, n_desc, idx);
                        const batch_desc &bd = table[desc];
                        ls = scene_of(bd.scene);
                        alpha_on = bd.scene.alpha != 0;
                        ray_id = idx - bd.first;
                        r = load_ray(bd.rays, ray_id);
                    } else {
                        ray_id = idx;
                        if (gather.aos) r = load_ray_aos(gather.aos, gather.perm ? gather.perm[idx] : idx, gather.stride,
                                                         gather.fixed_t != 0);
                        else r = load_ray(rays, idx);
                    }
                    cur = ls.root;
                    pc = 0;
                    st.reset();
                    h.found = false;
                    h.t = ANY ? INFINITY : FLT_MAX;
                    finite = ray_is_finite(r);
                    alive = true;
                }
            }
        }
        if (__ballot(alive) == 0) {
            // nothing fetched: try the next chunk, or leave once every chunk is dry
            if (exhausted) break;
            continue;
        }
        exact = __ballot(alive && !finite) != 0 || (MULTI ? false : !sc.nan_free);
        if (MULTI && !exact) exact = __ballot(alive && !table[desc].scene.nan_free) != 0;

        if (alive) {
            // ---- traversal, while-while ----
            for (;;) {
                bool searching = true;
                while (cur >= 0 && cur != kInvalidRef) {
                    if (COUNT) ++c_nodes;
                    if (BVH == 2) {
                        const float4 *np =
                            reinterpret_cast<const float4 *>(static_cast<const node2 *>(ls.nodes) + cur);
                        float4 n0 = np[0], n1 = np[1], n2 = np[2], n3 = np[3];
                        float dl, dr;
                        bool hl = box_test<ANY>(exact, n0.x, n0.y, n0.z, n0.w, n2.x, n2.y, r, h.t, dl);
                        bool hr = box_test<ANY>(exact, n1.x, n1.y, n1.z, n1.w, n2.z, n2.w, r, h.t, dr);
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
                    } else {
                        const float4 *np =
                            reinterpret_cast<const float4 *>(static_cast<const node4 *>(ls.nodes) + cur);
                        float4 lx = np[0], hx = np[1], ly = np[2], hy = np[3], lz = np[4], hz = np[5], rf = np[6];
                        float k0, k1, k2, k3;
                        int32_t c0 = __float_as_int(rf.x), c1 = __float_as_int(rf.y), c2 = __float_as_int(rf.z),
                                c3 = __float_as_int(rf.w);
                        bool b0 = box_test<ANY>(exact, lx.x, hx.x, ly.x, hy.x, lz.x, hz.x, r, h.t, k0) && c0 != kInvalidRef;
                        bool b1 = box_test<ANY>(exact, lx.y, hx.y, ly.y, hy.y, lz.y, hz.y, r, h.t, k1) && c1 != kInvalidRef;
                        bool b2 = box_test<ANY>(exact, lx.z, hx.z, ly.z, hy.z, lz.z, hz.z, r, h.t, k2) && c2 != kInvalidRef;
                        bool b3 = box_test<ANY>(exact, lx.w, hx.w, ly.w, hy.w, lz.w, hz.w, r, h.t, k3) && c3 != kInvalidRef;
                        int nh = (int)b0 + (int)b1 + (int)b2 + (int)b3;
                        // misses sort after hits; a NaN entry distance (axis-parallel
                        // ray on a slab plane) sorts last among the hits
                        k0 = b0 ? ((k0 == k0) ? k0 : FLT_MAX) : INFINITY;
                        k1 = b1 ? ((k1 == k1) ? k1 : FLT_MAX) : INFINITY;
                        k2 = b2 ? ((k2 == k2) ? k2 : FLT_MAX) : INFINITY;
                        k3 = b3 ? ((k3 == k3) ? k3 : FLT_MAX) : INFINITY;
                        cswap(k0, c0, k1, c1);
                        cswap(k2, c2, k3, c3);
                        cswap(k0, c0, k2, c2);
                        cswap(k1, c1, k3, c3);
                        cswap(k1, c1, k2, c2);
                        if (nh == 0) {
                            cur = st.pop();
                        } else {
                            if (nh == 4) st.push(c3);
                            if (nh >= 3) st.push(c2);
                            if (nh >= 2) st.push(c1);
                            cur = c0;
                        }
                    }
                    if (SPEC) {
                        // park leaves (up to PARK), then keep descending
                        while (cur < 0 && pc < park) {
                            if (PARK > 3) p3 = pc == 3 ? cur : p3;
                            if (PARK > 2) p2 = pc == 2 ? cur : p2;
                            if (PARK > 1) p1 = pc == 1 ? cur : p1;
                            p0 = pc == 0 ? cur : p0;
                            ++pc;
                            cur = st.pop();
                        }
                        searching = pc == 0;
                        if (__ballot(searching) == 0) break;
                    }
                }
                // ---- leaves: the parked ones, oldest first, then a leaf left in
                // cur and the leaves that follow it on the stack (the reference
                // order when nothing is parked) ----
                for (;;) {
                    int32_t lf;
                    if (SPEC && pc > 0) {
                        lf = p0;
                        if (PARK > 1) p0 = p1;
                        if (PARK > 2) p1 = p2;
                        if (PARK > 3) p2 = p3;
                        --pc;
                    } else if (cur < 0) {
                        lf = cur;
                        cur = st.pop();
                    } else {
                        break;
                    }
                    uint32_t slot = leaf_slot(lf), cnt = leaf_count(lf);
                    for (uint32_t k = 0; k < cnt; ++k) {
                        float t, u, v;
                        if (COUNT) ++c_tris;
                        if (test_rec<TRI>(ls, slot + k, r, h.t, t, u, v) &&
                            (!alpha_on || alpha_opaque(MULTI ? table[desc].scene : sc,
                                                       rec_tri_id<TRI>(ls, slot + k), u, v))) {
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
                        pc = 0;
                        break;
                    }
                }
                if (cur == kInvalidRef) break;
                if (REFILL && __popcll(__ballot(1)) < refill_thr) break;
            }
            if (cur == kInvalidRef) {
                // ---- terminated: write the result ----
                hit_soa ho = hits;
                uint8_t *occ = occluded;
                if (MULTI) {
                    ho = table[desc].hits;
                    occ = table[desc].occluded;
                }
                if (st.over && !(ANY && h.found)) {
                    store_lost(ho, ANY ? occ : nullptr, ray_id);
                } else if (ANY) {
                    if (occ) occ[ray_id] = h.found ? 1 : 0;
                    if (ho.t)
                        store_hit(ho, ray_id, h.found ? 4294967295.0f : FLT_MAX, 0.0f, 0.0f,
                                  h.found ? 0xffffffffu : 0u);
                } else if (h.found) {
                    store_hit(ho, ray_id, h.t, h.u, h.v, rec_tri_id<TRI>(ls, h.slot));
                } else {
                    store_hit(ho, ray_id, FLT_MAX, 0.0f, 0.0f, 0u);
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

// Stack configurations: (VGPR entries, LDS ring entries) per lane, for both
// tree widths.  A push or pop of the VGPR part shifts every register entry
// (static indices keep it out of scratch), so VGPR entries cost VALU moves;
// LDS entries cost LDS space, which bounds the blocks per CU.  Entries beyond
// both go to the global spill column, so a configuration never limits the
// tree depth; it only decides how often a deep traversal spills.
//   0: 4 VGPR + 28 LDS   28 KB of LDS per block
//   1: 8 VGPR + 16 LDS   16 KB
//   2: 2 VGPR + 20 LDS   20 KB
//   3: 0 VGPR + 24 LDS   24 KB
//   4: 0 VGPR + 20 LDS   20 KB, registers capped for 8 waves per SIMD
constexpr int reg_for(int stk) { return stk == 0 ? kRegStack : stk == 1 ? 8 : stk == 2 ? 2 : 0; }
constexpr int lds_for(int bvh, int stk) {
    (void)bvh;
    return stk == 0 ? kLdsStack : stk == 1 ? 16 : stk == 2 ? 20 : stk == 3 ? 24 : 20;
}

// Blocks of one instantiation that fit the device at once (occupancy x CUs),
// asked once per device.
template <int TRI, bool ANY, bool SPEC, bool REFILL, int BVH, bool MULTI, int STK, bool RING>
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
        e = hipOccupancyMaxActiveBlocksPerMultiprocessor(
            &per_cu,
            k_trace_persistent<TRI, ANY, SPEC, REFILL, reg_for(STK), lds_for(BVH, STK), false, BVH, MULTI, RING>, 256,
            0);
        if (e != hipSuccess) return e;
        cached[dev] = cus * (per_cu > 0 ? per_cu : 1);
    }
    *blocks = cached[dev];
    return hipSuccess;
}

struct launch_args {
    scene_desc sc;
    ray_soa rays;
    ray_gather gather;
    hit_soa hits;
    uint8_t *occluded;
    uint32_t n;
    const uint32_t *count_ptr;
    uint32_t *counter;
    const batch_desc *table;
    uint32_t n_desc;
};

__host__ __device__ constexpr uint32_t chunks_of(int c) { return c == 8 ? 8u : 1u; }

// Refill threshold (1..64) and parked-leaf limit (1..kParkClosest) packed
// into the kernel's lane_limits argument.
inline int lane_limits_of(const persistent_cfg &cfg) {
    int thr = cfg.refill_thr < 1 ? 1 : (cfg.refill_thr > 64 ? 64 : cfg.refill_thr);
    int park = cfg.park < 1 ? kParkClosest : (cfg.park > kParkClosest ? kParkClosest : cfg.park);
    return thr | (park << 8);
}

template <int TRI, bool ANY, bool SPEC, bool REFILL, int BVH, bool MULTI, int STK, bool RING>
hipError_t launch_ring(const launch_args &a, const persistent_cfg &cfg, hipStream_t stream, bool grid_only,
                       int *grid_out) {
    int blocks = cfg.blocks;
    if (blocks <= 0) {
        hipError_t e = grid_of<TRI, ANY, SPEC, REFILL, BVH, MULTI, STK, RING>(&blocks);
        if (e != hipSuccess) return e;
    }
    if (grid_out) *grid_out = blocks;
    if (grid_only) return hipSuccess;
    // the spill column of every lane of this grid must fit the buffer
    if (cfg.spill_cap && (!cfg.spill || cfg.spill_bytes < persistent_spill_bytes(blocks, cfg.spill_cap)))
        return hipErrorInvalidValue;
    constexpr int R = reg_for(STK), L = lds_for(BVH, STK);
    if constexpr (!MULTI && RING) {
        if (cfg.stats) {
            hipLaunchKernelGGL((k_trace_persistent<TRI, ANY, SPEC, REFILL, R, L, true, BVH, MULTI, true>),
                               dim3(blocks), dim3(256), 0, stream, a.sc, a.rays, a.hits, a.occluded, a.n, a.count_ptr,
                               a.counter, lane_limits_of(cfg), cfg.stats, a.table, a.n_desc, chunks_of(cfg.chunks),
                               cfg.spill, cfg.spill_cap, a.gather);
            return hipGetLastError();
        }
    }
    hipLaunchKernelGGL((k_trace_persistent<TRI, ANY, SPEC, REFILL, R, L, false, BVH, MULTI, RING>), dim3(blocks),
                       dim3(256), 0, stream, a.sc, a.rays, a.hits, a.occluded, a.n, a.count_ptr, a.counter,
                       lane_limits_of(cfg), cfg.stats, a.table, a.n_desc, chunks_of(cfg.chunks), cfg.spill, cfg.spill_cap,
                       a.gather);
    return hipGetLastError();
}

// The plain LDS stack runs single-scene launches whose tree fits the
// configuration; descriptor-table launches, counting launches and trees
// that spill use the ring.
template <int TRI, bool ANY, bool SPEC, bool REFILL, int BVH, bool MULTI, int STK>
hipError_t launch_variant(const launch_args &a, const persistent_cfg &cfg, hipStream_t stream, bool grid_only,
                          int *grid_out) {
    if constexpr (!MULTI) {
        if (!cfg.stats && cfg.spill_cap == 0)
            return launch_ring<TRI, ANY, SPEC, REFILL, BVH, MULTI, STK, false>(a, cfg, stream, grid_only, grid_out);
    }
    return launch_ring<TRI, ANY, SPEC, REFILL, BVH, MULTI, STK, true>(a, cfg, stream, grid_only, grid_out);
}

template <int TRI, bool ANY, int BVH, bool MULTI, int STK>
hipError_t dispatch_flags(const launch_args &a, const persistent_cfg &cfg, hipStream_t s, bool grid_only, int *g) {
    if (cfg.speculative) {
        if (cfg.refill) return launch_variant<TRI, ANY, true, true, BVH, MULTI, STK>(a, cfg, s, grid_only, g);
        return launch_variant<TRI, ANY, true, false, BVH, MULTI, STK>(a, cfg, s, grid_only, g);
    }
    if (cfg.refill) return launch_variant<TRI, ANY, false, true, BVH, MULTI, STK>(a, cfg, s, grid_only, g);
    return launch_variant<TRI, ANY, false, false, BVH, MULTI, STK>(a, cfg, s, grid_only, g);
}

template <int TRI, bool ANY, int BVH, bool MULTI>
hipError_t dispatch_stack(const launch_args &a, const persistent_cfg &cfg, hipStream_t s, bool grid_only, int *g) {
    switch (cfg.stack) {
    case 1: return dispatch_flags<TRI, ANY, BVH, MULTI, 1>(a, cfg, s, grid_only, g);
    case 2: return dispatch_flags<TRI, ANY, BVH, MULTI, 2>(a, cfg, s, grid_only, g);
    case 3: return dispatch_flags<TRI, ANY, BVH, MULTI, 3>(a, cfg, s, grid_only, g);
    case 4: return dispatch_flags<TRI, ANY, BVH, MULTI, 4>(a, cfg, s, grid_only, g);
    default: return dispatch_flags<TRI, ANY, BVH, MULTI, 0>(a, cfg, s, grid_only, g);
    }
}

template <bool MULTI>
hipError_t dispatch(int tri, bool any, int bvh, const launch_args &a, const persistent_cfg &cfg, hipStream_t s,
                    bool grid_only, int *g) {
    if (tri == TRI_EXACT) {
        if (bvh == BVH2)
            return any ? dispatch_stack<TRI_EXACT, true, 2, MULTI>(a, cfg, s, grid_only, g)
                       : dispatch_stack<TRI_EXACT, false, 2, MULTI>(a, cfg, s, grid_only, g);
        return any ? dispatch_stack<TRI_EXACT, true, 4, MULTI>(a, cfg, s, grid_only, g)
                   : dispatch_stack<TRI_EXACT, false, 4, MULTI>(a, cfg, s, grid_only, g);
    }
    if (bvh == BVH2)
        return any ? dispatch_stack<TRI_WOOP, true, 2, MULTI>(a, cfg, s, grid_only, g)
                   : dispatch_stack<TRI_WOOP, false, 2, MULTI>(a, cfg, s, grid_only, g);
    return any ? dispatch_stack<TRI_WOOP, true, 4, MULTI>(a, cfg, s, grid_only, g)
               : dispatch_stack<TRI_WOOP, false, 4, MULTI>(a, cfg, s, grid_only, g);
}

} // namespace

int persistent_stack_cap(int bvh, int stack) {
    int b = bvh == BVH4 ? 4 : 2;
    int k = (stack >= 0 && stack < kStackConfigs) ? stack : 0;
    return reg_for(k) + lds_for(b, k);
}

uint32_t persistent_spill_entries(int bvh, int stack, int need) {
    int cap = persistent_stack_cap(bvh, stack);
    return need > cap ? (uint32_t)(need - cap) : 0u;
}

size_t persistent_spill_bytes(int blocks, uint32_t entries) {
    return (size_t)(blocks > 0 ? blocks : 0) * 256u * entries * 4u;
}

hipError_t persistent_grid(const scene_desc &sc, bool any, const persistent_cfg &cfg, int *blocks, bool multi) {
    launch_args a{};
    if (multi) return dispatch<true>(sc.tri, any, sc.bvh, a, cfg, nullptr, true, blocks);
    return dispatch<false>(sc.tri, any, sc.bvh, a, cfg, nullptr, true, blocks);
}

hipError_t launch_trace_persistent(const scene_desc &sc, const ray_soa &rays, const hit_soa &hits,
                                   uint8_t *occluded, uint32_t n, const uint32_t *count_ptr, uint32_t *counter,
                                   bool any, const persistent_cfg &cfg, hipStream_t stream, ray_gather gather) {
    hipError_t e = launch_fill_u32(counter, 0u, kMaxChunks * kCounterStride, stream);
    if (e != hipSuccess) return e;
    launch_args a{sc, rays, gather, hits, occluded, n, count_ptr, counter, nullptr, 0};
    return dispatch<false>(sc.tri, any, sc.bvh, a, cfg, stream, false, nullptr);
}

hipError_t launch_trace_batch(const batch_desc *d_table, uint32_t n_desc, uint32_t total, int tri, int bvh, bool any,
                              uint32_t *counter, const persistent_cfg &cfg, hipStream_t stream) {
    hipError_t e = launch_fill_u32(counter, 0u, kMaxChunks * kCounterStride, stream);
    if (e != hipSuccess) return e;
    launch_args a{};
    a.n = total;
    a.counter = counter;
    a.table = d_table;
    a.n_desc = n_desc;
    return dispatch<true>(tri, any, bvh, a, cfg, stream, false, nullptr);
}

} // namespace dev
} // namespace rt
