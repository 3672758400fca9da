// Host-callable launchers of the gfx950 kernels (implemented in the .hip
// files of this directory).  All launches are asynchronous on `stream`.
#pragma once

#include <hip/hip_runtime.h>

#include "../core/layout.h"

#include <cstdint>

namespace rt {
namespace dev {

// Private-memory stack of the plain kernels; the host refuses BVHs deeper
// than kPlainStackCap - 1 inner levels for them.
constexpr int kPlainStackCap = 64;

// Persistent kernels, stack configuration 0: per-lane stack of kRegStack
// entries in VGPRs backed by a ring of kLdsStack entries per lane in LDS and,
// for deeper traversals, a per-lane column of a global spill buffer.
constexpr int kRegStack = 4;
constexpr int kLdsStack = 28;
constexpr int kPersistentWavesPerBlock = 4;
constexpr int kStackConfigs = 5;
// Leaves a closest-hit lane of the speculative persistent tracers parks
// before it stops descending: registers for at most 4; persistent_cfg::park
// picks 1..4 at run time (default 4).
constexpr int kParkClosest = 4;
// Work counters of the persistent kernels: one per chunk, each in its own
// 256-byte line; kCounterBytes of device memory hold all of them.
constexpr uint32_t kMaxChunks = 8;
constexpr uint32_t kCounterStride = 64;
constexpr size_t kCounterBytes = kMaxChunks * kCounterStride * 4;

// Stack entries the persistent kernels hold in VGPRs and LDS for a BVH of
// the given format and stack configuration (0: 4 VGPR + 28 LDS entries,
// 1: 8 + 16, 2: 2 + 20, 3: 0 + 24, 4: 0 + 20 with registers capped for 8
// waves per SIMD on single-scene binary trees).
int persistent_stack_cap(int bvh, int stack = 0);

// Spill entries per lane a persistent launch needs for a tree whose
// traversal can hold `need` stack entries (0 when VGPRs and LDS hold them),
// and the bytes of spill buffer a grid of `blocks` needs for that many.
uint32_t persistent_spill_entries(int bvh, int stack, int need);
size_t persistent_spill_bytes(int blocks, uint32_t entries);

// One entry of a multi-scene batch: a scene and its own ray and hit buffers;
// its rays occupy [first, first + n) of the launch's index space.
struct batch_desc {
    scene_desc scene;
    ray_soa rays;
    hit_soa hits;
    uint8_t *occluded;
    uint32_t first, n;
};

// With `stats` non-null the counting variant runs and adds the number of inner
// nodes fetched to stats[0] and of triangle records tested to stats[1].
hipError_t launch_trace_plain(const scene_desc &sc, const ray_soa &rays, const hit_soa &hits, uint8_t *occluded,
                              uint32_t n, bool any, hipStream_t stream, unsigned long long *stats = nullptr);

// Rays of a single-scene persistent launch read from records of `stride`
// float4 (at least two) that begin (ox, oy, oz, tmin | dx, dy, dz, tmax)
// instead of planes: ray idx is record perm[idx] (record idx when perm is
// null); hits are still written at idx.  With fixed_t set every ray spans
// [kRayEps, FLT_MAX] and the records' tmin/tmax words are not read.
struct ray_gather {
    const float4 *aos = nullptr;
    const uint32_t *perm = nullptr;
    uint32_t stride = 2;
    uint32_t fixed_t = 0;
};

// Tunables of the persistent kernel.
struct persistent_cfg {
    int speculative;      // leaf postponement
    int refill;           // fetch new rays when active lanes drop below refill_thr
    int refill_thr;       // active-lane threshold for refill (out of 64)
    int blocks;           // grid size; 0 = occupancy-derived default
    unsigned long long *stats = nullptr;   // non-null: counting variant (inner nodes, triangle tests)
    int stack = 0;        // stack configuration 0..kStackConfigs-1, see persistent_stack_cap
    int chunks = 8;       // 1: one work counter; 8: one contiguous chunk of rays per XCD
    int park = kParkClosest;   // parked leaves per closest-hit lane of the speculative tracers (1..4)
    // Global spill buffer: spill_cap entries per lane of the grid
    // (persistent_spill_bytes); spill_cap 0 when the tree fits VGPRs + LDS.
    int32_t *spill = nullptr;
    uint32_t spill_cap = 0;
    size_t spill_bytes = 0;
};

// `counter` points to kCounterBytes of device memory (cleared by the launch);
// `n` is the number of rays, read from device memory when count_ptr is
// non-null (so a captured graph can trace a batch whose size an earlier
// kernel produced).  The launch fails with hipErrorInvalidValue when
// cfg.spill_cap is set and the spill buffer is smaller than the grid needs.
hipError_t launch_trace_persistent(const scene_desc &sc, const ray_soa &rays, const hit_soa &hits,
                                   uint8_t *occluded, uint32_t n, const uint32_t *count_ptr, uint32_t *counter,
                                   bool any, const persistent_cfg &cfg, hipStream_t stream,
                                   ray_gather gather = ray_gather{});

// One persistent launch over a descriptor table (device memory, sorted by
// `first`, first entry at 0, `total` rays in all).  All scenes must share the
// triangle and BVH formats.
hipError_t launch_trace_batch(const batch_desc *d_table, uint32_t n_desc, uint32_t total, int tri, int bvh, bool any,
                              uint32_t *counter, const persistent_cfg &cfg, hipStream_t stream);

// Grid size (blocks) of the persistent kernel instantiation for these
// settings (`multi`: the descriptor-table variant; only the scene's
// triangle and BVH formats are read): cfg.blocks when positive, else blocks
// resident per CU times the CU count.
hipError_t persistent_grid(const scene_desc &sc, bool any, const persistent_cfg &cfg, int *blocks,
                           bool multi = false);

// Fills `n` uint32 words with `value` (used instead of hipMemsetAsync so the
// clearing can be captured into graphs).
hipError_t launch_fill_u32(uint32_t *p, uint32_t value, uint32_t n, hipStream_t stream);

} // namespace dev
} // namespace rt
