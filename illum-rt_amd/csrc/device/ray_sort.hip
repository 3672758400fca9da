// Stable LSD radix sort of 32-bit keys with their slot index, eight key bits
// per pass, for the path tracer's bounce queues and the adjoint's triangle
// keys.
//
// Per pass, over tiles of kSortTile consecutive items:
//   k_radix_count   digit histogram of each tile (LDS; one atomic for a wave
//                   whose items share a digit), written digit-major
//                   (counts[digit * n_tiles + tile]);
//   k_radix_rows    one block per digit: exclusive scan of its row over the
//                   tiles, in place (four consecutive tiles per thread, wave
//                   scans, one barrier per 4096 tiles), and the row total;
//   k_radix_scatter each wave of a tile ranks its own contiguous quarter of
//                   the tile: 64 items at a time, lanes of the same digit
//                   found with eight ballots (one when all share a digit),
//                   each digit's running count in the wave's LDS row (no
//                   block barrier between rounds).
//                   The waves' counts then give every item its position in
//                   the tile sorted by digit (digit, then input order); the
//                   tile is staged there in LDS and written out in that
//                   order, so consecutive threads store consecutive slots of
//                   a digit's run instead of one scattered word per item.
// Input order decides ties (waves' quarters in order, rounds in order, lanes
// in order), so every pass is stable and the passes compose into a sort by
// all the bits.  Blocks are dealt to the 8 XCDs round robin; block b sorts
// tile about (b mod 8) * n_tiles / 8 + b / 8 (sort_tile), so one XCD holds a
// contiguous range of tiles and writes the adjacent slots of each digit's
// run through its own L2.  The item count is read from device memory, so a
// captured graph sorts queues whose length an earlier kernel produced; tiles
// past it write zero counts and scatter nothing.  No memset, no host
// synchronization: every launch is capturable.
#include "ray_sort.h"

namespace rt {
namespace dev {

namespace {

constexpr int kSortThreads = 256;
constexpr int kSortWaves = kSortThreads / 64;
constexpr int kSortRounds = kSortTile / kSortThreads;   // 64-item rounds in one wave's quarter of a tile
constexpr int kDigits = 256;
constexpr int kRowThreads = 1024;
static_assert(kSortTile % kSortThreads == 0, "tile shape");
static_assert(kDigits == kSortThreads, "one thread per digit");

// Tile of block b: XCD b mod 8 takes the b-th eighth of the tiles (the first
// n_tiles mod 8 XCDs one tile more), in block order.  A bijection on
// [0, n_tiles) for any n_tiles.
__device__ __forceinline__ uint32_t sort_tile(uint32_t b, uint32_t n_tiles) {
    const uint32_t x = b & 7u, k = b >> 3, per = n_tiles >> 3, extra = n_tiles & 7u;
    return x * per + min(x, extra) + k;
}

__device__ __forceinline__ uint32_t wave_scan_inclusive(uint32_t v, uint32_t lane) {
#pragma unroll
    for (uint32_t off = 1; off < 64; off <<= 1) {
        const uint32_t u = __shfl_up(v, off, 64);
        if (lane >= off) v += u;
    }
    return v;
}

// Lanes of the wave whose (valid) digit equals this lane's.
__device__ __forceinline__ uint64_t same_digit_lanes(uint32_t d, bool valid) {
    uint64_t same = __ballot(valid);
#pragma unroll
    for (int b = 0; b < 8; ++b) {
        const bool bit = ((d >> b) & 1u) != 0u;
        const uint64_t ones = __ballot(valid && bit);
        same &= bit ? ones : ~ones;
    }
    return same;
}

__global__ __launch_bounds__(kSortThreads) void k_radix_count(const uint32_t *keys, const uint32_t *count_ptr,
                                                              uint32_t n_max, uint32_t shift, uint32_t dmask,
                                                              uint32_t *counts, uint32_t n_tiles) {
    __shared__ uint32_t hist[kDigits];
    const uint32_t n = min(*count_ptr, n_max);
    const uint32_t tile = sort_tile(blockIdx.x, n_tiles), t = threadIdx.x;
    hist[t] = 0u;
    __syncthreads();
    const uint64_t first = (uint64_t)tile * kSortTile;
    if (first < n) {   // every load in flight before the first atomic
        uint32_t k[kSortRounds];
#pragma unroll
        for (int s = 0; s < kSortRounds; ++s) {
            const uint64_t i = first + (uint64_t)s * kSortThreads + t;
            k[s] = i < n ? keys[i] : 0u;
        }
        // a wave whose valid lanes share one digit (common in the later
        // passes, once lower digits have grouped the keys) adds once instead
        // of 64 times to the same LDS word
#pragma unroll
        for (int s = 0; s < kSortRounds; ++s) {
            const bool valid = first + (uint64_t)s * kSortThreads + t < n;
            const uint32_t d = (k[s] >> shift) & dmask;
            const uint32_t d0 = __builtin_amdgcn_readfirstlane(d);
            const uint64_t vm = __ballot(valid), same = __ballot(valid && d == d0);
            if (same == vm) {
                if ((t & 63u) == 0u && vm) atomicAdd(&hist[d0], (uint32_t)__popcll(vm));
            } else if (valid) {
                atomicAdd(&hist[d], 1u);
            }
        }
    }
    __syncthreads();
    counts[(size_t)t * n_tiles + tile] = hist[t];
}

__global__ __launch_bounds__(kRowThreads) void k_radix_rows(uint32_t *counts, uint32_t n_tiles, uint32_t *totals) {
    constexpr int kRowWaves = kRowThreads / 64;
    constexpr uint32_t kPer = 4;                 // consecutive tiles per thread: four loads in flight
    constexpr uint32_t kChunk = kRowThreads * kPer;
    __shared__ uint32_t wsum[2][kRowWaves];   // wave totals, double-buffered over chunks
    const uint32_t d = blockIdx.x, t = threadIdx.x, lane = t & 63u, wave = t >> 6;
    uint32_t *row = counts + (size_t)d * n_tiles;
    uint32_t carry = 0;
    int buf = 0;
    for (uint32_t base = 0; base < n_tiles; base += kChunk) {
        const uint32_t j0 = base + kPer * t;
        uint32_t v[kPer], mine = 0;
#pragma unroll
        for (uint32_t k = 0; k < kPer; ++k) v[k] = j0 + k < n_tiles ? row[j0 + k] : 0u;
#pragma unroll
        for (uint32_t k = 0; k < kPer; ++k) mine += v[k];
        const uint32_t incl = wave_scan_inclusive(mine, lane);
        if (lane == 63) wsum[buf][wave] = incl;
        __syncthreads();
        uint32_t before = 0, chunk = 0;
#pragma unroll
        for (int w = 0; w < kRowWaves; ++w) {
            const uint32_t s = wsum[buf][w];
            before += (uint32_t)w < wave ? s : 0u;
            chunk += s;
        }
        uint32_t run = carry + before + incl - mine;   // exclusive prefix of tile j0
#pragma unroll
        for (uint32_t k = 0; k < kPer; ++k) {
            if (j0 + k < n_tiles) row[j0 + k] = run;
            run += v[k];
        }
        carry += chunk;
        // the next chunk writes the other buffer; the one after it writes this
        // buffer only after the next chunk's barrier, which every wave passes
        // after its reads here
        buf ^= 1;
    }
    if (t == 0) totals[d] = carry;
}

__global__ __launch_bounds__(kSortThreads) void k_radix_scatter(const uint32_t *keys_in, const uint32_t *vals_in,
                                                                const uint32_t *count_ptr, uint32_t n_max,
                                                                uint32_t shift, uint32_t dmask,
                                                                const uint32_t *counts, const uint32_t *totals,
                                                                uint32_t n_tiles, uint32_t *keys_out,
                                                                uint32_t *vals_out) {
    constexpr uint32_t kQuarter = kSortTile / kSortWaves;
    __shared__ uint32_t slot[kDigits];              // queue slot of tile position 0 counted in each digit's run
    __shared__ uint32_t start[kDigits];             // tile position of each digit's first item
    __shared__ uint32_t wrun[kSortWaves][kDigits];  // items of each digit per wave, then the earlier waves' sum
    __shared__ uint32_t wsum[2][kSortWaves];
    __shared__ uint32_t skey[kSortTile], sval[kSortTile];
    const uint32_t n = min(*count_ptr, n_max);
    const uint32_t tile = sort_tile(blockIdx.x, n_tiles), t = threadIdx.x;
    const uint64_t first = (uint64_t)tile * kSortTile;
    if (first >= n) return;   // uniform: the whole tile is past the queue
    const uint32_t lane = t & 63u, wave = t >> 6;
#pragma unroll
    for (int w = 0; w < kSortWaves; ++w) wrun[w][t] = 0u;
    const uint64_t quarter = first + (uint64_t)wave * kQuarter;
    uint32_t key[kSortRounds], val[kSortRounds], loc[kSortRounds];
#pragma unroll
    for (int r = 0; r < kSortRounds; ++r) {
        const uint64_t i = quarter + (uint64_t)r * 64u + lane;
        const bool valid = i < n;
        key[r] = valid ? keys_in[i] : 0u;
        val[r] = valid ? (vals_in ? vals_in[i] : (uint32_t)i) : 0u;
    }
    __syncthreads();
    // rank within the wave's quarter: items of the same digit in earlier
    // rounds (the wave's LDS row, read before its leader adds this round's
    // count) plus same-digit lanes below this one
    const uint64_t below = (1ull << lane) - 1ull;
#pragma unroll
    for (int r = 0; r < kSortRounds; ++r) {
        const bool valid = quarter + (uint64_t)r * 64u + lane < n;
        const uint32_t d = (key[r] >> shift) & dmask;
        // one ballot settles a round whose valid lanes share a digit
        const uint32_t d0 = __builtin_amdgcn_readfirstlane(d);
        const uint64_t vm = __ballot(valid), one = __ballot(valid && d == d0);
        const uint64_t same = one == vm ? vm : same_digit_lanes(d, valid);
        const uint32_t rank = (uint32_t)__popcll(same & below);
        const uint32_t prior = valid ? wrun[wave][d] : 0u;
        loc[r] = prior + rank;
        if (valid && rank == 0) wrun[wave][d] = prior + (uint32_t)__popcll(same);
        __builtin_amdgcn_wave_barrier();
    }
    __syncthreads();
    // thread t = digit t: the tile's count of it, earlier waves' counts, its
    // first position in the sorted tile, its first slot in the sorted queue
    uint32_t c = 0;
#pragma unroll
    for (int w = 0; w < kSortWaves; ++w) {
        const uint32_t x = wrun[w][t];
        wrun[w][t] = c;
        c += x;
    }
    const uint32_t total = totals[t];
    const uint32_t incl = wave_scan_inclusive(c, lane);
    const uint32_t tincl = wave_scan_inclusive(total, lane);
    if (lane == 63) {
        wsum[0][wave] = incl;
        wsum[1][wave] = tincl;
    }
    __syncthreads();
    uint32_t pos_tile = incl - c, pos_queue = tincl - total;
#pragma unroll
    for (int w = 0; w < kSortWaves; ++w) {
        if ((uint32_t)w < wave) {
            pos_tile += wsum[0][w];
            pos_queue += wsum[1][w];
        }
    }
    start[t] = pos_tile;
    slot[t] = pos_queue + counts[(size_t)t * n_tiles + tile] - pos_tile;   // mod 2^32; + a position >= start
    __syncthreads();
#pragma unroll
    for (int r = 0; r < kSortRounds; ++r) {
        if (quarter + (uint64_t)r * 64u + lane < n) {
            const uint32_t d = (key[r] >> shift) & dmask;
            const uint32_t p = start[d] + wrun[wave][d] + loc[r];
            skey[p] = key[r];
            sval[p] = val[r];
        }
    }
    __syncthreads();
    const uint32_t tn = (uint32_t)min((uint64_t)kSortTile, (uint64_t)n - first);
#pragma unroll
    for (int r = 0; r < kSortRounds; ++r) {
        const uint32_t j = (uint32_t)r * kSortThreads + t;
        if (j < tn) {
            const uint32_t k = skey[j];
            const uint32_t pos = slot[(k >> shift) & dmask] + j;
            if (pos < n) {   // always true when the counts match the keys
                if (keys_out) keys_out[pos] = k;   // the last pass writes only the permutation
                vals_out[pos] = sval[j];
            }
        }
    }
}

} // namespace

uint32_t radix_tiles(uint32_t n_max) { return (uint32_t)(((uint64_t)n_max + kSortTile - 1) / kSortTile); }

hipError_t launch_radix_sort(const radix_sort_buffers &b, uint32_t n_max, const uint32_t *count_ptr, int bits,
                             hipStream_t s, bool sorted_keys) {
    const uint32_t tiles = radix_tiles(n_max);
    if (tiles == 0) return hipSuccess;
    const int passes = (bits + 7) / 8;
    if (passes < 1 || passes > 4) return hipErrorInvalidValue;
    // pass p reads (src keys, src values) and writes (dst keys, dst values);
    // the last pass writes b.vals_out (and the keys only when the caller
    // reads them); the first reads the slot index itself
    uint32_t *kbuf[2] = {b.keys_tmp, b.keys_in};
    for (int p = 0; p < passes; ++p) {
        const uint32_t *kin = p == 0 ? b.keys_in : kbuf[(p - 1) & 1];
        uint32_t *kout = (p + 1 < passes || sorted_keys) ? kbuf[p & 1] : nullptr;
        const uint32_t *vin = p == 0 ? nullptr : ((passes - p) & 1 ? b.vals_tmp : b.vals_out);
        uint32_t *vout = ((passes - 1 - p) & 1) ? b.vals_tmp : b.vals_out;
        const uint32_t shift = 8u * (uint32_t)p;
        const int left = bits - 8 * p;   // key bits from `shift` on that take part
        const uint32_t dmask = left >= 8 ? 255u : (1u << left) - 1u;
        hipLaunchKernelGGL(k_radix_count, dim3(tiles), dim3(kSortThreads), 0, s, kin, count_ptr, n_max, shift,
                           dmask, b.counts, tiles);
        hipError_t e = hipGetLastError();
        if (e != hipSuccess) return e;
        hipLaunchKernelGGL(k_radix_rows, dim3(kDigits), dim3(kRowThreads), 0, s, b.counts, tiles, b.totals);
        e = hipGetLastError();
        if (e != hipSuccess) return e;
        hipLaunchKernelGGL(k_radix_scatter, dim3(tiles), dim3(kSortThreads), 0, s, kin, vin, count_ptr, n_max,
                           shift, dmask, b.counts, b.totals, tiles, kout, vout);
        e = hipGetLastError();
        if (e != hipSuccess) return e;
    }
    return hipSuccess;
}

} // namespace dev
} // namespace rt
