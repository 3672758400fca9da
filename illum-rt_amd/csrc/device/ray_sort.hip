// Stable LSD radix sort of 32-bit keys with their slot index, eight key bits
// per pass, for the path tracer's bounce queues.
//
// Per pass, over tiles of kSortTile items (256 threads x kSortTile/256 items):
//   k_radix_count   digit histogram of each tile (LDS), written digit-major
//                   (counts[digit * n_tiles + tile]);
//   k_radix_rows    one block per digit: exclusive scan of its row over the
//                   tiles, in place, and the row total;
//   k_radix_scatter each tile scans the 256 row totals itself (a digit's
//                   queue base), adds its row prefix, and writes every item
//                   to base + items of the same digit before it in the tile.
//                   The in-tile rank keeps input order: sub-tiles of 256 in
//                   order, waves in order within a sub-tile, lanes in order
//                   within a wave (same-digit lanes found with eight
//                   ballots), so each pass is stable and the passes compose
//                   into a sort by all the bits.
// The item count is read from device memory, so a captured graph sorts
// queues whose length an earlier kernel produced; tiles past it write zero
// counts and scatter nothing.  No memset, no host synchronization: every
// launch is capturable.
#include "ray_sort.h"

namespace rt {
namespace dev {

namespace {

constexpr int kSortThreads = 256;
constexpr int kSortPer = kSortTile / kSortThreads;   // items per thread (sub-tiles)
constexpr int kDigits = 256;
constexpr int kRowThreads = 1024;
static_assert(kSortTile % kSortThreads == 0, "tile shape");

__global__ __launch_bounds__(kSortThreads) void k_radix_count(const uint32_t *keys, const uint32_t *count_ptr,
                                                              uint32_t n_max, uint32_t shift, uint32_t *counts,
                                                              uint32_t n_tiles) {
    __shared__ uint32_t hist[kDigits];
    const uint32_t n = min(*count_ptr, n_max);
    const uint32_t tile = blockIdx.x, t = threadIdx.x;
    hist[t] = 0u;
    __syncthreads();
    const uint64_t first = (uint64_t)tile * kSortTile;
    if (first < n) {
        for (int s = 0; s < kSortPer; ++s) {
            uint64_t i = first + (uint64_t)s * kSortThreads + t;
            if (i < n) atomicAdd(&hist[(keys[i] >> shift) & 255u], 1u);
        }
    }
    __syncthreads();
    counts[(size_t)t * n_tiles + tile] = hist[t];
}

// Inclusive scan over one block's values in `buf` (blockDim.x of them).
template <int N>
__device__ __forceinline__ void block_scan_inclusive(uint32_t *buf, uint32_t t) {
    for (uint32_t off = 1; off < (uint32_t)N; off <<= 1) {
        uint32_t v = t >= off ? buf[t - off] : 0u;
        __syncthreads();
        buf[t] += v;
        __syncthreads();
    }
}

__global__ __launch_bounds__(kRowThreads) void k_radix_rows(uint32_t *counts, uint32_t n_tiles, uint32_t *totals) {
    __shared__ uint32_t buf[kRowThreads];
    const uint32_t d = blockIdx.x, t = threadIdx.x;
    uint32_t *row = counts + (size_t)d * n_tiles;
    uint32_t carry = 0;
    for (uint32_t base = 0; base < n_tiles; base += kRowThreads) {
        uint32_t j = base + t;
        uint32_t v = j < n_tiles ? row[j] : 0u;
        buf[t] = v;
        __syncthreads();
        block_scan_inclusive<kRowThreads>(buf, t);
        if (j < n_tiles) row[j] = carry + buf[t] - v;
        uint32_t chunk = buf[kRowThreads - 1];
        __syncthreads();   // buf is rewritten by the next chunk
        carry += chunk;
    }
    if (t == 0) totals[d] = carry;
}

__global__ __launch_bounds__(kSortThreads) void k_radix_scatter(const uint32_t *keys_in, const uint32_t *vals_in,
                                                                const uint32_t *count_ptr, uint32_t n_max,
                                                                uint32_t shift, const uint32_t *counts,
                                                                const uint32_t *totals, uint32_t n_tiles,
                                                                uint32_t *keys_out, uint32_t *vals_out) {
    constexpr int kWaves = kSortThreads / 64;
    __shared__ uint32_t base[kDigits];           // queue slot of the tile's first item of each digit
    __shared__ uint32_t run[kDigits];            // items of each digit in earlier sub-tiles
    __shared__ uint32_t wcnt[kWaves][kDigits];   // items of each digit per wave in this sub-tile
    const uint32_t n = min(*count_ptr, n_max);
    const uint32_t tile = blockIdx.x, t = threadIdx.x;
    const uint64_t first = (uint64_t)tile * kSortTile;
    if (first >= n) return;   // uniform: the whole tile is past the queue
    base[t] = totals[t];
    run[t] = 0u;
    for (int w = 0; w < kWaves; ++w) wcnt[w][t] = 0u;
    __syncthreads();
    block_scan_inclusive<kDigits>(base, t);
    uint32_t excl = base[t] - totals[t];
    __syncthreads();
    base[t] = excl + counts[(size_t)t * n_tiles + tile];
    __syncthreads();
    const uint32_t lane = t & 63u, wave = t >> 6;
    const uint64_t below = (1ull << lane) - 1ull;
    for (int s = 0; s < kSortPer; ++s) {
        uint64_t i = first + (uint64_t)s * kSortThreads + t;
        bool valid = i < n;
        uint32_t key = valid ? keys_in[i] : 0u;
        uint32_t val = valid ? (vals_in ? vals_in[i] : (uint32_t)i) : 0u;
        uint32_t d = (key >> shift) & 255u;
        uint64_t same = __ballot(valid);
        for (int b = 0; b < 8; ++b) {
            bool bit = ((d >> b) & 1u) != 0u;
            uint64_t ones = __ballot(valid && bit);
            same &= bit ? ones : ~ones;
        }
        uint32_t rank = (uint32_t)__popcll(same & below);
        if (valid && rank == 0) wcnt[wave][d] = (uint32_t)__popcll(same);
        __syncthreads();
        if (valid) {
            uint32_t pos = base[d] + run[d] + rank;
            for (uint32_t w = 0; w < wave; ++w) pos += wcnt[w][d];
            if (pos < n) {   // always true when the counts match the keys
                if (keys_out) keys_out[pos] = key;   // the last pass writes only the permutation
                vals_out[pos] = val;
            }
        }
        __syncthreads();
        uint32_t add = 0;
        for (int w = 0; w < kWaves; ++w) {
            add += wcnt[w][t];
            wcnt[w][t] = 0u;
        }
        run[t] += add;
        __syncthreads();
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
        hipLaunchKernelGGL(k_radix_count, dim3(tiles), dim3(kSortThreads), 0, s, kin, count_ptr, n_max, shift,
                           b.counts, tiles);
        hipError_t e = hipGetLastError();
        if (e != hipSuccess) return e;
        hipLaunchKernelGGL(k_radix_rows, dim3(kDigits), dim3(kRowThreads), 0, s, b.counts, tiles, b.totals);
        e = hipGetLastError();
        if (e != hipSuccess) return e;
        hipLaunchKernelGGL(k_radix_scatter, dim3(tiles), dim3(kSortThreads), 0, s, kin, vin, count_ptr, n_max,
                           shift, b.counts, b.totals, tiles, kout, vout);
        e = hipGetLastError();
        if (e != hipSuccess) return e;
    }
    return hipSuccess;
}

} // namespace dev
} // namespace rt
