// Stable radix sort of 32-bit keys with their slot index (device/ray_sort.hip).
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

namespace rt {
namespace dev {

// Items per block and pass (16 per thread); on MI355X 4096 gave 0.5% faster
// 64-sample path-tracing frames than 2048 and 1.3% faster than 1024.
constexpr uint32_t kSortTile = 4096;

// Device buffers of one sort: keys_in, keys_tmp, vals_tmp and vals_out hold
// n_max words each, counts 256 * radix_tiles(n_max), totals 256.  keys_in is
// overwritten; vals_out receives, for sorted position j < *count_ptr, the
// slot of keys_in that sorts there (ties keep slot order).  With
// `sorted_keys` the sorted keys end in keys_tmp after an odd number of passes
// (ceil(bits / 8)), in keys_in after an even one; without, the last pass
// writes the permutation only.
struct radix_sort_buffers {
    uint32_t *keys_in, *keys_tmp, *vals_tmp, *vals_out, *counts, *totals;
};

uint32_t radix_tiles(uint32_t n_max);

// Sorts the first *count_ptr (<= n_max) keys by their low `bits` bits
// (1..32, eight per pass).  Asynchronous on `s`, capturable into a graph.
hipError_t launch_radix_sort(const radix_sort_buffers &b, uint32_t n_max, const uint32_t *count_ptr, int bits,
                             hipStream_t s, bool sorted_keys = false);

} // namespace dev
} // namespace rt
