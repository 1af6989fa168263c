// Stable LSD radix sort of (32-bit key, 32-bit value) pairs in device memory (8-bit digits).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace sesdf
{
    constexpr uint32_t SortTile = 2048; // items per workgroup and digit pass

    // Words of the histogram workspace for n items: one count per digit and tile, and the digit totals.
    constexpr size_t sort_hist_words( uint64_t n ) { return 256 * ( ( n + SortTile - 1 ) / SortTile ) + 256; }

    // Sorts n pairs by the low `bits` bits of their keys (stable: equal keys keep their order).
    // (k0, v0) hold the input; (k1, v1) are as large. Returns 0 when the sorted pairs end in
    // (k0, v0), 1 when in (k1, v1). hist holds sort_hist_words(n) words; it needs no clearing.
    int radix_sort_pairs( uint32_t * k0, uint32_t * v0, uint32_t * k1, uint32_t * v1, uint32_t * hist, uint32_t n, int bits,
                          hipStream_t stream );
} // namespace sesdf
