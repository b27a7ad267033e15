// radix_sort.hpp -- stable LSD radix sort of (u32 key, u32 value) pairs.
#pragma once

#include "common.hpp"

namespace wsim {

// Histogram scratch (u32 elements) needed to sort n pairs.
size_t radix_hist_elems(size_t n);

// Sorts n pairs held in (k0, v0) by the low `bits` bits of the keys, stably
// (equal keys keep their input order), ping-ponging through (k1, v1).
// 8-bit digits, one upsweep / scan / LDS downsweep triple per digit.
// Returns 0 if the sorted pairs end in (k0, v0), 1 if in (k1, v1).
int radix_sort_pairs(u32* k0, u32* v0, u32* k1, u32* v1, size_t n, int bits, u32* hist,
                     hipStream_t stream);

// Number of key bits needed to represent values in [0, max_key].
inline int key_bits(u32 max_key) {
    int b = 0;
    while (b < 32 && (max_key >> b) != 0) ++b;
    return b;
}

}  // namespace wsim
