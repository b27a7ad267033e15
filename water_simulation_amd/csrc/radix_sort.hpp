// radix_sort.hpp -- stable LSD radix sort of (u32 key, u32 value) pairs.
#pragma once

#include "common.hpp"

namespace wsim {

// Histogram scratch (u32 elements) needed to sort n pairs.
size_t radix_hist_elems(size_t n);

// Sweep shape: workgroups of kRadixThreads threads over tiles of kRadixTile
// consecutive keys.  The per-tile digit counts are digit-major:
// hist[d * radix_tiles(n) + t] counts digit d in tile t.
constexpr int kRadixThreads = 1024;
constexpr int kRadixTile = 2048;
inline u32 radix_tiles(size_t n) { return div_up(n ? n : 1, kRadixTile); }
// Mask of the first (lowest) digit of a `bits`-bit key.
inline u32 radix_first_digit_mask(int bits) { return (1u << (bits < 8 ? bits : 8)) - 1u; }

// Sorts n pairs held in (k0, v0) by the low `bits` bits of the keys, stably
// (equal keys keep their input order), ping-ponging through (k1, v1).
// 8-bit digits, one upsweep / scan / LDS downsweep triple per digit.  With
// first_counted the caller's kernel that wrote the keys has already written
// the first digit's counts into hist (radix_count_* below) and the first
// upsweep is skipped.  Returns 0 if the sorted pairs end in (k0, v0), 1 if
// in (k1, v1).
int radix_sort_pairs(u32* k0, u32* v0, u32* k1, u32* v1, size_t n, int bits, u32* hist,
                     hipStream_t stream, bool first_counted = false);

// Digit counting of one tile, for the sort's upsweep and for a kernel that
// writes the keys it sorts: a workgroup of kRadixThreads threads whose
// blockIdx.x is the tile index calls begin, counts each key of its tile
// (indices [t * kRadixTile, (t + 1) * kRadixTile) in any order among the
// workgroup's threads) once, then calls end.  cnt is 256 words of LDS.
// One LDS atomic per key: the counts do not depend on the order they land in.
__device__ inline void radix_count_begin(u32* cnt) {
    if (threadIdx.x < 256) cnt[threadIdx.x] = 0;
    __syncthreads();
}
__device__ inline void radix_count_key(u32* cnt, u32 digit) { atomicAdd(&cnt[digit], 1u); }
__device__ inline void radix_count_end(const u32* cnt, u32 dmask, u32* hist, u32 tiles) {
    __syncthreads();
    if (threadIdx.x <= dmask)
        hist[static_cast<u64>(threadIdx.x) * tiles + blockIdx.x] = cnt[threadIdx.x];
}

// Number of key bits needed to represent values in [0, max_key].
inline int key_bits(u32 max_key) {
    int b = 0;
    while (b < 32 && (max_key >> b) != 0) ++b;
    return b;
}

}  // namespace wsim
