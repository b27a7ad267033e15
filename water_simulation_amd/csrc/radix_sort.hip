// radix_sort.hip -- stable LSD radix sort of (u32 key, u32 value) pairs for
// gfx950, 8-bit digits.
//
// Replaces cub::DeviceRadixSort::SortPairs at core/src/cell_grid.cu:98-110
// and core/src/spatial_accel.cu:163-171.  With values = iota the output
// permutation is the lexicographic (key, original index) order, i.e. exactly
// the permutation a stable radix sort produces, bit for bit.
//
// Per digit pass (reduce-then-scan):
//   rs_upsweep   -- per tile digit histogram, one LDS atomic per key (counts
//                   do not depend on the order the atomics land in);
//   rs_scan_rows -- one workgroup per digit scans that digit's per-tile counts
//                   and records the digit total;
//   rs_downsweep -- each wave ranks its keys stably (ballot match + per-wave
//                   LDS counters), the tile is reordered in LDS, and written
//                   out digit-run by digit-run so consecutive lanes store to
//                   consecutive addresses.
// The sweeps run 1024-thread workgroups over 2048-key tiles: at a million
// keys a pass is bound by launch and latency, not bytes, and the wide tiles
// halve the per-tile histogram matrix and its strided accesses (against
// 1024-key tiles) while keeping each lane's serial ranking work short.
#include "device_math.hpp"
#include "radix_sort.hpp"

namespace wsim {
namespace {

constexpr int RS_THREADS = kRadixThreads;           // sweep workgroup (1024)
constexpr int RS_WAVES = RS_THREADS / kWave;
constexpr int RS_ITEMS = 2;
constexpr int RS_WAVE_TILE = kWave * RS_ITEMS;      // 128 keys per wave
constexpr int RS_TILE = RS_THREADS * RS_ITEMS;      // 2048 keys per workgroup
static_assert(RS_TILE == kRadixTile, "tile of the header's counting helpers");
constexpr int RS_DIGITS = 256;                      // threads of the digit scans
constexpr int RS_ROW_PER = 4;  // tiles per thread and round of the row scan
static_assert(RS_THREADS >= RS_DIGITS, "threads [0, 256) own one digit each");

// Lanes of the wave whose digit equals this lane's digit (nbits <= 8).
__device__ inline u64 match_digit(u32 d, int nbits) {
    u64 m = ~0ull;
    for (int b = 0; b < nbits; ++b) {
        const bool bit = (d >> b) & 1u;
        const u64 bal = __ballot(bit);
        m &= bit ? bal : ~bal;
    }
    return m;
}

// Exclusive scan of one u32 per thread over threads [0, 256) of the
// workgroup.  Every thread of the workgroup calls it (it holds two
// barriers); threads from 256 on pass anything and ignore the result.
// s_tmp holds 4 entries.
__device__ inline u32 block_exclusive_scan256(u32 v, u32* s_tmp) {
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    constexpr int kW = RS_DIGITS / kWave;
    u32 x = v;
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) {
        const u32 y = __shfl_up(x, o);
        if (lane >= o) x += y;
    }
    if (lane == 63 && w < kW) s_tmp[w] = x;
    __syncthreads();
    u32 off = 0;
    for (int i = 0; i < w && i < kW; ++i) off += s_tmp[i];
    __syncthreads();
    return off + x - v;
}

__global__ __launch_bounds__(RS_THREADS) void rs_upsweep(const u32* __restrict__ keys, u32 n,
                                                         int shift, int nbits,
                                                         u32* __restrict__ hist, u32 nblocks) {
    __shared__ u32 cnt[RS_DIGITS];
    const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
    const u32 dmask = (1u << nbits) - 1u;
    radix_count_begin(cnt);
    const u32 base = blockIdx.x * RS_TILE + w * RS_WAVE_TILE;
#pragma unroll
    for (int r = 0; r < RS_ITEMS; ++r) {
        const u32 idx = base + r * kWave + lane;
        if (idx < n) radix_count_key(cnt, (keys[idx] >> shift) & dmask);
    }
    radix_count_end(cnt, dmask, hist, nblocks);
}

// Row scan of the digit-major histogram matrix: workgroup d turns row d (the
// per-tile counts of digit d) into row-local exclusive offsets, in place, and
// writes the digit's total.  Each thread owns up to RS_ROW_PER consecutive
// tiles of the row; rows longer than RS_THREADS * RS_ROW_PER loop with a
// carry.  The downsweep adds the exclusive scan of the digit totals.
__global__ __launch_bounds__(RS_DIGITS) void rs_scan_rows(u32* __restrict__ hist, u32 nblocks,
                                                          u32* __restrict__ totals) {
    __shared__ u32 s_tmp[RS_DIGITS / kWave];
    __shared__ u32 s_carry;
    u32* row = hist + static_cast<u64>(blockIdx.x) * nblocks;
    const int tid = threadIdx.x;
    if (tid == 0) s_carry = 0;
    __syncthreads();
    for (u32 base = 0; base < nblocks; base += RS_DIGITS * RS_ROW_PER) {
        const u32 first = base + static_cast<u32>(tid) * RS_ROW_PER;
        u32 v[RS_ROW_PER];
        u32 t = 0;
#pragma unroll
        for (int k = 0; k < RS_ROW_PER; ++k) {
            v[k] = first + k < nblocks ? row[first + k] : 0u;
            t += v[k];
        }
        u32 run = s_carry + block_exclusive_scan256(t, s_tmp);
#pragma unroll
        for (int k = 0; k < RS_ROW_PER; ++k) {
            if (first + k < nblocks) row[first + k] = run;
            run += v[k];
        }
        __syncthreads();
        if (tid == RS_DIGITS - 1) s_carry = run;
        __syncthreads();
    }
    if (tid == 0) totals[blockIdx.x] = s_carry;
}

__global__ __launch_bounds__(RS_THREADS) void rs_downsweep(
    const u32* __restrict__ kin, const u32* __restrict__ vin, u32* __restrict__ kout,
    u32* __restrict__ vout, u32 n, int shift, int nbits, const u32* __restrict__ offs,
    u32 nblocks, const u32* __restrict__ totals) {
    __shared__ u32 s_wc[RS_WAVES][RS_DIGITS];
    __shared__ u32 s_start[RS_DIGITS];
    __shared__ u32 s_goff[RS_DIGITS];
    __shared__ u32 s_tmp[RS_DIGITS / kWave];
    __shared__ u32 s_k[RS_TILE];
    __shared__ u32 s_v[RS_TILE];

    const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
    const u32 dmask = (1u << nbits) - 1u;
    // this tile's offset in each digit's run and the digit totals (threads
    // [0, 256), one digit each), loaded with the keys
    const bool digit_thread = static_cast<u32>(tid) <= dmask;
    const u32 goff = digit_thread ? offs[static_cast<u64>(tid) * nblocks + blockIdx.x] : 0u;
    const u32 dtot = digit_thread ? totals[tid] : 0u;
    for (int i = tid; i < RS_WAVES * RS_DIGITS; i += RS_THREADS) (&s_wc[0][0])[i] = 0;
    __syncthreads();

    const u32 tile0 = blockIdx.x * RS_TILE;
    const u32 base = tile0 + w * RS_WAVE_TILE;
    const u64 lanes_below = (1ull << lane) - 1ull;

    u32 key[RS_ITEMS], val[RS_ITEMS], rank[RS_ITEMS];
#pragma unroll
    for (int r = 0; r < RS_ITEMS; ++r) {
        const u32 idx = base + r * kWave + lane;
        const bool valid = idx < n;
        key[r] = valid ? kin[idx] : 0u;
        val[r] = valid ? vin[idx] : 0u;
    }
    // Stable ranks inside the wave's sub-tile: positions are ordered by
    // (round, lane), so a key's rank among equal digits is the running count
    // of its digit in this wave plus the number of lower lanes with it now.
#pragma unroll
    for (int r = 0; r < RS_ITEMS; ++r) {
        const u32 idx = base + r * kWave + lane;
        const bool valid = idx < n;
        const u32 d = (key[r] >> shift) & dmask;
        const u64 m = match_digit(d, nbits) & __ballot(valid);
        const u32 below = static_cast<u32>(__popcll(m & lanes_below));
        u32 cur = 0;
        if (valid) cur = s_wc[w][d];
        __builtin_amdgcn_wave_barrier();
        if (valid && below == 0) s_wc[w][d] = cur + static_cast<u32>(__popcll(m));
        __builtin_amdgcn_wave_barrier();
        rank[r] = cur + below;
    }
    __syncthreads();

    // Per digit (threads [0, 256), one digit each): exclusive prefix over
    // waves, tile total, this tile's offset in the digit's run.
    u32 total = 0;
    if (tid < RS_DIGITS) {
        const int d = tid;
#pragma unroll
        for (int ww = 0; ww < RS_WAVES; ++ww) {
            const u32 c = s_wc[ww][d];
            s_wc[ww][d] = total;
            total += c;
        }
        s_goff[d] = goff;
    }
    // digit base: scan of the digit totals
    const u32 dbase = block_exclusive_scan256(dtot, s_tmp);
    const u32 start = block_exclusive_scan256(total, s_tmp);
    if (tid < RS_DIGITS) {
        s_goff[tid] += dbase;
        s_start[tid] = start;
    }
    __syncthreads();

#pragma unroll
    for (int r = 0; r < RS_ITEMS; ++r) {
        const u32 idx = base + r * kWave + lane;
        if (idx < n) {
            const u32 d = (key[r] >> shift) & dmask;
            const u32 pos = s_start[d] + s_wc[w][d] + rank[r];
            s_k[pos] = key[r];
            s_v[pos] = val[r];
        }
    }
    __syncthreads();

    const u32 cnt = (n - tile0 < static_cast<u32>(RS_TILE)) ? n - tile0 : static_cast<u32>(RS_TILE);
    for (u32 e = tid; e < cnt; e += RS_THREADS) {
        const u32 k = s_k[e];
        const u32 d = (k >> shift) & dmask;
        const u32 o = s_goff[d] + (e - s_start[d]);
        kout[o] = k;
        vout[o] = s_v[e];
    }
}

}  // namespace

// Per-tile digit counts plus the 256 digit totals.
size_t radix_hist_elems(size_t n) {
    return static_cast<size_t>(RS_DIGITS) * div_up(n ? n : 1, RS_TILE) + RS_DIGITS;
}

int radix_sort_pairs(u32* k0, u32* v0, u32* k1, u32* v1, size_t n, int bits, u32* hist,
                     hipStream_t stream, bool first_counted) {
    if (n <= 1 || bits <= 0) return 0;
    if (n > 0xFFFFF000ull) fail(kInvalidArgument, "radix_sort_pairs: too many elements");
    const u32 nn = static_cast<u32>(n);
    const u32 nblocks = div_up(n, RS_TILE);
    u32* kin = k0;
    u32* vin = v0;
    u32* kout = k1;
    u32* vout = v1;
    int where = 0;
    for (int shift = 0; shift < bits; shift += 8) {
        const int nbits = (bits - shift) < 8 ? (bits - shift) : 8;
        if (shift > 0 || !first_counted) {
            rs_upsweep<<<nblocks, RS_THREADS, 0, stream>>>(kin, nn, shift, nbits, hist, nblocks);
            WSIM_LAUNCH_CHECK();
        }
        u32* totals = hist + static_cast<size_t>(RS_DIGITS) * nblocks;
        rs_scan_rows<<<1u << nbits, RS_DIGITS, 0, stream>>>(hist, nblocks, totals);
        WSIM_LAUNCH_CHECK();
        rs_downsweep<<<nblocks, RS_THREADS, 0, stream>>>(kin, vin, kout, vout, nn, shift, nbits,
                                                         hist, nblocks, totals);
        WSIM_LAUNCH_CHECK();
        u32* tk = kin;
        u32* tv = vin;
        kin = kout;
        vin = vout;
        kout = tk;
        vout = tv;
        where ^= 1;
    }
    return where;
}

}  // namespace wsim
