// grid.hip -- uniform cell grid for the fixed-radius neighbour search.
//
// Replaces CellGrid::build (core/src/cell_grid.cu:73-138):
//   k_cell_keys     -- key = (cx*Dy + cy)*Dz + cz from the host-rounded
//                      inverse cell length, sentinel C outside the grid
//                      (cell_grid.cu:10-30); value = original index;
//   radix sort      -- stable LDS radix sort of (key, index) over
//                      ceil(log2(C+1)) bits (cell_grid.cu:98-110 sorts all 32);
//   k_cell_start    -- cell_start[c] = lower_bound(sorted keys, c) for every
//                      c in [0, C].  The reference fills with the sentinel,
//                      scatters run starts and then runs a single-thread
//                      right-to-left min over all C+1 cells
//                      (cell_grid.cu:119-137); here each workgroup owns a
//                      tile of 4096 cells, finds its key range with two
//                      64-ary wave searches, marks run starts in LDS,
//                      suffix-min scans the tile and stores it coalesced.
// Per-build allocations of the reference (two temporaries per call,
// cell_grid.cu:94-95) are replaced by buffers sized once by reserve().
//
// A segmented build (batched scenes) offsets each scene's keys by
// scene * (C + 1): one sort then orders the particles scene-major and
// cell-sorted within each scene, and one cell_start pass over
// scenes * (C + 1) + 1 entries serves every scene.
#include "radix_sort.hpp"
#include "wsim_internal.hpp"

namespace wsim {
namespace {

constexpr int CS_THREADS = 256;

__device__ inline u32 cs_pad(u32 i) { return i + (i >> 4); }  // odd stride: no bank conflicts

// First index i in [0, n] with keys[i] >= target (keys sorted ascending).
// All 64 lanes probe evenly spaced positions per round; the ballot of
// "key < target" is a prefix of lanes, which narrows [lo, hi] 64-fold.
__device__ inline u32 wave_lower_bound(const u32* __restrict__ keys, u32 n, u32 target) {
    const int lane = threadIdx.x & 63;
    u32 lo = 0, hi = n;
    while (lo < hi) {
        const u32 len = hi - lo;
        const u32 p = lo + static_cast<u32>((static_cast<u64>(len) * lane) >> 6);
        const bool pred = keys[p] < target;
        const int cnt = __popcll(__ballot(pred));
        const u32 p_before = __shfl(p, cnt > 0 ? cnt - 1 : 0);
        const u32 p_at = __shfl(p, cnt < 64 ? cnt : 63);
        const u32 nlo = cnt > 0 ? p_before + 1 : lo;
        const u32 nhi = cnt < 64 ? p_at : hi;
        lo = nlo;
        hi = nhi;
    }
    return lo;
}

__global__ __launch_bounds__(256) void k_cell_keys(const float* __restrict__ xyz, u32 n, GridGeom g,
                                                   u32* __restrict__ keys, u32* __restrict__ vals,
                                                   u32* span) {
    advance_span(span);
    const u32 i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const u64 o = 3ull * i;
    keys[i] = g.key_of(xyz[o], xyz[o + 1], xyz[o + 2]);
    vals[i] = i;
}

// Scene of original index i: the last s with starts[s] <= i (empty scenes
// share their start with the next scene and are skipped).
__device__ inline u32 scene_of_index(const u32* __restrict__ starts, u32 scenes, u32 i) {
    u32 lo = 0, hi = scenes - 1;
    while (lo < hi) {
        const u32 mid = (lo + hi + 1) >> 1;
        if (starts[mid] <= i) lo = mid;
        else hi = mid - 1;
    }
    return lo;
}

__global__ __launch_bounds__(256) void k_seg_keys(const float* __restrict__ xyz, u32 n, GridGeom g,
                                                  const u32* __restrict__ starts, u32 scenes,
                                                  u32* __restrict__ keys, u32* __restrict__ vals) {
    const u32 i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const u64 o = 3ull * i;
    const u32 s = scene_of_index(starts, scenes, i);
    keys[i] = s * (g.C + 1u) + g.key_of(xyz[o], xyz[o + 1], xyz[o + 2]);
    vals[i] = i;
}

// Each workgroup owns CPT * 256 consecutive cells.
template <int CPT>
__global__ __launch_bounds__(CS_THREADS) void k_cell_start(const u32* __restrict__ keys, u32 n,
                                                           u32 ncs, u32* __restrict__ cs,
                                                           u32* span) {
    constexpr int CS_PER_THREAD = CPT;
    constexpr u32 CS_TILE = CS_THREADS * CPT;
    __shared__ u32 s_mark[CS_TILE + CS_TILE / 16];
    __shared__ u32 s_k[2];
    __shared__ u32 s_wmin[CS_THREADS / 64];
    const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
    const u32 c0 = blockIdx.x * CS_TILE;
    const u32 c1 = (ncs - c0 < CS_TILE) ? ncs : c0 + CS_TILE;
    if (span && n > 0) {
        // this build's slot and the previous build's (written by the
        // previous launch, read-only here); all workgroups record the same
        const u32 par = span[kSpanCounter] & 1u;
        u32* cur = span + 4 * par;
        const u32* prev = span + 4 * (par ^ 1u);
        const u32 kmin = keys[0], kmax = keys[n - 1];
        if (tid == 0) {
            cur[0] = kmin;
            cur[1] = kmax;
            cur[2] = n;
            cur[3] = 1u;
        }
        if (prev[3] == 1u && prev[2] == n) {
            const u32 lo = min(kmin, prev[0]), hi = max(kmax, prev[1]);
            // cells c <= lo hold 0 and cells c > hi hold n in both tables
            if (c1 - 1u <= lo || c0 > hi) return;  // block-uniform
        }
    }
    if (w < 2) {
        const u32 r = wave_lower_bound(keys, n, w == 0 ? c0 : c1);
        if (lane == 0) s_k[w] = r;
    }
    __syncthreads();
    const u32 k0 = s_k[0], k1 = s_k[1];
    for (u32 i = tid; i < CS_TILE; i += CS_THREADS) s_mark[cs_pad(i)] = k1;
    __syncthreads();
    // keys[k0, k1) lie in [c0, c1): mark where each occupied cell's run starts
    for (u32 s = k0 + tid; s < k1; s += CS_THREADS) {
        const u32 key = keys[s];
        if (s == k0 || keys[s - 1] != key) s_mark[cs_pad(key - c0)] = s;
    }
    __syncthreads();
    // suffix minimum over the tile: empty cells take the next run start
    u32 m = k1;
#pragma unroll
    for (int j = CS_PER_THREAD - 1; j >= 0; --j) {
        const u32 idx = cs_pad(tid * CS_PER_THREAD + j);
        m = min(m, s_mark[idx]);
        s_mark[idx] = m;
    }
    u32 x = m;
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) {
        const u32 y = __shfl_down(x, o);
        if (lane + o < 64) x = min(x, y);
    }
    if (lane == 0) s_wmin[w] = x;
    __syncthreads();
    u32 after = k1;
    for (int ww = w + 1; ww < CS_THREADS / 64; ++ww) after = min(after, s_wmin[ww]);
    u32 excl = __shfl_down(x, 1);
    if (lane == 63) excl = k1;
    excl = min(excl, after);
#pragma unroll
    for (int j = 0; j < CS_PER_THREAD; ++j) {
        const u32 idx = cs_pad(tid * CS_PER_THREAD + j);
        s_mark[idx] = min(s_mark[idx], excl);
    }
    __syncthreads();
    for (u32 i = tid; i < c1 - c0; i += CS_THREADS) cs[c0 + i] = s_mark[cs_pad(i)];
}

}  // namespace

void build_cell_start(const u32* sorted_keys, u32 n, u32 ncs, u32* cell_start, hipStream_t stream,
                      u32 cpt, u32* span) {
    switch (cpt) {
    case 4:
        k_cell_start<4><<<div_up(ncs, CS_THREADS * 4), CS_THREADS, 0, stream>>>(sorted_keys, n, ncs,
                                                                                cell_start, span);
        break;
    case 8:
        k_cell_start<8><<<div_up(ncs, CS_THREADS * 8), CS_THREADS, 0, stream>>>(sorted_keys, n, ncs,
                                                                                cell_start, span);
        break;
    case 16:
        k_cell_start<16><<<div_up(ncs, CS_THREADS * 16), CS_THREADS, 0, stream>>>(
            sorted_keys, n, ncs, cell_start, span);
        break;
    default:
        fail(kInvalidArgument, "build_cell_start: cells per thread must be 4, 8 or 16");
    }
    WSIM_LAUNCH_CHECK();
}

CellGrid::CellGrid(const float origin[3], const int32_t dims[3], float cell_length,
                   hipStream_t stream)
    : cell_(cell_length), stream_(stream) {
    if (dims[0] <= 0 || dims[1] <= 0 || dims[2] <= 0)
        fail(kInvalidArgument, "CellGrid: cells_per_axis must be positive");
    if (!(cell_length > 0.0f)) fail(kInvalidArgument, "CellGrid: cell_length must be positive");
    const u64 C = static_cast<u64>(dims[0]) * static_cast<u64>(dims[1]) * static_cast<u64>(dims[2]);
    if (C + 1 >= 0xFFFFF000ull) fail(kInvalidArgument, "CellGrid: too many cells");
    g_.ox = origin[0];
    g_.oy = origin[1];
    g_.oz = origin[2];
    g_.inv = 1.0f / cell_length;  // cell_grid.cu:81
    g_.dx = dims[0];
    g_.dy = dims[1];
    g_.dz = dims[2];
    g_.C = static_cast<u32>(C);
    cs_ = DeviceBuffer<u32>(C + 1);
    cs_.zero(stream_);  // cell_grid.cu:63
    span_ = DeviceBuffer<u32>(kSpanWords);
    reset_span();
}

// Both slots invalid: the next build writes the whole table.
void CellGrid::reset_span() { span_.zero(stream_); }

void CellGrid::set_track_span(bool on) {
    if (on && segments_ != 1) fail(kInvalidArgument, "CellGrid: span tracking needs one scene");
    track_span_ = on;
    reset_span();
}

void CellGrid::reserve(size_t capacity) {
    if (capacity <= capacity_) return;
    if (capacity > 0xFFFFF000ull) fail(kInvalidArgument, "CellGrid: capacity too large");
    capacity_ = capacity;
    for (int b = 0; b < 2; ++b) {
        k_[b] = DeviceBuffer<u32>(capacity);
        v_[b] = DeviceBuffer<u32>(capacity);
        // zeroed once so every slot always holds a valid particle index
        k_[b].zero(stream_);
        v_[b].zero(stream_);
    }
    hist_ = DeviceBuffer<u32>(radix_hist_elems(capacity));
    cur_ = 0;
    reset_span();
}

void CellGrid::build(const float* positions_xyz, size_t count) {
    if (count == 0) {  // cell_grid.cu:74-77
        count_ = 0;
        return;
    }
    if (count > capacity_) reserve(count);
    count_ = count;
    const u32 n = static_cast<u32>(count);
    k_cell_keys<<<div_up(n, 256), 256, 0, stream_>>>(positions_xyz, n, g_, k_[0].data(), v_[0].data(),
                                                     span());
    WSIM_LAUNCH_CHECK();
    cur_ = radix_sort_pairs(k_[0].data(), v_[0].data(), k_[1].data(), v_[1].data(), count,
                            key_bits(g_.C), hist_.data(), stream_);
    build_cell_start(k_[cur_].data(), n, g_.C + 1, cs_.data(), stream_, cpt_, span());
}

u32 CellGrid::first_digit_mask() const { return radix_first_digit_mask(key_bits(g_.C)); }

void CellGrid::build_staged(size_t count, bool first_counted) {
    if (count > capacity_) fail(kInvalidArgument, "CellGrid: staged keys exceed the capacity");
    count_ = count;
    if (count == 0) return;
    const u32 n = static_cast<u32>(count);
    cur_ = radix_sort_pairs(k_[0].data(), v_[0].data(), k_[1].data(), v_[1].data(), count,
                            key_bits(g_.C), hist_.data(), stream_, first_counted);
    build_cell_start(k_[cur_].data(), n, g_.C + 1, cs_.data(), stream_, cpt_, span());
}

void CellGrid::reserve_segments(u32 scenes) {
    if (scenes == 0) fail(kInvalidArgument, "CellGrid: at least one scene");
    if (scenes == segments_) return;
    const u64 ncs = static_cast<u64>(scenes) * (static_cast<u64>(g_.C) + 1) + 1;
    if (ncs >= 0xFFFFF000ull) fail(kInvalidArgument, "CellGrid: too many cells over all scenes");
    cs_ = DeviceBuffer<u32>(ncs);
    cs_.zero(stream_);
    segments_ = scenes;
    if (scenes != 1) track_span_ = false;
    reset_span();
}

void CellGrid::build_segmented(const float* positions_xyz, size_t count, const u32* starts,
                               u32 scenes) {
    if (scenes != segments_) fail(kInvalidArgument, "CellGrid: scene count differs from reserve_segments");
    if (scenes == 1) {
        build(positions_xyz, count);
        return;
    }
    count_ = count;
    if (count == 0) return;
    if (count > capacity_) reserve(count);
    const u32 n = static_cast<u32>(count);
    k_seg_keys<<<div_up(n, 256), 256, 0, stream_>>>(positions_xyz, n, g_, starts, scenes,
                                                    k_[0].data(), v_[0].data());
    WSIM_LAUNCH_CHECK();
    const u32 ncs = scenes * (g_.C + 1) + 1;
    cur_ = radix_sort_pairs(k_[0].data(), v_[0].data(), k_[1].data(), v_[1].data(), count,
                            key_bits(ncs - 2), hist_.data(), stream_);
    build_cell_start(k_[cur_].data(), n, ncs, cs_.data(), stream_, cpt_);
}

void CellGrid::set_cells_per_thread(u32 cpt) {
    if (cpt != 4 && cpt != 8 && cpt != 16)
        fail(kInvalidArgument, "CellGrid: cells per thread must be 4, 8 or 16");
    cpt_ = cpt;
}

}  // namespace wsim
