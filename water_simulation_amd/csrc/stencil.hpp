// stencil.hpp -- device-side neighbour search shared by the solver and its
// adjoint: the 27-cell stencil walk of ITER_NEIGHBORS
// (solvers/dfsph/src/dfsph_kernels.cu:13-35) over cell-sorted working slots,
// the scene lookup of a slot, and the per-particle pair lists.
#pragma once

#include <type_traits>

#include "wsim_internal.hpp"

namespace wsim {
namespace dev {

constexpr u32 kNone = 0xFFFFFFFFu;
constexpr int NS = 256;  // workgroup size of the streaming kernels

// Register arrays of staged or batched loads use clang vector types, which
// stay in registers when the loops over them are unrolled.
typedef float f4v __attribute__((ext_vector_type(4)));
typedef float f2v __attribute__((ext_vector_type(2)));

// Neighbour-search and scene inputs shared by the per-particle kernels.
struct Nbr {
    GridGeom gf;
    const u32* csf;    // fluid cell_start; scene s's table starts at key_base
    GridGeom gb;
    const u32* csb;
    const float4* xb;  // cell-sorted boundary (x, y, z, mass)
    u32 nb;            // 0: no boundary
    Spline sp;
    float mass;
    u32 nslots;        // working-set slots (multiple of kTile)
    const SceneSeg* seg;
    const u32* tile_scene;  // scene of every tile; nullptr: one scene
    int remap;              // XCD-aware workgroup order
    const u32* keys;        // fluid cell keys in sorted order (scene key base included)
    const u32* bnear;       // boundary proximity mask over the boundary grid, or nullptr
};

// 27-cell stencil of ITER_NEIGHBORS (dfsph_kernels.cu:13-35): dz, dy, dx in
// {-1, 0, 1}, out-of-grid cells skipped, then every slot of the cell.  Sorted
// positions map to working slots by adding `shift`.
template <class F>
__device__ __forceinline__ void for_stencil(const GridGeom& g, const u32* __restrict__ cs,
                                            u32 shift, float px, float py, float pz, F&& f) {
    const int cx = g.cx(px), cy = g.cy(py), cz = g.cz(pz);
    for (int dz = -1; dz <= 1; ++dz) {
        const int nz = cz + dz;
        if (nz < 0 || nz >= g.dz) continue;
        for (int dy = -1; dy <= 1; ++dy) {
            const int ny = cy + dy;
            if (ny < 0 || ny >= g.dy) continue;
            for (int dx = -1; dx <= 1; ++dx) {
                const int nx = cx + dx;
                if (nx < 0 || nx >= g.dx) continue;
                const u32 c =
                    (static_cast<u32>(nx) * static_cast<u32>(g.dy) + static_cast<u32>(ny)) *
                        static_cast<u32>(g.dz) +
                    static_cast<u32>(nz);
                const u32 b = cs[c], e = cs[c + 1];
                for (u32 s = b; s < e; ++s) f(s + shift);
            }
        }
    }
}

// A walk's callback takes the slot, or the slot and the candidate's x
// offset in the stencil as 0, 1, 2 (dx + 1): the staged range of its tile
// that holds it (see stage_ranges).
template <class F>
__device__ __forceinline__ void visit_slot(F& f, u32 s, int xk) {
    if constexpr (std::is_invocable_v<F&, u32, int>) f(s, xk);
    else f(s);
}

// for_stencil with the cell-table loads of each (dz, dy) row issued
// together: the three cells' [start, end) pairs are loaded before the first
// slot of the row is visited, so a row waits on one table latency instead of
// three.  Same slots, same order.
template <class F>
__device__ __forceinline__ void for_stencil_rows(const GridGeom& g, const u32* __restrict__ cs,
                                                 u32 shift, float px, float py, float pz, F&& f) {
    const int cx = g.cx(px), cy = g.cy(py), cz = g.cz(pz);
    for (int dz = -1; dz <= 1; ++dz) {
        const int nz = cz + dz;
        if (nz < 0 || nz >= g.dz) continue;
        for (int dy = -1; dy <= 1; ++dy) {
            const int ny = cy + dy;
            if (ny < 0 || ny >= g.dy) continue;
            u32 b[3], e[3];
#pragma unroll
            for (int k = 0; k < 3; ++k) {
                const int nx = cx + k - 1;
                const bool in = nx >= 0 && nx < g.dx;
                const u32 c = in ? (static_cast<u32>(nx) * static_cast<u32>(g.dy) +
                                    static_cast<u32>(ny)) *
                                           static_cast<u32>(g.dz) +
                                       static_cast<u32>(nz)
                                 : 0u;
                const u32 bk = cs[c], ek = cs[c + 1];  // c = 0 (in bounds) for skipped cells
                b[k] = in ? bk : 0u;
                e[k] = in ? ek : 0u;
            }
#pragma unroll
            for (int k = 0; k < 3; ++k)
                for (u32 s = b[k]; s < e[k]; ++s) visit_slot(f, s + shift, k);
        }
    }
}

// The same 27 cells walked as nine z-contiguous rows (dx, dy in {-1, 0, 1},
// cells cz - 1 .. cz + 1 of a row are consecutive keys, so each row is one
// slot range): the same candidates in another order, for the tolerance
// mode, whose sums need no fixed order.  A wave's trip count per row is the
// maximum over its lanes of the row's candidates, which varies far less
// than the per-cell maximum of the reference-order walk (a lattice at the
// cell spacing leaves 0, 1, 2, 4 or 8 particles per cell).
template <class F>
__device__ __forceinline__ void for_stencil_zrows(const GridGeom& g, const u32* __restrict__ cs,
                                                  u32 shift, float px, float py, float pz, F&& f) {
    const int cx = g.cx(px), cy = g.cy(py), cz = g.cz(pz);
    const int lo = cz - 1 < 0 ? 0 : cz - 1;
    const int hi = cz + 1 > g.dz - 1 ? g.dz - 1 : cz + 1;
    u32 b[9], e[9];
#pragma unroll
    for (int r = 0; r < 9; ++r) {
        const int nx = cx + r / 3 - 1, ny = cy + r % 3 - 1;
        const bool in = lo <= hi && nx >= 0 && nx < g.dx && ny >= 0 && ny < g.dy;
        const u32 row = in ? (static_cast<u32>(nx) * static_cast<u32>(g.dy) + static_cast<u32>(ny)) *
                                 static_cast<u32>(g.dz)
                           : 0u;
        const u32 bk = cs[in ? row + static_cast<u32>(lo) : 0u];
        const u32 ek = cs[in ? row + static_cast<u32>(hi) + 1u : 0u];
        b[r] = in ? bk : 0u;
        e[r] = in ? ek : 0u;
    }
#pragma unroll
    for (int r = 0; r < 9; ++r)
        for (u32 s = b[r]; s < e[r]; ++s) visit_slot(f, s + shift, r / 3);
}

// Boundary proximity: bit k of the mask is set when the 27-cell stencil of
// boundary-grid cell k holds a boundary particle.  A fluid particle whose
// home cell has a clear bit has no boundary candidates, so the boundary walk
// can be skipped (it would add nothing).  Home cells outside the grid walk.
__device__ __forceinline__ bool boundary_near(const GridGeom& g, const u32* __restrict__ mask,
                                              float px, float py, float pz) {
    if (!mask) return true;
    const int x = g.cx(px), y = g.cy(py), z = g.cz(pz);
    const u32 k = g.key(x, y, z);
    if (k >= g.C) return true;
    return (mask[k >> 5] >> (k & 31u)) & 1u;
}

template <int NT>
__device__ __forceinline__ u32 slot_index(int remap) {
    const u32 t = remap ? xcd_tile(blockIdx.x, gridDim.x) : blockIdx.x;
    return t * NT + threadIdx.x;
}

// Scene of working slot i (< nslots).  Waves never straddle tiles, so the
// result is wave-uniform and the scene table is read with scalar loads.
__device__ __forceinline__ u32 scene_of_slot(const Nbr& P, u32 i) {
    if (!P.tile_scene) return 0u;
    const u32 t = __builtin_amdgcn_readfirstlane(i / kTile);
    return P.tile_scene[t];
}

// Per-particle pair lists: entry k of particle i holds (neighbour slot, bits
// of c) with grad W_ij = (x_i - x_j) * c, for every pair the reference's
// stencil visits with c != 0, in the reference's summation order.  Positions
// are frozen from the grid build to advection, so the iteration kernels
// recompute r and g = r * c from the list and get the same binary32 values
// the reference gets from grad_W (zero-gradient pairs add +-0 to sums that
// never hold -0, so leaving them out is exact).  Particles with more pairs
// than the list depth use the stencil walk.
//
// Wave-blocked ELL: the K entries of the 64 particles of one wave occupy one
// contiguous block, row k of it holding entry k of each lane.
__device__ __forceinline__ u64 pair_at(u32 k, u32 i, u32 depth) {
    return (static_cast<u64>(i >> 6) * depth + k) * 64u + (i & 63u);
}

// List entries are streamed once per traversal: load them non-temporally so
// they do not evict the gathered neighbour data from L2.
__device__ __forceinline__ uint2 ld_pair(const uint2* p) {
    const u64 v = __builtin_nontemporal_load(reinterpret_cast<const u64*>(p));
    return make_uint2(static_cast<u32>(v), static_cast<u32>(v >> 32));
}

__device__ __forceinline__ bool finite3(float a, float b, float c) {
    return isfinite(a) && isfinite(b) && isfinite(c);
}

// ---- LDS-staged neighbourhoods ------------------------------------------
// Keys grow fastest along z, so the stencil cells of a tile of consecutive
// cell-sorted particles, whose keys span [kmin, kmax], lie for each x offset
// dx in the key range [kmin + dx Dy Dz - Dz - 1, kmax + dx Dy Dz + Dz + 1]:
// three contiguous slot ranges hold every fluid neighbour of the tile.  A
// tile stages them in LDS when they are disjoint and fit; otherwise (sparse
// tiles, particles outside the grid) it gathers from global memory.
constexpr u32 kStageCap = 1536;  // staged fluid particles per 256-slot tile
static_assert(kStageCap % kTile == 0, "whole staging rounds per thread");

// (StageRanges is declared in wsim_internal.hpp; an empty range sits where
// the previous one ends.)

// Computed by one thread for the tile of working slots [t0, t0 + kTile).
__device__ inline StageRanges stage_ranges(const Nbr& P, const SceneSeg& G, u32 t0) {
    StageRanges R{0, 0, 0, 0, 0, 0};
    const u32 l0 = t0 - G.off;
    if (l0 >= G.n) return R;
    const u32 l1 = (G.n - l0 < kTile ? G.n : l0 + kTile) - 1;
    const u32 C = P.gf.C;
    const u32 kmin = P.keys[G.start + l0] - G.key_base;
    const u32 kmax = P.keys[G.start + l1] - G.key_base;
    if (kmax >= C) return R;  // particles outside the grid: their stencils are elsewhere
    const long long dz = P.gf.dz, dyz = static_cast<long long>(P.gf.dy) * P.gf.dz;
    const u32* cs = P.csf + G.key_base;
    const u32 shift = G.off - G.start;
    u32 prev_end = 0;
    bool overlap = false;
    // slot range [s, e) of the stencil columns at x offset dx
    auto range = [&](long long dx, u32& s, u32& e) {
        long long klo = static_cast<long long>(kmin) + dx * dyz - dz - 1;
        long long khi = static_cast<long long>(kmax) + dx * dyz + dz + 1;
        if (klo < 0) klo = 0;
        if (khi > static_cast<long long>(C) - 1) khi = static_cast<long long>(C) - 1;
        s = e = prev_end;
        if (klo <= khi) {
            const u32 a = cs[klo] + shift, b = cs[khi + 1] + shift;
            if (b > a) {
                if (a < prev_end) overlap = true;
                s = a;
                e = b;
                prev_end = b;
            }
        }
    };
    u32 e0, e1, e2;
    range(-1, R.s0, e0);
    range(0, R.s1, e1);
    range(1, R.s2, e2);
    if (overlap) return StageRanges{0, 0, 0, 0, 0, 0};
    R.b1 = e0 - R.s0;
    R.b2 = R.b1 + (e1 - R.s1);
    const u32 total = R.b2 + (e2 - R.s2);
    R.total = total <= kStageCap ? total : 0u;
    return R;
}

// LDS entry of neighbour slot j, a stencil candidate of x offset
// dx = xk - 1 of a particle of a staged tile: every stencil cell at that
// offset of a particle whose key k lies in [kmin, kmax] has a key in
// [k + dx Dy Dz - Dz - 1, k + dx Dy Dz + Dz + 1], inside range xk, so the
// entry is j plus that range's offset.  The three offsets are wave-uniform
// (a wave's particles share one tile).
struct StageOffsets {
    u32 off[3];
    __device__ explicit StageOffsets(const StageRanges& R)
        : off{uniform(0u - R.s0), uniform(R.b1 - R.s1), uniform(R.b2 - R.s2)} {}
    __device__ static u32 uniform(u32 v) {
        return static_cast<u32>(__builtin_amdgcn_readfirstlane(static_cast<int>(v)));
    }
    __device__ __forceinline__ u32 entry(u32 j, int xk) const { return j + off[xk]; }
};
// Working slot of staged entry e.
__device__ __forceinline__ u32 stage_source(const StageRanges& R, u32 e) {
    return e >= R.b2 ? R.s2 + (e - R.b2) : (e >= R.b1 ? R.s1 + (e - R.b1) : R.s0 + e);
}

// Fluid list entry k of particle i.  A tile whose neighbourhood is staged
// (plan total != 0) keeps only the 2-byte LDS entry of each neighbour, whose
// working slot is stage_source(R, entry); other tiles keep (slot, bits of c).
// Positions are frozen between the list build and advection, so c of a
// staged entry is recomputed by pair_c with the expression that produced it
// in the list build: the same binary32 value either way.
//
// The 2-byte entries are grouped by four: entries 4g..4g+3 of a lane form
// one 8-byte word, and word g of the 64 lanes of a wave are contiguous, so a
// lane reads four entries with one load.
__device__ __forceinline__ u64 quad_at(u32 g, u32 i, u32 depth) {
    return (static_cast<u64>(i >> 6) * (depth >> 2) + g) * 64u + (i & 63u);
}
__device__ __forceinline__ u64 stage_entry_at(u32 k, u32 i, u32 depth) {
    return quad_at(k >> 2, i, depth) * 4u + (k & 3u);
}
__device__ __forceinline__ u64 ld_quad(const PairList& L, u32 g, u32 i) {
    return __builtin_nontemporal_load(reinterpret_cast<const u64*>(L.fi) + quad_at(g, i, L.kf));
}
__device__ __forceinline__ u32 fluid_stage_entry(const PairList& L, u32 k, u32 i) {
    return __builtin_nontemporal_load(L.fi + stage_entry_at(k, i, L.kf));
}
// c is left 0 for a staged entry.
__device__ __forceinline__ void fluid_entry(const PairList& L, const StageRanges& R, u32 k, u32 i,
                                            u32& j, float& c) {
    if (R.total) {
        j = stage_source(R, fluid_stage_entry(L, k, i));
        c = 0.0f;
    } else {
        const uint2 e = ld_pair(L.f + pair_at(k, i, L.kf));
        j = e.x;
        c = __uint_as_float(e.y);
    }
}
// Visits the cf staged entries of particle i in order, four per call:
// f(e0, e1, e2, e3, m) with m in [1, 4] valid entries; the invalid ones of
// the last word repeat e0, so every lane can load and compute all four and
// drop the terms past m with a select (adding -0.0f, or subtracting +0.0f,
// leaves every binary32 value unchanged).  One load per four entries; the
// next word is requested before the current one is used.
template <class F>
__device__ __forceinline__ void for_stage_quads(const PairList& L, u32 i, u32 cf, F&& f) {
    const u32 nq = (cf + 3u) >> 2;
    u64 next = nq ? ld_quad(L, 0, i) : 0ull;
    for (u32 g = 0; g < nq; ++g) {
        const u64 cur = next;
        if (g + 1 < nq) next = ld_quad(L, g + 1, i);
        const u32 m = cf - 4u * g < 4u ? cf - 4u * g : 4u;
        const u32 e0 = static_cast<u32>(cur & 0xFFFFu);
        const u32 e1 = m > 1u ? static_cast<u32>((cur >> 16) & 0xFFFFu) : e0;
        const u32 e2 = m > 2u ? static_cast<u32>((cur >> 32) & 0xFFFFu) : e0;
        const u32 e3 = m > 3u ? static_cast<u32>(cur >> 48) : e0;
        f(e0, e1, e2, e3, m);
    }
}

__device__ __forceinline__ float pair_c(const float4& p, const float4& q, const Spline& sp) {
    return spline_gradc_near(pair_sqrt(dist2(p.x - q.x, p.y - q.y, p.z - q.z), sp), sp);
}
// c of a pair of the fluid list in either pair-math mode, with the
// expression of the list build, so a recomputed c equals the recorded one
// (its c was not 0 when the list was built, so the tests for coincident
// particles are left out: the same bits).
template <bool F>
__device__ __forceinline__ float pair_c_listed(const float4& p, const float4& q, const Spline& sp) {
    return PairMath<F>::c_listed(PairMath<F>::r2(p.x - q.x, p.y - q.y, p.z - q.z), sp);
}
// Gradient factor of a stencil candidate inside the support (r2 < far2).
template <bool F>
__device__ __forceinline__ float stencil_c(float r2, const Spline& sp) {
    if (F) return PairMath<true>::c(r2, sp);
    return spline_gradc(pair_sqrt(r2, sp), sp);
}

// The staging entries of one thread: tid, tid + 256, tid + 512, ...  src
// holds their working slots (entries past the plan's end repeat its last
// entry, so every load is in bounds and all of them can be issued before
// the first LDS store).  Register arrays of the staged data use clang vector
// types, which stay in registers when the loops are unrolled.
constexpr int kStagePer = kStageCap / kTile;
struct StageLanes {
    u32 src[kStagePer];
    u32 total;
    __device__ __forceinline__ explicit StageLanes(const StageRanges& R) : total(R.total) {
#pragma unroll
        for (int u = 0; u < kStagePer; ++u)
            src[u] = stage_source(R, min(entry(u), R.total - 1u));
    }
    __device__ __forceinline__ static u32 entry(int u) {
        return threadIdx.x + static_cast<u32>(u) * kTile;
    }
    __device__ __forceinline__ bool has(int u) const { return entry(u) < total; }
};
__device__ __forceinline__ f4v ld4v(const float4* p, u32 i) {
    return reinterpret_cast<const f4v*>(p)[i];
}

// Host-side assembly of the kernel inputs.
inline Nbr make_nbr(const CellGrid& fg, const CellGrid* bg, const float4* xb, const u32* bnear,
                    size_t nb, const DFSPHConfig& cfg, float mass, u32 nslots,
                    const SceneSeg* seg, const u32* tile_scene, u32 remap) {
    Nbr P;
    P.gf = fg.geom();
    P.csf = fg.cell_start();
    if (bg && nb) {
        P.gb = bg->geom();
        P.csb = bg->cell_start();
        P.xb = xb;
        P.nb = static_cast<u32>(nb);
        P.bnear = bnear;
    } else {
        P.gb = fg.geom();
        P.csb = nullptr;
        P.xb = nullptr;
        P.nb = 0;
        P.bnear = nullptr;
    }
    P.sp = Spline::make(cfg.smoothing_length);
    P.mass = mass;
    P.nslots = nslots;
    P.seg = seg;
    P.tile_scene = tile_scene;
    P.remap = static_cast<int>(remap);
    P.keys = fg.sorted_keys();
    return P;
}

}  // namespace dev
}  // namespace wsim
