// stencil.hpp -- device-side neighbour search shared by the solver and its
// adjoint: the 27-cell stencil walk of ITER_NEIGHBORS
// (solvers/dfsph/src/dfsph_kernels.cu:13-35) over cell-sorted working slots,
// the scene lookup of a slot, and the per-particle pair lists.
#pragma once

#include "wsim_internal.hpp"

namespace wsim {
namespace dev {

constexpr u32 kNone = 0xFFFFFFFFu;
constexpr int NS = 256;  // workgroup size of the streaming kernels

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

// Host-side assembly of the kernel inputs.
inline Nbr make_nbr(const CellGrid& fg, const CellGrid* bg, const float4* xb, size_t nb,
                    const DFSPHConfig& cfg, float mass, u32 nslots, const SceneSeg* seg,
                    const u32* tile_scene, u32 remap) {
    Nbr P;
    P.gf = fg.geom();
    P.csf = fg.cell_start();
    if (bg && nb) {
        P.gb = bg->geom();
        P.csb = bg->cell_start();
        P.xb = xb;
        P.nb = static_cast<u32>(nb);
    } else {
        P.gb = fg.geom();
        P.csb = nullptr;
        P.xb = nullptr;
        P.nb = 0;
    }
    P.sp = Spline::make(cfg.smoothing_length);
    P.mass = mass;
    P.nslots = nslots;
    P.seg = seg;
    P.tile_scene = tile_scene;
    P.remap = static_cast<int>(remap);
    return P;
}

}  // namespace dev
}  // namespace wsim
