// solver.hip -- DFSPH substep on cell-sorted particle data for gfx950.
//
// Replaces DFSPHSolver::step (solvers/dfsph/src/dfsph.cu:120-133) and its 11
// kernels (solvers/dfsph/src/dfsph_kernels.cu:40-400), for one scene or for
// a batch of independent scenes of uneven sizes.
//
// Data flow per substep (all launches on the solver's stream):
//   grid build         keys -> stable radix sort -> dense cell_start      (grid.hip)
//   k_reorder          gather position/velocity into cell-sorted float4
//                      arrays; gravity + damping applied on the way
//                      (dfsph_kernels.cu:347-355; density and alpha do not
//                      read velocity, so applying it first is equivalent)
//   k_density_alpha    rho and alpha in one stencil walk (dfsph_kernels.cu:68-169)
//   k_xsph             Jacobi XSPH into the working velocity (dfsph_kernels.cu:305-341)
//   density solve      k_change<0> / k_apply per iteration; binary64 error
//                      partials per tile, reduced by the first wave of each
//                      scene in the following k_apply, which also decides the
//                      scene's early exit on the device
//                      (dfsph.cu:215-260, dfsph_kernels.cu:176-301, 389-394)
//   k_advect           clamp, integrate, domain clamp, scatter position,
//                      velocity and the solver attributes back to the caller's
//                      store in original order, max |v| and NaN count
//                      (dfsph_kernels.cu:363-386, 396-400; dfsph.cu:303-326);
//                      a single scene uses k_advect_gather, one thread per
//                      original index reading through the inverse permutation
//                      that k_reorder records
//   [divergence solve] grid, reorder, density/alpha, k_change<1> / k_apply,
//                      k_div_writeback (docs/superpowers/plans/2026-04-24-water-simulation-v2-phase2-dfsph-solver.md:1858-2003)
//
// Scenes.  Scene s owns store indices [start_s, start_s + n_s) and working
// slots [off_s, off_s + n_s), off_s a multiple of kTile, so every tile and
// every wave belongs to one scene (looked up per tile, read with scalar
// loads).  The segmented grid gives scene s the keys s*(C+1) + key; its
// cells' sorted ranges shift by off_s - start_s into working slots.  dt,
// damping factor, gravity step and XSPH viscosity are per scene; the error
// sum, the early exit and max |v| are per scene.  A plain solver is the one
// scene case.
//
// Every per-pair expression keeps the reference's operand order, the stencil
// is walked dz, dy, dx with ascending original id inside each cell (the
// sorted slots of a cell are in id order), and the code is compiled without
// FP contraction and with IEEE divide/sqrt: each field equals the binary32
// oracle bit for bit, for every launch shape and in every batch.
#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstring>
#include <initializer_list>
#include <limits>

#include "radix_sort.hpp"
#include "stencil.hpp"
#include "wsim_internal.hpp"

namespace wsim {
namespace {

using namespace dev;

// parts: 1 positions, 2 velocities, 3 both.
// inv (optional, with parts & 1): inv[original index] = working slot.
// tag_ids (with parts & 1): the position's w holds the original index's bits
// (advance() carries them through its slot-order state); otherwise 0.
__global__ __launch_bounds__(NS) void k_reorder(Nbr P, const u32* __restrict__ perm,
                                                const float* __restrict__ x,
                                                const float* __restrict__ v,
                                                float4* __restrict__ xs, float4* __restrict__ vs,
                                                int gravity, int parts, u32* __restrict__ inv,
                                                int tag_ids = 0) {
    const u32 i = blockIdx.x * NS + threadIdx.x;
    if (i >= P.nslots) return;
    const SceneSeg& G = P.seg[scene_of_slot(P, i)];
    const u32 l = i - G.off;
    if (l >= G.n) return;
    const u32 id = perm[G.start + l];
    const u64 o = 3ull * id;
    if (parts & 1) {
        const float3 p = *reinterpret_cast<const float3*>(x + o);
        xs[i] = make_float4(p.x, p.y, p.z, tag_ids ? __uint_as_float(id) : 0.0f);
        if (inv) inv[id] = i;
    }
    if (!(parts & 2)) return;
    const float3 w = *reinterpret_cast<const float3*>(v + o);
    float a = w.x, b = w.y, c = w.z;
    if (gravity) {  // v * exp(-damping*dt) + g*dt (dfsph_kernels.cu:351-354)
        a = a * G.decay + G.gdx;
        b = b * G.decay + G.gdy;
        c = c * G.decay + G.gdz;
    }
    vs[i] = make_float4(a, b, c, 0.0f);
}

__global__ __launch_bounds__(NS) void k_gather_boundary(const u32* __restrict__ perm,
                                                        const float* __restrict__ x, u32 nb,
                                                        float4* __restrict__ xb) {
    const u32 s = blockIdx.x * NS + threadIdx.x;
    if (s >= nb) return;
    const u64 o = 3ull * perm[s];
    xb[s] = make_float4(x[o], x[o + 1], x[o + 2], 0.0f);
}

// Boundary proximity mask: one bit per boundary-grid cell, set when a cell
// of its 27-cell stencil is occupied.  Built once with the boundary grid.
__global__ __launch_bounds__(NS) void k_boundary_near(GridGeom g, const u32* __restrict__ cs,
                                                      u32* __restrict__ mask) {
    const u32 k = blockIdx.x * NS + threadIdx.x;
    bool near = false;
    if (k < g.C) {
        const u32 dyz = static_cast<u32>(g.dy) * static_cast<u32>(g.dz);
        const int x = static_cast<int>(k / dyz);
        const int y = static_cast<int>((k % dyz) / static_cast<u32>(g.dz));
        const int z = static_cast<int>(k % static_cast<u32>(g.dz));
        for (int a = -1; a <= 1; ++a)
            for (int b = -1; b <= 1; ++b)
                for (int c = -1; c <= 1; ++c) {
                    const u32 n = g.key(x + a, y + b, z + c);
                    if (n < g.C && cs[n + 1] > cs[n]) near = true;
                }
    }
    const u64 bits = __ballot(near);
    const u32 w = (blockIdx.x * NS + (threadIdx.x & ~63u)) >> 5;  // first word of this wave
    if ((threadIdx.x & 63u) == 0) mask[w] = static_cast<u32>(bits);
    if ((threadIdx.x & 63u) == 32) mask[w + 1] = static_cast<u32>(bits >> 32);
}

// Staging plan of every 256-slot tile (one thread per tile), computed once
// per grid build and read by the list builder and the iteration kernels.
__global__ __launch_bounds__(NS) void k_stage_plan(Nbr P, StageRanges* __restrict__ plan) {
    const u32 t = blockIdx.x * NS + threadIdx.x;
    if (t >= P.nslots / kTile) return;
    const u32 sc = P.tile_scene ? P.tile_scene[t] : 0u;
    plan[t] = stage_ranges(P, P.seg[sc], t * kTile);
}

// Akinci boundary mass, in cell-sorted order (dfsph_kernels.cu:40-65).
__global__ __launch_bounds__(NS) void k_boundary_mass(GridGeom g, const u32* __restrict__ cs,
                                                      const float4* __restrict__ xb, u32 nb,
                                                      Spline sp, float rho0,
                                                      float* __restrict__ mb) {
    const u32 i = blockIdx.x * NS + threadIdx.x;
    if (i >= nb) return;
    const float4 p = xb[i];
    float sum_w = 0.0f;
    for_stencil(g, cs, 0u, p.x, p.y, p.z, [&](u32 s) {
        const float4 q = xb[s];
        const float r2 = dist2(p.x - q.x, p.y - q.y, p.z - q.z);
        if (r2 >= sp.far2) return;  // W = 0
        sum_w += spline_W_near(pair_sqrt(r2, sp), sp);
    });
    mb[i] = rho0 / fmaxf(sum_w, 1e-9f);
}

__global__ __launch_bounds__(NS) void k_pack_boundary(u32 nb, const u32* __restrict__ perm,
                                                      const float* __restrict__ mb,
                                                      float4* __restrict__ xb,
                                                      float* __restrict__ store_mb) {
    const u32 s = blockIdx.x * NS + threadIdx.x;
    if (s >= nb) return;
    xb[s].w = mb[s];
    store_mb[perm[s]] = mb[s];
}

// Neighbour counts of working slots to original order (as k_reorder maps
// slots to particles).
__global__ __launch_bounds__(NS) void k_scatter_counts(Nbr P, const u32* __restrict__ perm,
                                                       const uint4* __restrict__ counts,
                                                       u32* __restrict__ cand_f,
                                                       u32* __restrict__ nbr_f,
                                                       u32* __restrict__ cand_b,
                                                       u32* __restrict__ nbr_b) {
    const u32 i = blockIdx.x * NS + threadIdx.x;
    if (i >= P.nslots) return;
    const SceneSeg& G = P.seg[scene_of_slot(P, i)];
    const u32 l = i - G.off;
    if (l >= G.n) return;
    const u32 id = perm[G.start + l];
    const uint4 c = counts[i];
    if (cand_f) cand_f[id] = c.x;
    if (nbr_f) nbr_f[id] = c.y;
    if (cand_b) cand_b[id] = c.z;
    if (nbr_b) nbr_b[id] = c.w;
}

__device__ __forceinline__ u32 pack_counts(u32 cf, u32 cb) {
    return (cf < 0xFFFFu ? cf : 0xFFFFu) | ((cb < 0xFFFFu ? cb : 0xFFFFu) << 16);
}

// Density and alpha sums of one pair with mass m (dfsph_kernels.cu:95-107,
// 145-163): rho += m W, G += m grad W, Q += |m grad W|^2.
template <bool F>
__device__ __forceinline__ void grad_terms(float& rho, float& sx, float& sy, float& sz, float& sq,
                                           float m, float wv, float c, float rx, float ry,
                                           float rz) {
    if (F) {
        rho = __builtin_fmaf(m, wv, rho);
        const float mc = m * c;
        const float mx = rx * mc, my = ry * mc, mz = rz * mc;
        sx += mx;
        sy += my;
        sz += mz;
        sq = __builtin_fmaf(mx, mx, __builtin_fmaf(my, my, __builtin_fmaf(mz, mz, sq)));
    } else {
        rho += m * wv;
        const float mx = m * (rx * c), my = m * (ry * c), mz = m * (rz * c);
        sx += mx;
        sy += my;
        sz += mz;
        sq += mx * mx + my * my + mz * mz;
    }
}

// rho (dfsph_kernels.cu:68-112) and alpha (dfsph_kernels.cu:117-169) in one
// stencil walk: they share r, |r| and q for every pair.  Also records the
// pair lists used by every iteration of the pressure solves.  With COUNT the
// walk also writes, per working slot, the neighbour counts of ITER_NEIGHBORS
// (dfsph_kernels.cu:13-35) for the fluid and the boundary store: candidates
// (every slot of the in-grid stencil cells, self included) and neighbours
// (candidates with sqrtf(r^2) / l < 2, i.e. r^2 < far2).
// Stable partition of a workgroup's slots, tile by tile (groups of at most
// 256 threads, so a wave never mixes tiles): slots whose particle is near a
// wall first, in slot order, then the others.  Returns the slot this thread
// processes.  Only which lane computes which particle changes.
template <int NT>
__device__ __forceinline__ u32 near_slots_first(u32 slot, bool near) {
    constexpr int GS = NT < kTile ? NT : kTile;
    constexpr int GW = GS / kWave;
    __shared__ u32 s_cnt[NT / kWave];
    __shared__ u32 s_slot[NT];
    const int t = threadIdx.x, lane = t & 63, w = t >> 6;
    const int g0 = (t / GS) * GS;  // first thread of this thread's group
    const int w0 = g0 / kWave;
    const u64 bal = __ballot(near);
    if (lane == 0) s_cnt[w] = static_cast<u32>(__popcll(bal));
    __syncthreads();
    u32 before = 0, total = 0;
#pragma unroll
    for (int k = 0; k < GW; ++k) {
        const u32 c = s_cnt[w0 + k];
        if (w0 + k < w) before += c;
        total += c;
    }
    const u32 below = before + static_cast<u32>(__popcll(bal & ((1ull << lane) - 1ull)));
    const u32 r = static_cast<u32>(t - g0);
    s_slot[g0 + (near ? below : total + (r - below))] = slot;
    __syncthreads();
    return s_slot[t];
}

// With `open`, after an advection (advance(), one scene) thread 0 also opens
// the next density solve (k_reset_solve's work): the reorder before this
// kernel was the last to test the advected flag.
template <int NT, bool F, bool COUNT = false>
__global__ __launch_bounds__(NT) void k_density_alpha(Nbr P, PairList L,
                                                      const float4* __restrict__ xs,
                                                      float* __restrict__ rho_out,
                                                      float* __restrict__ alpha_out,
                                                      uint4* __restrict__ counts = nullptr,
                                                      DeviceState* open = nullptr) {
    if (open && blockIdx.x == 0 && threadIdx.x == 0 && open->advected) {
        open->conv_iter = P.seg[0].n ? kNone : 0u;
        open->advected = 0u;
    }
    // With one tile per workgroup, a tile whose neighbourhood fits the
    // staging plan reads its candidates' positions from LDS.
    constexpr bool kStage = NT == kTile;
    __shared__ float4 s_x[kStage ? kStageCap : 1];
    bool lds = false;
    if constexpr (kStage) {
        const u32 tile = P.remap ? xcd_tile(blockIdx.x, gridDim.x) : blockIdx.x;
        const StageRanges Rt = L.plan[tile];
        if (Rt.total) {
            // every load of a thread is issued before its LDS stores
            const StageLanes E(Rt);
            f4v x[kStagePer];
#pragma unroll
            for (int u = 0; u < kStagePer; ++u) x[u] = ld4v(xs, E.src[u]);
#pragma unroll
            for (int u = 0; u < kStagePer; ++u)
                if (E.has(u)) reinterpret_cast<f4v*>(s_x)[E.entry(u)] = x[u];
        }
        lds = Rt.total != 0u;
        __syncthreads();
    }
    // Particles near a wall also walk the boundary store.  Cell order puts
    // the start of every z-row next to a wall, so most waves would hold a few
    // such lanes and all wait through their walk; grouping them first keeps
    // the boundary walk to the waves that need it.
    const u32 i0 = slot_index<NT>(P.remap);
    bool near0 = false;
    if (P.nb && i0 < P.nslots) {
        const SceneSeg& G0 = P.seg[scene_of_slot(P, i0)];
        if (i0 - G0.off < G0.n) {
            const float4 q = xs[i0];
            near0 = boundary_near(P.gb, P.bnear, q.x, q.y, q.z);
        }
    }
    const u32 i = near_slots_first<NT>(i0, near0);
    if (i >= P.nslots) return;
    const SceneSeg& G = P.seg[scene_of_slot(P, i)];
    if (i - G.off >= G.n) return;
    const u32* cs = P.csf + G.key_base;
    const u32 shift = G.off - G.start;
    const float4 p = xs[i];
    float rho = 0.0f, sx = 0.0f, sy = 0.0f, sz = 0.0f, sq = 0.0f;
    u32 cf = 0, cb = 0;
    bool unlisted = false;  // an in-support fluid pair with a zero gradient
    uint4 nc = make_uint4(0u, 0u, 0u, 0u);  // COUNT: candidates, neighbours (fluid, boundary)
    // the tile's plan is wave-uniform: a wave's particles share one tile
    const StageRanges R = L.plan[i / kTile];
    const bool staged = __builtin_amdgcn_readfirstlane(R.total) != 0u;
    const StageOffsets O(R);
    // entry cf of this particle's list: stage_entry_at(cf, i, kf) and
    // pair_at(cf, i, kf) from their values at cf = 0
    uint16_t* const fi = L.fi + stage_entry_at(0u, i, L.kf);
    uint2* const f8 = L.f + pair_at(0u, i, L.kf);
    auto fluid_pair = [&](u32 s, int xk, const float4& q) {
        const float rx = p.x - q.x, ry = p.y - q.y, rz = p.z - q.z;
        const float r2 = PairMath<F>::r2(rx, ry, rz);
        if (COUNT) {
            ++nc.x;
            nc.y += r2 < P.sp.far2 ? 1u : 0u;
        }
        if (r2 >= P.sp.far2) return;  // W = 0 and c = 0
        // the self pair (d = 0) has c = 0: its terms add +0, and it is not listed
        float wv, c;
        PairMath<F>::wc(r2, P.sp, wv, c);
        grad_terms<F>(rho, sx, sy, sz, sq, P.mass, wv, c, rx, ry, rz);
        if (c != 0.0f) {
            if (cf < L.kf) {
                if (staged)
                    fi[((cf >> 2) << 8) | (cf & 3u)] = static_cast<uint16_t>(O.entry(s, xk));
                else
                    f8[cf << 6] = make_uint2(s, __float_as_uint(c));
            }
            ++cf;
        } else if (s != i) {
            unlisted = true;
        }
    };
    auto from_lds = [&](u32 s, int xk) { fluid_pair(s, xk, s_x[O.entry(s, xk)]); };
    auto from_global = [&](u32 s, int xk) { fluid_pair(s, xk, xs[s]); };
    if (lds) {
        if (F) for_stencil_zrows(P.gf, cs, shift, p.x, p.y, p.z, from_lds);
        else for_stencil_rows(P.gf, cs, shift, p.x, p.y, p.z, from_lds);
    } else {
        if (F) for_stencil_zrows(P.gf, cs, shift, p.x, p.y, p.z, from_global);
        else for_stencil_rows(P.gf, cs, shift, p.x, p.y, p.z, from_global);
    }
    // XSPH sums over every in-support pair (W != 0) and reads the fluid list;
    // a pair the list leaves out (coincident particles) sends this particle
    // to the stencil walk in every list kernel
    if (unlisted) cf = 0xFFFFu;
    float gbx = 0.0f, gby = 0.0f, gbz = 0.0f;  // tolerance mode: sum_b m_b grad W_ib
    if (P.nb && boundary_near(P.gb, P.bnear, p.x, p.y, p.z)) {
        auto boundary_pair = [&](u32 s) {
            const float4 q = P.xb[s];
            const float rx = p.x - q.x, ry = p.y - q.y, rz = p.z - q.z;
            const float r2 = PairMath<F>::r2(rx, ry, rz);
            if (COUNT) {
                ++nc.z;
                nc.w += r2 < P.sp.far2 ? 1u : 0u;
            }
            if (r2 >= P.sp.far2) return;
            float wv, c;
            PairMath<F>::wc(r2, P.sp, wv, c);
            grad_terms<F>(rho, sx, sy, sz, sq, q.w, wv, c, rx, ry, rz);
            if (F && L.gb) {
                const float mc = q.w * c;
                gbx = __builtin_fmaf(rx, mc, gbx);
                gby = __builtin_fmaf(ry, mc, gby);
                gbz = __builtin_fmaf(rz, mc, gbz);
            } else if (c != 0.0f) {
                if (cb < L.kb) L.b[pair_at(cb, i, L.kb)] = make_uint2(s, __float_as_uint(c));
                ++cb;
            }
        };
        if (F) for_stencil_zrows(P.gb, P.csb, 0u, p.x, p.y, p.z, boundary_pair);
        else for_stencil_rows(P.gb, P.csb, 0u, p.x, p.y, p.z, boundary_pair);
    }
    if (F && L.gb) L.gb[i] = make_float4(gbx, gby, gbz, 0.0f);
    L.cnt[i] = pack_counts(cf, cb);
    if (COUNT) counts[i] = nc;
    rho_out[i] = rho;
    const float denom = F ? __builtin_fmaf(sx, sx, __builtin_fmaf(sy, sy, __builtin_fmaf(sz, sz, sq)))
                          : sx * sx + sy * sy + sz * sz + sq;
    alpha_out[i] = (denom > 1e-9f) ? (rho / denom) : 0.0f;
}

// One XSPH pair term (dfsph_kernels.cu:326-337) for a pair at squared
// distance r2 inside the support.
template <bool F>
__device__ __forceinline__ void xsph_term_r2(float& dx, float& dy, float& dz, float r2,
                                             const float4& vi, float rho_i, const float4& vj,
                                             float rho_j, float eps, const Nbr& P,
                                             bool keep = true) {
    const float w = PairMath<F>::w(r2, P.sp);
    const float rho_avg = 0.5f * (rho_i + rho_j);
    const float c = eps * PairMath<F>::div(P.mass, fmaxf(rho_avg, 1e-9f)) * w;
    if (F) {
        const float k = keep ? c : 0.0f;
        dx = __builtin_fmaf(k, vj.x - vi.x, dx);
        dy = __builtin_fmaf(k, vj.y - vi.y, dy);
        dz = __builtin_fmaf(k, vj.z - vi.z, dz);
    } else {
        // a dropped term adds -0.0f, which leaves every value unchanged
        dx += keep ? c * (vj.x - vi.x) : -0.0f;
        dy += keep ? c * (vj.y - vi.y) : -0.0f;
        dz += keep ? c * (vj.z - vi.z) : -0.0f;
    }
}
template <bool F>
__device__ __forceinline__ void xsph_term(float& dx, float& dy, float& dz, const float4& p,
                                          const float4& vi, float rho_i, const float4& q,
                                          const float4& vj, float rho_j, float eps,
                                          const Nbr& P, bool keep = true) {
    xsph_term_r2<F>(dx, dy, dz, PairMath<F>::r2(p.x - q.x, p.y - q.y, p.z - q.z), vi, rho_i, vj,
                    rho_j, eps, P, keep);
}

// One XSPH term of a pair from the fluid list (r2 > 0).  Binary32 path:
// xsph_term into (dx, dy, dz).  Tolerance mode: the term without its
// pair-independent factor 2 eps m sigma, summed in (tx, ty, tz), which
// xsph_listed_scale applies once per particle:
//   (W / sigma) / max(rho_i + rho_j, 2e-9) (v_j - v_i).
template <bool F>
__device__ __forceinline__ void xsph_listed(float& dx, float& dy, float& dz, float& tx, float& ty,
                                            float& tz, const float4& p, const float4& vi,
                                            float rho_i, const float4& q, const float4& vj,
                                            float rho_j, float eps, const Nbr& P,
                                            bool keep = true) {
    if (!F) {
        xsph_term<F>(dx, dy, dz, p, vi, rho_i, q, vj, rho_j, eps, P, keep);
        return;
    }
    const float r2 = PairMath<true>::r2(p.x - q.x, p.y - q.y, p.z - q.z);
    const float s = r2 * __builtin_amdgcn_rsqf(r2) * P.sp.inv_l;
    const float inner = __builtin_fmaf(s * s, __builtin_fmaf(0.75f, s, -1.5f), 1.0f);
    const float t = 2.0f - s;
    const float w = s < 1.0f ? inner : 0.25f * (t * t * t);
    const float k = keep ? w * __builtin_amdgcn_rcpf(fmaxf(rho_i + rho_j, 2.0f * 1e-9f)) : 0.0f;
    tx = __builtin_fmaf(k, vj.x - vi.x, tx);
    ty = __builtin_fmaf(k, vj.y - vi.y, ty);
    tz = __builtin_fmaf(k, vj.z - vi.z, tz);
}
template <bool F>
__device__ __forceinline__ void xsph_listed_scale(float& dx, float& dy, float& dz, float tx,
                                                  float ty, float tz, float eps, const Nbr& P) {
    if (!F) return;
    const float K = 2.0f * eps * P.mass * P.sp.sigma;
    dx = K * tx;
    dy = K * ty;
    dz = K * tz;
}

// XSPH as a Jacobi sweep: every neighbour velocity is the pre-XSPH one.  A
// scene without viscosity copies its velocities (dfsph.cu:198 skips XSPH).
// The fluid pair list holds every in-support pair in stencil order (pairs
// outside the support add +-0); PU pairs per trip have their loads issued
// together.  Particles without a complete list walk the stencil.
template <int NT, int PU, bool F>
__global__ __launch_bounds__(NT) void k_xsph(Nbr P, PairList L, const float4* __restrict__ xs,
                                             const float4* __restrict__ vin,
                                             const float* __restrict__ rho,
                                             float4* __restrict__ vout) {
    const u32 i = slot_index<NT>(P.remap);
    if (i >= P.nslots) return;
    const SceneSeg& G = P.seg[scene_of_slot(P, i)];
    if (i - G.off >= G.n) return;
    const float4 vi = vin[i];
    const float eps = G.eps;
    if (!(eps > 0.0f)) {
        vout[i] = vi;
        return;
    }
    const float4 p = xs[i];
    const float rho_i = rho[i];
    float dx = 0.0f, dy = 0.0f, dz = 0.0f;
    float tx = 0.0f, ty = 0.0f, tz = 0.0f;  // tolerance mode: see xsph_listed
    const u32 cf = L.cnt[i] & 0xFFFFu;
    if (cf <= L.kf) {
        const StageRanges R = L.plan[i / kTile];
        float c_unused;
        u32 k = 0;
        for (; k + PU <= cf; k += PU) {
            u32 j[PU];
            float4 q[PU], vj[PU];
            float rj[PU];
#pragma unroll
            for (int u = 0; u < PU; ++u) fluid_entry(L, R, k + u, i, j[u], c_unused);
#pragma unroll
            for (int u = 0; u < PU; ++u) {
                q[u] = xs[j[u]];
                vj[u] = vin[j[u]];
                rj[u] = rho[j[u]];
            }
#pragma unroll
            for (int u = 0; u < PU; ++u)
                xsph_listed<F>(dx, dy, dz, tx, ty, tz, p, vi, rho_i, q[u], vj[u], rj[u], eps, P);
        }
        for (; k < cf; ++k) {
            u32 j;
            fluid_entry(L, R, k, i, j, c_unused);
            xsph_listed<F>(dx, dy, dz, tx, ty, tz, p, vi, rho_i, xs[j], vin[j], rho[j], eps, P);
        }
        xsph_listed_scale<F>(dx, dy, dz, tx, ty, tz, eps, P);
        vout[i] = make_float4(vi.x + dx, vi.y + dy, vi.z + dz, 0.0f);
        return;
    }
    const u32* cs = P.csf + G.key_base;
    const u32 shift = G.off - G.start;
    for_stencil(P.gf, cs, shift, p.x, p.y, p.z, [&](u32 s) {
        if (s == i) return;
        const float4 q = xs[s];
        const float r2 = PairMath<F>::r2(p.x - q.x, p.y - q.y, p.z - q.z);
        if (r2 >= P.sp.far2) return;  // W = 0
        xsph_term_r2<F>(dx, dy, dz, r2, vi, rho_i, vin[s], rho[s], eps, P);
    });
    vout[i] = make_float4(vi.x + dx, vi.y + dy, vi.z + dz, 0.0f);
}

// k_xsph with the tile's fluid neighbourhood staged in LDS: positions, and
// pre-XSPH velocities with the density in w.  The same terms in the same
// order; tiles without a staging plan gather from global memory.
template <int PU, bool F>
__global__ __launch_bounds__(kTile) void k_xsph_st(Nbr P, PairList L, const float4* __restrict__ xs,
                                                   const float4* __restrict__ vin,
                                                   const float* __restrict__ rho,
                                                   float4* __restrict__ vout) {
    // 24 B per staged neighbour (as k_change_st): rho_j is gathered from
    // global memory, which keeps four workgroups per CU instead of three
    __shared__ float4 s_a[kStageCap];  // x, y, z, pre-XSPH vx
    __shared__ float2 s_b[kStageCap];  // pre-XSPH vy, vz
    const u32 tile = P.remap ? xcd_tile(blockIdx.x, gridDim.x) : blockIdx.x;
    const u32 t0 = tile * kTile;
    const u32 i = t0 + threadIdx.x;
    const SceneSeg& G = P.seg[scene_of_slot(P, t0)];
    const u32 l = i - G.off;
    const float eps = G.eps;
    if (!(eps > 0.0f)) {  // scene-uniform
        if (l < G.n) vout[i] = vin[i];
        return;
    }
    const StageRanges R = L.plan[tile];
    // this slot's own inputs, loaded with the staging loads (slots
    // [0, nslots) are allocated)
    const float4 vi = vin[i];
    const float4 p = xs[i];
    const float rho_i = rho[i];
    const u32 cf = L.cnt[i] & 0xFFFFu;
    if (R.total) {
        // every load of a thread is issued before its LDS stores
        const StageLanes E(R);
        f4v x[kStagePer], v[kStagePer];
#pragma unroll
        for (int u = 0; u < kStagePer; ++u) {
            x[u] = ld4v(xs, E.src[u]);
            v[u] = ld4v(vin, E.src[u]);
        }
#pragma unroll
        for (int u = 0; u < kStagePer; ++u)
            if (E.has(u)) {
                reinterpret_cast<f4v*>(s_a)[E.entry(u)] = f4v{x[u].x, x[u].y, x[u].z, v[u].x};
                reinterpret_cast<f2v*>(s_b)[E.entry(u)] = f2v{v[u].y, v[u].z};
            }
    }
    __syncthreads();
    if (l >= G.n) return;
    float dx = 0.0f, dy = 0.0f, dz = 0.0f;
    float tx = 0.0f, ty = 0.0f, tz = 0.0f;  // tolerance mode: see xsph_listed
    if (cf <= L.kf) {
        u32 k = 0;
        if (R.total) {
            for_stage_quads(L, i, cf, [&](u32 a0, u32 a1, u32 a2, u32 a3, u32 m) {
                const float r0 = rho[stage_source(R, a0)], r1 = rho[stage_source(R, a1)];
                const float r2 = rho[stage_source(R, a2)], r3 = rho[stage_source(R, a3)];
                const float4 q0 = s_a[a0], q1 = s_a[a1], q2 = s_a[a2], q3 = s_a[a3];
                const float2 b0 = s_b[a0], b1 = s_b[a1], b2 = s_b[a2], b3 = s_b[a3];
                const float4 w0 = make_float4(q0.w, b0.x, b0.y, 0.0f);
                const float4 w1 = make_float4(q1.w, b1.x, b1.y, 0.0f);
                const float4 w2 = make_float4(q2.w, b2.x, b2.y, 0.0f);
                const float4 w3 = make_float4(q3.w, b3.x, b3.y, 0.0f);
                xsph_listed<F>(dx, dy, dz, tx, ty, tz, p, vi, rho_i, q0, w0, r0, eps, P);
                xsph_listed<F>(dx, dy, dz, tx, ty, tz, p, vi, rho_i, q1, w1, r1, eps, P, m > 1u);
                xsph_listed<F>(dx, dy, dz, tx, ty, tz, p, vi, rho_i, q2, w2, r2, eps, P, m > 2u);
                xsph_listed<F>(dx, dy, dz, tx, ty, tz, p, vi, rho_i, q3, w3, r3, eps, P, m > 3u);
            });
        } else {
            // four pairs per trip: their gathers are issued together, the
            // terms added in list order
            for (; k + 4u <= cf; k += 4u) {
                u32 j[4];
                float4 q[4], w[4];
                float rj[4];
#pragma unroll
                for (int u = 0; u < 4; ++u) j[u] = ld_pair(L.f + pair_at(k + u, i, L.kf)).x;
#pragma unroll
                for (int u = 0; u < 4; ++u) {
                    q[u] = xs[j[u]];
                    w[u] = vin[j[u]];
                    rj[u] = rho[j[u]];
                }
#pragma unroll
                for (int u = 0; u < 4; ++u)
                    xsph_listed<F>(dx, dy, dz, tx, ty, tz, p, vi, rho_i, q[u], w[u], rj[u], eps, P);
            }
            for (; k < cf; ++k) {
                const u32 j = ld_pair(L.f + pair_at(k, i, L.kf)).x;
                xsph_listed<F>(dx, dy, dz, tx, ty, tz, p, vi, rho_i, xs[j], vin[j], rho[j], eps, P);
            }
        }
        xsph_listed_scale<F>(dx, dy, dz, tx, ty, tz, eps, P);
        vout[i] = make_float4(vi.x + dx, vi.y + dy, vi.z + dz, 0.0f);
        return;
    }
    const u32* cs = P.csf + G.key_base;
    const u32 shift = G.off - G.start;
    for_stencil(P.gf, cs, shift, p.x, p.y, p.z, [&](u32 s) {
        if (s == i) return;
        const float4 q = xs[s];
        const float r2 = PairMath<F>::r2(p.x - q.x, p.y - q.y, p.z - q.z);
        if (r2 >= P.sp.far2) return;  // W = 0
        xsph_term_r2<F>(dx, dy, dz, r2, vi, rho_i, vin[s], rho[s], eps, P);
    });
    vout[i] = make_float4(vi.x + dx, vi.y + dy, vi.z + dz, 0.0f);
}

// ---- surface tension ------------------------------------------------------
// Akinci et al. 2013 cohesion + curvature after XSPH, with the reference's
// cohesion kernel C(r, L) (sph_kernels.cuh:44-51) on the kernel support
// L = 2l; the reference parses the coefficient (loader.cpp:50,
// main.cpp:132) but applies no force.  The powers of L are evaluated once
// on the host.
//   n_i = L sum_{j != i} (m / rho_j) grad W_ij
//   a_i = sum_{j != i, |r| < L} K_ij (s_ij r_ij - gamma (n_i - n_j)),
//   K_ij = 2 rho0 / (rho_i + rho_j),  s_ij = -gamma m C(|r|) / |r| (0 at r = 0)
//   v_i += dt a_i

__device__ __forceinline__ float cohesion_C(float r, const TensionArgs& T) {
    if (r <= 0.0f || r > T.L) return 0.0f;
    const float lm = T.L - r;
    const float t = lm * lm * lm * r * r * r;
    if (2.0f * r > T.L) return T.sig * t;
    return T.sig * (2.0f * t - T.c6);
}

template <int NT>
__global__ __launch_bounds__(NT) void k_tension_normals(Nbr P, PairList L,
                                                        const float4* __restrict__ xs,
                                                        const float* __restrict__ rho,
                                                        TensionArgs T, float4* __restrict__ nrm) {
    const u32 i = slot_index<NT>(P.remap);
    if (i >= P.nslots) return;
    const SceneSeg& G = P.seg[scene_of_slot(P, i)];
    if (i - G.off >= G.n) return;
    const float4 p = xs[i];
    float sx = 0.0f, sy = 0.0f, sz = 0.0f;
    // zero-gradient pairs add +-0 to sums that never hold -0: left out
    auto add = [&](const float4& q, float c, float rj) {
        const float gx = (p.x - q.x) * c, gy = (p.y - q.y) * c, gz = (p.z - q.z) * c;
        const float f = P.mass / rj;
        sx += f * gx;
        sy += f * gy;
        sz += f * gz;
    };
    const u32 cf = L.cnt[i] & 0xFFFFu;
    if (cf <= L.kf) {
        const StageRanges R = L.plan[i / kTile];
        for (u32 k = 0; k < cf; ++k) {
            u32 j;
            float c;
            fluid_entry(L, R, k, i, j, c);
            const float4 q = xs[j];
            add(q, R.total ? pair_c(p, q, P.sp) : c, rho[j]);
        }
    } else {
        for_stencil(P.gf, P.csf + G.key_base, G.off - G.start, p.x, p.y, p.z, [&](u32 s) {
            if (s == i) return;
            const float4 q = xs[s];
            const float r2 = dist2(p.x - q.x, p.y - q.y, p.z - q.z);
            if (r2 >= P.sp.far2) return;
            const float c = spline_gradc(pair_sqrt(r2, P.sp), P.sp);
            if (c != 0.0f) add(q, c, rho[s]);
        });
    }
    nrm[i] = make_float4(T.L * sx, T.L * sy, T.L * sz, 0.0f);
}

template <int NT>
__global__ __launch_bounds__(NT) void k_tension_force(Nbr P, PairList L,
                                                      const float4* __restrict__ xs,
                                                      const float* __restrict__ rho,
                                                      const float4* __restrict__ nrm,
                                                      TensionArgs T, float4* __restrict__ vs) {
    const u32 i = slot_index<NT>(P.remap);
    if (i >= P.nslots) return;
    const SceneSeg& G = P.seg[scene_of_slot(P, i)];
    if (i - G.off >= G.n) return;
    const float4 p = xs[i];
    const float4 ni = nrm[i];
    const float rho_i = rho[i];
    const float mg = -T.gamma * P.mass;
    float ax = 0.0f, ay = 0.0f, az = 0.0f;
    auto add = [&](const float4& q, float rj, const float4& nj) {
        const float rx = p.x - q.x, ry = p.y - q.y, rz = p.z - q.z;
        const float d = sqrtf(dist2(rx, ry, rz));
        const float K = 2.0f * T.rho0 / (rho_i + rj);
        const float sc = d > 0.0f ? mg * cohesion_C(d, T) / d : 0.0f;
        ax += K * (sc * rx + -T.gamma * (ni.x - nj.x));
        ay += K * (sc * ry + -T.gamma * (ni.y - nj.y));
        az += K * (sc * rz + -T.gamma * (ni.z - nj.z));
    };
    const u32 cf = L.cnt[i] & 0xFFFFu;
    if (cf <= L.kf) {
        // a complete list holds every in-support pair but the self pair
        const StageRanges R = L.plan[i / kTile];
        for (u32 k = 0; k < cf; ++k) {
            u32 j;
            float c;
            fluid_entry(L, R, k, i, j, c);
            add(xs[j], rho[j], nrm[j]);
        }
    } else {
        for_stencil(P.gf, P.csf + G.key_base, G.off - G.start, p.x, p.y, p.z, [&](u32 s) {
            if (s == i) return;
            const float4 q = xs[s];
            if (dist2(p.x - q.x, p.y - q.y, p.z - q.z) >= P.sp.far2) return;
            add(q, rho[s], nrm[s]);
        });
    }
    const float4 v = vs[i];
    const float dt = G.dt;
    vs[i] = make_float4(v.x + dt * ax, v.y + dt * ay, v.z + dt * az, 0.0f);
}

// rho_dot terms (dfsph_kernels.cu:210, 222) for one pair with gradient scalar c.
// v . Gb (tolerance mode's boundary term of the density change)
__device__ __forceinline__ float dot_gb(const float4& v, const float4& g) {
    return __builtin_fmaf(v.x, g.x, __builtin_fmaf(v.y, g.y, v.z * g.z));
}

template <bool F>
__device__ __forceinline__ void rho_dot_fluid(float& rd, const float4& p, const float4& vi,
                                              const float4& q, const float4& vj, float c,
                                              float mass, bool keep = true) {
    const float ux = vi.x - vj.x, uy = vi.y - vj.y, uz = vi.z - vj.z;
    if (F) {  // m (u . r) c
        const float dot = __builtin_fmaf(ux, p.x - q.x,
                                         __builtin_fmaf(uy, p.y - q.y, uz * (p.z - q.z)));
        rd = __builtin_fmaf(keep ? mass * c : 0.0f, dot, rd);
    } else {
        const float gx = (p.x - q.x) * c, gy = (p.y - q.y) * c, gz = (p.z - q.z) * c;
        rd += keep ? mass * (ux * gx + uy * gy + uz * gz) : -0.0f;  // -0.0f: no change
    }
}
template <bool F>
__device__ __forceinline__ void rho_dot_boundary(float& rd, const float4& p, const float4& vi,
                                                 const float4& q, float c) {
    if (F) {
        const float dot = __builtin_fmaf(vi.x, p.x - q.x,
                                         __builtin_fmaf(vi.y, p.y - q.y, vi.z * (p.z - q.z)));
        rd = __builtin_fmaf(q.w * c, dot, rd);
    } else {
        const float gx = (p.x - q.x) * c, gy = (p.y - q.y) * c, gz = (p.z - q.z) * c;
        rd += q.w * (vi.x * gx + vi.y * gy + vi.z * gz);
    }
}

// One predictor sweep.  MODE 0: density change + kappa (dfsph_kernels.cu:
// 176-232).  MODE 1: divergence change + kappa_v (docs/superpowers/plans/2026-04-24-water-simulation-v2-phase2-dfsph-solver.md:1862-1907, with
// the boundary term of the density change).  Writes (x, kappa-hat) for the
// following k_apply and one binary64 error partial per tile: the four waves
// of a tile are summed in a fixed order, so the partials do not depend on
// the workgroup size.  Scenes that converged before `iter` skip the work.
template <int MODE, int NT, int PU, bool F>
__global__ __launch_bounds__(NT) void k_change(Nbr P, PairList L, const float4* __restrict__ xs,
                                               const float4* __restrict__ vs,
                                               const float* __restrict__ rho,
                                               const float* __restrict__ alpha, float rho0,
                                               int pform, u32 iter, const DeviceState* st,
                                               double* __restrict__ part,
                                               float* __restrict__ rho_adv,
                                               float* __restrict__ kappa,
                                               float4* __restrict__ xk) {
    static_assert(NT % kTile == 0, "k_change workgroups hold whole tiles");
    __shared__ double s_red[NT / 64];
    const u32 i = slot_index<NT>(P.remap);
    double contrib = 0.0;
    bool active = false;  // wave-uniform
    if (i < P.nslots) {
        const u32 sc = scene_of_slot(P, i);
        const u32 conv = MODE == 0 ? st[sc].conv_iter : st[sc].div_conv_iter;
        active = !(conv < iter);
        const SceneSeg& G = P.seg[sc];
        if (active && i - G.off < G.n) {
            const u32* cs = P.csf + G.key_base;
            const u32 shift = G.off - G.start;
            const float4 p = xs[i];
            const float4 vi = vs[i];
            float rd = 0.0f;
            // tolerance mode: the fluid terms sum c (u . r) in rf, times m once
            float rf = 0.0f;
            float& racc = F ? rf : rd;
            const float mf = F ? 1.0f : P.mass;
            const u32 cw = L.cnt[i];
            const u32 cf = cw & 0xFFFFu, cb = cw >> 16;
            if (cf <= L.kf && cb <= L.kb) {
                // PU pairs per trip: loads issued together, terms summed in order
                const StageRanges R = L.plan[i / kTile];
                u32 k = 0;
                for (; k + PU <= cf; k += PU) {
                    u32 j[PU];
                    float c[PU];
                    float4 q[PU], w[PU];
#pragma unroll
                    for (int u = 0; u < PU; ++u) fluid_entry(L, R, k + u, i, j[u], c[u]);
#pragma unroll
                    for (int u = 0; u < PU; ++u) {
                        q[u] = xs[j[u]];
                        w[u] = vs[j[u]];
                    }
#pragma unroll
                    for (int u = 0; u < PU; ++u)
                        rho_dot_fluid<F>(racc, p, vi, q[u], w[u], R.total ? pair_c_listed<F>(p, q[u], P.sp) : c[u],
                                      mf);
                }
                for (; k < cf; ++k) {
                    u32 j;
                    float c;
                    fluid_entry(L, R, k, i, j, c);
                    const float4 q = xs[j];
                    rho_dot_fluid<F>(racc, p, vi, q, vs[j], R.total ? pair_c_listed<F>(p, q, P.sp) : c, mf);
                }
                k = 0;
                for (; k + PU <= cb; k += PU) {
                    uint2 e[PU];
                    float4 q[PU];
#pragma unroll
                    for (int u = 0; u < PU; ++u) e[u] = ld_pair(L.b + pair_at(k + u, i, L.kb));
#pragma unroll
                    for (int u = 0; u < PU; ++u) q[u] = P.xb[e[u].x];
#pragma unroll
                    for (int u = 0; u < PU; ++u)
                        rho_dot_boundary<F>(rd, p, vi, q[u], __uint_as_float(e[u].y));
                }
                for (; k < cb; ++k) {
                    const uint2 e = ld_pair(L.b + pair_at(k, i, L.kb));
                    rho_dot_boundary<F>(rd, p, vi, P.xb[e.x], __uint_as_float(e.y));
                }
            } else {
                for_stencil(P.gf, cs, shift, p.x, p.y, p.z, [&](u32 s) {
                    if (s == i) return;
                    const float4 q = xs[s];
                    const float r2 = PairMath<F>::r2(p.x - q.x, p.y - q.y, p.z - q.z);
                    if (r2 >= P.sp.far2) return;
                    const float c = stencil_c<F>(r2, P.sp);
                    rho_dot_fluid<F>(racc, p, vi, q, vs[s], c, mf);
                });
                if (P.nb && !(F && L.gb)) {
                    for_stencil(P.gb, P.csb, 0u, p.x, p.y, p.z, [&](u32 s) {
                        const float4 q = P.xb[s];
                        const float r2 = PairMath<F>::r2(p.x - q.x, p.y - q.y, p.z - q.z);
                        if (r2 >= P.sp.far2) return;
                        const float c = stencil_c<F>(r2, P.sp);
                        rho_dot_boundary<F>(rd, p, vi, q, c);
                    });
                }
            }
            if (F) rd = __builtin_fmaf(P.mass, rf, rd);
            if (F && L.gb) rd += dot_gb(vi, L.gb[i]);
            const float rho_i = rho[i];
            float kap, err;
            if (MODE == 0) {
                const float rho_pred = rho_i + G.dt * rd;
                rho_adv[i] = rho_pred;
                err = fmaxf(0.0f, rho_pred - rho0);
                kap = err * alpha[i] / G.dt2;
            } else {
                rho_adv[i] = rd;
                err = fmaxf(0.0f, rd);
                kap = err * alpha[i] / G.dt;
            }
            kappa[i] = kap;
            const float khat =
                pform ? kap / fmaxf(rho_i, 1e-9f) : kap / fmaxf(rho_i * rho_i, 1e-9f);
            xk[i] = make_float4(p.x, p.y, p.z, khat);
            contrib = static_cast<double>(err);
        }
    }
    // tile partial; the next launch (k_apply) finishes the scene's sum, so no
    // cross-workgroup fence is needed inside this kernel
    const double ws = wave_sum(contrib);
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    if (lane == 0) s_red[w] = ws;
    __syncthreads();
    if (lane == 0 && (w & 3) == 0 && active)
        part[i / kTile] = ((s_red[w] + s_red[w + 1]) + s_red[w + 2]) + s_red[w + 3];
}

// k_change with the tile's fluid neighbourhood staged in LDS (positions and
// velocities); the same arithmetic in the same order as k_change.
template <int MODE, int PU, bool F>
__global__ __launch_bounds__(kTile) void k_change_st(Nbr P, PairList L,
                                                     const float4* __restrict__ xs,
                                                     const float4* __restrict__ vs,
                                                     const float* __restrict__ rho,
                                                     const float* __restrict__ alpha, float rho0,
                                                     int pform, u32 iter, const DeviceState* st,
                                                     double* __restrict__ part,
                                                     float* __restrict__ rho_adv,
                                                     float* __restrict__ kappa,
                                                     float4* __restrict__ xk) {
    __shared__ float4 s_a[kStageCap];  // x, y, z, vx
    __shared__ float2 s_b[kStageCap];  // vy, vz
    __shared__ double s_red[kTile / 64];
    const u32 tile = P.remap ? xcd_tile(blockIdx.x, gridDim.x) : blockIdx.x;
    const u32 t0 = tile * kTile;
    const u32 i = t0 + threadIdx.x;
    const u32 sc = scene_of_slot(P, t0);
    // the plan's load is issued with the convergence test's
    const StageRanges R = L.plan[tile];
    const u32 conv = MODE == 0 ? st[sc].conv_iter : st[sc].div_conv_iter;
    if (conv < iter) return;  // the whole tile belongs to one scene
    const SceneSeg& G = P.seg[sc];
    // this slot's own inputs, loaded with the staging loads (slots
    // [0, nslots) are allocated)
    const float4 p = xs[i];
    const float4 vi = vs[i];
    const u32 cw = L.cnt[i];
    if (R.total) {
        // every load of a thread is issued before its LDS stores
        const StageLanes E(R);
        f4v x[kStagePer], v[kStagePer];
#pragma unroll
        for (int u = 0; u < kStagePer; ++u) {
            x[u] = ld4v(xs, E.src[u]);
            v[u] = ld4v(vs, E.src[u]);
        }
#pragma unroll
        for (int u = 0; u < kStagePer; ++u)
            if (E.has(u)) {
                reinterpret_cast<f4v*>(s_a)[E.entry(u)] = f4v{x[u].x, x[u].y, x[u].z, v[u].x};
                reinterpret_cast<f2v*>(s_b)[E.entry(u)] = f2v{v[u].y, v[u].z};
            }
    }
    __syncthreads();
    const bool staged = R.total != 0;
    double contrib = 0.0;
    if (i - G.off < G.n) {
        const u32* cs = P.csf + G.key_base;
        const u32 shift = G.off - G.start;
        float rd = 0.0f;
        // tolerance mode: the fluid terms sum c (u . r) in rf, times m once
        float rf = 0.0f;
        float& racc = F ? rf : rd;
        const float mf = F ? 1.0f : P.mass;
        const u32 cf = cw & 0xFFFFu, cb = cw >> 16;
        if (cf <= L.kf && cb <= L.kb) {
            u32 k = 0;
            if (staged) {
                // compact list: LDS entries only; c is recomputed from the staged
                // positions with the expression that produced it
                auto term = [&](const float4& q, const float4& w, bool keep) {
                    rho_dot_fluid<F>(racc, p, vi, q, w, pair_c_listed<F>(p, q, P.sp), mf, keep);
                };
                auto at = [&](u32 a, float4& q, float4& w) {
                    const float4 x = s_a[a];
                    const float2 y = s_b[a];
                    q = make_float4(x.x, x.y, x.z, 0.0f);
                    w = make_float4(x.w, y.x, y.y, 0.0f);
                };
                for_stage_quads(L, i, cf, [&](u32 a0, u32 a1, u32 a2, u32 a3, u32 m) {
                    float4 q0, q1, q2, q3, w0, w1, w2, w3;
                    at(a0, q0, w0);
                    at(a1, q1, w1);
                    at(a2, q2, w2);
                    at(a3, q3, w3);
                    term(q0, w0, true);
                    term(q1, w1, m > 1u);
                    term(q2, w2, m > 2u);
                    term(q3, w3, m > 3u);
                });
            } else {
                for (; k + PU <= cf; k += PU) {
                    uint2 e[PU];
                    float4 q[PU], w[PU];
#pragma unroll
                    for (int u = 0; u < PU; ++u) e[u] = ld_pair(L.f + pair_at(k + u, i, L.kf));
#pragma unroll
                    for (int u = 0; u < PU; ++u) {
                        q[u] = xs[e[u].x];
                        w[u] = vs[e[u].x];
                    }
#pragma unroll
                    for (int u = 0; u < PU; ++u)
                        rho_dot_fluid<F>(racc, p, vi, q[u], w[u], __uint_as_float(e[u].y), mf);
                }
                for (; k < cf; ++k) {
                    const uint2 e = ld_pair(L.f + pair_at(k, i, L.kf));
                    rho_dot_fluid<F>(racc, p, vi, xs[e.x], vs[e.x], __uint_as_float(e.y), mf);
                }
            }
            k = 0;
            for (; k + PU <= cb; k += PU) {
                uint2 e[PU];
                float4 q[PU];
#pragma unroll
                for (int u = 0; u < PU; ++u) e[u] = ld_pair(L.b + pair_at(k + u, i, L.kb));
#pragma unroll
                for (int u = 0; u < PU; ++u) q[u] = P.xb[e[u].x];
#pragma unroll
                for (int u = 0; u < PU; ++u)
                    rho_dot_boundary<F>(rd, p, vi, q[u], __uint_as_float(e[u].y));
            }
            for (; k < cb; ++k) {
                const uint2 e = ld_pair(L.b + pair_at(k, i, L.kb));
                rho_dot_boundary<F>(rd, p, vi, P.xb[e.x], __uint_as_float(e.y));
            }
        } else {
            for_stencil(P.gf, cs, shift, p.x, p.y, p.z, [&](u32 s) {
                if (s == i) return;
                const float4 q = xs[s];
                const float r2 = PairMath<F>::r2(p.x - q.x, p.y - q.y, p.z - q.z);
                if (r2 >= P.sp.far2) return;
                const float c = stencil_c<F>(r2, P.sp);
                rho_dot_fluid<F>(racc, p, vi, q, vs[s], c, mf);
            });
            if (P.nb && !(F && L.gb)) {
                for_stencil(P.gb, P.csb, 0u, p.x, p.y, p.z, [&](u32 s) {
                    const float4 q = P.xb[s];
                    const float r2 = PairMath<F>::r2(p.x - q.x, p.y - q.y, p.z - q.z);
                    if (r2 >= P.sp.far2) return;
                    const float c = stencil_c<F>(r2, P.sp);
                    rho_dot_boundary<F>(rd, p, vi, q, c);
                });
            }
        }
        if (F) rd = __builtin_fmaf(P.mass, rf, rd);
        if (F && L.gb) rd += dot_gb(vi, L.gb[i]);
        const float rho_i = rho[i];
        float kap, err;
        if (MODE == 0) {
            const float rho_pred = rho_i + G.dt * rd;
            rho_adv[i] = rho_pred;
            err = fmaxf(0.0f, rho_pred - rho0);
            kap = err * alpha[i] / G.dt2;
        } else {
            rho_adv[i] = rd;
            err = fmaxf(0.0f, rd);
            kap = err * alpha[i] / G.dt;
        }
        kappa[i] = kap;
        const float khat = pform ? kap / fmaxf(rho_i, 1e-9f) : kap / fmaxf(rho_i * rho_i, 1e-9f);
        xk[i] = make_float4(p.x, p.y, p.z, khat);
        contrib = static_cast<double>(err);
    }
    const double ws = wave_sum(contrib);
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    if (lane == 0) s_red[w] = ws;
    __syncthreads();
    if (threadIdx.x == 0) part[tile] = ((s_red[0] + s_red[1]) + s_red[2]) + s_red[3];
}

// The first wave of each scene in k_apply(iter) sums the scene's tile
// partials written by k_change(iter), in a fixed order, and records the
// error and the convergence decision (dfsph.cu:254-258; the divergence solve
// tests every iteration).
__device__ __forceinline__ void finish_error(int mode, u32 iter, double thr, const SceneSeg& G,
                                             const double* __restrict__ part, DeviceState* st) {
    const int lane = threadIdx.x & 63;
    double s = 0.0;
#pragma unroll 8
    for (u32 t = lane; t < G.ntiles; t += 64) s += part[G.tile0 + t];
    const double tot = wave_sum(s);
    if (lane == 0) {
        const double e = tot / static_cast<double>(G.n);
        if (mode == 0) {
            st->err = e;
            if (iter >= 1 && e < thr && st->conv_iter == kNone) st->conv_iter = iter;
        } else {
            st->div_err = e;
            if (e < thr && st->div_conv_iter == kNone) st->div_conv_iter = iter;
        }
    }
}

// Pressure velocity terms (dfsph_kernels.cu:275-278, 291-294) for one pair.
template <bool F>
__device__ __forceinline__ void kappa_term(float& dx, float& dy, float& dz, const float4& p,
                                           const float4& q, float c, float c2, float dt,
                                           bool keep = true) {
    if (F) {  // -dt c2 c r
        const float k = keep ? -(dt * c2 * c) : 0.0f;
        dx = __builtin_fmaf(k, p.x - q.x, dx);
        dy = __builtin_fmaf(k, p.y - q.y, dy);
        dz = __builtin_fmaf(k, p.z - q.z, dz);
    } else {
        const float gx = (p.x - q.x) * c, gy = (p.y - q.y) * c, gz = (p.z - q.z) * c;
        // a dropped term subtracts +0.0f, which leaves every value unchanged
        dx -= keep ? dt * c2 * gx : 0.0f;
        dy -= keep ? dt * c2 * gy : 0.0f;
        dz -= keep ? dt * c2 * gz : 0.0f;
    }
}

// Pressure velocity update (dfsph_kernels.cu:237-301); kappa-hat of every
// particle was precomputed by k_change with the same expression the
// reference evaluates per pair.
template <int NT, int PU, bool F>
__global__ __launch_bounds__(NT) void k_apply(Nbr P, PairList L, const float4* __restrict__ xk,
                                              const float* __restrict__ kappa, float rho0,
                                              int pform, u32 iter, int mode, double thr,
                                              const double* __restrict__ part, DeviceState* st,
                                              const float4* vin, float4* vout) {
    const u32 i = slot_index<NT>(P.remap);
    if (i >= P.nslots) return;
    const u32 sc = scene_of_slot(P, i);
    DeviceState* sst = st + sc;
    const u32 conv = mode == 0 ? sst->conv_iter : sst->div_conv_iter;
    if (conv < iter) return;  // scene-uniform: this iteration's k_change did not run either
    const SceneSeg& G = P.seg[sc];
    const u32 l = i - G.off;
    if (l < 64) finish_error(mode, iter, thr, G, part, sst);
    if (l >= G.n) return;
    const u32* cs = P.csf + G.key_base;
    const u32 shift = G.off - G.start;
    const float dt = G.dt;
    const float4 p = xk[i];
    const float ki = p.w;
    float dx = 0.0f, dy = 0.0f, dz = 0.0f;
    // tolerance mode: the fluid terms sum -(k_i + k_j) c r in a, times dt m once
    float ax = 0.0f, ay = 0.0f, az = 0.0f;
    float& fx = F ? ax : dx;
    float& fy = F ? ay : dy;
    float& fz = F ? az : dz;
    const float fdt = F ? 1.0f : dt;
    auto mk = [&](float kj) { return F ? ki + kj : P.mass * (ki + kj); };
    float bf = 0.0f;
    if (P.nb) {
        const float k_i = kappa[i];
        bf = pform ? k_i / fmaxf(rho0, 1e-9f) : k_i / fmaxf(rho0 * rho0, 1e-9f);
    }
    const u32 cw = L.cnt[i];
    const u32 cf = cw & 0xFFFFu, cb = cw >> 16;
    if (cf <= L.kf && cb <= L.kb) {
        const StageRanges R = L.plan[i / kTile];
        u32 k = 0;
        for (; k + PU <= cf; k += PU) {
            u32 j[PU];
            float c[PU];
            float4 q[PU];
#pragma unroll
            for (int u = 0; u < PU; ++u) fluid_entry(L, R, k + u, i, j[u], c[u]);
#pragma unroll
            for (int u = 0; u < PU; ++u) q[u] = xk[j[u]];
#pragma unroll
            for (int u = 0; u < PU; ++u)
                kappa_term<F>(fx, fy, fz, p, q[u], R.total ? pair_c_listed<F>(p, q[u], P.sp) : c[u],
                           mk(q[u].w), fdt);
        }
        for (; k < cf; ++k) {
            u32 j;
            float c;
            fluid_entry(L, R, k, i, j, c);
            const float4 q = xk[j];
            kappa_term<F>(fx, fy, fz, p, q, R.total ? pair_c_listed<F>(p, q, P.sp) : c, mk(q.w), fdt);
        }
        k = 0;
        for (; k + PU <= cb; k += PU) {
            uint2 e[PU];
            float4 q[PU];
#pragma unroll
            for (int u = 0; u < PU; ++u) e[u] = ld_pair(L.b + pair_at(k + u, i, L.kb));
#pragma unroll
            for (int u = 0; u < PU; ++u) q[u] = P.xb[e[u].x];
#pragma unroll
            for (int u = 0; u < PU; ++u)
                kappa_term<F>(dx, dy, dz, p, q[u], __uint_as_float(e[u].y), q[u].w * (ki + bf), dt);
        }
        for (; k < cb; ++k) {
            const uint2 e = ld_pair(L.b + pair_at(k, i, L.kb));
            const float4 q = P.xb[e.x];
            kappa_term<F>(dx, dy, dz, p, q, __uint_as_float(e.y), q.w * (ki + bf), dt);
        }
    } else {
        for_stencil(P.gf, cs, shift, p.x, p.y, p.z, [&](u32 s) {
            if (s == i) return;
            const float4 q = xk[s];
            const float r2 = PairMath<F>::r2(p.x - q.x, p.y - q.y, p.z - q.z);
            if (r2 >= P.sp.far2) return;
            const float c = stencil_c<F>(r2, P.sp);
            kappa_term<F>(fx, fy, fz, p, q, c, mk(q.w), fdt);
        });
        if (P.nb && !(F && L.gb)) {
            for_stencil(P.gb, P.csb, 0u, p.x, p.y, p.z, [&](u32 s) {
                const float4 q = P.xb[s];
                const float r2 = PairMath<F>::r2(p.x - q.x, p.y - q.y, p.z - q.z);
                if (r2 >= P.sp.far2) return;
                const float c = stencil_c<F>(r2, P.sp);
                kappa_term<F>(dx, dy, dz, p, q, c, q.w * (ki + bf), dt);
            });
        }
    }
    if (F) {
        const float sc = dt * P.mass;
        dx = __builtin_fmaf(sc, ax, dx);
        dy = __builtin_fmaf(sc, ay, dy);
        dz = __builtin_fmaf(sc, az, dz);
    }
    if (F && L.gb) {  // sum_b m_b (k_i + k_b) grad W_ib = (k_i + k_b) Gb_i
        const float4 g = L.gb[i];
        const float k = -(dt * (ki + bf));
        dx = __builtin_fmaf(k, g.x, dx);
        dy = __builtin_fmaf(k, g.y, dy);
        dz = __builtin_fmaf(k, g.z, dz);
    }
    float4 v = vin[i];
    v.x += dx;
    v.y += dy;
    v.z += dz;
    vout[i] = v;
}

// k_apply with the tile's fluid (x, kappa-hat) neighbourhood staged in LDS.
template <int PU, bool F>
__global__ __launch_bounds__(kTile) void k_apply_st(Nbr P, PairList L,
                                                    const float4* __restrict__ xk,
                                                    const float* __restrict__ kappa, float rho0,
                                                    int pform, u32 iter, int mode, double thr,
                                                    const double* __restrict__ part,
                                                    DeviceState* st, const float4* vin,
                                                    float4* vout) {
    __shared__ float4 s_x[kStageCap];
    const u32 tile = P.remap ? xcd_tile(blockIdx.x, gridDim.x) : blockIdx.x;
    const u32 t0 = tile * kTile;
    const u32 i = t0 + threadIdx.x;
    const u32 sc = scene_of_slot(P, t0);
    DeviceState* sst = st + sc;
    // the plan's load is issued with the convergence test's
    const StageRanges R = L.plan[tile];
    const u32 conv = mode == 0 ? sst->conv_iter : sst->div_conv_iter;
    if (conv < iter) return;  // tile-uniform
    const SceneSeg& G = P.seg[sc];
    // this slot's own inputs, loaded with the staging loads (slots
    // [0, nslots) are allocated)
    const float4 p = xk[i];
    const u32 cw = L.cnt[i];
    if (R.total) {
        // every load of a thread is issued before its LDS stores
        const StageLanes E(R);
        f4v x[kStagePer];
#pragma unroll
        for (int u = 0; u < kStagePer; ++u) x[u] = ld4v(xk, E.src[u]);
#pragma unroll
        for (int u = 0; u < kStagePer; ++u)
            if (E.has(u)) reinterpret_cast<f4v*>(s_x)[E.entry(u)] = x[u];
    }
    __syncthreads();
    const bool staged = R.total != 0;
    const u32 l = i - G.off;
    if (l < 64) finish_error(mode, iter, thr, G, part, sst);
    if (l >= G.n) return;
    const u32* cs = P.csf + G.key_base;
    const u32 shift = G.off - G.start;
    const float dt = G.dt;
    const float ki = p.w;
    float dx = 0.0f, dy = 0.0f, dz = 0.0f;
    // tolerance mode: the fluid terms sum -(k_i + k_j) c r in a, times dt m once
    float ax = 0.0f, ay = 0.0f, az = 0.0f;
    float& fx = F ? ax : dx;
    float& fy = F ? ay : dy;
    float& fz = F ? az : dz;
    const float fdt = F ? 1.0f : dt;
    auto mk = [&](float kj) { return F ? ki + kj : P.mass * (ki + kj); };
    float bf = 0.0f;
    if (P.nb) {
        const float k_i = kappa[i];
        bf = pform ? k_i / fmaxf(rho0, 1e-9f) : k_i / fmaxf(rho0 * rho0, 1e-9f);
    }
    const u32 cf = cw & 0xFFFFu, cb = cw >> 16;
    if (cf <= L.kf && cb <= L.kb) {
        u32 k = 0;
        if (staged) {
            auto term = [&](const float4& q, bool keep) {
                kappa_term<F>(fx, fy, fz, p, q, pair_c_listed<F>(p, q, P.sp), mk(q.w), fdt, keep);
            };
            for_stage_quads(L, i, cf, [&](u32 a0, u32 a1, u32 a2, u32 a3, u32 m) {
                const float4 q0 = s_x[a0], q1 = s_x[a1], q2 = s_x[a2], q3 = s_x[a3];
                term(q0, true);
                term(q1, m > 1u);
                term(q2, m > 2u);
                term(q3, m > 3u);
            });
        } else {
            for (; k + PU <= cf; k += PU) {
                uint2 e[PU];
                float4 q[PU];
#pragma unroll
                for (int u = 0; u < PU; ++u) e[u] = ld_pair(L.f + pair_at(k + u, i, L.kf));
#pragma unroll
                for (int u = 0; u < PU; ++u) q[u] = xk[e[u].x];
#pragma unroll
                for (int u = 0; u < PU; ++u)
                    kappa_term<F>(fx, fy, fz, p, q[u], __uint_as_float(e[u].y), mk(q[u].w), fdt);
            }
            for (; k < cf; ++k) {
                const uint2 e = ld_pair(L.f + pair_at(k, i, L.kf));
                const float4 q = xk[e.x];
                kappa_term<F>(fx, fy, fz, p, q, __uint_as_float(e.y), mk(q.w), fdt);
            }
        }
        k = 0;
        for (; k + PU <= cb; k += PU) {
            uint2 e[PU];
            float4 q[PU];
#pragma unroll
            for (int u = 0; u < PU; ++u) e[u] = ld_pair(L.b + pair_at(k + u, i, L.kb));
#pragma unroll
            for (int u = 0; u < PU; ++u) q[u] = P.xb[e[u].x];
#pragma unroll
            for (int u = 0; u < PU; ++u)
                kappa_term<F>(dx, dy, dz, p, q[u], __uint_as_float(e[u].y), q[u].w * (ki + bf), dt);
        }
        for (; k < cb; ++k) {
            const uint2 e = ld_pair(L.b + pair_at(k, i, L.kb));
            const float4 q = P.xb[e.x];
            kappa_term<F>(dx, dy, dz, p, q, __uint_as_float(e.y), q.w * (ki + bf), dt);
        }
    } else {
        for_stencil(P.gf, cs, shift, p.x, p.y, p.z, [&](u32 s) {
            if (s == i) return;
            const float4 q = xk[s];
            const float r2 = PairMath<F>::r2(p.x - q.x, p.y - q.y, p.z - q.z);
            if (r2 >= P.sp.far2) return;
            const float c = stencil_c<F>(r2, P.sp);
            kappa_term<F>(fx, fy, fz, p, q, c, mk(q.w), fdt);
        });
        if (P.nb && !(F && L.gb)) {
            for_stencil(P.gb, P.csb, 0u, p.x, p.y, p.z, [&](u32 s) {
                const float4 q = P.xb[s];
                const float r2 = PairMath<F>::r2(p.x - q.x, p.y - q.y, p.z - q.z);
                if (r2 >= P.sp.far2) return;
                const float c = stencil_c<F>(r2, P.sp);
                kappa_term<F>(dx, dy, dz, p, q, c, q.w * (ki + bf), dt);
            });
        }
    }
    if (F) {
        const float sc = dt * P.mass;
        dx = __builtin_fmaf(sc, ax, dx);
        dy = __builtin_fmaf(sc, ay, dy);
        dz = __builtin_fmaf(sc, az, dz);
    }
    if (F && L.gb) {  // sum_b m_b (k_i + k_b) grad W_ib = (k_i + k_b) Gb_i
        const float4 g = L.gb[i];
        const float k = -(dt * (ki + bf));
        dx = __builtin_fmaf(k, g.x, dx);
        dy = __builtin_fmaf(k, g.y, dy);
        dz = __builtin_fmaf(k, g.z, dz);
    }
    float4 v = vin[i];
    v.x += dx;
    v.y += dy;
    v.z += dz;
    vout[i] = v;
}

// Advection of a scene runs once per substep, as soon as its density solve
// has concluded: speculatively after an iteration chunk (force = 0: only
// scenes whose error test passed) or unconditionally after the last chunk
// (force = 1).  Scene-uniform.
__device__ __forceinline__ bool advect_due(const DeviceState& S, int force) {
    return !S.advected && (force || S.conv_iter != kNone);
}

// Per-workgroup partials of the substep statistics (one 256-slot tile or
// 256 original indices per workgroup); k_finish_vstats reduces them per
// scene in a fixed order.

// One particle's contribution after the substep: |v|, non-finite state,
// escape (position on or outside the domain clamp box, a NaN never counts)
// and, for a finite state, m, m v and m |v|^2 / 2 in binary64.
struct PStat {
    float mag = 0.0f;
    u32 bad = 0, esc = 0;
    double m = 0.0, px = 0.0, py = 0.0, pz = 0.0, ke = 0.0;
};
__device__ __forceinline__ PStat particle_stat(float mass, float x, float y, float z, float vx,
                                               float vy, float vz, const float* lo,
                                               const float* hi) {
    PStat r;
    r.mag = norm3(vx, vy, vz);
    r.bad = finite3(x, y, z) && finite3(vx, vy, vz) ? 0u : 1u;
    r.esc = (x <= lo[0] || y <= lo[1] || z <= lo[2] || x >= hi[0] || y >= hi[1] || z >= hi[2])
                ? 1u : 0u;
    if (!r.bad) {
        const double m = static_cast<double>(mass), a = vx, b = vy, c = vz;
        r.m = m;
        r.px = m * a;
        r.py = m * b;
        r.pz = m * c;
        r.ke = 0.5 * m * ((a * a + b * b) + c * c);
    }
    return r;
}

// The seven reductions of a workgroup's statistics in one pass: each wave
// reduces all of them with shuffles, one barrier, then thread 0 combines the
// waves in index order (the same trees as block_max / block_sum, one barrier
// instead of fourteen).
template <int NT> struct StatReduce {
    float f[NT / 64];
    u32 c[NT / 64], e[NT / 64];
    double d[kMomParts][NT / 64];
};
template <int NT>
__device__ __forceinline__ bool stat_reduce(float& m, u32& c, u32& e, double (&d)[kMomParts],
                                            StatReduce<NT>& R) {
    m = wave_max(m);
    c = wave_sum_u32(c);
    e = wave_sum_u32(e);
#pragma unroll
    for (int k = 0; k < kMomParts; ++k) d[k] = wave_sum(d[k]);
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    if (lane == 0) {
        R.f[w] = m;
        R.c[w] = c;
        R.e[w] = e;
#pragma unroll
        for (int k = 0; k < kMomParts; ++k) R.d[k][w] = d[k];
    }
    __syncthreads();
    if (threadIdx.x != 0) return false;
    m = 0.0f;
    c = e = 0u;
#pragma unroll
    for (int k = 0; k < kMomParts; ++k) d[k] = 0.0;
    for (int i = 0; i < NT / 64; ++i) {
        m = fmaxf(m, R.f[i]);
        c += R.c[i];
        e += R.e[i];
#pragma unroll
        for (int k = 0; k < kMomParts; ++k) d[k] += R.d[k][i];
    }
    return true;
}

__device__ __forceinline__ void vstats_partial(const PStat& q, const VStats& V) {
    __shared__ StatReduce<NS> R;
    float m = q.mag;
    u32 c = q.bad, e = q.esc;
    double d[kMomParts] = {q.m, q.px, q.py, q.pz, q.ke};
    if (stat_reduce<NS>(m, c, e, d, R)) {
        V.pmax[blockIdx.x] = m;
        V.pnan[blockIdx.x] = c;
        V.pesc[blockIdx.x] = e;
        double* o = V.pmom + kMomParts * static_cast<u64>(blockIdx.x);
#pragma unroll
        for (int k = 0; k < kMomParts; ++k) o[k] = d[k];
    }
}

// The CFL step of sim_cli's substep loop on the device (main.cpp:176-177,
// timestep.h:23-28): dt = next_dt(max(max|v|, v_floor), r, remaining), with
// the comparisons of std::max and std::min over {dt_max, dt_cfl, remaining,
// max(dt_cfl, dt_floor)} in their order, then the substep parameters the
// host otherwise uploads (dt, dt^2, the velocity decay, gravity * dt).  The
// host evaluates the same rule on the published state and checks that both
// agree bit for bit.
__device__ inline void cfl_step(const FrameParams& f, SceneSeg& g, DeviceState& st) {
    const float mv = st.max_v;
    const float a = (mv < f.v_floor) ? f.v_floor : mv;  // std::max(max|v|, v_floor)
    const float v = (a < 1e-6f) ? 1e-6f : a;             // std::max(v, 1e-6f)
    const float dt_cfl = f.cfl_lambda * f.particle_radius / v;
    const float dt_floor = f.frame_dt / static_cast<float>(f.max_substeps);
    const float rem = st.t_rem;
    float d = f.dt_max;
    if (dt_cfl < d) d = dt_cfl;
    if (rem < d) d = rem;
    const float hi = (dt_cfl < dt_floor) ? dt_floor : dt_cfl;
    if (hi < d) d = hi;
    g.dt = d;
    g.dt2 = d * d;
    g.decay = damping_decay(f.damping, d);
    g.gdx = f.gravity[0] * d;
    g.gdy = f.gravity[1] * d;
    g.gdz = f.gravity[2] * d;
    st.dt = d;
}

// A scene's state into mapped host memory, then a new sequence number with
// release ordering at system scope: the host that sees the number sees the
// state.
__device__ inline void publish_state(const DeviceState& S, u32* seq, StatePub* pub) {
    pub->st = S;
    const u32 k = *seq + 1u;
    *seq = k;
    __hip_atomic_store(&pub->seq, k, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_SYSTEM);
}

// advance()'s per-substep device step, folded into k_finish_vstats: the
// time-step rule, the substep's published state and its sequence word.
struct FrameStep {
    const FrameParams* fp;  // null outside advance()
    StatePub* pub;
    u32* seq;
};

// One workgroup per scene.  guarded = 1: only scenes the advection just
// processed (the advect_due test, evaluated before this kernel marks them).
// With F.fp (advance(), one scene): publishes the state whether or not the
// solve finished, and after an advection first takes the substep's dt off
// the time left and then computes the next substep's dt.
// The statistics step of scene sc by one workgroup of NT threads (every
// thread calls it), over `parts` partials from the scene's first tile (0:
// one per tile of the scene).
template <int NT>
__device__ __forceinline__ void finish_vstats(SceneSeg* __restrict__ seg, VStats V,
                                              DeviceState* st, int guarded, int force,
                                              FrameStep F, u32 sc, u32 parts) {
    __shared__ StatReduce<NT> R;
    const SceneSeg& G = seg[sc];
    DeviceState& S = st[sc];
    if (guarded && !advect_due(S, force)) {  // block-uniform
        if (F.fp && threadIdx.x == 0) publish_state(S, F.seq, F.pub);
        return;
    }
    float m = 0.0f;
    u32 c = 0, e = 0;
    double d[kMomParts] = {0.0, 0.0, 0.0, 0.0, 0.0};
    // each thread adds its tiles t, t + NT, t + 2 NT, ... in that order; the
    // loads of four of them are issued before the first is added
    constexpr int kRound = 4;
    const u32 np = parts ? parts : G.ntiles;
    for (u32 t0 = threadIdx.x; t0 < np; t0 += kRound * NT) {
        float pm[kRound];
        u32 pc[kRound], pe[kRound];
        double pd[kRound][kMomParts];
#pragma unroll
        for (int r = 0; r < kRound; ++r) {
            const u32 t = t0 + static_cast<u32>(r) * NT;
            if (t < np) {
                const u32 b = G.tile0 + t;
                pm[r] = V.pmax[b];
                pc[r] = V.pnan[b];
                pe[r] = V.pesc[b];
#pragma unroll
                for (int k = 0; k < kMomParts; ++k) pd[r][k] = V.pmom[kMomParts * static_cast<u64>(b) + k];
            }
        }
#pragma unroll
        for (int r = 0; r < kRound; ++r) {
            if (t0 + static_cast<u32>(r) * NT >= np) break;
            m = fmaxf(m, pm[r]);
            c += pc[r];
            e += pe[r];
#pragma unroll
            for (int k = 0; k < kMomParts; ++k) d[k] += pd[r][k];
        }
    }
    if (stat_reduce<NT>(m, c, e, d, R)) {
        S.max_v = m;
        S.nan_count = c;
        S.escapees = e;
        S.mass = d[0];
        S.mom[0] = d[1];
        S.mom[1] = d[2];
        S.mom[2] = d[3];
        S.ke = d[4];
        if (guarded) S.advected = 1u;
        if (F.fp) {
            S.t_rem = S.t_rem - S.dt;
            publish_state(S, F.seq, F.pub);
            cfl_step(*F.fp, seg[sc], S);
        }
    }
}
template <int NT>
__global__ __launch_bounds__(NT) void k_finish_vstats(SceneSeg* __restrict__ seg, VStats V,
                                                      DeviceState* st, int guarded, int force,
                                                      FrameStep F, u32 parts = 0) {
    finish_vstats<NT>(seg, V, st, guarded, force, F, blockIdx.x, parts);
}

struct StoreOut {
    float* x;
    float* v;
    float* rho;
    float* alpha;
    float* kappa;
    float* rho_adv;
};

// Advection with the |v| clamp and the domain clamp (dfsph_kernels.cu:363-386),
// scatter back to original order, then max |v| (dfsph_kernels.cu:396-400).
__global__ __launch_bounds__(NS) void k_advect(Nbr P, const float4* __restrict__ xs,
                                               const float4* __restrict__ vs,
                                               const u32* __restrict__ perm, float vmax,
                                               float lx, float ly, float lz, float hx, float hy,
                                               float hz, StoreOut out,
                                               const float* __restrict__ rho,
                                               const float* __restrict__ alpha,
                                               const float* __restrict__ kappa,
                                               const float* __restrict__ rho_adv, int wb_solve,
                                               const DeviceState* __restrict__ st, int force,
                                               VStats V) {
    const u32 i = blockIdx.x * NS + threadIdx.x;
    PStat q;
    {
        const u32 sc = scene_of_slot(P, i);
        if (!advect_due(st[sc], force)) return;  // the workgroup is one tile of one scene
        const SceneSeg& G = P.seg[sc];
        const u32 l = i - G.off;
        if (l < G.n) {
            const float dt = G.dt;
            const float4 x4 = xs[i];
            float4 v = vs[i];
            const float sp = norm3(v.x, v.y, v.z);
            if (sp > vmax) {
                const float f = vmax / sp;
                v.x *= f;
                v.y *= f;
                v.z *= f;
            }
            float px = x4.x + v.x * dt, py = x4.y + v.y * dt, pz = x4.z + v.z * dt;
            if (px < lx) { px = lx; if (v.x < 0) v.x = 0; }
            if (py < ly) { py = ly; if (v.y < 0) v.y = 0; }
            if (pz < lz) { pz = lz; if (v.z < 0) v.z = 0; }
            if (px > hx) { px = hx; if (v.x > 0) v.x = 0; }
            if (py > hy) { py = hy; if (v.y > 0) v.y = 0; }
            if (pz > hz) { pz = hz; if (v.z > 0) v.z = 0; }
            const u32 id = perm[G.start + l];
            const u64 o = 3ull * id;
            out.x[o] = px;
            out.x[o + 1] = py;
            out.x[o + 2] = pz;
            out.v[o] = v.x;
            out.v[o + 1] = v.y;
            out.v[o + 2] = v.z;
            out.rho[id] = rho[i];
            out.alpha[id] = alpha[i];
            if (wb_solve) {
                out.kappa[id] = kappa[i];
                out.rho_adv[id] = rho_adv[i];
            }
            const float lo[3] = {lx, ly, lz}, hi[3] = {hx, hy, hz};
            q = particle_stat(P.mass, px, py, pz, v.x, v.y, v.z, lo, hi);
        }
    }
    vstats_partial(q, V);
}

// k_advect for a single scene, one thread per original index: the sorted
// state is gathered through inv (working slot of each original index) and
// the store is written in order, which keeps the writes coalesced where
// the scatter of k_advect is not.  The same arithmetic per particle; the
// max |v| and NaN partials are per block of original indices, which
// k_finish_vstats reduces over the same block range.
__global__ __launch_bounds__(NS) void k_advect_gather(
    Nbr P, const float4* __restrict__ xs, const float4* __restrict__ vs,
    const u32* __restrict__ inv, float vmax, float lx, float ly, float lz, float hx, float hy,
    float hz, StoreOut out, const float* __restrict__ rho, const float* __restrict__ alpha,
    const float* __restrict__ kappa, const float* __restrict__ rho_adv, int wb_solve,
    const DeviceState* __restrict__ st, int force, VStats V) {
    const u32 id = blockIdx.x * NS + threadIdx.x;
    PStat q;
    if (!advect_due(st[0], force)) return;  // grid-uniform
    const SceneSeg& G = P.seg[0];
    if (id < G.n) {
        const float dt = G.dt;
        const u32 i = inv[id];
        const float4 x4 = xs[i];
        float4 v = vs[i];
        const float sp = norm3(v.x, v.y, v.z);
        if (sp > vmax) {
            const float f = vmax / sp;
            v.x *= f;
            v.y *= f;
            v.z *= f;
        }
        float px = x4.x + v.x * dt, py = x4.y + v.y * dt, pz = x4.z + v.z * dt;
        if (px < lx) { px = lx; if (v.x < 0) v.x = 0; }
        if (py < ly) { py = ly; if (v.y < 0) v.y = 0; }
        if (pz < lz) { pz = lz; if (v.z < 0) v.z = 0; }
        if (px > hx) { px = hx; if (v.x > 0) v.x = 0; }
        if (py > hy) { py = hy; if (v.y > 0) v.y = 0; }
        if (pz > hz) { pz = hz; if (v.z > 0) v.z = 0; }
        const u64 o = 3ull * id;
        // one 12-byte store per vector: a wave writes 768 contiguous bytes
        *reinterpret_cast<float3*>(out.x + o) = make_float3(px, py, pz);
        *reinterpret_cast<float3*>(out.v + o) = make_float3(v.x, v.y, v.z);
        out.rho[id] = rho[i];
        out.alpha[id] = alpha[i];
        if (wb_solve) {
            out.kappa[id] = kappa[i];
            out.rho_adv[id] = rho_adv[i];
        }
        const float lo[3] = {lx, ly, lz}, hi[3] = {hx, hy, hz};
        q = particle_stat(P.mass, px, py, pz, v.x, v.y, v.z, lo, hi);
    }
    vstats_partial(q, V);
}

// ---- advance(): state kept in working-slot order between substeps ----------
//
// Inside advance() a substep's advection writes the advected state in its
// own working-slot order (xa, va) instead of scattering it to the caller's
// store; the next substep gathers from there (positions, and velocities with
// gravity, in one reorder) and the store is written once, when advance()
// returns.  The grid, the permutation and every value are the ones the store
// round trip gives.  The stable sort needs the cell keys in original index
// order, and they gather at random (through inv); its values are the
// particles' previous slots, and each position's w carries the particle's
// original index, so the reorder gathers from the previous slots, which
// stay nearly sequential because slots move little between substeps, and
// only inv's update lands at random.

// k_advect on one scene's working slots, writing xa/va in slot order (xa's w
// keeps the original index that xs carries).  The
// statistics partials cover the same 256-slot blocks as k_advect_gather's
// 256-index blocks (binary64 momentum and energy sums then add in another
// order; max |v|, NaN and escapee counts are order-free).
// Also writes each slot's cell key of the advected position (ka), which the
// next grid's key kernel gathers in original index order: 4 bytes a particle
// from a 4 MB array at a million particles instead of its 16-byte position.
// Each thread advects kAdvectPer slots (i, i + NS, ... of a kAdvectSlots
// block) and adds their statistics before the workgroup's reduction, so the
// partials cover kAdvectSlots slots.
constexpr int kAdvectPer = 4;
constexpr u32 kAdvectSlots = kAdvectPer * NS;
__device__ __forceinline__ PStat stat_add(const PStat& a, const PStat& b) {
    PStat r;
    r.mag = fmaxf(a.mag, b.mag);
    r.bad = a.bad + b.bad;
    r.esc = a.esc + b.esc;
    r.m = a.m + b.m;
    r.px = a.px + b.px;
    r.py = a.py + b.py;
    r.pz = a.pz + b.pz;
    r.ke = a.ke + b.ke;
    return r;
}
__global__ __launch_bounds__(NS) void k_advect_sorted(
    Nbr P, const float4* __restrict__ xs, const float4* __restrict__ vs, float vmax, float lx,
    float ly, float lz, float hx, float hy, float hz, float4* __restrict__ xa,
    float4* __restrict__ va, u32* __restrict__ ka, const DeviceState* __restrict__ st, int force,
    VStats V) {
    const u32 i0 = blockIdx.x * kAdvectSlots + threadIdx.x;
    PStat q[kAdvectPer];
    // the slots' loads are issued with the state's (slots [0, nslots) are
    // allocated)
    float4 x4[kAdvectPer], v4[kAdvectPer];
#pragma unroll
    for (int u = 0; u < kAdvectPer; ++u) {
        const u32 i = i0 + u * NS;
        if (i < P.nslots) {
            x4[u] = xs[i];
            v4[u] = vs[i];
        }
    }
    if (!advect_due(st[0], force)) return;  // grid-uniform
    const SceneSeg& G = P.seg[0];
#pragma unroll
    for (int u = 0; u < kAdvectPer; ++u) {
        const u32 i = i0 + u * NS;
        if (i >= G.n) continue;
        float4 v = v4[u];
        const float4 xx = x4[u];
        const float dt = G.dt;
        const float sp = norm3(v.x, v.y, v.z);
        if (sp > vmax) {
            const float f = vmax / sp;
            v.x *= f;
            v.y *= f;
            v.z *= f;
        }
        float px = xx.x + v.x * dt, py = xx.y + v.y * dt, pz = xx.z + v.z * dt;
        if (px < lx) { px = lx; if (v.x < 0) v.x = 0; }
        if (py < ly) { py = ly; if (v.y < 0) v.y = 0; }
        if (pz < lz) { pz = lz; if (v.z < 0) v.z = 0; }
        if (px > hx) { px = hx; if (v.x > 0) v.x = 0; }
        if (py > hy) { py = hy; if (v.y > 0) v.y = 0; }
        if (pz > hz) { pz = hz; if (v.z > 0) v.z = 0; }
        xa[i] = make_float4(px, py, pz, xx.w);  // w: the original index
        va[i] = make_float4(v.x, v.y, v.z, 0.0f);
        ka[i] = P.gf.key_of(px, py, pz);
        const float lo[3] = {lx, ly, lz}, hi[3] = {hx, hy, hz};
        q[u] = particle_stat(P.mass, px, py, pz, v.x, v.y, v.z, lo, hi);
    }
    PStat t = q[0];
#pragma unroll
    for (int u = 1; u < kAdvectPer; ++u) t = stat_add(t, q[u]);
    vstats_partial(t, V);
}

// Cell keys (cell_grid.cu:24-30) in original index order, each index reading
// its particle through inv (its slot in the substep just advected): the key
// the advection computed (ka), or, when that substep's solve has not
// finished (nothing advected), the key of the position it started from,
// which rebuilds that substep's own grid.  The sort's value of each index
// is that slot.  Workgroup t writes the sort's tile t and counts the first
// digit of its keys for the sort (radix_count_*).  With `fold` (a grid enqueued right behind the advection)
// workgroup 0 first runs the advection's statistics step, which marks the
// scene advected; every workgroup takes the same decision from the flags it
// reads, before or after that mark.
__global__ __launch_bounds__(kRadixThreads) void k_keys_sorted(
    u32 n, GridGeom g, const u32* __restrict__ inv, const u32* __restrict__ ka,
    const float4* __restrict__ xs, DeviceState* st, u32* __restrict__ keys,
    u32* __restrict__ vals, u32* span, u32 dmask, u32* __restrict__ hist, u32 tiles, int fold,
    int force, SceneSeg* __restrict__ seg, VStats V, FrameStep F, u32 parts) {
    __shared__ u32 cnt[256];
    constexpr int kItems = kRadixTile / kRadixThreads;
    // the indices' slots, loaded with the state's
    u32 slot[kItems];
#pragma unroll
    for (int r = 0; r < kItems; ++r) {
        const u32 id = blockIdx.x * kRadixTile + r * kRadixThreads + threadIdx.x;
        slot[r] = id < n ? inv[id] : 0u;
    }
    const bool adv = st[0].advected != 0u || (fold && advect_due(st[0], force));
    if (fold && blockIdx.x == 0) {
        __syncthreads();  // every thread has read the flags
        finish_vstats<kRadixThreads>(seg, V, st, 1, force, F, 0u, parts);
    }
    advance_span(span);
    radix_count_begin(cnt);
#pragma unroll
    for (int r = 0; r < kItems; ++r) {
        const u32 id = blockIdx.x * kRadixTile + r * kRadixThreads + threadIdx.x;
        if (id < n) {
            const u32 o = slot[r];
            u32 key;
            if (adv) {
                key = ka[o];
            } else {
                const float4 p = xs[o];
                key = g.key_of(p.x, p.y, p.z);
            }
            keys[id] = key;
            vals[id] = o;
            radix_count_key(cnt, key & dmask);
        }
    }
    radix_count_end(cnt, dmask, hist, tiles);
}

// The new grid's working slots gathered from the advected slots (prev: the
// sorted values, each new slot's previous slot): positions with the
// original index in w, and velocities with gravity and damping
// (dfsph_kernels.cu:347-355, the expression of k_reorder) for the next
// substep, whose time step the statistics step has already computed; inv
// moves to the new slots.  Thread
// 0 of each workgroup (one 256-slot tile) also writes the tile's staging
// plan, as k_stage_plan does.  Without an advection the grid is the previous
// one: nothing moves and the plans stand.
__global__ __launch_bounds__(NS) void k_reorder_sorted(Nbr P, const u32* __restrict__ prev,
                                                       const float4* __restrict__ xa,
                                                       const float4* __restrict__ va,
                                                       const DeviceState* __restrict__ st,
                                                       u32* __restrict__ inv,
                                                       float4* __restrict__ xs,
                                                       float4* __restrict__ vs,
                                                       StageRanges* __restrict__ plan, u32 n) {
    static_assert(NS == kTile, "one workgroup per tile");
    const u32 s = blockIdx.x * NS + threadIdx.x;
    // the slot's source is loaded with the state (n: the scene's particle
    // count)
    const u32 o = s < n ? prev[s] : 0u;
    if (!st[0].advected) return;  // grid-uniform
    if (threadIdx.x == 0) plan[blockIdx.x] = stage_ranges(P, P.seg[0], blockIdx.x * kTile);
    const SceneSeg& G = P.seg[0];
    if (s >= n) return;
    const float4 x = xa[o];
    const float4 w = va[o];
    xs[s] = x;
    vs[s] = make_float4(w.x * G.decay + G.gdx, w.y * G.decay + G.gdy, w.z * G.decay + G.gdz, 0.0f);
    inv[__float_as_uint(x.w)] = s;
}

// The state of advance()'s last substep to the caller's store, in original
// order: positions and velocities from xa/va, and the substep's attributes.
__global__ __launch_bounds__(NS) void k_write_back_sorted(
    Nbr P, const float4* __restrict__ xa, const float4* __restrict__ va, StoreOut out,
    const float* __restrict__ rho, const float* __restrict__ alpha,
    const float* __restrict__ kappa, const float* __restrict__ rho_adv, int wb_solve) {
    const u32 s = blockIdx.x * NS + threadIdx.x;
    if (s >= P.seg[0].n) return;
    const float4 x = xa[s], v = va[s];
    const u32 id = __float_as_uint(x.w);
    const u64 o = 3ull * id;
    *reinterpret_cast<float3*>(out.x + o) = make_float3(x.x, x.y, x.z);
    *reinterpret_cast<float3*>(out.v + o) = make_float3(v.x, v.y, v.z);
    out.rho[id] = rho[s];
    out.alpha[id] = alpha[s];
    if (wb_solve) {
        out.kappa[id] = kappa[s];
        out.rho_adv[id] = rho_adv[s];
    }
}

// End of the divergence solve: velocities and the recomputed attributes back
// to original order, then max |v|.
__global__ __launch_bounds__(NS) void k_div_writeback(Nbr P, const float4* __restrict__ xs,
                                                      const float4* __restrict__ vs,
                                                      const u32* __restrict__ perm, StoreOut out,
                                                      const float* __restrict__ rho,
                                                      const float* __restrict__ alpha,
                                                      const float* __restrict__ kappa_v,
                                                      const float* __restrict__ rho_dot,
                                                      int wb_solve, float lx, float ly, float lz,
                                                      float hx, float hy, float hz, VStats V) {
    const u32 i = blockIdx.x * NS + threadIdx.x;
    PStat q;
    {
        const SceneSeg& G = P.seg[scene_of_slot(P, i)];
        const u32 l = i - G.off;
        if (l < G.n) {
            const float4 x4 = xs[i];
            const float4 v = vs[i];
            const u32 id = perm[G.start + l];
            const u64 o = 3ull * id;
            out.v[o] = v.x;
            out.v[o + 1] = v.y;
            out.v[o + 2] = v.z;
            out.rho[id] = rho[i];
            out.alpha[id] = alpha[i];
            if (wb_solve) {
                out.kappa[id] = kappa_v[i];
                out.rho_adv[id] = rho_dot[i];
            }
            const float lo[3] = {lx, ly, lz}, hi[3] = {hx, hy, hz};
            q = particle_stat(P.mass, x4.x, x4.y, x4.z, v.x, v.y, v.z, lo, hi);
        }
    }
    vstats_partial(q, V);
}

// Opens a solve: every non-empty scene unconverged, empty scenes done.
__global__ __launch_bounds__(NS) void k_reset_solve(const SceneSeg* __restrict__ seg,
                                                    DeviceState* st, u32 scenes, int which) {
    const u32 s = blockIdx.x * NS + threadIdx.x;
    if (s >= scenes) return;
    const u32 v = seg[s].n ? kNone : 0u;
    if (which == 0) {
        st[s].conv_iter = v;
        st[s].advected = 0u;
    } else {
        st[s].div_conv_iter = v;
    }
}

// advance(): the frame's time-step rule and budget into device memory, and
// the first substep's time step.
__global__ void k_frame_begin(FrameParams fp, FrameParams* __restrict__ dst,
                              SceneSeg* __restrict__ seg, DeviceState* __restrict__ st) {
    if (threadIdx.x != 0) return;
    *dst = fp;
    st->t_rem = fp.time;
    cfl_step(fp, seg[0], *st);
}

// ---- launch-shape dispatch ------------------------------------------------
u32 grid_for(u32 nslots, u32 nt) { return div_up(nslots, nt); }

template <bool F>
void density_alpha_mode(u32 nt, hipStream_t s, u32 nslots, const Nbr& P, const PairList& L,
                        const float4* xs, float* rho, float* alpha, DeviceState* open) {
    switch (nt) {
    case 64: k_density_alpha<64, F><<<grid_for(nslots, 64), 64, 0, s>>>(P, L, xs, rho, alpha, nullptr, open); break;
    case 128: k_density_alpha<128, F><<<grid_for(nslots, 128), 128, 0, s>>>(P, L, xs, rho, alpha, nullptr, open); break;
    case 256: k_density_alpha<256, F><<<grid_for(nslots, 256), 256, 0, s>>>(P, L, xs, rho, alpha, nullptr, open); break;
    case 512: k_density_alpha<512, F><<<grid_for(nslots, 512), 512, 0, s>>>(P, L, xs, rho, alpha, nullptr, open); break;
    default: fail(kInvalidArgument, "density_alpha workgroup size");
    }
    WSIM_LAUNCH_CHECK();
}
// fast: the tolerance-mode pair arithmetic (PairMath<true>); open: see
// k_density_alpha
void launch_density_alpha(u32 nt, hipStream_t s, u32 nslots, const Nbr& P, const PairList& L,
                          const float4* xs, float* rho, float* alpha, bool fast,
                          DeviceState* open = nullptr) {
    if (fast) density_alpha_mode<true>(nt, s, nslots, P, L, xs, rho, alpha, open);
    else density_alpha_mode<false>(nt, s, nslots, P, L, xs, rho, alpha, open);
}

template <bool F>
void xsph_mode(u32 nt, hipStream_t s, u32 nslots, const Nbr& P, const PairList& L,
               const float4* xs, const float4* vin, const float* rho, float4* vout) {
    constexpr int XPU = 4;
    switch (nt) {
    case 0: k_xsph_st<XPU, F><<<nslots / kTile, kTile, 0, s>>>(P, L, xs, vin, rho, vout); break;
    case 64: k_xsph<64, XPU, F><<<grid_for(nslots, 64), 64, 0, s>>>(P, L, xs, vin, rho, vout); break;
    case 128: k_xsph<128, XPU, F><<<grid_for(nslots, 128), 128, 0, s>>>(P, L, xs, vin, rho, vout); break;
    case 256: k_xsph<256, XPU, F><<<grid_for(nslots, 256), 256, 0, s>>>(P, L, xs, vin, rho, vout); break;
    case 512: k_xsph<512, XPU, F><<<grid_for(nslots, 512), 512, 0, s>>>(P, L, xs, vin, rho, vout); break;
    default: fail(kInvalidArgument, "xsph workgroup size");
    }
    WSIM_LAUNCH_CHECK();
}
void launch_xsph(u32 nt, hipStream_t s, u32 nslots, const Nbr& P, const PairList& L,
                 const float4* xs, const float4* vin, const float* rho, float4* vout, bool fast) {
    if (fast) xsph_mode<true>(nt, s, nslots, P, L, xs, vin, rho, vout);
    else xsph_mode<false>(nt, s, nslots, P, L, xs, vin, rho, vout);
}

struct ChangeArgs {
    const float4* xs;
    const float4* vs;
    const float* rho;
    const float* alpha;
    float rho0;
    int pform;
    u32 iter;
    const DeviceState* st;
    double* part;
    float* rho_adv;
    float* kappa;
    float4* xk;
    bool fast;  // tolerance-mode pair arithmetic
};

template <int MODE, int NT, int PU, bool F>
void change_one(hipStream_t s, u32 nslots, const Nbr& P, const PairList& L, const ChangeArgs& a) {
    k_change<MODE, NT, PU, F><<<grid_for(nslots, NT), NT, 0, s>>>(
        P, L, a.xs, a.vs, a.rho, a.alpha, a.rho0, a.pform, a.iter, a.st, a.part, a.rho_adv,
        a.kappa, a.xk);
}

template <int MODE, int NT, bool F>
void change_pu(u32 pu, hipStream_t s, u32 nslots, const Nbr& P, const PairList& L,
               const ChangeArgs& a) {
    switch (pu) {
    case 2: change_one<MODE, NT, 2, F>(s, nslots, P, L, a); break;
    case 4: change_one<MODE, NT, 4, F>(s, nslots, P, L, a); break;
    case 8: change_one<MODE, NT, 8, F>(s, nslots, P, L, a); break;
    default: fail(kInvalidArgument, "change pairs per trip");
    }
}

template <int MODE, int PU, bool F>
void change_staged(hipStream_t s, u32 nslots, const Nbr& P, const PairList& L, const ChangeArgs& a) {
    k_change_st<MODE, PU, F><<<nslots / kTile, kTile, 0, s>>>(
        P, L, a.xs, a.vs, a.rho, a.alpha, a.rho0, a.pform, a.iter, a.st, a.part, a.rho_adv,
        a.kappa, a.xk);
}

// nt == 0 selects the LDS-staged kernel (one 256-slot tile per workgroup).
template <int MODE, bool F>
void change_mode(u32 nt, u32 pu, hipStream_t s, u32 nslots, const Nbr& P, const PairList& L,
                 const ChangeArgs& a) {
    if (nt == 0) {
        switch (pu) {
        case 2: change_staged<MODE, 2, F>(s, nslots, P, L, a); break;
        case 4: change_staged<MODE, 4, F>(s, nslots, P, L, a); break;
        case 8: change_staged<MODE, 8, F>(s, nslots, P, L, a); break;
        default: fail(kInvalidArgument, "change pairs per trip");
        }
        WSIM_LAUNCH_CHECK();
        return;
    }
    switch (nt) {
    case 256: change_pu<MODE, 256, F>(pu, s, nslots, P, L, a); break;
    case 512: change_pu<MODE, 512, F>(pu, s, nslots, P, L, a); break;
    case 1024: change_pu<MODE, 1024, F>(pu, s, nslots, P, L, a); break;
    default: fail(kInvalidArgument, "change workgroup size");
    }
    WSIM_LAUNCH_CHECK();
}
template <int MODE>
void launch_change(u32 nt, u32 pu, hipStream_t s, u32 nslots, const Nbr& P, const PairList& L,
                   const ChangeArgs& a) {
    if (a.fast) change_mode<MODE, true>(nt, pu, s, nslots, P, L, a);
    else change_mode<MODE, false>(nt, pu, s, nslots, P, L, a);
}

struct ApplyArgs {
    const float4* xk;
    const float* kappa;
    float rho0;
    int pform;
    u32 iter;
    int mode;
    double thr;
    const double* part;
    DeviceState* st;
    const float4* vin;
    float4* vout;
    bool fast;  // tolerance-mode pair arithmetic
};

template <int NT, int PU, bool F>
void apply_one(hipStream_t s, u32 nslots, const Nbr& P, const PairList& L, const ApplyArgs& a) {
    k_apply<NT, PU, F><<<grid_for(nslots, NT), NT, 0, s>>>(P, L, a.xk, a.kappa, a.rho0, a.pform,
                                                           a.iter, a.mode, a.thr, a.part, a.st,
                                                           a.vin, a.vout);
}

template <int NT, bool F>
void apply_pu(u32 pu, hipStream_t s, u32 nslots, const Nbr& P, const PairList& L,
              const ApplyArgs& a) {
    switch (pu) {
    case 2: apply_one<NT, 2, F>(s, nslots, P, L, a); break;
    case 4: apply_one<NT, 4, F>(s, nslots, P, L, a); break;
    case 8: apply_one<NT, 8, F>(s, nslots, P, L, a); break;
    default: fail(kInvalidArgument, "apply pairs per trip");
    }
}

template <int PU, bool F>
void apply_staged(hipStream_t s, u32 nslots, const Nbr& P, const PairList& L, const ApplyArgs& a) {
    k_apply_st<PU, F><<<nslots / kTile, kTile, 0, s>>>(P, L, a.xk, a.kappa, a.rho0, a.pform,
                                                       a.iter, a.mode, a.thr, a.part, a.st, a.vin,
                                                       a.vout);
}

// nt == 0 selects the LDS-staged kernel.
template <bool F>
void apply_mode(u32 nt, u32 pu, hipStream_t s, u32 nslots, const Nbr& P, const PairList& L,
                const ApplyArgs& a) {
    if (nt == 0) {
        switch (pu) {
        case 2: apply_staged<2, F>(s, nslots, P, L, a); break;
        case 4: apply_staged<4, F>(s, nslots, P, L, a); break;
        case 8: apply_staged<8, F>(s, nslots, P, L, a); break;
        default: fail(kInvalidArgument, "apply pairs per trip");
        }
        WSIM_LAUNCH_CHECK();
        return;
    }
    switch (nt) {
    case 64: apply_pu<64, F>(pu, s, nslots, P, L, a); break;
    case 128: apply_pu<128, F>(pu, s, nslots, P, L, a); break;
    case 256: apply_pu<256, F>(pu, s, nslots, P, L, a); break;
    case 512: apply_pu<512, F>(pu, s, nslots, P, L, a); break;
    default: fail(kInvalidArgument, "apply workgroup size");
    }
    WSIM_LAUNCH_CHECK();
}
void launch_apply(u32 nt, u32 pu, hipStream_t s, u32 nslots, const Nbr& P, const PairList& L,
                  const ApplyArgs& a) {
    if (a.fast) apply_mode<true>(nt, pu, s, nslots, P, L, a);
    else apply_mode<false>(nt, pu, s, nslots, P, L, a);
}

// dfsph.cu:55-65
CellGrid make_grid(const DFSPHConfig& c, hipStream_t s) {
    float origin[3];
    int32_t dims[3];
    for (int a = 0; a < 3; ++a) {
        dims[a] = static_cast<int32_t>(std::ceil((c.domain_max[a] - c.domain_min[a]) /
                                                 c.smoothing_length)) + 2;
        origin[a] = c.domain_min[a] - c.smoothing_length;
    }
    return CellGrid(origin, dims, c.smoothing_length, s);
}

hipStream_t own_stream_if_null(hipStream_t s, hipStream_t* own) {
    if (s) return s;
    // a blocking stream: ordered with work the caller puts on the null stream
    WSIM_HIP(hipStreamCreate(own));
    return *own;
}

u32 round_tile(u64 n) { return static_cast<u32>((n + kTile - 1) / kTile * kTile); }

enum PhaseKey : u64 {
    kPhPrefix = 1,
    kPhDivPrefix = 2,
    kPhAdvect = 3,
    kPhDivTail = 4,
    kPhDensityIters = 5,
    kPhDivIters = 6,
    kPhPrefixGrid = 7,
    kPhPrefixSolve = 8,
    kPhSubstep = 9,
    kPhContinue = 10,
};
u64 phase_key(u64 kind, u64 b = 0, u64 e = 0) { return (kind << 48) | (b << 24) | e; }

}  // namespace

const char* kernel_kind_name(int k) {
    static const char* names[kKNumKinds] = {"grid_build",     "reorder",     "density_alpha",
                                            "xsph",           "density_change", "apply_kappa",
                                            "advect",         "div_change",  "div_apply",
                                            "div_writeback",  "surface_tension"};
    return (k >= 0 && k < kKNumKinds) ? names[k] : "?";
}

bool valid_shapes(const KernelShapes& s) {
    auto in = [](u32 v, std::initializer_list<u32> set) {
        return std::find(set.begin(), set.end(), v) != set.end();
    };
    return in(s.density_alpha_nt, {64, 128, 256, 512}) && in(s.xsph_nt, {0, 64, 128, 256, 512}) &&
           in(s.change_nt, {0, 256, 512, 1024}) && in(s.change_pu, {2, 4, 8}) &&
           in(s.apply_nt, {0, 64, 128, 256, 512}) && in(s.apply_pu, {2, 4, 8}) &&
           in(s.cell_cpt, {4, 8, 16}) && in(s.remap, {0, 1});
}

// ---------------------------------------------------------------------------
KernelTimer::~KernelTimer() {
    for (auto& p : pending_) {
        (void)hipEventDestroy(p.a);
        (void)hipEventDestroy(p.b);
    }
    for (auto e : pool_) (void)hipEventDestroy(e);
    if (open_) (void)hipEventDestroy(open_);
}

hipEvent_t KernelTimer::take() {
    if (!pool_.empty()) {
        hipEvent_t e = pool_.back();
        pool_.pop_back();
        return e;
    }
    hipEvent_t e;
    WSIM_HIP(hipEventCreate(&e));
    return e;
}

void KernelTimer::begin(int kind, int iter, hipStream_t s) {
    open_ = take();
    open_kind_ = kind;
    open_iter_ = iter;
    WSIM_HIP(hipEventRecord(open_, s));
}

void KernelTimer::end(hipStream_t s) {
    hipEvent_t b = take();
    WSIM_HIP(hipEventRecord(b, s));
    pending_.push_back({open_, b, open_kind_, open_iter_});
    open_ = nullptr;
}

void KernelTimer::collect(const int* valid_iters) {
    for (auto& p : pending_) {
        float ms = 0.0f;
        WSIM_HIP(hipEventElapsedTime(&ms, p.a, p.b));
        if (p.iter < valid_iters[p.kind]) {
            acc[p.kind].count += 1;
            acc[p.kind].total_ms += ms;
        }
        pool_.push_back(p.a);
        pool_.push_back(p.b);
    }
    pending_.clear();
}

void KernelTimer::reset() { acc = {}; }

// ---------------------------------------------------------------------------
DFSPHSolver::DFSPHSolver(ParticleStore& fluid, ParticleStore* boundary, const DFSPHConfig& cfg,
                         hipStream_t stream)
    : fluid_(fluid), boundary_(boundary), cfg_(cfg),
      stream_(own_stream_if_null(stream, &own_stream_)), fluid_grid_(make_grid(cfg, stream_)) {
    init({}, {});
}

DFSPHSolver::DFSPHSolver(ParticleStore& fluid, ParticleStore* boundary, const DFSPHConfig& cfg,
                         hipStream_t stream, const std::vector<u32>& counts,
                         const std::vector<float>& viscosity)
    : fluid_(fluid), boundary_(boundary), cfg_(cfg),
      stream_(own_stream_if_null(stream, &own_stream_)), fluid_grid_(make_grid(cfg, stream_)) {
    batched_ = true;
    use_graphs_ = true;
    init(counts, viscosity);
}

DFSPHSolver::~DFSPHSolver() {
    (void)hipStreamSynchronize(stream_);
    drop_graphs();
    if (own_stream_) (void)hipStreamDestroy(own_stream_);
}

void DFSPHSolver::init(const std::vector<u32>& counts, const std::vector<float>& viscosity) {
    if (!(cfg_.smoothing_length > 0.0f))
        fail(kInvalidArgument, "DFSPHSolver: smoothing_length must be positive");
    size_t slots_cap = 0;
    u32 scenes = 1;
    if (batched_) {
        if (counts.empty()) fail(kInvalidArgument, "batched DFSPHSolver: at least one scene");
        if (counts.size() > 0xFFFFu) fail(kInvalidArgument, "batched DFSPHSolver: too many scenes");
        if (!viscosity.empty() && viscosity.size() != counts.size())
            fail(kInvalidArgument, "batched DFSPHSolver: one viscosity per scene");
        u64 total = 0;
        for (u32 c : counts) {
            total += c;
            slots_cap += round_tile(c);
        }
        if (total != fluid_.count())
            fail(kInvalidArgument, "batched DFSPHSolver: scene counts must add up to the store count");
        scenes = static_cast<u32>(counts.size());
        viscosity_ = viscosity.empty() ? std::vector<float>(counts.size(), cfg_.viscosity) : viscosity;
    } else {
        slots_cap = round_tile(fluid_.capacity());
        viscosity_ = {cfg_.viscosity};
    }
    if (slots_cap == 0) slots_cap = kTile;
    if (slots_cap >= 0xFFFFF000ull) fail(kInvalidArgument, "DFSPHSolver: too many particles");

    // dfsph.cu:76-78
    const float spacing = 2.0f * cfg_.particle_radius;
    const float v0 = spacing * spacing * spacing;
    mass_ = cfg_.rest_density * v0;

    // dfsph.cu:80-83
    h_density_ = fluid_.register_attribute("dfsph_density", kF32);
    h_alpha_ = fluid_.register_attribute("dfsph_alpha", kF32);
    h_kappa_ = fluid_.register_attribute("dfsph_kappa", kF32);
    h_density_adv_ = fluid_.register_attribute("dfsph_density_adv", kF32);
    if (cfg_.enable_divergence_solve) h_kappa_v_ = fluid_.register_attribute("dfsph_kappa_v", kF32);

    fluid_grid_.reserve(batched_ ? fluid_.count() : fluid_.capacity());
    fluid_grid_.reserve_segments(scenes);
    fluid_grid_.set_track_span(!batched_);
    xs_ = DeviceBuffer<float4>(slots_cap);
    vs_ = DeviceBuffer<float4>(slots_cap);
    vpre_ = DeviceBuffer<float4>(slots_cap);
    xk_ = DeviceBuffer<float4>(slots_cap);
    rho_ = DeviceBuffer<float>(slots_cap);
    alpha_ = DeviceBuffer<float>(slots_cap);
    kappa_ = DeviceBuffer<float>(slots_cap);
    rho_adv_ = DeviceBuffer<float>(slots_cap);
    pairs_f_ = DeviceBuffer<uint2>(static_cast<size_t>(kPairDepthFluid) * slots_cap);
    pairs_b_ = DeviceBuffer<uint2>(static_cast<size_t>(kPairDepthBoundary) * slots_cap);
    pair_cnt_ = DeviceBuffer<u32>(slots_cap);
    pairs_fi_ = DeviceBuffer<uint16_t>(static_cast<size_t>(kPairDepthFluid) * slots_cap);
    if (cfg_.enable_surface_tension) nrm_ = DeviceBuffer<float4>(slots_cap);
    inv_ = DeviceBuffer<u32>(slots_cap);
    if (!batched_) {
        xa_ = DeviceBuffer<float4>(slots_cap);
        va_ = DeviceBuffer<float4>(slots_cap);
        ka_ = DeviceBuffer<u32>(slots_cap);
    }
    state_sorted_ = grid_sorted_ = defer_store_ = false;
    plan_ = DeviceBuffer<StageRanges>(slots_cap / kTile);
    part_d_ = DeviceBuffer<double>(slots_cap / kTile);
    part_f_ = DeviceBuffer<float>(slots_cap / 64);
    part_u_ = DeviceBuffer<u32>(slots_cap / 64);
    part_e_ = DeviceBuffer<u32>(slots_cap / 64);
    part_m_ = DeviceBuffer<double>(kMomParts * (slots_cap / 64));
    hseg_ = PinnedBuffer<SceneSeg>(scenes);
    dseg_ = DeviceBuffer<SceneSeg>(scenes);
    dstate_ = DeviceBuffer<DeviceState>(scenes);
    dstate_.zero(stream_);
    if (!pub_.data()) {
        dframe_ = DeviceBuffer<FrameParams>(1);
        pubseq_ = DeviceBuffer<u32>(1);
        pubseq_.zero(stream_);
        pub_ = PinnedBuffer<StatePub>(1, hipHostMallocMapped | hipHostMallocCoherent);
        std::memset(pub_.data(), 0, sizeof(StatePub));
        void* d = nullptr;
        WSIM_HIP(hipHostGetDevicePointer(&d, pub_.data(), 0));
        pub_dev_ = static_cast<StatePub*>(d);
        pub_expect_ = 0;
    }
    hstate_ = PinnedBuffer<DeviceState>(scenes);
    stats_.assign(scenes, StepStats{});

    // dfsph.cu:90-99: boundary grid built once, Akinci masses computed once
    if (boundary_) {
        h_bmass_ = boundary_->register_attribute("dfsph_boundary_mass", kF32);
        boundary_grid_ = std::make_unique<CellGrid>(make_grid(cfg_, stream_));
        boundary_grid_->reserve(boundary_->capacity());
        nb_ = boundary_->count();
        if (nb_ > 0) {
            boundary_grid_->build(boundary_->positions(), nb_);
            xb_ = DeviceBuffer<float4>(nb_);
            DeviceBuffer<float> mb(nb_);
            const u32 nb = static_cast<u32>(nb_);
            const u32 blocks = div_up(nb, NS);
            k_gather_boundary<<<blocks, NS, 0, stream_>>>(boundary_grid_->sorted_indices(),
                                                          boundary_->positions(), nb, xb_.data());
            WSIM_LAUNCH_CHECK();
            k_boundary_mass<<<blocks, NS, 0, stream_>>>(boundary_grid_->geom(),
                                                        boundary_grid_->cell_start(), xb_.data(), nb,
                                                        Spline::make(cfg_.smoothing_length),
                                                        cfg_.rest_density, mb.data());
            WSIM_LAUNCH_CHECK();
            k_pack_boundary<<<blocks, NS, 0, stream_>>>(
                nb, boundary_grid_->sorted_indices(), mb.data(), xb_.data(),
                static_cast<float*>(boundary_->attribute_data(h_bmass_)));
            WSIM_LAUNCH_CHECK();
            const GridGeom bg = boundary_grid_->geom();
            const u32 mblocks = div_up(bg.C, NS);
            bnear_ = DeviceBuffer<u32>(static_cast<size_t>(mblocks) * (NS / 32));
            k_boundary_near<<<mblocks, NS, 0, stream_>>>(bg, boundary_grid_->cell_start(),
                                                         bnear_.data());
            WSIM_LAUNCH_CHECK();
            WSIM_HIP(hipStreamSynchronize(stream_));
        }
    }
    if (cfg_.fast_math && nb_) {
        // tolerance mode keeps per-particle boundary gradient sums instead of
        // boundary pair lists
        gb_ = DeviceBuffer<float4>(slots_cap);
        pairs_b_ = DeviceBuffer<uint2>();
    }
    if (batched_) layout(counts);
    WSIM_HIP(hipStreamSynchronize(stream_));
}

void DFSPHSolver::layout(const std::vector<u32>& counts) {
    drop_graphs();
    const u32 scenes = static_cast<u32>(counts.size());
    counts_ = counts;
    starts_.assign(scenes + 1, 0);
    u32 off = 0;
    any_xsph_ = false;
    std::vector<u32> tile_scene;
    for (u32 s = 0; s < scenes; ++s) {
        starts_[s + 1] = starts_[s] + counts[s];
        SceneSeg& g = hseg_[s];
        g = SceneSeg{};
        g.off = off;
        g.n = counts[s];
        g.start = starts_[s];
        g.tile0 = off / kTile;
        g.ntiles = div_up(counts[s], kTile);
        g.key_base = s * (fluid_grid_.total_cells() + 1u);
        g.eps = viscosity_[s];
        if (viscosity_[s] > 0.0f && counts[s] > 0) any_xsph_ = true;
        for (u32 t = 0; t < g.ntiles; ++t) tile_scene.push_back(s);
        off += g.ntiles * kTile;
    }
    total_ = starts_[scenes];
    nslots_ = off;
    if (static_cast<size_t>(nslots_) > xs_.size())
        fail(kRuntimeError, "DFSPHSolver: working set smaller than the scene layout");
    if (batched_) {
        tile_scene_ = DeviceBuffer<u32>(tile_scene.empty() ? 1 : tile_scene.size());
        dstarts_ = DeviceBuffer<u32>(scenes + 1);
        if (!tile_scene.empty())
            WSIM_HIP(hipMemcpy(tile_scene_.data(), tile_scene.data(), tile_scene.size() * sizeof(u32),
                               hipMemcpyHostToDevice));
        WSIM_HIP(hipMemcpy(dstarts_.data(), starts_.data(), (scenes + 1) * sizeof(u32),
                           hipMemcpyHostToDevice));
    }
}

void DFSPHSolver::drop_graphs() {
    for (auto& kv : graphs_) (void)hipGraphExecDestroy(kv.second);
    graphs_.clear();
}

void DFSPHSolver::set_graphs(bool on) {
    if (!on) drop_graphs();
    use_graphs_ = on;
}

void DFSPHSolver::set_shapes(const KernelShapes& s) {
    if (!valid_shapes(s)) fail(kInvalidArgument, "DFSPHSolver: invalid kernel shapes");
    shapes_ = s;
    fluid_grid_.set_cells_per_thread(s.cell_cpt);
    drop_graphs();
}

const StepStats& DFSPHSolver::scene_stats(size_t s) const {
    if (s >= stats_.size()) fail(kOutOfRange, "DFSPHSolver: scene index out of range");
    return stats_[s];
}

float DFSPHSolver::max_velocity() const {
    float m = 0.0f;
    for (const auto& s : stats_) m = std::max(m, s.max_velocity);
    return m;
}

PairList DFSPHSolver::pair_list() {
    PairList L;
    L.f = pairs_f_.data();
    L.b = pairs_b_.data();
    L.cnt = pair_cnt_.data();
    L.fi = pairs_fi_.data();
    L.plan = plan_.data();
    L.kf = kPairDepthFluid;
    L.kb = (nb_ && pairs_b_.data()) ? kPairDepthBoundary : 0u;
    L.gb = (cfg_.fast_math && nb_) ? gb_.data() : nullptr;
    return L;
}

// Per-substep scene parameters, computed on the host once per scene exactly
// as the plain solver computes them (dfsph_kernels.cu:351 exp(-damping*dt)).
void DFSPHSolver::set_substep(const float* dt) {
    const u32 scenes = static_cast<u32>(counts_.size());
    for (u32 s = 0; s < scenes; ++s) {
        SceneSeg& g = hseg_[s];
        const float d = dt[s];
        g.dt = d;
        g.dt2 = d * d;
        g.decay = damping_decay(cfg_.damping, d);
        g.gdx = cfg_.gravity[0] * d;
        g.gdy = cfg_.gravity[1] * d;
        g.gdz = cfg_.gravity[2] * d;
        g.eps = viscosity_[s];
    }
    WSIM_HIP(hipMemcpyAsync(dseg_.data(), hseg_.data(), scenes * sizeof(SceneSeg),
                            hipMemcpyHostToDevice, stream_));
}

void DFSPHSolver::step(float dt) {
    std::vector<float> d(batched_ ? counts_.size() : 1, dt);
    step_scenes(d.data());
}

// sim_cli's CFL substep loop (main.cpp:171-178) in the library.  With the
// divergence solve off, each substep is one launch sequence: gravity, XSPH,
// the first chunk of density iterations and advection, whose statistics
// kernel publishes the state to mapped host memory and computes the next
// substep's time step from the new max |v| and the time left (the first one
// comes from k_frame_begin); then the next substep's grid, density, alpha
// and pair lists, which depend on the positions only.  The host waits for
// the published state (its read overlaps that grid phase), evaluates the
// same time-step rule for its own record and checks it against the
// device's, and only continues the density solve when the first chunk did
// not finish it.  Between substeps the state stays in working-slot order
// (xa_, va_); the store is written when the call returns.
u32 DFSPHSolver::advance(const TimeStepConfig& ts, float v_floor, float time, u32 max_substeps,
                         float* dts, StepStats* stats) {
    if (batched_) fail(kInvalidArgument, "advance: a batched solver steps with step_scenes");
    const size_t nsz = fluid_.count();
    const bool pipe = !cfg_.enable_divergence_solve && !timing_ && nsz > 0;
    if (pipe && (counts_.empty() || counts_[0] != nsz)) layout({static_cast<u32>(nsz)});
    float rem = time;
    u32 sub = 0;
    if (pipe && rem > 1e-6f && max_substeps > 0) {
        hseg_[0].eps = viscosity_[0];
        WSIM_HIP(hipMemcpyAsync(dseg_.data(), hseg_.data(), sizeof(SceneSeg),
                                hipMemcpyHostToDevice, stream_));
        FrameParams fp{};
        fp.frame_dt = ts.frame_dt;
        fp.dt_max = ts.dt_max;
        fp.cfl_lambda = ts.cfl_lambda;
        fp.max_substeps = ts.max_substeps;
        fp.v_floor = v_floor;
        fp.particle_radius = cfg_.particle_radius;
        fp.damping = cfg_.damping;
        for (int a = 0; a < 3; ++a) fp.gravity[a] = cfg_.gravity[a];
        fp.time = time;
        k_frame_begin<<<1, 64, 0, stream_>>>(fp, dframe_.data(), dseg_.data(), dstate_.data());
        WSIM_LAUNCH_CHECK();
    }
    bool grid_ready = false;  // the next substep's grid phase is enqueued
    grid_fold_ = false;
    // between the substeps of this call the state stays in working-slot
    // order; the store is written when the call returns (or throws)
    defer_store_ = pipe;
    try {
        while (rem > 1e-6f && sub < max_substeps) {
            const float dt =
                next_dt(ts, std::max(max_velocity(), v_floor), cfg_.particle_radius, rem);
            if (!pipe) {
                step(dt);
            } else {
                const bool next = rem - dt > 1e-6f && sub + 1 < max_substeps;
                if (!grid_ready) {
                    grid_sorted_ = state_sorted_;
                    run_phase(phase_key(kPhPrefixGrid, state_sorted_ ? 1 : 0),
                              [&] { phase_prefix_grid(); });
                }
                grid_ready = substep_device(next);
                const float dd = hstate_[0].dt;
                if (std::memcmp(&dd, &dt, sizeof(float)) != 0)
                    fail(kRuntimeError, "advance: the device time step differs from the host's");
                hseg_[0].dt = dt;
                hseg_[0].dt2 = dt * dt;
            }
            if (dts) dts[sub] = dt;
            if (stats) stats[sub] = stats_[0];
            rem -= dt;
            ++sub;
        }
    } catch (...) {
        defer_store_ = false;
        try {
            write_back_sorted();
        } catch (...) {
        }
        throw;
    }
    defer_store_ = false;
    write_back_sorted();
    return sub;
}

// Waits for the state k_finish_vstats publishes to mapped host memory and
// copies it to hstate_.  A stream that fails, or finishes without publishing, ends the
// wait with an error.
// The wait polls only the published word for its first 20 ms (a substep
// takes well under that; querying the stream meanwhile competes with the
// runtime's own submission of the queued launches), then also the stream.
void DFSPHSolver::wait_published() {
    const u32 want = ++pub_expect_;
    StatePub* p = pub_.data();
    const auto t0 = std::chrono::steady_clock::now();
    for (u64 spin = 1;; ++spin) {
        if (__atomic_load_n(&p->seq, __ATOMIC_ACQUIRE) == want) break;
        __builtin_ia32_pause();
        if ((spin & 4095u) != 0) continue;
        const auto waited = std::chrono::steady_clock::now() - t0;
        if (waited < std::chrono::milliseconds(20)) continue;
        const hipError_t e = hipStreamQuery(stream_);
        if (e != hipSuccess && e != hipErrorNotReady) WSIM_HIP(e);
        if (e == hipSuccess && __atomic_load_n(&p->seq, __ATOMIC_ACQUIRE) != want)
            fail(kDeviceError, "advance: the substep finished without publishing its state");
        if (waited > std::chrono::seconds(120))
            fail(kDeviceError, "advance: timed out waiting for the substep state");
    }
    std::memcpy(hstate_.data(), &p->st, sizeof(DeviceState));
}

bool DFSPHSolver::substep_device(bool next) {
    const u32 maxit = cfg_.max_density_iters;
    u32 chunk = first_density_chunk();
    u32 begin = 0, end = std::min(chunk, maxit);
    // two instances of each graph, used in turn: a graph is not launched
    // again while its previous launch is still running its tail
    parity_ ^= 1u;
    for (bool first = true;; first = false) {
        const bool last = end >= maxit;
        // a continuation chunk leaves the next grid phase to the host: its
        // tail would rebuild this substep's grid again whenever the solve
        // goes on past it
        const bool tail = next && first;
        const u32 flags =
            (tail ? 1u : 0u) | (last ? 2u : 0u) | (parity_ << 2) | (grid_sorted_ ? 8u : 0u);
        const u64 key = first ? phase_key(kPhSubstep, end, flags)
                              : phase_key(kPhContinue, begin, end | (flags << 16));
        run_phase(key, [&] {
            if (first) phase_prefix_solve(true);
            if (end > begin) phase_iterations(0, begin, end);
            // advection; its statistics step publishes the state and, once
            // the substep is done, computes the next substep's time step (in
            // the grid phase's key kernel when that phase follows here)
            phase_advect(last, tail);
        });
        // the advection above writes xa_/va_: later grids gather from there
        state_sorted_ = true;
        // behind an unfinished solve this rebuilds exactly this substep's
        // grid, lists, density and alpha from the unchanged store positions
        // (functions of the positions only).  A graph of its own: launching
        // the next substep while a large graph still runs its tail stalls
        // that tail's dispatch.
        if (tail) {
            grid_sorted_ = state_sorted_;
            grid_fold_ = true;
            grid_fold_force_ = last;
            run_phase(phase_key(kPhPrefixGrid, (state_sorted_ ? 1u : 0u) | 2u | (last ? 4u : 0u)),
                      [&] { phase_prefix_grid(); });
            grid_fold_ = false;
        }
        wait_published();
        if (hstate_[0].advected) {
            take_state();
            return tail;
        }
        chunk = std::min(chunk * 2u, 32u);
        begin = end;
        end = std::min(end + chunk, maxit);
    }
}

void DFSPHSolver::step_scenes(const float* dt) {
    if (!batched_) {
        const size_t nsz = fluid_.count();
        if (nsz == 0) return;  // dfsph.cu:121-122
        if (counts_.empty() || counts_[0] != nsz) layout({static_cast<u32>(nsz)});
    } else {
        if (fluid_.count() != total_)
            fail(kInvalidArgument, "batched DFSPHSolver: the store count changed");
        if (total_ == 0) return;
    }
    set_substep(dt);
    run_substep();
}

void DFSPHSolver::run_phase(u64 key, const std::function<void()>& body) {
    if (!use_graphs_ || timing_) {
        body();
        return;
    }
    auto it = graphs_.find(key);
    if (it == graphs_.end()) {
        hipGraph_t g = nullptr;
        WSIM_HIP(hipStreamBeginCapture(stream_, hipStreamCaptureModeThreadLocal));
        capturing_ = true;
        try {
            body();
        } catch (...) {
            capturing_ = false;
            (void)hipStreamEndCapture(stream_, &g);
            if (g) (void)hipGraphDestroy(g);
            throw;
        }
        capturing_ = false;
        WSIM_HIP(hipStreamEndCapture(stream_, &g));
        hipGraphExec_t exec = nullptr;
        const hipError_t e = hipGraphInstantiate(&exec, g, nullptr, nullptr, 0);
        (void)hipGraphDestroy(g);
        WSIM_HIP(e);
        it = graphs_.emplace(key, exec).first;
    }
    WSIM_HIP(hipGraphLaunch(it->second, stream_));
}

void DFSPHSolver::phase_prefix(bool gravity) {
    phase_prefix_grid();
    phase_prefix_solve(gravity);
}

// The part of a substep's prefix that does not depend on its time step:
// grid, sorted positions, staging plans, density, alpha and pair lists.
// The grid phase gathers from xa_/va_ when state_sorted_ (advance()); the
// caller sets grid_sorted_ = state_sorted_ before enqueuing it (outside any
// captured body, so graph replays see the same), which selects the matching
// velocity gather in phase_prefix_solve.
void DFSPHSolver::phase_prefix_grid() {
    const u32 scenes = static_cast<u32>(counts_.size());
    if (state_sorted_) {
        // advance(): keys, positions and velocities from the advected slots
        // (ka_, xa_, va_ through inv_) or, behind an unfinished solve, the
        // keys of this substep's own positions (and nothing moves)
        const u32 n = total_;
        tbegin(kKGrid);
        const u32 tiles = radix_tiles(n);
        // behind the first chunk of a substep the key kernel also runs the
        // advection's statistics step (phase_advect left it out)
        const FrameStep F{dframe_.data(), pub_dev_, pubseq_.data()};
        k_keys_sorted<<<tiles, kRadixThreads, 0, stream_>>>(
            n, fluid_grid_.geom(), inv_.data(), ka_.data(), xs_.data(), dstate_.data(),
            fluid_grid_.staged_keys(), fluid_grid_.staged_values(), fluid_grid_.span(),
            fluid_grid_.first_digit_mask(), fluid_grid_.sort_hist(), tiles,
            grid_fold_ ? 1 : 0, grid_fold_force_ ? 1 : 0, dseg_.data(), vstats(), F,
            div_up(nslots_, kAdvectSlots));
        WSIM_LAUNCH_CHECK();
        fluid_grid_.build_staged(n, true);
        tend();
        const Nbr P = nbr();
        tbegin(kKReorder);
        // the velocity target of the next prefix (the pre-XSPH velocities
        // when any scene has XSPH)
        k_reorder_sorted<<<nslots_ / NS, NS, 0, stream_>>>(
            P, fluid_grid_.sorted_indices(), xa_.data(), va_.data(), dstate_.data(), inv_.data(),
            xs_.data(), any_xsph_ ? vpre_.data() : vs_.data(), plan_.data(), total_);
        WSIM_LAUNCH_CHECK();
        tend();
    } else {
        tbegin(kKGrid);
        if (batched_) fluid_grid_.build_segmented(fluid_.positions(), total_, dstarts_.data(), scenes);
        else fluid_grid_.build(fluid_.positions(), total_);
        tend();
        const Nbr P = nbr();
        tbegin(kKReorder);
        k_reorder<<<nslots_ / NS, NS, 0, stream_>>>(P, fluid_grid_.sorted_indices(),
                                                     fluid_.positions(), fluid_.velocities(),
                                                     xs_.data(), nullptr, 0, 1,
                                                     batched_ ? nullptr : inv_.data(),
                                                     defer_store_ ? 1 : 0);
        WSIM_LAUNCH_CHECK();
        tend();
    }
    const Nbr P = nbr();
    if (!state_sorted_) {  // (the sorted reorder writes the plans itself)
        k_stage_plan<<<div_up(nslots_ / kTile, NS), NS, 0, stream_>>>(P, plan_.data());
        WSIM_LAUNCH_CHECK();
    }
    tbegin(kKDensityAlpha);
    launch_density_alpha(shapes_.density_alpha_nt, stream_, nslots_, P, pair_list(), xs_.data(),
                         rho_.data(), alpha_.data(), cfg_.fast_math != 0,
                         state_sorted_ ? dstate_.data() : nullptr);
    tend();
}

// The rest: sorted velocities with gravity and damping (dfsph_kernels.cu:
// 347-355), XSPH, and the opening of the solve.
void DFSPHSolver::phase_prefix_solve(bool gravity) {
    const u32 scenes = static_cast<u32>(counts_.size());
    const Nbr P = nbr();
    const bool xsph = gravity && any_xsph_;
    float4* vtarget = xsph ? vpre_.data() : vs_.data();
    // A grid built from the advected state (advance()) has gathered the
    // velocities with gravity into vtarget already, and its density kernel
    // opened the solve.
    if (grid_sorted_ && !gravity)
        fail(kRuntimeError, "DFSPHSolver: a slot-order grid is only followed by a gravity prefix");
    if (!grid_sorted_) {
        tbegin(kKReorder);
        k_reorder<<<nslots_ / NS, NS, 0, stream_>>>(P, fluid_grid_.sorted_indices(),
                                                     fluid_.positions(), fluid_.velocities(),
                                                     xs_.data(), vtarget, gravity ? 1 : 0, 2,
                                                     nullptr);
        WSIM_LAUNCH_CHECK();
        tend();
    }
    if (xsph) {
        tbegin(kKXsph);
        launch_xsph(shapes_.xsph_nt, stream_, nslots_, P, pair_list(), xs_.data(), vpre_.data(),
                    rho_.data(), vs_.data(), cfg_.fast_math != 0);
        tend();
    }
    if (gravity && cfg_.enable_surface_tension) {
        const float Lsup = 2.0f * cfg_.smoothing_length;
        const TensionArgs T{Lsup, 32.0f / (3.14159265358979323846f * std::pow(Lsup, 9.0f)),
                            std::pow(Lsup, 6.0f) / 64.0f, cfg_.surface_tension, cfg_.rest_density};
        tbegin(kKTension);
        k_tension_normals<NS><<<nslots_ / NS, NS, 0, stream_>>>(P, pair_list(), xs_.data(),
                                                                rho_.data(), T, nrm_.data());
        WSIM_LAUNCH_CHECK();
        k_tension_force<NS><<<nslots_ / NS, NS, 0, stream_>>>(P, pair_list(), xs_.data(),
                                                              rho_.data(), nrm_.data(), T,
                                                              vs_.data());
        WSIM_LAUNCH_CHECK();
        tend();
    }
    if (!grid_sorted_) {
        k_reset_solve<<<div_up(scenes, NS), NS, 0, stream_>>>(dseg_.data(), dstate_.data(), scenes,
                                                              gravity ? 0 : 1);
        WSIM_LAUNCH_CHECK();
    }
}

void DFSPHSolver::neighbor_counts(u32* cand_f, u32* nbr_f, u32* cand_b, u32* nbr_b) {
    // the substep's grid phase on the store's current positions, with the
    // counting walk; the next substep rebuilds all of it from the store
    if (!batched_) {
        const size_t nsz = fluid_.count();
        if (nsz == 0) return;
        if (counts_.empty() || counts_[0] != nsz) layout({static_cast<u32>(nsz)});
    } else {
        if (fluid_.count() != total_)
            fail(kInvalidArgument, "batched DFSPHSolver: the store count changed");
        if (total_ == 0) return;
    }
    // the scene table's layout part (a solver that never stepped has not
    // uploaded it yet; the per-substep part is not read by this walk)
    WSIM_HIP(hipMemcpyAsync(dseg_.data(), hseg_.data(), counts_.size() * sizeof(SceneSeg),
                            hipMemcpyHostToDevice, stream_));
    if (batched_) fluid_grid_.build_segmented(fluid_.positions(), total_, dstarts_.data(),
                                              static_cast<u32>(counts_.size()));
    else fluid_grid_.build(fluid_.positions(), total_);
    const Nbr P = nbr();
    k_reorder<<<nslots_ / NS, NS, 0, stream_>>>(P, fluid_grid_.sorted_indices(), fluid_.positions(),
                                                 fluid_.velocities(), xs_.data(), nullptr, 0, 1,
                                                 nullptr);
    WSIM_LAUNCH_CHECK();
    k_stage_plan<<<div_up(nslots_ / kTile, NS), NS, 0, stream_>>>(P, plan_.data());
    WSIM_LAUNCH_CHECK();
    DeviceBuffer<uint4> counts(std::max<u32>(nslots_, 1u));
    k_density_alpha<NS, false, true><<<nslots_ / NS, NS, 0, stream_>>>(P, pair_list(), xs_.data(),
                                                                rho_.data(), alpha_.data(),
                                                                counts.data());
    WSIM_LAUNCH_CHECK();
    k_scatter_counts<<<nslots_ / NS, NS, 0, stream_>>>(P, fluid_grid_.sorted_indices(),
                                                        counts.data(), cand_f, nbr_f, cand_b,
                                                        nbr_b);
    WSIM_LAUNCH_CHECK();
    WSIM_HIP(hipStreamSynchronize(stream_));
}

VStats DFSPHSolver::vstats() {
    return VStats{part_f_.data(), part_u_.data(), part_e_.data(), part_m_.data()};
}

Nbr DFSPHSolver::nbr() const {
    return make_nbr(fluid_grid_, boundary_grid_.get(), xb_.data(), bnear_.data(), nb_, cfg_, mass_,
                    nslots_, dseg_.data(), batched_ ? tile_scene_.data() : nullptr, shapes_.remap);
}

void DFSPHSolver::phase_iterations(int mode, u32 begin, u32 end) {
    const Nbr P = make_nbr(fluid_grid_, boundary_grid_.get(), xb_.data(), bnear_.data(), nb_, cfg_, mass_, nslots_,
                           dseg_.data(), batched_ ? tile_scene_.data() : nullptr, shapes_.remap);
    const PairList L = pair_list();
    const double thr = mode == 0 ? static_cast<double>(cfg_.eta_density * cfg_.rest_density)
                                 : static_cast<double>(cfg_.eta_divergence * cfg_.rest_density);
    ChangeArgs ca{xs_.data(), vs_.data(), rho_.data(), alpha_.data(), cfg_.rest_density,
                  static_cast<int>(cfg_.pressure_form), 0, dstate_.data(), part_d_.data(),
                  rho_adv_.data(), kappa_.data(), xk_.data(), cfg_.fast_math != 0};
    ApplyArgs aa{xk_.data(), kappa_.data(), cfg_.rest_density, static_cast<int>(cfg_.pressure_form),
                 0, mode, thr, part_d_.data(), dstate_.data(), vs_.data(), vs_.data(),
                 cfg_.fast_math != 0};
    for (u32 it = begin; it < end; ++it) {
        ca.iter = aa.iter = it;
        tbegin(mode == 0 ? kKDensityChange : kKDivChange, static_cast<int>(it));
        if (mode == 0) launch_change<0>(shapes_.change_nt, shapes_.change_pu, stream_, nslots_, P, L, ca);
        else launch_change<1>(shapes_.change_nt, shapes_.change_pu, stream_, nslots_, P, L, ca);
        tend();
        tbegin(mode == 0 ? kKApplyKappa : kKDivApply, static_cast<int>(it));
        launch_apply(shapes_.apply_nt, shapes_.apply_pu, stream_, nslots_, P, L, aa);
        tend();
    }
}

static StoreOut store_out(ParticleStore& f, size_t hd, size_t ha, size_t hk, size_t hr) {
    StoreOut out;
    out.x = f.positions();
    out.v = f.velocities();
    out.rho = static_cast<float*>(f.attribute_data(hd));
    out.alpha = static_cast<float*>(f.attribute_data(ha));
    out.kappa = static_cast<float*>(f.attribute_data(hk));
    out.rho_adv = static_cast<float*>(f.attribute_data(hr));
    return out;
}

void DFSPHSolver::phase_advect(bool force, bool fold_finish) {
    const Nbr P = make_nbr(fluid_grid_, boundary_grid_.get(), xb_.data(), bnear_.data(), nb_, cfg_, mass_, nslots_,
                           dseg_.data(), batched_ ? tile_scene_.data() : nullptr, shapes_.remap);
    const StoreOut out = store_out(fluid_, h_density_, h_alpha_, h_kappa_, h_density_adv_);
    const float eps = 1e-3f;  // dfsph_kernels.cu:377
    tbegin(kKAdvect);
    if (defer_store_) {
        k_advect_sorted<<<div_up(nslots_, kAdvectSlots), NS, 0, stream_>>>(
            P, xs_.data(), vs_.data(), 5.0f, cfg_.domain_min[0] + eps, cfg_.domain_min[1] + eps,
            cfg_.domain_min[2] + eps, cfg_.domain_max[0] - eps, cfg_.domain_max[1] - eps,
            cfg_.domain_max[2] - eps, xa_.data(), va_.data(), ka_.data(), dstate_.data(), force ? 1 : 0,
            vstats());
    } else if (!batched_) {
        k_advect_gather<<<nslots_ / NS, NS, 0, stream_>>>(
            P, xs_.data(), vs_.data(), inv_.data(), 5.0f, cfg_.domain_min[0] + eps,
            cfg_.domain_min[1] + eps, cfg_.domain_min[2] + eps, cfg_.domain_max[0] - eps,
            cfg_.domain_max[1] - eps, cfg_.domain_max[2] - eps, out, rho_.data(), alpha_.data(),
            kappa_.data(), rho_adv_.data(), cfg_.max_density_iters > 0 ? 1 : 0, dstate_.data(),
            force ? 1 : 0, vstats());
    } else {
        k_advect<<<nslots_ / NS, NS, 0, stream_>>>(
            P, xs_.data(), vs_.data(), fluid_grid_.sorted_indices(), 5.0f, cfg_.domain_min[0] + eps,
            cfg_.domain_min[1] + eps, cfg_.domain_min[2] + eps, cfg_.domain_max[0] - eps,
            cfg_.domain_max[1] - eps, cfg_.domain_max[2] - eps, out, rho_.data(), alpha_.data(),
            kappa_.data(), rho_adv_.data(), cfg_.max_density_iters > 0 ? 1 : 0, dstate_.data(),
            force ? 1 : 0, vstats());
    }
    WSIM_LAUNCH_CHECK();
    tend();
    // also marks the scenes advected (after the divergence solve, the
    // writeback reports max |v| instead); with fold_finish the next grid's
    // key kernel, enqueued right behind, runs it
    if (fold_finish) return;
    const FrameStep F = defer_store_ ? FrameStep{dframe_.data(), pub_dev_, pubseq_.data()}
                                     : FrameStep{nullptr, nullptr, nullptr};
    // (the slot-order advection writes one partial per kAdvectSlots slots)
    k_finish_vstats<1024><<<static_cast<u32>(counts_.size()), 1024, 0, stream_>>>(
        dseg_.data(), vstats(), dstate_.data(), 1, force ? 1 : 0, F,
        defer_store_ ? div_up(nslots_, kAdvectSlots) : 0u);
    WSIM_LAUNCH_CHECK();
}

void DFSPHSolver::write_back_sorted() {
    if (!state_sorted_) return;
    const Nbr P = nbr();
    const StoreOut out = store_out(fluid_, h_density_, h_alpha_, h_kappa_, h_density_adv_);
    k_write_back_sorted<<<nslots_ / NS, NS, 0, stream_>>>(
        P, xa_.data(), va_.data(), out, rho_.data(), alpha_.data(), kappa_.data(),
        rho_adv_.data(), cfg_.max_density_iters > 0 ? 1 : 0);
    WSIM_LAUNCH_CHECK();
    state_sorted_ = false;
}

void DFSPHSolver::phase_div_tail() {
    const Nbr P = make_nbr(fluid_grid_, boundary_grid_.get(), xb_.data(), bnear_.data(), nb_, cfg_, mass_, nslots_,
                           dseg_.data(), batched_ ? tile_scene_.data() : nullptr, shapes_.remap);
    const StoreOut out = store_out(fluid_, h_density_, h_alpha_, h_kappa_v_, h_density_adv_);
    const float eps = 1e-3f;  // the advection's clamp box (dfsph_kernels.cu:377)
    tbegin(kKDivWriteback);
    k_div_writeback<<<nslots_ / NS, NS, 0, stream_>>>(
        P, xs_.data(), vs_.data(), fluid_grid_.sorted_indices(), out, rho_.data(), alpha_.data(),
        kappa_.data(), rho_adv_.data(), cfg_.max_divergence_iters > 0 ? 1 : 0,
        cfg_.domain_min[0] + eps, cfg_.domain_min[1] + eps, cfg_.domain_min[2] + eps,
        cfg_.domain_max[0] - eps, cfg_.domain_max[1] - eps, cfg_.domain_max[2] - eps, vstats());
    WSIM_LAUNCH_CHECK();
    tend();
    k_finish_vstats<1024><<<static_cast<u32>(counts_.size()), 1024, 0, stream_>>>(
        dseg_.data(), vstats(), dstate_.data(), 0, 1, FrameStep{nullptr, nullptr, nullptr});
    WSIM_LAUNCH_CHECK();
}

bool DFSPHSolver::all_converged(int mode) {
    const u32 scenes = static_cast<u32>(counts_.size());
    WSIM_HIP(hipMemcpyAsync(hstate_.data(), dstate_.data(), scenes * sizeof(DeviceState),
                            hipMemcpyDeviceToHost, stream_));
    WSIM_HIP(hipStreamSynchronize(stream_));
    for (u32 s = 0; s < scenes; ++s) {
        const u32 c = mode == 0 ? hstate_[s].conv_iter : hstate_[s].div_conv_iter;
        if (c == kNone) return false;
    }
    return true;
}

// First chunk of density iterations: the previous substep's iteration
// count, plus one when its solve ended above half the tolerance (the count
// then tends to grow), and at least the two iterations every density solve
// runs (it tests from its second iteration on).  A chunk that falls short
// costs a continuation chunk (in advance(), behind a superseded grid phase);
// one iteration too many costs two early-exit launches.
u32 DFSPHSolver::first_density_chunk() const {
    const u32 extra = last_err_ratio_ > 0.5 ? 1u : 0u;
    return std::max(2u, std::min(last_iters_[0] + extra, 64u));
}

// Iterations run in chunks between host checks; inside a chunk, scenes that
// converged turn the remaining launches into no-ops.  The density solve's
// first chunk is first_density_chunk(); the divergence solve's is the
// previous substep's count rounded up to a power of two.  Chunks then double
// up to 32.
void DFSPHSolver::solve(int mode) {
    const u32 maxit = mode == 0 ? cfg_.max_density_iters : cfg_.max_divergence_iters;
    u32 it = 0, chunk = 1;
    if (mode == 0) chunk = first_density_chunk();
    else while (chunk < last_iters_[1] && chunk < maxit) chunk <<= 1;
    while (it < maxit) {
        const u32 end = (maxit - it < chunk) ? maxit : it + chunk;
        run_phase(phase_key(mode == 0 ? kPhDensityIters : kPhDivIters, it, end),
                  [&] { phase_iterations(mode, it, end); });
        it = end;
        if (it >= maxit) break;
        if (all_converged(mode)) break;
        chunk = chunk * 2 > 32 ? 32 : chunk * 2;
    }
}

void DFSPHSolver::run_substep() {
    grid_sorted_ = false;  // step() paths gather from the store
    run_phase(phase_key(kPhPrefix), [&] { phase_prefix(true); });
    if (cfg_.enable_divergence_solve) {
        solve(0);
        run_phase(phase_key(kPhAdvect, 1), [&] { phase_advect(true); });
        run_phase(phase_key(kPhDivPrefix), [&] { phase_prefix(false); });
        solve(1);
        run_phase(phase_key(kPhDivTail), [&] { phase_div_tail(); });
        read_state();
        return;
    }
    // Without the divergence solve the substep ends at advection, so each
    // chunk is followed by a speculative advection of the scenes whose solve
    // converged in it, and one state read both tests convergence and
    // delivers the substep's statistics: one host round trip per substep
    // when the first chunk suffices.
    const u32 maxit = cfg_.max_density_iters;
    u32 it = 0, chunk = first_density_chunk();
    while (it < maxit) {
        const u32 end = (maxit - it < chunk) ? maxit : it + chunk;
        run_phase(phase_key(kPhDensityIters, it, end), [&] { phase_iterations(0, it, end); });
        it = end;
        if (it >= maxit) break;
        run_phase(phase_key(kPhAdvect, 0), [&] { phase_advect(false); });
        fetch_state();
        bool all = true;
        for (size_t s = 0; s < counts_.size(); ++s) all = all && hstate_[s].advected;
        if (all) {
            take_state();
            return;
        }
        chunk = chunk * 2 > 32 ? 32 : chunk * 2;
    }
    run_phase(phase_key(kPhAdvect, 1), [&] { phase_advect(true); });
    read_state();
}

void DFSPHSolver::read_state() {
    fetch_state();
    take_state();
}

void DFSPHSolver::fetch_state() {
    const u32 scenes = static_cast<u32>(counts_.size());
    WSIM_HIP(hipMemcpyAsync(hstate_.data(), dstate_.data(), scenes * sizeof(DeviceState),
                            hipMemcpyDeviceToHost, stream_));
    WSIM_HIP(hipStreamSynchronize(stream_));
}

void DFSPHSolver::take_state() {
    const u32 scenes = static_cast<u32>(counts_.size());
    u32 max_d = 0, max_v = 0;
    double ratio = 0.0;
    const double tol = static_cast<double>(cfg_.eta_density) * cfg_.rest_density;
    for (u32 s = 0; s < scenes; ++s) {
        const DeviceState& h = hstate_[s];
        StepStats& st = stats_[s];
        const bool empty = counts_[s] == 0;
        const u32 maxit = cfg_.max_density_iters;
        st.density_iters =
            (maxit == 0 || empty) ? 0 : (h.conv_iter == kNone ? maxit : h.conv_iter + 1);
        st.mean_density_error = (maxit == 0 || empty) ? 0.0 : h.err;
        if (cfg_.enable_divergence_solve) {
            const u32 mv = cfg_.max_divergence_iters;
            st.divergence_iters =
                (mv == 0 || empty) ? 0 : (h.div_conv_iter == kNone ? mv : h.div_conv_iter + 1);
            st.mean_divergence_error = (mv == 0 || empty) ? 0.0 : h.div_err;
        }
        st.max_velocity = h.max_v;
        st.nan_count = h.nan_count;
        st.escapees = empty ? 0u : h.escapees;
        st.total_mass = empty ? 0.0 : h.mass;
        for (int a = 0; a < 3; ++a) st.momentum[a] = empty ? 0.0 : h.mom[a];
        st.kinetic_energy = empty ? 0.0 : h.ke;
        st.substeps += 1;
        max_d = std::max(max_d, st.density_iters);
        max_v = std::max(max_v, st.divergence_iters);
        ratio = std::max(ratio, tol > 0.0 ? st.mean_density_error / tol : 1.0);
    }
    last_iters_[0] = max_d;
    last_iters_[1] = max_v;
    last_err_ratio_ = ratio;
    if (timing_) {
        int valid[kKNumKinds];
        for (int k = 0; k < kKNumKinds; ++k) valid[k] = std::numeric_limits<int>::max();
        valid[kKDensityChange] = valid[kKApplyKappa] = static_cast<int>(max_d);
        valid[kKDivChange] = valid[kKDivApply] = static_cast<int>(max_v);
        timer_.collect(valid);
    }
}

float DFSPHSolver::attribute_at(int which, size_t i) const {
    const ParticleStore* st = (which == 2) ? boundary_ : &fluid_;
    if (!st) return 0.0f;  // dfsph.cu:347
    if (i >= st->count()) fail(kOutOfRange, "DFSPHSolver: particle index out of range");
    const size_t h = which == 0 ? h_density_ : which == 1 ? h_alpha_ : h_bmass_;
    float v = 0.0f;
    auto* s = const_cast<ParticleStore*>(st);
    WSIM_HIP(hipMemcpyAsync(&v, static_cast<const float*>(s->attribute_data(h)) + i, sizeof(float),
                            hipMemcpyDeviceToHost, stream_));
    WSIM_HIP(hipStreamSynchronize(stream_));
    return v;
}

// ---------------------------------------------------------------------------
// Taped substep for the adjoint: the same kernels as step(), with every
// iteration reading its velocity from slab k of the tape and writing slab
// k+1, and its (x, kappa-hat) and rho* into slab k.
void DFSPHSolver::iterations_taped(u32 begin, u32 end, SubstepTape& t) {
    const Nbr P = make_nbr(fluid_grid_, boundary_grid_.get(), xb_.data(), bnear_.data(), nb_, cfg_, mass_, nslots_,
                           dseg_.data(), nullptr, shapes_.remap);
    const PairList L = pair_list();
    const double thr = static_cast<double>(cfg_.eta_density * cfg_.rest_density);
    const size_t ns = nslots_;
    ChangeArgs ca{xs_.data(), nullptr, rho_.data(), alpha_.data(), cfg_.rest_density,
                  static_cast<int>(cfg_.pressure_form), 0, dstate_.data(), part_d_.data(),
                  nullptr, kappa_.data(), nullptr, false};
    ApplyArgs aa{nullptr, kappa_.data(), cfg_.rest_density, static_cast<int>(cfg_.pressure_form), 0, 0,
                 thr, part_d_.data(), dstate_.data(), nullptr, nullptr, false};
    for (u32 it = begin; it < end; ++it) {
        ca.iter = aa.iter = it;
        ca.vs = t.w.data() + it * ns;
        ca.rho_adv = t.p.data() + it * ns;
        ca.xk = t.xk.data() + it * ns;
        aa.xk = ca.xk;
        aa.vin = ca.vs;
        aa.vout = t.w.data() + (it + 1) * ns;
        tbegin(kKDensityChange, static_cast<int>(it));
        launch_change<0>(shapes_.change_nt, shapes_.change_pu, stream_, nslots_, P, L, ca);
        tend();
        tbegin(kKApplyKappa, static_cast<int>(it));
        launch_apply(shapes_.apply_nt, shapes_.apply_pu, stream_, nslots_, P, L, aa);
        tend();
    }
}

void DFSPHSolver::step_taped(float dt, SubstepTape& t) {
    if (batched_) fail(kInvalidArgument, "step_taped: plain solver only");
    if (cfg_.enable_divergence_solve || cfg_.enable_surface_tension)
        fail(kInvalidArgument,
             "step_taped: the adjoint covers the density solve without surface tension");
    if (cfg_.fast_math)
        fail(kInvalidArgument,
             "step_taped: the adjoint differentiates the bit-exact pair arithmetic (fast_math 0)");
    const size_t nsz = fluid_.count();
    if (nsz == 0) fail(kInvalidArgument, "step_taped: empty store");
    if (counts_.empty() || counts_[0] != nsz) layout({static_cast<u32>(nsz)});
    const u32 maxit = cfg_.max_density_iters;
    const size_t ns = nslots_;
    t.x.ensure(3 * nsz);
    t.v.ensure(3 * nsz);
    t.w.ensure((maxit + 1) * ns);
    t.xk.ensure(std::max<u32>(maxit, 1) * ns);
    t.p.ensure(std::max<u32>(maxit, 1) * ns);
    WSIM_HIP(hipMemcpyAsync(t.x.data(), fluid_.positions(), 3 * nsz * sizeof(float),
                            hipMemcpyDeviceToDevice, stream_));
    WSIM_HIP(hipMemcpyAsync(t.v.data(), fluid_.velocities(), 3 * nsz * sizeof(float),
                            hipMemcpyDeviceToDevice, stream_));
    set_substep(&dt);
    grid_sorted_ = false;
    phase_prefix(true);
    WSIM_HIP(hipMemcpyAsync(t.w.data(), vs_.data(), ns * sizeof(float4), hipMemcpyDeviceToDevice,
                            stream_));
    u32 it = 0, chunk = 2;
    while (it < maxit) {
        const u32 end = (maxit - it < chunk) ? maxit : it + chunk;
        iterations_taped(it, end, t);
        it = end;
        if (it >= maxit) break;
        if (all_converged(0)) break;
        chunk = chunk * 2 > 32 ? 32 : chunk * 2;
    }
    u32 k = 0;
    if (maxit > 0) {
        all_converged(0);
        k = hstate_[0].conv_iter == kNone ? maxit : hstate_[0].conv_iter + 1;
    }
    WSIM_HIP(hipMemcpyAsync(vs_.data(), t.w.data() + k * ns, ns * sizeof(float4),
                            hipMemcpyDeviceToDevice, stream_));
    if (k > 0)
        WSIM_HIP(hipMemcpyAsync(rho_adv_.data(), t.p.data() + (k - 1) * ns, ns * sizeof(float),
                                hipMemcpyDeviceToDevice, stream_));
    phase_advect(true);
    read_state();
    t.dt = dt;
    t.iters = k;
    t.n = static_cast<u32>(nsz);
    t.nslots = nslots_;
}

// ---------------------------------------------------------------------------
// Autotuner.  Every candidate computes the same bits, so the search runs on
// the solver's own data: one substep prefix on the current store state (no
// gravity step is written back, the store is not touched), then each kernel
// is launched `reps` times per candidate between HIP events and the median
// launch time decides.  Outputs of the timed launches land in solver buffers
// that the next substep rebuilds from the store.
void DFSPHSolver::autotune(int reps) {
    if (reps < 1) fail(kInvalidArgument, "autotune: reps must be positive");
    if (!batched_) {
        const size_t nsz = fluid_.count();
        if (nsz == 0) return;
        if (counts_.empty() || counts_[0] != nsz) layout({static_cast<u32>(nsz)});
    }
    if (total_ == 0) return;
    drop_graphs();
    const bool saved_timing = timing_;
    timing_ = false;
    const u32 scenes = static_cast<u32>(counts_.size());
    std::vector<float> dt(scenes);
    for (u32 s = 0; s < scenes; ++s) dt[s] = hseg_[s].dt > 0.0f ? hseg_[s].dt : 1e-3f;
    set_substep(dt.data());
    grid_sorted_ = false;
    phase_prefix(true);

    hipEvent_t ev[2];
    WSIM_HIP(hipEventCreate(&ev[0]));
    WSIM_HIP(hipEventCreate(&ev[1]));
    auto time_ms = [&](const std::function<void()>& launch) {
        launch();  // warm-up
        std::vector<float> t(reps);
        for (int r = 0; r < reps; ++r) {
            WSIM_HIP(hipEventRecord(ev[0], stream_));
            launch();
            WSIM_HIP(hipEventRecord(ev[1], stream_));
            WSIM_HIP(hipEventSynchronize(ev[1]));
            WSIM_HIP(hipEventElapsedTime(&t[r], ev[0], ev[1]));
        }
        std::nth_element(t.begin(), t.begin() + reps / 2, t.end());
        return static_cast<double>(t[reps / 2]);
    };
    tune_log_.clear();
    KernelShapes best = shapes_;
    const PairList L = pair_list();
    const u32* tsc = batched_ ? tile_scene_.data() : nullptr;
    auto nbr = [&](u32 remap) {
        return make_nbr(fluid_grid_, boundary_grid_.get(), xb_.data(), bnear_.data(), nb_, cfg_, mass_, nslots_,
                        dseg_.data(), tsc, remap);
    };
    ChangeArgs ca{xs_.data(), vs_.data(), rho_.data(), alpha_.data(), cfg_.rest_density,
                  static_cast<int>(cfg_.pressure_form), 0, dstate_.data(), part_d_.data(),
                  rho_adv_.data(), kappa_.data(), xk_.data(), cfg_.fast_math != 0};
    ApplyArgs aa{xk_.data(), kappa_.data(), cfg_.rest_density, static_cast<int>(cfg_.pressure_form),
                 0, 0, static_cast<double>(cfg_.eta_density * cfg_.rest_density), part_d_.data(),
                 dstate_.data(), vs_.data(), vs_.data(), cfg_.fast_math != 0};
    try {
        // cell_start pass
        // cell_start entries: C + 1 for one scene, scenes * (C + 1) + 1 segmented
        const u32 ncs = batched_ ? scenes * (fluid_grid_.total_cells() + 1u) + 1u
                                 : fluid_grid_.total_cells() + 1u;
        double bt = 1e300;
        for (u32 cpt : {4u, 8u, 16u}) {
            const double ms = time_ms([&] {
                build_cell_start(fluid_grid_.sorted_keys(), total_, ncs,
                                 const_cast<u32*>(fluid_grid_.cell_start()), stream_, cpt);
            });
            tune_log_.push_back({kKGrid, 256, 0, 0, cpt, ms});
            if (ms < bt) bt = ms, best.cell_cpt = cpt;
        }
        double best_total = 0.0;
        for (u32 remap : {0u, 1u}) {
            const Nbr P = nbr(remap);
            double b_da = 1e300, b_xs = 1e300, b_ch = 1e300, b_ap = 1e300;
            KernelShapes cand = best;
            for (u32 nt : {64u, 128u, 256u, 512u}) {
                const double ms = time_ms([&] {
                    launch_density_alpha(nt, stream_, nslots_, P, L, xs_.data(), rho_.data(),
                                         alpha_.data(), cfg_.fast_math != 0);
                });
                tune_log_.push_back({kKDensityAlpha, nt, 0, remap, 0, ms});
                if (ms < b_da) b_da = ms, cand.density_alpha_nt = nt;
            }
            if (any_xsph_) {
                for (u32 nt : {0u, 64u, 128u, 256u, 512u}) {
                    const double ms = time_ms([&] {
                        launch_xsph(nt, stream_, nslots_, P, L, xs_.data(), vpre_.data(),
                                    rho_.data(), vs_.data(), cfg_.fast_math != 0);
                    });
                    tune_log_.push_back({kKXsph, nt, 0, remap, 0, ms});
                    if (ms < b_xs) b_xs = ms, cand.xsph_nt = nt;
                }
            } else {
                b_xs = 0.0;
            }
            for (u32 nt : {0u, 256u, 512u, 1024u})
                for (u32 pu : {2u, 4u, 8u}) {
                    const double ms = time_ms(
                        [&] { launch_change<0>(nt, pu, stream_, nslots_, P, L, ca); });
                    tune_log_.push_back({kKDensityChange, nt, pu, remap, 0, ms});
                    if (ms < b_ch) b_ch = ms, cand.change_nt = nt, cand.change_pu = pu;
                }
            for (u32 nt : {0u, 64u, 128u, 256u, 512u})
                for (u32 pu : {2u, 4u, 8u}) {
                    const double ms =
                        time_ms([&] { launch_apply(nt, pu, stream_, nslots_, P, L, aa); });
                    tune_log_.push_back({kKApplyKappa, nt, pu, remap, 0, ms});
                    if (ms < b_ap) b_ap = ms, cand.apply_nt = nt, cand.apply_pu = pu;
                }
            const double total = b_da + b_xs + b_ch + b_ap;
            if (remap == 0 || total < best_total) {
                best_total = total;
                cand.remap = remap;
                cand.cell_cpt = best.cell_cpt;
                best = cand;
            }
        }
    } catch (...) {
        (void)hipEventDestroy(ev[0]);
        (void)hipEventDestroy(ev[1]);
        timing_ = saved_timing;
        throw;
    }
    (void)hipEventDestroy(ev[0]);
    (void)hipEventDestroy(ev[1]);
    timing_ = saved_timing;
    WSIM_HIP(hipStreamSynchronize(stream_));
    set_shapes(best);
}

}  // namespace wsim
