// solver.hip -- DFSPH substep on cell-sorted particle data for gfx950.
//
// Replaces DFSPHSolver::step (solvers/dfsph/src/dfsph.cu:120-133) and its 11
// kernels (solvers/dfsph/src/dfsph_kernels.cu:40-400).
//
// Data flow per substep (all launches on the caller's stream):
//   grid build         keys -> stable radix sort -> dense cell_start      (grid.hip)
//   k_reorder          gather position/velocity into cell-sorted float4
//                      arrays; gravity + damping applied on the way
//                      (dfsph_kernels.cu:347-355; density and alpha do not
//                      read velocity, so applying it first is equivalent)
//   k_density_alpha    rho and alpha in one stencil walk (dfsph_kernels.cu:68-169)
//   k_xsph             Jacobi XSPH into a second velocity buffer (dfsph_kernels.cu:305-341)
//   density solve      k_change<0> / k_apply per iteration; the mean error is
//                      reduced in binary64 by the last workgroup of k_change
//                      and the early exit is decided on the device
//                      (dfsph.cu:215-260, dfsph_kernels.cu:176-301, 389-394)
//   k_advect           clamp, integrate, domain clamp, scatter position,
//                      velocity and the solver attributes back to the caller's
//                      store in original order, max |v| and NaN count
//                      (dfsph_kernels.cu:363-386, 396-400; dfsph.cu:303-326)
//   [divergence solve] grid, reorder, density/alpha, k_change<1> / k_apply,
//                      k_div_writeback (phase-2 plan text :1858-2003)
//
// Every per-pair expression keeps the reference's operand order, the stencil
// is walked dz, dy, dx with ascending original id inside each cell (the
// sorted slots of a cell are in id order), and the code is compiled without
// FP contraction and with IEEE divide/sqrt: each field equals the binary32
// oracle bit for bit.
#include <cmath>
#include <limits>

#include "radix_sort.hpp"
#include "wsim_internal.hpp"

namespace wsim {
namespace {

constexpr int NT = 256;
constexpr u32 kNone = 0xFFFFFFFFu;
constexpr int PU = 4;  // pairs per trip of the list loops (loads issued together)

// Neighbour-search inputs shared by the stencil kernels.
struct Nbr {
    GridGeom gf;
    const u32* csf;
    GridGeom gb;
    const u32* csb;
    const float4* xb;  // cell-sorted boundary (x, y, z, mass)
    u32 nb;            // 0: no boundary
    Spline sp;
    float mass;
    u32 n;
};

// 27-cell stencil of ITER_NEIGHBORS (dfsph_kernels.cu:13-35): dz, dy, dx in
// {-1, 0, 1}, out-of-grid cells skipped, then every slot of the cell.
template <class F>
__device__ __forceinline__ void for_stencil(const GridGeom& g, const u32* __restrict__ cs, float px,
                                            float py, float pz, F&& f) {
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
                for (u32 s = b; s < e; ++s) f(s);
            }
        }
    }
}

__device__ __forceinline__ u32 tile_index(u32 n) {
    return xcd_tile(blockIdx.x, gridDim.x) * NT + threadIdx.x;
}

__global__ __launch_bounds__(NT) void k_reorder(const u32* __restrict__ perm,
                                                const float* __restrict__ x,
                                                const float* __restrict__ v, u32 n,
                                                float4* __restrict__ xs, float4* __restrict__ vs,
                                                int gravity, float decay, float gdx, float gdy,
                                                float gdz) {
    const u32 s = blockIdx.x * NT + threadIdx.x;
    if (s >= n) return;
    const u64 o = 3ull * perm[s];
    xs[s] = make_float4(x[o], x[o + 1], x[o + 2], 0.0f);
    if (vs) {
        float a = v[o], b = v[o + 1], c = v[o + 2];
        if (gravity) {  // v * exp(-damping*dt) + g*dt (dfsph_kernels.cu:351-354)
            a = a * decay + gdx;
            b = b * decay + gdy;
            c = c * decay + gdz;
        }
        vs[s] = make_float4(a, b, c, 0.0f);
    }
}

// Akinci boundary mass, in cell-sorted order (dfsph_kernels.cu:40-65).
__global__ __launch_bounds__(NT) void k_boundary_mass(GridGeom g, const u32* __restrict__ cs,
                                                      const float4* __restrict__ xb, u32 nb,
                                                      Spline sp, float rho0,
                                                      float* __restrict__ mb) {
    const u32 i = tile_index(nb);
    if (i >= nb) return;
    const float4 p = xb[i];
    float sum_w = 0.0f;
    for_stencil(g, cs, p.x, p.y, p.z, [&](u32 s) {
        const float4 q = xb[s];
        sum_w += spline_W(norm3(p.x - q.x, p.y - q.y, p.z - q.z), sp);
    });
    mb[i] = rho0 / fmaxf(sum_w, 1e-9f);
}

__global__ __launch_bounds__(NT) void k_pack_boundary(u32 nb, const u32* __restrict__ perm,
                                                      const float* __restrict__ mb,
                                                      float4* __restrict__ xb,
                                                      float* __restrict__ store_mb) {
    const u32 s = blockIdx.x * NT + threadIdx.x;
    if (s >= nb) return;
    xb[s].w = mb[s];
    store_mb[perm[s]] = mb[s];
}

// Per-particle pair lists, ELL layout: entry k of particle i at [k * n + i]
// holds (neighbour slot, bits of c) with grad W_ij = (x_i - x_j) * c, for
// every pair the reference's stencil visits with c != 0, in the reference's
// summation order.  Positions are frozen from the grid build to advection, so
// the iteration kernels recompute r and g = r * c from the list and get the
// same binary32 values the reference gets from grad_W (zero-gradient pairs
// add +-0 to sums that never hold -0, so leaving them out is exact).
// Particles with more pairs than the list depth use the stencil walk.
// (PairList itself is declared in wsim_internal.hpp.)

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

__device__ __forceinline__ u32 pack_counts(u32 cf, u32 cb) {
    return (cf < 0xFFFFu ? cf : 0xFFFFu) | ((cb < 0xFFFFu ? cb : 0xFFFFu) << 16);
}

// rho (dfsph_kernels.cu:68-112) and alpha (dfsph_kernels.cu:117-169) in one
// stencil walk: they share r, |r| and q for every pair.  Also records the
// pair lists used by every iteration of the pressure solves.
__global__ __launch_bounds__(NT) void k_density_alpha(Nbr P, PairList L,
                                                      const float4* __restrict__ xs,
                                                      float* __restrict__ rho_out,
                                                      float* __restrict__ alpha_out) {
    const u32 i = tile_index(P.n);
    if (i >= P.n) return;
    const float4 p = xs[i];
    float rho = 0.0f, sx = 0.0f, sy = 0.0f, sz = 0.0f, sq = 0.0f;
    u32 cf = 0, cb = 0;
    for_stencil(P.gf, P.csf, p.x, p.y, p.z, [&](u32 s) {
        const float4 q = xs[s];
        const float rx = p.x - q.x, ry = p.y - q.y, rz = p.z - q.z;
        const float d = norm3(rx, ry, rz);
        rho += P.mass * spline_W(d, P.sp);
        if (s != i) {
            const float c = spline_gradc(d, P.sp);
            const float mx = P.mass * (rx * c), my = P.mass * (ry * c), mz = P.mass * (rz * c);
            sx += mx;
            sy += my;
            sz += mz;
            sq += mx * mx + my * my + mz * mz;
            if (c != 0.0f) {
                if (cf < L.kf) L.f[pair_at(cf, i, L.kf)] = make_uint2(s, __float_as_uint(c));
                ++cf;
            }
        }
    });
    if (P.nb) {
        for_stencil(P.gb, P.csb, p.x, p.y, p.z, [&](u32 s) {
            const float4 q = P.xb[s];
            const float rx = p.x - q.x, ry = p.y - q.y, rz = p.z - q.z;
            const float d = norm3(rx, ry, rz);
            rho += q.w * spline_W(d, P.sp);
            const float c = spline_gradc(d, P.sp);
            const float mx = q.w * (rx * c), my = q.w * (ry * c), mz = q.w * (rz * c);
            sx += mx;
            sy += my;
            sz += mz;
            sq += mx * mx + my * my + mz * mz;
            if (c != 0.0f) {
                if (cb < L.kb) L.b[pair_at(cb, i, L.kb)] = make_uint2(s, __float_as_uint(c));
                ++cb;
            }
        });
    }
    L.cnt[i] = pack_counts(cf, cb);
    rho_out[i] = rho;
    const float denom = sx * sx + sy * sy + sz * sz + sq;
    alpha_out[i] = (denom > 1e-9f) ? (rho / denom) : 0.0f;
}

// XSPH as a Jacobi sweep: every neighbour velocity is the pre-XSPH one.
__global__ __launch_bounds__(NT) void k_xsph(Nbr P, const float4* __restrict__ xs,
                                             const float4* __restrict__ vin,
                                             const float* __restrict__ rho, float eps,
                                             float4* __restrict__ vout) {
    const u32 i = tile_index(P.n);
    if (i >= P.n) return;
    const float4 p = xs[i];
    const float4 vi = vin[i];
    const float rho_i = rho[i];
    float dx = 0.0f, dy = 0.0f, dz = 0.0f;
    for_stencil(P.gf, P.csf, p.x, p.y, p.z, [&](u32 s) {
        if (s == i) return;
        const float4 q = xs[s];
        const float4 vj = vin[s];
        const float w = spline_W(norm3(p.x - q.x, p.y - q.y, p.z - q.z), P.sp);
        const float rho_avg = 0.5f * (rho_i + rho[s]);
        const float c = eps * (P.mass / fmaxf(rho_avg, 1e-9f)) * w;
        dx += c * (vj.x - vi.x);
        dy += c * (vj.y - vi.y);
        dz += c * (vj.z - vi.z);
    });
    vout[i] = make_float4(vi.x + dx, vi.y + dy, vi.z + dz, 0.0f);
}

// rho_dot terms (dfsph_kernels.cu:210, 222) for one pair with gradient scalar c.
__device__ __forceinline__ void rho_dot_fluid(float& rd, const float4& p, const float4& vi,
                                              const float4& q, const float4& vj, float c,
                                              float mass) {
    const float gx = (p.x - q.x) * c, gy = (p.y - q.y) * c, gz = (p.z - q.z) * c;
    const float ux = vi.x - vj.x, uy = vi.y - vj.y, uz = vi.z - vj.z;
    rd += mass * (ux * gx + uy * gy + uz * gz);
}
__device__ __forceinline__ void rho_dot_boundary(float& rd, const float4& p, const float4& vi,
                                                 const float4& q, float c) {
    const float gx = (p.x - q.x) * c, gy = (p.y - q.y) * c, gz = (p.z - q.z) * c;
    rd += q.w * (vi.x * gx + vi.y * gy + vi.z * gz);
}

// One predictor sweep.  MODE 0: density change + kappa (dfsph_kernels.cu:
// 176-232).  MODE 1: divergence change + kappa_v (plan text :1862-1907, with
// the boundary term of the density change).  Writes (x, kappa-hat) for the
// following k_apply and reduces the mean error in binary64; the last
// workgroup decides convergence on the device.
template <int MODE>
__global__ __launch_bounds__(NT) void k_change(Nbr P, PairList L, const float4* __restrict__ xs,
                                               const float4* __restrict__ vs,
                                               const float* __restrict__ rho,
                                               const float* __restrict__ alpha, float dt,
                                               float dt2, float rho0, int pform, u32 iter,
                                               const DeviceState* st,
                                               double* __restrict__ part,
                                               float* __restrict__ rho_adv,
                                               float* __restrict__ kappa,
                                               float4* __restrict__ xk) {
    __shared__ double s_red[NT / 64];
    const u32 conv = MODE == 0 ? st->conv_iter : st->div_conv_iter;
    if (conv < iter) return;  // converged earlier: grid-uniform no-op
    const u32 i = tile_index(P.n);
    double contrib = 0.0;
    if (i < P.n) {
        const float4 p = xs[i];
        const float4 vi = vs[i];
        float rd = 0.0f;
        const u32 cw = L.cnt[i];
        const u32 cf = cw & 0xFFFFu, cb = cw >> 16;
        if (cf <= L.kf && cb <= L.kb) {
            // PU pairs per trip: loads issued together, terms summed in order
            u32 k = 0;
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
                    rho_dot_fluid(rd, p, vi, q[u], w[u], __uint_as_float(e[u].y), P.mass);
            }
            for (; k < cf; ++k) {
                const uint2 e = ld_pair(L.f + pair_at(k, i, L.kf));
                rho_dot_fluid(rd, p, vi, xs[e.x], vs[e.x], __uint_as_float(e.y), P.mass);
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
                for (int u = 0; u < PU; ++u) rho_dot_boundary(rd, p, vi, q[u], __uint_as_float(e[u].y));
            }
            for (; k < cb; ++k) {
                const uint2 e = ld_pair(L.b + pair_at(k, i, L.kb));
                rho_dot_boundary(rd, p, vi, P.xb[e.x], __uint_as_float(e.y));
            }
        } else {
            for_stencil(P.gf, P.csf, p.x, p.y, p.z, [&](u32 s) {
                if (s == i) return;
                const float4 q = xs[s];
                const float c = spline_gradc(norm3(p.x - q.x, p.y - q.y, p.z - q.z), P.sp);
                rho_dot_fluid(rd, p, vi, q, vs[s], c, P.mass);
            });
            if (P.nb) {
                for_stencil(P.gb, P.csb, p.x, p.y, p.z, [&](u32 s) {
                    const float4 q = P.xb[s];
                    const float c = spline_gradc(norm3(p.x - q.x, p.y - q.y, p.z - q.z), P.sp);
                    rho_dot_boundary(rd, p, vi, q, c);
                });
            }
        }
        const float rho_i = rho[i];
        float kap, err;
        if (MODE == 0) {
            const float rho_pred = rho_i + dt * rd;
            rho_adv[i] = rho_pred;
            err = fmaxf(0.0f, rho_pred - rho0);
            kap = err * alpha[i] / dt2;
        } else {
            rho_adv[i] = rd;
            err = fmaxf(0.0f, rd);
            kap = err * alpha[i] / dt;
        }
        kappa[i] = kap;
        const float khat = pform ? kap / fmaxf(rho_i, 1e-9f) : kap / fmaxf(rho_i * rho_i, 1e-9f);
        xk[i] = make_float4(p.x, p.y, p.z, khat);
        contrib = static_cast<double>(err);
    }
    // per-workgroup partial; the next launch (k_apply) finishes the sum, so
    // no cross-workgroup fence is needed inside this kernel
    const double bsum = block_sum<NT>(contrib, s_red);
    if (threadIdx.x == 0) part[blockIdx.x] = bsum;
}

// Workgroup 0 of k_apply(iter) reduces the partials k_change(iter) wrote, in
// a fixed order, and records the error and the convergence decision
// (dfsph.cu:254-258; the divergence solve tests every iteration).
__device__ __forceinline__ void finish_error(int mode, u32 iter, double thr, u32 n,
                                             const double* __restrict__ part, u32 nparts,
                                             DeviceState* st) {
    __shared__ double s_red[NT / 64];
    double s = 0.0;
    for (u32 b = threadIdx.x; b < nparts; b += NT) s += part[b];
    const double tot = block_sum<NT>(s, s_red);
    if (threadIdx.x == 0) {
        const double e = tot / static_cast<double>(n);
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
__device__ __forceinline__ void kappa_term(float& dx, float& dy, float& dz, const float4& p,
                                           const float4& q, float c, float c2, float dt) {
    const float gx = (p.x - q.x) * c, gy = (p.y - q.y) * c, gz = (p.z - q.z) * c;
    dx -= dt * c2 * gx;
    dy -= dt * c2 * gy;
    dz -= dt * c2 * gz;
}

// Pressure velocity update (dfsph_kernels.cu:237-301); kappa-hat of every
// particle was precomputed by k_change with the same expression the
// reference evaluates per pair.
__global__ __launch_bounds__(NT) void k_apply(Nbr P, PairList L, const float4* __restrict__ xk,
                                              const float* __restrict__ kappa, float dt,
                                              float rho0, int pform, u32 iter, int mode,
                                              double thr, const double* __restrict__ part,
                                              DeviceState* st, float4* __restrict__ vs) {
    const u32 conv = mode == 0 ? st->conv_iter : st->div_conv_iter;
    if (conv < iter) return;  // grid-uniform: this iteration's k_change did not run either
    if (blockIdx.x == 0) finish_error(mode, iter, thr, P.n, part, gridDim.x, st);
    const u32 i = tile_index(P.n);
    if (i >= P.n) return;
    const float4 p = xk[i];
    const float ki = p.w;
    float dx = 0.0f, dy = 0.0f, dz = 0.0f;
    float bf = 0.0f;
    if (P.nb) {
        const float k_i = kappa[i];
        bf = pform ? k_i / fmaxf(rho0, 1e-9f) : k_i / fmaxf(rho0 * rho0, 1e-9f);
    }
    const u32 cw = L.cnt[i];
    const u32 cf = cw & 0xFFFFu, cb = cw >> 16;
    if (cf <= L.kf && cb <= L.kb) {
        u32 k = 0;
        for (; k + PU <= cf; k += PU) {
            uint2 e[PU];
            float4 q[PU];
#pragma unroll
            for (int u = 0; u < PU; ++u) e[u] = ld_pair(L.f + pair_at(k + u, i, L.kf));
#pragma unroll
            for (int u = 0; u < PU; ++u) q[u] = xk[e[u].x];
#pragma unroll
            for (int u = 0; u < PU; ++u)
                kappa_term(dx, dy, dz, p, q[u], __uint_as_float(e[u].y), P.mass * (ki + q[u].w), dt);
        }
        for (; k < cf; ++k) {
            const uint2 e = ld_pair(L.f + pair_at(k, i, L.kf));
            const float4 q = xk[e.x];
            kappa_term(dx, dy, dz, p, q, __uint_as_float(e.y), P.mass * (ki + q.w), dt);
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
                kappa_term(dx, dy, dz, p, q[u], __uint_as_float(e[u].y), q[u].w * (ki + bf), dt);
        }
        for (; k < cb; ++k) {
            const uint2 e = ld_pair(L.b + pair_at(k, i, L.kb));
            const float4 q = P.xb[e.x];
            kappa_term(dx, dy, dz, p, q, __uint_as_float(e.y), q.w * (ki + bf), dt);
        }
    } else {
        for_stencil(P.gf, P.csf, p.x, p.y, p.z, [&](u32 s) {
            if (s == i) return;
            const float4 q = xk[s];
            const float c = spline_gradc(norm3(p.x - q.x, p.y - q.y, p.z - q.z), P.sp);
            kappa_term(dx, dy, dz, p, q, c, P.mass * (ki + q.w), dt);
        });
        if (P.nb) {
            for_stencil(P.gb, P.csb, p.x, p.y, p.z, [&](u32 s) {
                const float4 q = P.xb[s];
                const float c = spline_gradc(norm3(p.x - q.x, p.y - q.y, p.z - q.z), P.sp);
                kappa_term(dx, dy, dz, p, q, c, q.w * (ki + bf), dt);
            });
        }
    }
    float4 v = vs[i];
    v.x += dx;
    v.y += dy;
    v.z += dz;
    vs[i] = v;
}

__device__ __forceinline__ bool finite3(float a, float b, float c) {
    return isfinite(a) && isfinite(b) && isfinite(c);
}

// Per-workgroup (max |v|, NaN count) partials; k_finish_vstats reduces them.
__device__ __forceinline__ void vstats_partial(float mag, u32 bad, float* __restrict__ pmax,
                                               u32* __restrict__ pnan) {
    __shared__ float s_f[NT / 64];
    __shared__ u32 s_u[NT / 64];
    const float bm = block_max<NT>(mag, s_f);
    const u32 bn = block_sum_u32<NT>(bad, s_u);
    if (threadIdx.x == 0) {
        pmax[blockIdx.x] = bm;
        pnan[blockIdx.x] = bn;
    }
}

__global__ __launch_bounds__(NT) void k_finish_vstats(const float* __restrict__ pmax,
                                                      const u32* __restrict__ pnan, u32 nparts,
                                                      DeviceState* st) {
    __shared__ float s_f[NT / 64];
    __shared__ u32 s_u[NT / 64];
    float m = 0.0f;
    u32 c = 0;
    for (u32 b = threadIdx.x; b < nparts; b += NT) {
        m = fmaxf(m, pmax[b]);
        c += pnan[b];
    }
    m = block_max<NT>(m, s_f);
    c = block_sum_u32<NT>(c, s_u);
    if (threadIdx.x == 0) {
        st->max_v = m;
        st->nan_count = c;
    }
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
__global__ __launch_bounds__(NT) void k_advect(u32 n, const float4* __restrict__ xs,
                                               const float4* __restrict__ vs,
                                               const u32* __restrict__ perm, float dt, float vmax,
                                               float lx, float ly, float lz, float hx, float hy,
                                               float hz, StoreOut out,
                                               const float* __restrict__ rho,
                                               const float* __restrict__ alpha,
                                               const float* __restrict__ kappa,
                                               const float* __restrict__ rho_adv, int wb_solve,
                                               float* __restrict__ pmax,
                                               u32* __restrict__ pnan) {
    const u32 s = blockIdx.x * NT + threadIdx.x;
    float mag = 0.0f;
    u32 bad = 0;
    if (s < n) {
        const float4 x4 = xs[s];
        float4 v = vs[s];
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
        const u32 id = perm[s];
        const u64 o = 3ull * id;
        out.x[o] = px;
        out.x[o + 1] = py;
        out.x[o + 2] = pz;
        out.v[o] = v.x;
        out.v[o + 1] = v.y;
        out.v[o + 2] = v.z;
        out.rho[id] = rho[s];
        out.alpha[id] = alpha[s];
        if (wb_solve) {
            out.kappa[id] = kappa[s];
            out.rho_adv[id] = rho_adv[s];
        }
        mag = norm3(v.x, v.y, v.z);
        bad = finite3(px, py, pz) && finite3(v.x, v.y, v.z) ? 0u : 1u;
    }
    vstats_partial(mag, bad, pmax, pnan);
}

// End of the divergence solve: velocities and the recomputed attributes back
// to original order, then max |v|.
__global__ __launch_bounds__(NT) void k_div_writeback(u32 n, const float4* __restrict__ xs,
                                                      const float4* __restrict__ vs,
                                                      const u32* __restrict__ perm, StoreOut out,
                                                      const float* __restrict__ rho,
                                                      const float* __restrict__ alpha,
                                                      const float* __restrict__ kappa_v,
                                                      const float* __restrict__ rho_dot,
                                                      int wb_solve,
                                                      float* __restrict__ pmax,
                                                      u32* __restrict__ pnan) {
    const u32 s = blockIdx.x * NT + threadIdx.x;
    float mag = 0.0f;
    u32 bad = 0;
    if (s < n) {
        const float4 x4 = xs[s];
        const float4 v = vs[s];
        const u32 id = perm[s];
        const u64 o = 3ull * id;
        out.v[o] = v.x;
        out.v[o + 1] = v.y;
        out.v[o + 2] = v.z;
        out.rho[id] = rho[s];
        out.alpha[id] = alpha[s];
        if (wb_solve) {
            out.kappa[id] = kappa_v[s];
            out.rho_adv[id] = rho_dot[s];
        }
        mag = norm3(v.x, v.y, v.z);
        bad = finite3(x4.x, x4.y, x4.z) && finite3(v.x, v.y, v.z) ? 0u : 1u;
    }
    vstats_partial(mag, bad, pmax, pnan);
}

__global__ void k_reset_solve(DeviceState* st, int which) {
    if (threadIdx.x == 0) {
        if (which == 0) st->conv_iter = kNone;
        else st->div_conv_iter = kNone;
    }
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

}  // namespace

const char* kernel_kind_name(int k) {
    static const char* names[kKNumKinds] = {"grid_build",     "reorder",     "density_alpha",
                                            "xsph",           "density_change", "apply_kappa",
                                            "advect",         "div_change",  "div_apply",
                                            "div_writeback"};
    return (k >= 0 && k < kKNumKinds) ? names[k] : "?";
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
    : fluid_(fluid), boundary_(boundary), cfg_(cfg), stream_(stream),
      fluid_grid_(make_grid(cfg, stream)) {
    if (!(cfg_.smoothing_length > 0.0f))
        fail(kInvalidArgument, "DFSPHSolver: smoothing_length must be positive");
    // dfsph.cu:76-78
    const float spacing = 2.0f * cfg_.particle_radius;
    const float v0 = spacing * spacing * spacing;
    mass_ = cfg_.rest_density * v0;

    // dfsph.cu:80-83
    h_density_ = fluid_.register_attribute("dfsph_density", kF32);
    h_alpha_ = fluid_.register_attribute("dfsph_alpha", kF32);
    h_kappa_ = fluid_.register_attribute("dfsph_kappa", kF32);
    h_density_adv_ = fluid_.register_attribute("dfsph_density_adv", kF32);
    if (cfg_.enable_divergence_solve) {
        h_kappa_v_ = fluid_.register_attribute("dfsph_kappa_v", kF32);
        has_kappa_v_ = true;
    }

    const size_t cap = fluid_.capacity();
    fluid_grid_.reserve(cap);
    const size_t c1 = cap ? cap : 1;
    xs_ = DeviceBuffer<float4>(c1);
    vs_ = DeviceBuffer<float4>(c1);
    vs2_ = DeviceBuffer<float4>(c1);
    xk_ = DeviceBuffer<float4>(c1);
    rho_ = DeviceBuffer<float>(c1);
    alpha_ = DeviceBuffer<float>(c1);
    kappa_ = DeviceBuffer<float>(c1);
    rho_adv_ = DeviceBuffer<float>(c1);
    const size_t c64 = (c1 + 63) / 64 * 64;  // the wave-blocked lists pad to whole waves
    pairs_f_ = DeviceBuffer<uint2>(static_cast<size_t>(kPairDepthFluid) * c64);
    pairs_b_ = DeviceBuffer<uint2>(static_cast<size_t>(kPairDepthBoundary) * c64);
    pair_cnt_ = DeviceBuffer<u32>(c1);
    const size_t nblk = div_up(c1, NT);
    part_d_ = DeviceBuffer<double>(nblk);
    part_f_ = DeviceBuffer<float>(nblk);
    part_u_ = DeviceBuffer<u32>(nblk);
    dstate_ = DeviceBuffer<DeviceState>(1);
    dstate_.zero(stream_);
    hstate_ = PinnedBuffer<DeviceState>(1);

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
            const u32 blocks = div_up(nb, NT);
            k_reorder<<<blocks, NT, 0, stream_>>>(boundary_grid_->sorted_indices(),
                                                  boundary_->positions(), nullptr, nb, xb_.data(),
                                                  nullptr, 0, 0.0f, 0.0f, 0.0f, 0.0f);
            WSIM_LAUNCH_CHECK();
            k_boundary_mass<<<blocks, NT, 0, stream_>>>(boundary_grid_->geom(),
                                                        boundary_grid_->cell_start(), xb_.data(), nb,
                                                        Spline::make(cfg_.smoothing_length),
                                                        cfg_.rest_density, mb.data());
            WSIM_LAUNCH_CHECK();
            k_pack_boundary<<<blocks, NT, 0, stream_>>>(
                nb, boundary_grid_->sorted_indices(), mb.data(), xb_.data(),
                static_cast<float*>(boundary_->attribute_data(h_bmass_)));
            WSIM_LAUNCH_CHECK();
            WSIM_HIP(hipStreamSynchronize(stream_));
        }
    }
    WSIM_HIP(hipStreamSynchronize(stream_));
}

void DFSPHSolver::sort_and_reorder(bool gravity, float dt) {
    const u32 n = static_cast<u32>(fluid_.count());
    tbegin(kKGrid);
    fluid_grid_.build(fluid_.positions(), n);
    tend();
    float decay = 1.0f, gdx = 0.0f, gdy = 0.0f, gdz = 0.0f;
    if (gravity) {
        decay = std::exp(-cfg_.damping * dt);  // dfsph_kernels.cu:351, once on the host
        gdx = cfg_.gravity[0] * dt;
        gdy = cfg_.gravity[1] * dt;
        gdz = cfg_.gravity[2] * dt;
    }
    tbegin(kKReorder);
    k_reorder<<<div_up(n, NT), NT, 0, stream_>>>(fluid_grid_.sorted_indices(), fluid_.positions(),
                                                 fluid_.velocities(), n, xs_.data(), vs_.data(),
                                                 gravity ? 1 : 0, decay, gdx, gdy, gdz);
    WSIM_LAUNCH_CHECK();
    tend();
}

static Nbr make_nbr(const CellGrid& fg, const CellGrid* bg, const float4* xb, size_t nb,
                    const DFSPHConfig& cfg, float mass, u32 n) {
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
    P.n = n;
    return P;
}

PairList DFSPHSolver::pair_list() {
    PairList L;
    L.f = pairs_f_.data();
    L.b = pairs_b_.data();
    L.cnt = pair_cnt_.data();
    L.kf = kPairDepthFluid;
    L.kb = nb_ ? kPairDepthBoundary : 0u;
    return L;
}

void DFSPHSolver::density_solve(float dt) {
    const u32 n = static_cast<u32>(fluid_.count());
    const u32 blocks = div_up(n, NT);
    const Nbr P = make_nbr(fluid_grid_, boundary_grid_.get(), xb_.data(), nb_, cfg_, mass_, n);
    const PairList L = pair_list();
    const double thr = static_cast<double>(cfg_.eta_density * cfg_.rest_density);
    const float dt2 = dt * dt;
    DeviceState* st = dstate_.data();
    k_reset_solve<<<1, 64, 0, stream_>>>(st, 0);
    WSIM_LAUNCH_CHECK();
    const u32 maxit = cfg_.max_density_iters;
    u32 it = 0, chunk = 2;
    while (it < maxit) {
        const u32 end = (maxit - it < chunk) ? maxit : it + chunk;
        for (; it < end; ++it) {
            tbegin(kKDensityChange, static_cast<int>(it));
            k_change<0><<<blocks, NT, 0, stream_>>>(P, L, xs_.data(), vs_.data(), rho_.data(),
                                                    alpha_.data(), dt, dt2, cfg_.rest_density,
                                                    static_cast<int>(cfg_.pressure_form), it,
                                                    st, part_d_.data(), rho_adv_.data(),
                                                    kappa_.data(), xk_.data());
            WSIM_LAUNCH_CHECK();
            tend();
            tbegin(kKApplyKappa, static_cast<int>(it));
            k_apply<<<blocks, NT, 0, stream_>>>(P, L, xk_.data(), kappa_.data(), dt,
                                                cfg_.rest_density,
                                                static_cast<int>(cfg_.pressure_form), it, 0, thr,
                                                part_d_.data(), st, vs_.data());
            WSIM_LAUNCH_CHECK();
            tend();
        }
        if (it >= maxit) break;
        WSIM_HIP(hipMemcpyAsync(&hstate_[0].conv_iter, &st->conv_iter, sizeof(u32),
                                hipMemcpyDeviceToHost, stream_));
        WSIM_HIP(hipStreamSynchronize(stream_));
        if (hstate_[0].conv_iter != kNone) break;
        chunk = chunk * 2 > 32 ? 32 : chunk * 2;
    }
}

void DFSPHSolver::divergence_solve(float dt) {
    const u32 n = static_cast<u32>(fluid_.count());
    const u32 blocks = div_up(n, NT);
    const Nbr P = make_nbr(fluid_grid_, boundary_grid_.get(), xb_.data(), nb_, cfg_, mass_, n);
    const PairList L = pair_list();
    const double thr = static_cast<double>(cfg_.eta_divergence * cfg_.rest_density);
    DeviceState* st = dstate_.data();
    k_reset_solve<<<1, 64, 0, stream_>>>(st, 1);
    WSIM_LAUNCH_CHECK();
    const u32 maxit = cfg_.max_divergence_iters;
    u32 it = 0, chunk = 1;
    while (it < maxit) {
        const u32 end = (maxit - it < chunk) ? maxit : it + chunk;
        for (; it < end; ++it) {
            tbegin(kKDivChange, static_cast<int>(it));
            k_change<1><<<blocks, NT, 0, stream_>>>(P, L, xs_.data(), vs_.data(), rho_.data(),
                                                    alpha_.data(), dt, dt * dt, cfg_.rest_density,
                                                    static_cast<int>(cfg_.pressure_form), it,
                                                    st, part_d_.data(), rho_adv_.data(),
                                                    kappa_.data(), xk_.data());
            WSIM_LAUNCH_CHECK();
            tend();
            tbegin(kKDivApply, static_cast<int>(it));
            k_apply<<<blocks, NT, 0, stream_>>>(P, L, xk_.data(), kappa_.data(), dt,
                                                cfg_.rest_density,
                                                static_cast<int>(cfg_.pressure_form), it, 1, thr,
                                                part_d_.data(), st, vs_.data());
            WSIM_LAUNCH_CHECK();
            tend();
        }
        if (it >= maxit) break;
        WSIM_HIP(hipMemcpyAsync(&hstate_[0].div_conv_iter, &st->div_conv_iter, sizeof(u32),
                                hipMemcpyDeviceToHost, stream_));
        WSIM_HIP(hipStreamSynchronize(stream_));
        if (hstate_[0].div_conv_iter != kNone) break;
        chunk = chunk * 2 > 32 ? 32 : chunk * 2;
    }
}

void DFSPHSolver::step(float dt) {
    const size_t nsz = fluid_.count();
    if (nsz == 0) return;  // dfsph.cu:121-122
    const u32 n = static_cast<u32>(nsz);
    const u32 blocks = div_up(n, NT);

    sort_and_reorder(true, dt);
    Nbr P = make_nbr(fluid_grid_, boundary_grid_.get(), xb_.data(), nb_, cfg_, mass_, n);
    tbegin(kKDensityAlpha);
    k_density_alpha<<<blocks, NT, 0, stream_>>>(P, pair_list(), xs_.data(), rho_.data(),
                                                alpha_.data());
    WSIM_LAUNCH_CHECK();
    tend();
    if (cfg_.viscosity > 0.0f) {  // dfsph.cu:198
        tbegin(kKXsph);
        k_xsph<<<blocks, NT, 0, stream_>>>(P, xs_.data(), vs_.data(), rho_.data(), cfg_.viscosity,
                                           vs2_.data());
        WSIM_LAUNCH_CHECK();
        tend();
        std::swap(vs_, vs2_);
    }
    density_solve(dt);

    StoreOut out;
    out.x = fluid_.positions();
    out.v = fluid_.velocities();
    out.rho = static_cast<float*>(fluid_.attribute_data(h_density_));
    out.alpha = static_cast<float*>(fluid_.attribute_data(h_alpha_));
    out.kappa = static_cast<float*>(fluid_.attribute_data(h_kappa_));
    out.rho_adv = static_cast<float*>(fluid_.attribute_data(h_density_adv_));
    const float eps = 1e-3f;  // dfsph_kernels.cu:377
    tbegin(kKAdvect);
    k_advect<<<blocks, NT, 0, stream_>>>(
        n, xs_.data(), vs_.data(), fluid_grid_.sorted_indices(), dt, 5.0f,
        cfg_.domain_min[0] + eps, cfg_.domain_min[1] + eps, cfg_.domain_min[2] + eps,
        cfg_.domain_max[0] - eps, cfg_.domain_max[1] - eps, cfg_.domain_max[2] - eps, out,
        rho_.data(), alpha_.data(), kappa_.data(), rho_adv_.data(),
        cfg_.max_density_iters > 0 ? 1 : 0, part_f_.data(), part_u_.data());
    WSIM_LAUNCH_CHECK();
    tend();
    if (!cfg_.enable_divergence_solve) {
        k_finish_vstats<<<1, NT, 0, stream_>>>(part_f_.data(), part_u_.data(), blocks, dstate_.data());
        WSIM_LAUNCH_CHECK();
    }

    if (cfg_.enable_divergence_solve) {
        sort_and_reorder(false, dt);
        P = make_nbr(fluid_grid_, boundary_grid_.get(), xb_.data(), nb_, cfg_, mass_, n);
        tbegin(kKDensityAlpha);
        k_density_alpha<<<blocks, NT, 0, stream_>>>(P, pair_list(), xs_.data(), rho_.data(),
                                                    alpha_.data());
        WSIM_LAUNCH_CHECK();
        tend();
        divergence_solve(dt);
        out.kappa = static_cast<float*>(fluid_.attribute_data(h_kappa_v_));
        tbegin(kKDivWriteback);
        k_div_writeback<<<blocks, NT, 0, stream_>>>(
            n, xs_.data(), vs_.data(), fluid_grid_.sorted_indices(), out, rho_.data(),
            alpha_.data(), kappa_.data(), rho_adv_.data(), cfg_.max_divergence_iters > 0 ? 1 : 0,
            part_f_.data(), part_u_.data());
        WSIM_LAUNCH_CHECK();
        tend();
        k_finish_vstats<<<1, NT, 0, stream_>>>(part_f_.data(), part_u_.data(), blocks, dstate_.data());
        WSIM_LAUNCH_CHECK();
    }
    read_state();
}

void DFSPHSolver::read_state() {
    WSIM_HIP(hipMemcpyAsync(hstate_.data(), dstate_.data(), sizeof(DeviceState),
                            hipMemcpyDeviceToHost, stream_));
    WSIM_HIP(hipStreamSynchronize(stream_));
    const DeviceState& h = hstate_[0];
    const u32 maxit = cfg_.max_density_iters;
    stats_.density_iters = maxit == 0 ? 0 : (h.conv_iter == kNone ? maxit : h.conv_iter + 1);
    stats_.mean_density_error = maxit == 0 ? 0.0 : h.err;
    if (cfg_.enable_divergence_solve) {
        const u32 mv = cfg_.max_divergence_iters;
        stats_.divergence_iters = mv == 0 ? 0 : (h.div_conv_iter == kNone ? mv : h.div_conv_iter + 1);
        stats_.mean_divergence_error = mv == 0 ? 0.0 : h.div_err;
    }
    stats_.max_velocity = h.max_v;
    stats_.nan_count = h.nan_count;
    stats_.substeps += 1;
    if (timing_) {
        int valid[kKNumKinds];
        for (int k = 0; k < kKNumKinds; ++k) valid[k] = std::numeric_limits<int>::max();
        valid[kKDensityChange] = valid[kKApplyKappa] = static_cast<int>(stats_.density_iters);
        valid[kKDivChange] = valid[kKDivApply] = static_cast<int>(stats_.divergence_iters);
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

}  // namespace wsim
