// adjoint.hip -- reverse-mode derivative of the DFSPH substep.
//
// The substep of solver.hip (solvers/dfsph/src/dfsph.cu:120-133) as a map
// (x, v, viscosity) -> (x', v') with its discrete structure frozen: the
// neighbour sets of the grid built from x, the number of density iterations
// K, the max(0, .) masks of kappa and the |v| and domain clamps of advection,
// all as the recorded forward pass took them.  With r = x_i - x_j,
// grad W = r c(|r|) and H = d(grad W)/dx_i = c I + (c'/|r|) r r^T:
//
//   forward                                  adjoint (reverse order)
//   v0 = v decay + g dt                      v_bar = decay v0_bar
//   rho_i = sum m_j W_ij                     x_bar_i += sum m (rho_bar_i + rho_bar_j) grad W_ij
//   alpha_i = rho_i / (|G_i|^2 + Q_i)        G_bar = 2 D_bar G, Q_bar = D_bar; H terms into x_bar
//   v1 = v0 + sum e_ij (v0_j - v0_i)         v0_bar = v1_bar + sum e_ij (v1_bar_j - v1_bar_i);
//     e_ij = eps m W_ij / avg(rho)             e_ij's rho, x and eps sensitivities
//   K Jacobi sweeps                          K transposed sweeps, last first:
//     rho*_i = rho_i + dt rho_dot_i            kappa-hat_bar_i = -dt sum m (w_bar_i - w_bar_j).grad W
//     kappa_i = max(0, rho*_i - rho0)           kappa_bar -> rho*_bar (mask) -> rho_dot_bar = dt rho*_bar
//       alpha_i / dt^2                         w_bar_i += sum m (rho_dot_bar_i + rho_dot_bar_j) grad W
//     w_i -= dt sum m (k_i + k_j) grad W       plus the H terms of both sweeps into x_bar
//   advect with clamps                       masks of the recorded pass
//
// Boundary particles are fixed; their terms enter the fluid particle's own
// sums.  Adjoint variables and every accumulation are binary64; per-pair
// geometry is evaluated in binary64 from the binary32 positions of the
// recorded pass.  Pair sets are symmetric (every fluid particle lies inside
// the grid, which the domain clamp guarantees), so each particle gathers the
// transposed terms of its own pairs and no atomics are needed.
#include <algorithm>
#include <cmath>
#include <vector>

#include "stencil.hpp"
#include "wsim_internal.hpp"

namespace wsim {
namespace {

using namespace dev;

constexpr int NA = 256;  // workgroup size of the adjoint kernels (one tile)
constexpr int kAdjointLanes = 4;  // lanes per particle in the per-iteration kernels

struct D3 {
    double x, y, z;
};
__device__ __forceinline__ D3 d3(const double4& a) { return {a.x, a.y, a.z}; }
__device__ __forceinline__ D3 d3f(const float4& a) { return {a.x, a.y, a.z}; }
__device__ __forceinline__ double dot(const D3& a, const D3& b) {
    return a.x * b.x + a.y * b.y + a.z * b.z;
}
__device__ __forceinline__ D3 sub(const D3& a, const D3& b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
__device__ __forceinline__ void axpy(D3& acc, double s, const D3& v) {
    acc.x += s * v.x;
    acc.y += s * v.y;
    acc.z += s * v.z;
}

// Pair geometry in binary64 from the binary32 difference the forward pass
// forms: grad W = r c, H v = c v + h2 (r.v) r.
struct Geo {
    D3 r;
    double c, h2;
    __device__ __forceinline__ D3 grad() const { return {r.x * c, r.y * c, r.z * c}; }
    __device__ __forceinline__ D3 hess(const D3& v) const {
        const double t = h2 * dot(r, v);
        return {c * v.x + t * r.x, c * v.y + t * r.y, c * v.z + t * r.z};
    }
};

__device__ __forceinline__ Geo pair_geo(const float4& p, const float4& q, const Spline& sp) {
    Geo g;
    g.r = {static_cast<double>(p.x - q.x), static_cast<double>(p.y - q.y),
           static_cast<double>(p.z - q.z)};
    const double d = sqrt(dot(g.r, g.r));
    const double l = sp.l, sigma = sp.sigma;
    const double q_ = d / l;
    g.c = 0.0;
    g.h2 = 0.0;
    if (d < 1e-12 || q_ >= 2.0) return g;
    double dw, d2w;
    if (q_ < 1.0) {
        dw = sigma * (-3.0 * q_ + 2.25 * q_ * q_);
        d2w = sigma * (-3.0 + 4.5 * q_);
    } else {
        const double t = 2.0 - q_;
        dw = -0.75 * sigma * t * t;
        d2w = 1.5 * sigma * t;
    }
    g.c = dw / (l * d);
    g.h2 = (d2w / (l * l) - g.c) / (d * d);
    return g;
}

// Every pair of particle i with a non-zero kernel gradient, in the forward
// pass's order: ff(slot, c) for fluid neighbours, fb(slot, c) for boundary
// particles; from the pair lists, or the stencil walk for overflowed lists.
template <class FF, class FB>
__device__ __forceinline__ void for_pairs(const Nbr& P, const PairList& L, const SceneSeg& G,
                                          u32 i, const float4& p, const float4* __restrict__ xs,
                                          FF&& ff, FB&& fb) {
    const u32 cw = L.cnt[i];
    const u32 cf = cw & 0xFFFFu, cb = cw >> 16;
    if (cf <= L.kf && cb <= L.kb) {
        const StageRanges R = L.plan[i / kTile];
        for (u32 k = 0; k < cf; ++k) {
            u32 j;
            float c;
            fluid_entry(L, R, k, i, j, c);
            ff(j, R.total ? pair_c(p, xs[j], P.sp) : c);
        }
        for (u32 k = 0; k < cb; ++k) {
            const uint2 e = ld_pair(L.b + pair_at(k, i, L.kb));
            fb(e.x, __uint_as_float(e.y));
        }
        return;
    }
    for_stencil(P.gf, P.csf + G.key_base, G.off - G.start, p.x, p.y, p.z, [&](u32 s) {
        if (s == i) return;
        const float4 q = xs[s];
        const float r2 = dist2(p.x - q.x, p.y - q.y, p.z - q.z);
        if (r2 >= P.sp.far2) return;
        const float c = spline_gradc(pair_sqrt(r2, P.sp), P.sp);
        if (c != 0.0f) ff(s, c);
    });
    if (P.nb) {
        for_stencil(P.gb, P.csb, 0u, p.x, p.y, p.z, [&](u32 s) {
            const float4 q = P.xb[s];
            const float r2 = dist2(p.x - q.x, p.y - q.y, p.z - q.z);
            if (r2 >= P.sp.far2) return;
            const float c = spline_gradc(pair_sqrt(r2, P.sp), P.sp);
            if (c != 0.0f) fb(s, c);
        });
    }
}

// Pairs of particle i split over a group of G lanes: lane g takes pairs
// g, g + G, ... of the lists; an overflowed list is walked by lane 0 alone.
template <int G, class FF, class FB>
__device__ __forceinline__ void for_pairs_split(const Nbr& P, const PairList& L, const SceneSeg& Sg,
                                                u32 i, u32 g, const float4& p,
                                                const float4* __restrict__ xs, FF&& ff, FB&& fb) {
    const u32 cw = L.cnt[i];
    const u32 cf = cw & 0xFFFFu, cb = cw >> 16;
    if (cf <= L.kf && cb <= L.kb) {
        const StageRanges R = L.plan[i / kTile];
        for (u32 k = g; k < cf; k += G) {
            u32 j;
            float c;
            fluid_entry(L, R, k, i, j, c);
            ff(j, R.total ? pair_c(p, xs[j], P.sp) : c);
        }
        for (u32 k = g; k < cb; k += G) {
            const uint2 e = ld_pair(L.b + pair_at(k, i, L.kb));
            fb(e.x, __uint_as_float(e.y));
        }
    } else if (g == 0) {
        for_pairs(P, L, Sg, i, p, xs, ff, fb);
    }
}

// Sum over the G lanes of a group (fixed butterfly order).
template <int G> __device__ __forceinline__ double group_sum(double v) {
#pragma unroll
    for (int o = 1; o < G; o <<= 1) v += __shfl_xor(v, o);
    return v;
}
template <int G> __device__ __forceinline__ void group_sum3(D3& a) {
    a.x = group_sum<G>(a.x);
    a.y = group_sum<G>(a.y);
    a.z = group_sum<G>(a.z);
}

// acc[k] += part[k]: viscosity partials accumulated over substeps on device.
__global__ __launch_bounds__(NS) void k_add_parts(const double* __restrict__ part, u32 n,
                                                  double* __restrict__ acc) {
    const u32 k = blockIdx.x * NS + threadIdx.x;
    if (k < n) acc[k] += part[k];
}

// Advection (dfsph_kernels.cu:363-386) transposed with the recorded clamp
// masks; gathers the output adjoints from original order and opens the
// per-particle accumulators.
__global__ __launch_bounds__(NA) void k_adj_advect(Nbr P, const u32* __restrict__ perm,
                                                   const float4* __restrict__ xs,
                                                   const float4* __restrict__ wk, float vmax,
                                                   float lx, float ly, float lz, float hx,
                                                   float hy, float hz,
                                                   const double* __restrict__ gx,
                                                   const double* __restrict__ gv,
                                                   double4* __restrict__ xbar,
                                                   double4* __restrict__ wbar,
                                                   double* __restrict__ rhobar,
                                                   double* __restrict__ alphabar) {
    const u32 i = blockIdx.x * NA + threadIdx.x;
    if (i >= P.nslots) return;
    const SceneSeg& G = P.seg[scene_of_slot(P, i)];
    const u32 l = i - G.off;
    if (l >= G.n) return;
    const float dt = G.dt;
    const u32 id = perm[G.start + l];
    const u64 o = 3ull * id;
    const D3 yb = {gx[o], gx[o + 1], gx[o + 2]};
    const D3 vb = {gv[o], gv[o + 1], gv[o + 2]};
    const float4 x4 = xs[i];
    const float4 w = wk[i];
    const float sp = norm3(w.x, w.y, w.z);
    const bool mclamp = sp > vmax;
    float4 v = w;
    if (mclamp) {
        const float f = vmax / sp;
        v.x *= f;
        v.y *= f;
        v.z *= f;
    }
    float px = x4.x + v.x * dt, py = x4.y + v.y * dt, pz = x4.z + v.z * dt;
    bool cx = false, cy = false, cz = false, zx = false, zy = false, zz = false;
    if (px < lx) { px = lx; cx = true; if (v.x < 0) { v.x = 0; zx = true; } }
    if (py < ly) { py = ly; cy = true; if (v.y < 0) { v.y = 0; zy = true; } }
    if (pz < lz) { pz = lz; cz = true; if (v.z < 0) { v.z = 0; zz = true; } }
    if (px > hx) { px = hx; cx = true; if (v.x > 0) { v.x = 0; zx = true; } }
    if (py > hy) { py = hy; cy = true; if (v.y > 0) { v.y = 0; zy = true; } }
    if (pz > hz) { pz = hz; cz = true; if (v.z > 0) { v.z = 0; zz = true; } }
    const double ddt = dt;
    xbar[i] = make_double4(cx ? 0.0 : yb.x, cy ? 0.0 : yb.y, cz ? 0.0 : yb.z, 0.0);
    D3 wc = {(cx ? 0.0 : ddt * yb.x) + (zx ? 0.0 : vb.x), (cy ? 0.0 : ddt * yb.y) + (zy ? 0.0 : vb.y),
             (cz ? 0.0 : ddt * yb.z) + (zz ? 0.0 : vb.z)};
    if (mclamp) {  // w_c = vmax w / |w|: Jacobian (vmax/s)(I - w^ w^T)
        const D3 wd = d3f(w);
        const double s = sqrt(dot(wd, wd));
        const D3 u = {wd.x / s, wd.y / s, wd.z / s};
        const double k = static_cast<double>(vmax) / s, t = dot(u, wc);
        wc = {k * (wc.x - t * u.x), k * (wc.y - t * u.y), k * (wc.z - t * u.z)};
    }
    wbar[i] = make_double4(wc.x, wc.y, wc.z, 0.0);
    rhobar[i] = 0.0;
    alphabar[i] = 0.0;
}

// Transposed apply-kappa of one iteration (dfsph_kernels.cu:237-301) and the
// pointwise part of its kappa (dfsph_kernels.cu:226-231): from w_bar^(k+1)
// to rho_dot_bar^k, with the rho, alpha and H-term sensitivities.  G lanes
// per particle share its pairs.
template <int G>
__global__ __launch_bounds__(NA) void k_adj_kappa(Nbr P, PairList L, const float4* __restrict__ xs,
                                                  const double4* __restrict__ wbar,
                                                  const float4* __restrict__ xk,
                                                  const float* __restrict__ pk,
                                                  const float* __restrict__ rho,
                                                  const float* __restrict__ alpha, float rho0,
                                                  int pform, double* __restrict__ rdbar,
                                                  double* __restrict__ rhobar,
                                                  double* __restrict__ alphabar,
                                                  double4* __restrict__ xbar) {
    const u32 t = slot_index<NA>(P.remap);
    const u32 i = t / G, g = t % G;
    double kbar = 0.0, bbar = 0.0;
    D3 xa = {0.0, 0.0, 0.0};
    bool valid = false;
    const SceneSeg* Gp = nullptr;
    float4 p = make_float4(0.f, 0.f, 0.f, 0.f);
    float bf = 0.0f;
    if (i < P.nslots) {
        Gp = &P.seg[scene_of_slot(P, i)];
        valid = i - Gp->off < Gp->n;
    }
    if (valid) {
        const SceneSeg& S = *Gp;
        const double ddt = S.dt;
        p = xk[i];
        const float kap = fmaxf(0.0f, pk[i] - rho0) * alpha[i] / S.dt2;
        bf = pform ? kap / fmaxf(rho0, 1e-9f) : kap / fmaxf(rho0 * rho0, 1e-9f);
        const double ki = p.w;
        const D3 wi = d3(wbar[i]);
        const double m = P.mass;
        for_pairs_split<G>(
            P, L, S, i, g, p, xs,
            [&](u32 j, float) {
                const float4 q = xk[j];
                const Geo ge = pair_geo(p, q, P.sp);
                const D3 dw = sub(wi, d3(wbar[j]));
                kbar -= ddt * m * dot(dw, ge.grad());
                axpy(xa, -ddt * m * (ki + static_cast<double>(q.w)), ge.hess(dw));
            },
            [&](u32 b, float) {
                const float4 q = P.xb[b];
                const Geo ge = pair_geo(p, q, P.sp);
                const double mb = q.w;
                const double tt = -ddt * mb * dot(wi, ge.grad());
                kbar += tt;
                bbar += tt;
                axpy(xa, -ddt * mb * (ki + static_cast<double>(bf)), ge.hess(wi));
            });
    }
    kbar = group_sum<G>(kbar);
    bbar = group_sum<G>(bbar);
    group_sum3<G>(xa);
    if (!valid || g != 0) return;
    const SceneSeg& S = *Gp;
    const float rho_i = rho[i];
    const float a_i = alpha[i];
    const float pr = pk[i];
    const float over = fmaxf(0.0f, pr - rho0);
    const float kap = over * a_i / S.dt2;
    // kappa-hat = kappa / max(rho^2, 1e-9) (or / max(rho, 1e-9)); boundary factor
    // kappa / max(rho0^2, 1e-9) (or / max(rho0, 1e-9))
    const double dr = rho_i;
    double kapbar, rb = 0.0;
    if (pform) {
        const float den = fmaxf(rho_i, 1e-9f);
        kapbar = kbar / den + bbar / fmaxf(rho0, 1e-9f);
        if (rho_i > 1e-9f) rb -= kbar * kap / (dr * dr);
    } else {
        const float den = fmaxf(rho_i * rho_i, 1e-9f);
        kapbar = kbar / den + bbar / fmaxf(rho0 * rho0, 1e-9f);
        if (rho_i * rho_i > 1e-9f) rb -= 2.0 * kbar * kap / (dr * dr * dr);
    }
    double pbar = 0.0, ab = 0.0;
    if (pr - rho0 > 0.0f) {
        const double dt2 = S.dt2;
        pbar = kapbar * static_cast<double>(a_i) / dt2;
        ab = kapbar * static_cast<double>(over) / dt2;
    }
    rdbar[i] = static_cast<double>(S.dt) * pbar;
    rhobar[i] += rb + pbar;
    alphabar[i] += ab;
    const double4 x0 = xbar[i];
    xbar[i] = make_double4(x0.x + xa.x, x0.y + xa.y, x0.z + xa.z, 0.0);
}

// Transposed density change (dfsph_kernels.cu:196-224): w_bar^k from
// w_bar^(k+1) and rho_dot_bar^k, with its H terms.  G lanes per particle.
template <int G>
__global__ __launch_bounds__(NA) void k_adj_rhodot(Nbr P, PairList L, const float4* __restrict__ xs,
                                                   const float4* __restrict__ wk,
                                                   const double* __restrict__ rdbar,
                                                   double4* __restrict__ wbar,
                                                   double4* __restrict__ xbar) {
    const u32 t = slot_index<NA>(P.remap);
    const u32 i = t / G, g = t % G;
    D3 acc = {0.0, 0.0, 0.0}, xa = {0.0, 0.0, 0.0};
    bool valid = false;
    if (i < P.nslots) {
        const SceneSeg& S = P.seg[scene_of_slot(P, i)];
        valid = i - S.off < S.n;
        if (valid) {
            const float4 p = xs[i];
            const D3 wi = d3f(wk[i]);
            const double ri = rdbar[i];
            const double m = P.mass;
            for_pairs_split<G>(
                P, L, S, i, g, p, xs,
                [&](u32 j, float) {
                    const Geo ge = pair_geo(p, xs[j], P.sp);
                    const double sc = m * (ri + rdbar[j]);
                    axpy(acc, sc, ge.grad());
                    axpy(xa, sc, ge.hess(sub(wi, d3f(wk[j]))));
                },
                [&](u32 b, float) {
                    const float4 q = P.xb[b];
                    const Geo ge = pair_geo(p, q, P.sp);
                    const double sc = static_cast<double>(q.w) * ri;
                    axpy(acc, sc, ge.grad());
                    axpy(xa, sc, ge.hess(wi));
                });
        }
    }
    group_sum3<G>(acc);
    group_sum3<G>(xa);
    if (!valid || g != 0) return;
    const double4 w0 = wbar[i];
    wbar[i] = make_double4(w0.x + acc.x, w0.y + acc.y, w0.z + acc.z, 0.0);
    const double4 x0 = xbar[i];
    xbar[i] = make_double4(x0.x + xa.x, x0.y + xa.y, x0.z + xa.z, 0.0);
}

// Transposed Jacobi XSPH (dfsph_kernels.cu:305-341) over the pairs of the
// forward pass (every pair with W != 0: the fluid list, or the stencil walk
// for particles whose list is incomplete), G lanes per particle.  Writes
// v0_bar, adds the rho and x sensitivities of e_ij and one binary64
// viscosity partial per workgroup.  With the viscosity <= 0 the forward pass
// skipped XSPH (dfsph.cu:198): v0_bar = v1_bar and no viscosity derivative.
template <int G>
__global__ __launch_bounds__(NA) void k_adj_xsph(Nbr P, PairList L, const float4* __restrict__ xs,
                                                 const float4* __restrict__ v0,
                                                 const double4* __restrict__ v1bar,
                                                 const float* __restrict__ rho, float eps,
                                                 double4* __restrict__ v0bar,
                                                 double* __restrict__ rhobar,
                                                 double4* __restrict__ xbar,
                                                 double* __restrict__ part) {
    __shared__ double s_red[NA / 64];
    const u32 t = slot_index<NA>(P.remap);
    const u32 i = t / G, g = t % G;
    double ep = 0.0;
    D3 acc = {0.0, 0.0, 0.0}, xa = {0.0, 0.0, 0.0};
    double rb = 0.0;
    bool valid = false;
    D3 bi = {0.0, 0.0, 0.0};
    if (i < P.nslots) {
        const SceneSeg& S = P.seg[scene_of_slot(P, i)];
        valid = i - S.off < S.n;
        if (valid) {
            const float4 p = xs[i];
            const D3 vi = d3f(v0[i]);
            bi = d3(v1bar[i]);
            const float rho_i = rho[i];
            const bool on = eps > 0.0f;
            const double de = eps, m = P.mass;
            auto term = [&](u32 s, const float4& q, float r2) {
                const float w = spline_W(pair_sqrt(r2, P.sp), P.sp);
                if (w == 0.0f) return;
                const float avg = 0.5f * (rho_i + rho[s]);
                const float am = fmaxf(avg, 1e-9f);
                const double mw = static_cast<double>(P.mass / am) * static_cast<double>(w);
                const D3 dv = sub(d3f(v0[s]), vi);
                if (on) {
                    ep += dot(bi, dv) * mw;
                    const D3 bj = d3(v1bar[s]);
                    const double e = de * mw;
                    axpy(acc, e, sub(bj, bi));
                    const double sd = dot(sub(bi, bj), dv);
                    if (avg > 1e-9f) rb -= sd * e / (2.0 * static_cast<double>(am));
                    const Geo gg = pair_geo(p, q, P.sp);
                    axpy(xa, sd * de * m / static_cast<double>(am), gg.grad());
                }
            };
            const u32 cf = L.cnt[i] & 0xFFFFu;
            if (cf <= L.kf) {
                // a complete fluid list holds every in-support pair but the self
                // pair (the forward XSPH reads the same list)
                const StageRanges R = L.plan[i / kTile];
                for (u32 k = g; k < cf; k += G) {
                    u32 j;
                    float c;
                    fluid_entry(L, R, k, i, j, c);
                    const float4 q = xs[j];
                    term(j, q, dist2(p.x - q.x, p.y - q.y, p.z - q.z));
                }
            } else if (g == 0) {
                for_stencil(P.gf, P.csf + S.key_base, S.off - S.start, p.x, p.y, p.z, [&](u32 s) {
                    if (s == i) return;
                    const float4 q = xs[s];
                    const float r2 = dist2(p.x - q.x, p.y - q.y, p.z - q.z);
                    if (r2 >= P.sp.far2) return;
                    term(s, q, r2);
                });
            }
        }
    }
    group_sum3<G>(acc);
    group_sum3<G>(xa);
    rb = group_sum<G>(rb);
    if (valid && g == 0) {
        v0bar[i] = make_double4(bi.x + acc.x, bi.y + acc.y, bi.z + acc.z, 0.0);
        rhobar[i] += rb;
        const double4 x0 = xbar[i];
        xbar[i] = make_double4(x0.x + xa.x, x0.y + xa.y, x0.z + xa.z, 0.0);
    }
    // one viscosity partial per workgroup, summed in a fixed order
    const double bs = block_sum<NA>(ep, s_red);
    if (threadIdx.x == 0) part[blockIdx.x] = bs;
}

// alpha = rho / (|G|^2 + Q) (dfsph_kernels.cu:117-169): recomputes G and the
// denominator exactly as the forward pass and turns alpha_bar into
// G_bar = 2 D_bar G, Q_bar = D_bar (and rho_bar += alpha_bar / D).
__global__ __launch_bounds__(NA) void k_adj_alpha(Nbr P, PairList L, const float4* __restrict__ xs,
                                                  const float* __restrict__ rho,
                                                  const double* __restrict__ alphabar,
                                                  double* __restrict__ rhobar,
                                                  double4* __restrict__ gbar) {
    const u32 i = slot_index<NA>(P.remap);
    if (i >= P.nslots) return;
    const SceneSeg& G = P.seg[scene_of_slot(P, i)];
    if (i - G.off >= G.n) return;
    const float4 p = xs[i];
    float sx = 0.0f, sy = 0.0f, sz = 0.0f, sq = 0.0f;
    for_pairs(
        P, L, G, i, p, xs,
        [&](u32 j, float c) {
            const float4 q = xs[j];
            const float rx = p.x - q.x, ry = p.y - q.y, rz = p.z - q.z;
            const float mx = P.mass * (rx * c), my = P.mass * (ry * c), mz = P.mass * (rz * c);
            sx += mx;
            sy += my;
            sz += mz;
            sq += mx * mx + my * my + mz * mz;
        },
        [&](u32 b, float c) {
            const float4 q = P.xb[b];
            const float rx = p.x - q.x, ry = p.y - q.y, rz = p.z - q.z;
            const float mx = q.w * (rx * c), my = q.w * (ry * c), mz = q.w * (rz * c);
            sx += mx;
            sy += my;
            sz += mz;
            sq += mx * mx + my * my + mz * mz;
        });
    const float denom = sx * sx + sy * sy + sz * sz + sq;
    const double ab = alphabar[i];
    if (denom > 1e-9f && ab != 0.0) {
        const double D = denom;
        rhobar[i] += ab / D;
        const double Db = -ab * static_cast<double>(rho[i]) / (D * D);
        gbar[i] = make_double4(2.0 * Db * sx, 2.0 * Db * sy, 2.0 * Db * sz, Db);
    } else {
        gbar[i] = make_double4(0.0, 0.0, 0.0, 0.0);
    }
}

// Position sensitivities of rho, G and Q (dfsph_kernels.cu:68-169), then the
// input adjoints in original order: x_bar, and v_bar = decay v0_bar
// (dfsph_kernels.cu:351-354).
template <int G>
__global__ __launch_bounds__(NA) void k_adj_geom(Nbr P, PairList L, const u32* __restrict__ perm,
                                                 const float4* __restrict__ xs,
                                                 const double* __restrict__ rhobar,
                                                 const double4* __restrict__ gbar,
                                                 const double4* __restrict__ xbar,
                                                 const double4* __restrict__ v0bar,
                                                 double* __restrict__ gx,
                                                 double* __restrict__ gv) {
    const u32 t = slot_index<NA>(P.remap);
    const u32 i = t / G, g = t % G;
    bool valid = false;
    const SceneSeg* Sp = nullptr;
    D3 xa = {0.0, 0.0, 0.0};
    if (i < P.nslots) {
        Sp = &P.seg[scene_of_slot(P, i)];
        valid = i - Sp->off < Sp->n;
    }
    if (valid) {
        const SceneSeg& S = *Sp;
        const float4 p = xs[i];
        const double ri = rhobar[i];
        const double4 gi4 = gbar[i];
        const D3 gi = d3(gi4);
        const double qi = gi4.w;
        const double m = P.mass;
        for_pairs_split<G>(
            P, L, S, i, g, p, xs,
            [&](u32 j, float) {
                const Geo ge = pair_geo(p, xs[j], P.sp);
                const double4 gj4 = gbar[j];
                const D3 gr = ge.grad();
                axpy(xa, m * (ri + rhobar[j]), gr);
                const D3 hg = ge.hess(sub(gi, d3(gj4)));
                axpy(xa, m, hg);
                axpy(xa, 2.0 * m * m * (qi + gj4.w), ge.hess(gr));
            },
            [&](u32 b, float) {
                const float4 q = P.xb[b];
                const Geo ge = pair_geo(p, q, P.sp);
                const double mb = q.w;
                const D3 gr = ge.grad();
                axpy(xa, mb * ri, gr);
                axpy(xa, mb, ge.hess(gi));
                axpy(xa, 2.0 * mb * mb * qi, ge.hess(gr));
            });
    }
    group_sum3<G>(xa);
    if (!valid || g != 0) return;
    const SceneSeg& S = *Sp;
    const D3 x0 = d3(xbar[i]);
    const u32 id = perm[S.start + (i - S.off)];
    const u64 o = 3ull * id;
    gx[o] = x0.x + xa.x;
    gx[o + 1] = x0.y + xa.y;
    gx[o + 2] = x0.z + xa.z;
    const double4 vb = v0bar[i];
    const double dec = S.decay;
    gv[o] = dec * vb.x;
    gv[o + 1] = dec * vb.y;
    gv[o + 2] = dec * vb.z;
}

}  // namespace

void DFSPHSolver::set_viscosity(float eps) {
    if (batched_) fail(kInvalidArgument, "set_viscosity: plain solver only");
    cfg_.viscosity = eps;
    viscosity_.assign(1, eps);
    if (!counts_.empty()) layout(counts_);
    drop_graphs();
}

double DFSPHSolver::adjoint_step(const SubstepTape& t, const double* gx_out, const double* gv_out,
                                 double* gx_in, double* gv_in, double* visc_acc) {
    if (batched_) fail(kInvalidArgument, "adjoint_step: plain solver only");
    const size_t nsz = fluid_.count();
    if (nsz == 0 || t.n != nsz) fail(kInvalidArgument, "adjoint_step: tape and store sizes differ");
    if (counts_.empty() || counts_[0] != nsz) layout({static_cast<u32>(nsz)});
    if (t.nslots != nslots_) fail(kInvalidArgument, "adjoint_step: tape layout differs");
    if (!gx_out || !gv_out || !gx_in || !gv_in) fail(kInvalidArgument, "adjoint_step: NULL adjoint");
    // recompute the substep's grid, pair lists, rho, alpha and v0/v1 from its input state
    WSIM_HIP(hipMemcpyAsync(fluid_.positions(), t.x.data(), 3 * nsz * sizeof(float),
                            hipMemcpyDeviceToDevice, stream_));
    WSIM_HIP(hipMemcpyAsync(fluid_.velocities(), t.v.data(), 3 * nsz * sizeof(float),
                            hipMemcpyDeviceToDevice, stream_));
    set_substep(&t.dt);
    phase_prefix(true);

    const size_t ns = nslots_;
    adj_.xbar.ensure(ns);
    adj_.wbar.ensure(ns);
    adj_.v0bar.ensure(ns);
    adj_.gbar.ensure(ns);
    adj_.rdbar.ensure(ns);
    adj_.rhobar.ensure(ns);
    adj_.alphabar.ensure(ns);
    const size_t nparts = ns * kAdjointLanes / NA;  // viscosity partials (one per workgroup)
    adj_.part.ensure(nparts);
    if (adj_.hpart.size() < nparts) adj_.hpart = PinnedBuffer<double>(nparts);

    const Nbr P = make_nbr(fluid_grid_, boundary_grid_.get(), xb_.data(), bnear_.data(), nb_, cfg_, mass_, nslots_,
                           dseg_.data(), nullptr, shapes_.remap);
    const PairList L = pair_list();
    const u32 blocks = static_cast<u32>(ns / NA);
    const u32 K = t.iters;
    const float eps = 1e-3f;  // dfsph_kernels.cu:377
    k_adj_advect<<<blocks, NA, 0, stream_>>>(
        P, fluid_grid_.sorted_indices(), xs_.data(), t.w.data() + K * ns, 5.0f,
        cfg_.domain_min[0] + eps, cfg_.domain_min[1] + eps, cfg_.domain_min[2] + eps,
        cfg_.domain_max[0] - eps, cfg_.domain_max[1] - eps, cfg_.domain_max[2] - eps, gx_out, gv_out,
        adj_.xbar.data(), adj_.wbar.data(), adj_.rhobar.data(), adj_.alphabar.data());
    WSIM_LAUNCH_CHECK();
    const int pform = static_cast<int>(cfg_.pressure_form);
    constexpr int GL = kAdjointLanes;
    const u32 gblocks = static_cast<u32>(ns * GL / NA);
    for (u32 k = K; k-- > 0;) {
        k_adj_kappa<GL><<<gblocks, NA, 0, stream_>>>(P, L, xs_.data(), adj_.wbar.data(),
                                                      t.xk.data() + k * ns, t.p.data() + k * ns,
                                                      rho_.data(), alpha_.data(), cfg_.rest_density,
                                                      pform, adj_.rdbar.data(), adj_.rhobar.data(),
                                                      adj_.alphabar.data(), adj_.xbar.data());
        WSIM_LAUNCH_CHECK();
        k_adj_rhodot<GL><<<gblocks, NA, 0, stream_>>>(P, L, xs_.data(), t.w.data() + k * ns,
                                                       adj_.rdbar.data(), adj_.wbar.data(),
                                                       adj_.xbar.data());
        WSIM_LAUNCH_CHECK();
    }
    const float4* v0 = any_xsph_ ? vpre_.data() : vs_.data();
    k_adj_xsph<GL><<<gblocks, NA, 0, stream_>>>(P, L, xs_.data(), v0, adj_.wbar.data(),
                                                 rho_.data(), viscosity_[0], adj_.v0bar.data(),
                                                 adj_.rhobar.data(), adj_.xbar.data(),
                                                 adj_.part.data());
    WSIM_LAUNCH_CHECK();
    k_adj_alpha<<<blocks, NA, 0, stream_>>>(P, L, xs_.data(), rho_.data(), adj_.alphabar.data(),
                                             adj_.rhobar.data(), adj_.gbar.data());
    WSIM_LAUNCH_CHECK();
    k_adj_geom<GL><<<gblocks, NA, 0, stream_>>>(P, L, fluid_grid_.sorted_indices(), xs_.data(),
                                                 adj_.rhobar.data(), adj_.gbar.data(),
                                                 adj_.xbar.data(), adj_.v0bar.data(), gx_in, gv_in);
    WSIM_LAUNCH_CHECK();
    if (visc_acc) {
        // the caller sums the partials once, after many substeps
        k_add_parts<<<div_up(gblocks, NS), NS, 0, stream_>>>(adj_.part.data(), gblocks, visc_acc);
        WSIM_LAUNCH_CHECK();
        return 0.0;
    }
    WSIM_HIP(hipMemcpyAsync(adj_.hpart.data(), adj_.part.data(), gblocks * sizeof(double),
                            hipMemcpyDeviceToHost, stream_));
    WSIM_HIP(hipStreamSynchronize(stream_));
    double g = 0.0;
    for (u32 k = 0; k < gblocks; ++k) g += adj_.hpart[k];
    return g;
}

size_t DFSPHSolver::adjoint_partials() const { return nslots_ * kAdjointLanes / NA; }

// ---------------------------------------------------------------------------
Rollout::Rollout(DFSPHSolver& solver, ParticleStore& fluid, u32 every)
    : solver_(solver), fluid_(fluid), every_(every) {
    if (every == 0) fail(kInvalidArgument, "Rollout: checkpoint interval must be positive");
    if (solver.batched()) fail(kInvalidArgument, "Rollout: plain solver only");
    if (solver.config().enable_divergence_solve || solver.config().enable_surface_tension)
        fail(kInvalidArgument,
             "Rollout: the adjoint covers the density solve without surface tension");
    if (solver.config().fast_math)
        fail(kInvalidArgument,
             "Rollout: the adjoint differentiates the bit-exact pair arithmetic (fast_math 0)");
}

void Rollout::forward(const float* x0, const float* v0, const float* dts, u32 substeps) {
    n_ = fluid_.count();
    if (n_ == 0) fail(kInvalidArgument, "Rollout: empty store");
    if (!x0 || !v0 || (!dts && substeps)) fail(kInvalidArgument, "Rollout: NULL input");
    hipStream_t s = solver_.stream();
    const size_t bytes = 3 * n_ * sizeof(float);
    WSIM_HIP(hipMemcpyAsync(fluid_.positions(), x0, bytes, hipMemcpyDeviceToDevice, s));
    WSIM_HIP(hipMemcpyAsync(fluid_.velocities(), v0, bytes, hipMemcpyDeviceToDevice, s));
    dts_.assign(dts, dts + substeps);
    const u32 nck = (substeps + every_ - 1) / every_;
    if (ckx_.size() < nck) {
        ckx_.resize(nck);
        ckv_.resize(nck);
    }
    for (u32 k = 0; k < substeps; ++k) {
        if (k % every_ == 0) {
            ckx_[k / every_].ensure(3 * n_);
            ckv_[k / every_].ensure(3 * n_);
            WSIM_HIP(hipMemcpyAsync(ckx_[k / every_].data(), fluid_.positions(), bytes,
                                    hipMemcpyDeviceToDevice, s));
            WSIM_HIP(hipMemcpyAsync(ckv_[k / every_].data(), fluid_.velocities(), bytes,
                                    hipMemcpyDeviceToDevice, s));
        }
        solver_.step(dts_[k]);
    }
    WSIM_HIP(hipStreamSynchronize(s));
}

double Rollout::backward(const double* gx, const double* gv, double* gx0, double* gv0) {
    const u32 T = substeps();
    if (n_ == 0 || fluid_.count() != n_) fail(kInvalidArgument, "Rollout: run forward first");
    if (!gx || !gv || !gx0 || !gv0) fail(kInvalidArgument, "Rollout: NULL adjoint");
    hipStream_t s = solver_.stream();
    const size_t db = 3 * n_ * sizeof(double), fb = 3 * n_ * sizeof(float);
    for (auto& b : g_) b.ensure(3 * n_);
    WSIM_HIP(hipMemcpyAsync(g_[0].data(), gx, db, hipMemcpyDeviceToDevice, s));
    WSIM_HIP(hipMemcpyAsync(g_[1].data(), gv, db, hipMemcpyDeviceToDevice, s));
    int cur = 0;
    double gvisc = 0.0;
    if (tapes_.size() < std::min<u32>(every_, T)) tapes_.resize(std::min<u32>(every_, T));
    const size_t nparts = solver_.adjoint_partials();
    visc_acc_.ensure(nparts);
    WSIM_HIP(hipMemsetAsync(visc_acc_.data(), 0, nparts * sizeof(double), s));
    const u32 nck = (T + every_ - 1) / every_;
    for (u32 seg = nck; seg-- > 0;) {
        const u32 start = seg * every_, end = std::min(T, start + every_);
        WSIM_HIP(hipMemcpyAsync(fluid_.positions(), ckx_[seg].data(), fb, hipMemcpyDeviceToDevice, s));
        WSIM_HIP(hipMemcpyAsync(fluid_.velocities(), ckv_[seg].data(), fb, hipMemcpyDeviceToDevice, s));
        for (u32 k = start; k < end; ++k) solver_.step_taped(dts_[k], tapes_[k - start]);
        for (u32 k = end; k-- > start;) {
            const int nxt = cur ^ 2;
            solver_.adjoint_step(tapes_[k - start], g_[cur].data(), g_[cur + 1].data(),
                                 g_[nxt].data(), g_[nxt + 1].data(), visc_acc_.data());
            cur = nxt;
        }
    }
    WSIM_HIP(hipMemcpyAsync(gx0, g_[cur].data(), db, hipMemcpyDeviceToDevice, s));
    WSIM_HIP(hipMemcpyAsync(gv0, g_[cur + 1].data(), db, hipMemcpyDeviceToDevice, s));
    std::vector<double> acc(nparts);
    WSIM_HIP(hipMemcpyAsync(acc.data(), visc_acc_.data(), nparts * sizeof(double),
                            hipMemcpyDeviceToHost, s));
    WSIM_HIP(hipStreamSynchronize(s));
    for (double a : acc) gvisc += a;
    return gvisc;
}

}  // namespace wsim
