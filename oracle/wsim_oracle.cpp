// wsim_oracle.cpp -- CPU restatement of the reference DFSPH hot path.
//
// TEST INFRASTRUCTURE ONLY: loaded by tests/, __graft_entry__.smoke() and the
// cpu_baseline leg of bench.py.  The product library never links it.
//
// The code follows the reference's data flow literally: particle arrays stay
// in original order, neighbour loops gather through the cell grid's sorted
// index array, the 27-cell stencil is walked dz, dy, dx with ascending original
// id inside a cell, and every arithmetic expression keeps the reference's
// operand order.  Compiled with -ffp-contract=off, the binary32 instance is the
// bit-level parity oracle for the HIP kernels.  The binary64 instance runs the
// same algorithm with the same (binary32-rounded) constants.
//
// Deviations from the reference, all deliberate and documented:
//  * the reference compiles device code with -use_fast_math
//    (CMakeLists.txt:56); here every divide, sqrt and exp is IEEE, and
//    exp(-damping*dt) is evaluated once on the host;
//  * XSPH is a Jacobi sweep (all neighbours read pre-XSPH velocities); the
//    reference updates velocities in place inside the same launch
//    (dfsph_kernels.cu:317,338-340), which is schedule dependent;
//  * a NaN or far out-of-range coordinate maps to the sentinel cell (the
//    reference's float->int cast is undefined there, cell_grid.cu:11-13);
//  * convergence errors are accumulated in binary64 (reference: CUB float sum,
//    dfsph.cu:288-300).
#include "wsim_oracle.h"

// Mutation hooks for tests/test_oracle_mutations.py: a build with
// -DORC_MUTANT=k plants mistake k (a plausible slip in one formula) so that
// the suite can show that the pinning tests reject it.  0 (the default) is
// the restatement itself.
#ifndef ORC_MUTANT
#define ORC_MUTANT 0
#endif

#include <algorithm>
#include <cmath>
#include <cstring>
#include <limits>
#include <vector>
#include <type_traits>
#include <omp.h>

namespace {

using u32 = uint32_t;

template <class R> struct V3 { R x, y, z; };

// Forward-mode tangent number for the gradient checks.  The value part is
// computed in binary32 by exactly the operations, in exactly the order, of
// the float instance, so a DualF solver follows the binary32 trajectory bit
// for bit; the tangent part carries d(value)/d(seed) in binary64 by the chain
// rule at those binary32 values.  Branches (stencil membership, clamps,
// max(0, .), the iteration count) follow the value, which is the frozen
// structure the adjoint differentiates.
struct DualF {
    float v = 0.0f;
    double d = 0.0;
    DualF() = default;
    DualF(float a) : v(a) {}
    DualF(double a) : v(static_cast<float>(a)) {}
    DualF(int a) : v(static_cast<float>(a)) {}
    DualF(float a, double t) : v(a), d(t) {}
    explicit operator int() const { return static_cast<int>(v); }
    explicit operator float() const { return v; }
    explicit operator double() const { return v; }
};
inline DualF operator+(DualF a, DualF b) { return {a.v + b.v, a.d + b.d}; }
inline DualF operator-(DualF a, DualF b) { return {a.v - b.v, a.d - b.d}; }
inline DualF operator-(DualF a) { return {-a.v, -a.d}; }
inline DualF operator*(DualF a, DualF b) {
    return {a.v * b.v, a.d * double(b.v) + double(a.v) * b.d};
}
inline DualF operator/(DualF a, DualF b) {
    const double bv = b.v;
    return {a.v / b.v, (a.d * bv - double(a.v) * b.d) / (bv * bv)};
}
inline DualF& operator+=(DualF& a, DualF b) { return a = a + b; }
inline DualF& operator-=(DualF& a, DualF b) { return a = a - b; }
inline DualF& operator*=(DualF& a, DualF b) { return a = a * b; }
inline bool operator<(DualF a, DualF b) { return a.v < b.v; }
inline bool operator<=(DualF a, DualF b) { return a.v <= b.v; }
inline bool operator>(DualF a, DualF b) { return a.v > b.v; }
inline bool operator>=(DualF a, DualF b) { return a.v >= b.v; }
inline bool operator==(DualF a, DualF b) { return a.v == b.v; }
inline bool operator!=(DualF a, DualF b) { return a.v != b.v; }

// Math used by the solver templates, for float, double and DualF.
inline float o_sqrt(float a) { return std::sqrt(a); }
inline double o_sqrt(double a) { return std::sqrt(a); }
inline DualF o_sqrt(DualF a) {
    return {std::sqrt(a.v), a.v > 0.0f ? a.d / (2.0 * std::sqrt(double(a.v))) : 0.0};
}
inline float o_fmax(float a, float b) { return std::fmax(a, b); }
inline double o_fmax(double a, double b) { return std::fmax(a, b); }
inline DualF o_fmax(DualF a, DualF b) {
    // the operand std::fmax returns (NaN loses; ties keep the first)
    return (a.v >= b.v || std::isnan(b.v)) ? a : b;
}
inline float o_floor(float a) { return std::floor(a); }
inline double o_floor(double a) { return std::floor(a); }
inline DualF o_floor(DualF a) { return {std::floor(a.v), 0.0}; }
inline bool o_isfinite(float a) { return std::isfinite(a); }
inline bool o_isfinite(double a) { return std::isfinite(a); }
inline bool o_isfinite(DualF a) { return std::isfinite(a.v); }

// Instances whose values are binary32 (float itself and DualF).
template <class R> struct F32Values { static constexpr bool value = std::is_same<R, float>::value; };
template <> struct F32Values<DualF> { static constexpr bool value = true; };

// Frozen structure for the finite-difference checks: every grid build's
// per-particle home cells and every solve's iteration count, in call order.
// Recorded by one solver and replayed by others (any precision), so the
// perturbed runs keep the recording run's neighbour sets and iteration
// counts: the map the adjoint differentiates.
struct StructTape {
    std::vector<std::vector<int16_t>> homes;
    std::vector<u32> iters;
    // per particle: max(0, .) branch of every sweep; clamp decisions of every advection
    std::vector<std::vector<uint8_t>> masks;
    std::vector<std::vector<uint16_t>> clamps;
};

// Clamp decisions of one particle's advection (dfsph_kernels.cu:363-386).
enum : uint16_t {
    kClampSpeed = 1,        // |v| > vmax: rescaled
    kClampLo = 2,           // << axis: position clamped at the low bound
    kClampHi = 16,          // << axis: position clamped at the high bound
    kZeroLo = 128,          // << axis: velocity component zeroed at the low bound
    kZeroHi = 1024,         // << axis: at the high bound
};

// Test-only switch: XSPH as an in-place sequential sweep in original index
// order (one of the schedules the reference's racy in-place kernel can
// produce, dfsph_kernels.cu:317,338-340) instead of the Jacobi sweep.
int g_xsph_in_place = 0;

// Analysis-only switches for the study of the adjoint's sensitivity growth
// (tests/test_oracle_adjoint_growth.py); both 0 in every other use.
//   g_freeze_geometry: for R = DualF, the kernel values and gradients carry
//     no tangent, so positions reach the next state only through advection
//     (the pressure, XSPH and boundary terms act on the velocity tangents
//     with the geometry held at its values);
//   g_free_advect: bit 0 drops the |v| clamp, bit 1 the domain clamp.
int g_freeze_geometry = 0;
int g_free_advect = 0;
template <class R> inline R geom(R a) { return a; }
template <> inline DualF geom(DualF a) { return g_freeze_geometry ? DualF(a.v) : a; }

// Constants are the reference's binary32 literals, widened for R = double.
template <class R> inline R kpi() { return R(3.14159265358979323846f); }
template <class R> inline R k1em9() { return R(1e-9f); }
template <class R> inline R k1em12() { return R(1e-12f); }

// ---------------------------------------------------------------------------
// sph_kernels.cuh:12-22  cubic spline W(r, l), support [0, 2l)
template <class R> inline R W(R r_in, R l) {
    const R r = geom(r_in);
    const R q = r / l;
    if (q >= R(2)) return R(0);
    const R sigma = R(1) / (kpi<R>() * l * l * l);
    if (q < R(1)) return sigma * (R(1) - R(1.5) * q * q + R(0.75) * q * q * q);
    const R t = R(2) - q;
    if (ORC_MUTANT == 1) return sigma * R(0.25) * t * t;  // mutant: (2 - q)^2
    return sigma * R(0.25) * t * t * t;
}

// sph_kernels.cuh:25-41  grad_i W(x_i - x_j, l)
template <class R> inline V3<R> gradW(const V3<R>& rij_in, R l) {
    const V3<R> rij{geom(rij_in.x), geom(rij_in.y), geom(rij_in.z)};
    const R r2 = rij.x * rij.x + rij.y * rij.y + rij.z * rij.z;
    const R r = o_sqrt(r2);
    if (r < k1em12<R>()) return {R(0), R(0), R(0)};
    const R q = r / l;
    if (q >= R(2)) return {R(0), R(0), R(0)};
    const R sigma = R(1) / (kpi<R>() * l * l * l);
    R dWdq;
    if (q < R(1)) {
        dWdq = sigma * (R(-3) * q + R(2.25) * q * q);
    } else {
        const R t = R(2) - q;
        dWdq = sigma * R(-0.75) * t * t;
    }
    if (ORC_MUTANT == 2 && q < R(1)) dWdq = sigma * (R(-3) * q + R(2.5) * q * q);  // mutant: 2.5 q^2
    const R c = dWdq / (l * r);
    return {rij.x * c, rij.y * c, rij.z * c};
}

// ---------------------------------------------------------------------------
// Grid: cell_grid.cu:10-22 (key), :93-117 (stable sort by key, ties by id),
// :119-137 (cell_start = lower_bound of each cell in the sorted keys).

// (int)floorf((p - o) * inv) (cell_grid.cu:11-13, dfsph_kernels.cu:15-17),
// made total: NaN and coordinates beyond the stencil's reach are clamped to
// values whose whole stencil lies outside the grid.
//
// Cells are integer structure: every instance computes them in binary32 from
// the binary32-rounded position, as the reference does, so the float, double
// and dual solvers see the same grids for the same inputs (wall particles sit
// exactly on cell faces, where a binary64 evaluation can pick the neighbour).
inline int cell_coord_f(float p, float o, float inv, int D) {
    const float f = std::floor((p - o) * inv);
    if (!(f >= -2.0f)) return -2;
    if (f > float(D + 1)) return D + 1;
    return static_cast<int>(f);
}
template <class R> inline int cell_coord(R p, float o, float inv, int D) {
    return cell_coord_f(static_cast<float>(p), o, inv, D);
}

inline u32 cell_key(int cx, int cy, int cz, const int32_t D[3]) {
    if (cx < 0 || cx >= D[0] || cy < 0 || cy >= D[1] || cz < 0 || cz >= D[2])
        return u32(D[0]) * u32(D[1]) * u32(D[2]);
    return (u32(cx) * u32(D[1]) + u32(cy)) * u32(D[2]) + u32(cz);
}

template <class R> struct Grid {
    float ox = 0, oy = 0, oz = 0, inv = 0;
    int32_t D[3] = {0, 0, 0};
    u32 C = 0;
    std::vector<u32> keys, sorted_keys, perm, cs;

    void init(const float origin[3], const int32_t dims[3], float cell) {
        ox = origin[0]; oy = origin[1]; oz = origin[2];
        inv = 1.0f / cell;  // cell_grid.cu:81: inverse cell length computed once on the host
        std::memcpy(D, dims, sizeof(D));
        C = u32(D[0]) * u32(D[1]) * u32(D[2]);
        cs.assign(size_t(C) + 1, 0);
    }

    u32 key_of(const V3<R>& p) const {
        return cell_key(cell_coord(p.x, ox, inv, D[0]), cell_coord(p.y, oy, inv, D[1]),
                        cell_coord(p.z, oz, inv, D[2]), D);
    }

    // home: optional taped home cells (3 per particle) used instead of the
    // positions' cells (frozen-structure replays).
    void build(const std::vector<V3<R>>& x, size_t n, const int16_t* home = nullptr) {
        keys.resize(n);
        sorted_keys.resize(n);
        perm.resize(n);
#pragma omp parallel for schedule(static)
        for (long long i = 0; i < (long long)n; ++i)
            keys[i] = home ? cell_key(home[3 * i], home[3 * i + 1], home[3 * i + 2], D)
                           : key_of(x[i]);
        // Stable LSD radix sort of (key, id) pairs, 8-bit digits, as the
        // reference's SortPairs over (key, iota) (cell_grid.cu:98-110).  Each
        // pass is parallel and stable: thread t counts and then scatters its
        // own contiguous chunk, after the chunks of lower threads.
        int bits = 0;
        while (bits < 32 && (C >> bits) != 0) ++bits;
        std::vector<u32> k2(n), p2(n);
        sorted_keys = keys;
#pragma omp parallel for schedule(static)
        for (long long i = 0; i < (long long)n; ++i) perm[i] = u32(i);
        const int T = omp_get_max_threads();
        std::vector<size_t> off(size_t(T) * 256);
        for (int shift = 0; shift < bits; shift += 8) {
            std::fill(off.begin(), off.end(), size_t(0));
#pragma omp parallel num_threads(T)
            {
                const int t = omp_get_thread_num(), nt = omp_get_num_threads();
                const size_t lo = n * size_t(t) / size_t(nt), hi = n * size_t(t + 1) / size_t(nt);
                size_t* h = off.data() + size_t(t) * 256;
                for (size_t i = lo; i < hi; ++i) ++h[(sorted_keys[i] >> shift) & 255u];
#pragma omp barrier
#pragma omp single
                {
                    size_t run = 0;
                    for (int d = 0; d < 256; ++d)
                        for (int u = 0; u < nt; ++u) {
                            const size_t c = off[size_t(u) * 256 + d];
                            off[size_t(u) * 256 + d] = run;
                            run += c;
                        }
                }
                for (size_t i = lo; i < hi; ++i) {
                    const size_t o = h[(sorted_keys[i] >> shift) & 255u]++;
                    k2[o] = sorted_keys[i];
                    p2[o] = perm[i];
                }
            }
            sorted_keys.swap(k2);
            perm.swap(p2);
        }
        // cell_start[c] = number of particles with key < c, for c in [0, C]
        // (fill + scatter + right-to-left min, cell_grid.cu:119-137): slot i
        // that starts a run of key k fills (previous key, k]; the cells after
        // the last key take n.  Every entry is written once.
#pragma omp parallel for schedule(dynamic, 4096)
        for (long long i = 0; i < (long long)n; ++i) {
            const u32 k = sorted_keys[i];
            if (i > 0 && sorted_keys[i - 1] == k) continue;
            const size_t c0 = i > 0 ? size_t(sorted_keys[i - 1]) + 1 : 0;
            for (size_t c = c0; c <= k; ++c) cs[c] = u32(i);
        }
        const size_t tail0 = n ? size_t(sorted_keys[n - 1]) + 1 : 0;
#pragma omp parallel for schedule(static)
        for (long long c = (long long)tail0; c <= (long long)C; ++c) cs[c] = u32(n);
    }

    void home_of(const V3<R>& p, int16_t* out) const {
        out[0] = int16_t(cell_coord(p.x, ox, inv, D[0]));
        out[1] = int16_t(cell_coord(p.y, oy, inv, D[1]));
        out[2] = int16_t(cell_coord(p.z, oz, inv, D[2]));
    }

    // ITER_NEIGHBORS (dfsph_kernels.cu:13-35): dz, dy, dx in {-1,0,1}.  The
    // query's home cell comes from its position or, in a replay, from `hc`.
    template <class F> void for_neighbors(const V3<R>& p, F&& f, const int16_t* hc = nullptr) const {
        const int cx = hc ? hc[0] : cell_coord(p.x, ox, inv, D[0]);
        const int cy = hc ? hc[1] : cell_coord(p.y, oy, inv, D[1]);
        const int cz = hc ? hc[2] : cell_coord(p.z, oz, inv, D[2]);
        for (int dz = -1; dz <= 1; ++dz)
            for (int dy = -1; dy <= 1; ++dy)
                for (int dx = -1; dx <= (ORC_MUTANT == 3 ? 0 : 1); ++dx) {  // mutant 3: no +x column
                    const int nx = cx + dx, ny = cy + dy, nz = cz + dz;
                    if (nx < 0 || nx >= D[0] || ny < 0 || ny >= D[1] || nz < 0 || nz >= D[2])
                        continue;
                    const u32 c = (u32(nx) * u32(D[1]) + u32(ny)) * u32(D[2]) + u32(nz);
                    if (c >= C) continue;
                    for (u32 s = cs[c]; s < cs[c + 1]; ++s) f(perm[s]);
                }
    }
};

// dfsph.cu:55-65
void make_grid_params(const float dmin[3], const float dmax[3], float h, float origin[3],
                      int32_t dims[3]) {
    for (int a = 0; a < 3; ++a) {
        dims[a] = static_cast<int32_t>(std::ceil((dmax[a] - dmin[a]) / h)) + 2;
        origin[a] = dmin[a] - h;
    }
}

template <class R> inline V3<R> load3(const float* p, size_t i) {
    return {R(p[3 * i]), R(p[3 * i + 1]), R(p[3 * i + 2])};
}

// ---------------------------------------------------------------------------
// DFSPHSolver restatement (dfsph.cu:71-133, dfsph_kernels.cu:40-400)
template <class R> struct Solver {
    orc_config cfg;
    size_t n = 0, nb = 0;
    std::vector<V3<R>> x, v, vtmp, xb;
    std::vector<R> rho, alpha, kappa, rho_adv, kappa_v, mb;
    Grid<R> fg, bg;
    R mass = 0, h = 0;
    R eps_param = 0;  // XSPH viscosity (a tangent seed for DualF)
    orc_stats st{};
    // frozen-structure tape: 0 off, 1 record, 2 replay
    StructTape* tape = nullptr;
    int tape_mode = 0;
    size_t tape_home = 0, tape_iter = 0, tape_mask = 0, tape_clamp = 0;
    std::vector<uint8_t> mask_buf;
    std::vector<uint16_t> clamp_buf;

    // max(0, a) with its branch recorded or replayed (frozen-structure tapes).
    const uint8_t* mask_in = nullptr;
    uint8_t* mask_out = nullptr;
    void open_masks() {
        mask_in = nullptr;
        mask_out = nullptr;
        if (tape_mode == 1) {
            mask_buf.assign(n, 0);
            mask_out = mask_buf.data();
        } else if (tape_mode == 2 && tape_mask < tape->masks.size()) {
            mask_in = tape->masks[tape_mask++].data();
        }
    }
    void close_masks() {
        if (tape_mode == 1) tape->masks.push_back(mask_buf);
    }
    R relu(long long i, R a) const {
        if (mask_in) return mask_in[i] ? a : R(0);
        const R r = o_fmax(R(0), a);
        if (mask_out) mask_out[i] = (r != R(0)) ? 1 : 0;
        return r;
    }
    const int16_t* home = nullptr;  // home cells of the current grid build
    std::vector<int16_t> home_buf;

    const int16_t* hc(long long i) const { return home ? home + 3 * i : nullptr; }

    Solver(const orc_config& c, const float* px, const float* pv, size_t n_, const float* pxb,
           size_t nb_)
        : cfg(c), n(n_), nb(nb_) {
        x.resize(n);
        v.resize(n);
        vtmp.resize(n);
        for (size_t i = 0; i < n; ++i) {
            x[i] = load3<R>(px, i);
            v[i] = pv ? load3<R>(pv, i) : V3<R>{0, 0, 0};
        }
        rho.assign(n, 0); alpha.assign(n, 0); kappa.assign(n, 0);
        rho_adv.assign(n, 0); kappa_v.assign(n, 0);
        h = R(cfg.smoothing_length);
        eps_param = R(cfg.viscosity);
        // dfsph.cu:76-78
        if constexpr (F32Values<R>::value) {
            const float s = 2.0f * cfg.particle_radius;
            const float v0 = s * s * s;
            mass = R(cfg.rest_density * v0);
        } else {
            const R s = R(2) * R(cfg.particle_radius);
            const R v0 = s * s * s;
            mass = R(cfg.rest_density) * v0;
        }
        float origin[3];
        int32_t dims[3];
        make_grid_params(cfg.domain_min, cfg.domain_max, cfg.smoothing_length, origin, dims);
        fg.init(origin, dims, cfg.smoothing_length);
        bg.init(origin, dims, cfg.smoothing_length);
        if (nb > 0 && pxb) {
            xb.resize(nb);
            for (size_t i = 0; i < nb; ++i) xb[i] = load3<R>(pxb, i);
            mb.assign(nb, 0);
            bg.build(xb, nb);  // dfsph.cu:96-97, built once
            boundary_mass();   // dfsph.cu:98
        } else {
            nb = 0;
        }
    }

    // dfsph_kernels.cu:40-65
    void boundary_mass() {
        const R rho0 = R(cfg.rest_density);
#pragma omp parallel for schedule(static)
        for (long long i = 0; i < (long long)nb; ++i) {
            const V3<R> pi = xb[i];
            R sum_w = 0;
            bg.for_neighbors(pi, [&](u32 j) {
                const V3<R> pj = xb[j];
                const V3<R> r{pi.x - pj.x, pi.y - pj.y, pi.z - pj.z};
                const R d = o_sqrt(r.x * r.x + r.y * r.y + r.z * r.z);
                sum_w += W(d, h);
            });
            mb[i] = rho0 / o_fmax(sum_w, k1em9<R>());
        }
    }

    // dfsph_kernels.cu:68-112
    void density() {
#pragma omp parallel for schedule(static)
        for (long long i = 0; i < (long long)n; ++i) {
            const V3<R> pi = x[i];
            R r_sum = 0;
            fg.for_neighbors(pi, [&](u32 j) {
                const V3<R> pj = x[j];
                const V3<R> r{pi.x - pj.x, pi.y - pj.y, pi.z - pj.z};
                const R d = o_sqrt(r.x * r.x + r.y * r.y + r.z * r.z);
                r_sum += mass * W(d, h);
            }, hc(i));
            if (nb > 0) {
                bg.for_neighbors(pi, [&](u32 j) {
                    const V3<R> pj = xb[j];
                    const V3<R> r{pi.x - pj.x, pi.y - pj.y, pi.z - pj.z};
                    const R d = o_sqrt(r.x * r.x + r.y * r.y + r.z * r.z);
                    r_sum += mb[j] * W(d, h);
                }, hc(i));
            }
            rho[i] = r_sum;
        }
    }

    // dfsph_kernels.cu:117-169
    void alpha_factor() {
#pragma omp parallel for schedule(static)
        for (long long i = 0; i < (long long)n; ++i) {
            const V3<R> pi = x[i];
            V3<R> sg{0, 0, 0};
            R sq = 0;
            fg.for_neighbors(pi, [&](u32 j) {
                if (j == u32(i)) return;
                const V3<R> pj = x[j];
                const V3<R> g = gradW(V3<R>{pi.x - pj.x, pi.y - pj.y, pi.z - pj.z}, h);
                const V3<R> mg{mass * g.x, mass * g.y, mass * g.z};
                sg.x += mg.x; sg.y += mg.y; sg.z += mg.z;
                sq += mg.x * mg.x + mg.y * mg.y + mg.z * mg.z;
            }, hc(i));
            if (nb > 0) {
                bg.for_neighbors(pi, [&](u32 j) {
                    const V3<R> pj = xb[j];
                    const V3<R> g = gradW(V3<R>{pi.x - pj.x, pi.y - pj.y, pi.z - pj.z}, h);
                    const V3<R> mg{mb[j] * g.x, mb[j] * g.y, mb[j] * g.z};
                    sg.x += mg.x; sg.y += mg.y; sg.z += mg.z;
                    sq += mg.x * mg.x + mg.y * mg.y + mg.z * mg.z;
                }, hc(i));
            }
            const R denom = (ORC_MUTANT == 4 ? R(0) : sg.x * sg.x + sg.y * sg.y + sg.z * sg.z) + sq;
            alpha[i] = (denom > k1em9<R>()) ? (rho[i] / denom) : R(0);
        }
    }

    // dfsph_kernels.cu:347-355 with exp(-damping*dt) evaluated once per substep
    void gravity(float dt) {
        R decay;
        if constexpr (F32Values<R>::value) decay = R(orc_damping_decay(cfg.damping, dt));
        else decay = std::exp(-R(cfg.damping) * R(dt));
        const R gx = R(cfg.gravity[0]), gy = R(cfg.gravity[1]), gz = R(cfg.gravity[2]);
        const R dtr = R(dt);
#pragma omp parallel for schedule(static)
        for (long long i = 0; i < (long long)n; ++i) {
            v[i].x = v[i].x * decay + gx * dtr;
            v[i].y = v[i].y * decay + gy * dtr;
            v[i].z = v[i].z * decay + gz * dtr;
        }
    }

    // dfsph_kernels.cu:305-341, as a Jacobi sweep
    void xsph() {
        if (cfg.viscosity <= 0.0f) return;  // dfsph.cu:198
        const R eps = eps_param;
        if (g_xsph_in_place) vtmp = v;
#pragma omp parallel for schedule(static) if (!g_xsph_in_place)
        for (long long i = 0; i < (long long)n; ++i) {
            const V3<R> pi = x[i];
            const V3<R> vi = v[i];
            const R rho_i = rho[i];
            V3<R> dl{0, 0, 0};
            fg.for_neighbors(pi, [&](u32 j) {
                if (j == u32(i)) return;
                const V3<R> pj = x[j];
                const V3<R> vj = g_xsph_in_place ? vtmp[j] : v[j];
                const V3<R> r{pi.x - pj.x, pi.y - pj.y, pi.z - pj.z};
                const R d = o_sqrt(r.x * r.x + r.y * r.y + r.z * r.z);
                const R w = W(d, h);
                const R rho_avg = R(0.5) * (rho_i + rho[j]);
                const R c = eps * (mass / o_fmax(rho_avg, k1em9<R>())) * w;
                if (ORC_MUTANT == 8) { dl.x += c * vj.x; dl.y += c * vj.y; dl.z += c * vj.z; return; }
                dl.x += c * (vj.x - vi.x);
                dl.y += c * (vj.y - vi.y);
                dl.z += c * (vj.z - vi.z);
            }, hc(i));
            vtmp[i] = {vi.x + dl.x, vi.y + dl.y, vi.z + dl.z};
        }
        v.swap(vtmp);
    }

    // Akinci et al. 2013 surface tension (cohesion + curvature), with the
    // reference's cohesion kernel C(r, L) (sph_kernels.cuh:44-51) on the
    // kernel support L = 2l; the reference parses the coefficient
    // (loader.cpp:50, main.cpp:132) but applies no force.  After XSPH:
    //   n_i = L sum_{j != i} (m / rho_j) grad W_ij
    //   a_i = sum_{j != i, |r| < L} K_ij (s_ij r_ij - gamma (n_i - n_j)),
    //   K_ij = 2 rho0 / (rho_i + rho_j),  s_ij = -gamma m C(|r|) / |r| (0 at r = 0)
    //   v_i += dt a_i
    // powf(L, 9) and powf(L, 6) are evaluated once on the host.
    void surface_tension(float dt) {
        if (!cfg.enable_surface_tension) return;
        const float Lf = 2.0f * cfg.smoothing_length;
        R Lr, sig, c6;
        if constexpr (F32Values<R>::value) {
            Lr = R(Lf);
            sig = R(32.0f / (3.14159265358979323846f * std::pow(Lf, 9.0f)));
            c6 = R(std::pow(Lf, 6.0f) / 64.0f);
        } else {
            Lr = R(2) * h;
            sig = R(32) / (kpi<R>() * std::pow(Lr, R(9)));
            c6 = std::pow(Lr, R(6)) / R(64);
        }
        const R gam = R(cfg.surface_tension), rho0 = R(cfg.rest_density), dtr = R(dt);
        std::vector<V3<R>> nrm(n);
#pragma omp parallel for schedule(static)
        for (long long i = 0; i < (long long)n; ++i) {
            const V3<R> pi = x[i];
            V3<R> sn{0, 0, 0};
            fg.for_neighbors(pi, [&](u32 j) {
                if (j == u32(i)) return;
                const V3<R> pj = x[j];
                const V3<R> g = gradW(V3<R>{pi.x - pj.x, pi.y - pj.y, pi.z - pj.z}, h);
                const R f = mass / rho[j];
                sn.x += f * g.x;
                sn.y += f * g.y;
                sn.z += f * g.z;
            }, hc(i));
            nrm[i] = {Lr * sn.x, Lr * sn.y, Lr * sn.z};
        }
        auto cohesion = [&](R r) -> R {
            if (r <= R(0) || r > Lr) return R(0);
            const R lm = Lr - r;
            const R t = lm * lm * lm * r * r * r;
            if (R(2) * r > Lr) return sig * t;
            return sig * (R(2) * t - c6);
        };
#pragma omp parallel for schedule(static)
        for (long long i = 0; i < (long long)n; ++i) {
            const V3<R> pi = x[i];
            const V3<R> ni = nrm[i];
            const R rho_i = rho[i];
            V3<R> a{0, 0, 0};
            fg.for_neighbors(pi, [&](u32 j) {
                if (j == u32(i)) return;
                const V3<R> pj = x[j];
                const V3<R> r{pi.x - pj.x, pi.y - pj.y, pi.z - pj.z};
                const R d = o_sqrt(r.x * r.x + r.y * r.y + r.z * r.z);
                if (!(d / h < R(2))) return;
                const R K = R(2) * rho0 / (rho_i + rho[j]);
                const R sc = d > R(0) ? -gam * mass * cohesion(d) / d : R(0);
                const V3<R> nj = nrm[j];
                a.x += K * (sc * r.x + -gam * (ni.x - nj.x));
                a.y += K * (sc * r.y + -gam * (ni.y - nj.y));
                a.z += K * (sc * r.z + -gam * (ni.z - nj.z));
            }, hc(i));
            vtmp[i] = {v[i].x + dtr * a.x, v[i].y + dtr * a.y, v[i].z + dtr * a.z};
        }
        v.swap(vtmp);
    }

    // rho_dot_i (dfsph_kernels.cu:196-224)
    R rho_dot(long long i) const {
        const V3<R> pi = x[i];
        const V3<R> vi = v[i];
        R rd = 0;
        fg.for_neighbors(pi, [&](u32 j) {
            if (j == u32(i)) return;
            const V3<R> pj = x[j];
            const V3<R> vj = v[j];
            const V3<R> g = gradW(V3<R>{pi.x - pj.x, pi.y - pj.y, pi.z - pj.z}, h);
            const V3<R> dv = ORC_MUTANT == 5 ? V3<R>{vj.x - vi.x, vj.y - vi.y, vj.z - vi.z}
                                             : V3<R>{vi.x - vj.x, vi.y - vj.y, vi.z - vj.z};
            rd += mass * (dv.x * g.x + dv.y * g.y + dv.z * g.z);
        }, hc(i));
        if (nb > 0) {
            bg.for_neighbors(pi, [&](u32 j) {
                const V3<R> pj = xb[j];
                const V3<R> g = gradW(V3<R>{pi.x - pj.x, pi.y - pj.y, pi.z - pj.z}, h);
                rd += mb[j] * (vi.x * g.x + vi.y * g.y + vi.z * g.z);
            }, hc(i));
        }
        return rd;
    }

    // dfsph_kernels.cu:176-232
    void density_change(float dt) {
        const R dtr = R(dt), rho0 = R(cfg.rest_density);
        open_masks();
#pragma omp parallel for schedule(static)
        for (long long i = 0; i < (long long)n; ++i) {
            const R rd = rho_dot(i);
            const R rho_pred = rho[i] + dtr * rd;
            rho_adv[i] = rho_pred;
            const R over = relu(i, rho_pred - rho0);
            kappa[i] = over * alpha[i] / (dtr * dtr);
        }
        close_masks();
    }

    // divergence_change_kernel (docs/superpowers/plans/2026-04-24-water-simulation-v2-phase2-dfsph-solver.md:1862-1907); boundary
    // term as in the shipped density_change_kernel (dfsph_kernels.cu:213-224).
    void divergence_change(float dt) {
        const R dtr = R(dt);
        open_masks();
#pragma omp parallel for schedule(static)
        for (long long i = 0; i < (long long)n; ++i) {
            const R rd = rho_dot(i);
            rho_adv[i] = rd;
            const R s = relu(i, rd);
            kappa_v[i] = s * alpha[i] / dtr;
        }
        close_masks();
    }

    // dfsph_kernels.cu:237-301 (pressure_form 1 replaces kappa/rho^2 by kappa/rho)
    void apply_kappa(const std::vector<R>& kap, float dt) {
        const R dtr = R(dt), rho0 = R(cfg.rest_density);
        const bool bk = cfg.pressure_form == 1;
#pragma omp parallel for schedule(static)
        for (long long i = 0; i < (long long)n; ++i) {
            const V3<R> pi = x[i];
            const R rho_i = rho[i];
            const R k_i = kap[i];
            const R ki = (bk || ORC_MUTANT == 6) ? k_i / o_fmax(rho_i, k1em9<R>())
                                                 : k_i / o_fmax(rho_i * rho_i, k1em9<R>());
            V3<R> dl{0, 0, 0};
            fg.for_neighbors(pi, [&](u32 j) {
                if (j == u32(i)) return;
                const V3<R> pj = x[j];
                const R rho_j = rho[j];
                const R k_j = kap[j];
                const V3<R> g = gradW(V3<R>{pi.x - pj.x, pi.y - pj.y, pi.z - pj.z}, h);
                const R kj = bk ? k_j / o_fmax(rho_j, k1em9<R>())
                                : k_j / o_fmax(rho_j * rho_j, k1em9<R>());
                const R c2 = mass * (ki + kj);
                dl.x -= dtr * c2 * g.x;
                dl.y -= dtr * c2 * g.y;
                dl.z -= dtr * c2 * g.z;
            }, hc(i));
            if (nb > 0) {
                const R bf = ORC_MUTANT == 7 ? R(0)
                             : bk ? k_i / o_fmax(rho0, k1em9<R>())
                                  : k_i / o_fmax(rho0 * rho0, k1em9<R>());
                bg.for_neighbors(pi, [&](u32 j) {
                    const V3<R> pj = xb[j];
                    const V3<R> g = gradW(V3<R>{pi.x - pj.x, pi.y - pj.y, pi.z - pj.z}, h);
                    const R c2 = mb[j] * (ki + bf);
                    dl.x -= dtr * c2 * g.x;
                    dl.y -= dtr * c2 * g.y;
                    dl.z -= dtr * c2 * g.z;
                }, hc(i));
            }
            v[i].x += dl.x;
            v[i].y += dl.y;
            v[i].z += dl.z;
        }
    }

    // mean(max(0, a_i - ref)), accumulated in binary64 (dfsph.cu:277-301)
    double mean_excess(const std::vector<R>& a, R ref) const {
        double s = 0;
        for (size_t i = 0; i < n; ++i) s += double(o_fmax(R(0), a[i] - ref));
        return s / double(n);
    }

    // dfsph.cu:215-260
    // Iteration count of a solve in a replay (kNone: decide by the error test).
    static constexpr u32 kFree = 0xFFFFFFFFu;
    u32 replay_iters() {
        if (tape_mode != 2) return kFree;
        return tape_iter < tape->iters.size() ? tape->iters[tape_iter++] : 0u;
    }
    void record_iters(u32 k) {
        if (tape_mode == 1) tape->iters.push_back(k);
    }

    void density_solve(float dt) {
        const double thr = double(cfg.eta_density * cfg.rest_density);
        const u32 fixed = replay_iters();
        st.density_iters = 0;
        for (u32 it = 0; it < cfg.max_density_iters; ++it) {
            if (fixed != kFree && it >= fixed) break;
            density_change(dt);
            apply_kappa(kappa, dt);
            st.density_iters = it + 1;
            const double err = mean_excess(rho_adv, R(cfg.rest_density));
            st.mean_density_error = err;
            if (fixed == kFree && it >= 1 && err < thr) break;
        }
        record_iters(st.density_iters);
    }

    // divergence_solve (docs/superpowers/plans/2026-04-24-water-simulation-v2-phase2-dfsph-solver.md:1967-1990)
    void divergence_solve(float dt) {
        const double thr = double(cfg.eta_divergence * cfg.rest_density);
        const u32 fixed = replay_iters();
        st.divergence_iters = 0;
        for (u32 it = 0; it < cfg.max_divergence_iters; ++it) {
            if (fixed != kFree && it >= fixed) break;
            divergence_change(dt);
            apply_kappa(kappa_v, dt);
            st.divergence_iters = it + 1;
            const double err = mean_excess(rho_adv, R(0));
            st.mean_divergence_error = err;
            if (fixed == kFree && err < thr) break;
        }
        record_iters(st.divergence_iters);
    }

    // dfsph_kernels.cu:363-386 (dfsph.cu:262-273: |v| clamp 5 m/s)
    void advect(float dt) {
        const R dtr = R(dt), vmax = R(5.0f), eps = R(1e-3f);
        const R lox = R(cfg.domain_min[0]) + eps, loy = R(cfg.domain_min[1]) + eps,
                loz = R(cfg.domain_min[2]) + eps;
        const R hix = R(cfg.domain_max[0]) - eps, hiy = R(cfg.domain_max[1]) - eps,
                hiz = R(cfg.domain_max[2]) - eps;
        const uint16_t* cin = nullptr;
        uint16_t* cout = nullptr;
        if (tape_mode == 1) {
            clamp_buf.assign(n, 0);
            cout = clamp_buf.data();
        } else if (tape_mode == 2 && tape_clamp < tape->clamps.size()) {
            cin = tape->clamps[tape_clamp++].data();
        }
        const R lo[3] = {lox, loy, loz}, hi[3] = {hix, hiy, hiz};
#pragma omp parallel for schedule(static)
        for (long long i = 0; i < (long long)n; ++i) {
            V3<R> vv = v[i];
            R* vc[3] = {&vv.x, &vv.y, &vv.z};
            uint16_t f = 0;
            const R sp = o_sqrt(vv.x * vv.x + vv.y * vv.y + vv.z * vv.z);
            const bool speed_on = !(g_free_advect & 1), box_on = !(g_free_advect & 2);
            if (speed_on && (cin ? (cin[i] & kClampSpeed) : (sp > vmax))) {
                const R s = vmax / sp;
                vv.x *= s; vv.y *= s; vv.z *= s;
                f |= kClampSpeed;
            }
            V3<R> p{x[i].x + vv.x * dtr, x[i].y + vv.y * dtr, x[i].z + vv.z * dtr};
            R* pc[3] = {&p.x, &p.y, &p.z};
            for (int a = 0; a < 3; ++a) {
                if (box_on && (cin ? (cin[i] & (kClampLo << a)) : (*pc[a] < lo[a]))) {
                    *pc[a] = lo[a];
                    f |= kClampLo << a;
                    if (cin ? (cin[i] & (kZeroLo << a)) : (*vc[a] < R(0))) {
                        *vc[a] = R(0);
                        f |= kZeroLo << a;
                    }
                }
            }
            for (int a = 0; a < 3; ++a) {
                if (box_on && (cin ? (cin[i] & (kClampHi << a)) : (*pc[a] > hi[a]))) {
                    *pc[a] = hi[a];
                    f |= kClampHi << a;
                    if (cin ? (cin[i] & (kZeroHi << a)) : (*vc[a] > R(0))) {
                        *vc[a] = R(0);
                        f |= kZeroHi << a;
                    }
                }
            }
            if (cout) cout[i] = f;
            v[i] = vv;
            x[i] = p;
        }
        if (tape_mode == 1) tape->clamps.push_back(clamp_buf);
    }

    // dfsph_kernels.cu:396-400 + CUB max (dfsph.cu:303-326), plus the
    // regression diagnostics: escapees (a position on or outside the
    // advection's clamp box, dfsph_kernels.cu:376-385: the particle reached
    // the domain limit) and binary64 mass, momentum and kinetic energy of the
    // particles with a finite state, summed in index order.
    void finish_stats() {
        R m = 0;
        u32 nan_count = 0, esc = 0;
        double mass_sum = 0, px = 0, py = 0, pz = 0, ke = 0;
        const R eps = R(1e-3f);
        const R lo[3] = {R(cfg.domain_min[0]) + eps, R(cfg.domain_min[1]) + eps,
                         R(cfg.domain_min[2]) + eps};
        const R hi[3] = {R(cfg.domain_max[0]) - eps, R(cfg.domain_max[1]) - eps,
                         R(cfg.domain_max[2]) - eps};
        const double md = double(mass);
        for (size_t i = 0; i < n; ++i) {
            const V3<R>& vv = v[i];
            const R s = o_sqrt(vv.x * vv.x + vv.y * vv.y + vv.z * vv.z);
            m = o_fmax(m, s);
            const V3<R>& p = x[i];
            if (p.x <= lo[0] || p.y <= lo[1] || p.z <= lo[2] || p.x >= hi[0] || p.y >= hi[1] ||
                p.z >= hi[2])
                ++esc;
            if (!o_isfinite(x[i].x) || !o_isfinite(x[i].y) || !o_isfinite(x[i].z) ||
                !o_isfinite(vv.x) || !o_isfinite(vv.y) || !o_isfinite(vv.z)) {
                ++nan_count;
                continue;
            }
            const double a = double(vv.x), b = double(vv.y), c = double(vv.z);
            mass_sum += md;
            px += md * a;
            py += md * b;
            pz += md * c;
            ke += 0.5 * md * ((a * a + b * b) + c * c);
        }
        st.max_velocity = float(m);
        st.nan_count = nan_count;
        st.escapees = esc;
        st.total_mass = mass_sum;
        st.momentum[0] = px;
        st.momentum[1] = py;
        st.momentum[2] = pz;
        st.kinetic_energy = ke;
    }

    void density_alpha() {
        home = nullptr;
        if (tape_mode == 1) {
            home_buf.resize(3 * n);
            for (size_t i = 0; i < n; ++i) fg.home_of(x[i], home_buf.data() + 3 * i);
            tape->homes.push_back(home_buf);
            home = home_buf.data();
        } else if (tape_mode == 2 && tape_home < tape->homes.size()) {
            home = tape->homes[tape_home++].data();
        }
        fg.build(x, n, home);
        density();
        alpha_factor();
    }

    // dfsph.cu:120-133 (+ divergence solve after advect, docs/superpowers/plans/2026-04-24-water-simulation-v2-phase2-dfsph-solver.md:1992-2003)
    void step(float dt) {
        if (n == 0) return;
        density_alpha();
        gravity(dt);
        xsph();
        surface_tension(dt);
        density_solve(dt);
        advect(dt);
        if (cfg.enable_divergence_solve) {
            density_alpha();
            divergence_solve(dt);
        }
        finish_stats();
    }
};

template <class R> void get_field(Solver<R>* s, int field, double* out) {
    auto put3 = [&](const std::vector<V3<R>>& a) {
        for (size_t i = 0; i < a.size(); ++i) {
            out[3 * i] = double(a[i].x);
            out[3 * i + 1] = double(a[i].y);
            out[3 * i + 2] = double(a[i].z);
        }
    };
    auto put1 = [&](const std::vector<R>& a) {
        for (size_t i = 0; i < a.size(); ++i) out[i] = double(a[i]);
    };
    switch (field) {
        case ORC_FIELD_X: put3(s->x); break;
        case ORC_FIELD_V: put3(s->v); break;
        case ORC_FIELD_RHO: put1(s->rho); break;
        case ORC_FIELD_ALPHA: put1(s->alpha); break;
        case ORC_FIELD_KAPPA: put1(s->kappa); break;
        case ORC_FIELD_RHO_ADV: put1(s->rho_adv); break;
        case ORC_FIELD_BMASS: put1(s->mb); break;
        case ORC_FIELD_XB: put3(s->xb); break;
        default: break;
    }
}

struct Handle {
    int precision;  // ORC_F32, ORC_F64 or ORC_DUAL
    void* impl;
};

// Bit spread of a 10-bit value: bit b moves to bit 3b (the expand step of
// spatial_accel.cu:11-17, restated as a plain loop).
u32 spread3(u32 v) {
    u32 out = 0;
    for (int b = 0; b < 10; ++b) out |= ((v >> b) & 1u) << (3 * b);
    return out;
}

}  // namespace

extern "C" {

void orc_set_xsph_in_place(int on) { g_xsph_in_place = on; }
void orc_set_analysis(int freeze_geometry, int free_advect) {
    g_freeze_geometry = freeze_geometry;
    g_free_advect = free_advect;
}

void orc_set_threads(int n) { omp_set_num_threads(n > 0 ? n : 1); }
int orc_get_threads(void) { return omp_get_max_threads(); }

float orc_W_f32(float r, float l) { return W<float>(r, l); }
double orc_W_f64(double r, double l) { return W<double>(r, l); }
void orc_gradW_f32(const float r[3], float l, float out[3]) {
    const V3<float> g = gradW<float>({r[0], r[1], r[2]}, l);
    out[0] = g.x; out[1] = g.y; out[2] = g.z;
}
void orc_gradW_f64(const double r[3], double l, double out[3]) {
    const V3<double> g = gradW<double>({r[0], r[1], r[2]}, l);
    out[0] = g.x; out[1] = g.y; out[2] = g.z;
}

void orc_cell_keys(const float* xyz, size_t n, const float origin[3], const int32_t dims[3],
                   float cell, uint32_t* keys) {
    Grid<float> g;
    g.init(origin, dims, cell);
    for (size_t i = 0; i < n; ++i) keys[i] = g.key_of(load3<float>(xyz, i));
}

void orc_grid_build(const float* xyz, size_t n, const float origin[3], const int32_t dims[3],
                    float cell, uint32_t* sorted_keys, uint32_t* perm, uint32_t* cell_start) {
    Grid<float> g;
    g.init(origin, dims, cell);
    std::vector<V3<float>> x(n);
    for (size_t i = 0; i < n; ++i) x[i] = load3<float>(xyz, i);
    g.build(x, n);
    if (n) {
        std::memcpy(sorted_keys, g.sorted_keys.data(), n * sizeof(u32));
        std::memcpy(perm, g.perm.data(), n * sizeof(u32));
    }
    std::memcpy(cell_start, g.cs.data(), g.cs.size() * sizeof(u32));
}

void orc_make_grid(const float dmin[3], const float dmax[3], float h, float origin[3],
                   int32_t dims[3]) {
    make_grid_params(dmin, dmax, h, origin, dims);
}

void orc_neighbor_counts(const float* xq, size_t nq, const float* xs, size_t ns,
                         const float origin[3], const int32_t dims[3], float cell, float h,
                         uint32_t* cand, uint32_t* nbr) {
    Grid<float> g;
    g.init(origin, dims, cell);
    std::vector<V3<float>> src(ns);
    for (size_t i = 0; i < ns; ++i) src[i] = load3<float>(xs, i);
    g.build(src, ns);
#pragma omp parallel for schedule(static)
    for (long long i = 0; i < (long long)nq; ++i) {
        const V3<float> p = load3<float>(xq, i);
        u32 c = 0, k = 0;
        g.for_neighbors(p, [&](u32 j) {
            ++c;
            const V3<float> r{p.x - src[j].x, p.y - src[j].y, p.z - src[j].z};
            const float d = std::sqrt(r.x * r.x + r.y * r.y + r.z * r.z);
            if (d / h < 2.0f) ++k;
        });
        cand[i] = c;
        nbr[i] = k;
    }
}

// spatial_accel.cu:30-40 (codes), :139-148 (host inverse extent), :158-171 (sort)
void orc_morton(const float* xyz, size_t n, const float bmin[3], const float bmax[3],
                uint32_t* codes, uint32_t* sorted_codes, uint32_t* perm) {
    float inv[3];
    for (int a = 0; a < 3; ++a) {
        const float e = bmax[a] - bmin[a];
        inv[a] = e > 0 ? 1.0f / e : 0.0f;
    }
    std::vector<u32> c(n);
    for (size_t i = 0; i < n; ++i) {
        u32 q[3];
        for (int a = 0; a < 3; ++a) {
            const float t = (xyz[3 * i + a] - bmin[a]) * inv[a];
            const float s = std::fmin(std::fmax(t * 1024.0f, 0.0f), 1023.0f);
            q[a] = static_cast<u32>(s);
        }
        c[i] = (spread3(q[0]) << 2) | (spread3(q[1]) << 1) | spread3(q[2]);
        if (codes) codes[i] = c[i];
    }
    std::vector<u32> idx(n);
    for (size_t i = 0; i < n; ++i) idx[i] = u32(i);
    std::stable_sort(idx.begin(), idx.end(), [&](u32 a, u32 b) { return c[a] < c[b]; });
    for (size_t i = 0; i < n; ++i) {
        if (perm) perm[i] = idx[i];
        if (sorted_codes) sorted_codes[i] = c[idx[i]];
    }
}

// Karras 2012 radix tree over the sorted codes (spatial_accel.cu:44-100):
// common prefix of leaves i, j in the total order (code, index); direction,
// range end and split of every internal node.  Then the boxes: leaves are
// their points (spatial_accel.cu:102-113, radius 0) and internal nodes the
// union of their children, filled children-first by an explicit post-order.
static int lbvh_prefix(const uint32_t* c, long long n, long long i, long long j) {
    if (j < 0 || j >= n) return -1;
    if (c[i] == c[j]) {
        const uint32_t x = static_cast<uint32_t>(i ^ j);
        return 32 + (x ? __builtin_clz(x) : 32);
    }
    return __builtin_clz(c[i] ^ c[j]);
}

void orc_lbvh(const uint32_t* codes, const float* xyz, const uint32_t* perm, size_t count,
              uint32_t* left, uint32_t* right, uint32_t* parent, float* boxes) {
    const long long n = static_cast<long long>(count), ni = n - 1;
    if (n == 0) return;
    parent[0] = 0xFFFFFFFFu;
    for (long long i = 0; i < ni; ++i) {
        const long long d = lbvh_prefix(codes, n, i, i + 1) - lbvh_prefix(codes, n, i, i - 1) > 0 ? 1 : -1;
        const int dmin = lbvh_prefix(codes, n, i, i - d);
        long long lmax = 2;
        while (lbvh_prefix(codes, n, i, i + lmax * d) > dmin) lmax *= 2;
        long long l = 0;
        for (long long t = lmax / 2; t > 0; t /= 2)
            if (lbvh_prefix(codes, n, i, i + (l + t) * d) > dmin) l += t;
        const long long j = i + l * d;
        const int dnode = lbvh_prefix(codes, n, i, j);
        long long s = 0;
        for (long long t = (l + 1) / 2; t > 0; t = (t == 1) ? 0 : (t + 1) / 2)
            if (lbvh_prefix(codes, n, i, i + (s + t) * d) > dnode) s += t;
        const long long g = i + s * d + (d < 0 ? -1 : 0);
        const long long lo = std::min(i, j), hi = std::max(i, j);
        left[i] = static_cast<uint32_t>(lo == g ? ni + g : g);
        right[i] = static_cast<uint32_t>(hi == g + 1 ? ni + g + 1 : g + 1);
        parent[left[i]] = static_cast<uint32_t>(i);
        parent[right[i]] = static_cast<uint32_t>(i);
    }
    for (long long k = 0; k < n; ++k) {
        float* b = boxes + 6 * (ni + k);
        for (int a = 0; a < 3; ++a) b[a] = b[a + 3] = xyz[3 * size_t(perm[k]) + a];
    }
    if (ni == 0) return;
    std::vector<std::pair<uint32_t, bool>> stack{{0u, false}};
    while (!stack.empty()) {
        auto [node, done] = stack.back();
        stack.pop_back();
        if (node >= ni) continue;  // a leaf
        if (!done) {
            stack.push_back({node, true});
            stack.push_back({left[node], false});
            stack.push_back({right[node], false});
            continue;
        }
        const float* a = boxes + 6 * size_t(left[node]);
        const float* b = boxes + 6 * size_t(right[node]);
        float* c = boxes + 6 * size_t(node);
        for (int k = 0; k < 3; ++k) {
            c[k] = std::fmin(a[k], b[k]);
            c[k + 3] = std::fmax(a[k + 3], b[k + 3]);
        }
    }
}

// boundary_sampler.cpp:6-32: six faces (-X,+X,-Y,+Y,-Z,+Z), n_layers inward
// layers each, binary32 accumulating loops over the two in-face axes.
size_t orc_sample_aabb_boundary(const float lo[3], const float hi[3], float spacing,
                                int n_layers, float* out, size_t cap) {
    size_t count = 0;
    const int axes[6] = {0, 0, 1, 1, 2, 2};
    for (int f = 0; f < 6; ++f) {
        const int axis = axes[f];
        const bool high = (f & 1) != 0;
        const float face = high ? hi[axis] : lo[axis];
        const float step = high ? -spacing : spacing;
        const int ua = (axis + 1) % 3, va = (axis + 2) % 3;
        for (int layer = 0; layer < n_layers; ++layer) {
            const float apos = face + layer * step;
            for (float u = lo[ua]; u <= hi[ua] + 0.5f * spacing; u += spacing)
                for (float w = lo[va]; w <= hi[va] + 0.5f * spacing; w += spacing) {
                    if (out && count < cap) {
                        float* p = out + 3 * count;
                        p[axis] = apos;
                        p[ua] = u;
                        p[va] = w;
                    }
                    ++count;
                }
        }
    }
    return count;
}

// main.cpp:65-74 / test_settled_puddle.cu:24-32: x-major, binary32 accumulation
size_t orc_lattice_block(const float lo[3], const float hi[3], float spacing, float* out,
                         size_t cap) {
    size_t count = 0;
    for (float a = lo[0]; a <= hi[0]; a += spacing)
        for (float b = lo[1]; b <= hi[1]; b += spacing)
            for (float c = lo[2]; c <= hi[2]; c += spacing) {
                if (out && count < cap) {
                    out[3 * count] = a;
                    out[3 * count + 1] = b;
                    out[3 * count + 2] = c;
                }
                ++count;
            }
    return count;
}

// exp(-damping * dt) rounded to binary32 (dfsph_kernels.cu:351; the
// reference takes fast-math expf per particle): the product in binary32,
// then exp in binary64 by k = round(x / ln2), r = x - k ln2 (ln2 in a 32-bit
// head and a tail), a degree-13 Taylor polynomial of e^r in Horner form,
// 2^k applied exactly, one rounding to binary32.  The library evaluates
// the same sequence of IEEE operations on the host and on the device.
float orc_damping_decay(float damping, float dt) {
    const float xf = -damping * dt;
    const double x = static_cast<double>(xf);
    if (!(x > -745.0)) return std::isnan(x) ? xf : 0.0f;
    if (x > 709.0) return std::numeric_limits<float>::infinity();
    const double kd = x * 1.4426950408889634;
    const int k = static_cast<int>(kd + (kd >= 0.0 ? 0.5 : -0.5));
    const double r = (x - static_cast<double>(k) * 6.93147180369123816490e-01) -
                     static_cast<double>(k) * 1.90821492927058770002e-10;
    double p = 1.0;
    for (int j = 13; j >= 1; --j) p = 1.0 + p * (r / static_cast<double>(j));
    const int k1 = k / 2, k2 = k - k1;
    return static_cast<float>(std::ldexp(std::ldexp(p, k1), k2));
}

// timestep.h:23-28
float orc_next_dt(float frame_dt, float dt_max, float cfl_lambda, uint32_t max_substeps,
                  float max_velocity, float particle_radius, float remaining) {
    const float v = std::max(max_velocity, 1e-6f);
    const float dt_cfl = cfl_lambda * particle_radius / v;
    const float dt_min = frame_dt / static_cast<float>(max_substeps);
    return std::min({dt_max, dt_cfl, remaining, std::max(dt_cfl, dt_min)});
}

// dfsph.h:19-33 defaults
void orc_config_default(orc_config* c) {
    std::memset(c, 0, sizeof(*c));
    c->rest_density = 1000.0f;
    c->particle_radius = 0.0015f;
    c->smoothing_length = 0.003f;
    c->viscosity = 1e-3f;
    c->damping = 0.5f;
    c->surface_tension = 0.0f;
    c->gravity[0] = 0.0f; c->gravity[1] = -9.81f; c->gravity[2] = 0.0f;
    c->max_density_iters = 100;
    c->max_divergence_iters = 100;
    c->eta_density = 1e-3f;
    c->eta_divergence = 1e-3f;
    c->domain_min[0] = c->domain_min[1] = c->domain_min[2] = 0.0f;
    c->domain_max[0] = c->domain_max[1] = c->domain_max[2] = 1.0f;
    c->enable_divergence_solve = 0;
    c->pressure_form = 0;
    c->enable_surface_tension = 0;
}

void* orc_solver_create(const orc_config* cfg, int precision, const float* x, const float* v,
                        size_t n, const float* xb, size_t nb) {
    Handle* h = new Handle{precision, nullptr};
    if (precision == ORC_F64) h->impl = new Solver<double>(*cfg, x, v, n, xb, nb);
    else if (precision == ORC_DUAL) h->impl = new Solver<DualF>(*cfg, x, v, n, xb, nb);
    else h->impl = new Solver<float>(*cfg, x, v, n, xb, nb);
    return h;
}

void orc_solver_destroy(void* s) {
    Handle* h = static_cast<Handle*>(s);
    if (!h) return;
    if (h->precision == ORC_F64) delete static_cast<Solver<double>*>(h->impl);
    else if (h->precision == ORC_DUAL) delete static_cast<Solver<DualF>*>(h->impl);
    else delete static_cast<Solver<float>*>(h->impl);
    delete h;
}

#define ORC_DISPATCH(s, call)                                         \
    do {                                                              \
        Handle* h_ = static_cast<Handle*>(s);                         \
        if (h_->precision == ORC_F64) {                               \
            auto* S = static_cast<Solver<double>*>(h_->impl);         \
            call;                                                     \
        } else if (h_->precision == ORC_DUAL) {                       \
            auto* S = static_cast<Solver<DualF>*>(h_->impl);          \
            call;                                                     \
        } else {                                                      \
            auto* S = static_cast<Solver<float>*>(h_->impl);          \
            call;                                                     \
        }                                                             \
    } while (0)

void orc_solver_seed(void* s, const double* dx, const double* dv, double d_viscosity) {
    Handle* h = static_cast<Handle*>(s);
    if (h->precision != ORC_DUAL) return;
    auto* S = static_cast<Solver<DualF>*>(h->impl);
    for (size_t i = 0; i < S->n; ++i) {
        if (dx) { S->x[i].x.d = dx[3 * i]; S->x[i].y.d = dx[3 * i + 1]; S->x[i].z.d = dx[3 * i + 2]; }
        if (dv) { S->v[i].x.d = dv[3 * i]; S->v[i].y.d = dv[3 * i + 1]; S->v[i].z.d = dv[3 * i + 2]; }
    }
    S->eps_param.d = d_viscosity;
}

void orc_solver_get_tangent(void* s, int field, double* out) {
    Handle* h = static_cast<Handle*>(s);
    if (h->precision != ORC_DUAL) return;
    auto* S = static_cast<Solver<DualF>*>(h->impl);
    auto put3 = [&](const std::vector<V3<DualF>>& a) {
        for (size_t i = 0; i < a.size(); ++i) {
            out[3 * i] = a[i].x.d;
            out[3 * i + 1] = a[i].y.d;
            out[3 * i + 2] = a[i].z.d;
        }
    };
    auto put1 = [&](const std::vector<DualF>& a) {
        for (size_t i = 0; i < a.size(); ++i) out[i] = a[i].d;
    };
    switch (field) {
        case ORC_FIELD_X: put3(S->x); break;
        case ORC_FIELD_V: put3(S->v); break;
        case ORC_FIELD_RHO: put1(S->rho); break;
        case ORC_FIELD_ALPHA: put1(S->alpha); break;
        case ORC_FIELD_KAPPA: put1(S->kappa); break;
        case ORC_FIELD_RHO_ADV: put1(S->rho_adv); break;
        default: break;
    }
}

void* orc_tape_create(void) { return new StructTape(); }
void orc_tape_destroy(void* t) { delete static_cast<StructTape*>(t); }
size_t orc_tape_builds(void* t) { return static_cast<StructTape*>(t)->homes.size(); }

void orc_solver_set_tape(void* s, void* tape, int mode) {
    ORC_DISPATCH(s, {
        S->tape = static_cast<StructTape*>(tape);
        S->tape_mode = tape ? mode : 0;
        S->tape_home = 0;
        S->tape_iter = 0;
        S->tape_mask = 0;
        S->tape_clamp = 0;
    });
}

void orc_solver_step(void* s, float dt) { ORC_DISPATCH(s, S->step(dt)); }
void orc_solver_density_alpha(void* s) { ORC_DISPATCH(s, S->density_alpha()); }
void orc_solver_surface_tension(void* s, float dt) { ORC_DISPATCH(s, S->surface_tension(dt)); }

void orc_solver_set_state(void* s, const float* x, const float* v) {
    ORC_DISPATCH(s, {
        using RR = std::remove_reference_t<decltype(S->mass)>;
        for (size_t i = 0; i < S->n; ++i) {
            if (x) S->x[i] = load3<RR>(x, i);
            if (v) S->v[i] = load3<RR>(v, i);
        }
    });
}

void orc_solver_set_state_f64(void* s, const double* x, const double* v) {
    ORC_DISPATCH(s, {
        using RR = std::remove_reference_t<decltype(S->mass)>;
        for (size_t i = 0; i < S->n; ++i) {
            if (x) {
                S->x[i].x = RR(x[3 * i]);
                S->x[i].y = RR(x[3 * i + 1]);
                S->x[i].z = RR(x[3 * i + 2]);
            }
            if (v) {
                S->v[i].x = RR(v[3 * i]);
                S->v[i].y = RR(v[3 * i + 1]);
                S->v[i].z = RR(v[3 * i + 2]);
            }
        }
    });
}

void orc_solver_set_viscosity_f64(void* s, double viscosity) {
    ORC_DISPATCH(s, {
        using RR = std::remove_reference_t<decltype(S->mass)>;
        S->cfg.viscosity = float(viscosity);
        S->eps_param = RR(viscosity);
    });
}

void orc_solver_get(void* s, int field, double* out) { ORC_DISPATCH(s, get_field(S, field, out)); }

void orc_solver_get_f32(void* s, int field, float* out) {
    Handle* h = static_cast<Handle*>(s);
    if (h->precision != ORC_F32) return;
    auto* S = static_cast<Solver<float>*>(h->impl);
    auto put3 = [&](const std::vector<V3<float>>& a) {
        std::memcpy(out, a.data(), a.size() * sizeof(V3<float>));
    };
    auto put1 = [&](const std::vector<float>& a) {
        std::memcpy(out, a.data(), a.size() * sizeof(float));
    };
    switch (field) {
        case ORC_FIELD_X: put3(S->x); break;
        case ORC_FIELD_V: put3(S->v); break;
        case ORC_FIELD_RHO: put1(S->rho); break;
        case ORC_FIELD_ALPHA: put1(S->alpha); break;
        case ORC_FIELD_KAPPA: put1(S->kappa); break;
        case ORC_FIELD_RHO_ADV: put1(S->rho_adv); break;
        case ORC_FIELD_BMASS: put1(S->mb); break;
        case ORC_FIELD_XB: put3(S->xb); break;
        default: break;
    }
}

void orc_solver_stats(void* s, orc_stats* out) { ORC_DISPATCH(s, *out = S->st); }
float orc_solver_particle_mass(void* s) {
    float m = 0;
    ORC_DISPATCH(s, m = float(S->mass));
    return m;
}
size_t orc_solver_count(void* s) {
    size_t n = 0;
    ORC_DISPATCH(s, n = S->n);
    return n;
}

}  // extern "C"
