// device_math.hpp -- device helpers shared by the grid, sort and solver kernels.
//
// Every floating-point expression keeps the operand order of the reference
// (core/include/water/core/sph_kernels.cuh, solvers/dfsph/src/dfsph_kernels.cu)
// and the library is compiled with -ffp-contract=off and IEEE divide/sqrt, so
// each value equals the binary32 result of the same expression on the host.
#pragma once

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdint>

namespace wsim {

using u32 = uint32_t;
using u64 = uint64_t;

constexpr int kWave = 64;
constexpr int kXcds = 8;  // MI355X: 8 accelerator complex dies, one L2 each

// Geometry of a uniform grid (cell_grid.h:16-50): origin, host-computed
// inverse cell length (cell_grid.cu:81) and cells per axis.
struct GridGeom {
    float ox, oy, oz;
    float inv;  // 1.0f / cell_length, IEEE, computed once on the host
    int dx, dy, dz;
    u32 C;      // total cells == sentinel key

    // (int)floorf((p - o) * inv) (cell_grid.cu:11-13), made total: NaN and
    // coordinates beyond the stencil's reach clamp to a value whose whole
    // 3x3x3 stencil lies outside the grid.
    __host__ __device__ static inline int coord(float p, float o, float inv, int D) {
        const float f = floorf((p - o) * inv);
        if (!(f >= -2.0f)) return -2;
        if (f > static_cast<float>(D + 1)) return D + 1;
        return static_cast<int>(f);
    }
    __host__ __device__ inline int cx(float p) const { return coord(p, ox, inv, dx); }
    __host__ __device__ inline int cy(float p) const { return coord(p, oy, inv, dy); }
    __host__ __device__ inline int cz(float p) const { return coord(p, oz, inv, dz); }

    // Linear x-major key (cx*Dy + cy)*Dz + cz; out of range -> sentinel C
    // (cell_grid.cu:14-21).
    __host__ __device__ inline u32 key(int x, int y, int z) const {
        if (x < 0 || x >= dx || y < 0 || y >= dy || z < 0 || z >= dz) return C;
        return (static_cast<u32>(x) * static_cast<u32>(dy) + static_cast<u32>(y)) *
                   static_cast<u32>(dz) +
               static_cast<u32>(z);
    }
    __host__ __device__ inline u32 key_of(float px, float py, float pz) const {
        return key(cx(px), cy(py), cz(pz));
    }
};

// exp(-damping * dt) rounded to binary32, the velocity decay of a substep
// (dfsph_kernels.cu:351 evaluates it per particle with fast-math expf).  The
// same operations run on the host and on the device: the product in
// binary32, then exp in binary64 by reduction x = k ln2 + r (ln2 split in a
// 32-bit head and a tail, |r| <= ln2/2), a degree-13 Taylor polynomial in
// Horner form (truncation error below 1e-16 relative), scaling by 2^k and
// one rounding to binary32.  Every step is an IEEE operation without
// contraction, so a time step computed on the device gives the decay the
// host computes for the same step, bit for bit.
__host__ __device__ inline float damping_decay(float damping, float dt) {
    const float xf = -damping * dt;
    const double x = static_cast<double>(xf);
    if (!(x > -745.0)) return (x != x) ? xf : 0.0f;  // NaN stays NaN; underflow
    if (x > 709.0) return static_cast<float>(1e300 * 1e300);  // +inf
    const double kd = x * 1.4426950408889634;  // x / ln2
    const int k = static_cast<int>(kd + (kd >= 0.0 ? 0.5 : -0.5));
    const double ln2_hi = 6.93147180369123816490e-01;  // 32 significant bits
    const double ln2_lo = 1.90821492927058770002e-10;
    const double r = (x - static_cast<double>(k) * ln2_hi) - static_cast<double>(k) * ln2_lo;
    double p = 1.0;
    for (int j = 13; j >= 1; --j) p = 1.0 + p * (r / static_cast<double>(j));
    // times 2^k (k in [-1075, 1024]) in two exact scalings
    const int k1 = k / 2, k2 = k - k1;
    return static_cast<float>(ldexp(ldexp(p, k1), k2));
}

// Constants of the cubic spline for one smoothing length l, hoisted out of
// the pair loop.  sigma = 1/(pi*l*l*l) is the same binary32 value the
// reference recomputes per call (sph_kernels.cuh:15,31).
struct Spline {
    float l;
    float sigma;
    float sigma_m075;  // sigma * (-0.75f), the left product of sph_kernels.cuh:37
    float sigma_025;   // sigma * 0.25f, the left product of sph_kernels.cuh:20
    // Smallest squared distance r2 with sqrtf(r2) / l >= 2 in binary32: every
    // r2 >= far2 has W = 0 and grad W = 0 (sqrtf and the IEEE quotient are
    // monotone), so stencil candidates that far are skipped without the
    // square root and divisions.  Their terms would add +-0 to sums that
    // never hold -0, so skipping them changes no bit.
    float far2;
    float inv_l;  // RN(1 / l)
    int fast;     // the pair-path sqrt and quotients below apply (pair_ops_in_range)
    // Tolerance-mode constants (PairMath<true>): grad W = r * c with
    // c = ck_in * (2.25 q - 3) for q < 1 (sigma (-3 q + 2.25 q^2) / (l d),
    // q / d = 1 / l) and c = ck_out * (2 - q)^2 / d for 1 <= q < 2.
    float ck_in;   // sigma / l^2
    float ck_out;  // -0.75 sigma / l
    static Spline make(float l);
};

// ---- correctly rounded sqrt and quotients on the pair path ----------------
// The compiler's IEEE sqrt and divide wrap the hardware approximations
// (v_sqrt_f32, v_rcp_f32) in range scaling and special-case fix-ups.  The
// pair path keeps the rounding steps and drops what its operands never need:
//
//  * pair_sqrt: v_sqrt_f32 and the same two fma residual corrections, i.e.
//    the compiler's sqrtf sequence without its scaling for x < 2^-96 and its
//    +-0 / +inf pass-through, so it equals sqrtf for +0, NaN, +inf and every
//    x >= 2^-96.  Below 2^-96 (d < 2^-48) it returns an approximation, and
//    the callers' results do not depend on d there: W = sigma * 1 exactly
//    once q < 2^-13 (l >= 2^-30), and the gradient factor is 0 for d < 1e-12.
//  * pair_q: d / l as q0 = d * RN(1/l), r = fma(-q0, l, d) (exact),
//    fma(r, RN(1/l), q0).  RN(1/l) is within half an ulp of 1/l and q0
//    within one ulp of d/l, so the corrected quotient is the correctly
//    rounded one (Markstein's theorem); scaling d by a power of two scales
//    every step exactly while no step is subnormal (d >= 2^-48 here).
//  * pair_div: a / b as the compiler's Newton-Raphson and residual steps
//    without v_div_scale / v_div_fmas scaling / v_div_fixup, which are
//    identities when a, b, 1/b and a/b are normal and the exponents of a and
//    b differ by less than 96.  It is used for the gradient factor only,
//    whose operands pair_ops_in_range bounds from l: |dW/dq| in
//    [0.75 sigma min(q_min, 2^-46), sigma] with q_min = 1e-12 / l, and
//    l * d in [1e-12 l, 2 l^2].
//
// Spline::fast selects them when pair_ops_in_range(l) holds; otherwise every
// kernel uses the IEEE operations.  wsim_pair_ops_check compares all three
// with the IEEE operations on the device for every operand a given l can
// produce.
__device__ __forceinline__ float pair_sqrt_fast(float x) {
    const float s = __builtin_amdgcn_sqrtf(x);
    const float dn = __uint_as_float(__float_as_uint(s) - 1u);
    const float up = __uint_as_float(__float_as_uint(s) + 1u);
    const float t = __builtin_fmaf(-dn, s, x) <= 0.0f ? dn : s;
    return __builtin_fmaf(-up, s, x) > 0.0f ? up : t;
}
__device__ __forceinline__ float pair_q_fast(float d, float l, float inv_l) {
    const float q0 = d * inv_l;
    return __builtin_fmaf(__builtin_fmaf(-q0, l, d), inv_l, q0);
}
__device__ __forceinline__ float pair_div_fast(float a, float b) {
    float y = __builtin_amdgcn_rcpf(b);
    y = __builtin_fmaf(__builtin_fmaf(-b, y, 1.0f), y, y);
    float q = a * y;
    q = __builtin_fmaf(__builtin_fmaf(-b, q, a), y, q);
    return __builtin_fmaf(__builtin_fmaf(-b, q, a), y, q);
}
__device__ __forceinline__ float pair_sqrt(float x, const Spline& sp) {
    if (sp.fast) return pair_sqrt_fast(x);
    return sqrtf(x);
}
__device__ __forceinline__ float pair_q(float d, const Spline& sp) {
    if (sp.fast) return pair_q_fast(d, sp.l, sp.inv_l);
    return d / sp.l;
}
__device__ __forceinline__ float pair_div(float a, float b, const Spline& sp) {
    if (sp.fast) return pair_div_fast(a, b);
    return a / b;
}

// Host: the operand bounds above, with margins, in binary64.
inline bool pair_ops_in_range(float l) {
    if (!(l >= 0x1p-30f && l <= 0x1p30f)) return false;
    const double L = l, sigma = 1.0 / (3.14159265358979323846 * L * L * L);
    const double a_lo = 0.75 * sigma * std::min(1e-12 / L, std::ldexp(1.0, -46)) * 0.999;
    const double a_hi = sigma * 1.001, b_lo = 1e-12 * L * 0.999, b_hi = 2.0 * L * L * 1.001;
    const double lo = std::ldexp(1.0, -100), hi = std::ldexp(1.0, 100);
    return a_lo > lo && a_hi < hi && b_lo > lo && b_hi < hi && a_hi / b_lo < std::ldexp(1.0, 90) &&
           a_lo / b_hi > lo;
}

inline Spline Spline::make(float l) {
    Spline s;
    s.l = l;
    s.sigma = 1.0f / (3.14159265358979323846f * l * l * l);
    s.sigma_m075 = s.sigma * (-0.75f);
    s.sigma_025 = s.sigma * 0.25f;
    // start a little below (2l)^2 and walk up ulp by ulp
    float x = (2.0f * l) * (2.0f * l) * 0.99999f;
    for (int k = 0; k < 4096 && !(sqrtf(x) / l >= 2.0f); ++k) x = nextafterf(x, 3.0e38f);
    s.far2 = (sqrtf(x) / l >= 2.0f) ? x : 3.0e38f;
    s.inv_l = 1.0f / l;
    s.fast = pair_ops_in_range(l) ? 1 : 0;
    s.ck_in = static_cast<float>(static_cast<double>(s.sigma) / (static_cast<double>(l) * l));
    s.ck_out = static_cast<float>(-0.75 * static_cast<double>(s.sigma) / l);
    return s;
}

// W(d, l) for d = sqrtf(r.r) (sph_kernels.cuh:12-22).
__device__ inline float spline_W(float d, const Spline& sp) {
    const float q = pair_q(d, sp);
    if (q >= 2.0f) return 0.0f;
    if (q < 1.0f) return sp.sigma * (1.0f - 1.5f * q * q + 0.75f * q * q * q);
    const float t = 2.0f - q;
    return sp.sigma_025 * t * t * t;
}

// Scalar factor c of grad W = r * c (sph_kernels.cuh:25-41) given d = sqrtf(r.r);
// returns 0 where the reference returns the zero vector.
__device__ inline float spline_gradc(float d, const Spline& sp) {
    if (d < 1e-12f) return 0.0f;
    const float q = pair_q(d, sp);
    if (q >= 2.0f) return 0.0f;
    float dWdq;
    if (q < 1.0f) {
        dWdq = sp.sigma * (-3.0f * q + 2.25f * q * q);
    } else {
        const float t = 2.0f - q;
        dWdq = sp.sigma_m075 * t * t;
    }
    return pair_div(dWdq, sp.l * d, sp);
}

// W and the gradient factor for a candidate known to lie inside the support
// (r2 < Spline::far2, so q < 2): both polynomial pieces are evaluated and one
// is selected, which gives the same bits as spline_W / spline_gradc without
// their branches.
__device__ inline float spline_W_near(float d, const Spline& sp) {
    const float q = pair_q(d, sp);
    const float inner = sp.sigma * (1.0f - 1.5f * q * q + 0.75f * q * q * q);
    const float t = 2.0f - q;
    const float outer = sp.sigma_025 * t * t * t;
    return q < 1.0f ? inner : outer;
}
__device__ inline float spline_gradc_near(float d, const Spline& sp) {
    const float q = pair_q(d, sp);
    const float inner = sp.sigma * (-3.0f * q + 2.25f * q * q);
    const float t = 2.0f - q;
    const float outer = sp.sigma_m075 * t * t;
    const float c = pair_div(q < 1.0f ? inner : outer, sp.l * d, sp);
    return d < 1e-12f ? 0.0f : c;
}
// spline_gradc_near for a listed pair: its factor was not 0, so d >= 1e-12
// and the test is left out (same bits).
__device__ inline float spline_gradc_listed(float d, const Spline& sp) {
    const float q = pair_q(d, sp);
    const float inner = sp.sigma * (-3.0f * q + 2.25f * q * q);
    const float t = 2.0f - q;
    const float outer = sp.sigma_m075 * t * t;
    return pair_div(q < 1.0f ? inner : outer, sp.l * d, sp);
}

__device__ inline float norm3(float x, float y, float z) { return sqrtf(x * x + y * y + z * z); }
// The radicand of norm3, in the same operation order.
__device__ inline float dist2(float x, float y, float z) { return x * x + y * y + z * z; }

// ---- pair arithmetic of the neighbour kernels, two modes -----------------
// PairMath<false> (default): the reference's expressions in its operand
// order, no contraction, correctly rounded sqrt and quotients: every field
// equals the binary32 oracle bit for bit.
// PairMath<true> (tolerance mode, DFSPHConfig::fast_math): the same
// formulas evaluated the way the reference's own build does it
// (-use_fast_math, CMakeLists.txt:56): one v_rsq_f32 per pair gives 1/d
// and d = r2 / d, q = d * RN(1/l) without correction, the gradient factor
// without a divide (q / d = 1 / l inside the core, 1 / d = rsq outside),
// and products summed with fma.  Its fields differ from the binary32
// oracle in the last bits and are checked against the binary64 oracle
// within the tolerance the binary32 path itself meets.
template <bool F> struct PairMath;

template <> struct PairMath<false> {
    static constexpr bool fast = false;
    __device__ static inline float r2(float x, float y, float z) { return dist2(x, y, z); }
    // W for a candidate inside the support (r2 < far2)
    __device__ static inline float w(float r2, const Spline& sp) {
        return spline_W_near(pair_sqrt(r2, sp), sp);
    }
    __device__ static inline float c(float r2, const Spline& sp) {
        return spline_gradc_near(pair_sqrt(r2, sp), sp);
    }
    // c of a listed pair (c != 0 when the list was built)
    __device__ static inline float c_listed(float r2, const Spline& sp) {
        return spline_gradc_listed(pair_sqrt(r2, sp), sp);
    }
    __device__ static inline void wc(float r2, const Spline& sp, float& wv, float& cv) {
        const float d = pair_sqrt(r2, sp);
        wv = spline_W_near(d, sp);
        cv = spline_gradc_near(d, sp);
    }
    // acc + a * b, rounded twice (the reference's acc += a * b)
    __device__ static inline float mad(float a, float b, float acc) { return acc + a * b; }
    __device__ static inline float div(float a, float b) { return a / b; }
};

template <> struct PairMath<true> {
    static constexpr bool fast = true;
    __device__ static inline float r2(float x, float y, float z) {
        return __builtin_fmaf(x, x, __builtin_fmaf(y, y, z * z));
    }
    // d and 1/d from r2 > 0; d = 0 (and 1/d = 0) for r2 = 0, where the
    // gradient factor is 0 (coincident particles, the self pair)
    __device__ static inline void dist(float r2, float& d, float& inv_d) {
        const float rs = __builtin_amdgcn_rsqf(r2);
        const bool pos = r2 > 0.0f;
        d = pos ? r2 * rs : 0.0f;
        inv_d = pos ? rs : 0.0f;
    }
    __device__ static inline float w_of_q(float q, const Spline& sp) {
        const float q2 = q * q;
        const float inner = sp.sigma * __builtin_fmaf(q2, __builtin_fmaf(0.75f, q, -1.5f), 1.0f);
        const float t = 2.0f - q;
        const float outer = sp.sigma_025 * (t * t * t);
        return q < 1.0f ? inner : outer;
    }
    __device__ static inline float c_of_q(float q, float d, float inv_d, const Spline& sp) {
        const float inner = sp.ck_in * __builtin_fmaf(2.25f, q, -3.0f);
        const float t = 2.0f - q;
        const float outer = sp.ck_out * (t * t) * inv_d;
        const float c = q < 1.0f ? inner : outer;
        return d < 1e-12f ? 0.0f : c;
    }
    __device__ static inline float w(float r2, const Spline& sp) {
        float d, inv_d;
        dist(r2, d, inv_d);
        return w_of_q(d * sp.inv_l, sp);
    }
    __device__ static inline float c(float r2, const Spline& sp) {
        float d, inv_d;
        dist(r2, d, inv_d);
        return c_of_q(d * sp.inv_l, d, inv_d, sp);
    }
    // c of a listed pair: its factor was not 0, so r2 > 0 and d >= 1e-12,
    // and the tests for both are left out (same bits)
    __device__ static inline float c_listed(float r2, const Spline& sp) {
        const float inv_d = __builtin_amdgcn_rsqf(r2);
        const float d = r2 * inv_d;
        const float q = d * sp.inv_l;
        const float inner = sp.ck_in * __builtin_fmaf(2.25f, q, -3.0f);
        const float t = 2.0f - q;
        const float outer = sp.ck_out * (t * t) * inv_d;
        return q < 1.0f ? inner : outer;
    }
    __device__ static inline void wc(float r2, const Spline& sp, float& wv, float& cv) {
        float d, inv_d;
        dist(r2, d, inv_d);
        const float q = d * sp.inv_l;
        wv = w_of_q(q, sp);
        cv = c_of_q(q, d, inv_d, sp);
    }
    __device__ static inline float mad(float a, float b, float acc) {
        return __builtin_fmaf(a, b, acc);
    }
    __device__ static inline float div(float a, float b) {
        return a * __builtin_amdgcn_rcpf(b);
    }
};

// Workgroup id remap so that each XCD walks a contiguous run of tiles: the
// dispatcher hands workgroup b to XCD (b % 8), so neighbouring tiles (which
// gather the same cell-sorted particles) land in the same L2.
__device__ inline u32 xcd_tile(u32 b, u32 nblocks) {
    const u32 xcd = b % kXcds, k = b / kXcds;
    const u32 q = nblocks / kXcds, rem = nblocks % kXcds;
    return xcd < rem ? xcd * (q + 1) + k : rem * (q + 1) + (xcd - rem) * q + k;
}

// ---- wave64 reductions (shuffle based) ----
__device__ inline double wave_sum(double v) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o);
    return v;
}
__device__ inline float wave_max(float v) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v = fmaxf(v, __shfl_xor(v, o));
    return v;
}
__device__ inline u32 wave_sum_u32(u32 v) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o);
    return v;
}

// Block reductions for blockDim.x == NT (multiple of 64).  Deterministic:
// fixed lane/wave tree.  Result valid in thread 0.
template <int NT> __device__ inline double block_sum(double v, double* s) {
    v = wave_sum(v);
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    if (lane == 0) s[w] = v;
    __syncthreads();
    double r = 0.0;
    if (threadIdx.x == 0) {
        for (int i = 0; i < NT / 64; ++i) r += s[i];
    }
    __syncthreads();
    return r;
}
template <int NT> __device__ inline float block_max(float v, float* s) {
    v = wave_max(v);
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    if (lane == 0) s[w] = v;
    __syncthreads();
    float r = 0.0f;
    if (threadIdx.x == 0) {
        for (int i = 0; i < NT / 64; ++i) r = fmaxf(r, s[i]);
    }
    __syncthreads();
    return r;
}
template <int NT> __device__ inline u32 block_sum_u32(u32 v, u32* s) {
    v = wave_sum_u32(v);
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    if (lane == 0) s[w] = v;
    __syncthreads();
    u32 r = 0;
    if (threadIdx.x == 0) {
        for (int i = 0; i < NT / 64; ++i) r += s[i];
    }
    __syncthreads();
    return r;
}

}  // namespace wsim
