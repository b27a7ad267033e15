// pair_ops.hip -- exhaustive device check of the pair-path sqrt and quotients.
//
// For one smoothing length l, every binary32 operand the solver kernels can
// hand to pair_sqrt_fast, pair_q_fast and pair_div_fast is run through them
// and through the compiler's IEEE operations, and the results that differ
// in any bit are counted:
//   sqrt:     every x in [2^-96, 4 * far2] (callers reject r2 >= far2, and
//             beyond 4 * far2 both square roots put q well past 2);
//   d / l:    every d in [2^-48, 2l (1 + 2^-20)];
//   gradient: the factor's quotient dW/dq / (l d) for every d in
//             [1e-12, 2l (1 + 2^-20)] with q < 2 (the kernels return 0 before
//             dividing at q >= 2), its numerator formed as in
//             spline_gradc_near.
#include <cmath>
#include <cstring>

#include "common.hpp"
#include "device_math.hpp"
#include "wsim_internal.hpp"

namespace wsim {
namespace {

constexpr int kCheckThreads = 256;

__device__ __forceinline__ u32 differs(float a, float b) {
    return __float_as_uint(a) != __float_as_uint(b) ? 1u : 0u;
}

__global__ __launch_bounds__(kCheckThreads) void k_pair_ops_check(Spline sp, u32 x_lo, u32 x_hi,
                                                                  u32 d_lo, u32 d_hi, u32 c_lo,
                                                                  unsigned long long* bad) {
    const u32 stride = gridDim.x * blockDim.x;
    const u32 t = blockIdx.x * blockDim.x + threadIdx.x;
    u32 nb_s = 0, nb_q = 0, nb_c = 0;
    for (u32 b = x_lo + t; b < x_hi; b += stride) {
        const float x = __uint_as_float(b);
        nb_s += differs(pair_sqrt_fast(x), sqrtf(x));
    }
    for (u32 b = d_lo + t; b < d_hi; b += stride) {
        const float d = __uint_as_float(b);
        const float q = d / sp.l;
        nb_q += differs(pair_q_fast(d, sp.l, sp.inv_l), q);
        if (b >= c_lo && q < 2.0f) {
            const float inner = sp.sigma * (-3.0f * q + 2.25f * q * q);
            const float tq = 2.0f - q;
            const float outer = sp.sigma_m075 * tq * tq;
            const float a = q < 1.0f ? inner : outer;
            const float den = sp.l * d;
            nb_c += differs(pair_div_fast(a, den), a / den);
        }
    }
    nb_s = wave_sum_u32(nb_s);
    nb_q = wave_sum_u32(nb_q);
    nb_c = wave_sum_u32(nb_c);
    if ((threadIdx.x & (kWave - 1)) == 0) {
        if (nb_s) atomicAdd(bad + 0, static_cast<unsigned long long>(nb_s));
        if (nb_q) atomicAdd(bad + 1, static_cast<unsigned long long>(nb_q));
        if (nb_c) atomicAdd(bad + 2, static_cast<unsigned long long>(nb_c));
    }
}

u32 bits_of(float f) {
    u32 u;
    std::memcpy(&u, &f, sizeof(u));
    return u;
}

}  // namespace

PairOpsReport pair_ops_check(float l) {
    if (!(l > 0.0f) || !std::isfinite(l)) fail(kInvalidArgument, "pair_ops_check: l must be positive");
    const Spline sp = Spline::make(l);
    PairOpsReport r{};
    r.fast = sp.fast;
    const float xmax = sp.far2 < 3.0e38f / 4.0f ? 4.0f * sp.far2 : 3.0e38f;
    const float dmax = 2.0f * l * (1.0f + 0x1p-20f);
    r.x_lo = bits_of(0x1p-96f);
    r.x_hi = bits_of(xmax) + 1u;
    r.d_lo = bits_of(0x1p-48f);
    r.d_hi = bits_of(dmax) + 1u;
    r.c_lo = bits_of(1e-12f);
    if (r.x_hi < r.x_lo) r.x_hi = r.x_lo;
    if (r.d_hi < r.d_lo) r.d_hi = r.d_lo;
    DeviceBuffer<unsigned long long> bad(3);
    WSIM_HIP(hipMemset(bad.data(), 0, 3 * sizeof(unsigned long long)));
    k_pair_ops_check<<<8192, kCheckThreads, 0, nullptr>>>(sp, r.x_lo, r.x_hi, r.d_lo, r.d_hi, r.c_lo,
                                                          bad.data());
    WSIM_LAUNCH_CHECK();
    unsigned long long h[3];
    WSIM_HIP(hipMemcpy(h, bad.data(), sizeof(h), hipMemcpyDeviceToHost));
    r.sqrt_mismatches = h[0];
    r.quotient_mismatches = h[1];
    r.gradient_mismatches = h[2];
    return r;
}

}  // namespace wsim
