// morton.hip -- 30-bit Morton codes on a scene box and their stable sort.
//
// Replaces the Morton half of SpatialAccel::build (core/src/spatial_accel.cu:
// 11-40 codes, 139-148 host inverse extent, 158-179 sort on bits [0,30)).
// The Karras hierarchy built on top of the codes (spatial_accel.cu:44-107)
// is not on the solver path and is not built here.
#include "radix_sort.hpp"
#include "wsim_internal.hpp"

namespace wsim {
namespace {

// Bit b of a 10-bit value moves to bit 3b.
__device__ inline u32 spread10(u32 v) {
    v &= 0x3FFu;
    v = (v | (v << 16)) & 0x030000FFu;
    v = (v | (v << 8)) & 0x0300F00Fu;
    v = (v | (v << 4)) & 0x030C30C3u;
    v = (v | (v << 2)) & 0x09249249u;
    return v;
}

// Quantise one normalised coordinate to [0, 1023] (spatial_accel.cu:21-26).
__device__ inline u32 quant10(float p, float lo, float inv) {
    const float t = (p - lo) * inv;
    return static_cast<u32>(fminf(fmaxf(t * 1024.0f, 0.0f), 1023.0f));
}

__global__ __launch_bounds__(256) void k_morton(const float* __restrict__ xyz, u32 n, float mx,
                                                float my, float mz, float ix, float iy, float iz,
                                                u32* __restrict__ codes, u32* __restrict__ vals) {
    const u32 i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const u64 o = 3ull * i;
    const u32 qx = quant10(xyz[o], mx, ix);
    const u32 qy = quant10(xyz[o + 1], my, iy);
    const u32 qz = quant10(xyz[o + 2], mz, iz);
    codes[i] = (spread10(qx) << 2) | (spread10(qy) << 1) | spread10(qz);
    vals[i] = i;
}

}  // namespace

void SpatialAccel::build(const float* positions_xyz, size_t count, const float bmin[3],
                         const float bmax[3], hipStream_t stream) {
    leaf_count_ = count;
    if (count == 0) return;
    if (count > 0xFFFFF000ull) fail(kInvalidArgument, "SpatialAccel: too many particles");
    k_[0].ensure(count);
    k_[1].ensure(count);
    v_[0].ensure(count);
    v_[1].ensure(count);
    hist_.ensure(radix_hist_elems(count));
    float inv[3];
    for (int a = 0; a < 3; ++a) {
        const float e = bmax[a] - bmin[a];
        inv[a] = e > 0 ? 1.0f / e : 0.0f;  // spatial_accel.cu:144-148
    }
    const u32 n = static_cast<u32>(count);
    k_morton<<<div_up(n, 256), 256, 0, stream>>>(positions_xyz, n, bmin[0], bmin[1], bmin[2], inv[0],
                                                 inv[1], inv[2], k_[0].data(), v_[0].data());
    WSIM_LAUNCH_CHECK();
    cur_ = radix_sort_pairs(k_[0].data(), v_[0].data(), k_[1].data(), v_[1].data(), count, 30,
                            hist_.data(), stream);
}

}  // namespace wsim
