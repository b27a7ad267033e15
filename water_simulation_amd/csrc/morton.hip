// morton.hip -- SpatialAccel::build: 30-bit Morton codes on a scene box,
// their stable sort, and the linear BVH over the sorted codes.
//
// Replaces core/src/spatial_accel.cu: codes 11-40, host inverse extent
// 139-148, sort on bits [0,30) 158-179, Karras internal nodes 44-100, leaf
// boxes 102-113 (radius 0, as build() uses).  The internal boxes, which the
// reference leaves for later, are filled here as well.
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

// Length of the common prefix of sorted leaves i and j in the total order
// (code, index): equal codes compare their indices, so every prefix is
// unique (Karras 2012; spatial_accel.cu:44-55).  -1 for j outside [0, n).
__device__ inline int prefix_len(const u32* __restrict__ codes, long long n, long long i,
                                 long long j) {
    if (j < 0 || j >= n) return -1;
    const u32 a = codes[i], b = codes[j];
    if (a == b) return 32 + __clz(static_cast<u32>(i ^ j));
    return __clz(a ^ b);
}

// Internal node i of the binary radix tree over n sorted codes (Karras 2012,
// as spatial_accel.cu:57-100): the direction of its range, the range's far
// end by exponential then binary search, the split by binary search.
// Children >= n - 1 are leaves (n - 1 + leaf index); the root's parent is
// ~0u.
__global__ __launch_bounds__(256) void k_karras(const u32* __restrict__ codes, u32 n,
                                                u32* __restrict__ left, u32* __restrict__ right,
                                                u32* __restrict__ parent, uint2* __restrict__ range) {
    const long long i = static_cast<long long>(blockIdx.x) * blockDim.x + threadIdx.x;
    const long long nn = n, ni = nn - 1;
    if (i >= ni) return;
    if (i == 0) parent[0] = ~0u;
    const long long d = prefix_len(codes, nn, i, i + 1) - prefix_len(codes, nn, i, i - 1) > 0 ? 1 : -1;
    const int dmin = prefix_len(codes, nn, i, i - d);
    long long lmax = 2;
    while (prefix_len(codes, nn, i, i + lmax * d) > dmin) lmax *= 2;
    long long l = 0;
    for (long long t = lmax / 2; t > 0; t /= 2)
        if (prefix_len(codes, nn, i, i + (l + t) * d) > dmin) l += t;
    const long long j = i + l * d;
    const int dnode = prefix_len(codes, nn, i, j);
    long long s = 0;
    for (long long t = (l + 1) / 2; t > 0; t = (t == 1) ? 0 : (t + 1) / 2)
        if (prefix_len(codes, nn, i, i + (s + t) * d) > dnode) s += t;
    const long long gamma = i + s * d + (d < 0 ? -1 : 0);
    const long long lo = i < j ? i : j, hi = i < j ? j : i;
    const u32 a = static_cast<u32>(lo == gamma ? ni + gamma : gamma);
    const u32 b = static_cast<u32>(hi == gamma + 1 ? ni + gamma + 1 : gamma + 1);
    left[i] = a;
    right[i] = b;
    parent[a] = static_cast<u32>(i);
    parent[b] = static_cast<u32>(i);
    range[i] = make_uint2(static_cast<u32>(lo), static_cast<u32>(hi));
}

// Leaf boxes: the point itself (spatial_accel.cu:102-113 with radius 0).
__global__ __launch_bounds__(256) void k_leaf_boxes(const float* __restrict__ xyz,
                                                    const u32* __restrict__ sorted, u32 n,
                                                    float* __restrict__ boxes) {
    const u32 i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const u64 o = 3ull * sorted[i];
    float* b = boxes + 6ull * (n - 1 + i);
    b[0] = b[3] = xyz[o];
    b[1] = b[4] = xyz[o + 1];
    b[2] = b[5] = xyz[o + 2];
}

__device__ __forceinline__ void box_union(float* c, const float* a, const float* b) {
    for (int k = 0; k < 3; ++k) {
        c[k] = fminf(a[k], b[k]);
        c[k + 3] = fmaxf(a[k + 3], b[k + 3]);
    }
}

// Internal boxes without inter-thread ordering: a segment tree over the
// sorted leaves is built level by level (one launch per level, so each level
// reads a complete one), then every internal node unions the O(log n)
// segment-tree boxes that tile its leaf range [lo, hi].  min and max are
// exact, so the result equals the bottom-up union of the children.
__global__ __launch_bounds__(256) void k_seg_level(const float* __restrict__ below, u32 m_below,
                                                   float* __restrict__ level, u32 m) {
    const u32 t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= m) return;
    const float* a = below + 12ull * t;
    float* c = level + 6ull * t;
    if (2 * t + 1 < m_below) {
        box_union(c, a, a + 6);
    } else {
        for (int k = 0; k < 6; ++k) c[k] = a[k];
    }
}

__global__ __launch_bounds__(256) void k_internal_boxes(const uint2* __restrict__ range, u32 n,
                                                        const float* __restrict__ seg,
                                                        const u64* __restrict__ seg_off,
                                                        float* __restrict__ boxes) {
    const u32 i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i + 1 >= n) return;
    const uint2 rg = range[i];
    float acc[6] = {INFINITY, INFINITY, INFINITY, -INFINITY, -INFINITY, -INFINITY};
    u32 l = rg.x, r = rg.y + 1;
    for (int k = 0; l < r; ++k, l >>= 1, r >>= 1) {
        const float* lev = seg + 6ull * seg_off[k];
        if (l & 1u) box_union(acc, acc, lev + 6ull * l++);
        if (r & 1u) box_union(acc, acc, lev + 6ull * --r);
    }
    float* c = boxes + 6ull * i;
    for (int k = 0; k < 6; ++k) c[k] = acc[k];
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
    boxes_.ensure(6 * (2 * count - 1));
    parent_.ensure(2 * count - 1);
    k_leaf_boxes<<<div_up(n, 256), 256, 0, stream>>>(positions_xyz, v_[cur_].data(), n,
                                                     boxes_.data());
    WSIM_LAUNCH_CHECK();
    if (count == 1) {
        WSIM_HIP(hipMemsetAsync(parent_.data(), 0xFF, sizeof(u32), stream));
        return;
    }
    left_.ensure(count - 1);
    right_.ensure(count - 1);
    range_.ensure(count - 1);
    k_karras<<<div_up(n - 1, 256), 256, 0, stream>>>(k_[cur_].data(), n, left_.data(), right_.data(),
                                                     parent_.data(), range_.data());
    WSIM_LAUNCH_CHECK();
    // segment tree over the leaf boxes: level 0 is the leaf part of boxes_
    std::vector<u64> off{0};
    std::vector<u32> m{n};
    while (m.back() > 1) {
        off.push_back(off.back() + m.back());
        m.push_back((m.back() + 1) / 2);
    }
    seg_.ensure(6 * (off.back() + 1));
    seg_off_.ensure(off.size());
    WSIM_HIP(hipMemcpyAsync(seg_.data(), boxes_.data() + 6ull * (n - 1), 24ull * n,
                            hipMemcpyDeviceToDevice, stream));
    for (size_t k = 1; k < m.size(); ++k) {
        k_seg_level<<<div_up(m[k], 256), 256, 0, stream>>>(seg_.data() + 6 * off[k - 1], m[k - 1],
                                                           seg_.data() + 6 * off[k], m[k]);
        WSIM_LAUNCH_CHECK();
    }
    seg_off_host_.assign(off.begin(), off.end());
    WSIM_HIP(hipMemcpyAsync(seg_off_.data(), seg_off_host_.data(), off.size() * sizeof(u64),
                            hipMemcpyHostToDevice, stream));
    k_internal_boxes<<<div_up(n - 1, 256), 256, 0, stream>>>(range_.data(), n, seg_.data(),
                                                             seg_off_.data(), boxes_.data());
    WSIM_LAUNCH_CHECK();
    // the host offsets must outlive the asynchronous copy
    WSIM_HIP(hipStreamSynchronize(stream));
}

}  // namespace wsim
