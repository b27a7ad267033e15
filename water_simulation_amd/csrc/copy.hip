// copy.hip -- plain streaming device copy (16-byte vectors, grid-stride).
//
// Used as the known-size reference when HBM byte counters (FETCH_SIZE,
// WRITE_SIZE) are converted into bytes for the solver kernels.
#include "common.hpp"

namespace wsim {
namespace {

__global__ __launch_bounds__(256) void k_copy16(const uint4* __restrict__ src,
                                                uint4* __restrict__ dst, u64 n16) {
    const u64 stride = static_cast<u64>(gridDim.x) * blockDim.x;
    for (u64 i = static_cast<u64>(blockIdx.x) * blockDim.x + threadIdx.x; i < n16; i += stride)
        dst[i] = src[i];
}

__global__ void k_copy1(const unsigned char* __restrict__ src, unsigned char* __restrict__ dst,
                        u64 n) {
    const u64 i = static_cast<u64>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i < n) dst[i] = src[i];
}

}  // namespace

void device_copy(void* dst, const void* src, size_t bytes, hipStream_t stream) {
    const u64 n16 = bytes / 16;
    if (n16) {
        const u64 want = (n16 + 255) / 256;
        const u32 blocks = static_cast<u32>(want < 65536 ? want : 65536);
        k_copy16<<<blocks, 256, 0, stream>>>(static_cast<const uint4*>(src),
                                             static_cast<uint4*>(dst), n16);
        WSIM_LAUNCH_CHECK();
    }
    const u64 tail = bytes - n16 * 16;
    if (tail) {
        k_copy1<<<1, 256, 0, stream>>>(static_cast<const unsigned char*>(src) + n16 * 16,
                                       static_cast<unsigned char*>(dst) + n16 * 16, tail);
        WSIM_LAUNCH_CHECK();
    }
}

}  // namespace wsim
