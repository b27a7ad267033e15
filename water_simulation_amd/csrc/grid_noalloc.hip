// grid_noalloc.hip -- the cell-grid build (keys, stable sort, cell_start)
// into caller arrays, with every scratch buffer in a fixed device array
// instead of an allocation: for callers that cannot allocate device memory
// (code under HIP graph capture, tight serving loops).
//
// The kernels are CellGrid::build's (grid.hip, radix_sort.hip); only where
// the sort ping-pongs and keeps its digit histograms differs.  The scratch
// array is one per device, allocated with the code object and held until the
// process ends; calls that overlap on one device race on it.
#include <mutex>

#include "radix_sort.hpp"
#include "wsim_internal.hpp"

namespace wsim {
namespace {

// Largest particle count of one call.
constexpr u32 kNoAllocMaxCount = 1u << 21;  // 2,097,152
// Ping-pong keys and values of the sort (2 x count) plus its histograms.
constexpr size_t kHistWords = 256u * ((kNoAllocMaxCount + 1023u) / 1024u) + 256u;
constexpr size_t kScratchWords = 2u * kNoAllocMaxCount + kHistWords;  // 18 MiB

static __device__ __align__(256) u32 scratch[kScratchWords];

constexpr int kMaxDevices = 64;
std::mutex g_scratch_mutex;
u32* g_scratch_ptr[kMaxDevices] = {};

// Device address of this device's copy of `scratch`, looked up once per
// device with that device current.  In host code the array's name is only a
// stand-in for the symbol: every kernel argument uses this address.
u32* scratch_on_current_device() {
    int dev = 0;
    WSIM_HIP(hipGetDevice(&dev));
    if (dev < 0 || dev >= kMaxDevices) fail(kRuntimeError, "cell sort (no allocation): device id");
    std::lock_guard<std::mutex> lock(g_scratch_mutex);
    if (!g_scratch_ptr[dev]) {
        void* p = nullptr;
        const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !p)
            fail(kDeviceError, std::string("cell sort (no allocation): hipGetSymbolAddress: ") +
                                   hipGetErrorString(e));
        g_scratch_ptr[dev] = static_cast<u32*>(p);
    }
    return g_scratch_ptr[dev];
}

__global__ __launch_bounds__(256) void k_noalloc_keys(const float* __restrict__ xyz, u32 n,
                                                      GridGeom g, u32* __restrict__ keys,
                                                      u32* __restrict__ vals) {
    const u32 i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const u64 o = 3ull * i;
    keys[i] = g.key_of(xyz[o], xyz[o + 1], xyz[o + 2]);  // cell_grid.cu:24-30
    vals[i] = i;
}

__global__ __launch_bounds__(256) void k_noalloc_copy(const u32* __restrict__ ks,
                                                      const u32* __restrict__ vs, u32 n,
                                                      u32* __restrict__ kd, u32* __restrict__ vd) {
    const u32 i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    kd[i] = ks[i];
    vd[i] = vs[i];
}

}  // namespace

void cell_sort_noalloc(const float* xyz, size_t count, const float origin[3],
                       const int32_t dims[3], float cell_length, u32* sorted_keys,
                       u32* sorted_indices, u32* cell_start, hipStream_t stream) {
    if (dims[0] <= 0 || dims[1] <= 0 || dims[2] <= 0)
        fail(kInvalidArgument, "cell sort (no allocation): cells_per_axis must be positive");
    if (!(cell_length > 0.0f))
        fail(kInvalidArgument, "cell sort (no allocation): cell_length must be positive");
    const u64 C = static_cast<u64>(dims[0]) * static_cast<u64>(dims[1]) * static_cast<u64>(dims[2]);
    if (C + 1 >= 0xFFFFF000ull) fail(kInvalidArgument, "cell sort (no allocation): too many cells");
    if (count > kNoAllocMaxCount)
        fail(kInvalidArgument, "cell sort (no allocation): " + std::to_string(count) +
                                   " particles exceed the maximum of " +
                                   std::to_string(kNoAllocMaxCount));
    if (!cell_start) fail(kInvalidArgument, "cell sort (no allocation): cell_start is null");
    if (count && (!xyz || !sorted_keys || !sorted_indices))
        fail(kInvalidArgument, "cell sort (no allocation): null positions or outputs");
    GridGeom g{};
    g.ox = origin[0];
    g.oy = origin[1];
    g.oz = origin[2];
    g.inv = 1.0f / cell_length;  // cell_grid.cu:81, on the host
    g.dx = dims[0];
    g.dy = dims[1];
    g.dz = dims[2];
    g.C = static_cast<u32>(C);
    u32* scr = scratch_on_current_device();
    u32* k1 = scr;
    u32* v1 = scr + kNoAllocMaxCount;
    u32* hist = scr + 2u * static_cast<size_t>(kNoAllocMaxCount);
    const u32 n = static_cast<u32>(count);
    if (n) {
        k_noalloc_keys<<<div_up(n, 256u), 256, 0, stream>>>(xyz, n, g, sorted_keys,
                                                             sorted_indices);
        WSIM_LAUNCH_CHECK();
        const int where = radix_sort_pairs(sorted_keys, sorted_indices, k1, v1, n,
                                           key_bits(g.C), hist, stream);
        if (where == 1) {
            k_noalloc_copy<<<div_up(n, 256u), 256, 0, stream>>>(k1, v1, n, sorted_keys,
                                                                sorted_indices);
            WSIM_LAUNCH_CHECK();
        }
    }
    // every entry of cell_start is written (an empty set gives all zeros)
    build_cell_start(sorted_keys, n, g.C + 1, cell_start, stream, 16);
}

}  // namespace wsim
