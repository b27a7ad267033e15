// surface.hip -- fluid surface mesh from particle positions on the GPU.
//
// Replaces the splat method of tools/particles_to_mesh.py
// (reconstruct_splat, :77-138 with load_frame :46-58), the consumer of the
// frame format downstream of sim_cli:
//   1. bounds of the points (optionally with y and z swapped, the
//      reference's sim Y-up -> Blender Z-up), padded by 4 r;
//   2. splat: +1 in the voxel (p - lo) / extent * res, truncated and
//      clipped, res = max(ceil(extent / (factor r)), 8) per axis;
//   3. Gaussian blur of the voxel counts, the three 1-D passes of
//      scipy.ndimage.gaussian_filter (truncate 4, mode 'reflect', binary64
//      sums in its order for symmetric weights, binary32 between passes);
//   4. iso = max * iso_factor in binary32 (numpy 2 scalar rules);
//   5. marching cubes with shared vertices: one vertex per grid edge whose
//      end values straddle the iso level (linear interpolation), triangles
//      from a table built from the cube faces (mc_table).
// Before the bounds, the reference's k-nearest-neighbour position smoothing
// (smooth_positions_knn :61-74) runs smooth_pre_iters times; after marching
// cubes, smooth_post_iters passes of the Laplacian mesh smoothing it asks of
// Open3D (:138-139, filter_smooth_laplacian, lambda 0.5).
#include <algorithm>
#include <cmath>
#include <cstring>
#include <type_traits>
#include <vector>

#include "common.hpp"
#include "device_math.hpp"
#include "radix_sort.hpp"
#include "wsim_internal.hpp"

namespace wsim {
namespace {

constexpr int kMeshThreads = 256;
constexpr int kMaxTri = 8;  // triangles per cube configuration (the table has at most 5)

// ---- marching-cubes table ---------------------------------------------------
// Corner c of a cube sits at offset (c & 1, c >> 1 & 1, c >> 2 & 1) along the
// grid axes (0, 1, 2).  Edge a * 4 + b runs along axis a from the corner
// with bit a clear; b holds that corner's other two bits (lower axis first).
// A corner is inside when its value exceeds the iso level.  On every face,
// each run of consecutive inside corners (going round the face; diagonal
// inside corners are separate runs) is bounded by two crossing edges, which
// a surface segment joins.  The segments close into loops around the
// edge-connected inside corners, and each loop becomes a triangle fan.
// Neighbouring cubes build the same segments on their shared face, so the
// surface is closed; segments are oriented so that triangle normals point
// from the inside corners outwards.
struct McTable {
    uint8_t ntri[256];
    uint8_t tri[256][kMaxTri * 3];
};

int edge_between(int c0, int c1) {
    const int a = __builtin_ctz(static_cast<unsigned>(c0 ^ c1));
    const int base = c0 & ~(1 << a);
    int b = 0, k = 0;
    for (int ax = 0; ax < 3; ++ax) {
        if (ax == a) continue;
        b |= ((base >> ax) & 1) << k;
        ++k;
    }
    return a * 4 + b;
}

McTable build_mc_table() {
    McTable t{};
    for (int cfg = 0; cfg < 256; ++cfg) {
        auto inside = [&](int c) { return ((cfg >> c) & 1) != 0; };
        int next[12];
        for (int& e : next) e = -1;
        for (int a = 0; a < 3; ++a) {
            const int u = (a + 1) % 3, v = (a + 2) % 3;  // (u, v, a) right-handed
            for (int s = 0; s < 2; ++s) {
                int cyc[4];
                const int du[4] = {0, 1, 1, 0}, dv[4] = {0, 0, 1, 1};
                for (int m = 0; m < 4; ++m) cyc[m] = (s << a) | (du[m] << u) | (dv[m] << v);
                if (s == 0) std::swap(cyc[1], cyc[3]);  // counter-clockwise seen from outside
                for (int m = 0; m < 4; ++m) {
                    if (!inside(cyc[m]) || inside(cyc[(m + 3) % 4])) continue;  // run start
                    int k = m;
                    while (inside(cyc[(k + 1) % 4])) k = (k + 1) % 4;
                    const int e_in = edge_between(cyc[(m + 3) % 4], cyc[m]);
                    const int e_out = edge_between(cyc[k], cyc[(k + 1) % 4]);
                    next[e_in] = e_out;
                }
            }
        }
        bool seen[12] = {};
        int nt = 0;
        for (int e0 = 0; e0 < 12; ++e0) {
            if (next[e0] < 0 || seen[e0]) continue;
            int loop[12], len = 0;
            for (int e = e0; !seen[e]; e = next[e]) {
                seen[e] = true;
                loop[len++] = e;
            }
            for (int k = 1; k + 1 < len; ++k) {
                if (nt >= kMaxTri) fail(kRuntimeError, "marching-cubes table overflow");
                t.tri[cfg][3 * nt + 0] = static_cast<uint8_t>(loop[0]);
                t.tri[cfg][3 * nt + 1] = static_cast<uint8_t>(loop[k]);
                t.tri[cfg][3 * nt + 2] = static_cast<uint8_t>(loop[k + 1]);
                ++nt;
            }
        }
        t.ntri[cfg] = static_cast<uint8_t>(nt);
    }
    return t;
}

const McTable& mc_table() {
    static const McTable t = build_mc_table();
    return t;
}

// ---- kernels ----------------------------------------------------------------
struct MeshGeom {
    int r0, r1, r2;
    double lo0, lo1, lo2;
    double ext0, ext1, ext2;
    double sp0, sp1, sp2;
};

// binary64 points as load_frame makes them (:46-58): astype(float64), then
// (x, y, z) -> (x, z, y) when swap is set
__global__ __launch_bounds__(kMeshThreads) void k_mesh_points(const float* __restrict__ xyz, u64 n,
                                                              int swap, double* __restrict__ P) {
    const u64 i = static_cast<u64>(blockIdx.x) * kMeshThreads + threadIdx.x;
    if (i >= n) return;
    const double x = xyz[3 * i], y = xyz[3 * i + 1], z = xyz[3 * i + 2];
    P[3 * i] = x;
    P[3 * i + 1] = swap ? z : y;
    P[3 * i + 2] = swap ? y : z;
}

// Per-workgroup partials: min x, y, z, max x, y, z and the number of
// non-finite coordinates (fmin/fmax skip NaN, so the count is what lets the
// host refuse a point set holding one before any k-NN search runs).
constexpr int kBoundsPart = 7;
__global__ __launch_bounds__(kMeshThreads) void k_mesh_bounds(const double* __restrict__ P, u64 n,
                                                              double* __restrict__ part) {
    __shared__ double s[kBoundsPart][kMeshThreads];
    double mn[3] = {INFINITY, INFINITY, INFINITY}, mx[3] = {-INFINITY, -INFINITY, -INFINITY};
    double bad = 0.0;
    const u64 stride = static_cast<u64>(gridDim.x) * kMeshThreads;
    for (u64 i = static_cast<u64>(blockIdx.x) * kMeshThreads + threadIdx.x; i < n; i += stride) {
        for (int a = 0; a < 3; ++a) {
            const double c = P[3 * i + a];
            mn[a] = fmin(mn[a], c);
            mx[a] = fmax(mx[a], c);
            bad += isfinite(c) ? 0.0 : 1.0;
        }
    }
    for (int a = 0; a < 3; ++a) {
        s[a][threadIdx.x] = mn[a];
        s[3 + a][threadIdx.x] = mx[a];
    }
    s[6][threadIdx.x] = bad;
    __syncthreads();
    if (threadIdx.x < kBoundsPart) {
        double r = s[threadIdx.x][0];
        for (int k = 1; k < kMeshThreads; ++k) {
            const double o = s[threadIdx.x][k];
            r = threadIdx.x < 3 ? fmin(r, o) : (threadIdx.x < 6 ? fmax(r, o) : r + o);
        }
        part[kBoundsPart * blockIdx.x + threadIdx.x] = r;
    }
}

// numpy's ((p - lo) / extent * res).astype(int), then clip to [0, res - 1]
__device__ __forceinline__ int voxel_of(double p, double lo, double ext, int res) {
    const double f = (p - lo) / ext * static_cast<double>(res);
    long long k = f == f ? static_cast<long long>(f) : 0;
    if (k < 0) k = 0;
    if (k > res - 1) k = res - 1;
    return static_cast<int>(k);
}

__global__ __launch_bounds__(kMeshThreads) void k_mesh_splat(const double* __restrict__ P, u64 n,
                                                             MeshGeom G, u32* __restrict__ count) {
    const u64 i = static_cast<u64>(blockIdx.x) * kMeshThreads + threadIdx.x;
    if (i >= n) return;
    const int a = voxel_of(P[3 * i], G.lo0, G.ext0, G.r0);
    const int b = voxel_of(P[3 * i + 1], G.lo1, G.ext1, G.r1);
    const int c = voxel_of(P[3 * i + 2], G.lo2, G.ext2, G.r2);
    atomicAdd(count + (static_cast<u64>(a) * G.r1 + b) * G.r2 + c, 1u);
}

// ---- k-nearest-neighbour position smoothing (smooth_positions_knn, :61-74)
// Each point becomes the mean of its k nearest points (itself included), in
// binary64, summed in ascending (distance^2, index) order and divided by k,
// as numpy's mean over cKDTree.query's rows.  Exact kNN on a uniform grid:
// rings of cells around the home cell are scanned until the k-th distance is
// below the ring's reach (points outside rings 0..R lie at least R cells away).

struct KnnGrid {
    double lo0, lo1, lo2, cell;
    int d0, d1, d2;
};

__device__ __forceinline__ int knn_cell(double p, double lo, double cell, int d) {
    const double f = (p - lo) / cell;
    int k = f >= 0.0 ? static_cast<int>(f) : 0;
    return k > d - 1 ? d - 1 : k;
}

__global__ __launch_bounds__(kMeshThreads) void k_knn_keys(const double* __restrict__ P, u32 n,
                                                           KnnGrid g, u32* __restrict__ keys,
                                                           u32* __restrict__ vals) {
    const u32 i = blockIdx.x * kMeshThreads + threadIdx.x;
    if (i >= n) return;
    const int a = knn_cell(P[3ull * i], g.lo0, g.cell, g.d0);
    const int b = knn_cell(P[3ull * i + 1], g.lo1, g.cell, g.d1);
    const int c = knn_cell(P[3ull * i + 2], g.lo2, g.cell, g.d2);
    keys[i] = (static_cast<u32>(a) * g.d1 + b) * g.d2 + c;
    vals[i] = i;
}

// Points in cell order, so that a cell's candidates are read contiguously.
__global__ __launch_bounds__(kMeshThreads) void k_knn_gather(const double* __restrict__ P, u32 n,
                                                             const u32* __restrict__ sorted,
                                                             double* __restrict__ Ps) {
    const u32 s = blockIdx.x * kMeshThreads + threadIdx.x;
    if (s >= n) return;
    const u32 j = sorted[s];
    Ps[3ull * s] = P[3ull * j];
    Ps[3ull * s + 1] = P[3ull * j + 1];
    Ps[3ull * s + 2] = P[3ull * j + 2];
}

// Candidates s in [lo, hi) of the cell order into the ascending (d^2, id)
// list of the K nearest so far, the next candidate loaded while the current
// one is compared.
template <int K>
__device__ __forceinline__ void knn_scan(u32 lo, u32 hi, double px, double py, double pz,
                                         const double* __restrict__ Ps,
                                         const u32* __restrict__ sorted, double (&bd)[K],
                                         u32 (&bi)[K]) {
    if (lo >= hi) return;
    double nx = Ps[3ull * lo], ny = Ps[3ull * lo + 1], nz = Ps[3ull * lo + 2];
    u32 nj = sorted[lo];
    for (u32 s = lo; s < hi; ++s) {
        const double dx = nx - px, dy = ny - py, dz = nz - pz;
        u32 id = nj;
        if (s + 1 < hi) {
            nx = Ps[3ull * s + 3], ny = Ps[3ull * s + 4], nz = Ps[3ull * s + 5];
            nj = sorted[s + 1];
        }
        double d = (dx * dx + dy * dy) + dz * dz;
        // most candidates do not beat the current k-th: one test
        if (!(d < bd[K - 1] || (d == bd[K - 1] && id < bi[K - 1]))) continue;
        // insert (d, id) into the ascending list (compile-time indices)
#pragma unroll
        for (int k = 0; k < K; ++k) {
            const bool less = d < bd[k] || (d == bd[k] && id < bi[k]);
            const double td = bd[k];
            const u32 ti = bi[k];
            bd[k] = less ? d : td;
            bi[k] = less ? id : ti;
            d = less ? td : d;
            id = less ? ti : id;
        }
    }
}

// Rings a lane walks on its own; points whose k-th neighbour lies further
// out go to a list that k_knn_far searches a wave per point.
constexpr int kKnnLaneRings = 3;

template <int K>
__global__ __launch_bounds__(kMeshThreads) void k_knn_smooth(const double* __restrict__ P, u32 n,
                                                             KnnGrid g,
                                                             const u32* __restrict__ sorted,
                                                             const double* __restrict__ Ps,
                                                             const u32* __restrict__ cs,
                                                             double* __restrict__ out,
                                                             u32* __restrict__ far_count,
                                                             u32* __restrict__ far) {
    // one thread per point in cell order: a wave's points share cells
    const u32 q = blockIdx.x * kMeshThreads + threadIdx.x;
    if (q >= n) return;
    const u32 i = sorted[q];
    const double px = Ps[3ull * q], py = Ps[3ull * q + 1], pz = Ps[3ull * q + 2];
    const int h0 = knn_cell(px, g.lo0, g.cell, g.d0);
    const int h1 = knn_cell(py, g.lo1, g.cell, g.d1);
    const int h2 = knn_cell(pz, g.lo2, g.cell, g.d2);
    double bd[K];
    u32 bi[K];
#pragma unroll
    for (int k = 0; k < K; ++k) {
        bd[k] = INFINITY;
        bi[k] = 0xFFFFFFFFu;
    }
    auto scan = [&](u32 lo, u32 hi) { knn_scan<K>(lo, hi, px, py, pz, Ps, sorted, bd, bi); };
    const int rmax = max(g.d0, max(g.d1, g.d2));
    const int rlim = min(rmax, kKnnLaneRings);
    bool done = false;
    for (int R = 0; R <= rlim; ++R) {
        const int c0 = max(h2 - R, 0), c1 = min(h2 + R, g.d2 - 1);
        for (int a = max(h0 - R, 0); a <= min(h0 + R, g.d0 - 1); ++a) {
            for (int b = max(h1 - R, 0); b <= min(h1 + R, g.d1 - 1); ++b) {
                const u32 row = (static_cast<u32>(a) * g.d1 + b) * g.d2;
                if (R == 0 || a == h0 - R || a == h0 + R || b == h1 - R || b == h1 + R) {
                    // a face of the shell: the whole row, one contiguous range
                    scan(cs[row + c0], cs[row + c1 + 1]);
                } else {  // inside the shell's faces: its two ends in c
                    if (h2 - R >= 0) scan(cs[row + h2 - R], cs[row + h2 - R + 1]);
                    if (h2 + R < g.d2) scan(cs[row + h2 + R], cs[row + h2 + R + 1]);
                }
            }
        }
        const double reach = static_cast<double>(R) * g.cell * (1.0 - 1e-9);
        if (bi[K - 1] != 0xFFFFFFFFu && bd[K - 1] < reach * reach) {
            done = true;
            break;
        }
    }
    if (!done && rlim < rmax) {  // rings beyond rlim hold cells: leave it to k_knn_far
        far[atomicAdd(far_count, 1u)] = q;
        return;
    }
    double sx = 0.0, sy = 0.0, sz = 0.0;
#pragma unroll
    for (int k = 0; k < K; ++k) {
        const u32 j = bi[k];
        if (j == 0xFFFFFFFFu) continue;  // never with finite input (the host refuses others)
        sx = sx + P[3ull * j];
        sy = sy + P[3ull * j + 1];
        sz = sz + P[3ull * j + 2];
    }
    out[3ull * i] = sx / static_cast<double>(K);
    out[3ull * i + 1] = sy / static_cast<double>(K);
    out[3ull * i + 2] = sz / static_cast<double>(K);
}

// The listed points, one wave each: the lanes take the (a, b) rows of every
// shell in turn, each keeping the K nearest of its own rows; the shells stop
// once K candidates of the union lie inside the reach (then the union's K-th
// does), and K rounds of the wave minimum by (d^2, id) give the K nearest in
// ascending order for the sum.
template <int K>
__global__ __launch_bounds__(kMeshThreads) void k_knn_far(const double* __restrict__ P,
                                                          KnnGrid g,
                                                          const u32* __restrict__ sorted,
                                                          const double* __restrict__ Ps,
                                                          const u32* __restrict__ cs,
                                                          const u32* __restrict__ far_count,
                                                          const u32* __restrict__ far,
                                                          double* __restrict__ out) {
    const int lane = threadIdx.x & 63;
    const u32 wave = (blockIdx.x * kMeshThreads + threadIdx.x) >> 6;
    const u32 nwaves = (gridDim.x * kMeshThreads) >> 6;
    const u32 nfar = *far_count;
    const int rmax = max(g.d0, max(g.d1, g.d2));
    for (u32 f = wave; f < nfar; f += nwaves) {
        const u32 q = far[f];
        const double px = Ps[3ull * q], py = Ps[3ull * q + 1], pz = Ps[3ull * q + 2];
        const int h0 = knn_cell(px, g.lo0, g.cell, g.d0);
        const int h1 = knn_cell(py, g.lo1, g.cell, g.d1);
        const int h2 = knn_cell(pz, g.lo2, g.cell, g.d2);
        double bd[K];
        u32 bi[K];
#pragma unroll
        for (int k = 0; k < K; ++k) {
            bd[k] = INFINITY;
            bi[k] = 0xFFFFFFFFu;
        }
        for (int R = 0; R <= rmax; ++R) {
            // the shell's (a, b) rows inside the grid: na * nb <= d0 * d1 <= 2^26
            const int a0 = max(h0 - R, 0), na = min(h0 + R, g.d0 - 1) - a0 + 1;
            const int b0 = max(h1 - R, 0), nb = min(h1 + R, g.d1 - 1) - b0 + 1;
            const int c0 = max(h2 - R, 0), c1 = min(h2 + R, g.d2 - 1);
            for (int t = lane; t < na * nb; t += 64) {
                const int a = a0 + t / nb, b = b0 + t % nb;
                const u32 row = (static_cast<u32>(a) * g.d1 + b) * g.d2;
                if (R == 0 || a == h0 - R || a == h0 + R || b == h1 - R || b == h1 + R) {
                    knn_scan<K>(cs[row + c0], cs[row + c1 + 1], px, py, pz, Ps, sorted, bd, bi);
                } else {
                    if (h2 - R >= 0)
                        knn_scan<K>(cs[row + h2 - R], cs[row + h2 - R + 1], px, py, pz, Ps,
                                    sorted, bd, bi);
                    if (h2 + R < g.d2)
                        knn_scan<K>(cs[row + h2 + R], cs[row + h2 + R + 1], px, py, pz, Ps,
                                    sorted, bd, bi);
                }
            }
            const double reach = static_cast<double>(R) * g.cell * (1.0 - 1e-9);
            u32 inside = 0;
#pragma unroll
            for (int k = 0; k < K; ++k) inside += bd[k] < reach * reach ? 1u : 0u;
            for (int o = 32; o > 0; o >>= 1) inside += __shfl_xor(inside, o, 64);
            if (inside >= static_cast<u32>(K)) break;  // the same on every lane
        }
        double sx = 0.0, sy = 0.0, sz = 0.0;
        for (int r = 0; r < K; ++r) {
            double d = bd[0];
            u32 id = bi[0];
            for (int o = 32; o > 0; o >>= 1) {
                const double od = __shfl_xor(d, o, 64);
                const u32 oi = __shfl_xor(id, o, 64);
                if (od < d || (od == d && oi < id)) {
                    d = od;
                    id = oi;
                }
            }
            // every candidate sits in one lane's list: its owner pops it
            const bool mine = bi[0] == id;
#pragma unroll
            for (int k = 0; k + 1 < K; ++k) {
                bd[k] = mine ? bd[k + 1] : bd[k];
                bi[k] = mine ? bi[k + 1] : bi[k];
            }
            bd[K - 1] = mine ? INFINITY : bd[K - 1];
            bi[K - 1] = mine ? 0xFFFFFFFFu : bi[K - 1];
            if (id != 0xFFFFFFFFu) {  // always, with finite input
                sx = sx + P[3ull * id];
                sy = sy + P[3ull * id + 1];
                sz = sz + P[3ull * id + 2];
            }
        }
        if (lane == 0) {
            const u32 i = sorted[q];
            out[3ull * i] = sx / static_cast<double>(K);
            out[3ull * i + 1] = sy / static_cast<double>(K);
            out[3ull * i + 2] = sz / static_cast<double>(K);
        }
    }
}

__global__ __launch_bounds__(kMeshThreads) void k_mesh_to_float(const u32* __restrict__ count,
                                                                float* __restrict__ grid, u64 nv) {
    const u64 i = static_cast<u64>(blockIdx.x) * kMeshThreads + threadIdx.x;
    if (i < nv) grid[i] = static_cast<float>(count[i]);
}

// scipy.ndimage 'reflect' extension: (d c b a | a b c d | d c b a), repeated
__device__ __forceinline__ int reflect_index(int k, int n) {
    const int p = 2 * n;
    k %= p;
    if (k < 0) k += p;
    return k >= n ? p - 1 - k : k;
}

// One 1-D pass along `axis` with 2R+1 symmetric weights w (w[R] the centre):
// acc = x[i] w[R] + sum_{j=1..R} (x[i-j] + x[i+j]) w[R+j] in binary64.
__global__ __launch_bounds__(kMeshThreads) void k_mesh_blur(const float* __restrict__ in,
                                                            float* __restrict__ out, int r0, int r1,
                                                            int r2, int axis,
                                                            const double* __restrict__ w, int R) {
    const u64 nv = static_cast<u64>(r0) * r1 * r2;
    const u64 i = static_cast<u64>(blockIdx.x) * kMeshThreads + threadIdx.x;
    if (i >= nv) return;
    const int c = static_cast<int>(i % r2);
    const int b = static_cast<int>((i / r2) % r1);
    const int a = static_cast<int>(i / (static_cast<u64>(r2) * r1));
    int pos, len;
    u64 step;
    if (axis == 0) {
        pos = a, len = r0, step = static_cast<u64>(r1) * r2;
    } else if (axis == 1) {
        pos = b, len = r1, step = static_cast<u64>(r2);
    } else {
        pos = c, len = r2, step = 1;
    }
    const u64 base = i - static_cast<u64>(pos) * step;
    double acc = static_cast<double>(in[i]) * w[R];
    if (pos >= R && pos + R < len) {  // no reflection: consecutive taps
        for (int j = 1; j <= R; ++j) {
            const double lo = in[i - static_cast<u64>(j) * step];
            const double hi = in[i + static_cast<u64>(j) * step];
            acc += (lo + hi) * w[R + j];
        }
    } else {
        for (int j = 1; j <= R; ++j) {
            const double lo = in[base + static_cast<u64>(reflect_index(pos - j, len)) * step];
            const double hi = in[base + static_cast<u64>(reflect_index(pos + j, len)) * step];
            acc += (lo + hi) * w[R + j];
        }
    }
    out[i] = static_cast<float>(acc);
}

// Tiled blur passes.  A strided pass (axis 0 or 1) stages, per workgroup,
// 64 consecutive c values x (kBlurTile + 2R) positions along the axis in LDS,
// with the 'reflect' ends resolved while loading, so every input is read from
// HBM about (kBlurTile + 2R) / kBlurTile times instead of 2R + 1 times; each
// of the 4 wave rows then forms kBlurTile / 4 outputs, taps outermost so each
// weight is loaded once.  Sums and their order are k_mesh_blur's.
constexpr int kBlurTile = 64;
constexpr int kBlurMaxR = 64;  // LDS (kBlurTile + 2R) x 64 floats, 48 KiB at most

__device__ __forceinline__ float blur_in(const float* p, u64 i) { return p[i]; }
__device__ __forceinline__ float blur_in(const u32* p, u64 i) { return static_cast<float>(p[i]); }

// position k of an axis of length n, 'reflect' beyond the ends
__device__ __forceinline__ int reflect_fast(int k, int n) {
    return k >= 0 && k < n ? k : reflect_index(k, n);
}

template <typename In>
__global__ __launch_bounds__(256) void k_blur_strided(const In* __restrict__ in,
                                                      float* __restrict__ out, int r0, int r1,
                                                      int r2, int axis,
                                                      const double* __restrict__ w, int R) {
    extern __shared__ float sb[];  // [(kBlurTile + 2R) * 64]
    const int lx = threadIdx.x & 63, ly = threadIdx.x >> 6;
    const int c = blockIdx.x * 64 + lx;
    const int p0 = blockIdx.y * kBlurTile;
    const int len = axis == 0 ? r0 : r1;
    const u64 step = axis == 0 ? static_cast<u64>(r1) * r2 : static_cast<u64>(r2);
    // position 0 of this column: (0, b, c) for axis 0, (a, 0, c) for axis 1
    const u64 base = axis == 0 ? static_cast<u64>(blockIdx.z) * r2 + c
                               : static_cast<u64>(blockIdx.z) * r1 * r2 + c;
    const bool cin = c < r2;
    const int rows = kBlurTile + 2 * R;
    for (int k = ly; k < rows; k += 4)
        sb[k * 64 + lx] =
            cin ? blur_in(in, base + static_cast<u64>(reflect_fast(p0 - R + k, len)) * step)
                : 0.0f;
    __syncthreads();
    constexpr int M = kBlurTile / 4;
    double acc[M];
    const double w0 = w[R];
#pragma unroll
    for (int m = 0; m < M; ++m) acc[m] = static_cast<double>(sb[(ly + 4 * m + R) * 64 + lx]) * w0;
    for (int j = 1; j <= R; ++j) {
        const double wj = w[R + j];
#pragma unroll
        for (int m = 0; m < M; ++m) {
            const int k = ly + 4 * m + R;
            acc[m] += (static_cast<double>(sb[(k - j) * 64 + lx]) +
                       static_cast<double>(sb[(k + j) * 64 + lx])) * wj;
        }
    }
    if (!cin) return;
#pragma unroll
    for (int m = 0; m < M; ++m) {
        const int p = p0 + ly + 4 * m;
        if (p < len) out[base + static_cast<u64>(p) * step] = static_cast<float>(acc[m]);
    }
}

// The contiguous pass (axis 2): a workgroup takes kBlurRows (a, b) rows in
// turn, 256 consecutive c plus 2R halo in LDS at a time, one output per
// thread, and writes the maximum of its outputs to part[] (blurred densities
// are >= 0, so 0 is neutral).
// RT > 0 fixes the tap radius at compile time (RT == R); RT == 0 takes R.
constexpr int kBlurRows = 4;  // (a, b) rows per workgroup of the contiguous pass

template <int RT>
__global__ __launch_bounds__(256) void k_blur_rows(const float* __restrict__ in,
                                                   float* __restrict__ out, int r1, int r2,
                                                   const double* __restrict__ w, int Rrt,
                                                   float* __restrict__ part) {
    const int R = RT > 0 ? RT : Rrt;
    constexpr int kTapUnroll = RT > 0 ? RT : 1;
    extern __shared__ float sb[];  // [256 + 2R]
    __shared__ float wmax[4];
    float vmax = 0.0f;
    const int b1 = min((static_cast<int>(blockIdx.x) + 1) * kBlurRows, r1);
    for (int b = blockIdx.x * kBlurRows; b < b1; ++b) {
        const u64 row = (static_cast<u64>(blockIdx.y) * r1 + b) * r2;
        for (int c0 = 0; c0 < r2; c0 += 256) {
            __syncthreads();  // the previous segment's taps are read
            for (int k = threadIdx.x; k < 256 + 2 * R; k += 256)
                sb[k] = in[row + reflect_fast(c0 - R + k, r2)];
            __syncthreads();
            const int k = threadIdx.x + R;
            double acc = static_cast<double>(sb[k]) * w[R];
#pragma unroll kTapUnroll
            for (int j = 1; j <= R; ++j)
                acc += (static_cast<double>(sb[k - j]) + static_cast<double>(sb[k + j])) * w[R + j];
            const int c = c0 + threadIdx.x;
            const float v = static_cast<float>(acc);
            if (c < r2) {
                out[row + c] = v;
                vmax = fmaxf(vmax, v);
            }
        }
    }
    for (int o = 32; o > 0; o >>= 1) vmax = fmaxf(vmax, __shfl_xor(vmax, o, 64));
    if ((threadIdx.x & 63) == 0) wmax[threadIdx.x >> 6] = vmax;
    __syncthreads();
    if (threadIdx.x == 0)
        part[static_cast<u64>(blockIdx.y) * gridDim.x + blockIdx.x] =
            fmaxf(fmaxf(wmax[0], wmax[1]), fmaxf(wmax[2], wmax[3]));
}

__global__ __launch_bounds__(kMeshThreads) void k_mesh_max(const float* __restrict__ g, u64 nv,
                                                           float* __restrict__ part) {
    __shared__ float s[kMeshThreads];
    float m = -INFINITY;
    const u64 stride = static_cast<u64>(gridDim.x) * kMeshThreads;
    for (u64 i = static_cast<u64>(blockIdx.x) * kMeshThreads + threadIdx.x; i < nv; i += stride)
        m = fmaxf(m, g[i]);
    s[threadIdx.x] = m;
    __syncthreads();
    if (threadIdx.x == 0) {
        for (int k = 1; k < kMeshThreads; ++k) m = fmaxf(m, s[k]);
        part[blockIdx.x] = m;
    }
}

struct McDev {
    uint8_t ntri[256];
    uint8_t tri[256][kMaxTri * 3];
};

// Marching cubes over wave segments.  A segment is 64 consecutive c of one
// (a, b) row, handled by one wave; segments in linear order cover the grid
// in grid-point order, so an exclusive scan of per-segment counts plus a
// prefix across the wave numbers every vertex by (grid point, axis) and
// every triangle by (cube, table order).  A grid point owns the crossing
// edges from it along +a, +b, +c (edge mask bits 0, 1, 2); a cube starts at
// its lowest corner.
constexpr int kMcPlanes = 16;     // a planes walked by each k_mc_segments thread
constexpr int kMcEmitPlanes = 64;  // a planes (segments) per wave of the emit passes: one per lane

// exclusive prefix over the wave of small per-lane counts (< 2^bits)
__device__ __forceinline__ u32 wave_prefix(u32 v, int bits) {
    const unsigned long long below = (1ull << (threadIdx.x & 63u)) - 1ull;
    u32 r = 0;
    for (int k = 0; k < bits; ++k)
        r += static_cast<u32>(__popcll(__ballot((v >> k) & 1u) & below)) << k;
    return r;
}

// corner k of a cube sits at (a + (k & 1), b + (k >> 1 & 1), c + (k >> 2 & 1))
__device__ __forceinline__ int mc_config(const float* __restrict__ g, u64 i, u64 sa, u64 sb,
                                         float iso) {
    int cfg = 0;
    for (int k = 0; k < 8; ++k) {
        const u64 j = i + (k & 1) * sa + ((k >> 1) & 1) * sb + ((k >> 2) & 1);
        cfg |= (g[j] > iso ? 1 : 0) << k;
    }
    return cfg;
}

// Vertex and triangle counts per segment.  Each thread walks kMcPlanes
// planes of a at one (b, c), keeping the inside bits of the (b, c), (b+1, c),
// (b, c+1), (b+1, c+1) points of the current plane (bits 0..3), so it loads
// four grid values per grid point.
__global__ __launch_bounds__(256) void k_mc_segments(const float* __restrict__ g, int r0, int r1,
                                                     int r2, float iso,
                                                     const McDev* __restrict__ T, u32 ntx,
                                                     u32* __restrict__ segE,
                                                     u32* __restrict__ segT) {
    const int lane = threadIdx.x & 63;
    const int b = blockIdx.y * 4 + (threadIdx.x >> 6);
    if (b >= r1) return;  // whole wave
    const int c = blockIdx.x * 64 + lane;
    const int a0 = blockIdx.z * kMcPlanes;
    const int a1 = min(a0 + kMcPlanes, r0);
    const bool cv = c < r2, bn = b + 1 < r1, cn = c + 1 < r2;
    auto plane = [&](int a) -> u32 {
        if (!cv) return 0u;
        const u64 i = (static_cast<u64>(a) * r1 + b) * r2 + c;
        u32 q = g[i] > iso ? 1u : 0u;
        if (bn && g[i + r2] > iso) q |= 2u;
        if (cn && g[i + 1] > iso) q |= 4u;
        if (bn && cn && g[i + r2 + 1] > iso) q |= 8u;
        return q;
    };
    u32 cur = plane(a0);
    for (int a = a0; a < a1; ++a) {
        const bool an = a + 1 < r0;
        const u32 nxt = an ? plane(a + 1) : 0u;
        u32 e = 0, t = 0;
        if (cv) {
            const u32 in0 = cur & 1u;
            u32 m = 0;
            if (an && (nxt & 1u) != in0) m |= 1u;
            if (bn && ((cur >> 1) & 1u) != in0) m |= 2u;
            if (cn && ((cur >> 2) & 1u) != in0) m |= 4u;
            e = static_cast<u32>(__popc(m));
            if (an && bn && cn) {
                // plane bit j = db + 2 dc is corner 2j (this plane) and 2j + 1 (next)
                u32 cfg = 0;
                for (int j = 0; j < 4; ++j)
                    cfg |= (((cur >> j) & 1u) << (2 * j)) | (((nxt >> j) & 1u) << (2 * j + 1));
                t = T->ntri[cfg];
            }
        }
        u32 p = e | (t << 16);  // at most 3 x 64 and 8 x 64 per wave
        for (int o = 32; o > 0; o >>= 1) p += __shfl_xor(p, o, 64);
        if (lane == 0) {
            const u64 seg = (static_cast<u64>(a) * r1 + b) * ntx + blockIdx.x;
            segE[seg] = p & 0xFFFFu;
            segT[seg] = p >> 16;
        }
        cur = nxt;
    }
}

// Vertices of the segments that have any; each grid point with crossing
// edges also records its edge mask and first vertex for k_mc_triangles.
__global__ __launch_bounds__(256) void k_mc_vertices(const float* __restrict__ g, MeshGeom G,
                                                     float iso, u32 ntx,
                                                     const u32* __restrict__ segE,
                                                     const u32* __restrict__ segEoff,
                                                     uint8_t* __restrict__ emask,
                                                     u32* __restrict__ eoff,
                                                     float* __restrict__ verts,
                                                     u32* __restrict__ vpt) {
    const int r0 = G.r0, r1 = G.r1, r2 = G.r2;
    const int b = blockIdx.y * 4 + (threadIdx.x >> 6);
    if (b >= r1) return;
    const int lane = threadIdx.x & 63;
    const int c = blockIdx.x * 64 + lane;
    // the wave's segments with vertices: lane l reads plane a0 + l's count
    const int a0 = blockIdx.z * kMcEmitPlanes;
    const u32 cnt = a0 + lane < r0
                        ? segE[(static_cast<u64>(a0 + lane) * r1 + b) * ntx + blockIdx.x]
                        : 0u;
    for (unsigned long long act = __ballot(cnt != 0); act; act &= act - 1) {
        const int a = a0 + __ffsll(act) - 1;
        const u64 seg = (static_cast<u64>(a) * r1 + b) * ntx + blockIdx.x;
        const u64 st[3] = {static_cast<u64>(r1) * r2, static_cast<u64>(r2), 1};
        const u64 i = (static_cast<u64>(a) * r1 + b) * r2 + c;
        u32 m = 0;
        float g0 = 0.0f;
        if (c < r2) {
            g0 = g[i];
            const bool in0 = g0 > iso;
            if (a + 1 < r0 && ((g[i + st[0]] > iso) != in0)) m |= 1u;
            if (b + 1 < r1 && ((g[i + st[1]] > iso) != in0)) m |= 2u;
            if (c + 1 < r2 && ((g[i + st[2]] > iso) != in0)) m |= 4u;
        }
        u32 v = segEoff[seg] + wave_prefix(static_cast<u32>(__popc(m)), 2);
        if (!m) continue;
        emask[i] = static_cast<uint8_t>(m);
        eoff[i] = v;
        const int idx[3] = {a, b, c};
        const double lo[3] = {G.lo0, G.lo1, G.lo2}, sp[3] = {G.sp0, G.sp1, G.sp2};
        const double v0 = static_cast<double>(g0);
        for (int ax = 0; ax < 3; ++ax) {
            if (!((m >> ax) & 1u)) continue;
            const double v1 = static_cast<double>(g[i + st[ax]]);
            const double t = (static_cast<double>(iso) - v0) / (v1 - v0);
            for (int k = 0; k < 3; ++k) {
                const double f = k == ax ? static_cast<double>(idx[k]) + t : static_cast<double>(idx[k]);
                verts[3 * static_cast<u64>(v) + k] = static_cast<float>(f * sp[k] + lo[k]);
            }
            if (vpt) vpt[v] = static_cast<u32>(i) << 2 | static_cast<u32>(ax);
            ++v;
        }
    }
}

// Triangles of the segments that have any: each vertex is the owner grid
// point's first vertex plus the crossing edges it owns on lower axes.
__global__ __launch_bounds__(256) void k_mc_triangles(const float* __restrict__ g, int r0, int r1,
                                                      int r2, float iso,
                                                      const McDev* __restrict__ T, u32 ntx,
                                                      const u32* __restrict__ segT,
                                                      const u32* __restrict__ segToff,
                                                      const uint8_t* __restrict__ emask,
                                                      const u32* __restrict__ eoff,
                                                      u32* __restrict__ tris) {
    const int b = blockIdx.y * 4 + (threadIdx.x >> 6);
    if (b >= r1) return;
    const int lane = threadIdx.x & 63;
    const int c = blockIdx.x * 64 + lane;
    // the wave's segments with triangles: lane l reads plane a0 + l's count
    const int a0 = blockIdx.z * kMcEmitPlanes;
    const u32 cnt = a0 + lane < r0
                        ? segT[(static_cast<u64>(a0 + lane) * r1 + b) * ntx + blockIdx.x]
                        : 0u;
    for (unsigned long long act = __ballot(cnt != 0); act; act &= act - 1) {
        const int a = a0 + __ffsll(act) - 1;
        const u64 seg = (static_cast<u64>(a) * r1 + b) * ntx + blockIdx.x;
        const u64 st[3] = {static_cast<u64>(r1) * r2, static_cast<u64>(r2), 1};
        const u64 i = (static_cast<u64>(a) * r1 + b) * r2 + c;
        u32 nt = 0;
        int cfg = 0;
        if (c + 1 < r2 && a + 1 < r0 && b + 1 < r1) {
            cfg = mc_config(g, i, st[0], st[1], iso);
            nt = T->ntri[cfg];
        }
        const u32 t0 = segToff[seg] + wave_prefix(nt, 4);
        for (u32 t = 0; t < nt; ++t) {
            for (int q = 0; q < 3; ++q) {
                const int e = T->tri[cfg][3 * t + q];
                const int ax = e >> 2, bb = e & 3;
                // owner corner: bit ax clear, the other two bits from bb (lower axis first)
                int off[3] = {0, 0, 0}, k = 0;
                for (int x = 0; x < 3; ++x) {
                    if (x == ax) continue;
                    off[x] = (bb >> k) & 1;
                    ++k;
                }
                const u64 o = i + off[0] * st[0] + off[1] * st[1] + off[2] * st[2];
                const u32 rank = static_cast<u32>(__popc(emask[o] & ((1u << ax) - 1u)));
                tris[3 * static_cast<u64>(t0 + t) + q] = eoff[o] + rank;
            }
        }
    }
}

// ---- Laplacian mesh smoothing ---------------------------------------------
// Open3D's TriangleMesh::filter_smooth_laplacian (the reference's
// mesh.filter_smooth_laplacian(number_of_iterations=n), default lambda 0.5),
// as its documentation states it: every vertex moves to
// v + lambda (sum_n w_n v_n / sum_n w_n - v), w_n = 1 / (|v - v_n| + 1e-12),
// over the vertices that share a triangle with it, all from the previous
// pass, in binary64.  The sums here run over the neighbours in ascending
// index (Open3D's hash-set order is unspecified).
// Neighbours per vertex: a cube's table has at most 5 triangles, so a vertex
// sits on one fan of at most 7 vertices per cube and has at most 6
// neighbours there, 24 over the four cubes around its edge.
constexpr int kAdjCap = 24;

// Neighbours of every vertex, from the (up to) four cubes around its grid
// edge: the other vertices of the table triangles that contain the edge.
__global__ __launch_bounds__(256) void k_mc_adjacency(const float* __restrict__ g, int r0, int r1,
                                                      int r2, float iso,
                                                      const McDev* __restrict__ T, u64 nvert,
                                                      const u32* __restrict__ vpt,
                                                      const uint8_t* __restrict__ emask,
                                                      const u32* __restrict__ eoff,
                                                      u32* __restrict__ adj,
                                                      u32* __restrict__ deg,
                                                      u32* __restrict__ over) {
    __shared__ __align__(16) McDev sT;  // the table, read per lane below
    for (u32 k = threadIdx.x; k < sizeof(McDev) / 4; k += 256)
        reinterpret_cast<u32*>(&sT)[k] = reinterpret_cast<const u32*>(T)[k];
    __syncthreads();
    const u64 v = static_cast<u64>(blockIdx.x) * 256 + threadIdx.x;
    if (v >= nvert) return;
    // this vertex's grid edge: from grid point i (< 2^30) along axis ax
    const u32 code = vpt[v];
    const u32 i32 = code >> 2;
    const int ax = static_cast<int>(code & 3u);
    const u32 plane = static_cast<u32>(r1) * static_cast<u32>(r2);
    const int idx[3] = {static_cast<int>(i32 / plane), static_cast<int>((i32 % plane) / r2),
                        static_cast<int>(i32 % static_cast<u32>(r2))};
    const u64 i = i32;
    const int r[3] = {r0, r1, r2};
    const u64 st[3] = {plane, static_cast<u64>(r2), 1};
    // ascending list in registers, 0xFFFFFFFF past its end
    u32 L[kAdjCap];
#pragma unroll
    for (int k = 0; k < kAdjCap; ++k) L[k] = 0xFFFFFFFFu;
    u32 n = 0;
    const int u = ax == 0 ? 1 : 0, w = ax == 2 ? 1 : 2;  // the other axes, lower first
    for (int o = 0; o < 4; ++o) {
        const int ou = o & 1, ow = o >> 1;
        // the cube whose lowest corner is this point - ou e_u - ow e_w
        if (idx[u] - ou < 0 || idx[w] - ow < 0 || idx[u] - ou + 1 >= r[u] ||
            idx[w] - ow + 1 >= r[w])
            continue;
        const u64 qi = i - ou * st[u] - ow * st[w];
        const int cfg = mc_config(g, qi, st[0], st[1], iso);
        const int mine = 4 * ax + (ou | (ow << 1));  // this edge within that cube
        for (u32 t = 0; t < sT.ntri[cfg]; ++t) {
            const uint8_t* tri = sT.tri[cfg] + 3 * t;
            if (tri[0] != mine && tri[1] != mine && tri[2] != mine) continue;
            for (int q = 0; q < 3; ++q) {
                const int e = tri[q];
                if (e == mine) continue;
                const int eax = e >> 2, bb = e & 3;
                int off[3] = {0, 0, 0}, k = 0;
                for (int x = 0; x < 3; ++x) {
                    if (x == eax) continue;
                    off[x] = (bb >> k) & 1;
                    ++k;
                }
                const u64 oi = qi + off[0] * st[0] + off[1] * st[1] + off[2] * st[2];
                u32 y = eoff[oi] + static_cast<u32>(__popc(emask[oi] & ((1u << eax) - 1u)));
                bool seen = false;
#pragma unroll
                for (int k2 = 0; k2 < kAdjCap; ++k2) seen |= L[k2] == y;
                if (seen) continue;
                if (n == kAdjCap) {
                    atomicOr(over, 1u);
                    continue;
                }
#pragma unroll
                for (int k2 = 0; k2 < kAdjCap; ++k2) {  // sorted insert
                    const bool less = y < L[k2];
                    const u32 tk = L[k2];
                    L[k2] = less ? y : tk;
                    y = less ? tk : y;
                }
                ++n;
            }
        }
    }
    u32* row = adj + v * kAdjCap;
#pragma unroll
    for (int k = 0; k < kAdjCap; ++k)
        if (static_cast<u32>(k) < n) row[k] = L[k];
    deg[v] = n;
}

__global__ __launch_bounds__(kMeshThreads) void k_verts_widen(const float* __restrict__ in,
                                                              u64 n, double* __restrict__ out) {
    const u64 i = static_cast<u64>(blockIdx.x) * kMeshThreads + threadIdx.x;
    if (i < n) out[i] = static_cast<double>(in[i]);
}

__global__ __launch_bounds__(kMeshThreads) void k_verts_narrow(const double* __restrict__ in,
                                                               u64 n, float* __restrict__ out) {
    const u64 i = static_cast<u64>(blockIdx.x) * kMeshThreads + threadIdx.x;
    if (i < n) out[i] = static_cast<float>(in[i]);
}

__global__ __launch_bounds__(kMeshThreads) void k_mesh_laplacian(const double* __restrict__ V,
                                                                 u64 nv,
                                                                 const u32* __restrict__ adj,
                                                                 const u32* __restrict__ deg,
                                                                 double* __restrict__ out) {
    const u64 v = static_cast<u64>(blockIdx.x) * kMeshThreads + threadIdx.x;
    if (v >= nv) return;
    const double px = V[3 * v], py = V[3 * v + 1], pz = V[3 * v + 2];
    double sx = 0.0, sy = 0.0, sz = 0.0, tw = 0.0;
    const u32* L = adj + v * kAdjCap;
    const u32 n = deg[v];
    for (u32 k = 0; k < n; ++k) {
        const u64 j = L[k];
        const double qx = V[3 * j], qy = V[3 * j + 1], qz = V[3 * j + 2];
        const double dx = px - qx, dy = py - qy, dz = pz - qz;
        const double wgt = 1.0 / (sqrt(dx * dx + dy * dy + dz * dz) + 1e-12);
        tw = tw + wgt;
        sx = sx + wgt * qx;
        sy = sy + wgt * qy;
        sz = sz + wgt * qz;
    }
    out[3 * v] = px + 0.5 * (sx / tw - px);
    out[3 * v + 1] = py + 0.5 * (sy / tw - py);
    out[3 * v + 2] = pz + 0.5 * (sz / tw - pz);
}

// ---- exclusive scan of u32 counts (1024 per block) ----------------------------
constexpr int kScanPer = 4;
constexpr int kScanBlock = kMeshThreads * kScanPer;

__global__ __launch_bounds__(kMeshThreads) void k_scan_local(const u32* __restrict__ in,
                                                             u32* __restrict__ out, u64 n,
                                                             u32* __restrict__ sums) {
    __shared__ u32 s[kMeshThreads];
    const u64 base = static_cast<u64>(blockIdx.x) * kScanBlock + threadIdx.x * kScanPer;
    u32 v[kScanPer], tot = 0;
    for (int k = 0; k < kScanPer; ++k) {
        v[k] = base + k < n ? in[base + k] : 0u;
        tot += v[k];
    }
    s[threadIdx.x] = tot;
    __syncthreads();
    for (int o = 1; o < kMeshThreads; o <<= 1) {
        const u32 add = threadIdx.x >= static_cast<u32>(o) ? s[threadIdx.x - o] : 0u;
        __syncthreads();
        s[threadIdx.x] += add;
        __syncthreads();
    }
    u32 run = s[threadIdx.x] - tot;
    for (int k = 0; k < kScanPer; ++k) {
        if (base + k < n) out[base + k] = run;
        run += v[k];
    }
    if (threadIdx.x == kMeshThreads - 1) sums[blockIdx.x] = s[threadIdx.x];
}

__global__ __launch_bounds__(kMeshThreads) void k_scan_add(u32* __restrict__ out, u64 n,
                                                           const u32* __restrict__ offs) {
    const u64 base = static_cast<u64>(blockIdx.x) * kScanBlock + threadIdx.x * kScanPer;
    const u32 add = offs[blockIdx.x];
    for (int k = 0; k < kScanPer; ++k)
        if (base + k < n) out[base + k] += add;
}

}  // namespace

// Exclusive scan of n counts into out; returns the total (synchronous).
static u64 scan_u32(const u32* in, u32* out, u64 n, hipStream_t s) {
    if (n == 0) return 0;
    const u64 nb = (n + kScanBlock - 1) / kScanBlock;
    DeviceBuffer<u32> sums(nb), offs(nb);
    k_scan_local<<<static_cast<u32>(nb), kMeshThreads, 0, s>>>(in, out, n, sums.data());
    WSIM_LAUNCH_CHECK();
    std::vector<u32> h(nb);
    WSIM_HIP(hipMemcpyAsync(h.data(), sums.data(), nb * sizeof(u32), hipMemcpyDeviceToHost, s));
    WSIM_HIP(hipStreamSynchronize(s));
    u64 run = 0;
    for (u64 k = 0; k < nb; ++k) {
        const u64 t = h[k];
        h[k] = static_cast<u32>(run);
        run += t;
        if (run > 0xFFFFFFFFull) fail(kOutOfRange, "SurfaceMesher: more than 2^32 - 1 elements");
    }
    WSIM_HIP(hipMemcpyAsync(offs.data(), h.data(), nb * sizeof(u32), hipMemcpyHostToDevice, s));
    k_scan_add<<<static_cast<u32>(nb), kMeshThreads, 0, s>>>(out, n, offs.data());
    WSIM_LAUNCH_CHECK();
    WSIM_HIP(hipStreamSynchronize(s));
    return run;
}

// numpy's float64 sum: sequential below 8 elements, eight running sums
// combined pairwise up to 128, halves (a multiple of 8 first) beyond.
static double numpy_sum(const double* a, size_t n) {
    if (n < 8) {
        double r = 0.0;
        for (size_t i = 0; i < n; ++i) r = r + a[i];
        return r;
    }
    if (n <= 128) {
        double r[8];
        for (int j = 0; j < 8; ++j) r[j] = a[j];
        size_t i = 8;
        for (; i < n - n % 8; i += 8)
            for (int j = 0; j < 8; ++j) r[j] = r[j] + a[i + j];
        double res = ((r[0] + r[1]) + (r[2] + r[3])) + ((r[4] + r[5]) + (r[6] + r[7]));
        for (; i < n; ++i) res = res + a[i];
        return res;
    }
    size_t n2 = n / 2;
    n2 -= n2 % 8;
    return numpy_sum(a, n2) + numpy_sum(a + n2, n - n2);
}

std::vector<double> gaussian_weights(double sigma, int* radius) {
    // scipy.ndimage._gaussian_kernel1d, order 0, truncate 4.0: phi / phi.sum()
    const int R = static_cast<int>(4.0 * sigma + 0.5);
    std::vector<double> w(2 * R + 1);
    const double s2 = sigma * sigma;
    for (int k = -R; k <= R; ++k) {
        w[k + R] = std::exp(-0.5 / s2 * static_cast<double>(k * k));
    }
    const double sum = numpy_sum(w.data(), w.size());
    for (double& x : w) x = x / sum;
    *radius = R;
    return w;
}

const uint8_t* mc_table_counts() { return mc_table().ntri; }
const uint8_t* mc_table_triangles() { return &mc_table().tri[0][0]; }

SurfaceMesher::SurfaceMesher(hipStream_t stream) : stream_(stream) {}

void SurfaceMesher::run(const float* xyz, size_t n, const MeshParams& P) {
    if (!xyz && n) fail(kInvalidArgument, "SurfaceMesher: NULL positions");
    if (n == 0) fail(kInvalidArgument, "SurfaceMesher: no points");
    if (!(P.particle_radius > 0.0) || !(P.grid_voxel_factor > 0.0) || !(P.smooth_sigma >= 0.0))
        fail(kInvalidArgument, "SurfaceMesher: radius, voxel factor and sigma must be positive");
    if (P.smooth_pre_iters < 0 || P.smooth_post_iters < 0)
        fail(kInvalidArgument, "SurfaceMesher: negative smoothing iterations");
    if (P.smooth_pre_iters > 0) {
        const int k = P.k_neighbors;
        if (k != 8 && k != 10 && k != 15 && k != 20 && k != 30)
            fail(kInvalidArgument, "SurfaceMesher: k_neighbors must be 8, 10, 15, 20 or 30");
        if (n < static_cast<size_t>(k))
            fail(kInvalidArgument, "SurfaceMesher: fewer points than k_neighbors");
    }
    if (n > 0xFFFFF000ull) fail(kInvalidArgument, "SurfaceMesher: too many points");
    hipStream_t s = stream_;
    const u32 nn = static_cast<u32>(n);
    const u32 gp = div_up(nn, static_cast<u32>(kMeshThreads));
    if (pts_[0].size() < 3 * n) {
        pts_[0] = DeviceBuffer<double>(3 * n);
        pts_[1] = DeviceBuffer<double>(3 * n);
    }
    k_mesh_points<<<gp, kMeshThreads, 0, s>>>(xyz, n, P.swap_yz ? 1 : 0, pts_[0].data());
    WSIM_LAUNCH_CHECK();
    const u32 nbb = std::min<u32>(gp, 1024);
    DeviceBuffer<double> part(kBoundsPart * static_cast<size_t>(nbb));
    auto bounds = [&](const double* pts, double* mn, double* mx) {
        k_mesh_bounds<<<nbb, kMeshThreads, 0, s>>>(pts, n, part.data());
        WSIM_LAUNCH_CHECK();
        std::vector<double> hp(kBoundsPart * static_cast<size_t>(nbb));
        WSIM_HIP(hipMemcpyAsync(hp.data(), part.data(), hp.size() * sizeof(double),
                                hipMemcpyDeviceToHost, s));
        WSIM_HIP(hipStreamSynchronize(s));
        for (int a = 0; a < 3; ++a) {
            mn[a] = INFINITY;
            mx[a] = -INFINITY;
        }
        double bad = 0.0;
        for (u32 k = 0; k < nbb; ++k) {
            for (int a = 0; a < 3; ++a) {
                mn[a] = std::fmin(mn[a], hp[kBoundsPart * k + a]);
                mx[a] = std::fmax(mx[a], hp[kBoundsPart * k + 3 + a]);
            }
            bad += hp[kBoundsPart * k + 6];
        }
        if (bad != 0.0) fail(kInvalidArgument, "SurfaceMesher: non-finite positions");
        for (int a = 0; a < 3; ++a)
            if (!std::isfinite(mn[a]) || !std::isfinite(mx[a]))
                fail(kInvalidArgument, "SurfaceMesher: non-finite positions");
    };
    // 1. k-nearest-neighbour position smoothing (:61-74)
    int cur = 0;
    double mn[3], mx[3];
    for (int it = 0; it < P.smooth_pre_iters; ++it) {
        bounds(pts_[cur].data(), mn, mx);
        KnnGrid kg{};
        kg.lo0 = mn[0], kg.lo1 = mn[1], kg.lo2 = mn[2];
        kg.cell = 4.0 * P.particle_radius;
        u64 cells = 0;
        for (;;) {
            kg.d0 = static_cast<int>((mx[0] - mn[0]) / kg.cell) + 1;
            kg.d1 = static_cast<int>((mx[1] - mn[1]) / kg.cell) + 1;
            kg.d2 = static_cast<int>((mx[2] - mn[2]) / kg.cell) + 1;
            cells = static_cast<u64>(kg.d0) * kg.d1 * kg.d2;
            if (cells <= (1ull << 26)) break;
            kg.cell *= 2.0;
        }
        if (knn_keys_[0].size() < n) {
            for (auto& b : knn_keys_) b = DeviceBuffer<u32>(n);
            for (auto& b : knn_vals_) b = DeviceBuffer<u32>(n);
            knn_hist_ = DeviceBuffer<u32>(radix_hist_elems(n));
        }
        if (knn_cs_.size() < cells + 1) knn_cs_ = DeviceBuffer<u32>(cells + 1);
        k_knn_keys<<<gp, kMeshThreads, 0, s>>>(pts_[cur].data(), nn, kg, knn_keys_[0].data(),
                                               knn_vals_[0].data());
        WSIM_LAUNCH_CHECK();
        const int w = radix_sort_pairs(knn_keys_[0].data(), knn_vals_[0].data(), knn_keys_[1].data(),
                                       knn_vals_[1].data(), n, key_bits(static_cast<u32>(cells - 1)),
                                       knn_hist_.data(), s);
        build_cell_start(knn_keys_[w].data(), nn, static_cast<u32>(cells + 1), knn_cs_.data(), s);
        const u32* sorted = knn_vals_[w].data();
        double* out = pts_[cur ^ 1].data();
        if (knn_ps_.size() < 3 * n) knn_ps_ = DeviceBuffer<double>(3 * n);
        k_knn_gather<<<gp, kMeshThreads, 0, s>>>(pts_[cur].data(), nn, sorted, knn_ps_.data());
        WSIM_LAUNCH_CHECK();
        const double* in = pts_[cur].data();
        const double* ps = knn_ps_.data();
        const u32* cs = knn_cs_.data();
        if (knn_far_.size() < n) knn_far_ = DeviceBuffer<u32>(n);
        if (knn_farc_.size() < 1) knn_farc_ = DeviceBuffer<u32>(1);
        u32* fc = knn_farc_.data();
        u32* fl = knn_far_.data();
        WSIM_HIP(hipMemsetAsync(fc, 0, sizeof(u32), s));
        const u32 gf = std::min<u32>(gp, 256);  // 4 waves per block walk the list
        switch (P.k_neighbors) {
        case 8:
            k_knn_smooth<8><<<gp, kMeshThreads, 0, s>>>(in, nn, kg, sorted, ps, cs, out, fc, fl);
            k_knn_far<8><<<gf, kMeshThreads, 0, s>>>(in, kg, sorted, ps, cs, fc, fl, out);
            break;
        case 10:
            k_knn_smooth<10><<<gp, kMeshThreads, 0, s>>>(in, nn, kg, sorted, ps, cs, out, fc, fl);
            k_knn_far<10><<<gf, kMeshThreads, 0, s>>>(in, kg, sorted, ps, cs, fc, fl, out);
            break;
        case 15:
            k_knn_smooth<15><<<gp, kMeshThreads, 0, s>>>(in, nn, kg, sorted, ps, cs, out, fc, fl);
            k_knn_far<15><<<gf, kMeshThreads, 0, s>>>(in, kg, sorted, ps, cs, fc, fl, out);
            break;
        case 20:
            k_knn_smooth<20><<<gp, kMeshThreads, 0, s>>>(in, nn, kg, sorted, ps, cs, out, fc, fl);
            k_knn_far<20><<<gf, kMeshThreads, 0, s>>>(in, kg, sorted, ps, cs, fc, fl, out);
            break;
        default:
            k_knn_smooth<30><<<gp, kMeshThreads, 0, s>>>(in, nn, kg, sorted, ps, cs, out, fc, fl);
            k_knn_far<30><<<gf, kMeshThreads, 0, s>>>(in, kg, sorted, ps, cs, fc, fl, out);
            break;
        }
        WSIM_LAUNCH_CHECK();
        cur ^= 1;
    }
    const double* pts = pts_[cur].data();
    bounds(pts, mn, mx);
    // 2. grid (tools/particles_to_mesh.py:104-109), binary64 as numpy
    const double pad = P.particle_radius * 4.0;
    const double voxel = P.particle_radius * P.grid_voxel_factor;
    MeshGeom G{};
    int res[3];
    double lo[3], ext[3];
    for (int a = 0; a < 3; ++a) {
        lo[a] = mn[a] - pad;
        const double hi = mx[a] + pad;
        ext[a] = hi - lo[a];
        const double r = std::ceil(ext[a] / voxel);
        if (!(r < 65536.0)) fail(kOutOfRange, "SurfaceMesher: grid too large");
        res[a] = std::max(static_cast<int>(r), 8);
    }
    G.r0 = res[0], G.r1 = res[1], G.r2 = res[2];
    G.lo0 = lo[0], G.lo1 = lo[1], G.lo2 = lo[2];
    G.ext0 = ext[0], G.ext1 = ext[1], G.ext2 = ext[2];
    G.sp0 = ext[0] / res[0], G.sp1 = ext[1] / res[1], G.sp2 = ext[2] / res[2];
    const u64 nv = static_cast<u64>(res[0]) * res[1] * res[2];
    if (nv > (1ull << 31)) fail(kOutOfRange, "SurfaceMesher: grid too large");
    info_ = MeshInfo{};
    for (int a = 0; a < 3; ++a) {
        info_.dims[a] = res[a];
        info_.lo[a] = lo[a];
        info_.extent[a] = ext[a];
    }
    info_.spacing[0] = G.sp0, info_.spacing[1] = G.sp1, info_.spacing[2] = G.sp2;
    const u32 gv = static_cast<u32>((nv + kMeshThreads - 1) / kMeshThreads);
    // grid-sized work buffers are kept between runs and grow as needed
    auto fit = [](auto& b, u64 want) {
        if (b.size() < want) b = std::remove_reference_t<decltype(b)>(want);
    };
    fit(grid_, nv);
    fit(tmp_, nv);
    fit(count_, nv);
    fit(eoff_, nv);
    fit(emask_, nv);
    // 3. splat
    u32* count = count_.data();
    WSIM_HIP(hipMemsetAsync(count, 0, nv * sizeof(u32), s));
    k_mesh_splat<<<gp, kMeshThreads, 0, s>>>(pts, n, G, count);
    WSIM_LAUNCH_CHECK();
    // 4. blur and 5. the maximum for the iso level
    int R = 0;
    std::vector<double> w;
    DeviceBuffer<double> dw;
    if (P.smooth_sigma > 0.0) {
        w = gaussian_weights(P.smooth_sigma, &R);
        if (P.weights) std::memcpy(w.data(), P.weights, w.size() * sizeof(double));
        dw = DeviceBuffer<double>(w.size());
        WSIM_HIP(hipMemcpyAsync(dw.data(), w.data(), w.size() * sizeof(double),
                                hipMemcpyHostToDevice, s));
    }
    auto max_of = [&](const float* a, u64 na) {
        const u32 nbm = static_cast<u32>(std::min<u64>((na + kMeshThreads - 1) / kMeshThreads, 1024));
        DeviceBuffer<float> mpart(nbm);
        k_mesh_max<<<nbm, kMeshThreads, 0, s>>>(a, na, mpart.data());
        WSIM_LAUNCH_CHECK();
        std::vector<float> hm(nbm);
        WSIM_HIP(hipMemcpyAsync(hm.data(), mpart.data(), nbm * sizeof(float),
                                hipMemcpyDeviceToHost, s));
        WSIM_HIP(hipStreamSynchronize(s));
        float m = -INFINITY;
        for (float v : hm) m = std::fmax(m, v);
        return m;
    };
    float gmax;
    if (P.smooth_sigma > 0.0 && R <= kBlurMaxR) {
        // counts -> grid_ (axis 0) -> tmp_ (axis 1) -> grid_ (axis 2)
        const size_t lds = static_cast<size_t>(kBlurTile + 2 * R) * 64 * sizeof(float);
        const u32 tx = static_cast<u32>((res[2] + 63) / 64);
        const dim3 g0(tx, static_cast<u32>((res[0] + kBlurTile - 1) / kBlurTile), static_cast<u32>(res[1]));
        k_blur_strided<u32><<<g0, 256, lds, s>>>(count, grid_.data(), res[0], res[1], res[2], 0,
                                                 dw.data(), R);
        WSIM_LAUNCH_CHECK();
        const dim3 g1(tx, static_cast<u32>((res[1] + kBlurTile - 1) / kBlurTile), static_cast<u32>(res[0]));
        k_blur_strided<float><<<g1, 256, lds, s>>>(grid_.data(), tmp_.data(), res[0], res[1],
                                                   res[2], 1, dw.data(), R);
        WSIM_LAUNCH_CHECK();
        const dim3 g2(static_cast<u32>((res[1] + kBlurRows - 1) / kBlurRows),
                      static_cast<u32>(res[0]));
        const u64 nparts = static_cast<u64>(g2.x) * g2.y;
        fit(part_, nparts);
        const size_t lr = (256 + 2 * R) * sizeof(float);
        if (R == 10)  // sigma 2.5, the tool's default: unrolled taps
            k_blur_rows<10><<<g2, 256, lr, s>>>(tmp_.data(), grid_.data(), res[1], res[2],
                                                dw.data(), R, part_.data());
        else
            k_blur_rows<0><<<g2, 256, lr, s>>>(tmp_.data(), grid_.data(), res[1], res[2],
                                               dw.data(), R, part_.data());
        WSIM_LAUNCH_CHECK();
        gmax = max_of(part_.data(), nparts);
    } else {
        k_mesh_to_float<<<gv, kMeshThreads, 0, s>>>(count, grid_.data(), nv);
        WSIM_LAUNCH_CHECK();
        if (P.smooth_sigma > 0.0) {  // taps beyond the tiles' LDS: one load per tap
            float* src = grid_.data();
            float* dst = tmp_.data();
            for (int axis = 0; axis < 3; ++axis) {
                k_mesh_blur<<<gv, kMeshThreads, 0, s>>>(src, dst, res[0], res[1], res[2], axis,
                                                        dw.data(), R);
                WSIM_LAUNCH_CHECK();
                std::swap(src, dst);
            }
            if (src != grid_.data()) std::swap(grid_, tmp_);
        }
        gmax = max_of(grid_.data(), nv);
    }
    info_.max_density = gmax;
    // numpy 2: float32 scalar * Python float stays float32
    const float iso = gmax * static_cast<float>(P.iso_factor);
    info_.iso = iso;
    info_.n_vertices = info_.n_triangles = 0;
    if (!(gmax > 0.0f)) return;
    // 6. marching cubes
    const McTable& T = mc_table();
    static_assert(sizeof(McDev) == sizeof(McTable), "table layout");
    DeviceBuffer<McDev> dT(1);
    WSIM_HIP(hipMemcpyAsync(dT.data(), &T, sizeof(McTable), hipMemcpyHostToDevice, s));
    const u32 ntx = static_cast<u32>((res[2] + 63) / 64);
    const u64 nseg = static_cast<u64>(res[0]) * res[1] * ntx;
    fit(segE_, nseg);
    fit(segT_, nseg);
    fit(segEoff_, nseg);
    fit(segToff_, nseg);
    const u32 by = static_cast<u32>((res[1] + 3) / 4);
    const dim3 gs(ntx, by, static_cast<u32>((res[0] + kMcPlanes - 1) / kMcPlanes));
    k_mc_segments<<<gs, 256, 0, s>>>(grid_.data(), res[0], res[1], res[2], iso, dT.data(), ntx,
                                     segE_.data(), segT_.data());
    WSIM_LAUNCH_CHECK();
    const u64 nvert = scan_u32(segE_.data(), segEoff_.data(), nseg, s);
    const u64 ntri = scan_u32(segT_.data(), segToff_.data(), nseg, s);
    fit(verts_, 3 * std::max<u64>(nvert, 1));
    fit(tris_, 3 * std::max<u64>(ntri, 1));
    const dim3 ge(ntx, by, static_cast<u32>((res[0] + kMcEmitPlanes - 1) / kMcEmitPlanes));
    const bool post = P.smooth_post_iters > 0 && nvert > 0;
    if (post) {
        if (nv > (1ull << 30)) fail(kOutOfRange, "SurfaceMesher: grid too large for smoothing");
        fit(vpt_, nvert);
    }
    k_mc_vertices<<<ge, 256, 0, s>>>(grid_.data(), G, iso, ntx, segE_.data(), segEoff_.data(),
                                     emask_.data(), eoff_.data(), verts_.data(),
                                     post ? vpt_.data() : nullptr);
    WSIM_LAUNCH_CHECK();
    k_mc_triangles<<<ge, 256, 0, s>>>(grid_.data(), res[0], res[1], res[2], iso, dT.data(), ntx,
                                      segT_.data(), segToff_.data(), emask_.data(), eoff_.data(),
                                      tris_.data());
    WSIM_LAUNCH_CHECK();
    // 7. Laplacian mesh smoothing (:138-139)
    if (post) {
        fit(adj_, nvert * kAdjCap);
        fit(deg_, nvert);
        fit(adj_over_, 1);
        for (auto& b : vd_) fit(b, 3 * nvert);
        WSIM_HIP(hipMemsetAsync(adj_over_.data(), 0, sizeof(u32), s));
        k_mc_adjacency<<<static_cast<u32>((nvert + 255) / 256), 256, 0, s>>>(
            grid_.data(), res[0], res[1], res[2], iso, dT.data(), nvert, vpt_.data(),
            emask_.data(), eoff_.data(), adj_.data(), deg_.data(), adj_over_.data());
        WSIM_LAUNCH_CHECK();
        u32 over = 0;
        WSIM_HIP(hipMemcpyAsync(&over, adj_over_.data(), sizeof(u32), hipMemcpyDeviceToHost, s));
        WSIM_HIP(hipStreamSynchronize(s));
        if (over) fail(kRuntimeError, "SurfaceMesher: a vertex with more than 24 neighbours");
        const u64 n3 = 3 * nvert;
        const u32 g3 = static_cast<u32>((n3 + kMeshThreads - 1) / kMeshThreads);
        const u32 g1 = static_cast<u32>((nvert + kMeshThreads - 1) / kMeshThreads);
        k_verts_widen<<<g3, kMeshThreads, 0, s>>>(verts_.data(), n3, vd_[0].data());
        WSIM_LAUNCH_CHECK();
        int cur = 0;
        for (int it = 0; it < P.smooth_post_iters; ++it) {
            k_mesh_laplacian<<<g1, kMeshThreads, 0, s>>>(vd_[cur].data(), nvert, adj_.data(),
                                                         deg_.data(), vd_[cur ^ 1].data());
            WSIM_LAUNCH_CHECK();
            cur ^= 1;
        }
        k_verts_narrow<<<g3, kMeshThreads, 0, s>>>(vd_[cur].data(), n3, verts_.data());
        WSIM_LAUNCH_CHECK();
    }
    WSIM_HIP(hipStreamSynchronize(s));
    info_.n_vertices = nvert;
    info_.n_triangles = ntri;
}

}  // namespace wsim
