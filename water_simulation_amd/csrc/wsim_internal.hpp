// wsim_internal.hpp -- the C++ objects behind the C ABI handles.
#pragma once

#include <array>
#include <memory>
#include <string>
#include <unordered_map>
#include <vector>

#include "common.hpp"
#include "device_math.hpp"

namespace wsim {

// ---------------------------------------------------------------------------
// ParticleStore (particle_store.h:15-127, particle_store.cu:5-22): named SoA
// device attributes of fixed capacity, zero-filled at registration;
// "position" and "velocity" (F32x3, 12-byte AoS records) always exist.
enum AttribType : int { kF32 = 0, kF32x3 = 1, kU32 = 2 };

class ParticleStore {
public:
    ParticleStore(size_t capacity, hipStream_t stream);
    size_t register_attribute(const std::string& name, int type);
    bool has_attribute(const std::string& name) const;
    size_t find_attribute(const std::string& name) const;
    void* attribute_data(size_t handle);
    int attribute_type(size_t handle) const;
    float* positions() { return static_cast<float*>(attribute_data(position_)); }
    float* velocities() { return static_cast<float*>(attribute_data(velocity_)); }
    size_t position_handle() const { return position_; }
    size_t velocity_handle() const { return velocity_; }
    void resize(size_t count);
    size_t count() const { return count_; }
    size_t capacity() const { return capacity_; }
    hipStream_t stream() const { return stream_; }

private:
    struct Attrib {
        std::string name;
        int type;
        size_t elem;
        DeviceBuffer<unsigned char> buf;
    };
    size_t capacity_;
    size_t count_ = 0;
    hipStream_t stream_;
    std::vector<Attrib> attribs_;
    std::unordered_map<std::string, size_t> index_;
    size_t position_ = 0, velocity_ = 0;
};

// ---------------------------------------------------------------------------
// CellGrid (cell_grid.h:16-50): cell keys, stable (key, id) sort, dense
// cell_start[C+1] with cell_start[c] = #particles whose key < c.
class CellGrid {
public:
    CellGrid(const float origin[3], const int32_t dims[3], float cell_length, hipStream_t stream);
    void reserve(size_t capacity);
    void build(const float* positions_xyz, size_t count);

    const u32* cell_start() const { return cs_.data(); }
    const u32* sorted_indices() const { return v_[cur_].data(); }
    const u32* sorted_keys() const { return k_[cur_].data(); }
    const GridGeom& geom() const { return g_; }
    float cell_length() const { return cell_; }
    u32 total_cells() const { return g_.C; }
    size_t count() const { return count_; }
    size_t capacity() const { return capacity_; }
    hipStream_t stream() const { return stream_; }

private:
    GridGeom g_;
    float cell_;
    hipStream_t stream_;
    size_t capacity_ = 0, count_ = 0;
    DeviceBuffer<u32> k_[2], v_[2], hist_, cs_;
    int cur_ = 0;
};

// Plain 16-byte streaming copy kernel (copy.hip).
void device_copy(void* dst, const void* src, size_t bytes, hipStream_t stream);

// Cell-start table from sorted keys (one pass over the C+1 cells).
void build_cell_start(const u32* sorted_keys, u32 n, u32 num_cells_plus_one, u32* cell_start,
                      hipStream_t stream);

// ---------------------------------------------------------------------------
// Morton path of SpatialAccel (spatial_accel.h:19-53, spatial_accel.cu:11-40,
// 111-179): 30-bit codes on the scene box and their stable sort.
class SpatialAccel {
public:
    void build(const float* positions_xyz, size_t count, const float bmin[3], const float bmax[3],
               hipStream_t stream);
    size_t leaf_count() const { return leaf_count_; }
    const u32* sorted_indices() const { return leaf_count_ ? v_[cur_].data() : nullptr; }
    const u32* morton_codes() const { return leaf_count_ ? k_[cur_].data() : nullptr; }

private:
    size_t leaf_count_ = 0;
    DeviceBuffer<u32> k_[2], v_[2], hist_;
    int cur_ = 0;
};

// ---------------------------------------------------------------------------
// DFSPH solver (dfsph.h:17-84, dfsph.cu, dfsph_kernels.cu).
struct DFSPHConfig {
    float rest_density = 1000.0f;
    float particle_radius = 0.0015f;
    float smoothing_length = 0.003f;
    float viscosity = 1e-3f;
    float damping = 0.5f;
    float surface_tension = 0.0f;
    float gravity[3] = {0.0f, -9.81f, 0.0f};
    u32 max_density_iters = 100;
    u32 max_divergence_iters = 100;
    float eta_density = 1e-3f;
    float eta_divergence = 1e-3f;
    float domain_min[3] = {0.0f, 0.0f, 0.0f};
    float domain_max[3] = {1.0f, 1.0f, 1.0f};
    u32 enable_divergence_solve = 0;
    u32 pressure_form = 0;  // 0: kappa/rho^2 (reference), 1: kappa/rho
};

struct StepStats {
    u32 density_iters = 0;
    u32 divergence_iters = 0;
    double mean_density_error = 0.0;
    double mean_divergence_error = 0.0;
    float max_velocity = 0.0f;
    u32 nan_count = 0;
    u64 substeps = 0;
};

// Device-resident solve state; counters self-reset so no memset is needed
// between launches (all of it is valid under graph capture).
struct DeviceState {
    u32 ticket[4];
    u32 conv_iter;      // density solve: iteration whose error test passed, or kNone
    u32 div_conv_iter;  // divergence solve: same
    u32 nan_count;
    float max_v;
    double err;         // mean error of the latest executed density iteration
    double div_err;     // same for the divergence solve
};

enum KernelKind : int {
    kKGrid = 0,
    kKReorder,
    kKDensityAlpha,
    kKXsph,
    kKDensityChange,
    kKApplyKappa,
    kKAdvect,
    kKDivChange,
    kKDivApply,
    kKDivWriteback,
    kKNumKinds
};
const char* kernel_kind_name(int k);

class KernelTimer {
public:
    ~KernelTimer();
    void begin(int kind, int iter, hipStream_t s);
    void end(hipStream_t s);
    // Must be called after the stream is idle.  Launches of the iteration
    // kernels with iter >= valid_iters[kind] were device-side no-ops and are
    // dropped.
    void collect(const int* valid_iters);
    void reset();
    struct Acc {
        u64 count = 0;
        double total_ms = 0.0;
    };
    std::array<Acc, kKNumKinds> acc{};

private:
    struct Pair {
        hipEvent_t a, b;
        int kind, iter;
    };
    std::vector<hipEvent_t> pool_;
    std::vector<Pair> pending_;
    hipEvent_t take();
    hipEvent_t open_ = nullptr;
    int open_kind_ = 0, open_iter_ = 0;
};

// Per-particle pair lists of the pressure solves (see solver.hip).
struct PairList {
    uint2* f;   // fluid pairs (slot, bits of c), ELL [k * n + i]
    uint2* b;   // boundary pairs
    u32* cnt;   // fluid count | boundary count << 16 (saturating)
    u32 kf, kb; // list depths
};

// Depth of the per-particle pair lists (pairs with a non-zero kernel
// gradient); a particle with more pairs walks the stencil instead.
constexpr u32 kPairDepthFluid = 48;
constexpr u32 kPairDepthBoundary = 48;

class DFSPHSolver {
public:
    DFSPHSolver(ParticleStore& fluid, ParticleStore* boundary, const DFSPHConfig& cfg,
                hipStream_t stream);
    void step(float dt);

    float max_velocity() const { return stats_.max_velocity; }
    float particle_mass() const { return mass_; }
    u32 last_density_iters() const { return stats_.density_iters; }
    const StepStats& stats() const { return stats_; }
    const CellGrid& fluid_grid() const { return fluid_grid_; }
    const CellGrid* boundary_grid() const { return boundary_grid_.get(); }
    float attribute_at(int which, size_t i) const;  // 0 density, 1 alpha, 2 boundary mass
    const DFSPHConfig& config() const { return cfg_; }

    void set_timing(bool on) { timing_ = on; }
    KernelTimer& timer() { return timer_; }

private:
    void density_solve(float dt);
    void divergence_solve(float dt);
    void sort_and_reorder(bool gravity, float dt);
    void read_state();
    PairList pair_list();
    void tbegin(int kind, int iter = 0) {
        if (timing_) timer_.begin(kind, iter, stream_);
    }
    void tend() {
        if (timing_) timer_.end(stream_);
    }

    ParticleStore& fluid_;
    ParticleStore* boundary_;
    DFSPHConfig cfg_;
    hipStream_t stream_;
    CellGrid fluid_grid_;
    std::unique_ptr<CellGrid> boundary_grid_;
    float mass_ = 0.0f;
    size_t nb_ = 0;
    StepStats stats_;

    size_t h_density_ = 0, h_alpha_ = 0, h_kappa_ = 0, h_density_adv_ = 0, h_kappa_v_ = 0;
    size_t h_bmass_ = 0;
    bool has_kappa_v_ = false;

    // cell-sorted working set
    DeviceBuffer<float4> xs_, vs_, vs2_, xk_, xb_;
    DeviceBuffer<uint2> pairs_f_, pairs_b_;
    DeviceBuffer<u32> pair_cnt_;
    DeviceBuffer<float> rho_, alpha_, kappa_, rho_adv_;
    DeviceBuffer<double> part_d_;
    DeviceBuffer<float> part_f_;
    DeviceBuffer<u32> part_u_;
    DeviceBuffer<DeviceState> dstate_;
    PinnedBuffer<DeviceState> hstate_;

    bool timing_ = false;
    KernelTimer timer_;
};

}  // namespace wsim
