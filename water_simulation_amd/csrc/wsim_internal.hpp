// wsim_internal.hpp -- the C++ objects behind the C ABI handles.
#pragma once

#include <array>
#include <functional>
#include <map>
#include <memory>
#include <string>
#include <unordered_map>
#include <vector>

#include "common.hpp"
#include "device_math.hpp"
#include "host_surface.hpp"

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
namespace dev {
struct Nbr;  // kernel view of the grids and stores (stencil.hpp)

// Per-workgroup partials of the substep statistics, reduced per scene after
// advection: max |v|, non-finite count, escapee count, and kMomParts
// binary64 sums (mass, momentum x, y, z, kinetic energy) per workgroup.
constexpr int kMomParts = 5;
struct VStats {
    float* pmax;
    u32* pnan;
    u32* pesc;
    double* pmom;
};
}

class CellGrid {
public:
    CellGrid(const float origin[3], const int32_t dims[3], float cell_length, hipStream_t stream);
    void reserve(size_t capacity);
    void build(const float* positions_xyz, size_t count);

    // Scene-segmented grid: particles [starts[s], starts[s+1]) (device array
    // of scenes+1 offsets) belong to scene s and take the key s*(C+1) + key,
    // scene-local sentinel included, so one sort orders every scene's cells
    // and cell_start holds scenes*(C+1)+1 entries.  One scene is build().
    void reserve_segments(u32 scenes);
    void build_segmented(const float* positions_xyz, size_t count, const u32* starts,
                         u32 scenes);
    u32 segments() const { return segments_; }
    // Cells per thread of the cell_start pass (4, 8 or 16); any value gives
    // the same table.
    void set_cells_per_thread(u32 cpt);
    u32 cells_per_thread() const { return cpt_; }

    // Keys and original indices written by the caller into staged_keys() and
    // staged_values() (count <= capacity() entries each), then sorted and
    // tabled as build() does with the keys it computes itself.
    u32* staged_keys() { return k_[0].data(); }
    u32* staged_values() { return v_[0].data(); }
    // The staged keys' kernel calls advance_span(span()) (one grid only).
    // With first_counted it has also counted the sort's first digit into
    // sort_hist() (radix_count_*, tile t of the keys by workgroup t, digit
    // mask first_digit_mask()), and the sort starts at its scan.
    u32* sort_hist() { return hist_.data(); }
    u32 first_digit_mask() const;
    void build_staged(size_t count, bool first_counted = false);

    // Rewrite only the cells of the cell_start table that can change: for
    // c <= min(smallest key) every table holds 0 and for c > max(largest
    // key) every table holds n, so with the same n only the cells between
    // the previous and current key spans are written.  For a grid rebuilt
    // every substep from few-cell moves (one scene; the span record lives on
    // the device, so graph replays keep it).
    void set_track_span(bool on);
    u32* span() { return track_span_ ? span_.data() : nullptr; }

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
    DeviceBuffer<u32> k_[2], v_[2], hist_, cs_, span_;
    bool track_span_ = false;
    void reset_span();
    int cur_ = 0;
    u32 segments_ = 1;
    u32 cpt_ = 16;
};

// Plain 16-byte streaming copy kernel (copy.hip).
void device_copy(void* dst, const void* src, size_t bytes, hipStream_t stream);

// CellGrid::build into caller arrays (sorted keys and indices: count
// entries; cell_start: C + 1) with its scratch in a fixed per-device array:
// allocates nothing.  At most 2^21 particles; calls that overlap on one
// device race on the scratch (grid_noalloc.hip).
void cell_sort_noalloc(const float* xyz, size_t count, const float origin[3],
                       const int32_t dims[3], float cell_length, u32* sorted_keys,
                       u32* sorted_indices, u32* cell_start, hipStream_t stream);

// Fluid surface mesh from particle positions (surface.hip): splat, Gaussian
// blur, iso level, marching cubes (tools/particles_to_mesh.py:77-138).
struct MeshParams {
    double particle_radius = 0.012;   // --radius
    double grid_voxel_factor = 0.5;   // voxel = factor * radius
    double smooth_sigma = 2.5;        // blur sigma in voxels (0: none)
    double iso_factor = 0.18;         // iso = factor * max density
    bool swap_yz = true;              // sim Y-up -> Blender Z-up (load_frame)
    const double* weights = nullptr;  // the 2R+1 blur weights, R = int(4 sigma + 0.5); NULL: exp()
    int smooth_pre_iters = 2;         // k-NN position smoothing passes (--smooth-pre)
    int k_neighbors = 15;             // --k-neighbors: 8, 10, 15, 20 or 30
    int smooth_post_iters = 2;        // Laplacian mesh smoothing passes (--smooth)
};
struct MeshInfo {
    int dims[3];
    double lo[3], extent[3], spacing[3];
    float max_density, iso;
    u64 n_vertices, n_triangles;
};
class SurfaceMesher {
public:
    explicit SurfaceMesher(hipStream_t stream);
    // positions: device binary32 xyz, n points
    void run(const float* xyz, size_t n, const MeshParams& p);
    const MeshInfo& info() const { return info_; }
    const float* grid() const { return grid_.data(); }      // blurred density, dims[0..2] C order
    const float* vertices() const { return verts_.data(); }  // n_vertices x 3
    const u32* triangles() const { return tris_.data(); }    // n_triangles x 3 vertex indices

private:
    hipStream_t stream_;
    MeshInfo info_{};
    DeviceBuffer<float> grid_, tmp_, verts_, part_;
    DeviceBuffer<u32> tris_, count_, eoff_;
    DeviceBuffer<u32> segE_, segT_, segEoff_, segToff_;  // per 64-point row segment
    DeviceBuffer<uint8_t> emask_;
    DeviceBuffer<double> pts_[2];  // binary64 points, smoothed in place
    DeviceBuffer<u32> knn_keys_[2], knn_vals_[2], knn_hist_, knn_cs_;
    DeviceBuffer<double> knn_ps_;       // points in cell order
    DeviceBuffer<u32> knn_far_, knn_farc_;  // points left to the wave-per-point search
    DeviceBuffer<u32> adj_, deg_, adj_over_;  // vertex neighbours for the Laplacian passes
    DeviceBuffer<u32> vpt_;                   // vertex -> grid point << 2 | axis
    DeviceBuffer<double> vd_[2];
};
// The marching-cubes table: triangle counts [256] and edge ids [256][24].
const uint8_t* mc_table_counts();
const uint8_t* mc_table_triangles();

// Exhaustive comparison of the pair-path sqrt and quotients with the IEEE
// operations for one smoothing length (pair_ops.hip); bit ranges are the
// binary32 patterns covered ([lo, hi)).
struct PairOpsReport {
    int fast;  // Spline::make(l).fast: the solver kernels use the pair-path operations
    u64 sqrt_mismatches, quotient_mismatches, gradient_mismatches;
    u32 x_lo, x_hi, d_lo, d_hi, c_lo;
};
PairOpsReport pair_ops_check(float l);

// Cell-start table from sorted keys (one pass over the C+1 cells).  With
// `span` (a CellGrid's key-span record, see CellGrid::set_track_span) only
// the cells that can differ from the previous table of the same record are
// written.
void build_cell_start(const u32* sorted_keys, u32 n, u32 num_cells_plus_one, u32* cell_start,
                      hipStream_t stream, u32 cells_per_thread = 16, u32* span = nullptr);

// Key-span record of a cell table: two slots of (smallest key, largest key,
// particle count, valid), the slot of the current build chosen by the parity
// of a build counter that the build's key kernel advances (kSpanCounter).
constexpr u32 kSpanWords = 9;
constexpr u32 kSpanCounter = 8;
__device__ inline void advance_span(u32* span) {
    if (span && blockIdx.x == 0 && threadIdx.x == 0) span[kSpanCounter] += 1u;
}

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
    // Linear BVH (spatial_accel.h:34-53): internal node i in [0, n-1) has
    // children left/right (>= n - 1: leaf n-1+k, the k-th sorted particle);
    // parent of every node (root: ~0u); boxes are 6 floats (min, max) per
    // node, internal nodes first, then leaves.
    const u32* internal_left() const { return leaf_count_ > 1 ? left_.data() : nullptr; }
    const u32* internal_right() const { return leaf_count_ > 1 ? right_.data() : nullptr; }
    const u32* internal_parent() const { return leaf_count_ ? parent_.data() : nullptr; }
    const float* node_aabbs() const { return leaf_count_ ? boxes_.data() : nullptr; }

private:
    size_t leaf_count_ = 0;
    DeviceBuffer<u32> k_[2], v_[2], hist_;
    DeviceBuffer<u32> left_, right_, parent_;
    DeviceBuffer<uint2> range_;      // leaf range [lo, hi] of every internal node
    DeviceBuffer<float> boxes_, seg_;  // node boxes; segment tree over the leaf boxes
    DeviceBuffer<u64> seg_off_;
    std::vector<u64> seg_off_host_;
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
    u32 enable_surface_tension = 0;  // Akinci 2013 cohesion + curvature, gamma = surface_tension
    // Tolerance mode of the neighbour kernels (PairMath<true>): hardware
    // rsqrt/rcp and fma contraction on the pair terms, as the reference's
    // -use_fast_math build; 0 keeps every field bit-exact to the oracle.
    u32 fast_math = 0;
};

struct StepStats {
    u32 density_iters = 0;
    u32 divergence_iters = 0;
    double mean_density_error = 0.0;
    double mean_divergence_error = 0.0;
    float max_velocity = 0.0f;
    u32 nan_count = 0;
    u64 substeps = 0;
    // after the substep, binary64 over the particles with a finite state:
    u32 escapees = 0;           // positions on or outside the domain clamp box
    double total_mass = 0.0;    // m * (particles with finite position and velocity)
    double momentum[3] = {0.0, 0.0, 0.0};  // sum m v
    double kinetic_energy = 0.0;           // sum m |v|^2 / 2
};

// Working-set slots are grouped in tiles of kTile; each scene starts on a
// tile boundary, so every tile (and every wave) belongs to one scene and the
// binary64 error partials are per tile whatever the workgroup size.
constexpr u32 kTile = 256;

// One scene of the packed working set.  The static part is set when the
// scene layout is fixed, the per-substep part before every substep; the
// device copy is read with scalar loads by every kernel.
struct SceneSeg {
    u32 off;       // first working-set slot (multiple of kTile)
    u32 n;         // fluid particles
    u32 start;     // first original index == first position in the sorted order
    u32 tile0;     // off / kTile
    u32 ntiles;    // ceil(n / kTile)
    u32 key_base;  // scene * (C + 1)
    u32 pad0, pad1;
    float dt, dt2, decay;  // dt, dt*dt, exp(-damping*dt)
    float gdx, gdy, gdz;   // gravity * dt
    float eps;             // XSPH viscosity
    float pad2;
};
static_assert(sizeof(SceneSeg) == 64, "SceneSeg layout");

// Surface-tension constants: support L = 2l, cohesion normalisation
// 32 / (pi L^9), L^6 / 64, coefficient gamma, rest density.
struct TensionArgs {
    float L, sig, c6, gamma, rho0;
};

// Device-resident solve state of one scene; kernels keep it current, the
// host reads it once per substep (and between iteration chunks).
struct DeviceState {
    u32 conv_iter;      // density solve: iteration whose error test passed, or ~0u
    u32 div_conv_iter;  // divergence solve: same
    u32 nan_count;
    float max_v;
    double err;         // mean error of the latest executed density iteration
    double div_err;     // same for the divergence solve
    u32 advected;       // this substep's advection has run for the scene
    u32 escapees;       // see StepStats
    double mass;
    double mom[3];
    double ke;
    float t_rem;        // advance(): time left in the frame after this substep
    float dt;           // advance(): this substep's CFL time step
};

// The time-step rule of an advance() call, read by the device CFL step
// (timestep.h:8-34 and sim_cli's velocity floor, main.cpp:176-177).
struct FrameParams {
    float frame_dt, dt_max, cfl_lambda;
    u32 max_substeps;
    float v_floor, particle_radius, damping;
    float gravity[3];
    float time;  // the frame's time budget
};

// A substep's state as the device publishes it to mapped host memory: the
// state, then (released after it) a sequence number the host waits for.
struct StatePub {
    DeviceState st;
    u32 seq;
    u32 pad[15];
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
    kKTension,
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

// LDS staging plan of one 256-slot tile (see stencil.hpp): the tile's fluid
// neighbours are the working slots [s0, s0 + b1), [s1, s1 + b2 - b1) and
// [s2, s2 + total - b2), staged at LDS entries 0, b1 and b2.
struct StageRanges {
    u32 s0, s1, s2;  // first working slot of each range
    u32 b1, b2;      // first LDS entry of ranges 1 and 2 (range 0 starts at 0)
    u32 total;       // staged entries; 0: the tile gathers from global memory
};

// Per-particle pair lists of the pressure solves (see solver.hip).
struct PairList {
    uint2* f;            // fluid pairs (slot, bits of c), wave-blocked ELL
    uint2* b;            // boundary pairs
    u32* cnt;            // fluid count | boundary count << 16 (saturating)
    uint16_t* fi;        // fluid pairs as LDS entries of the tile's staging plan (staged tiles)
    const StageRanges* plan;  // one per 256-slot tile
    u32 kf, kb;          // list depths
    // Tolerance mode: per slot, sum_b m_b grad W_ib over the particle's
    // boundary neighbours (positions are frozen through the solves, so the
    // boundary terms of every pressure sweep are v_i . Gb_i and
    // (k_i + k_b) Gb_i); no boundary pair lists are kept then.
    float4* gb;
};

// Depth of the per-particle pair lists (pairs with a non-zero kernel
// gradient); a particle with more pairs walks the stencil instead.
constexpr u32 kPairDepthFluid = 48;
constexpr u32 kPairDepthBoundary = 48;

// Launch shapes of the stencil kernels.  Every choice computes the same bits;
// the autotuner times them on the solver's own data and keeps the fastest.
struct KernelShapes {
    u32 density_alpha_nt = 256;  // workgroup size: 64, 128, 256, 512
    u32 xsph_nt = 0;             // 0 (LDS-staged neighbourhood, 256), 64, 128, 256, 512
    u32 change_nt = 0;           // 0 (LDS-staged neighbourhood, 256), 256, 512, 1024
    u32 change_pu = 4;           // pairs per trip of the list loops: 2, 4, 8
    u32 apply_nt = 0;            // 0 (LDS-staged), 64, 128, 256, 512
    u32 apply_pu = 4;            // 2, 4, 8
    u32 cell_cpt = 16;           // cells per thread of the cell_start pass: 4, 8, 16
    u32 remap = 1;               // XCD-aware workgroup -> tile order
};
bool valid_shapes(const KernelShapes& s);

// Timing of one autotuner candidate (milliseconds per launch, median).
struct TuneRecord {
    int kernel;  // KernelKind (kKGrid stands for the cell_start pass)
    u32 nt, pu, remap, cpt;
    double ms;
};

// Forward record of one substep for the adjoint: the input state and every
// pressure iterate.  Grids, pair lists, rho and alpha are recomputed from the
// input state, which reproduces them bit for bit.
struct SubstepTape {
    DeviceBuffer<float> x, v;  // input state, original order (3n)
    DeviceBuffer<float4> w;    // K+1 velocity slabs: w[k] enters iteration k, w[K] advection
    DeviceBuffer<float4> xk;   // K slabs of (x, kappa-hat)
    DeviceBuffer<float> p;     // K slabs of rho*
    float dt = 0.0f;
    u32 iters = 0, n = 0, nslots = 0;
};

// Adjoint variables in working-slot order (binary64).
struct AdjointWork {
    DeviceBuffer<double4> xbar, wbar, v0bar, gbar;
    DeviceBuffer<double> rdbar, rhobar, alphabar, part;
    PinnedBuffer<double> hpart;
};

// DFSPH substep over one or many independent scenes.  A plain solver is one
// scene holding every particle of the fluid store; a batched solver packs
// scenes of uneven sizes (consecutive index ranges of the store) into one
// working set, one grid sort and one launch per phase, with per-scene dt,
// viscosity, convergence test and statistics.  Scene s of a batch produces
// the same bits as a plain solver run on that scene alone.
class DFSPHSolver {
public:
    DFSPHSolver(ParticleStore& fluid, ParticleStore* boundary, const DFSPHConfig& cfg,
                hipStream_t stream);
    // Batched: scene s holds `counts[s]` particles (store indices in order),
    // XSPH viscosity viscosity[s] (nullptr: cfg.viscosity for all).
    DFSPHSolver(ParticleStore& fluid, ParticleStore* boundary, const DFSPHConfig& cfg,
                hipStream_t stream, const std::vector<u32>& counts,
                const std::vector<float>& viscosity);
    ~DFSPHSolver();
    DFSPHSolver(const DFSPHSolver&) = delete;
    DFSPHSolver& operator=(const DFSPHSolver&) = delete;

    void step(float dt);               // every scene advances by dt
    void step_scenes(const float* dt);  // scene s advances by dt[s]
    // The CFL substep loop of sim_cli (main.cpp:171-178) for a plain solver:
    // before each substep dt = next_dt(ts, max(max_velocity(), v_floor), r,
    // remaining) (timestep.h:23-28), until the remaining time is at most
    // 1e-6 or max_substeps ran.  Writes each dt and each substep's statistics
    // to dts and stats (where not null) and returns the substep count.  Every
    // substep computes what step(dt) does.
    u32 advance(const TimeStepConfig& ts, float v_floor, float time, u32 max_substeps, float* dts,
                StepStats* stats);

    float max_velocity() const;  // over all scenes
    float particle_mass() const { return mass_; }
    u32 last_density_iters() const { return stats_[0].density_iters; }
    const StepStats& stats() const { return stats_[0]; }
    const StepStats& scene_stats(size_t s) const;
    size_t scene_count() const { return counts_.size(); }
    bool batched() const { return batched_; }
    const std::vector<u32>& scene_starts() const { return starts_; }
    const CellGrid& fluid_grid() const { return fluid_grid_; }
    const CellGrid* boundary_grid() const { return boundary_grid_.get(); }
    float attribute_at(int which, size_t i) const;  // 0 density, 1 alpha, 2 boundary mass
    const DFSPHConfig& config() const { return cfg_; }
    hipStream_t stream() const { return stream_; }
    ParticleStore& fluid_store() { return fluid_; }

    void set_timing(bool on) { timing_ = on; }
    KernelTimer& timer() { return timer_; }

    // Substep phases as captured HIP graphs (one per phase and per iteration
    // chunk), replayed while the scene layout and the shapes are unchanged.
    void set_graphs(bool on);
    bool graphs() const { return use_graphs_; }

    const KernelShapes& shapes() const { return shapes_; }
    void set_shapes(const KernelShapes& s);
    // Times every shape candidate of the stencil kernels on the current
    // particle state (solver buffers only: the store and the results of the
    // next step are untouched) and keeps the fastest per kernel.
    void autotune(int reps);
    const std::vector<TuneRecord>& tune_log() const { return tune_log_; }

    // Neighbour counts of the density/alpha walk on the store's current
    // positions, per particle in original order (device arrays, each may be
    // null): stencil candidates and in-support neighbours in the fluid and
    // the boundary store.  Synchronous; leaves the store and the results of
    // the next step unchanged.
    void neighbor_counts(u32* cand_f, u32* nbr_f, u32* cand_b, u32* nbr_b);

    // XSPH viscosity of a plain solver (the parameter the adjoint differentiates).
    void set_viscosity(float eps);
    // Plain solver only: a substep that records `tape` (eager launches).
    void step_taped(float dt, SubstepTape& tape);
    // Adjoint of the substep recorded in `tape` (solvers/dfsph/src/dfsph.cu:
    // 120-133 transposed, neighbour sets, iteration count and clamp masks of
    // the recorded pass): from d(loss)/d(output positions, velocities) to
    // d(loss)/d(input positions, velocities), all binary64 device arrays of
    // 3n in original order; returns this substep's d(loss)/d(viscosity).
    // Leaves the store at the substep's input state.
    // With visc_acc, the viscosity partials are added to visc_acc on the
    // device (adjoint_partials() entries) and 0 is returned; otherwise the
    // step waits for them and returns their sum.
    double adjoint_step(const SubstepTape& tape, const double* gx_out, const double* gv_out,
                        double* gx_in, double* gv_in, double* visc_acc = nullptr);
    size_t adjoint_partials() const;

private:
    void iterations_taped(u32 begin, u32 end, SubstepTape& t);
    void init(const std::vector<u32>& counts, const std::vector<float>& viscosity);
    void layout(const std::vector<u32>& counts);
    void set_substep(const float* dt);
    void run_substep();
    // One substep of advance() as one launch sequence: the device CFL step,
    // gravity/XSPH, the first chunk of density iterations, advection, the
    // state published to host memory and, with `next`, the following
    // substep's grid phase; an unfinished solve continues in chunks of the
    // same shape.  Returns whether the grid phase of the next substep stands.
    bool substep_device(bool next);
    void wait_published();
    void phase_prefix(bool gravity);
    void phase_prefix_grid();
    void phase_prefix_solve(bool gravity);
    dev::Nbr nbr() const;
    void phase_iterations(int mode, u32 begin, u32 end);
    void phase_advect(bool force, bool fold_finish = false);
    void write_back_sorted();  // the state in xa_/va_ (and rho, alpha, ...) to the store
    void fetch_state();
    void take_state();
    void phase_div_tail();
    void solve(int mode);
    u32 first_density_chunk() const;
    bool all_converged(int mode);
    void read_state();
    void run_phase(u64 key, const std::function<void()>& body);
    void drop_graphs();
    PairList pair_list();
    dev::VStats vstats();
    void tbegin(int kind, int iter = 0) {
        if (timing_ && !capturing_) timer_.begin(kind, iter, stream_);
    }
    void tend() {
        if (timing_ && !capturing_) timer_.end(stream_);
    }

    ParticleStore& fluid_;
    ParticleStore* boundary_;
    DFSPHConfig cfg_;
    hipStream_t own_stream_ = nullptr;  // created when the caller passes the null stream
    hipStream_t stream_;
    CellGrid fluid_grid_;
    std::unique_ptr<CellGrid> boundary_grid_;
    float mass_ = 0.0f;
    size_t nb_ = 0;
    bool batched_ = false;

    // scene layout
    std::vector<u32> counts_, starts_;
    std::vector<float> viscosity_;
    u32 nslots_ = 0, total_ = 0;
    bool any_xsph_ = false;
    PinnedBuffer<SceneSeg> hseg_;
    DeviceBuffer<SceneSeg> dseg_;
    DeviceBuffer<u32> dstarts_;
    DeviceBuffer<u32> tile_scene_;
    std::vector<StepStats> stats_;

    size_t h_density_ = 0, h_alpha_ = 0, h_kappa_ = 0, h_density_adv_ = 0, h_kappa_v_ = 0;
    size_t h_bmass_ = 0;

    // cell-sorted working set (nslots_ entries, scenes tile-aligned)
    DeviceBuffer<float4> xs_, vs_, vpre_, xk_, xb_;
    DeviceBuffer<uint2> pairs_f_, pairs_b_;
    DeviceBuffer<uint16_t> pairs_fi_;
    DeviceBuffer<float4> nrm_;  // surface-tension normals (when enabled)
    DeviceBuffer<float4> gb_;   // boundary gradient sums (tolerance mode with a boundary)
    DeviceBuffer<u32> inv_;     // working slot of each original index (plain solver)
    // advance(): the advected state in the working-slot order of the substep
    // that advected it, and the cell key of each advected position
    DeviceBuffer<float4> xa_, va_;
    DeviceBuffer<u32> ka_;
    bool defer_store_ = false;   // advection writes xa_/va_, not the store
    bool state_sorted_ = false;  // the current state is in xa_/va_ (the store is stale)
    bool grid_sorted_ = false;   // the current grid phase gathered from xa_/va_
    // the grid phase being enqueued follows the first chunk of a substep
    // directly and runs its statistics step (with the chunk's force flag)
    bool grid_fold_ = false, grid_fold_force_ = false;
    DeviceBuffer<u32> bnear_;  // boundary proximity mask (see stencil.hpp)
    DeviceBuffer<StageRanges> plan_;
    DeviceBuffer<u32> pair_cnt_;
    DeviceBuffer<float> rho_, alpha_, kappa_, rho_adv_;
    DeviceBuffer<double> part_d_;
    DeviceBuffer<float> part_f_;
    DeviceBuffer<u32> part_u_, part_e_;
    DeviceBuffer<double> part_m_;
    DeviceBuffer<DeviceState> dstate_;
    PinnedBuffer<DeviceState> hstate_;

    KernelShapes shapes_;
    std::vector<TuneRecord> tune_log_;
    u32 last_iters_[2] = {0, 0};  // iterations of the previous substep's solves
    double last_err_ratio_ = 1.0;  // its density solve's final error over eta rho0 (scene max)
    DeviceBuffer<FrameParams> dframe_;  // advance(): the time-step rule on the device
    DeviceBuffer<u32> pubseq_;          // sequence number of the last published state
    PinnedBuffer<StatePub> pub_;        // published state (mapped, coherent host memory)
    StatePub* pub_dev_ = nullptr;       // device address of pub_
    u32 pub_expect_ = 0;
    u32 parity_ = 0;  // which instance of the substep graph runs next
    bool use_graphs_ = false, capturing_ = false;
    std::map<u64, hipGraphExec_t> graphs_;

    bool timing_ = false;
    KernelTimer timer_;
    AdjointWork adj_;
};

// Differentiable rollout over a plain solver: forward substeps with a state
// checkpoint every `every` substeps; backward recomputes one segment at a
// time from its checkpoint, recording every substep's tape, and runs the
// substep adjoints in reverse.
class Rollout {
public:
    Rollout(DFSPHSolver& solver, ParticleStore& fluid, u32 every);
    // x0, v0: device, 3n, original order.  Leaves the final state in the store.
    void forward(const float* x0, const float* v0, const float* dts, u32 substeps);
    // gx, gv: d(loss)/d(final positions, velocities); writes d(loss)/d(x0, v0)
    // and returns d(loss)/d(viscosity).  Device binary64 arrays of 3n.
    double backward(const double* gx, const double* gv, double* gx0, double* gv0);
    u32 substeps() const { return static_cast<u32>(dts_.size()); }

private:
    DFSPHSolver& solver_;
    ParticleStore& fluid_;
    u32 every_;
    size_t n_ = 0;
    std::vector<float> dts_;
    std::vector<DeviceBuffer<float>> ckx_, ckv_;
    std::vector<SubstepTape> tapes_;
    DeviceBuffer<double> g_[4];
    DeviceBuffer<double> visc_acc_;  // viscosity partials summed over the backward pass
};

}  // namespace wsim
