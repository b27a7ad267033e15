// water/wsim.hpp -- header-only C++ surface over the libwsim C ABI.
//
// Mirrors the reference's C++ API (namespaces, class and member names,
// argument meaning and exception types) so code written against
// Qervas/water_simulation compiles against libwsim:
//   water::ParticleStore            core/include/water/core/particle_store.h:48-107
//   water::CellGrid                 core/include/water/core/cell_grid.h:16-50
//   water::SpatialAccel             core/include/water/core/spatial_accel.h:19-53
//   water::TimeStepper              core/include/water/core/timestep.h:8-34
//   water::solvers::DFSPHSolver     solvers/dfsph/include/water/solvers/dfsph.h:17-84
//   water::solvers::sample_aabb_boundary  boundary_sampler.h:24-25
//   water::scene::load_scene        scene/include/water/scene/scene.h:45
// Status codes from the C ABI are rethrown as std::out_of_range /
// std::invalid_argument / std::runtime_error.
#pragma once

#include <wsim.h>

#include <cstdint>
#include <stdexcept>
#include <string>
#include <string_view>
#include <vector>

namespace water {

using u8 = std::uint8_t;
using u16 = std::uint16_t;
using u32 = std::uint32_t;
using u64 = std::uint64_t;
using i8 = std::int8_t;
using i16 = std::int16_t;
using i32 = std::int32_t;
using i64 = std::int64_t;
using f32 = float;
using f64 = double;

struct Vec3f { f32 x, y, z; };
struct Vec3i { i32 x, y, z; };
struct Vec4f { f32 x, y, z, w; };
struct Vec3u { u32 x, y, z; };
inline constexpr Vec3f kZero3f{0.0f, 0.0f, 0.0f};
static_assert(sizeof(Vec3f) == 12, "positions are 12-byte AoS records");

using Stream = void*;  // hipStream_t

namespace detail {
inline void check(wsim_status s) {
    if (s == WSIM_OK) return;
    const std::string m = wsim_last_error();
    if (s == WSIM_ERR_OUT_OF_RANGE) throw std::out_of_range(m);
    if (s == WSIM_ERR_INVALID_ARGUMENT) throw std::invalid_argument(m);
    throw std::runtime_error(m);
}
}  // namespace detail

enum class AttribType { F32, F32x3, U32 };

template <typename T> class AttribHandle {
public:
    AttribHandle() = default;
    explicit AttribHandle(std::size_t index) : index_(index) {}
    std::size_t index() const noexcept { return index_; }
    bool valid() const noexcept { return index_ != npos; }

private:
    static constexpr std::size_t npos = static_cast<std::size_t>(-1);
    std::size_t index_ = npos;
};

class ParticleStore {
public:
    explicit ParticleStore(std::size_t capacity, Stream stream = nullptr) {
        detail::check(wsim_store_create(capacity, stream, &h_));
        stream_ = stream;
    }
    ~ParticleStore() { wsim_store_destroy(h_); }
    ParticleStore(const ParticleStore&) = delete;
    ParticleStore& operator=(const ParticleStore&) = delete;
    ParticleStore(ParticleStore&& o) noexcept : h_(o.h_), stream_(o.stream_) { o.h_ = nullptr; }
    ParticleStore& operator=(ParticleStore&& o) noexcept {
        if (this != &o) {
            wsim_store_destroy(h_);
            h_ = o.h_;
            stream_ = o.stream_;
            o.h_ = nullptr;
        }
        return *this;
    }

    AttribHandle<Vec3f> position_handle() const noexcept {
        return AttribHandle<Vec3f>(wsim_store_position_handle(h_));
    }
    AttribHandle<Vec3f> velocity_handle() const noexcept {
        return AttribHandle<Vec3f>(wsim_store_velocity_handle(h_));
    }
    Vec3f* positions() { return attribute_data(position_handle()); }
    Vec3f* velocities() { return attribute_data(velocity_handle()); }
    const Vec3f* positions() const { return const_cast<ParticleStore*>(this)->positions(); }
    const Vec3f* velocities() const { return const_cast<ParticleStore*>(this)->velocities(); }

    template <typename T> AttribHandle<T> register_attribute(std::string_view name, AttribType type) {
        std::size_t idx = 0;
        const std::string s{name};
        detail::check(wsim_store_register_attribute(h_, s.c_str(),
                                                    static_cast<wsim_attrib_type>(type), &idx));
        return AttribHandle<T>(idx);
    }
    template <typename T> T* attribute_data(AttribHandle<T> h) {
        if (!h.valid()) throw std::out_of_range("ParticleStore: invalid handle");
        void* p = nullptr;
        detail::check(wsim_store_attribute_data(h_, h.index(), &p));
        return static_cast<T*>(p);
    }
    template <typename T> const T* attribute_data(AttribHandle<T> h) const {
        return const_cast<ParticleStore*>(this)->attribute_data(h);
    }
    bool has_attribute(std::string_view name) const {
        const std::string s{name};
        return wsim_store_has_attribute(h_, s.c_str()) != 0;
    }
    void resize(std::size_t count) { detail::check(wsim_store_resize(h_, count)); }
    std::size_t count() const noexcept { return wsim_store_count(h_); }
    std::size_t capacity() const noexcept { return wsim_store_capacity(h_); }
    Stream stream() const noexcept { return stream_; }
    wsim_store native() const noexcept { return h_; }

private:
    wsim_store h_ = nullptr;
    Stream stream_ = nullptr;
};

class CellGrid {
public:
    CellGrid(Vec3f origin, Vec3i cells_per_axis, f32 cell_length, Stream stream = nullptr) {
        const float o[3] = {origin.x, origin.y, origin.z};
        const int32_t d[3] = {cells_per_axis.x, cells_per_axis.y, cells_per_axis.z};
        detail::check(wsim_grid_create(o, d, cell_length, stream, &h_));
        stream_ = stream;
    }
    ~CellGrid() {
        if (owned_) wsim_grid_destroy(h_);
    }
    CellGrid(const CellGrid&) = delete;
    CellGrid& operator=(const CellGrid&) = delete;

    void reserve(std::size_t capacity) { detail::check(wsim_grid_reserve(h_, capacity)); }
    void build(const Vec3f* positions, std::size_t count) {
        detail::check(wsim_grid_build(h_, reinterpret_cast<const float*>(positions), count));
    }
    const u32* cell_start() const noexcept { return wsim_grid_cell_start(h_); }
    const u32* sorted_indices() const noexcept { return wsim_grid_sorted_indices(h_); }
    Vec3f origin() const noexcept {
        float o[3];
        wsim_grid_origin(h_, o);
        return {o[0], o[1], o[2]};
    }
    Vec3i cells_per_axis() const noexcept {
        int32_t d[3];
        wsim_grid_cells_per_axis(h_, d);
        return {d[0], d[1], d[2]};
    }
    f32 cell_length() const noexcept { return wsim_grid_cell_length(h_); }
    u32 total_cells() const noexcept { return wsim_grid_total_cells(h_); }
    Stream stream() const noexcept { return stream_; }
    wsim_grid native() const noexcept { return h_; }

private:
    wsim_grid h_ = nullptr;
    Stream stream_ = nullptr;
    bool owned_ = true;
};

struct AABB {
    Vec3f min;
    Vec3f max;
};

class SpatialAccel {
public:
    SpatialAccel() { detail::check(wsim_accel_create(&h_)); }
    ~SpatialAccel() { wsim_accel_destroy(h_); }
    SpatialAccel(const SpatialAccel&) = delete;
    SpatialAccel& operator=(const SpatialAccel&) = delete;
    void build(const Vec3f* positions, std::size_t count, AABB scene_bounds, Stream stream = nullptr) {
        const float lo[3] = {scene_bounds.min.x, scene_bounds.min.y, scene_bounds.min.z};
        const float hi[3] = {scene_bounds.max.x, scene_bounds.max.y, scene_bounds.max.z};
        detail::check(wsim_accel_build(h_, reinterpret_cast<const float*>(positions), count, lo,
                                       hi, stream));
    }
    std::size_t leaf_count() const noexcept { return wsim_accel_leaf_count(h_); }
    const u32* sorted_indices() const noexcept { return wsim_accel_sorted_indices(h_); }
    const u32* morton_codes() const noexcept { return wsim_accel_morton_codes(h_); }
    const u32* internal_left() const noexcept { return wsim_accel_internal_left(h_); }
    const u32* internal_right() const noexcept { return wsim_accel_internal_right(h_); }
    const AABB* node_aabbs() const noexcept {
        return reinterpret_cast<const AABB*>(wsim_accel_node_aabbs(h_));
    }

private:
    wsim_accel h_ = nullptr;
};

class TimeStepper {
public:
    struct Config {
        f32 frame_dt = 1.0f / 24.0f;
        f32 dt_max = 1.0f / 24.0f;
        f32 cfl_lambda = 0.4f;
        u32 max_substeps = 16;
    };
    TimeStepper() = default;
    explicit TimeStepper(Config c) : cfg_(c) {}
    f32 next_dt(f32 max_velocity, f32 particle_radius, f32 remaining_in_frame) const {
        const wsim_timestep_config c{cfg_.frame_dt, cfg_.dt_max, cfg_.cfl_lambda, cfg_.max_substeps};
        return wsim_timestep_next_dt(&c, max_velocity, particle_radius, remaining_in_frame);
    }
    Config config() const noexcept { return cfg_; }

private:
    Config cfg_;
};

// Asynchronous frame_%04d.bin output (wsim_frame_writer_*); not in the
// reference, whose sim_cli writes frames synchronously.
class FrameWriter {
public:
    FrameWriter(const std::string& dir, std::size_t max_points) {
        detail::check(wsim_frame_writer_create(dir.c_str(), max_points, &h_));
    }
    ~FrameWriter() { wsim_frame_writer_destroy(h_); }
    FrameWriter(const FrameWriter&) = delete;
    FrameWriter& operator=(const FrameWriter&) = delete;
    void push(ParticleStore& store, int index, void* stream) {
        detail::check(wsim_frame_writer_push(h_, store.native(), index, stream));
    }
    void finish() { detail::check(wsim_frame_writer_finish(h_)); }

private:
    wsim_frame_writer h_ = nullptr;
};

// Surface mesh of device-resident positions: the splat method of
// tools/particles_to_mesh.py (reconstruct_splat :77-138) on the device.
class SurfaceMesher {
public:
    struct Params {
        double particle_radius = 0.012;
        double grid_voxel_factor = 0.5;
        double smooth_sigma = 2.5;
        double iso_factor = 0.18;
        bool swap_yz = true;               // sim Y-up -> Blender Z-up (load_frame :46-58)
        const double* weights = nullptr;   // 2R + 1 blur weights, R = int(4 sigma + 0.5)
        int smooth_pre_iters = 2;          // k-NN position smoothing passes
        int k_neighbors = 15;              // 8, 10, 15, 20 or 30
        int smooth_post_iters = 2;         // Laplacian mesh smoothing passes
    };
    explicit SurfaceMesher(void* stream = nullptr) { detail::check(wsim_mesher_create(stream, &h_)); }
    ~SurfaceMesher() { wsim_mesher_destroy(h_); }
    SurfaceMesher(const SurfaceMesher&) = delete;
    SurfaceMesher& operator=(const SurfaceMesher&) = delete;
    // Synchronous; vertices() / triangles() stay valid until the next run.
    wsim_mesh_info run(const float* positions_device, std::size_t count, const Params& p) {
        wsim_mesh_params c;
        wsim_mesh_params_default(&c);
        c.particle_radius = p.particle_radius;
        c.grid_voxel_factor = p.grid_voxel_factor;
        c.smooth_sigma = p.smooth_sigma;
        c.iso_factor = p.iso_factor;
        c.swap_yz = p.swap_yz ? 1u : 0u;
        c.weights = p.weights;
        c.smooth_pre_iters = p.smooth_pre_iters;
        c.k_neighbors = p.k_neighbors;
        c.smooth_post_iters = p.smooth_post_iters;
        detail::check(wsim_mesher_run(h_, positions_device, count, &c));
        wsim_mesh_info info;
        detail::check(wsim_mesher_info(h_, &info));
        return info;
    }
    wsim_mesh_info run(ParticleStore& store, const Params& p) {
        return run(reinterpret_cast<const float*>(store.positions()), store.count(), p);
    }
    const float* vertices() const {  // device, n_vertices x 3
        const float* v = nullptr;
        detail::check(wsim_mesher_vertices(h_, &v));
        return v;
    }
    const u32* triangles() const {  // device, n_triangles x 3
        const uint32_t* t = nullptr;
        detail::check(wsim_mesher_triangles(h_, &t));
        return t;
    }

private:
    wsim_mesher h_ = nullptr;
};

namespace solvers {

class DFSPHSolver {
public:
    struct Config {
        f32 rest_density = 1000.0f;
        f32 particle_radius = 0.0015f;
        f32 smoothing_length = 0.003f;
        f32 viscosity = 1e-3f;
        f32 damping = 0.5f;
        f32 surface_tension = 0.0f;
        Vec3f gravity{0.0f, -9.81f, 0.0f};
        u32 max_density_iters = 100;
        u32 max_divergence_iters = 100;
        f32 eta_density = 1e-3f;
        f32 eta_divergence = 1e-3f;
        Vec3f domain_min{0.0f, 0.0f, 0.0f};
        Vec3f domain_max{1.0f, 1.0f, 1.0f};
        bool enable_divergence_solve = false;  // extension
        u32 pressure_form = 0;                 // extension: 0 kappa/rho^2, 1 kappa/rho
        bool enable_surface_tension = false;   // extension: Akinci 2013 with surface_tension
        bool fast_math = false;                // extension: tolerance-mode pair arithmetic
    };

    DFSPHSolver(ParticleStore& fluid, ParticleStore* boundary, Config cfg, Stream stream = nullptr) {
        wsim_dfsph_config c;
        wsim_dfsph_config_default(&c);
        c.rest_density = cfg.rest_density;
        c.particle_radius = cfg.particle_radius;
        c.smoothing_length = cfg.smoothing_length;
        c.viscosity = cfg.viscosity;
        c.damping = cfg.damping;
        c.surface_tension = cfg.surface_tension;
        c.gravity[0] = cfg.gravity.x;
        c.gravity[1] = cfg.gravity.y;
        c.gravity[2] = cfg.gravity.z;
        c.max_density_iters = cfg.max_density_iters;
        c.max_divergence_iters = cfg.max_divergence_iters;
        c.eta_density = cfg.eta_density;
        c.eta_divergence = cfg.eta_divergence;
        c.domain_min[0] = cfg.domain_min.x;
        c.domain_min[1] = cfg.domain_min.y;
        c.domain_min[2] = cfg.domain_min.z;
        c.domain_max[0] = cfg.domain_max.x;
        c.domain_max[1] = cfg.domain_max.y;
        c.domain_max[2] = cfg.domain_max.z;
        c.enable_divergence_solve = cfg.enable_divergence_solve ? 1u : 0u;
        c.pressure_form = cfg.pressure_form;
        c.enable_surface_tension = cfg.enable_surface_tension ? 1u : 0u;
        c.fast_math = cfg.fast_math ? 1u : 0u;
        detail::check(wsim_solver_create(fluid.native(), boundary ? boundary->native() : nullptr,
                                         &c, stream, &h_));
    }
    ~DFSPHSolver() { wsim_solver_destroy(h_); }
    DFSPHSolver(const DFSPHSolver&) = delete;
    DFSPHSolver& operator=(const DFSPHSolver&) = delete;

    void step(f32 dt) { detail::check(wsim_solver_step(h_, dt)); }
    // sim_cli's substep loop (main.cpp:171-178) in the library: substeps of
    // dt = ts.next_dt(max(max_velocity(), v_floor), r, remaining) until `time`
    // is used up (to 1e-6) or max_substeps ran; returns the substep count.
    // Not in the reference's class.
    u32 advance(const TimeStepper& ts, f32 v_floor, f32 time, u32 max_substeps) {
        const TimeStepper::Config t = ts.config();
        const wsim_timestep_config c{t.frame_dt, t.dt_max, t.cfl_lambda, t.max_substeps};
        u32 n = 0;
        detail::check(wsim_solver_advance(h_, &c, v_floor, time, max_substeps, nullptr, nullptr, &n));
        return n;
    }
    void* stream() const noexcept { return wsim_solver_stream(h_); }
    f32 max_velocity() const noexcept { return wsim_solver_max_velocity(h_); }
    f32 particle_mass() const noexcept { return wsim_solver_particle_mass(h_); }
    u32 last_density_iters() const noexcept { return wsim_solver_last_density_iters(h_); }
    f32 density_at(std::size_t i) const { return at(&wsim_solver_density_at, i); }
    f32 alpha_at(std::size_t i) const { return at(&wsim_solver_alpha_at, i); }
    f32 boundary_mass_at(std::size_t i) const { return at(&wsim_solver_boundary_mass_at, i); }
    wsim_step_stats stats() const {
        wsim_step_stats s;
        detail::check(wsim_solver_stats(h_, &s));
        return s;
    }
    // Borrowed C handles of the solver's grids (fluid_grid()/boundary_grid()).
    wsim_grid fluid_grid_handle() const noexcept { return wsim_solver_fluid_grid(h_); }
    wsim_grid boundary_grid_handle() const noexcept { return wsim_solver_boundary_grid(h_); }
    wsim_solver native() const noexcept { return h_; }

private:
    f32 at(wsim_status (*fn)(wsim_solver, size_t, float*), std::size_t i) const {
        float v = 0.0f;
        detail::check(fn(h_, i, &v));
        return v;
    }
    wsim_solver h_ = nullptr;
};

inline std::vector<Vec3f> sample_aabb_boundary(Vec3f lo, Vec3f hi, f32 spacing, int n_layers = 3) {
    const float l[3] = {lo.x, lo.y, lo.z};
    const float h[3] = {hi.x, hi.y, hi.z};
    size_t n = 0;
    detail::check(wsim_sample_aabb_boundary(l, h, spacing, n_layers, nullptr, 0, &n));
    std::vector<Vec3f> pts(n);
    detail::check(wsim_sample_aabb_boundary(l, h, spacing, n_layers,
                                            reinterpret_cast<float*>(pts.data()), n, &n));
    return pts;
}

}  // namespace solvers

namespace scene {

struct Camera {
    Vec3f position{0.0f, 0.0f, 1.0f};
    Vec3f look_at{0.0f, 0.0f, 0.0f};
    Vec3f up{0.0f, 1.0f, 0.0f};
    f32 fov_y_deg = 30.0f;
};
struct OutputCfg {
    i32 width = 1920;
    i32 height = 1080;
    f32 fps = 24.0f;
    i32 frame_start = 0;
    i32 frame_end = 120;
    std::string format = "exr";
    std::string video = "mp4";
};
struct FluidCfg {
    std::string solver = "dfsph";
    f32 particle_radius = 0.0015f;
    f32 rest_density = 1000.0f;
    f32 viscosity = 1e-3f;
    f32 surface_tension = 0.0728f;
    Vec3f gravity{0.0f, -9.81f, 0.0f};
    Vec3f initial_block_min{0.4f, 0.05f, 0.4f};
    Vec3f initial_block_max{0.6f, 0.25f, 0.6f};
};
struct Scene {
    std::string name;
    OutputCfg output;
    Camera camera;
    FluidCfg fluid;
};

inline Scene load_scene(const std::string& path) {
    wsim_scene c;
    detail::check(wsim_scene_load(path.c_str(), &c));
    auto v3 = [](const float* a) { return Vec3f{a[0], a[1], a[2]}; };
    Scene s;
    s.name = c.name;
    s.output.width = c.width;
    s.output.height = c.height;
    s.output.fps = c.fps;
    s.output.frame_start = c.frame_start;
    s.output.frame_end = c.frame_end;
    s.output.format = c.format;
    s.output.video = c.video;
    s.camera.position = v3(c.cam_position);
    s.camera.look_at = v3(c.cam_look_at);
    s.camera.up = v3(c.cam_up);
    s.camera.fov_y_deg = c.cam_fov_y_deg;
    s.fluid.solver = c.solver;
    s.fluid.particle_radius = c.particle_radius;
    s.fluid.rest_density = c.rest_density;
    s.fluid.viscosity = c.viscosity;
    s.fluid.surface_tension = c.surface_tension;
    s.fluid.gravity = v3(c.gravity);
    s.fluid.initial_block_min = v3(c.initial_block_min);
    s.fluid.initial_block_max = v3(c.initial_block_max);
    return s;
}

}  // namespace scene
}  // namespace water
