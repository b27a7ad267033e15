// capi.cpp -- the extern "C" boundary of libwsim (include/wsim.h).
//
// Every entry point converts exceptions into status codes and a thread-local
// message; nothing C++ crosses the boundary.
#include <cstring>
#include <new>
#include <string>
#include <vector>

#include "../../include/wsim.h"
#include "frames.hpp"
#include "host_surface.hpp"
#include "wsim_internal.hpp"

struct wsim_store_t {
    std::unique_ptr<wsim::ParticleStore> impl;
};
struct wsim_grid_t {
    std::unique_ptr<wsim::CellGrid> owned;
    const wsim::CellGrid* view = nullptr;
    wsim::CellGrid* mut() { return owned.get(); }
};
struct wsim_accel_t {
    wsim::SpatialAccel impl;
};
struct wsim_solver_t {
    std::unique_ptr<wsim::DFSPHSolver> impl;
    wsim_grid_t fluid_view;
    wsim_grid_t boundary_view;
    std::unique_ptr<wsim::SubstepTape> tape;  // wsim_adjoint_step's record
};
struct wsim_rollout_t {
    std::unique_ptr<wsim::Rollout> impl;
};
struct wsim_mesher_t {
    std::unique_ptr<wsim::SurfaceMesher> impl;
};

namespace {

thread_local std::string g_last_error;

template <class F> wsim_status guarded(F&& f) {
    try {
        f();
        return WSIM_OK;
    } catch (const wsim::Error& e) {
        g_last_error = e.what();
        return static_cast<wsim_status>(e.status);
    } catch (const std::bad_alloc&) {
        g_last_error = "out of host memory";
        return WSIM_ERR_RUNTIME;
    } catch (const std::exception& e) {
        g_last_error = e.what();
        return WSIM_ERR_RUNTIME;
    } catch (...) {
        g_last_error = "unknown error";
        return WSIM_ERR_RUNTIME;
    }
}

inline hipStream_t as_stream(void* s) { return reinterpret_cast<hipStream_t>(s); }

void need(const void* p, const char* what) {
    if (!p) wsim::fail(wsim::kInvalidArgument, std::string(what) + " is NULL");
}

}  // namespace

extern "C" {

const char* wsim_last_error(void) { return g_last_error.c_str(); }
int wsim_abi_version(void) { return WSIM_ABI_VERSION; }

// ---- device memory ----
wsim_status wsim_device_count(int* count) {
    return guarded([&] {
        need(count, "count");
        int n = 0;
        if (hipGetDeviceCount(&n) != hipSuccess) n = 0;
        *count = n;
    });
}
wsim_status wsim_device_alloc(size_t bytes, void** out) {
    return guarded([&] {
        need(out, "out");
        *out = nullptr;
        if (bytes) WSIM_HIP(hipMalloc(out, bytes));
    });
}
wsim_status wsim_device_free(void* ptr) {
    return guarded([&] {
        if (ptr) WSIM_HIP(hipFree(ptr));
    });
}
wsim_status wsim_memcpy_h2d(void* dst, const void* src, size_t bytes, void* stream) {
    return guarded([&] {
        if (!bytes) return;
        need(dst, "dst");
        need(src, "src");
        WSIM_HIP(hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, as_stream(stream)));
        WSIM_HIP(hipStreamSynchronize(as_stream(stream)));
    });
}
wsim_status wsim_memcpy_d2h(void* dst, const void* src, size_t bytes, void* stream) {
    return guarded([&] {
        if (!bytes) return;
        need(dst, "dst");
        need(src, "src");
        WSIM_HIP(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, as_stream(stream)));
        WSIM_HIP(hipStreamSynchronize(as_stream(stream)));
    });
}
wsim_status wsim_memcpy_d2d(void* dst, const void* src, size_t bytes, void* stream) {
    return guarded([&] {
        if (!bytes) return;
        need(dst, "dst");
        need(src, "src");
        WSIM_HIP(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToDevice, as_stream(stream)));
    });
}
wsim_status wsim_device_copy(void* dst, const void* src, size_t bytes, void* stream) {
    return guarded([&] {
        if (!bytes) return;
        need(dst, "dst");
        need(src, "src");
        wsim::device_copy(dst, src, bytes, as_stream(stream));
    });
}
wsim_status wsim_pair_ops_check(float l, uint64_t out[4]) {
    return guarded([&] {
        need(out, "out");
        const wsim::PairOpsReport r = wsim::pair_ops_check(l);
        out[0] = static_cast<uint64_t>(r.fast);
        out[1] = r.sqrt_mismatches;
        out[2] = r.quotient_mismatches;
        out[3] = r.gradient_mismatches;
    });
}
wsim_status wsim_memset(void* dst, int value, size_t bytes, void* stream) {
    return guarded([&] {
        if (!bytes) return;
        need(dst, "dst");
        WSIM_HIP(hipMemsetAsync(dst, value, bytes, as_stream(stream)));
    });
}
wsim_status wsim_stream_create(void** out_stream) {
    return guarded([&] {
        need(out_stream, "out_stream");
        hipStream_t s;
        WSIM_HIP(hipStreamCreate(&s));
        *out_stream = reinterpret_cast<void*>(s);
    });
}
wsim_status wsim_stream_destroy(void* stream) {
    return guarded([&] {
        if (stream) WSIM_HIP(hipStreamDestroy(as_stream(stream)));
    });
}
wsim_status wsim_stream_synchronize(void* stream) {
    return guarded([&] { WSIM_HIP(hipStreamSynchronize(as_stream(stream))); });
}
wsim_status wsim_device_synchronize(void) {
    return guarded([&] { WSIM_HIP(hipDeviceSynchronize()); });
}

// ---- ParticleStore ----
wsim_status wsim_store_create(size_t capacity, void* stream, wsim_store* out) {
    return guarded([&] {
        need(out, "out");
        *out = nullptr;
        auto h = std::make_unique<wsim_store_t>();
        h->impl = std::make_unique<wsim::ParticleStore>(capacity, as_stream(stream));
        *out = h.release();
    });
}
void wsim_store_destroy(wsim_store store) { delete store; }
wsim_status wsim_store_resize(wsim_store store, size_t count) {
    return guarded([&] {
        need(store, "store");
        store->impl->resize(count);
    });
}
size_t wsim_store_count(wsim_store store) { return store ? store->impl->count() : 0; }
size_t wsim_store_capacity(wsim_store store) { return store ? store->impl->capacity() : 0; }
wsim_status wsim_store_register_attribute(wsim_store store, const char* name, wsim_attrib_type type,
                                          size_t* handle) {
    return guarded([&] {
        need(store, "store");
        need(name, "name");
        const size_t h = store->impl->register_attribute(name, static_cast<int>(type));
        if (handle) *handle = h;
    });
}
int wsim_store_has_attribute(wsim_store store, const char* name) {
    return (store && name) ? (store->impl->has_attribute(name) ? 1 : 0) : 0;
}
wsim_status wsim_store_find_attribute(wsim_store store, const char* name, size_t* handle) {
    return guarded([&] {
        need(store, "store");
        need(name, "name");
        need(handle, "handle");
        *handle = store->impl->find_attribute(name);
    });
}
wsim_status wsim_store_attribute_data(wsim_store store, size_t handle, void** out) {
    return guarded([&] {
        need(store, "store");
        need(out, "out");
        *out = store->impl->attribute_data(handle);
    });
}
wsim_status wsim_store_positions(wsim_store store, float** out) {
    return guarded([&] {
        need(store, "store");
        need(out, "out");
        *out = store->impl->positions();
    });
}
wsim_status wsim_store_velocities(wsim_store store, float** out) {
    return guarded([&] {
        need(store, "store");
        need(out, "out");
        *out = store->impl->velocities();
    });
}
size_t wsim_store_position_handle(wsim_store store) {
    return store ? store->impl->position_handle() : 0;
}
size_t wsim_store_velocity_handle(wsim_store store) {
    return store ? store->impl->velocity_handle() : 0;
}

// ---- CellGrid ----
wsim_status wsim_grid_create(const float origin[3], const int32_t dims[3], float cell_length,
                             void* stream, wsim_grid* out) {
    return guarded([&] {
        need(out, "out");
        need(origin, "origin");
        need(dims, "cells_per_axis");
        *out = nullptr;
        auto h = std::make_unique<wsim_grid_t>();
        h->owned = std::make_unique<wsim::CellGrid>(origin, dims, cell_length, as_stream(stream));
        h->view = h->owned.get();
        *out = h.release();
    });
}
void wsim_grid_destroy(wsim_grid grid) {
    if (grid && grid->owned) delete grid;
}
wsim_status wsim_grid_reserve(wsim_grid grid, size_t capacity) {
    return guarded([&] {
        need(grid, "grid");
        if (!grid->mut()) wsim::fail(wsim::kInvalidArgument, "grid is a read-only solver view");
        grid->mut()->reserve(capacity);
    });
}
wsim_status wsim_grid_build(wsim_grid grid, const float* positions_xyz, size_t count) {
    return guarded([&] {
        need(grid, "grid");
        if (!grid->mut()) wsim::fail(wsim::kInvalidArgument, "grid is a read-only solver view");
        if (count) need(positions_xyz, "positions");
        grid->mut()->build(positions_xyz, count);
    });
}
wsim_status wsim_cell_sort_noalloc(const float* positions_xyz, size_t count, const float origin[3],
                                   const int32_t cells_per_axis[3], float cell_length,
                                   uint32_t* sorted_keys, uint32_t* sorted_indices,
                                   uint32_t* cell_start, void* stream) {
    return guarded([&] {
        need(origin, "origin");
        need(cells_per_axis, "cells_per_axis");
        wsim::cell_sort_noalloc(positions_xyz, count, origin, cells_per_axis, cell_length,
                                sorted_keys, sorted_indices, cell_start, as_stream(stream));
    });
}
const uint32_t* wsim_grid_cell_start(wsim_grid g) { return g ? g->view->cell_start() : nullptr; }
const uint32_t* wsim_grid_sorted_indices(wsim_grid g) {
    return (g && g->view->capacity()) ? g->view->sorted_indices() : nullptr;
}
const uint32_t* wsim_grid_sorted_keys(wsim_grid g) {
    return (g && g->view->capacity()) ? g->view->sorted_keys() : nullptr;
}
void wsim_grid_origin(wsim_grid g, float out[3]) {
    if (!g || !out) return;
    out[0] = g->view->geom().ox;
    out[1] = g->view->geom().oy;
    out[2] = g->view->geom().oz;
}
void wsim_grid_cells_per_axis(wsim_grid g, int32_t out[3]) {
    if (!g || !out) return;
    out[0] = g->view->geom().dx;
    out[1] = g->view->geom().dy;
    out[2] = g->view->geom().dz;
}
float wsim_grid_cell_length(wsim_grid g) { return g ? g->view->cell_length() : 0.0f; }
uint32_t wsim_grid_total_cells(wsim_grid g) { return g ? g->view->total_cells() : 0u; }
size_t wsim_grid_count(wsim_grid g) { return g ? g->view->count() : 0; }

// ---- SpatialAccel (Morton) ----
wsim_status wsim_accel_create(wsim_accel* out) {
    return guarded([&] {
        need(out, "out");
        *out = new wsim_accel_t();
    });
}
void wsim_accel_destroy(wsim_accel a) { delete a; }
wsim_status wsim_accel_build(wsim_accel a, const float* positions_xyz, size_t count,
                             const float bmin[3], const float bmax[3], void* stream) {
    return guarded([&] {
        need(a, "accel");
        need(bmin, "bounds_min");
        need(bmax, "bounds_max");
        if (count) need(positions_xyz, "positions");
        a->impl.build(positions_xyz, count, bmin, bmax, as_stream(stream));
    });
}
size_t wsim_accel_leaf_count(wsim_accel a) { return a ? a->impl.leaf_count() : 0; }
const uint32_t* wsim_accel_sorted_indices(wsim_accel a) {
    return a ? a->impl.sorted_indices() : nullptr;
}
const uint32_t* wsim_accel_morton_codes(wsim_accel a) { return a ? a->impl.morton_codes() : nullptr; }
const uint32_t* wsim_accel_internal_left(wsim_accel a) { return a ? a->impl.internal_left() : nullptr; }
const uint32_t* wsim_accel_internal_right(wsim_accel a) {
    return a ? a->impl.internal_right() : nullptr;
}
const uint32_t* wsim_accel_internal_parent(wsim_accel a) {
    return a ? a->impl.internal_parent() : nullptr;
}
const float* wsim_accel_node_aabbs(wsim_accel a) { return a ? a->impl.node_aabbs() : nullptr; }

// ---- DFSPHSolver ----
void wsim_dfsph_config_default(wsim_dfsph_config* c) {
    if (!c) return;
    const wsim::DFSPHConfig d;
    c->rest_density = d.rest_density;
    c->particle_radius = d.particle_radius;
    c->smoothing_length = d.smoothing_length;
    c->viscosity = d.viscosity;
    c->damping = d.damping;
    c->surface_tension = d.surface_tension;
    for (int a = 0; a < 3; ++a) {
        c->gravity[a] = d.gravity[a];
        c->domain_min[a] = d.domain_min[a];
        c->domain_max[a] = d.domain_max[a];
    }
    c->max_density_iters = d.max_density_iters;
    c->max_divergence_iters = d.max_divergence_iters;
    c->eta_density = d.eta_density;
    c->eta_divergence = d.eta_divergence;
    c->enable_divergence_solve = d.enable_divergence_solve;
    c->pressure_form = d.pressure_form;
    c->enable_surface_tension = d.enable_surface_tension;
    c->fast_math = d.fast_math;
}

static wsim::DFSPHConfig to_config(const wsim_dfsph_config* c) {
    if (c->pressure_form > 1) wsim::fail(wsim::kInvalidArgument, "pressure_form must be 0 or 1");
    if (c->fast_math > 1) wsim::fail(wsim::kInvalidArgument, "fast_math must be 0 or 1");
    wsim::DFSPHConfig d;
    d.rest_density = c->rest_density;
    d.particle_radius = c->particle_radius;
    d.smoothing_length = c->smoothing_length;
    d.viscosity = c->viscosity;
    d.damping = c->damping;
    d.surface_tension = c->surface_tension;
    for (int a = 0; a < 3; ++a) {
        d.gravity[a] = c->gravity[a];
        d.domain_min[a] = c->domain_min[a];
        d.domain_max[a] = c->domain_max[a];
    }
    d.max_density_iters = c->max_density_iters;
    d.max_divergence_iters = c->max_divergence_iters;
    d.eta_density = c->eta_density;
    d.eta_divergence = c->eta_divergence;
    d.enable_divergence_solve = c->enable_divergence_solve;
    d.pressure_form = c->pressure_form;
    d.enable_surface_tension = c->enable_surface_tension;
    d.fast_math = c->fast_math;
    return d;
}

static void fill_stats(const wsim::StepStats& st, wsim_step_stats* out) {
    out->density_iters = st.density_iters;
    out->divergence_iters = st.divergence_iters;
    out->mean_density_error = st.mean_density_error;
    out->mean_divergence_error = st.mean_divergence_error;
    out->max_velocity = st.max_velocity;
    out->nan_count = st.nan_count;
    out->substeps = st.substeps;
    out->escapees = st.escapees;
    out->reserved = 0;
    out->total_mass = st.total_mass;
    for (int a = 0; a < 3; ++a) out->momentum[a] = st.momentum[a];
    out->kinetic_energy = st.kinetic_energy;
}

wsim_status wsim_solver_create(wsim_store fluid, wsim_store boundary, const wsim_dfsph_config* c,
                               void* stream, wsim_solver* out) {
    return guarded([&] {
        need(out, "out");
        need(fluid, "fluid");
        need(c, "cfg");
        *out = nullptr;
        const wsim::DFSPHConfig d = to_config(c);
        auto h = std::make_unique<wsim_solver_t>();
        h->impl = std::make_unique<wsim::DFSPHSolver>(
            *fluid->impl, boundary ? boundary->impl.get() : nullptr, d, as_stream(stream));
        h->fluid_view.view = &h->impl->fluid_grid();
        h->boundary_view.view = h->impl->boundary_grid();
        *out = h.release();
    });
}
void wsim_solver_destroy(wsim_solver s) { delete s; }
wsim_status wsim_solver_step(wsim_solver s, float dt) {
    return guarded([&] {
        need(s, "solver");
        s->impl->step(dt);
    });
}
float wsim_solver_max_velocity(wsim_solver s) { return s ? s->impl->max_velocity() : 0.0f; }
float wsim_solver_particle_mass(wsim_solver s) { return s ? s->impl->particle_mass() : 0.0f; }
uint32_t wsim_solver_last_density_iters(wsim_solver s) {
    return s ? s->impl->last_density_iters() : 0u;
}
static wsim_status attr_at(wsim_solver s, int which, size_t i, float* out) {
    return guarded([&] {
        need(s, "solver");
        need(out, "out");
        *out = s->impl->attribute_at(which, i);
    });
}
wsim_status wsim_solver_density_at(wsim_solver s, size_t i, float* out) { return attr_at(s, 0, i, out); }
wsim_status wsim_solver_alpha_at(wsim_solver s, size_t i, float* out) { return attr_at(s, 1, i, out); }
wsim_status wsim_solver_boundary_mass_at(wsim_solver s, size_t i, float* out) {
    return attr_at(s, 2, i, out);
}
wsim_status wsim_solver_stats(wsim_solver s, wsim_step_stats* out) {
    return guarded([&] {
        need(s, "solver");
        need(out, "out");
        fill_stats(s->impl->stats(), out);
    });
}
wsim_grid wsim_solver_fluid_grid(wsim_solver s) { return s ? &s->fluid_view : nullptr; }
wsim_grid wsim_solver_boundary_grid(wsim_solver s) {
    return (s && s->boundary_view.view) ? &s->boundary_view : nullptr;
}
wsim_status wsim_solver_set_timing(wsim_solver s, int enable) {
    return guarded([&] {
        need(s, "solver");
        s->impl->set_timing(enable != 0);
    });
}
wsim_status wsim_solver_reset_timing(wsim_solver s) {
    return guarded([&] {
        need(s, "solver");
        s->impl->timer().reset();
    });
}
int wsim_solver_kernel_kinds(void) { return wsim::kKNumKinds; }
const char* wsim_solver_kernel_name(int kind) { return wsim::kernel_kind_name(kind); }
wsim_status wsim_solver_kernel_time(wsim_solver s, int kind, uint64_t* launches, double* total_ms) {
    return guarded([&] {
        need(s, "solver");
        if (kind < 0 || kind >= wsim::kKNumKinds) wsim::fail(wsim::kOutOfRange, "kernel kind");
        const auto& a = s->impl->timer().acc[kind];
        if (launches) *launches = a.count;
        if (total_ms) *total_ms = a.total_ms;
    });
}

// ---- launch shapes, autotuner, graphs ----
static wsim::KernelShapes from_c(const wsim_kernel_shapes* c) {
    wsim::KernelShapes k;
    k.density_alpha_nt = c->density_alpha_nt;
    k.xsph_nt = c->xsph_nt;
    k.change_nt = c->change_nt;
    k.change_pu = c->change_pu;
    k.apply_nt = c->apply_nt;
    k.apply_pu = c->apply_pu;
    k.cell_cpt = c->cell_cpt;
    k.remap = c->remap;
    return k;
}
wsim_status wsim_solver_get_shapes(wsim_solver s, wsim_kernel_shapes* out) {
    return guarded([&] {
        need(s, "solver");
        need(out, "out");
        const wsim::KernelShapes& k = s->impl->shapes();
        out->density_alpha_nt = k.density_alpha_nt;
        out->xsph_nt = k.xsph_nt;
        out->change_nt = k.change_nt;
        out->change_pu = k.change_pu;
        out->apply_nt = k.apply_nt;
        out->apply_pu = k.apply_pu;
        out->cell_cpt = k.cell_cpt;
        out->remap = k.remap;
    });
}
wsim_status wsim_solver_set_shapes(wsim_solver s, const wsim_kernel_shapes* shapes) {
    return guarded([&] {
        need(s, "solver");
        need(shapes, "shapes");
        s->impl->set_shapes(from_c(shapes));
    });
}
wsim_status wsim_solver_autotune(wsim_solver s, int reps) {
    return guarded([&] {
        need(s, "solver");
        s->impl->autotune(reps);
    });
}
size_t wsim_solver_tune_count(wsim_solver s) { return s ? s->impl->tune_log().size() : 0; }
wsim_status wsim_solver_tune_entry(wsim_solver s, size_t i, uint32_t shape[5], double* ms) {
    return guarded([&] {
        need(s, "solver");
        need(shape, "shape");
        const auto& log = s->impl->tune_log();
        if (i >= log.size()) wsim::fail(wsim::kOutOfRange, "tune entry out of range");
        const wsim::TuneRecord& r = log[i];
        shape[0] = static_cast<uint32_t>(r.kernel);
        shape[1] = r.nt;
        shape[2] = r.pu;
        shape[3] = r.remap;
        shape[4] = r.cpt;
        if (ms) *ms = r.ms;
    });
}
wsim_status wsim_solver_set_graphs(wsim_solver s, int enable) {
    return guarded([&] {
        need(s, "solver");
        s->impl->set_graphs(enable != 0);
    });
}

// ---- batched scenes ----
wsim_status wsim_batch_create(wsim_store fluid, wsim_store boundary, const wsim_dfsph_config* c,
                              uint32_t num_scenes, const uint32_t* counts, const float* viscosity,
                              void* stream, wsim_solver* out) {
    return guarded([&] {
        need(out, "out");
        need(fluid, "fluid");
        need(c, "cfg");
        need(counts, "counts");
        *out = nullptr;
        if (num_scenes == 0) wsim::fail(wsim::kInvalidArgument, "num_scenes must be positive");
        const wsim::DFSPHConfig d = to_config(c);
        std::vector<uint32_t> cnt(counts, counts + num_scenes);
        std::vector<float> visc;
        if (viscosity) visc.assign(viscosity, viscosity + num_scenes);
        auto h = std::make_unique<wsim_solver_t>();
        h->impl = std::make_unique<wsim::DFSPHSolver>(*fluid->impl,
                                                      boundary ? boundary->impl.get() : nullptr, d,
                                                      as_stream(stream), cnt, visc);
        h->fluid_view.view = &h->impl->fluid_grid();
        h->boundary_view.view = h->impl->boundary_grid();
        *out = h.release();
    });
}
wsim_status wsim_batch_step(wsim_solver b, const float* dt) {
    return guarded([&] {
        need(b, "batch");
        need(dt, "dt");
        if (!b->impl->batched()) wsim::fail(wsim::kInvalidArgument, "not a batched solver");
        b->impl->step_scenes(dt);
    });
}
uint32_t wsim_batch_scene_count(wsim_solver b) {
    return b ? static_cast<uint32_t>(b->impl->scene_count()) : 0u;
}
wsim_status wsim_batch_scene_offsets(wsim_solver b, uint32_t* out) {
    return guarded([&] {
        need(b, "batch");
        need(out, "out");
        if (!b->impl->batched()) wsim::fail(wsim::kInvalidArgument, "not a batched solver");
        const auto& st = b->impl->scene_starts();
        std::memcpy(out, st.data(), st.size() * sizeof(uint32_t));
    });
}
wsim_status wsim_batch_scene_stats(wsim_solver b, uint32_t scene, wsim_step_stats* out) {
    return guarded([&] {
        need(b, "batch");
        need(out, "out");
        fill_stats(b->impl->scene_stats(scene), out);
    });
}

wsim_status wsim_solver_neighbor_counts(wsim_solver s, uint32_t* cand_fluid, uint32_t* nbr_fluid,
                                        uint32_t* cand_boundary, uint32_t* nbr_boundary) {
    return guarded([&] {
        need(s, "solver");
        s->impl->neighbor_counts(cand_fluid, nbr_fluid, cand_boundary, nbr_boundary);
    });
}

// ---- differentiation ----
wsim_status wsim_solver_set_viscosity(wsim_solver s, float viscosity) {
    return guarded([&] {
        need(s, "solver");
        s->impl->set_viscosity(viscosity);
    });
}
wsim_status wsim_adjoint_step(wsim_solver s, float dt, const double* gx_out, const double* gv_out,
                              double* gx_in, double* gv_in, double* g_viscosity) {
    return guarded([&] {
        need(s, "solver");
        if (!s->tape) s->tape = std::make_unique<wsim::SubstepTape>();
        s->impl->step_taped(dt, *s->tape);
        const double g = s->impl->adjoint_step(*s->tape, gx_out, gv_out, gx_in, gv_in);
        if (g_viscosity) *g_viscosity = g;
    });
}
wsim_status wsim_rollout_create(wsim_solver s, uint32_t every, wsim_rollout* out) {
    return guarded([&] {
        need(s, "solver");
        need(out, "out");
        *out = nullptr;
        auto h = std::make_unique<wsim_rollout_t>();
        h->impl = std::make_unique<wsim::Rollout>(*s->impl, s->impl->fluid_store(), every);
        *out = h.release();
    });
}
void wsim_rollout_destroy(wsim_rollout r) { delete r; }
wsim_status wsim_rollout_forward(wsim_rollout r, const float* x0, const float* v0, const float* dt,
                                 uint32_t substeps) {
    return guarded([&] {
        need(r, "rollout");
        r->impl->forward(x0, v0, dt, substeps);
    });
}
wsim_status wsim_rollout_backward(wsim_rollout r, const double* gx, const double* gv, double* gx0,
                                  double* gv0, double* g_viscosity) {
    return guarded([&] {
        need(r, "rollout");
        const double g = r->impl->backward(gx, gv, gx0, gv0);
        if (g_viscosity) *g_viscosity = g;
    });
}

// ---- host surface ----
wsim_status wsim_sample_aabb_boundary(const float lo[3], const float hi[3], float spacing,
                                      int n_layers, float* out, size_t capacity, size_t* count) {
    return guarded([&] {
        need(lo, "lo");
        need(hi, "hi");
        const auto pts = wsim::sample_aabb_boundary(lo, hi, spacing, n_layers);
        if (count) *count = pts.size();
        if (out) {
            const size_t m = pts.size() < capacity ? pts.size() : capacity;
            std::memcpy(out, pts.data(), m * sizeof(wsim::Vec3));
        }
    });
}
void wsim_timestep_config_default(wsim_timestep_config* c) {
    if (!c) return;
    const wsim::TimeStepConfig d;
    c->frame_dt = d.frame_dt;
    c->dt_max = d.dt_max;
    c->cfl_lambda = d.cfl_lambda;
    c->max_substeps = d.max_substeps;
}
float wsim_timestep_next_dt(const wsim_timestep_config* c, float max_velocity,
                            float particle_radius, float remaining) {
    wsim::TimeStepConfig d;
    if (c) {
        d.frame_dt = c->frame_dt;
        d.dt_max = c->dt_max;
        d.cfl_lambda = c->cfl_lambda;
        d.max_substeps = c->max_substeps;
    }
    return wsim::next_dt(d, max_velocity, particle_radius, remaining);
}
wsim_status wsim_solver_advance(wsim_solver s, const wsim_timestep_config* c, float v_floor,
                                float time, uint32_t max_substeps, float* dts,
                                wsim_step_stats* stats, uint32_t* substeps) {
    return guarded([&] {
        need(s, "solver");
        wsim::TimeStepConfig d;
        if (c) {
            d.frame_dt = c->frame_dt;
            d.dt_max = c->dt_max;
            d.cfl_lambda = c->cfl_lambda;
            d.max_substeps = c->max_substeps;
        }
        std::vector<wsim::StepStats> rec(stats ? max_substeps : 0u);
        const uint32_t n = s->impl->advance(d, v_floor, time, max_substeps, dts,
                                            stats ? rec.data() : nullptr);
        for (uint32_t k = 0; stats && k < n; ++k) fill_stats(rec[k], stats + k);
        if (substeps) *substeps = n;
    });
}
void* wsim_solver_stream(wsim_solver s) { return s ? static_cast<void*>(s->impl->stream()) : nullptr; }

struct wsim_frame_writer_t {
    std::unique_ptr<wsim::FrameWriter> impl;
};
// ---- surface mesh ----
void wsim_mesh_params_default(wsim_mesh_params* out) {
    if (!out) return;
    const wsim::MeshParams d;
    out->particle_radius = d.particle_radius;
    out->grid_voxel_factor = d.grid_voxel_factor;
    out->smooth_sigma = d.smooth_sigma;
    out->iso_factor = d.iso_factor;
    out->swap_yz = d.swap_yz ? 1u : 0u;
    out->reserved = 0;
    out->weights = nullptr;
    out->smooth_pre_iters = d.smooth_pre_iters;
    out->k_neighbors = d.k_neighbors;
    out->smooth_post_iters = d.smooth_post_iters;
    out->reserved2 = 0;
}
wsim_status wsim_mesher_create(void* stream, wsim_mesher* out) {
    return guarded([&] {
        need(out, "out");
        *out = nullptr;
        auto h = std::make_unique<wsim_mesher_t>();
        h->impl = std::make_unique<wsim::SurfaceMesher>(as_stream(stream));
        *out = h.release();
    });
}
void wsim_mesher_destroy(wsim_mesher m) { delete m; }
wsim_status wsim_mesher_run(wsim_mesher m, const float* xyz, size_t count,
                            const wsim_mesh_params* p) {
    return guarded([&] {
        need(m, "mesher");
        need(p, "params");
        wsim::MeshParams q;
        q.particle_radius = p->particle_radius;
        q.grid_voxel_factor = p->grid_voxel_factor;
        q.smooth_sigma = p->smooth_sigma;
        q.iso_factor = p->iso_factor;
        q.swap_yz = p->swap_yz != 0;
        q.weights = p->weights;
        q.smooth_pre_iters = p->smooth_pre_iters;
        q.k_neighbors = p->k_neighbors;
        q.smooth_post_iters = p->smooth_post_iters;
        m->impl->run(xyz, count, q);
    });
}
wsim_status wsim_mesher_info(wsim_mesher m, wsim_mesh_info* out) {
    return guarded([&] {
        need(m, "mesher");
        need(out, "out");
        const wsim::MeshInfo& i = m->impl->info();
        std::memset(out, 0, sizeof(*out));
        for (int a = 0; a < 3; ++a) {
            out->dims[a] = i.dims[a];
            out->lo[a] = i.lo[a];
            out->extent[a] = i.extent[a];
            out->spacing[a] = i.spacing[a];
        }
        out->max_density = i.max_density;
        out->iso = i.iso;
        out->n_vertices = i.n_vertices;
        out->n_triangles = i.n_triangles;
    });
}
wsim_status wsim_mesher_grid(wsim_mesher m, const float** out) {
    return guarded([&] {
        need(m, "mesher");
        need(out, "out");
        *out = m->impl->grid();
    });
}
wsim_status wsim_mesher_vertices(wsim_mesher m, const float** out) {
    return guarded([&] {
        need(m, "mesher");
        need(out, "out");
        *out = m->impl->vertices();
    });
}
wsim_status wsim_mesher_triangles(wsim_mesher m, const uint32_t** out) {
    return guarded([&] {
        need(m, "mesher");
        need(out, "out");
        *out = m->impl->triangles();
    });
}
wsim_status wsim_mc_table(uint8_t* counts, uint8_t* edges) {
    return guarded([&] {
        need(counts, "counts");
        need(edges, "edges");
        std::memcpy(counts, wsim::mc_table_counts(), 256);
        std::memcpy(edges, wsim::mc_table_triangles(), 256 * 24);
    });
}

wsim_status wsim_frame_writer_create(const char* dir, size_t max_points, wsim_frame_writer* out) {
    return guarded([&] {
        need(out, "out");
        need(dir, "dir");
        *out = nullptr;
        auto h = std::make_unique<wsim_frame_writer_t>();
        h->impl = std::make_unique<wsim::FrameWriter>(dir, max_points);
        *out = h.release();
    });
}
wsim_status wsim_frame_writer_push(wsim_frame_writer w, wsim_store store, int index, void* stream) {
    return guarded([&] {
        need(w, "writer");
        need(store, "store");
        w->impl->push(store->impl->positions(), store->impl->count(), index, as_stream(stream));
    });
}
wsim_status wsim_frame_writer_finish(wsim_frame_writer w) {
    return guarded([&] {
        need(w, "writer");
        w->impl->finish();
    });
}
void wsim_frame_writer_destroy(wsim_frame_writer w) { delete w; }

wsim_status wsim_lattice_block(const float lo[3], const float hi[3], float spacing, float* out,
                               size_t capacity, size_t* count) {
    return guarded([&] {
        need(lo, "lo");
        need(hi, "hi");
        const auto pts = wsim::lattice_block(lo, hi, spacing);
        if (count) *count = pts.size();
        if (out) {
            const size_t m = pts.size() < capacity ? pts.size() : capacity;
            std::memcpy(out, pts.data(), m * sizeof(wsim::Vec3));
        }
    });
}
wsim_status wsim_write_frame(const char* path, const float* xyz, size_t count) {
    return guarded([&] {
        need(path, "path");
        if (count) need(xyz, "positions");
        wsim::write_frame(path, xyz, count);
    });
}

static void copy_str(char* dst, size_t cap, const std::string& s) {
    const size_t m = s.size() < cap - 1 ? s.size() : cap - 1;
    std::memcpy(dst, s.data(), m);
    dst[m] = '\0';
}

wsim_status wsim_scene_load(const char* path, wsim_scene* out) {
    return guarded([&] {
        need(path, "path");
        need(out, "out");
        const wsim::Scene s = wsim::load_scene(path);
        std::memset(out, 0, sizeof(*out));
        copy_str(out->name, sizeof(out->name), s.name);
        out->width = s.width;
        out->height = s.height;
        out->fps = s.fps;
        out->frame_start = s.frame_start;
        out->frame_end = s.frame_end;
        copy_str(out->format, sizeof(out->format), s.format);
        copy_str(out->video, sizeof(out->video), s.video);
        for (int a = 0; a < 3; ++a) {
            out->cam_position[a] = s.cam_position[a];
            out->cam_look_at[a] = s.cam_look_at[a];
            out->cam_up[a] = s.cam_up[a];
            out->gravity[a] = s.gravity[a];
            out->initial_block_min[a] = s.initial_block_min[a];
            out->initial_block_max[a] = s.initial_block_max[a];
        }
        out->cam_fov_y_deg = s.cam_fov_y_deg;
        copy_str(out->solver, sizeof(out->solver), s.solver);
        out->particle_radius = s.particle_radius;
        out->rest_density = s.rest_density;
        out->viscosity = s.viscosity;
        out->surface_tension = s.surface_tension;
    });
}

}  // extern "C"
