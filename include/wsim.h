/*
 * wsim.h -- C ABI of libwsim, the MI355X (gfx950) DFSPH neighbour-search and
 * pressure-solve library.
 *
 * The reference (Qervas/water_simulation) exposes this path as a C++ class
 * API with no C layer; every entry point below names the reference member it
 * replaces (file:line under the reference tree).  Conventions:
 *   - every call returns wsim_status (0 = ok) unless it is a plain getter;
 *     wsim_last_error() returns a thread-local message for the last failure;
 *   - no C++ exception ever crosses this boundary; the status codes map onto
 *     the reference's exception types (see wsim_status);
 *   - handles are opaque; device buffers are plain device pointers;
 *   - streams are hipStream_t passed as void* (NULL = the null stream);
 *   - positions and velocities are N x 3 float, AoS (x, y, z per particle),
 *     in the caller's original particle order.
 */
#ifndef WSIM_H
#define WSIM_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define WSIM_ABI_VERSION 2

typedef enum {
    WSIM_OK = 0,
    WSIM_ERR_RUNTIME = 1,          /* reference: std::runtime_error */
    WSIM_ERR_OUT_OF_RANGE = 2,     /* reference: std::out_of_range */
    WSIM_ERR_INVALID_ARGUMENT = 3, /* reference: std::invalid_argument / bad input */
    WSIM_ERR_DEVICE = 4            /* reference: CUDA error -> std::runtime_error
                                      (core/include/water/core/cuda_check.h:10-36) */
} wsim_status;

/* Message of the last failed call on this thread ("" if none). */
const char* wsim_last_error(void);
int wsim_abi_version(void);

/* ---- device memory (core/include/water/core/device_buffer.h:21-89) ---- */
wsim_status wsim_device_count(int* count);
wsim_status wsim_device_alloc(size_t bytes, void** out);
wsim_status wsim_device_free(void* ptr);
wsim_status wsim_memcpy_h2d(void* dst_device, const void* src_host, size_t bytes, void* stream);
wsim_status wsim_memcpy_d2h(void* dst_host, const void* src_device, size_t bytes, void* stream);
wsim_status wsim_memcpy_d2d(void* dst_device, const void* src_device, size_t bytes, void* stream);
wsim_status wsim_memset(void* dst_device, int value, size_t bytes, void* stream);
/* Device-to-device copy by libwsim's own streaming kernel (no reference
 * counterpart): the known-size workload used to calibrate HBM byte counters. */
wsim_status wsim_device_copy(void* dst_device, const void* src_device, size_t bytes, void* stream);
/* Diagnostic (no reference counterpart).  The solver kernels evaluate the
 * pair distance sqrt, q = d / l and the kernel-gradient quotient with the
 * hardware approximations and their fma correction steps instead of the
 * compiler's range-scaled IEEE sequences, when l is in the range where the
 * two agree (out[0]: 1 if so).  This runs every binary32 operand the
 * kernels can produce for smoothing length l through both on the current
 * device and returns the counts that differ in any bit: out[1] sqrt,
 * out[2] d / l, out[3] gradient quotient (all 0 when the solver is
 * bit-exact to the IEEE evaluation).  Synchronous; takes about a second. */
wsim_status wsim_pair_ops_check(float smoothing_length, uint64_t out[4]);
wsim_status wsim_stream_create(void** out_stream);
wsim_status wsim_stream_destroy(void* stream);
wsim_status wsim_stream_synchronize(void* stream);
wsim_status wsim_device_synchronize(void);

/* ---- ParticleStore (core/include/water/core/particle_store.h:48-107) ---- */
typedef struct wsim_store_t* wsim_store;
typedef enum { WSIM_ATTRIB_F32 = 0, WSIM_ATTRIB_F32X3 = 1, WSIM_ATTRIB_U32 = 2 } wsim_attrib_type;

/* ParticleStore(capacity, stream): registers "position" and "velocity" (F32x3). */
wsim_status wsim_store_create(size_t capacity, void* stream, wsim_store* out);
void wsim_store_destroy(wsim_store store);
/* resize(count): WSIM_ERR_OUT_OF_RANGE if count > capacity (particle_store.cu:12-16). */
wsim_status wsim_store_resize(wsim_store store, size_t count);
size_t wsim_store_count(wsim_store store);
size_t wsim_store_capacity(wsim_store store);
/* register_attribute<T>(name, type): zero-filled; WSIM_ERR_RUNTIME on a duplicate name
 * (particle_store.h:109-127). */
wsim_status wsim_store_register_attribute(wsim_store store, const char* name,
                                          wsim_attrib_type type, size_t* handle);
int wsim_store_has_attribute(wsim_store store, const char* name);
/* Handle of a registered attribute by name; WSIM_ERR_OUT_OF_RANGE if absent.  The
 * reference keeps its name map private; this lets callers reach the attributes a
 * solver registers (dfsph_density, dfsph_alpha, ...). */
wsim_status wsim_store_find_attribute(wsim_store store, const char* name, size_t* handle);
/* attribute_data(handle): WSIM_ERR_OUT_OF_RANGE on an invalid handle (particle_store.h:85-96). */
wsim_status wsim_store_attribute_data(wsim_store store, size_t handle, void** out);
/* positions() / velocities() (particle_store.h:61-64): N x 3 float device arrays. */
wsim_status wsim_store_positions(wsim_store store, float** out);
wsim_status wsim_store_velocities(wsim_store store, float** out);
size_t wsim_store_position_handle(wsim_store store);
size_t wsim_store_velocity_handle(wsim_store store);

/* ---- CellGrid (core/include/water/core/cell_grid.h:16-50) ---- */
typedef struct wsim_grid_t* wsim_grid;
/* CellGrid(origin, cells_per_axis, cell_length, stream); cell_start zero-filled. */
wsim_status wsim_grid_create(const float origin[3], const int32_t cells_per_axis[3],
                             float cell_length, void* stream, wsim_grid* out);
void wsim_grid_destroy(wsim_grid grid);
wsim_status wsim_grid_reserve(wsim_grid grid, size_t capacity);
/* build(positions, count) (cell_grid.cu:73-138): device N x 3 positions. */
wsim_status wsim_grid_build(wsim_grid grid, const float* positions_xyz, size_t count);
/* cell_start(): total_cells()+1 entries; cell_start[c] = #particles with key < c. */
const uint32_t* wsim_grid_cell_start(wsim_grid grid);
/* sorted_indices(): sorted slot -> original particle index. */
const uint32_t* wsim_grid_sorted_indices(wsim_grid grid);
/* Sorted cell keys (not exposed by the reference; its cell_ids_ buffer). */
const uint32_t* wsim_grid_sorted_keys(wsim_grid grid);
void wsim_grid_origin(wsim_grid grid, float out[3]);
void wsim_grid_cells_per_axis(wsim_grid grid, int32_t out[3]);
float wsim_grid_cell_length(wsim_grid grid);
uint32_t wsim_grid_total_cells(wsim_grid grid);
size_t wsim_grid_count(wsim_grid grid);
/* No counterpart in the reference: for callers that cannot allocate device
 * memory (code under HIP graph capture, serving loops, harnesses that forbid
 * hipMalloc).  The result of wsim_grid_create + wsim_grid_build (cell keys,
 * stable (key, index) sort, cell_start) written to caller arrays:
 * sorted_keys and sorted_indices hold count entries, cell_start
 * cells_per_axis[0] * [1] * [2] + 1 (all zero for count = 0).  Allocates
 * nothing: the sort's scratch lives in a fixed per-device array (18 MiB,
 * held until the process ends).  At most 2,097,152 particles
 * (WSIM_ERR_INVALID_ARGUMENT above that).  Calls that overlap on one device
 * race on the scratch: serialize them (one stream, or events). */
wsim_status wsim_cell_sort_noalloc(const float* positions_xyz, size_t count,
                                   const float origin[3], const int32_t cells_per_axis[3],
                                   float cell_length, uint32_t* sorted_keys,
                                   uint32_t* sorted_indices, uint32_t* cell_start, void* stream);

/* ---- SpatialAccel (core/include/water/core/spatial_accel.h:19-53) ---- */
typedef struct wsim_accel_t* wsim_accel;
wsim_status wsim_accel_create(wsim_accel* out);
void wsim_accel_destroy(wsim_accel accel);
/* build(positions, count, scene_bounds, stream) (spatial_accel.cu:111-199):
 * 30-bit Morton codes, their stable sort, the Karras internal nodes and the
 * node boxes (leaves: the point; internal: the union of the children, which
 * the reference leaves for later). */
wsim_status wsim_accel_build(wsim_accel accel, const float* positions_xyz, size_t count,
                             const float bounds_min[3], const float bounds_max[3], void* stream);
size_t wsim_accel_leaf_count(wsim_accel accel);
const uint32_t* wsim_accel_sorted_indices(wsim_accel accel);
/* Sorted Morton codes (the reference keeps them in morton_codes_). */
const uint32_t* wsim_accel_morton_codes(wsim_accel accel);
/* internal_left() / internal_right() (spatial_accel.h:36-37): children of the
 * leaf_count - 1 internal nodes; a child >= leaf_count - 1 is the leaf
 * child - (leaf_count - 1), i.e. the sorted particle at that slot.  NULL for
 * fewer than 2 leaves. */
const uint32_t* wsim_accel_internal_left(wsim_accel accel);
const uint32_t* wsim_accel_internal_right(wsim_accel accel);
/* Parent of every node (2 leaf_count - 1 entries, internal nodes first); the
 * root's is 0xFFFFFFFF (the reference keeps internal_parent_ private). */
const uint32_t* wsim_accel_internal_parent(wsim_accel accel);
/* node_aabbs() (spatial_accel.h:38): per node 6 floats (min x, y, z, max x,
 * y, z), the leaf_count - 1 internal nodes first, then the leaves. */
const float* wsim_accel_node_aabbs(wsim_accel accel);

/* ---- DFSPHSolver (solvers/dfsph/include/water/solvers/dfsph.h:17-84) ---- */
typedef struct {
    float rest_density;          /* dfsph.h:19-33 defaults via wsim_dfsph_config_default */
    float particle_radius;
    float smoothing_length;
    float viscosity;             /* XSPH epsilon */
    float damping;
    float surface_tension;       /* gamma of the surface-tension force when enabled below */
    float gravity[3];
    uint32_t max_density_iters;
    uint32_t max_divergence_iters;
    float eta_density;
    float eta_divergence;
    float domain_min[3];
    float domain_max[3];
    uint32_t enable_divergence_solve; /* extension: 0 = the reference's step() */
    uint32_t pressure_form;           /* extension: 0 = reference kappa/rho^2, 1 = kappa/rho */
    uint32_t enable_surface_tension;  /* extension: Akinci 2013 cohesion + curvature with
                                         coefficient surface_tension (the reference parses the
                                         coefficient and applies no force) */
    uint32_t fast_math;               /* extension: 1 = tolerance mode of the neighbour kernels
                                         (density/alpha, XSPH, both pressure solves): the pair
                                         terms use the hardware reciprocal square root and
                                         reciprocal and fused multiply-adds, as the reference's
                                         own -use_fast_math build (CMakeLists.txt:56) does;
                                         fields then differ from the IEEE evaluation in the
                                         last bits.  0 (default) = bit-exact IEEE binary32.
                                         The adjoint requires 0. */
} wsim_dfsph_config;

typedef struct {
    uint32_t density_iters;      /* last_density_iters() (dfsph.h:51) */
    uint32_t divergence_iters;
    double mean_density_error;   /* binary64 mean of max(0, rho* - rho0), last iteration */
    double mean_divergence_error;
    float max_velocity;          /* max_velocity() (dfsph.h:39) */
    uint32_t nan_count;          /* particles with a non-finite position or velocity */
    uint64_t substeps;
    /* Regression diagnostics after the substep (the checks of
     * tests/regression/test_settled_puddle.cu:197-206), no reference
     * counterpart in the stats API: */
    uint32_t escapees;           /* particles whose position lies on or outside the
                                    advection's clamp box [domain_min + 1e-3,
                                    domain_max - 1e-3] (dfsph_kernels.cu:376-385):
                                    they reached the domain limit */
    uint32_t reserved;
    double total_mass;           /* binary64: m x (particles with a finite state) */
    double momentum[3];          /* binary64: sum of m v over those particles */
    double kinetic_energy;       /* binary64: sum of m |v|^2 / 2 */
} wsim_step_stats;

typedef struct wsim_solver_t* wsim_solver;
void wsim_dfsph_config_default(wsim_dfsph_config* cfg);
/* DFSPHSolver(fluid, boundary (may be NULL), cfg, stream) (dfsph.cu:71-100).
 * Registers dfsph_density/alpha/kappa/density_adv on the fluid store and
 * dfsph_boundary_mass on the boundary store; both stores must outlive the
 * solver. */
wsim_status wsim_solver_create(wsim_store fluid, wsim_store boundary, const wsim_dfsph_config* cfg,
                               void* stream, wsim_solver* out);
void wsim_solver_destroy(wsim_solver solver);
/* step(dt) (dfsph.cu:120-133). */
wsim_status wsim_solver_step(wsim_solver solver, float dt);
float wsim_solver_max_velocity(wsim_solver solver);
float wsim_solver_particle_mass(wsim_solver solver);
uint32_t wsim_solver_last_density_iters(wsim_solver solver);
/* density_at / alpha_at / boundary_mass_at (dfsph.cu:330-353). */
wsim_status wsim_solver_density_at(wsim_solver solver, size_t i, float* out);
wsim_status wsim_solver_alpha_at(wsim_solver solver, size_t i, float* out);
wsim_status wsim_solver_boundary_mass_at(wsim_solver solver, size_t i, float* out);
wsim_status wsim_solver_stats(wsim_solver solver, wsim_step_stats* out);
/* Neighbour counts of ITER_NEIGHBORS (solvers/dfsph/src/dfsph_kernels.cu:13-35)
 * as the density/alpha walk of the next substep sees them, on the fluid
 * store's current positions, per particle in original order: cand = slots
 * of the in-grid 27-cell stencil (self included), nbr = candidates with
 * sqrtf(|r|^2) / smoothing_length < 2 (binary32), for the fluid store and
 * for the boundary store (0 where the particle's stencil holds no boundary
 * cell).  Device arrays of count() entries, each may be NULL.  Synchronous;
 * the store and the results of the following steps are unchanged.  The
 * reference exposes these only implicitly, through its neighbour loops. */
wsim_status wsim_solver_neighbor_counts(wsim_solver solver, uint32_t* cand_fluid,
                                        uint32_t* nbr_fluid, uint32_t* cand_boundary,
                                        uint32_t* nbr_boundary);
/* fluid_grid() / boundary_grid() (dfsph.h:41-42): borrowed, valid while the solver lives;
 * boundary grid is NULL without a boundary store.  Do not destroy. */
wsim_grid wsim_solver_fluid_grid(wsim_solver solver);
wsim_grid wsim_solver_boundary_grid(wsim_solver solver);
/* Per-kernel timing with HIP events on the solver's stream (profiling only). */
wsim_status wsim_solver_set_timing(wsim_solver solver, int enable);
wsim_status wsim_solver_reset_timing(wsim_solver solver);
/* kind in [0, wsim_solver_kernel_kinds()); name of the kernel class, launches
 * that did work and their summed device time in milliseconds. */
int wsim_solver_kernel_kinds(void);
const char* wsim_solver_kernel_name(int kind);
wsim_status wsim_solver_kernel_time(wsim_solver solver, int kind, uint64_t* launches,
                                    double* total_ms);

/* ---- launch shapes, autotuner, graph replay ----
 * No counterpart in the reference (its kernels use fixed 128/256-thread
 * blocks, dfsph.cu:137-326).  Every shape computes the same bits. */
typedef struct {
    uint32_t density_alpha_nt;   /* workgroup size: 64, 128, 256, 512 */
    uint32_t xsph_nt;            /* 0 (staged in LDS), 64, 128, 256, 512 */
    uint32_t change_nt;          /* 0 (neighbourhood staged in LDS, 256 threads), 256, 512, 1024 */
    uint32_t change_pu;          /* pairs per trip of the list loops: 2, 4, 8 */
    uint32_t apply_nt;           /* 0 (staged in LDS), 64, 128, 256, 512 */
    uint32_t apply_pu;           /* 2, 4, 8 */
    uint32_t cell_cpt;           /* cells per thread of the cell_start pass: 4, 8, 16 */
    uint32_t remap;              /* 1: XCD-aware workgroup order */
} wsim_kernel_shapes;
wsim_status wsim_solver_get_shapes(wsim_solver solver, wsim_kernel_shapes* out);
wsim_status wsim_solver_set_shapes(wsim_solver solver, const wsim_kernel_shapes* shapes);
/* Times every candidate shape on the current particle state (reps launches
 * each, median) and keeps the fastest per kernel.  Leaves the stores and the
 * results of the following steps unchanged. */
wsim_status wsim_solver_autotune(wsim_solver solver, int reps);
/* Candidates timed by the last autotune: entry i gives
 * {kernel kind, workgroup size, pairs per trip, remap, cells per thread}
 * and the median milliseconds per launch. */
size_t wsim_solver_tune_count(wsim_solver solver);
wsim_status wsim_solver_tune_entry(wsim_solver solver, size_t i, uint32_t shape[5], double* ms);
/* Replays each substep phase as a captured HIP graph (one per phase and per
 * iteration chunk).  Off by default for a plain solver, on for a batch;
 * ignored while kernel timing is on. */
wsim_status wsim_solver_set_graphs(wsim_solver solver, int enable);

/* ---- batched scenes ----
 * No counterpart in the reference, which steps one scene per solver.
 * num_scenes independent scenes share the config (except the XSPH viscosity)
 * and the boundary store.  Scene s owns the fluid store indices
 * [offsets[s], offsets[s+1]), offsets being the prefix sums of counts; the
 * store count must equal their sum.  Every scene steps with its own dt,
 * convergence test and statistics and gets the same results, bit for bit, as
 * a plain solver stepping that scene alone with the same dts.  The handle is
 * a wsim_solver: destroy, stats (scene 0), shapes, autotune, timing and
 * graphs work as for a plain solver. */
wsim_status wsim_batch_create(wsim_store fluid, wsim_store boundary, const wsim_dfsph_config* cfg,
                              uint32_t num_scenes, const uint32_t* counts,
                              const float* viscosity /* num_scenes values, or NULL */,
                              void* stream, wsim_solver* out);
/* dt: num_scenes substep lengths. */
wsim_status wsim_batch_step(wsim_solver batch, const float* dt);
uint32_t wsim_batch_scene_count(wsim_solver batch);
/* out: num_scenes + 1 offsets. */
wsim_status wsim_batch_scene_offsets(wsim_solver batch, uint32_t* out);
wsim_status wsim_batch_scene_stats(wsim_solver batch, uint32_t scene, wsim_step_stats* out);

/* ---- differentiation ----
 * No counterpart in the reference.  Reverse-mode derivatives of plain-solver
 * substeps (density solve; the divergence solve must be off) with the
 * discrete structure of the forward pass frozen: neighbour sets, iteration
 * count, max(0, .) and clamp masks.  Adjoint arrays are binary64 device
 * arrays of 3n values in original particle order. */
wsim_status wsim_solver_set_viscosity(wsim_solver solver, float viscosity);
/* One substep from the store's current state, recorded, then transposed:
 * d(loss)/d(output x, v) -> d(loss)/d(input x, v) and d(loss)/d(viscosity).
 * The store is left at the substep's input state. */
wsim_status wsim_adjoint_step(wsim_solver solver, float dt, const double* gx_out,
                              const double* gv_out, double* gx_in, double* gv_in,
                              double* g_viscosity);
/* Checkpointed rollout: forward keeps the state every checkpoint_every
 * substeps; backward recomputes one segment at a time from its checkpoint,
 * recording every pressure iterate, and runs the substep adjoints last to
 * first. */
typedef struct wsim_rollout_t* wsim_rollout;
wsim_status wsim_rollout_create(wsim_solver solver, uint32_t checkpoint_every, wsim_rollout* out);
void wsim_rollout_destroy(wsim_rollout rollout);
/* x0, v0: device, 3n floats; dt: host, substeps values.  Leaves the final
 * state in the solver's fluid store. */
wsim_status wsim_rollout_forward(wsim_rollout rollout, const float* x0, const float* v0,
                                 const float* dt, uint32_t substeps);
/* gx, gv: d(loss)/d(final x, v); writes d(loss)/d(x0, v0) and
 * d(loss)/d(viscosity).  Leaves the store at the initial state. */
wsim_status wsim_rollout_backward(wsim_rollout rollout, const double* gx, const double* gv,
                                  double* gx0, double* gv0, double* g_viscosity);

/* ---- host-side surface ---- */
/* sample_aabb_boundary(lo, hi, spacing, n_layers)
 * (solvers/dfsph/include/water/solvers/boundary_sampler.h:24-25).  Writes up to
 * capacity points (3 floats each) to out (may be NULL) and the total count. */
wsim_status wsim_sample_aabb_boundary(const float lo[3], const float hi[3], float spacing,
                                      int n_layers, float* out, size_t capacity, size_t* count);
/* TimeStepper (core/include/water/core/timestep.h:8-34). */
typedef struct {
    float frame_dt;
    float dt_max;
    float cfl_lambda;
    uint32_t max_substeps;
} wsim_timestep_config;
void wsim_timestep_config_default(wsim_timestep_config* cfg);
float wsim_timestep_next_dt(const wsim_timestep_config* cfg, float max_velocity,
                            float particle_radius, float remaining_in_frame);
/* The substep loop of sim_cli's frame (apps/sim_cli/main.cpp:171-178) in one
 * call, for a plain (not batched) solver: while time remains (> 1e-6) and
 * fewer than max_substeps ran, dt = next_dt(ts, max(max_velocity, v_floor),
 * particle_radius, remaining), then step(dt).  Writes every dt to dts and
 * every substep's statistics to stats (each NULL, or room for max_substeps)
 * and the substep count to substeps.  Each substep computes the same bits as
 * step(dt); without the divergence solve the grid phase of the next substep
 * overlaps the host's per-substep state read.  No counterpart in the
 * reference, whose loop lives in sim_cli. */
wsim_status wsim_solver_advance(wsim_solver solver, const wsim_timestep_config* ts, float v_floor,
                                float time, uint32_t max_substeps, float* dts,
                                wsim_step_stats* stats, uint32_t* substeps);
/* The solver's stream (the one it was created with, or its own). */
void* wsim_solver_stream(wsim_solver solver);

/* Asynchronous frame_%04d.bin output (replaces the synchronous copy and
 * write of apps/sim_cli/main.cpp:163-166, 191-193).  push snapshots a
 * store's positions (original order) in stream order on `stream`, copies
 * them to one of two pinned host buffers on a copy stream, and a background
 * thread writes dir/frame_<index>.bin; it blocks only while both buffers hold
 * frames not yet written.  finish waits for every pushed frame and reports
 * the first write error. */
typedef struct wsim_frame_writer_t* wsim_frame_writer;
wsim_status wsim_frame_writer_create(const char* dir, size_t max_points, wsim_frame_writer* out);
wsim_status wsim_frame_writer_push(wsim_frame_writer writer, wsim_store store, int index,
                                   void* stream);
wsim_status wsim_frame_writer_finish(wsim_frame_writer writer);
void wsim_frame_writer_destroy(wsim_frame_writer writer);

/* ---- fluid surface mesh: the splat method of tools/particles_to_mesh.py
 * (reconstruct_splat :77-139, smooth_positions_knn :61-74, load_frame :46-58)
 * on the device ----
 * Bounds padded by 4 r, +1 per point in voxel ((p - lo) / extent * res)
 * with res = max(ceil(extent / (factor r)), 8), the Gaussian blur of
 * scipy.ndimage.gaussian_filter (truncate 4, 'reflect'), iso = max *
 * iso_factor, then marching cubes with one shared vertex per crossing grid
 * edge.  Its triangle table resolves ambiguous faces by separating the
 * inside corners (the reference calls skimage's marching_cubes).  Before
 * the bounds, smooth_pre_iters passes replace every point by the mean of its
 * k nearest points (smooth_positions_knn :61-74, exact kNN, ties by index);
 * after marching cubes, smooth_post_iters passes of Laplacian smoothing
 * (:138-139, Open3D's filter_smooth_laplacian with its default lambda 0.5):
 * v + 0.5 (sum w_n v_n / sum w_n - v), w_n = 1 / (|v - v_n| + 1e-12) over
 * the vertices sharing a triangle with v, binary64, all vertices from the
 * previous pass.  weights, when
 * not NULL, are the 2R + 1 blur weights (R = (int)(4 sigma + 0.5)) to use
 * instead of exp(). */
typedef struct wsim_mesher_t* wsim_mesher;
typedef struct {
    double particle_radius;   /* --radius (0.012) */
    double grid_voxel_factor; /* 0.5 */
    double smooth_sigma;      /* blur sigma in voxels (2.5; 0 = none) */
    double iso_factor;        /* 0.18 */
    uint32_t swap_yz;         /* 1: (x, y, z) -> (x, z, y) as load_frame */
    uint32_t reserved;
    const double* weights;
    int32_t smooth_pre_iters; /* k-NN position smoothing passes (--smooth-pre, 2) */
    int32_t k_neighbors;      /* --k-neighbors: 8, 10, 15 (default), 20 or 30 */
    int32_t smooth_post_iters; /* Laplacian mesh smoothing passes (--smooth, 2) */
    int32_t reserved2;
} wsim_mesh_params;
typedef struct {
    int32_t dims[3];
    int32_t reserved;
    double lo[3], extent[3], spacing[3];
    float max_density, iso;
    uint64_t n_vertices, n_triangles;
} wsim_mesh_info;
void wsim_mesh_params_default(wsim_mesh_params* out);
wsim_status wsim_mesher_create(void* stream, wsim_mesher* out);
void wsim_mesher_destroy(wsim_mesher mesher);
/* positions_device: count x 3 binary32 on the device (e.g. wsim_store_positions).
 * Synchronous; the outputs stay valid until the next run. */
wsim_status wsim_mesher_run(wsim_mesher mesher, const float* positions_device, size_t count,
                            const wsim_mesh_params* params);
wsim_status wsim_mesher_info(wsim_mesher mesher, wsim_mesh_info* out);
wsim_status wsim_mesher_grid(wsim_mesher mesher, const float** out_device);
wsim_status wsim_mesher_vertices(wsim_mesher mesher, const float** out_device);
wsim_status wsim_mesher_triangles(wsim_mesher mesher, const uint32_t** out_device);
/* The triangle table: counts[256], edges[256 * 24] (edge a * 4 + b along
 * axis a from the cube corner with bit a clear and other bits b). */
wsim_status wsim_mc_table(uint8_t* counts, uint8_t* edges);

/* Lattice fluid block of sim_cli (apps/sim_cli/main.cpp:65-74). */
wsim_status wsim_lattice_block(const float lo[3], const float hi[3], float spacing, float* out,
                               size_t capacity, size_t* count);
/* frame_%04d.bin writer (apps/sim_cli/main.cpp:76-83): u32 N, then N x 3 float LE. */
wsim_status wsim_write_frame(const char* path, const float* positions_host, size_t count);

/* ---- scene loader (scene/include/water/scene/scene.h:9-45) ---- */
typedef struct {
    char name[256];
    int32_t width, height;
    float fps;
    int32_t frame_start, frame_end;
    char format[32];
    char video[32];
    float cam_position[3], cam_look_at[3], cam_up[3];
    float cam_fov_y_deg;
    char solver[32];
    float particle_radius;
    float rest_density;
    float viscosity;
    float surface_tension;
    float gravity[3];
    float initial_block_min[3];
    float initial_block_max[3];
} wsim_scene;
/* load_scene(path) (scene/src/loader.cpp:58-83): WSIM_ERR_RUNTIME on a missing
 * file, malformed JSON, missing or unsupported schema_version. */
wsim_status wsim_scene_load(const char* path, wsim_scene* out);

#ifdef __cplusplus
}
#endif
#endif /* WSIM_H */
