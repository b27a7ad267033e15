/*
 * wsim_oracle.h -- C interface of the CPU oracle.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is a plain CPU restatement of the
 * reference DFSPH hot path (Qervas/water_simulation).  It is loaded by the
 * tests, by __graft_entry__.smoke() and by the cpu_baseline leg of bench.py,
 * and nowhere else: the product library (libwsim.so) never links or calls it.
 *
 * Every function cites the reference file:line it restates.  Two precisions:
 *   ORC_F32 -- IEEE binary32 in the reference's operation order (the parity
 *              oracle; the HIP kernels must match it bit for bit);
 *   ORC_F64 -- the same algorithm and constants in binary64 (used to state the
 *              FP32 rounding error of each field).
 *
 * Pinning status (what checks each function, see tests/test_oracle_*.py):
 *   cell keys, permutation, cell_start   -- numpy lexsort / searchsorted, reference KATs
 *   morton codes                         -- known answers + bitwise loop interleave
 *   W, grad W                            -- closed forms, quadrature (int W = 1), central FD
 *   boundary mass, density, alpha        -- lattice closed forms, reference KAT m*W(0)
 *   gravity                              -- geometric-series closed form
 *   XSPH, density change, apply kappa    -- momentum invariants, uniform-field invariants,
 *                                           the kappa/rho^2 vs kappa/rho one-sweep contraction factor (rho)
 *   advect, max |v|, density error       -- closed-form cases, numpy reductions
 *   boundary sampler, time stepper       -- bit-exact against the compiled reference (oracle/_ref)
 *   divergence solve                     -- invariants only (the reference has no
 *                                           implementation; docs/superpowers/plans/2026-04-24-water-simulation-v2-phase2-dfsph-solver.md only)
 */
#ifndef WSIM_ORACLE_H
#define WSIM_ORACLE_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

enum { ORC_F32 = 32, ORC_F64 = 64, ORC_DUAL = 33 };

typedef struct {
    float    rest_density;
    float    particle_radius;
    float    smoothing_length;
    float    viscosity;
    float    damping;
    float    surface_tension;
    float    gravity[3];
    uint32_t max_density_iters;
    uint32_t max_divergence_iters;
    float    eta_density;
    float    eta_divergence;
    float    domain_min[3];
    float    domain_max[3];
    uint32_t enable_divergence_solve;
    uint32_t pressure_form; /* 0 = reference kappa/rho^2, 1 = kappa/rho (Bender-Koschier) */
    uint32_t enable_surface_tension; /* Akinci 2013 cohesion + curvature, coefficient surface_tension */
} orc_config;

/* Fields readable from an oracle solver (values converted to double). */
enum {
    ORC_FIELD_X = 0,        /* 3n */
    ORC_FIELD_V = 1,        /* 3n */
    ORC_FIELD_RHO = 2,      /* n  */
    ORC_FIELD_ALPHA = 3,    /* n  */
    ORC_FIELD_KAPPA = 4,    /* n  */
    ORC_FIELD_RHO_ADV = 5,  /* n  */
    ORC_FIELD_BMASS = 6,    /* nb */
    ORC_FIELD_XB = 7        /* 3nb */
};

typedef struct {
    uint32_t density_iters;
    uint32_t divergence_iters;
    double   mean_density_error;     /* of the last density iteration's rho* */
    double   mean_divergence_error;
    float    max_velocity;
    uint32_t nan_count;
    /* after the substep, in binary64 over the particles with a finite state */
    uint32_t escapees;               /* position on or outside the advection clamp box */
    double   total_mass;             /* m x (particles with a finite state) */
    double   momentum[3];            /* sum m v */
    double   kinetic_energy;         /* sum m |v|^2 / 2 */
} orc_stats;

void orc_set_threads(int n);
/* Test-only: XSPH as an in-place sequential sweep (one schedule of the
 * reference's racy in-place kernel) instead of the Jacobi sweep. */
void orc_set_xsph_in_place(int on);
/* analysis-only: DualF kernel terms without position tangents; advection
   without the |v| clamp (bit 0) or the domain clamp (bit 1) */
void orc_set_analysis(int freeze_geometry, int free_advect);
int  orc_get_threads(void);

/* --- SPH kernel (core/include/water/core/sph_kernels.cuh:12-41) --- */
float  orc_W_f32(float r, float l);
double orc_W_f64(double r, double l);
void   orc_gradW_f32(const float r[3], float l, float out[3]);
void   orc_gradW_f64(const double r[3], double l, double out[3]);

/* --- uniform grid (core/src/cell_grid.cu:10-54,73-138) --- */
void orc_cell_keys(const float* xyz, size_t n, const float origin[3],
                   const int32_t dims[3], float cell, uint32_t* keys);
/* Writes sorted_keys[n], perm[n] (sorted slot -> original id), cell_start[C+1]. */
void orc_grid_build(const float* xyz, size_t n, const float origin[3],
                    const int32_t dims[3], float cell, uint32_t* sorted_keys,
                    uint32_t* perm, uint32_t* cell_start);
/* DFSPH grid layout (solvers/dfsph/src/dfsph.cu:55-65). */
void orc_make_grid(const float dmin[3], const float dmax[3], float h,
                   float origin[3], int32_t dims[3]);
/* Per-particle stencil candidate counts and q<2 neighbour counts (fluid set
 * queried against a (possibly different) point set on the same grid). */
void orc_neighbor_counts(const float* xq, size_t nq, const float* xs, size_t ns,
                         const float origin[3], const int32_t dims[3], float cell,
                         float h, uint32_t* cand, uint32_t* nbr);

/* --- Morton path of SpatialAccel (core/src/spatial_accel.cu:11-40,139-179) --- */
void orc_morton(const float* xyz, size_t n, const float bmin[3], const float bmax[3],
                uint32_t* codes, uint32_t* sorted_codes, uint32_t* perm);
void orc_lbvh(const uint32_t* sorted_codes, const float* xyz, const uint32_t* perm, size_t n,
              uint32_t* left, uint32_t* right, uint32_t* parent, float* boxes);

/* --- host-side surface --- */
/* solvers/dfsph/src/boundary_sampler.cpp:6-32; returns count, writes if out != NULL */
size_t orc_sample_aabb_boundary(const float lo[3], const float hi[3], float spacing,
                                int n_layers, float* out, size_t cap);
/* apps/sim_cli/main.cpp:65-74 and tests/regression/test_settled_puddle.cu:24-32 */
size_t orc_lattice_block(const float lo[3], const float hi[3], float spacing,
                         float* out, size_t cap);
/* exp(-damping*dt) in binary32 (dfsph_kernels.cu:351), the library's rule */
float orc_damping_decay(float damping, float dt);
/* core/include/water/core/timestep.h:23-28 */
float orc_next_dt(float frame_dt, float dt_max, float cfl_lambda, uint32_t max_substeps,
                  float max_velocity, float particle_radius, float remaining);

/* --- DFSPH solver (solvers/dfsph/src/dfsph.cu, dfsph_kernels.cu) --- */
void  orc_config_default(orc_config* c);
void* orc_solver_create(const orc_config* cfg, int precision, const float* x,
                        const float* v, size_t n, const float* xb, size_t nb);
void  orc_solver_destroy(void* s);
void  orc_solver_step(void* s, float dt);
/* Grid + density + alpha only, on the current positions (dfsph.cu:124-126). */
void  orc_solver_density_alpha(void* s);
/* the surface-tension phase alone on the current state (after density_alpha) */
void  orc_solver_surface_tension(void* s, float dt);
void  orc_solver_set_state(void* s, const float* x, const float* v);
/* Same in binary64 (rounded to binary32 for F32 and DUAL solvers). */
void  orc_solver_set_state_f64(void* s, const double* x, const double* v);
/* XSPH viscosity in binary64 (finite differences in the viscosity). */
void  orc_solver_set_viscosity_f64(void* s, double viscosity);
void  orc_solver_get(void* s, int field, double* out);
void  orc_solver_get_f32(void* s, int field, float* out); /* F32 solvers only */
void  orc_solver_stats(void* s, orc_stats* out);

/* --- gradient checks ---
 * ORC_DUAL solvers carry forward-mode tangents: values bit-identical to the
 * ORC_F32 instance, tangents d(value)/d(seed) in binary64.  Seeds: initial
 * positions, initial velocities (3n each, may be NULL) and the XSPH
 * viscosity. */
void  orc_solver_seed(void* s, const double* dx, const double* dv, double d_viscosity);
void  orc_solver_get_tangent(void* s, int field, double* out); /* X, V, RHO, ALPHA, KAPPA, RHO_ADV */
/* Frozen structure: mode 1 records every grid build's home cells, every
 * solve's iteration count, every sweep's max(0, .) branches and every
 * advection's clamp decisions; mode 2 replays a recorded tape (from its
 * start), so perturbed runs of any precision take the recording run's
 * branches -- the piecewise-smooth map the adjoint differentiates. */
void* orc_tape_create(void);
void  orc_tape_destroy(void* tape);
size_t orc_tape_builds(void* tape);
void  orc_solver_set_tape(void* s, void* tape, int mode);
float orc_solver_particle_mass(void* s);
size_t orc_solver_count(void* s);

#ifdef __cplusplus
}
#endif
#endif
