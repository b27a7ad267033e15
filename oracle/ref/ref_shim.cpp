// ref_shim.cpp -- C entry points over the reference's host-only sources.
//
// TEST INFRASTRUCTURE ONLY.  Compiled together with the reference's own
// solvers/dfsph/src/boundary_sampler.cpp (and its header-only
// core/include/water/core/timestep.h) into oracle/_ref/libref_host.so by
// oracle/ref/Makefile, so the oracle's restatements of the boundary sampler
// and the CFL time stepper can be checked bit for bit against the reference.
#include "water/core/timestep.h"
#include "water/solvers/boundary_sampler.h"

#include <cstddef>
#include <cstdint>

extern "C" size_t ref_sample_aabb_boundary(const float lo[3], const float hi[3], float spacing,
                                           int n_layers, float* out, size_t cap) {
    const auto pts = water::solvers::sample_aabb_boundary({lo[0], lo[1], lo[2]},
                                                          {hi[0], hi[1], hi[2]}, spacing, n_layers);
    if (out) {
        const size_t m = pts.size() < cap ? pts.size() : cap;
        for (size_t i = 0; i < m; ++i) {
            out[3 * i] = pts[i].x;
            out[3 * i + 1] = pts[i].y;
            out[3 * i + 2] = pts[i].z;
        }
    }
    return pts.size();
}

extern "C" float ref_next_dt(float frame_dt, float dt_max, float cfl_lambda, uint32_t max_substeps,
                             float max_velocity, float particle_radius, float remaining) {
    water::TimeStepper::Config c;
    c.frame_dt = frame_dt;
    c.dt_max = dt_max;
    c.cfl_lambda = cfl_lambda;
    c.max_substeps = max_substeps;
    return water::TimeStepper(c).next_dt(max_velocity, particle_radius, remaining);
}
