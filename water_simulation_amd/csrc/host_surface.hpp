// host_surface.hpp -- host-side parts of the solver surface: boundary
// sampler, CFL time stepper, lattice generator, frame writer, scene loader.
#pragma once

#include <cstdint>
#include <string>
#include <vector>

namespace wsim {

struct Vec3 {
    float x, y, z;
};

// solvers/dfsph/src/boundary_sampler.cpp:6-32
std::vector<Vec3> sample_aabb_boundary(const float lo[3], const float hi[3], float spacing,
                                       int n_layers);

// core/include/water/core/timestep.h:8-34
struct TimeStepConfig {
    float frame_dt = 1.0f / 24.0f;
    float dt_max = 1.0f / 24.0f;
    float cfl_lambda = 0.4f;
    uint32_t max_substeps = 16;
};
float next_dt(const TimeStepConfig& c, float max_velocity, float particle_radius,
              float remaining_in_frame);

// apps/sim_cli/main.cpp:65-74
std::vector<Vec3> lattice_block(const float lo[3], const float hi[3], float spacing);

// apps/sim_cli/main.cpp:76-83
void write_frame(const std::string& path, const float* xyz, size_t count);

// scene/include/water/scene/scene.h:9-45 (defaults) and scene/src/loader.cpp:12-83
struct Scene {
    std::string name;
    int32_t width = 1920, height = 1080;
    float fps = 24.0f;
    int32_t frame_start = 0, frame_end = 120;
    std::string format = "exr", video = "mp4";
    float cam_position[3] = {0.0f, 0.0f, 1.0f};
    float cam_look_at[3] = {0.0f, 0.0f, 0.0f};
    float cam_up[3] = {0.0f, 1.0f, 0.0f};
    float cam_fov_y_deg = 30.0f;
    std::string solver = "dfsph";
    float particle_radius = 0.0015f;
    float rest_density = 1000.0f;
    float viscosity = 1e-3f;
    float surface_tension = 0.0728f;
    float gravity[3] = {0.0f, -9.81f, 0.0f};
    float initial_block_min[3] = {0.4f, 0.05f, 0.4f};
    float initial_block_max[3] = {0.6f, 0.25f, 0.6f};
};
Scene load_scene(const std::string& path);

}  // namespace wsim
