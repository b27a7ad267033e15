// host_surface.cpp -- host-side solver surface (no device code).
#include "host_surface.hpp"

#include <algorithm>
#include <cstdio>
#include <fstream>
#include <sstream>

#include "common.hpp"
#include "json.hpp"

namespace wsim {

// Faces in the order -X, +X, -Y, +Y, -Z, +Z; each face emits n_layers layers
// stepping inward by `spacing`; the two in-face axes ((axis+1)%3, (axis+2)%3)
// are walked by binary32 accumulation up to hi + spacing/2, so edge and
// corner points are duplicated exactly as in boundary_sampler.cpp:6-32 (the
// Akinci mass calibration absorbs the duplicates).
std::vector<Vec3> sample_aabb_boundary(const float lo[3], const float hi[3], float spacing,
                                       int n_layers) {
    std::vector<Vec3> pts;
    if (!(spacing > 0.0f) || n_layers <= 0) return pts;
    for (int face = 0; face < 6; ++face) {
        const int axis = face >> 1;
        const float base = (face & 1) ? hi[axis] : lo[axis];
        const float step = (face & 1) ? -spacing : spacing;
        const int ua = (axis + 1) % 3, va = (axis + 2) % 3;
        const float ulim = hi[ua] + 0.5f * spacing;
        const float vlim = hi[va] + 0.5f * spacing;
        for (int layer = 0; layer < n_layers; ++layer) {
            const float a = base + layer * step;
            for (float u = lo[ua]; u <= ulim; u += spacing) {
                for (float w = lo[va]; w <= vlim; w += spacing) {
                    float p[3];
                    p[axis] = a;
                    p[ua] = u;
                    p[va] = w;
                    pts.push_back({p[0], p[1], p[2]});
                }
            }
        }
    }
    return pts;
}

// min{dt_max, dt_cfl, remaining, max(dt_cfl, frame_dt/max_substeps)} with
// dt_cfl = lambda*r/max(v, 1e-6) (timestep.h:23-28).
float next_dt(const TimeStepConfig& c, float max_velocity, float particle_radius,
              float remaining_in_frame) {
    const float v = std::max(max_velocity, 1e-6f);
    const float dt_cfl = c.cfl_lambda * particle_radius / v;
    const float dt_floor = c.frame_dt / static_cast<float>(c.max_substeps);
    return std::min({c.dt_max, dt_cfl, remaining_in_frame, std::max(dt_cfl, dt_floor)});
}

// x-major lattice by binary32 accumulation (main.cpp:67-72).
std::vector<Vec3> lattice_block(const float lo[3], const float hi[3], float spacing) {
    std::vector<Vec3> pts;
    if (!(spacing > 0.0f)) return pts;
    for (float x = lo[0]; x <= hi[0]; x += spacing)
        for (float y = lo[1]; y <= hi[1]; y += spacing)
            for (float z = lo[2]; z <= hi[2]; z += spacing) pts.push_back({x, y, z});
    return pts;
}

void write_frame(const std::string& path, const float* xyz, size_t count) {
    std::ofstream out(path, std::ios::binary);
    if (!out) fail(kRuntimeError, "cannot open " + path + " for writing");
    const uint32_t n = static_cast<uint32_t>(count);
    out.write(reinterpret_cast<const char*>(&n), sizeof(n));
    if (count) out.write(reinterpret_cast<const char*>(xyz), static_cast<std::streamsize>(12 * count));
    if (!out) fail(kRuntimeError, "write failed for " + path);
}

namespace {

void read_vec3(const json::Value& j, const char* key, float out[3]) {
    if (!j.contains(key)) return;
    const json::Value& a = j.at(key);
    out[0] = a.at(size_t(0)).as_float();
    out[1] = a.at(size_t(1)).as_float();
    out[2] = a.at(size_t(2)).as_float();
}

}  // namespace

// loader.cpp:58-83: schema_version must be the string "1.0"; "name",
// "output", "camera" and "scene.fluid" are optional and fall back to the
// scene.h defaults; unknown keys are ignored.
Scene load_scene(const std::string& path) {
    std::ifstream in(path);
    if (!in) fail(kRuntimeError, "scene::load_scene: cannot open " + path);
    std::stringstream ss;
    ss << in.rdbuf();
    const std::string text = ss.str();
    json::Value j;
    try {
        j = json::Parser(text).parse();
    } catch (const json::ParseError& e) {
        fail(kRuntimeError, "scene::load_scene: JSON parse failed in " + path + ": " + e.what());
    }
    try {
        if (!j.contains("schema_version"))
            fail(kRuntimeError, "scene::load_scene: missing schema_version in " + path);
        const std::string sv = j.at("schema_version").as_string();
        if (sv != "1.0") fail(kRuntimeError, "scene::load_scene: unsupported schema_version: " + sv);
        Scene s;
        if (j.contains("name")) s.name = j.at("name").as_string();
        if (j.contains("output")) {
            const json::Value& o = j.at("output");
            if (o.contains("resolution")) {
                s.width = o.at("resolution").at(size_t(0)).as_int();
                s.height = o.at("resolution").at(size_t(1)).as_int();
            }
            if (o.contains("fps")) s.fps = o.at("fps").as_float();
            if (o.contains("frame_range")) {
                s.frame_start = o.at("frame_range").at(size_t(0)).as_int();
                s.frame_end = o.at("frame_range").at(size_t(1)).as_int();
            }
            if (o.contains("format")) s.format = o.at("format").as_string();
            if (o.contains("video")) s.video = o.at("video").as_string();
        }
        if (j.contains("camera")) {
            const json::Value& c = j.at("camera");
            read_vec3(c, "position", s.cam_position);
            read_vec3(c, "look_at", s.cam_look_at);
            read_vec3(c, "up", s.cam_up);
            if (c.contains("fov_y_deg")) s.cam_fov_y_deg = c.at("fov_y_deg").as_float();
        }
        if (j.contains("scene") && j.at("scene").contains("fluid")) {
            const json::Value& f = j.at("scene").at("fluid");
            if (f.contains("solver")) s.solver = f.at("solver").as_string();
            if (f.contains("particle_radius")) s.particle_radius = f.at("particle_radius").as_float();
            if (f.contains("rest_density")) s.rest_density = f.at("rest_density").as_float();
            if (f.contains("viscosity")) s.viscosity = f.at("viscosity").as_float();
            if (f.contains("surface_tension")) s.surface_tension = f.at("surface_tension").as_float();
            read_vec3(f, "gravity", s.gravity);
            read_vec3(f, "initial_block_min", s.initial_block_min);
            read_vec3(f, "initial_block_max", s.initial_block_max);
        }
        return s;
    } catch (const json::TypeError& e) {
        fail(kRuntimeError, "scene::load_scene: " + std::string(e.what()) + " in " + path);
    }
}

}  // namespace wsim
