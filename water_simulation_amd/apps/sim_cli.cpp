// sim_cli -- scene driver writing frame_%04d.bin files.
//
// Same command line, substep loop and frame format as the reference driver
// (apps/sim_cli/main.cpp:21-202): lattice block from the scene, Akinci walls
// on the unit box (3 layers at 2r), CFL substeps (at most 64 per frame),
// frames of u32 N followed by N x 3 float32 in original particle order.
// The reference also creates a Vulkan device only to print its name
// (main.cpp:104-113); that renderer is outside this library, so --no-vulkan
// is accepted and the device line is skipped either way.
#include <water/wsim.hpp>

#include <algorithm>
#include <cstdio>
#include <cstdlib>
#include <filesystem>
#include <memory>
#include <string>
#include <vector>

namespace {

struct Args {
    std::string scene_path;
    int frames_start = -1;
    int frames_end = -1;
    bool record = false;
    bool skip_vulkan = false;
    std::string out_dir = "out";
    float damping = -1.0f;
    float viscosity = -1.0f;
};

Args parse(int argc, char** argv) {
    Args a;
    for (int i = 1; i < argc; ++i) {
        const std::string s = argv[i];
        if (s == "--scene" && i + 1 < argc) {
            a.scene_path = argv[++i];
        } else if (s == "--frames" && i + 1 < argc) {
            const std::string r = argv[++i];
            const auto colon = r.find(':');
            if (colon == std::string::npos) {
                std::fprintf(stderr, "--frames expects START:END\n");
                std::exit(2);
            }
            a.frames_start = std::stoi(r.substr(0, colon));
            a.frames_end = std::stoi(r.substr(colon + 1));
        } else if (s == "--record") {
            a.record = true;
        } else if (s == "--out" && i + 1 < argc) {
            a.out_dir = argv[++i];
        } else if (s == "--no-vulkan") {
            a.skip_vulkan = true;
        } else if (s == "--damping" && i + 1 < argc) {
            a.damping = std::stof(argv[++i]);
        } else if (s == "--viscosity" && i + 1 < argc) {
            a.viscosity = std::stof(argv[++i]);
        } else if (s == "-h" || s == "--help") {
            std::puts("Usage: sim_cli --scene PATH [--frames START:END] [--record] "
                      "[--out DIR] [--no-vulkan] [--damping F] [--viscosity F]");
            std::exit(0);
        } else {
            std::fprintf(stderr, "Unknown arg: %s\n", s.c_str());
            std::exit(2);
        }
    }
    if (a.scene_path.empty()) {
        std::fprintf(stderr, "Error: --scene required\n");
        std::exit(2);
    }
    return a;
}

void write_frame(const std::string& dir, int index, const std::vector<water::Vec3f>& pts) {
    char name[512];
    std::snprintf(name, sizeof(name), "%s/frame_%04d.bin", dir.c_str(), index);
    water::detail::check(
        wsim_write_frame(name, reinterpret_cast<const float*>(pts.data()), pts.size()));
}

}  // namespace

int main(int argc, char** argv) try {
    Args args = parse(argc, argv);
    const auto scene = water::scene::load_scene(args.scene_path);
    if (args.frames_start < 0) args.frames_start = scene.output.frame_start;
    if (args.frames_end < 0) args.frames_end = scene.output.frame_end;

    std::printf("=== water_sim sim_cli ===\n");
    std::printf("scene:        %s\n", scene.name.c_str());
    std::printf("frames:       %d..%d (%d total)\n", args.frames_start, args.frames_end,
                args.frames_end - args.frames_start);
    std::printf("solver:       %s\n", scene.fluid.solver.c_str());
    std::printf("record:       %s%s\n", args.record ? "yes -> " : "no",
                args.record ? args.out_dir.c_str() : "");

    const float r = scene.fluid.particle_radius;
    const float spacing = 2.0f * r;
    std::vector<water::Vec3f> initial;
    {
        const float lo[3] = {scene.fluid.initial_block_min.x, scene.fluid.initial_block_min.y,
                             scene.fluid.initial_block_min.z};
        const float hi[3] = {scene.fluid.initial_block_max.x, scene.fluid.initial_block_max.y,
                             scene.fluid.initial_block_max.z};
        size_t n = 0;
        water::detail::check(wsim_lattice_block(lo, hi, spacing, nullptr, 0, &n));
        initial.resize(n);
        water::detail::check(
            wsim_lattice_block(lo, hi, spacing, reinterpret_cast<float*>(initial.data()), n, &n));
    }
    if (initial.empty()) {
        std::fprintf(stderr, "no particles\n");
        return 1;
    }
    std::printf("particles:    %zu\n", initial.size());

    water::ParticleStore store(initial.size());
    store.resize(initial.size());
    water::detail::check(wsim_memcpy_h2d(store.positions(), initial.data(),
                                         sizeof(water::Vec3f) * initial.size(), nullptr));

    water::solvers::DFSPHSolver::Config cfg;
    cfg.rest_density = scene.fluid.rest_density;
    cfg.particle_radius = r;
    cfg.smoothing_length = 2.0f * r;
    cfg.viscosity = (args.viscosity > 0.0f) ? args.viscosity : std::max(scene.fluid.viscosity, 1e-2f);
    cfg.surface_tension = scene.fluid.surface_tension;
    cfg.gravity = scene.fluid.gravity;
    cfg.domain_min = {0.0f, 0.0f, 0.0f};
    cfg.domain_max = {1.0f, 1.0f, 1.0f};
    if (args.damping >= 0.0f) cfg.damping = args.damping;
    std::printf("solver cfg:   damping=%.3f viscosity=%.4f\n", cfg.damping, cfg.viscosity);

    const auto bpts = water::solvers::sample_aabb_boundary(cfg.domain_min, cfg.domain_max, spacing, 3);
    std::printf("boundary:     %zu particles (Akinci 2012, 3 layers)\n", bpts.size());
    water::ParticleStore boundary(bpts.size());
    boundary.resize(bpts.size());
    water::detail::check(wsim_memcpy_h2d(boundary.positions(), bpts.data(),
                                         sizeof(water::Vec3f) * bpts.size(), nullptr));

    water::solvers::DFSPHSolver solver(store, &boundary, cfg);
    water::TimeStepper ts;

    std::unique_ptr<water::FrameWriter> writer;
    if (args.record) {
        std::filesystem::create_directories(args.out_dir);
        write_frame(args.out_dir, args.frames_start, initial);
        writer = std::make_unique<water::FrameWriter>(args.out_dir, store.count());
    }

    // main.cpp:171-195: per frame, CFL substeps (velocity floor 0.5 m/s, at
    // most 64) until the frame time is used up, then the frame.  The substep
    // loop runs in the library (the next substep's grid overlaps the host's
    // state read) and frames are written in the background while the next
    // frame is simulated.
    const float frame_dt = 1.0f / scene.output.fps;
    for (int f = args.frames_start; f < args.frames_end; ++f) {
        const water::u32 sub = solver.advance(ts, 0.5f, frame_dt, 64);
        std::printf("frame %4d: substeps=%2u, max_v=%.4f m/s\n", f + 1, sub, solver.max_velocity());
        if (writer) writer->push(store, f + 1, solver.stream());
    }
    if (writer) writer->finish();
    water::detail::check(wsim_device_synchronize());
    std::printf("=== sim_cli OK ===\n");
    return 0;
} catch (const std::exception& e) {
    std::fprintf(stderr, "sim_cli error: %s\n", e.what());
    return 1;
}
