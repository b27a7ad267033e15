// mesh_frame: one frame_%04d.bin to a binary PLY surface mesh through
// water::SurfaceMesher -- the splat method of tools/particles_to_mesh.py
// (--frame/--out, :46-58, 61-74, 77-139, 194-264) from C++.  The frame is
// u32 N then N x 3 float, little-endian (apps/sim_cli/main.cpp:76-83); the
// PLY holds float x, y, z per vertex and uchar-counted uint faces.
//
//   mesh_frame --frame out/frame_0030.bin --out mesh.ply [--radius 0.012]
//              [--grid-voxel-factor 0.5] [--smooth-sigma 2.5] [--iso-factor 0.18]
//              [--smooth-pre 2] [--k-neighbors 15] [--smooth 2] [--no-swap]
#include <water/wsim.hpp>

#include <cstdint>
#include <cstdlib>
#include <exception>
#include <fstream>
#include <iostream>
#include <string>
#include <vector>

int main(int argc, char** argv) {
    std::string frame, out;
    water::SurfaceMesher::Params p;
    for (int i = 1; i < argc; ++i) {
        const std::string a = argv[i];
        auto value = [&]() -> const char* {
            if (i + 1 >= argc) {
                std::cerr << "mesh_frame: " << a << " needs a value\n";
                std::exit(2);
            }
            return argv[++i];
        };
        if (a == "--frame") frame = value();
        else if (a == "--out") out = value();
        else if (a == "--radius") p.particle_radius = std::atof(value());
        else if (a == "--grid-voxel-factor") p.grid_voxel_factor = std::atof(value());
        else if (a == "--smooth-sigma") p.smooth_sigma = std::atof(value());
        else if (a == "--iso-factor") p.iso_factor = std::atof(value());
        else if (a == "--smooth-pre") p.smooth_pre_iters = std::atoi(value());
        else if (a == "--k-neighbors") p.k_neighbors = std::atoi(value());
        else if (a == "--smooth") p.smooth_post_iters = std::atoi(value());
        else if (a == "--no-swap") p.swap_yz = false;
        else {
            std::cerr << "mesh_frame: unknown argument " << a << "\n";
            return 2;
        }
    }
    if (frame.empty() || out.empty()) {
        std::cerr << "usage: mesh_frame --frame FRAME.bin --out MESH.ply [options]\n";
        return 2;
    }
    std::ifstream in(frame, std::ios::binary);
    uint32_t n = 0;
    in.read(reinterpret_cast<char*>(&n), sizeof n);
    std::vector<float> xyz(3 * static_cast<std::size_t>(n));
    in.read(reinterpret_cast<char*>(xyz.data()), static_cast<std::streamsize>(xyz.size() * 4));
    if (!in) {
        std::cerr << "mesh_frame: cannot read " << frame << "\n";
        return 1;
    }
    try {
        water::ParticleStore store(n);
        store.resize(n);
        water::detail::check(
            wsim_memcpy_h2d(store.positions(), xyz.data(), xyz.size() * sizeof(float), nullptr));
        water::SurfaceMesher mesher;
        const wsim_mesh_info info = mesher.run(store, p);
        std::vector<float> v(3 * info.n_vertices);
        std::vector<uint32_t> t(3 * info.n_triangles);
        water::detail::check(
            wsim_memcpy_d2h(v.data(), mesher.vertices(), v.size() * sizeof(float), nullptr));
        water::detail::check(
            wsim_memcpy_d2h(t.data(), mesher.triangles(), t.size() * sizeof(uint32_t), nullptr));
        std::ofstream ply(out, std::ios::binary);
        ply << "ply\nformat binary_little_endian 1.0\nelement vertex " << info.n_vertices
            << "\nproperty float x\nproperty float y\nproperty float z\nelement face "
            << info.n_triangles << "\nproperty list uchar uint vertex_indices\nend_header\n";
        ply.write(reinterpret_cast<const char*>(v.data()),
                  static_cast<std::streamsize>(v.size() * sizeof(float)));
        for (std::size_t f = 0; f < info.n_triangles; ++f) {
            const unsigned char three = 3;
            ply.write(reinterpret_cast<const char*>(&three), 1);
            ply.write(reinterpret_cast<const char*>(&t[3 * f]), 3 * sizeof(uint32_t));
        }
        if (!ply) {
            std::cerr << "mesh_frame: cannot write " << out << "\n";
            return 1;
        }
        std::cout << "  " << frame << ": " << n << " particles -> " << info.n_vertices
                  << " verts, " << info.n_triangles << " tris\n";
    } catch (const std::exception& e) {
        std::cerr << "mesh_frame: " << e.what() << "\n";
        return 1;
    }
    return 0;
}
