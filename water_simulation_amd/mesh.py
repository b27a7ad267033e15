"""Fluid surface meshes from particle frames on the GPU.

The splat method of tools/particles_to_mesh.py (reconstruct_splat :77-138,
load_frame :46-58, the command line :194-264) on libwsim's SurfaceMesher:
splat, Gaussian blur (the weights scipy.ndimage uses), iso level, marching
cubes, after the k-nearest-neighbour position smoothing and followed by the
Laplacian mesh smoothing (Open3D's filter_smooth_laplacian, :138-139).  The
reference's Poisson method is not provided.

    python -m water_simulation_amd.mesh --frame out/frame_0030.bin --out mesh.ply --radius 0.012
    python -m water_simulation_amd.mesh --seq out/dam_break --out-dir out/mesh --stride 2
"""
from __future__ import annotations

import argparse
import ctypes as C
import sys
import time
from pathlib import Path
from typing import Optional

import numpy as np

from . import DeviceBuffer, ParticleStore, _stream, memcpy_d2h, read_frame
from . import _ffi
from ._ffi import MeshInfoC, MeshParamsC, check


def _L():
    return _ffi.lib()


def gaussian_weights(sigma: float, truncate: float = 4.0) -> np.ndarray:
    """scipy.ndimage's order-0 Gaussian kernel: 2R + 1 binary64 taps,
    R = int(truncate * sigma + 0.5), exp(-x^2 / (2 sigma^2)) normalised."""
    radius = int(truncate * float(sigma) + 0.5)
    x = np.arange(-radius, radius + 1)
    phi = np.exp(-0.5 / (sigma * sigma) * x ** 2)
    return np.ascontiguousarray(phi / phi.sum(), np.float64)


class SurfaceMesher:
    """Device surface reconstruction; outputs stay on the device until read."""

    def __init__(self, stream=None):
        h = C.c_void_p()
        check(_L().wsim_mesher_create(_stream(stream), C.byref(h)))
        self._h = h.value
        self._points: Optional[DeviceBuffer] = None

    def close(self) -> None:
        if self._h:
            _L().wsim_mesher_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def run(self, points, radius: float = 0.012, grid_voxel_factor: float = 0.5,
            smooth_sigma: float = 2.5, iso_factor: float = 0.18, swap_yz: bool = True,
            smooth_pre: int = 2, k_neighbors: int = 15, smooth_post: int = 2) -> dict:
        """points: (N, 3) binary32 array, a ParticleStore, or (device pointer, N)."""
        if isinstance(points, ParticleStore):
            ptr, n = points.positions(), points.count()
        elif isinstance(points, tuple):
            ptr, n = int(points[0]), int(points[1])
        else:
            a = np.ascontiguousarray(points, np.float32).reshape(-1, 3)
            self._points = DeviceBuffer.from_array(a)
            ptr, n = self._points.ptr, len(a)
        p = MeshParamsC()
        _L().wsim_mesh_params_default(C.byref(p))
        p.particle_radius = float(radius)
        p.grid_voxel_factor = float(grid_voxel_factor)
        p.smooth_sigma = float(smooth_sigma)
        p.iso_factor = float(iso_factor)
        p.swap_yz = 1 if swap_yz else 0
        p.smooth_pre_iters = int(smooth_pre)
        p.k_neighbors = int(k_neighbors)
        p.smooth_post_iters = int(smooth_post)
        w = gaussian_weights(smooth_sigma) if smooth_sigma > 0 else np.zeros(1)
        p.weights = w.ctypes.data
        check(_L().wsim_mesher_run(self._h, ptr, n, C.byref(p)))
        return self.info()

    def info(self) -> dict:
        i = MeshInfoC()
        check(_L().wsim_mesher_info(self._h, C.byref(i)))
        return {"dims": tuple(i.dims), "lo": np.array(i.lo[:]), "extent": np.array(i.extent[:]),
                "spacing": np.array(i.spacing[:]), "max_density": np.float32(i.max_density),
                "iso": np.float32(i.iso), "n_vertices": int(i.n_vertices),
                "n_triangles": int(i.n_triangles)}

    def _read(self, fn, dtype, count: int) -> np.ndarray:
        out = np.empty(count, dtype)
        if count:
            p = C.c_void_p()
            check(fn(self._h, C.byref(p)))
            memcpy_d2h(out, p.value)
        return out

    def grid(self) -> np.ndarray:
        d = self.info()["dims"]
        return self._read(_L().wsim_mesher_grid, np.float32, d[0] * d[1] * d[2]).reshape(d)

    def vertices(self) -> np.ndarray:
        n = self.info()["n_vertices"]
        return self._read(_L().wsim_mesher_vertices, np.float32, 3 * n).reshape(n, 3)

    def triangles(self) -> np.ndarray:
        n = self.info()["n_triangles"]
        return self._read(_L().wsim_mesher_triangles, np.uint32, 3 * n).reshape(n, 3)


def mc_table():
    """The marching-cubes table: (counts uint8 [256], edges uint8 [256, 24])."""
    counts = np.zeros(256, np.uint8)
    edges = np.zeros((256, 24), np.uint8)
    check(_L().wsim_mc_table(counts.ctypes.data, edges.ctypes.data))
    return counts, edges


def write_ply(path, vertices: np.ndarray, triangles: np.ndarray) -> None:
    """Binary little-endian PLY: float x, y, z per vertex; uchar-counted
    uint faces."""
    v = np.ascontiguousarray(vertices, "<f4").reshape(-1, 3)
    t = np.ascontiguousarray(triangles, "<u4").reshape(-1, 3)
    head = ("ply\nformat binary_little_endian 1.0\n"
            f"element vertex {len(v)}\nproperty float x\nproperty float y\nproperty float z\n"
            f"element face {len(t)}\nproperty list uchar uint vertex_indices\nend_header\n")
    faces = np.zeros(len(t), dtype=[("n", "u1"), ("i", "<u4", (3,))])
    faces["n"] = 3
    faces["i"] = t
    with open(path, "wb") as f:
        f.write(head.encode("ascii"))
        f.write(v.tobytes())
        f.write(faces.tobytes())


def main(argv=None) -> int:
    p = argparse.ArgumentParser(description=__doc__.split("\n\n")[0])
    p.add_argument("--frame", help="single binary frame path")
    p.add_argument("--seq", help="directory of frame_NNNN.bin")
    p.add_argument("--out", help="single output .ply (with --frame)")
    p.add_argument("--out-dir", help="output directory (with --seq)")
    p.add_argument("--radius", type=float, default=0.012, help="particle radius from sim")
    p.add_argument("--method", choices=("splat",), default="splat")
    p.add_argument("--grid-voxel-factor", type=float, default=0.5)
    p.add_argument("--smooth-sigma", type=float, default=2.5)
    p.add_argument("--iso-factor", type=float, default=0.18)
    p.add_argument("--smooth-pre", type=int, default=2,
                   help="k-NN position smoothing iterations (Zhu-Bridson)")
    p.add_argument("--k-neighbors", type=int, default=15,
                   help="K-NN count for position smoothing (8, 10, 15, 20 or 30)")
    p.add_argument("--smooth", type=int, default=2,
                   help="post-MC Laplacian mesh smoothing iterations (0 disables)")
    p.add_argument("--stride", type=int, default=1)
    args = p.parse_args(argv)
    mesher = SurfaceMesher()

    def one(frame, out):
        pts = read_frame(str(frame))
        t0 = time.time()
        info = mesher.run(pts, args.radius, args.grid_voxel_factor, args.smooth_sigma,
                          args.iso_factor, smooth_pre=args.smooth_pre,
                          k_neighbors=args.k_neighbors, smooth_post=args.smooth)
        v, t = mesher.vertices(), mesher.triangles()
        dt = time.time() - t0
        Path(out).parent.mkdir(parents=True, exist_ok=True)
        write_ply(out, v, t)
        return len(pts), info["n_vertices"], info["n_triangles"], dt

    if args.frame:
        if not args.out:
            p.error("--out required with --frame")
        n, nv, nt, dt = one(args.frame, args.out)
        print(f"  {args.frame}: {n} particles -> {nv} verts, {nt} tris ({dt:.2f}s)")
    elif args.seq:
        if not args.out_dir:
            p.error("--out-dir required with --seq")
        files = sorted(Path(args.seq).glob("frame_*.bin"))[::args.stride]
        for k, f in enumerate(files):
            out = Path(args.out_dir) / f"{f.stem}.ply"
            n, nv, nt, dt = one(f, out)
            print(f"[{k + 1}/{len(files)}] {f.name} -> {out.name}: {n} particles -> {nv} verts, "
                  f"{nt} tris ({dt:.2f}s)")
    else:
        p.error("must pass --frame or --seq")
    return 0


if __name__ == "__main__":
    sys.exit(main())
