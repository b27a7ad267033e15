"""CPU oracle of the surface mesher (test infrastructure only).

Follows the splat method of tools/particles_to_mesh.py: load_frame (:46-58,
the y/z swap), the k-nearest-neighbour position smoothing with scipy's
cKDTree as the tool calls it (:61-74), reconstruct_splat (:77-138) with the
Laplacian post-smoothing off: the grid set-up and the splat
(:104-115) in numpy exactly as written there, the blur by
scipy.ndimage.gaussian_filter itself (:118), iso = max * iso_factor
(:120-123, numpy 2 scalar rules keep it binary32).  The isosurface step
replaces skimage.measure.marching_cubes (scikit-image, absent here; the
reference pins no version) with the marching-cubes variant of the GPU path,
restated here: one vertex per grid edge whose end values straddle the iso
level, ordered by (grid point, axis); triangles per cube from a table that
joins, on every face, the two crossing edges bounding each run of inside
corners, ordered by (cube, table order).  The Laplacian post-smoothing
(:138-139) restates Open3D's TriangleMesh.filter_smooth_laplacian (Open3D,
absent here; the reference pins no version) as its documentation gives it,
with lambda 0.5: v + 0.5 (sum w_n v_n / sum w_n - v),
w_n = 1 / (|v - v_n| + 1e-12), over the vertices sharing a triangle with v,
binary64, neighbours in ascending index (Open3D iterates a hash set).

Pinned by tests/test_mesh.py: blur_restated equals scipy bit for bit;
every mesh of a field that stays inside the grid is closed and consistently
oriented; a sphere's area and enclosed volume converge to 4 pi r^2 and
4/3 pi r^3 with the outward orientation, Euler characteristic 2 (a torus 0);
complementary corner configurations cut the same edges; cKDTree's
neighbours equal a brute-force search ordered by (distance, index) on
inputs without distance ties (the GPU's order); the Laplacian step keeps a
plane's vertices on the plane, leaves a vertex whose neighbours sit
symmetrically around it in place, moves a vertex halfway to the
inverse-distance-weighted neighbour mean, and shrinks a sphere's mesh.
Parity with skimage's Lewiner tables and with Open3D's smoothing is
unpinned (neither is available), and so is cKDTree's order among exactly
equal distances.
"""
from __future__ import annotations

import numpy as np
import scipy.ndimage


def scipy_weights(sigma: float, truncate: float = 4.0) -> np.ndarray:
    """scipy.ndimage's order-0 Gaussian kernel (float64, 2R + 1 taps)."""
    radius = int(truncate * float(sigma) + 0.5)
    x = np.arange(-radius, radius + 1)
    phi = np.exp(-0.5 / (sigma * sigma) * x ** 2)
    return phi / phi.sum()


def blur_restated(grid: np.ndarray, sigma: float) -> np.ndarray:
    """The three 1-D passes as the GPU computes them: binary64 sums
    x[i] w[R] + sum_j (x[i-j] + x[i+j]) w[R+j], 'reflect' ends, binary32
    between passes."""
    w = scipy_weights(sigma)
    R = (len(w) - 1) // 2
    out = grid.astype(np.float32)
    for axis in range(3):
        a = np.moveaxis(out, axis, -1).astype(np.float64)
        n = a.shape[-1]
        idx = np.arange(n)

        def refl(k):
            k = np.mod(k, 2 * n)
            return np.where(k >= n, 2 * n - 1 - k, k)

        acc = a * w[R]
        for j in range(1, R + 1):
            acc = acc + (a[..., refl(idx - j)] + a[..., refl(idx + j)]) * w[R + j]
        out = np.moveaxis(acc.astype(np.float32), -1, axis)
    return out


def smooth_positions_knn(pts: np.ndarray, k: int = 15, iters: int = 2) -> np.ndarray:
    """tools/particles_to_mesh.py:61-74: every point becomes the mean of its
    k nearest points (itself included) found by scipy's cKDTree, twice."""
    from scipy.spatial import cKDTree

    smoothed = pts.copy()
    for _ in range(iters):
        tree = cKDTree(smoothed)
        _, idx = tree.query(smoothed, k=k)
        smoothed = smoothed[idx].mean(axis=1)
    return smoothed


def knn_brute(pts: np.ndarray, k: int) -> np.ndarray:
    """k nearest of every point by brute force, ordered by (squared distance
    with ((dx^2 + dy^2) + dz^2), index)."""
    d = pts[None, :, :] - pts[:, None, :]
    d2 = (d[..., 0] * d[..., 0] + d[..., 1] * d[..., 1]) + d[..., 2] * d[..., 2]
    order = np.lexsort((np.broadcast_to(np.arange(len(pts)), d2.shape), d2), axis=1)
    return order[:, :k]


def grid_setup(pts: np.ndarray, radius: float, voxel_factor: float):
    """tools/particles_to_mesh.py:104-109 (pts float64)."""
    pad = radius * 4.0
    lo = pts.min(0) - pad
    hi = pts.max(0) + pad
    extent = hi - lo
    voxel = radius * voxel_factor
    res = np.maximum(np.ceil(extent / voxel).astype(int), 8)
    return lo, extent, res


def splat(pts, lo, extent, res) -> np.ndarray:
    """tools/particles_to_mesh.py:111-115."""
    grid = np.zeros(tuple(res), dtype=np.float32)
    gc = ((pts - lo) / extent * res).astype(int)
    gc = np.clip(gc, 0, res - 1)
    np.add.at(grid, (gc[:, 0], gc[:, 1], gc[:, 2]), 1.0)
    return grid


# ---- marching-cubes table --------------------------------------------------
# corner c at offset (c & 1, c >> 1 & 1, c >> 2 & 1); edge a * 4 + b along
# axis a from the corner with bit a clear, b = its other two bits (lower
# axis first); inside = value > iso.

def _edge(c0: int, c1: int) -> int:
    a = (c0 ^ c1).bit_length() - 1
    base = c0 & ~(1 << a)
    others = [ax for ax in range(3) if ax != a]
    b = ((base >> others[0]) & 1) | (((base >> others[1]) & 1) << 1)
    return 4 * a + b


def mc_table():
    counts = np.zeros(256, np.uint8)
    tris = np.zeros((256, 24), np.uint8)
    for cfg in range(256):
        inside = [(cfg >> c) & 1 == 1 for c in range(8)]
        nxt = {}
        for a in range(3):
            u, v = (a + 1) % 3, (a + 2) % 3
            for s in (0, 1):
                quad = [(0, 0), (1, 0), (1, 1), (0, 1)]
                if s == 0:
                    quad = [quad[0], quad[3], quad[2], quad[1]]
                cyc = [(s << a) | (du << u) | (dv << v) for du, dv in quad]
                for m in range(4):
                    if not inside[cyc[m]] or inside[cyc[m - 1]]:
                        continue
                    k = m
                    while inside[cyc[(k + 1) % 4]]:
                        k = (k + 1) % 4
                    nxt[_edge(cyc[m - 1], cyc[m])] = _edge(cyc[k], cyc[(k + 1) % 4])
        seen, out = set(), []
        for e0 in range(12):
            if e0 not in nxt or e0 in seen:
                continue
            loop, e = [], e0
            while e not in seen:
                seen.add(e)
                loop.append(e)
                e = nxt[e]
            for k in range(1, len(loop) - 1):
                out.append((loop[0], loop[k], loop[k + 1]))
        counts[cfg] = len(out)
        for t, tri in enumerate(out):
            tris[cfg, 3 * t:3 * t + 3] = tri
    return counts, tris


_TABLE = None


def table():
    global _TABLE
    if _TABLE is None:
        _TABLE = mc_table()
    return _TABLE


def marching_cubes(grid: np.ndarray, iso, lo, spacing):
    """(vertices float32 (V, 3), triangles uint32 (F, 3)) in the GPU's order."""
    counts, tris = table()
    g = grid
    r = np.array(g.shape)
    inside = g > iso
    emask = np.zeros(g.shape, np.uint8)
    for ax in range(3):
        sl0 = [slice(None)] * 3
        sl1 = [slice(None)] * 3
        sl0[ax] = slice(0, -1)
        sl1[ax] = slice(1, None)
        cross = inside[tuple(sl0)] != inside[tuple(sl1)]
        emask[tuple(sl0)] |= (cross.astype(np.uint8) << ax)
    flat = emask.ravel()
    ecount = ((flat & 1) + ((flat >> 1) & 1) + ((flat >> 2) & 1)).astype(np.int64)
    eoff = np.concatenate([[0], np.cumsum(ecount)[:-1]])
    strides = np.array([r[1] * r[2], r[2], 1])
    gf = g.ravel().astype(np.float64)
    iso64 = np.float64(np.float32(iso))
    verts = np.zeros((int(ecount.sum()), 3), np.float32)
    pts = np.nonzero(flat)[0]
    coords = np.stack(np.unravel_index(pts, g.shape), -1)
    for p, ijk in zip(pts, coords):
        v = eoff[p]
        for ax in range(3):
            if not (flat[p] >> ax) & 1:
                continue
            v0, v1 = gf[p], gf[p + strides[ax]]
            t = (iso64 - v0) / (v1 - v0)
            for k in range(3):
                f = float(ijk[k]) + t if k == ax else float(ijk[k])
                verts[v, k] = np.float32(f * spacing[k] + lo[k])
            v += 1
    # cube configurations
    cfg = np.zeros(tuple(r - 1), np.int64)
    for c in range(8):
        o = [(c >> ax) & 1 for ax in range(3)]
        cfg |= inside[o[0]:r[0] - 1 + o[0], o[1]:r[1] - 1 + o[1], o[2]:r[2] - 1 + o[2]].astype(
            np.int64) << c
    out = []
    for cube in np.nonzero(counts[cfg.ravel()])[0]:
        cidx = np.unravel_index(cube, cfg.shape)
        p = int(np.dot(cidx, strides))
        c = int(cfg[cidx])
        for t in range(counts[c]):
            tri = []
            for q in range(3):
                e = int(tris[c, 3 * t + q])
                ax, b = e >> 2, e & 3
                others = [x for x in range(3) if x != ax]
                off = [0, 0, 0]
                off[others[0]] = b & 1
                off[others[1]] = (b >> 1) & 1
                o = p + int(np.dot(off, strides))
                rank = bin(int(flat[o]) & ((1 << ax) - 1)).count("1")
                tri.append(eoff[o] + rank)
            out.append(tri)
    return verts, np.array(out, np.uint32).reshape(-1, 3)


def vertex_neighbours(tris: np.ndarray, nv: int):
    """Every vertex's neighbours (the vertices sharing a triangle with it),
    ascending: (table (nv, D) int64 padded with 0, valid (nv, D) bool)."""
    t = np.asarray(tris, np.int64).reshape(-1, 3)
    pairs = np.concatenate([t[:, [0, 1]], t[:, [1, 2]], t[:, [2, 0]],
                            t[:, [1, 0]], t[:, [2, 1]], t[:, [0, 2]]])
    key = np.unique(pairs[:, 0] * nv + pairs[:, 1])
    src, dst = key // nv, key % nv
    deg = np.bincount(src, minlength=nv)
    start = np.concatenate([[0], np.cumsum(deg)])
    D = int(deg.max()) if len(deg) else 0
    table = np.zeros((nv, D), np.int64)
    valid = np.zeros((nv, D), bool)
    col = np.arange(len(src)) - start[src]
    table[src, col] = dst
    valid[src, col] = True
    return table, valid


def laplacian_smooth(verts: np.ndarray, tris: np.ndarray, iters: int, lam: float = 0.5):
    """tools/particles_to_mesh.py:138-139 (Open3D filter_smooth_laplacian,
    restated): binary64 passes over binary32 vertices, binary32 out."""
    V = np.asarray(verts, np.float32).astype(np.float64)
    if iters <= 0 or len(V) == 0:
        return np.asarray(verts, np.float32)
    table, valid = vertex_neighbours(tris, len(V))
    for _ in range(iters):
        tw = np.zeros(len(V))
        s = np.zeros_like(V)
        for k in range(table.shape[1]):  # sequential sums; padding adds exact zeros
            q = V[table[:, k]]
            d = V - q
            dist = np.sqrt(d[:, 0] * d[:, 0] + d[:, 1] * d[:, 1] + d[:, 2] * d[:, 2])
            w = np.where(valid[:, k], 1.0 / (dist + 1e-12), 0.0)
            tw = tw + w
            s = s + w[:, None] * q
        V = V + lam * (s / tw[:, None] - V)
    return V.astype(np.float32)


def reconstruct_splat(points, radius=0.012, voxel_factor=0.5, sigma=2.5, iso_factor=0.18,
                      swap_yz=True, smooth_pre=2, k=15, smooth_post=2):
    """The whole path on the CPU -> dict(lo, extent, res, grid, iso, vertices, triangles)."""
    pts = np.asarray(points, np.float32).astype(np.float64)
    if swap_yz:
        pts = pts[:, [0, 2, 1]]
    if smooth_pre > 0:
        pts = smooth_positions_knn(pts, k=k, iters=smooth_pre)
    lo, extent, res = grid_setup(pts, radius, voxel_factor)
    grid = splat(pts, lo, extent, res)
    if sigma > 0:
        grid = scipy.ndimage.gaussian_filter(grid, sigma=sigma)
    max_d = grid.max()
    iso = max_d * iso_factor
    out = dict(lo=lo, extent=extent, res=res, grid=grid, max_density=max_d, iso=iso)
    if max_d <= 0:
        out.update(vertices=np.zeros((0, 3), np.float32), triangles=np.zeros((0, 3), np.uint32))
        return out
    v, t = marching_cubes(grid, iso, lo, extent / res)
    v = laplacian_smooth(v, t, smooth_post)
    out.update(vertices=v, triangles=t)
    return out
