"""GPU parity of the neighbour-search front end: cell keys, the stable sort
permutation, cell_start, Morton codes -- bit-exact against the oracle.

Mirrors tests/unit/test_cell_grid.cu and tests/unit/test_spatial_accel.cu of
the reference and extends them to random clouds, out-of-grid and NaN points,
and the 1M-particle / 16.4M-cell grid of the large dam break.
"""
from __future__ import annotations

import numpy as np
import pytest

import water_simulation_amd as ws
from helpers import assert_bitexact
from oracle import oracle as O

pytestmark = pytest.mark.gpu


def gpu_grid(x, origin, dims, cell):
    g = ws.CellGrid(origin, dims, cell)
    g.reserve(max(len(x), 1))
    buf = ws.DeviceBuffer.from_array(np.asarray(x, np.float32).reshape(-1, 3))
    g.build(buf.ptr, len(x))
    ws.device_synchronize()
    return g, buf


def check_grid(x, origin, dims, cell, tag):
    g, _buf = gpu_grid(x, origin, dims, cell)
    keys, perm, cs = O.grid_build(x, origin, dims, cell)
    assert_bitexact(g.read_sorted_keys(), keys, f"{tag} sorted keys")
    assert_bitexact(g.read_sorted_indices(), perm, f"{tag} permutation")
    assert_bitexact(g.read_cell_start(), cs, f"{tag} cell_start")
    # independent closed forms: lexicographic (key, id) order, lower_bound table
    k_in = O.cell_keys(x, origin, dims, cell)
    order = np.lexsort((np.arange(len(x)), k_in))
    assert np.array_equal(order.astype(np.uint32), perm)
    assert np.array_equal(np.searchsorted(keys, np.arange(len(cs)), "left").astype(np.uint32), cs)


def test_eight_grid_aligned_points_map_to_distinct_cells():
    # tests/unit/test_cell_grid.cu:9-47
    pts = np.array([[a, b, c] for c in (0.25, 0.75) for b in (0.25, 0.75) for a in (0.25, 0.75)],
                   np.float32)
    g, _ = gpu_grid(pts, (0, 0, 0), (4, 4, 4), 0.5)
    perm = g.read_sorted_indices()
    assert sorted(perm.tolist()) == list(range(8))
    cs = g.read_cell_start()
    assert np.all(np.diff(cs.astype(np.int64)) >= 0)
    check_grid(pts, (0, 0, 0), (4, 4, 4), 0.5, "8pts")


def test_out_of_bounds_particle_goes_to_sentinel():
    # tests/unit/test_cell_grid.cu:49-74
    pts = np.array([[0.25, 0.25, 0.25], [99.0, 99.0, 99.0]], np.float32)
    g, _ = gpu_grid(pts, (0, 0, 0), (2, 2, 2), 0.5)
    cs = g.read_cell_start()
    assert cs[-1] == 1
    assert cs[0] == 0
    check_grid(pts, (0, 0, 0), (2, 2, 2), 0.5, "oob")


@pytest.mark.parametrize("n,seed", [(1, 0), (2, 1), (100, 2), (5000, 3), (100_000, 4)])
def test_random_clouds_match_oracle(n, seed):
    rng = np.random.default_rng(seed)
    origin, dims, cell = (-0.02, -0.02, -0.02), (52, 52, 52), 0.02
    x = rng.uniform(-0.1, 1.1, (n, 3)).astype(np.float32)  # some outside the grid
    if n >= 100:
        # exact cell faces, duplicates, huge values and NaN
        x[: n // 10] = np.round(x[: n // 10] / cell) * cell
        x[n // 10: n // 10 + 5] = x[0]
        x[-1] = [np.nan, 0.5, 0.5]
        x[-2] = [1e30, -1e30, 0.5]
    check_grid(x, origin, dims, cell, f"cloud{n}")


def test_dense_block_all_in_few_cells():
    # many particles per cell, long runs of equal keys across sort tiles
    rng = np.random.default_rng(7)
    x = rng.uniform(0.30, 0.34, (20_000, 3)).astype(np.float32)
    check_grid(x, (0, 0, 0), (10, 10, 10), 0.1, "dense")


def test_large_dam_break_grid():
    # 1,000,000 particles on the 402 x 202 x 202 grid of the large dam break
    i = np.arange(100, dtype=np.float32)
    g = np.stack(np.meshgrid(i, i, i, indexing="ij"), -1).reshape(-1, 3)
    x = (np.float32(0.015) + np.float32(0.005) * g).astype(np.float32)
    origin, dims = O.make_grid((0, 0, 0), (2, 1, 1), 0.005)
    check_grid(x, origin, dims, 0.005, "cfg2")


@pytest.mark.parametrize("n", [8, 1000, 300_000])
def test_morton_codes_and_sort_match_oracle(n):
    rng = np.random.default_rng(n)
    x = rng.uniform(-0.05, 1.05, (n, 3)).astype(np.float32)
    if n == 8:  # tests/unit/test_spatial_accel.cu:9-41
        x = np.array([[a, b, c] for c in (0.1, 0.9) for b in (0.1, 0.9) for a in (0.1, 0.9)],
                     np.float32)
    acc = ws.SpatialAccel()
    buf = ws.DeviceBuffer.from_array(x)
    acc.build(buf.ptr, n, (0, 0, 0), (1, 1, 1))
    ws.device_synchronize()
    assert acc.leaf_count() == n
    codes, sorted_codes, perm = O.morton(x, (0, 0, 0), (1, 1, 1))
    assert_bitexact(acc.read_sorted_indices(), perm, "morton permutation")
    assert_bitexact(acc.read_morton_codes(), sorted_codes, "morton codes")


def test_morton_single_particle():
    # tests/unit/test_spatial_accel.cu:43-55
    acc = ws.SpatialAccel()
    buf = ws.DeviceBuffer.from_array(np.array([[0.5, 0.5, 0.5]], np.float32))
    acc.build(buf.ptr, 1, (0, 0, 0), (1, 1, 1))
    ws.device_synchronize()
    assert acc.leaf_count() == 1
    assert acc.read_sorted_indices().tolist() == [0]


@pytest.mark.parametrize("case,n", [("random", 8), ("random", 5000), ("random", 300_000),
                                    ("duplicates", 20_000), ("identical", 513), ("two", 2)])
def test_lbvh_nodes_and_boxes_match_oracle(case, n):
    # spatial_accel.cu:44-113 (Karras internal nodes, leaf boxes) plus the
    # bottom-up internal boxes; integers and floats bit for bit
    rng = np.random.default_rng(n)
    if case == "random":
        x = rng.uniform(-0.05, 1.05, (n, 3)).astype(np.float32)
    elif case == "duplicates":
        x = (rng.integers(0, 6, (n, 3)) / 6 + 0.02).astype(np.float32)
    elif case == "identical":
        x = np.full((n, 3), 0.3, np.float32)
    else:
        x = np.array([[0.2, 0.2, 0.2], [0.8, 0.1, 0.3]], np.float32)
    acc = ws.SpatialAccel()
    buf = ws.DeviceBuffer.from_array(x)
    acc.build(buf.ptr, n, (0, 0, 0), (1, 1, 1))
    ws.device_synchronize()
    codes, sc, perm = O.morton(x, (0, 0, 0), (1, 1, 1))
    wl, wr, wp, wb = O.lbvh(x, sc, perm)
    gl, gr, gp = acc.read_internal_nodes()
    assert_bitexact(gl, wl, "left")
    assert_bitexact(gr, wr, "right")
    assert_bitexact(gp, wp, "parent")
    assert_bitexact(acc.read_node_aabbs(), wb, "boxes")
    # rebuilding into the same object clears the arrival counters
    acc.build(buf.ptr, n, (0, 0, 0), (1, 1, 1))
    ws.device_synchronize()
    assert_bitexact(acc.read_node_aabbs(), wb, "boxes, second build")


def test_lbvh_single_leaf():
    acc = ws.SpatialAccel()
    buf = ws.DeviceBuffer.from_array(np.array([[0.5, 0.25, 0.75]], np.float32))
    acc.build(buf.ptr, 1, (0, 0, 0), (1, 1, 1))
    ws.device_synchronize()
    l, r, p = acc.read_internal_nodes()
    assert len(l) == 0 and len(r) == 0 and p.tolist() == [0xFFFFFFFF]
    assert acc.read_node_aabbs().tolist() == [[[0.5, 0.25, 0.75], [0.5, 0.25, 0.75]]]
