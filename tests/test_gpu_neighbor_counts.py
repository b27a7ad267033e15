"""Neighbour counts of the HIP density/alpha walk against the oracle.

ITER_NEIGHBORS (solvers/dfsph/src/dfsph_kernels.cu:13-35) visits, for every
fluid particle, the slots of the in-grid cells of its 27-cell stencil in the
fluid grid and in the boundary grid.  Per particle and per store: cand =
visited slots (self included), nbr = visited slots with sqrtf(|r|^2) / l < 2
in binary32.  The solver's counting walk (wsim_solver_neighbor_counts, the
density/alpha kernel with its counters on) must equal orc_neighbor_counts
bit for bit; the oracle's counts are pinned against an O(N^2) brute force in
test_oracle_pinning.py.
"""
from __future__ import annotations

import numpy as np
import pytest

import water_simulation_amd as ws
from helpers import puddle_ref, random_box
from oracle import oracle as O

pytestmark = pytest.mark.gpu


def solver_for(x, v, xb, cfg):
    n = len(x)
    f = ws.ParticleStore(max(n, 1))
    f.resize(n)
    if n:
        f.set_positions(x)
        if v is not None:
            f.set_velocities(v)
    b = None
    if xb is not None:
        b = ws.ParticleStore(len(xb))
        b.resize(len(xb))
        b.set_positions(xb)
    return ws.DFSPHSolver(f, b, ws.DFSPHConfig(**cfg)), f, b


def oracle_counts(x, xb, cfg):
    h = cfg["smoothing_length"]
    origin, dims = O.make_grid(cfg.get("domain_min", (0, 0, 0)), cfg.get("domain_max", (1, 1, 1)), h)
    cf, nf = O.neighbor_counts(x, x, origin, dims, h, h)
    if xb is None:
        return cf, nf, np.zeros_like(cf), np.zeros_like(nf)
    cb, nb = O.neighbor_counts(x, xb, origin, dims, h, h)
    return cf, nf, cb, nb


def check(x, v, xb, cfg, tag):
    s, f, _ = solver_for(x, v, xb, cfg)
    got = s.neighbor_counts()
    want = oracle_counts(x, xb, cfg)
    names = ("cand_fluid", "nbr_fluid", "cand_boundary", "nbr_boundary")
    for k, w in zip(names, want):
        g = got[k]
        bad = np.nonzero(g != w)[0]
        assert bad.size == 0, (f"{tag} {k}: {bad.size}/{len(g)} differ; first particle {bad[0]} "
                               f"got {g[bad[0]]} want {w[bad[0]]}")
    # a particle inside the grid visits its own cell: nbr_fluid counts itself
    if len(x):
        h = cfg["smoothing_length"]
        origin, dims = O.make_grid(cfg.get("domain_min", (0, 0, 0)),
                                   cfg.get("domain_max", (1, 1, 1)), h)
        inside = O.cell_keys(x, origin, dims, h) < np.prod(dims)
        assert (got["nbr_fluid"][inside] >= 1).all(), tag
    return got


def test_counts_tolerance_mode_solver():
    # a tolerance-mode solver keeps no boundary pair lists; its counting walk
    # still counts both stores exactly, and its next substep is unaffected
    x, v, xb, cfg = random_box(n=3000, seed=5)
    check(x, v, xb, dict(cfg, fast_math=True), "tolerance mode")
    s, f, _ = solver_for(x, v, xb, dict(cfg, fast_math=True))
    s.neighbor_counts()
    s.step(0.002)
    t, g, _ = solver_for(x, v, xb, dict(cfg, fast_math=True))
    t.step(0.002)
    assert np.array_equal(f.get_positions().view(np.uint32), g.get_positions().view(np.uint32))


@pytest.mark.parametrize("seed", [0, 1, 2])
def test_counts_random_box(seed):
    x, v, xb, cfg = random_box(n=4000, seed=seed)
    got = check(x, v, xb, cfg, f"random box {seed}")
    print(f"\nrandom box {seed}: mean cand {got['cand_fluid'].mean():.2f}, "
          f"nbr {got['nbr_fluid'].mean():.2f}, boundary cand {got['cand_boundary'].mean():.2f}")


def test_counts_random_cloud_with_outliers():
    # clustered, coincident, face-exact and out-of-grid (clamped into the
    # domain only by advection, so here they sit outside the grid) points
    rng = np.random.default_rng(7)
    r = 0.01
    x = rng.uniform(0.05, 0.75, (3000, 3)).astype(np.float32)
    x[:300] = x[300:600]                                   # coincident pairs
    x[600:700] = np.round(x[600:700] / (2 * r)) * (2 * r)  # on cell faces
    x[700:720] = np.float32(0.4)                           # 20 at one point
    x[720:730, 0] = np.float32(5.0)                        # outside the grid
    x[730:735, 1] = np.float32(-3.0)
    _, v, xb, cfg = random_box(n=27, seed=0)
    check(x, None, xb, cfg, "outlier cloud")


def test_counts_no_boundary():
    x, v, _, cfg = random_box(n=3000, seed=4, boundary=False)
    got = check(x, v, None, cfg, "no boundary")
    assert not got["cand_boundary"].any()


def test_counts_settled_puddle_case():
    x, v, xb, cfg = puddle_ref()
    check(x, v, xb, cfg, "puddle start")
    # after some substeps (positions off the lattice), same comparison
    s, f, _ = solver_for(x, v, xb, cfg)
    for _ in range(40):
        s.step(1e-3)
    xt = f.get_positions()
    check(xt, f.get_velocities(), xb, cfg, "puddle after 40 substeps")


def test_counts_single_particle_and_empty():
    _, _, xb, cfg = puddle_ref()
    got = check(np.array([[0.5, 0.5, 0.5]], np.float32), None, xb, cfg, "one particle")
    assert got["cand_fluid"][0] == 1 and got["nbr_fluid"][0] == 1
    s, _, _ = solver_for(np.zeros((0, 3), np.float32), None, xb, cfg)
    got = s.neighbor_counts()
    assert all(len(a) == 0 for a in got.values())


def test_counts_full_size_dam_break():
    """The benchmark workload: 1,000,000 fluid particles and 1,209,618 wall
    particles on the 402 x 202 x 202 grid, at rest and after two substeps."""
    import sys
    from pathlib import Path

    sys.path.insert(0, str(Path(__file__).resolve().parents[1]))
    import bench

    x, xb, cfg = bench.make_workload(100)
    cfg = dict(cfg, pressure_form=1)
    got = check(x, None, xb, cfg, "1M start")
    s, f, _ = solver_for(x, None, xb, cfg)
    s.step(2e-4)
    s.step(2e-4)
    check(f.get_positions(), None, xb, cfg, "1M after 2 substeps")
    print(f"\n1M dam break: cand {got['cand_fluid'].mean():.2f} nbr {got['nbr_fluid'].mean():.2f} "
          f"boundary cand {got['cand_boundary'].mean():.3f} nbr {got['nbr_boundary'].mean():.3f}")
