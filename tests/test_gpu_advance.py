"""The CFL substep loop (sim_cli's frame loop, apps/sim_cli/main.cpp:171-178)
run by the library: DFSPHSolver.advance against the same loop written on the
host around step(dt).  The time steps (computed on the device by advance,
on the host by the loop), positions, velocities, the attributes advance()
writes back when it returns (it keeps the state in working-slot order
between substeps) and the statistics must agree bit for bit, with the next
substep's grid phase pipelined behind the state read, across first
iteration chunks that do and do not finish the solve, with and without graph
replay, with surface tension, and on the serial paths (divergence solve,
kernel timing, an empty store).
"""
from __future__ import annotations

import numpy as np
import pytest

import water_simulation_amd as ws
from helpers import assert_bitexact, random_box

pytestmark = pytest.mark.gpu


def _solver(x, v, xb, cfg, graphs):
    fluid = ws.ParticleStore(len(x))
    fluid.resize(len(x))
    fluid.set_positions(x)
    fluid.set_velocities(v)
    bnd = ws.ParticleStore(len(xb))
    bnd.resize(len(xb))
    bnd.set_positions(xb)
    s = ws.DFSPHSolver(fluid, bnd, ws.DFSPHConfig(**cfg))
    s.set_graphs(graphs)
    return fluid, bnd, s


def _host_loop(solver, ts, v_floor, time, max_sub, r, iters=None, errs=None):
    dts, rem = [], np.float32(time)
    while rem > np.float32(1e-6) and len(dts) < max_sub:
        dt = np.float32(ts.next_dt(max(solver.max_velocity(), v_floor), r, float(rem)))
        solver.step(float(dt))
        if iters is not None:
            iters.append(solver.stats()["density_iters"])
        if errs is not None:
            errs.append(solver.stats()["mean_density_error"])
        dts.append(dt)
        rem = np.float32(rem - dt)
    return np.array(dts, np.float32)


def _first_chunk(last, ratio):
    # the library's first chunk of density iterations: the previous substep's
    # count, one more when its solve ended above half the tolerance (ratio =
    # its final mean density error over eta rho0), at least 2
    return max(2, min(last + (1 if ratio > 0.5 else 0), 64))


def _misses(iters, errs, tol):
    # substeps whose solve needed more iterations than their first chunk held
    return [k for k in range(len(iters))
            if iters[k] > _first_chunk(iters[k - 1] if k else 0, errs[k - 1] / tol if k else 1.0)]


@pytest.mark.parametrize("graphs", [False, True])
@pytest.mark.parametrize("over", [
    dict(),                                   # iteration counts vary, chunks miss and hit
    dict(max_density_iters=2),                # the first chunk covers every iteration
    dict(enable_divergence_solve=True, max_divergence_iters=10),  # serial path
    dict(enable_surface_tension=True, surface_tension=0.5),      # tension kernels in the loop
])
def test_advance_matches_host_loop(graphs, over):
    x, v, xb, cfg = random_box(n=8000, seed=21, **over)
    ts = ws.TimeStepper(ws.TimeStepperConfig(frame_dt=1 / 24, max_substeps=16))
    r = cfg["particle_radius"]
    fa, ba, a = _solver(x, v, xb, cfg, graphs=False)
    fb, bb, b = _solver(x, v, xb, cfg, graphs=graphs)
    iters, errs = [], []
    for frame in range(2):
        n0 = len(iters)
        want = _host_loop(a, ts, 0.5, 1 / 24, 12, r, iters, errs)
        got, recs = b.advance(1 / 24, 12, ts, v_floor=0.5, with_stats=True)
        assert_bitexact(got, want, f"frame {frame} dts")
        assert [q["density_iters"] for q in recs] == iters[n0:]
        assert_bitexact(fb.get_positions(), fa.get_positions(), f"frame {frame} x")
        assert_bitexact(fb.get_velocities(), fa.get_velocities(), f"frame {frame} v")
        # the attributes advance() writes back when it returns
        for name in ("dfsph_density", "dfsph_alpha", "dfsph_kappa", "dfsph_density_adv"):
            assert_bitexact(fb.get_attribute(name), fa.get_attribute(name), f"frame {frame} {name}")
        sa, sb = a.stats(), b.stats()
        for k in ("density_iters", "divergence_iters", "nan_count"):
            assert sa[k] == sb[k], (frame, k)
        assert sa["mean_density_error"] == sb["mean_density_error"]
        assert sa["max_velocity"] == sb["max_velocity"]
    if not over:
        # some substeps needed more iterations than their first chunk holds
        # (the pipelined grid phase was then superseded), others did not
        misses = _misses(iters, errs, cfg["eta_density"] * cfg["rest_density"])
        assert 0 < len(misses) < len(iters), (iters, errs)


@pytest.mark.parametrize("fast_math", [False, True])
def test_advance_tiny_grid_matches_host_loop(fast_math):
    """A 5 x 5 x 5 cell grid (domain 3 smoothing lengths wide, plus the
    one-cell margin on each side; every length a power of two, so the cell
    count is exact): 7-bit cell keys, so the stable sort behind advance()'s
    key kernel, which counts the first digit itself, is a single pass of
    fewer than 8 bits."""
    rng = np.random.default_rng(3)
    h = 2.0 ** -6
    r = h / 2
    g = np.stack(np.meshgrid(*[np.arange(3)] * 3, indexing="ij"), -1).reshape(-1, 3)
    x = (r + 2 * r * g + rng.uniform(-0.3 * r, 0.3 * r, g.shape)).astype(np.float32)
    v = rng.normal(0.0, 0.3, x.shape).astype(np.float32)
    cfg = dict(rest_density=1000.0, particle_radius=r, smoothing_length=h, viscosity=0.01,
               damping=0.5, domain_min=(0, 0, 0), domain_max=(3 * h,) * 3,
               max_density_iters=20, eta_density=1e-3, fast_math=fast_math)
    stores, solvers = [], []
    for _ in range(2):
        f = ws.ParticleStore(len(x))
        f.resize(len(x))
        f.set_positions(x)
        f.set_velocities(v)
        stores.append(f)
        solvers.append(ws.DFSPHSolver(f, None, ws.DFSPHConfig(**cfg)))
    a, b = solvers
    ts = ws.TimeStepper(ws.TimeStepperConfig(frame_dt=1 / 60, max_substeps=16))
    for frame in range(2):
        want = _host_loop(a, ts, 0.5, 1 / 60, 12, r)
        got = b.advance(1 / 60, 12, ts, v_floor=0.5)
        assert len(got) > 1
        assert_bitexact(got, want, f"frame {frame} dts")
        assert_bitexact(stores[1].get_positions(), stores[0].get_positions(), f"frame {frame} x")
        assert_bitexact(stores[1].get_velocities(), stores[0].get_velocities(), f"frame {frame} v")


def test_advance_single_particle_matches_host_loop():
    """One particle: the sort behind advance()'s key kernel has nothing to
    order (its first-digit counts go unused) and the cell table holds one run."""
    h = 2.0 ** -6
    r = h / 2
    x = np.array([[1.5 * h, 1.5 * h, 1.5 * h]], np.float32)
    v = np.array([[0.2, -0.1, 0.05]], np.float32)
    cfg = dict(rest_density=1000.0, particle_radius=r, smoothing_length=h, viscosity=0.01,
               damping=0.5, domain_min=(0, 0, 0), domain_max=(3 * h,) * 3,
               max_density_iters=20, eta_density=1e-3)
    stores, solvers = [], []
    for _ in range(2):
        f = ws.ParticleStore(1)
        f.resize(1)
        f.set_positions(x)
        f.set_velocities(v)
        stores.append(f)
        solvers.append(ws.DFSPHSolver(f, None, ws.DFSPHConfig(**cfg)))
    a, b = solvers
    ts = ws.TimeStepper(ws.TimeStepperConfig(frame_dt=1 / 60, max_substeps=16))
    want = _host_loop(a, ts, 0.5, 1 / 60, 12, r)
    got = b.advance(1 / 60, 12, ts, v_floor=0.5)
    assert len(got) > 1
    assert_bitexact(got, want, "dts")
    assert_bitexact(stores[1].get_positions(), stores[0].get_positions(), "x")
    assert_bitexact(stores[1].get_velocities(), stores[0].get_velocities(), "v")


@pytest.mark.parametrize("fast_math", [True, False])
def test_advance_full_size_dam_break_matches_host_loop(fast_math):
    """The benchmark workload itself -- 1,000,000 fluid and 1,209,618 wall
    particles on the 16.4 M-cell grid (24-bit keys: three sort passes over
    489 tiles), kappa/rho scaling -- for 24 CFL substeps: advance() against
    the host loop around step(dt), bit for bit.  At the start about a quarter
    of the 256-slot tiles have neighbourhoods too large to stage, so both
    list formats are in use."""
    import sys
    from pathlib import Path

    sys.path.insert(0, str(Path(__file__).resolve().parents[1]))
    import bench

    x, xb, cfg = bench.make_workload(100)
    cfg = dict(cfg, pressure_form=1, fast_math=fast_math)
    v = np.zeros_like(x)
    ts = ws.TimeStepper()
    r = cfg["particle_radius"]
    fa, ba, a = _solver(x, v, xb, cfg, graphs=False)
    fb, bb, b = _solver(x, v, xb, cfg, graphs=False)
    iters, errs = [], []
    want = _host_loop(a, ts, 0.5, 1.0, 24, r, iters, errs)
    got, recs = b.advance(1.0, 24, ts, v_floor=0.5, with_stats=True)
    assert len(got) == 24
    assert_bitexact(got, want, "1M dts")
    assert [q["density_iters"] for q in recs] == iters
    assert_bitexact(fb.get_positions(), fa.get_positions(), "1M x")
    assert_bitexact(fb.get_velocities(), fa.get_velocities(), "1M v")
    for name in ("dfsph_density", "dfsph_alpha", "dfsph_kappa", "dfsph_density_adv"):
        assert_bitexact(fb.get_attribute(name), fa.get_attribute(name), f"1M {name}")
    sa, sb = a.stats(), b.stats()
    for k in ("density_iters", "nan_count", "escapees", "mean_density_error", "max_velocity"):
        assert sa[k] == sb[k], k
    # binary64 sums over blocks of slots (advance) or of indices (step)
    for k in ("total_mass", "kinetic_energy"):
        assert abs(sa[k] - sb[k]) <= 1e-12 * abs(sa[k]), k
    misses = _misses(iters, errs, cfg["eta_density"] * cfg["rest_density"])
    print(f"\n1M advance (fast_math={fast_math}): iterations {iters}, first-chunk misses {misses}")
    for s in (a, b):
        s.close()


def test_advance_stops_on_time_and_count():
    x, v, xb, cfg = random_box(n=1000, seed=4)
    _, _, s = _solver(x, v, xb, cfg, graphs=True)
    ts = ws.TimeStepper()
    dts = s.advance(1e-3, 100, ts, v_floor=0.5)
    assert 0 < len(dts) < 100
    assert abs(float(np.sum(dts.astype(np.float64))) - 1e-3) < 1e-6
    assert len(s.advance(1.0, 3, ts)) == 3
    assert len(s.advance(0.0, 3, ts)) == 0


def test_advance_with_timing_and_empty_store():
    x, v, xb, cfg = random_box(n=1000, seed=6)
    fa, _, a = _solver(x, v, xb, cfg, graphs=False)
    fb, _, b = _solver(x, v, xb, cfg, graphs=False)
    b.set_timing(True)
    ts = ws.TimeStepper()
    want = _host_loop(a, ts, 0.5, 5e-3, 6, cfg["particle_radius"])
    got = b.advance(5e-3, 6, ts, v_floor=0.5)
    assert_bitexact(got, want, "timed dts")
    assert_bitexact(fb.get_positions(), fa.get_positions(), "timed x")
    empty = ws.ParticleStore(10)
    s = ws.DFSPHSolver(empty, None, ws.DFSPHConfig(**cfg))
    assert len(s.advance(1e-3, 4, ts, v_floor=0.5)) >= 1
