"""Long-run regression: the reference's settled-puddle oracle
(tests/regression/test_settled_puddle.cu:119-222) on the HIP path, in the
reference's parameters and in the settled-puddle configuration of the
benchmark set (walls at 0.01 m, XSPH 0.01, eta 0.1 %, CFL dt <= 0.004 s),
plus a whole-run bitwise comparison with the binary32 oracle.
"""
from __future__ import annotations

import numpy as np
import pytest

import water_simulation_amd as ws
from helpers import Pair, assert_bitexact, puddle_ref
from oracle import oracle as O

pytestmark = pytest.mark.gpu


def run_puddle(x, xb, cfg, dt_rule, sim_time=1.5, check_every=250, pair=None):
    if pair is None:
        fluid = ws.ParticleStore(len(x))
        fluid.resize(len(x))
        fluid.set_positions(x)
        bnd = ws.ParticleStore(len(xb))
        bnd.resize(len(xb))
        bnd.set_positions(xb)
        solver = ws.DFSPHSolver(fluid, bnd, ws.DFSPHConfig(**cfg))
    else:
        fluid, solver = pair.fluid, pair.gpu
    t, steps, errs, iters = 0.0, 0, [], []
    while t < sim_time:
        dt = dt_rule(solver.max_velocity())
        if t + dt > sim_time:
            dt = sim_time - t
        solver.step(dt)
        if pair is not None:
            pair.ref.step(dt)
        t += dt
        steps += 1
        st = solver.stats()
        errs.append(st["mean_density_error"])
        iters.append(st["density_iters"])
        if pair is not None and steps % check_every == 0:
            pair.compare(f"puddle t={t:.3f}")
    return fluid, solver, steps, np.array(errs), np.array(iters)


def check_settled(fluid, n, r, wall_top=0.048):
    """Checks (1)-(3) of tests/regression/test_settled_puddle.cu:197-206 and the
    floor; the depth check uses the floor the three inward wall layers leave
    (top layer at y = 0.048, open floor (1 - 2*0.048)^2 m^2)."""
    p = fluid.get_positions()
    v = fluid.get_velocities()
    finite = np.isfinite(p).all(1)
    assert finite.sum() == n                        # (1) no NaN
    speed = np.sqrt((v.astype(np.float64) ** 2).sum(1))
    assert speed.max() < 5.0                        # (2) bounded
    assert p[:, 1].max() < 0.70 + 0.05              # (3) no escapees above the drop height
    assert p[:, 1].min() >= -0.01                   # floor held
    depth = n * (2 * r) ** 3 / (1.0 - 2 * wall_top) ** 2
    assert p[:, 1].mean() == pytest.approx(wall_top + depth / 2, rel=0.5)
    return p[:, 1].mean(), speed.max()


def reference_depth_band(n, r):
    # check (4) as written (test_settled_puddle.cu:208-221): floor area 1 m^2 at y = 0
    return n * (2 * r) ** 3 / 1.0 / 2.0


def puddle_dt(r):
    return lambda vmax: max(min(0.001, 0.3 * r / max(vmax, 0.5)), 1e-5)


def test_settled_puddle_reference_parameters():
    x, _, xb, cfg = puddle_ref()
    r = cfg["particle_radius"]
    fluid, solver, steps, errs, iters = run_puddle(x, xb, cfg, puddle_dt(r))
    ymean, vmax = check_settled(fluid, len(x), r)
    print(f"\nsettled puddle (reference parameters): {steps} substeps, mean y {ymean:.4f} "
          f"(reference band {reference_depth_band(len(x), r):.4f} +- 50 %), max |v| {vmax:.3f}, "
          f"mean density iters {iters.mean():.2f}, final mean density error {errs[-1]:.3f} "
          f"(eta*rho0 = {cfg['eta_density'] * 1000:.1f})")
    assert np.isfinite(errs).all()


@pytest.mark.xfail(strict=True, reason=(
    "check (4) of test_settled_puddle.cu (mean y within 50 % of N(2r)^3/2 over a 1 m^2 floor at "
    "y = 0) ignores the three inward wall layers that lift the floor to y = 0.048; the restated "
    "reference algorithm (bit-identical on GPU and CPU oracle, Jacobi or in-place XSPH alike) "
    "settles at mean y ~ 0.105"))
def test_settled_puddle_literal_depth_check():
    x, _, xb, cfg = puddle_ref()
    r = cfg["particle_radius"]
    fluid, _, _, _, _ = run_puddle(x, xb, cfg, puddle_dt(r))
    assert fluid.get_positions()[:, 1].mean() == pytest.approx(reference_depth_band(len(x), r),
                                                               rel=0.5)


def test_settled_puddle_bender_koschier_meets_eta():
    # kappa/rho pressure scaling plus the divergence solve: the density error
    # reaches eta, which the reference's kappa/rho^2 scaling cannot: each of its
    # sweeps removes about 1/rho of the density excess
    x, _, xb, cfg = puddle_ref()
    r = cfg["particle_radius"]
    cfg.update(pressure_form=1, enable_divergence_solve=True, max_divergence_iters=10,
               eta_divergence=5e-3)
    fluid, solver, steps, errs, iters = run_puddle(x, xb, cfg, puddle_dt(r))
    ymean, vmax = check_settled(fluid, len(x), r)
    eta_rho0 = cfg["eta_density"] * cfg["rest_density"]
    assert errs[-1] <= eta_rho0
    print(f"\nsettled puddle (kappa/rho + divergence solve): {steps} substeps, mean y {ymean:.4f}, "
          f"max |v| {vmax:.3f}, mean density iters {iters.mean():.2f}, final mean density error "
          f"{errs[-1]:.4f} <= {eta_rho0:.1f}")


def test_settled_puddle_benchmark_configuration():
    # 17^3 block (lattice 0.02 m, r = 0.01), walls at 0.01 m x 3 layers, XSPH
    # 0.01, eta 0.1 %, CFL dt <= 0.004 s (TimeStepper, v floor 0.5 m/s)
    r = 0.01
    i = np.arange(17, dtype=np.float32)
    g = np.stack(np.meshgrid(i, i, i, indexing="ij"), -1).reshape(-1, 3)
    x = (np.float32(0.34) + np.float32(0.02) * g).astype(np.float32)  # bottom 0.34 m
    xb = ws.sample_aabb_boundary((0, 0, 0), (1, 1, 1), 0.01, 3)
    assert len(xb) == 183618
    cfg = dict(rest_density=1000.0, particle_radius=r, smoothing_length=2 * r, viscosity=0.01,
               damping=0.5, max_density_iters=100, eta_density=1e-3)
    ts = ws.TimeStepper(ws.TimeStepperConfig(dt_max=0.004))
    fluid, solver, steps, errs, iters = run_puddle(
        x, xb, cfg, lambda vmax: ts.next_dt(max(vmax, 0.5), r, 1.0))
    p = fluid.get_positions()
    assert np.isfinite(p).all()
    assert p[:, 1].min() >= -0.01 and p[:, 1].max() < 1.0
    print(f"\nsettled puddle (benchmark configuration): {steps} substeps, mean density iters "
          f"{iters.mean():.2f}, final mean density error {errs[-1]:.4f} (eta*rho0 = 1.0), "
          f"max |v| {solver.max_velocity():.3f}")


def test_settled_puddle_whole_run_bitwise_against_oracle():
    x, v, xb, cfg = puddle_ref()
    r = cfg["particle_radius"]
    pair = Pair(x, v, xb, cfg)
    run_puddle(None, None, cfg, puddle_dt(r), sim_time=0.6, check_every=100, pair=pair)
    pair.compare("puddle end")


def test_full_size_dam_break_matches_oracle():
    """The benchmark workload itself -- 1,000,000 fluid and 1,209,618 wall
    particles on the 16.4 M-cell grid, kappa/rho scaling -- for two substeps
    on the HIP path (graph replay) and in the binary32 oracle: every field
    and statistic bit for bit."""
    import sys
    from pathlib import Path

    sys.path.insert(0, str(Path(__file__).resolve().parents[1]))
    import bench

    x, xb, cfg = bench.make_workload(100)
    cfg = dict(cfg, pressure_form=1)
    p = Pair(x, None, xb, cfg)
    p.gpu.set_graphs(True)
    for k, dt in enumerate((2e-4, 1.5e-4)):
        p.step(dt)
        p.compare(f"1M substep {k}")
