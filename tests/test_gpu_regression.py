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


def benchmark_puddle():
    """The settled puddle as the benchmark set states it: a 17^3 block
    (lattice 0.02 m, r = 0.01) at rest with its bottom 0.34 m above the
    floor, walls at 0.01 m x 3 layers, XSPH 0.01, damping 0.5 (the
    reference's default, dfsph.h:24), eta 0.1 %, CFL dt <= 0.004 s
    (TimeStepper, v floor 0.5 m/s), 1.5 s."""
    r = 0.01
    i = np.arange(17, dtype=np.float32)
    g = np.stack(np.meshgrid(i, i, i, indexing="ij"), -1).reshape(-1, 3)
    x = (np.float32(0.34) + np.float32(0.02) * g).astype(np.float32)  # bottom 0.34 m
    xb = ws.sample_aabb_boundary((0, 0, 0), (1, 1, 1), 0.01, 3)
    assert len(xb) == 183618
    cfg = dict(rest_density=1000.0, particle_radius=r, smoothing_length=2 * r, viscosity=0.01,
               damping=0.5, max_density_iters=100, eta_density=1e-3)
    ts = ws.TimeStepper(ws.TimeStepperConfig(dt_max=0.004))
    return x, xb, cfg, (lambda vmax: ts.next_dt(max(vmax, 0.5), r, 1.0))


def run_recorded(x, xb, cfg, dt_rule, sim_time):
    """run_puddle with every substep's statistics kept."""
    fluid = ws.ParticleStore(len(x))
    fluid.resize(len(x))
    fluid.set_positions(x)
    bnd = ws.ParticleStore(len(xb))
    bnd.resize(len(xb))
    bnd.set_positions(xb)
    solver = ws.DFSPHSolver(fluid, bnd, ws.DFSPHConfig(**cfg))
    solver.set_graphs(True)
    t, recs = 0.0, []
    while t < sim_time:
        dt = dt_rule(solver.max_velocity())
        if t + dt > sim_time:
            dt = sim_time - t
        solver.step(dt)
        t += dt
        recs.append(solver.stats())
    return fluid, solver, recs


def assert_regression_run(recs, eta_rho0, tag):
    """Per substep: no non-finite particle, no escapee (no particle reached
    the domain limit, i.e. got through the wall layers), and the density
    solve met eta."""
    nan = [k for k, q in enumerate(recs) if q["nan_count"]]
    esc = [k for k, q in enumerate(recs) if q["escapees"]]
    over = [k for k, q in enumerate(recs) if not q["mean_density_error"] <= eta_rho0]
    assert not nan, f"{tag}: non-finite particles from substep {nan[0]}"
    assert not esc, f"{tag}: escapees from substep {esc[0]} ({recs[esc[0]]['escapees']})"
    assert not over, (f"{tag}: mean density error above eta at {len(over)} substeps, first "
                      f"{over[0]}: {recs[over[0]]['mean_density_error']:.4f}")


def test_settled_puddle_benchmark_configuration():
    """The benchmark-set puddle with kappa/rho scaling and the divergence
    solve: the reference's checks (1)-(3) on the final state, and at every
    substep no non-finite particle, no escapee and the density error within
    eta.  The fluid is still sloshing at 1.5 s (damping 0.5/s: the drop's
    energy decays slowly), so the depth band of check (4) is asserted after
    the run is continued to 4 s."""
    x, xb, cfg, rule = benchmark_puddle()
    cfg.update(pressure_form=1, enable_divergence_solve=True, max_divergence_iters=100,
               eta_divergence=1e-3)
    eta_rho0 = cfg["eta_density"] * cfg["rest_density"]
    fluid, solver, recs = run_recorded(x, xb, cfg, rule, 1.5)
    assert_regression_run(recs, eta_rho0, "benchmark puddle")
    p = fluid.get_positions()
    v = fluid.get_velocities()
    speed = np.sqrt((v.astype(np.float64) ** 2).sum(1))
    assert np.isfinite(p).all()                 # (1)
    assert speed.max() < 5.0                    # (2)
    assert p[:, 1].max() < 0.70 + 0.05          # (3) no escapees above the drop height
    assert p[:, 1].min() >= 0.0                 # floor held
    its = np.array([q["density_iters"] for q in recs])
    dits = np.array([q["divergence_iters"] for q in recs])
    ke = [q["kinetic_energy"] for q in recs]
    print(f"\nbenchmark puddle (kappa/rho + divergence solve): {len(recs)} substeps, mean "
          f"density iters {its.mean():.2f} (max {its.max()}), divergence iters {dits.mean():.2f}, "
          f"final mean density error {recs[-1]['mean_density_error']:.4f} <= {eta_rho0:.1f}, "
          f"max |v| {speed.max():.3f}, mean y {p[:, 1].mean():.4f}, kinetic energy "
          f"{max(ke):.2f} J peak -> {ke[-1]:.3f} J")


def test_settled_puddle_benchmark_configuration_settles_by_4s():
    x, xb, cfg, rule = benchmark_puddle()
    cfg.update(pressure_form=1, enable_divergence_solve=True, max_divergence_iters=100,
               eta_divergence=1e-3)
    fluid, solver, recs = run_recorded(x, xb, cfg, rule, 4.0)
    assert_regression_run(recs, cfg["eta_density"] * cfg["rest_density"], "puddle to 4 s")
    ymean, vmax = check_settled(fluid, len(x), cfg["particle_radius"], wall_top=0.02)
    print(f"\nbenchmark puddle at 4 s: mean y {ymean:.4f}, max |v| {vmax:.3f}, kinetic energy "
          f"{recs[-1]['kinetic_energy']:.3f} J")


@pytest.mark.xfail(strict=True, reason=(
    "the reference's kappa/rho^2 scaling removes about 1/rho of the density excess per sweep "
    "(each substep runs into the 100-iteration cap): on this configuration the run ends with "
    "1215 substeps at 92.98 mean density iterations, a final mean density error of 17.28 "
    "kg/m^3 against eta*rho0 = 1.0 and max |v| at the 5.000 m/s clamp"))
def test_settled_puddle_benchmark_configuration_reference_pressure_form():
    x, xb, cfg, rule = benchmark_puddle()
    fluid, solver, recs = run_recorded(x, xb, cfg, rule, 1.5)
    print(f"\nbenchmark puddle (kappa/rho^2): {len(recs)} substeps, mean density iters "
          f"{np.mean([q['density_iters'] for q in recs]):.2f}, final mean density error "
          f"{recs[-1]['mean_density_error']:.4f}, max |v| {solver.max_velocity():.3f}")
    assert recs[-1]["mean_density_error"] <= cfg["eta_density"] * cfg["rest_density"]
    assert solver.max_velocity() < 5.0


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


def test_mid_dam_break_matches_oracle():
    """The mid dam break (128,000 fluid particles, 210,378 wall particles,
    162 x 92 x 82 cells), kappa/rho scaling, graph replay: two substeps on
    the HIP path and in the binary32 oracle, every field and statistic bit
    for bit."""
    import sys
    from pathlib import Path

    sys.path.insert(0, str(Path(__file__).resolve().parents[1]))
    import bench

    x, xb, cfg = bench.make_mid_workload()
    cfg = dict(cfg, pressure_form=1)
    p = Pair(x, None, xb, cfg)
    p.gpu.set_graphs(True)
    for k, dt in enumerate((1e-3, 8e-4)):
        p.step(dt)
        p.compare(f"128k substep {k}")


def test_mid_dam_break_two_seconds_regression():
    """2 s of the mid dam break (CFL substeps through the library's frame
    loop, kappa/rho scaling): at every substep no non-finite particle, no
    escapee and the density solve within eta; mass and momentum are finite
    and the mass is that of every particle.  (The surge hits the far wall at
    about 0.6 s and splashes: speeds reach the 5 m/s clamp of the advection,
    dfsph_kernels.cu:367-371, so the settled puddle's speed bound is not a
    check here.)"""
    import sys
    from pathlib import Path

    sys.path.insert(0, str(Path(__file__).resolve().parents[1]))
    import bench

    x, xb, cfg = bench.make_mid_workload()
    cfg = dict(cfg, pressure_form=1)
    fluid = ws.ParticleStore(len(x))
    fluid.resize(len(x))
    fluid.set_positions(x)
    bnd = ws.ParticleStore(len(xb))
    bnd.resize(len(xb))
    bnd.set_positions(xb)
    solver = ws.DFSPHSolver(fluid, bnd, ws.DFSPHConfig(**cfg))
    solver.set_graphs(True)
    ts = ws.TimeStepper()
    recs, t = [], 0.0
    frame = 1.0 / 24.0
    while t < 2.0 - 1e-6:
        dts, rec = solver.advance(min(frame, 2.0 - t), 64, ts, v_floor=0.5, with_stats=True)
        t += float(np.sum(dts.astype(np.float64)))
        recs += rec
    assert_regression_run(recs, cfg["eta_density"] * cfg["rest_density"], "mid dam break")
    v = fluid.get_velocities().astype(np.float64)
    speed = np.sqrt((v ** 2).sum(1))
    its = np.array([q["density_iters"] for q in recs])
    print(f"\nmid dam break, 2 s: {len(recs)} substeps, mean density iters {its.mean():.2f} "
          f"(max {its.max()}), final max |v| {speed.max():.3f}, kinetic energy "
          f"{recs[-1]['kinetic_energy']:.2f} J, momentum {recs[-1]['momentum']}")
    assert speed.max() <= 5.0 * (1 + 1e-6)
    assert np.isfinite(recs[-1]["momentum"]).all()
    assert recs[-1]["total_mass"] == pytest.approx(len(x) * solver.particle_mass(), rel=1e-9)
