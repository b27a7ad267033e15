"""Tolerance mode of the neighbour kernels (DFSPHConfig.fast_math) against
the binary64 oracle.

The reference builds its kernels with -use_fast_math (CMakeLists.txt:56);
the tolerance mode evaluates the same formulas that way (hardware reciprocal
square root and reciprocal, fused multiply-adds) instead of the bit-exact
IEEE binary32 sequence.  Its fields are checked per substep, from the same
input state, against the binary64 oracle:

  * density, alpha and positions within 1e-5 relative;
  * kappa, velocity and the predicted density, which carry cancellation,
    within 1e-4 of their field's scale -- the bound the bit-exact binary32
    evaluation itself meets -- and within 4x the binary32 oracle's own
    error from the same state plus 1e-6 of the scale;
  * the same density-solve iteration count as the binary64 oracle.

Whole runs in tolerance mode pass the same regression checks as the exact
path (no non-finite particle, no escapee, density error within eta).
"""
from __future__ import annotations

import numpy as np
import pytest

import water_simulation_amd as ws
from helpers import puddle_ref, random_box
from oracle import oracle as O

pytestmark = pytest.mark.gpu


def make_fast(x, v, xb, cfg):
    n = len(x)
    f = ws.ParticleStore(n)
    f.resize(n)
    f.set_positions(x)
    if v is not None:
        f.set_velocities(v)
    b = None
    if xb is not None:
        b = ws.ParticleStore(len(xb))
        b.resize(len(xb))
        b.set_positions(xb)
    s = ws.DFSPHSolver(f, b, ws.DFSPHConfig(**dict(cfg, fast_math=True)))
    return s, f, b


FIELDS = {"rho": (O.FIELD_RHO, "dfsph_density"), "alpha": (O.FIELD_ALPHA, "dfsph_alpha"),
          "kappa": (O.FIELD_KAPPA, "dfsph_kappa"), "rho_adv": (O.FIELD_RHO_ADV, "dfsph_density_adv")}


def compare_substeps(x, v, xb, cfg, dts, tag, check_iters=True):
    s, f, _ = make_fast(x, v, xb, cfg)
    o64 = O.Solver(O.config(**cfg), x, v, xb, O.F64)
    o32 = O.Solver(O.config(**cfg), x, v, xb, O.F32)
    worst = {}
    for k, dt in enumerate(dts):
        xin, vin = f.get_positions(), f.get_velocities()
        o64.set_state(xin, vin)
        o32.set_state(xin, vin)
        s.step(dt)
        o64.step(dt)
        o32.step(dt)
        got = {"x": f.get_positions().astype(np.float64), "v": f.get_velocities().astype(np.float64)}
        for name, (fid, attr) in FIELDS.items():
            got[name] = f.get_attribute(attr).astype(np.float64)
        ref = {"x": o64.get(O.FIELD_X), "v": o64.get(O.FIELD_V)}
        b32 = {"x": o32.get(O.FIELD_X).astype(np.float64), "v": o32.get(O.FIELD_V).astype(np.float64)}
        for name, (fid, _) in FIELDS.items():
            ref[name] = o64.get(fid)
            b32[name] = o32.get(fid).astype(np.float64)
        if check_iters:
            gs, rs = s.stats(), o64.stats()
            assert gs["density_iters"] == rs["density_iters"], (tag, k, gs, rs)
        for name in ("rho", "alpha", "x"):
            rel = np.abs(got[name] - ref[name]) / np.maximum(np.abs(ref[name]), 1e-30)
            worst[name] = max(worst.get(name, 0.0), float(rel.max()))
            assert rel.max() <= 1e-5, f"{tag} substep {k} {name}: relative error {rel.max():.3e}"
        for name in ("kappa", "v", "rho_adv"):
            scale = float(np.abs(ref[name]).max())
            if name == "rho_adv":  # around rho0: the scale of its deviation
                scale = max(float(np.abs(ref[name] - cfg["rest_density"]).max()), 1e-3)
            if scale == 0.0:
                continue
            err = float(np.abs(got[name] - ref[name]).max())
            err32 = float(np.abs(b32[name] - ref[name]).max())
            worst[name] = max(worst.get(name, 0.0), err / scale)
            assert err <= 1e-4 * scale, f"{tag} substep {k} {name}: {err:.3e} > 1e-4 x {scale:.3e}"
            assert err <= 4.0 * err32 + 1e-6 * scale, \
                f"{tag} substep {k} {name}: {err:.3e} vs binary32 {err32:.3e} (scale {scale:.3e})"
    print(f"\n{tag}: worst error vs binary64 over {len(dts)} substeps: " +
          ", ".join(f"{k} {v:.2e}" for k, v in worst.items()))


@pytest.mark.parametrize("seed", [0, 1, 2])
@pytest.mark.parametrize("pform", [0, 1])
def test_fast_math_random_box(seed, pform):
    x, v, xb, cfg = random_box(n=3000, seed=seed, pressure_form=pform)
    compare_substeps(x, v, xb, cfg, [1e-3] * 4, f"random box {seed} form {pform}")


@pytest.mark.parametrize("pform", [0, 1])
def test_fast_math_block_on_the_floor(pform):
    # the block rests on the wall layers: every bottom particle has boundary
    # neighbours, whose terms the tolerance mode sums once per substep
    x, v, xb, cfg = random_box(n=3000, seed=7, pressure_form=pform)
    x[:, 1] -= np.float32(0.14)
    v[:, 1] -= np.float32(0.3)
    compare_substeps(x, v, xb, cfg, [1e-3] * 4, f"block on the floor, form {pform}")


def test_fast_math_no_boundary_and_divergence_solve():
    x, v, _, cfg = random_box(n=3000, seed=5, boundary=False, pressure_form=1,
                              enable_divergence_solve=True, max_divergence_iters=10)
    compare_substeps(x, v, None, cfg, [1e-3] * 3, "no boundary + divergence solve")


def test_fast_math_settled_puddle_case():
    x, v, xb, cfg = puddle_ref()
    compare_substeps(x, v, xb, cfg, [1e-3] * 5, "puddle")


def test_fast_math_full_size_dam_break():
    import sys
    from pathlib import Path

    sys.path.insert(0, str(Path(__file__).resolve().parents[1]))
    import bench

    x, xb, cfg = bench.make_workload(100)
    cfg = dict(cfg, pressure_form=1)
    compare_substeps(x, None, xb, cfg, [2e-4, 1.5e-4], "1M dam break")


def test_fast_math_benchmark_puddle_regression():
    """The benchmark-set puddle (kappa/rho + divergence solve) in tolerance
    mode: the same per-substep and final checks as the exact path."""
    from test_gpu_regression import assert_regression_run, benchmark_puddle

    x, xb, cfg, rule = benchmark_puddle()
    cfg.update(pressure_form=1, enable_divergence_solve=True, max_divergence_iters=100,
               eta_divergence=1e-3, fast_math=True)
    s, f, _ = make_fast(x, None, xb, cfg)
    s.set_graphs(True)
    t, recs = 0.0, []
    while t < 1.5:
        dt = rule(s.max_velocity())
        dt = min(dt, 1.5 - t)
        s.step(dt)
        t += dt
        recs.append(s.stats())
    assert_regression_run(recs, cfg["eta_density"] * cfg["rest_density"], "fast puddle")
    p = f.get_positions()
    assert np.isfinite(p).all() and p[:, 1].max() < 0.75
    print(f"\nfast-math benchmark puddle: {len(recs)} substeps, final error "
          f"{recs[-1]['mean_density_error']:.4f}, max |v| {s.max_velocity():.3f}")


def test_fast_math_refused_by_the_adjoint():
    from water_simulation_amd import adjoint as A
    from water_simulation_amd._ffi import InvalidArgument

    x, v, xb, cfg = random_box(n=512, seed=3)
    s, f, _ = make_fast(x, v, xb, cfg)
    with pytest.raises(InvalidArgument):
        A.Rollout(s, checkpoint_every=4)
