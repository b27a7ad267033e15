"""Where the adjoint's sensitivity growth comes from, measured on the oracle.

The gradient of the inverse problem's loss with respect to the initial
velocities grows by ~1.35x per 2 ms substep on the 32^3 block (4.7e30 after
200 substeps).  The forward-mode (dual-number) oracle propagates a random
tangent of the initial positions and velocities through the frozen-structure
map that the adjoint differentiates (neighbour sets, iteration counts and
branch decisions replayed from a recording run) and reports the finite-time
growth per substep, (|J_k ... J_1 u| / |u|)^(1/k), with one term switched
off at a time:

  * growth stays with XSPH off, with the clamps off, and with the density
    solve converged to 1e-6 instead of 0.1 % (so it is neither XSPH, nor the
    advection's clamps, nor the truncated Jacobi iteration);
  * it disappears without the pressure solve, and with the pressure solve on
    but the kernel values and gradients held at their values (no position
    tangent): it is the position dependence of the pressure projection --
    density, alpha and grad W as functions of the particle positions.

That is a property of the forward dynamics (a positive finite-time Lyapunov
exponent of the particle system, about 150 per second at dt = 2 ms), not of
the adjoint code: the adjoint reproduces finite differences where binary64
still tracks the binary32 trajectory.  This file pins the attribution on a
16^3 block over 100 substeps, where it takes seconds.
"""
from __future__ import annotations

import numpy as np
import pytest

from oracle import oracle as O
from water_simulation_amd import adjoint as A

DT = 0.002
H = 0.01
K = 100


def _norm(dx, dv):
    return np.sqrt(((dx / H) ** 2).sum() + ((dv * DT / H) ** 2).sum())


def growth_per_substep(over=None, freeze=False, free_advect=0, side=16, k=K, seed=0):
    p = A.inverse_problem(side=side, substeps=k)
    cfg = O.config(**dict(p.cfg, **(over or {})))
    x0, v0 = p.x0, p.v_true.astype(np.float32)
    try:
        O.set_analysis(False, free_advect)
        tape = O.Tape()
        rec = O.Solver(cfg, x0, v0, p.xb, O.DUAL)
        rec.set_tape(tape, 1)
        for _ in range(k):
            rec.step(DT)
        d = O.Solver(cfg, x0, v0, p.xb, O.DUAL)
        d.set_tape(tape, 2)
        O.set_analysis(freeze, free_advect)
        rng = np.random.default_rng(seed)
        dx = rng.normal(size=(len(x0), 3)) * H * 1e-3
        dv = rng.normal(size=(len(x0), 3)) * H / DT * 1e-3
        d.seed(dx=dx, dv=dv)
        for _ in range(k):
            d.step(DT)
        g = _norm(d.tangent(O.FIELD_X), d.tangent(O.FIELD_V)) / _norm(dx, dv)
    finally:
        O.set_analysis(False, 0)
    return g ** (1.0 / k)


@pytest.fixture(scope="module")
def full():
    return growth_per_substep()


def test_full_map_grows(full):
    assert full > 1.12


@pytest.mark.parametrize("name,kw", [
    ("no XSPH", dict(over=dict(viscosity=0.0))),
    ("no clamps", dict(free_advect=3)),
    ("converged solve", dict(over=dict(eta_density=1e-6))),
])
def test_growth_survives(name, kw, full):
    assert growth_per_substep(**kw) > 1.12, name


@pytest.mark.parametrize("name,kw", [
    ("no pressure solve", dict(over=dict(max_density_iters=0))),
    ("pressure geometry frozen", dict(freeze=True)),
])
def test_growth_vanishes(name, kw, full):
    g = growth_per_substep(**kw)
    # what remains is the linear drift x += dt v of the velocity tangent
    assert g < 1.06, name
    assert full - g > 0.08, name


def test_switches_leave_values_alone():
    # the geometry switch changes tangents only; the trajectory is the same
    p = A.inverse_problem(side=8, substeps=5)
    cfg = O.config(**p.cfg)
    out = []
    for freeze in (False, True):
        try:
            O.set_analysis(freeze, 0)
            s = O.Solver(cfg, p.x0, p.v_true.astype(np.float32), p.xb, O.DUAL)
            for _ in range(5):
                s.step(DT)
            out.append(s.get(O.FIELD_X))
        finally:
            O.set_analysis(False, 0)
    assert np.array_equal(out[0], out[1])
