"""Pins the host reference gradient used for the adjoint: the forward-mode
DUAL oracle (binary32 values bit-identical to the F32 oracle, binary64
tangents) against central finite differences of the binary64 oracle with the
neighbour structure and iteration counts frozen by a recorded tape.  A wrong
tangent rule (a dropped product term, a sign, the sqrt at zero distance) fails
the FD comparison; the value check keeps the dual on the binary32 trajectory.
"""
from __future__ import annotations

import numpy as np
import pytest

from helpers import random_box
from oracle import oracle as O


def fd_directional(x, v, xb, cfg, dts, ux, uv, ue, tape, eps):
    outs = []
    for sgn in (1.0, -1.0):
        c = dict(cfg)
        s = O.Solver(O.config(**c), x, v, xb, O.F64)
        s.set_state_f64(x.astype(np.float64) + sgn * eps * ux, v.astype(np.float64) + sgn * eps * uv)
        s.set_viscosity_f64(cfg["viscosity"] + sgn * eps * ue)
        s.set_tape(tape, 2)
        for dt in dts:
            s.step(dt)
        outs.append((s.get(O.FIELD_X), s.get(O.FIELD_V)))
    return [(a - b) / (2 * eps) for a, b in zip(outs[0], outs[1])]


@pytest.mark.parametrize("iters,visc,pform", [(0, 0.0, 0), (4, 0.01, 0), (6, 0.02, 1)])
def test_dual_matches_frozen_fd(iters, visc, pform):
    x, v, xb, cfg = random_box(n=216, seed=3, max_density_iters=iters, viscosity=visc,
                               pressure_form=pform)
    dts = [0.002, 0.0015, 0.002, 0.001]
    rng = np.random.default_rng(7)
    ux = 1e-3 * rng.normal(size=x.shape)
    uv = rng.normal(size=x.shape)
    # XSPH off (viscosity <= 0, dfsph.cu:198) is a branch of its own: no viscosity derivative
    ue = 0.5 if visc > 0 else 0.0
    f32 = O.Solver(O.config(**cfg), x, v, xb, O.F32)
    tape = O.Tape()
    f32.set_tape(tape, 1)
    d = O.Solver(O.config(**cfg), x, v, xb, O.DUAL)
    d.seed(dx=ux, dv=uv, d_viscosity=ue)
    for dt in dts:
        f32.step(dt)
        d.step(dt)
    # values: the dual follows the binary32 trajectory exactly
    assert np.array_equal(d.get(O.FIELD_X), f32.get(O.FIELD_X).astype(np.float64))
    assert np.array_equal(d.get(O.FIELD_V), f32.get(O.FIELD_V).astype(np.float64))
    assert tape.builds() == len(dts)
    fx, fv = fd_directional(x, v, xb, cfg, dts, ux, uv, ue, tape, 1e-6)
    tx, tv = d.tangent(O.FIELD_X), d.tangent(O.FIELD_V)
    # the FD differentiates the binary64 map, the dual the binary32 one: they
    # differ by rounding of the trajectory (~1e-6 with near-zero kappa excesses)
    assert np.abs(fx - tx).max() <= 2e-5 * np.abs(tx).max()
    assert np.abs(fv - tv).max() <= 2e-5 * np.abs(tv).max()


def test_frozen_replay_reproduces_recording():
    # a replay of the recording run's own inputs takes the same path
    x, v, xb, cfg = random_box(n=216, seed=5, max_density_iters=8)
    a = O.Solver(O.config(**cfg), x, v, xb, O.F32)
    tape = O.Tape()
    a.set_tape(tape, 1)
    for _ in range(3):
        a.step(0.002)
    b = O.Solver(O.config(**cfg), x, v, xb, O.F32)
    b.set_tape(tape, 2)
    for _ in range(3):
        b.step(0.002)
    assert np.array_equal(a.get(O.FIELD_X), b.get(O.FIELD_X))
    assert a.stats()["density_iters"] == b.stats()["density_iters"]
