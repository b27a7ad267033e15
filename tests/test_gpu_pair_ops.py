"""The pair-path square root and quotients (hardware approximations with
their fma correction steps, without the compiler's range scaling) against
the IEEE operations, for every binary32 operand the solver kernels can
produce at a given smoothing length: every count of differing results must
be 0 wherever the solver selects the pair path.  The smoothing lengths are
those of the tests and benchmarks (2.5 mm and 5 mm particles, the random
boxes, the settled puddle) plus 1 m and an out-of-range one.
"""
from __future__ import annotations

import pytest

import water_simulation_amd as ws

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("l", [0.005, 0.01, 0.02, 0.024, 1.0])
def test_pair_ops_exact(l):
    r = ws.pair_ops_check(l)
    assert r["fast"], r
    assert r["sqrt"] == 0 and r["quotient"] == 0 and r["gradient"] == 0, r


def test_pair_ops_off_outside_range():
    # kernel-gradient quotients up to 1e12 / (pi l^4) leave the range where
    # the division needs no scaling: the solver keeps the IEEE operations
    r = ws.pair_ops_check(1e-6)
    assert not r["fast"], r


def test_pair_ops_rejects_bad_length():
    with pytest.raises(ws.InvalidArgument):
        ws.pair_ops_check(0.0)
