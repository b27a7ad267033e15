"""Pins the CPU oracle (oracle/wsim_oracle.cpp) independently of its own code.

Each function of the oracle is checked against something that does not
restate it: closed forms, quadrature, finite differences, numpy's sort and
search, brute force, invariants of the equations, the reference's own
known-answer tests, and the reference's compiled host sources (oracle/_ref).
A plausible slip -- a dropped term, a wrong sign, a transposed operand, the
wrong kappa scaling -- fails at least one check here (see
test_oracle_mutations.py).
"""
from __future__ import annotations

import math

import numpy as np
import pytest

from oracle import oracle as O

PI_F = float(np.float32(math.pi))


def sigma(l):
    return 1.0 / (PI_F * l ** 3)


# ---------------------------------------------------------------- kernels
@pytest.mark.parametrize("l", [0.003, 0.024, 0.05, 0.1])
def test_W_closed_form_values(l):
    s = sigma(l)
    assert O.W(0.0, l, O.F64) == pytest.approx(s, rel=1e-14)
    assert O.W(l, l, O.F64) == pytest.approx(0.25 * s, rel=1e-14)           # q = 1
    assert O.W(0.5 * l, l, O.F64) == pytest.approx(s * (1 - 1.5 * 0.25 + 0.75 * 0.125), rel=1e-14)
    assert O.W(1.5 * l, l, O.F64) == pytest.approx(0.25 * s * 0.125, rel=1e-14)
    assert O.W(2.0 * l, l, O.F64) == 0.0
    assert O.W(3.0 * l, l, O.F64) == 0.0


def test_W_reference_unit_properties():
    # tests/unit/test_sph_kernels.cu:36-56
    l = 0.05
    assert O.W(0.0, l) > 0
    assert O.W(2 * l, l) == pytest.approx(0.0, abs=1e-6)
    assert O.W(3 * l, l) == 0.0
    l = 0.1
    prev = O.W(0.0, l)
    for i in range(1, 21):
        cur = O.W(i * 0.1 * l, l)
        assert cur <= prev + 1e-7
        prev = cur


def test_W_integrates_to_one():
    # int_0^{2l} 4 pi r^2 W(r) dr = 1 (normalisation of the Monaghan spline);
    # the binary32-rounded pi in sigma leaves a 3e-8 relative offset
    l = 0.02
    r = np.linspace(0.0, 2 * l, 20001)
    w = np.array([O.W(x, l, O.F64) for x in r])
    integral = np.trapezoid(4 * math.pi * r * r * w, r)
    assert integral == pytest.approx(math.pi / PI_F, rel=1e-6)


@pytest.mark.parametrize("q", [0.1, 0.5, 0.99, 1.01, 1.5, 1.9])
def test_gradW_matches_central_difference_of_W(q):
    l = 0.03
    rng = np.random.default_rng(int(q * 100))
    d = rng.normal(size=3)
    d /= np.linalg.norm(d)
    r = q * l * d
    g = O.gradW(r, l, O.F64)
    h = 1e-7 * l
    fd = np.array([(O.W(np.linalg.norm(r + h * e), l, O.F64) -
                    O.W(np.linalg.norm(r - h * e), l, O.F64)) / (2 * h) for e in np.eye(3)])
    assert np.allclose(g, fd, rtol=1e-6, atol=1e-9 * np.abs(fd).max())


def test_gradW_reference_unit_properties():
    # tests/unit/test_sph_kernels.cu:58-71: zero at r=0 and beyond 2l; points against r
    l = 0.05
    assert np.all(O.gradW([0, 0, 0], l) == 0)
    assert np.all(O.gradW([2.5 * l, 0, 0], l) == 0)
    g = O.gradW([0.5 * l, 0, 0], l)
    assert g[0] < 0 and g[1] == 0 and g[2] == 0
    # antisymmetric, bit for bit
    rng = np.random.default_rng(1)
    for _ in range(50):
        r = rng.uniform(-2 * l, 2 * l, 3).astype(np.float32)
        assert np.array_equal(O.gradW(r, l), -O.gradW(-r, l))


# ---------------------------------------------------------------- grid
def test_cell_keys_known_answers():
    origin, dims, cell = (0.0, 0.0, 0.0), (4, 5, 6), 0.5
    # cell centres: key = (cx*Dy + cy)*Dz + cz, x-major
    pts, want = [], []
    for cx in range(4):
        for cy in range(5):
            for cz in range(6):
                pts.append([(cx + 0.5) * cell, (cy + 0.5) * cell, (cz + 0.5) * cell])
                want.append((cx * 5 + cy) * 6 + cz)
    assert O.cell_keys(pts, origin, dims, cell).tolist() == want
    # outside in any coordinate, NaN, huge -> sentinel C
    C = 4 * 5 * 6
    bad = [[-0.01, 1, 1], [1, 2.5, 1], [1, 1, 3.0], [float("nan"), 1, 1], [1e30, 1, 1], [1, -1e30, 1]]
    assert O.cell_keys(bad, origin, dims, cell).tolist() == [C] * len(bad)


def test_grid_reference_unit_cases():
    # tests/unit/test_cell_grid.cu:9-74
    pts = [[a, b, c] for c in (0.25, 0.75) for b in (0.25, 0.75) for a in (0.25, 0.75)]
    keys, perm, cs = O.grid_build(pts, (0, 0, 0), (4, 4, 4), 0.5)
    assert sorted(perm.tolist()) == list(range(8))
    assert np.all(np.diff(cs.astype(np.int64)) >= 0)
    keys, perm, cs = O.grid_build([[0.25, 0.25, 0.25], [99, 99, 99]], (0, 0, 0), (2, 2, 2), 0.5)
    assert cs[-1] == 1 and cs[0] == 0


@pytest.mark.parametrize("seed", range(4))
def test_grid_permutation_and_cell_start_against_numpy(seed):
    rng = np.random.default_rng(seed)
    n = 4000
    x = rng.uniform(-0.05, 0.6, (n, 3)).astype(np.float32)
    x[:200] = x[200:400]  # equal keys with distinct ids
    origin, dims, cell = (-0.02, -0.02, -0.02), (30, 28, 31), 0.02
    keys_in = O.cell_keys(x, origin, dims, cell)
    keys, perm, cs = O.grid_build(x, origin, dims, cell)
    order = np.lexsort((np.arange(n), keys_in))           # stable (key, id) order
    assert np.array_equal(perm, order.astype(np.uint32))
    assert np.array_equal(keys, keys_in[order])
    C = 30 * 28 * 31
    assert np.array_equal(cs, np.searchsorted(keys, np.arange(C + 1), "left").astype(np.uint32))
    assert cs[C] == np.count_nonzero(keys_in < C)


def test_make_grid_matches_reference_layout():
    # dfsph.cu:55-65: dims = ceil(extent/h) + 2, origin = min - h
    o, d = O.make_grid((0, 0, 0), (2, 1, 1), 0.005)
    assert d == (402, 202, 202)
    assert o == pytest.approx((-0.005, -0.005, -0.005))
    o, d = O.make_grid((0, 0, 0), (1, 1, 1), 0.024)
    assert d == (44, 44, 44)


def test_neighbor_counts_against_brute_force():
    rng = np.random.default_rng(3)
    h = 0.02
    x = rng.uniform(0.0, 0.2, (300, 3)).astype(np.float32)
    origin, dims = O.make_grid((0, 0, 0), (0.2, 0.2, 0.2), h)
    cand, nbr = O.neighbor_counts(x, x, origin, dims, h, h)
    inv = np.float32(1.0) / np.float32(h)
    c = np.floor((x - np.array(origin, np.float32)) * inv).astype(np.int64)
    for i in range(len(x)):
        in_stencil = np.all(np.abs(c - c[i]) <= 1, axis=1)
        d = np.sqrt(((x[i] - x) ** 2).sum(1, dtype=np.float32), dtype=np.float32)
        assert cand[i] == in_stencil.sum()
        assert nbr[i] == np.count_nonzero(in_stencil & (d / np.float32(h) < 2))


# ---------------------------------------------------------------- Morton
def _interleave(q):
    code = 0
    for b in range(10):
        for a in range(3):
            code |= ((int(q[a]) >> b) & 1) << (3 * b + (2 - a))
    return code


def test_morton_known_answers():
    lo, hi = (0, 0, 0), (1, 1, 1)
    pts = [[0, 0, 0], [1, 1, 1], [2, 2, 2], [-1, -1, -1], [1 / 1024 + 1e-6, 0, 0],
           [0, 1 / 1024 + 1e-6, 0], [0, 0, 1 / 1024 + 1e-6], [2 / 1024 + 1e-6, 0, 0]]
    codes, _, _ = O.morton(pts, lo, hi)
    assert codes.tolist() == [0, 0x3FFFFFFF, 0x3FFFFFFF, 0, 4, 2, 1, 32]


def test_morton_against_bit_loop_and_stable_sort():
    rng = np.random.default_rng(5)
    x = rng.uniform(-0.1, 1.1, (2000, 3)).astype(np.float32)
    x[1000:1100] = x[0]
    lo, hi = np.array([0, 0, 0], np.float32), np.array([1, 1, 1], np.float32)
    codes, sc, perm = O.morton(x, lo, hi)
    inv = np.float32(1.0) / (hi - lo)
    q = np.minimum(np.maximum(((x - lo) * inv) * np.float32(1024), 0), 1023).astype(np.uint32)
    want = np.array([_interleave(r) for r in q], np.uint32)
    assert np.array_equal(codes, want)
    order = np.lexsort((np.arange(len(x)), want))
    assert np.array_equal(perm, order.astype(np.uint32))
    assert np.array_equal(sc, want[order])


# ---------------------------------------------------------------- host surface
@pytest.mark.parametrize("args", [((0, 0, 0), (1, 1, 1), 0.05, 1), ((0, 0, 0), (1, 1, 1), 0.05, 2),
                                  ((0, 0, 0), (1, 1, 1), 0.024, 3), ((0, 0, 0), (1, 1, 1), 0.01, 3),
                                  ((0, 0, 0), (2, 1, 1), 0.005, 3), ((0, 0, 0), (1.6, 0.9, 0.8), 0.01, 3)])
def test_sampler_bit_exact_against_compiled_reference(args):
    if O.ref() is None:
        pytest.skip("oracle/_ref not built (reference tree absent)")
    a = O.sample_aabb_boundary(*args)
    b = O.sample_aabb_boundary(*args, use_reference=True)
    assert a.shape == b.shape
    assert np.array_equal(a.view(np.uint32), b.view(np.uint32))


def test_sampler_reference_unit_cases():
    # tests/unit/test_boundary_sampler.cu:7-31
    p1 = O.sample_aabb_boundary((0, 0, 0), (1, 1, 1), 0.05, 1)
    assert 2000 < len(p1) < 4000 and len(p1) == 2646
    p2 = O.sample_aabb_boundary((0, 0, 0), (1, 1, 1), 0.05, 2)
    mi = 0.05 * 2 + 1e-3
    near = ((p2 <= mi) | (p2 >= 1 - mi)).any(axis=1)
    assert near.all()
    assert 1.5 * len(p1) < len(p2) < 2.5 * len(p1)
    # wall-particle counts of the benchmark tanks (0.024 m and 0.005 m spacing, 3 layers)
    assert len(O.sample_aabb_boundary((0, 0, 0), (1, 1, 1), 0.024, 3)) == 33282
    assert len(O.sample_aabb_boundary((0, 0, 0), (2, 1, 1), 0.005, 3)) == 1209618


def test_next_dt_bit_exact_against_compiled_reference():
    if O.ref() is None:
        pytest.skip("oracle/_ref not built")
    rng = np.random.default_rng(0)
    for _ in range(500):
        args = (1 / 24, 1 / 24, 0.4, 16, float(rng.uniform(0, 10)), float(rng.uniform(1e-3, 0.02)),
                float(rng.uniform(0, 0.05)))
        assert np.float32(O.next_dt(*args)) == np.float32(O.next_dt(*args, use_reference=True))


def test_next_dt_closed_form():
    # min{dt_max, lambda r / max(v,1e-6), remaining, max(dt_cfl, frame/max_sub)}
    assert O.next_dt(1 / 24, 1 / 24, 0.4, 16, 0.5, 0.01, 1.0) == pytest.approx(0.008, rel=1e-6)
    assert O.next_dt(1 / 24, 1 / 24, 0.4, 16, 0.0, 0.01, 1.0) == pytest.approx(1 / 24, rel=1e-6)
    assert O.next_dt(1 / 24, 1 / 24, 0.4, 16, 0.5, 0.01, 0.001) == pytest.approx(0.001, rel=1e-6)


def test_damping_decay_is_exp_rounded_to_binary32():
    # exp(x) for x = RN32(-damping * dt): against numpy's binary64 exp rounded
    # once to binary32, which is the correctly rounded value except within
    # 1e-16 relative of a rounding boundary; the oracle may differ there by
    # one unit in the last place, and nowhere else
    rng = np.random.default_rng(3)
    damping = np.concatenate([rng.uniform(0, 5, 20000), [0.0, 0.5, 3.0, 60.0, 1e4]]).astype(np.float32)
    dt = np.concatenate([rng.uniform(0, 0.05, 20000), [1e-3, 1 / 24, 4e-3, 2.0, 1.0]]).astype(np.float32)
    got = np.array([O.damping_decay(float(a), float(b)) for a, b in zip(damping, dt)], np.float32)
    x = (-damping * dt).astype(np.float32)
    want = np.exp(x.astype(np.float64)).astype(np.float32)
    ulps = np.abs(got.view(np.int32).astype(np.int64) - want.view(np.int32).astype(np.int64))
    assert ulps.max() <= 1
    assert np.count_nonzero(ulps) <= 2
    assert O.damping_decay(0.0, 0.01) == 1.0
    assert O.damping_decay(1e4, 1.0) == 0.0        # underflow
    assert O.damping_decay(-100.0, 1.0) == np.inf  # exp(100) overflows binary32
    assert O.damping_decay(-50.0, 1.0) == np.float32(np.exp(50.0))


def test_lattice_counts_match_reference():
    # README.md:75-88 (single_drop: 180 particles); settled puddle 17^3
    assert len(O.lattice_block((0.45, 0.30, 0.45), (0.55, 0.40, 0.55), 0.02)) == 180
    assert len(O.lattice_block((0.3, 0.3, 0.3), (0.7, 0.7, 0.7), 0.024)) == 4913
    p = O.lattice_block((0.1, 0.2, 0.3), (0.2, 0.3, 0.4), 0.05)
    assert p[0].tolist() == pytest.approx([0.1, 0.2, 0.3])      # x-major ordering
    assert p[1].tolist() == pytest.approx([0.1, 0.2, 0.35])


@pytest.mark.parametrize("case", ["random", "duplicates", "identical", "two"])
def test_lbvh_against_top_down_radix_tree(case):
    # the oracle's per-node Karras search (spatial_accel.cu:44-100) against the
    # tree built top-down from its definition, and the node boxes against
    # brute-force unions over each node's leaf range
    from helpers import radix_tree_top_down
    rng = np.random.default_rng(9)
    if case == "random":
        x = rng.uniform(0, 1, (700, 3)).astype(np.float32)
    elif case == "duplicates":  # many equal codes: the index breaks the ties
        x = rng.integers(0, 4, (500, 3)).astype(np.float32) / 4 + 0.01
    elif case == "identical":
        x = np.full((37, 3), 0.25, np.float32)
    else:
        x = np.array([[0.2, 0.2, 0.2], [0.8, 0.1, 0.3]], np.float32)
    codes, sc, perm = O.morton(x, (0, 0, 0), (1, 1, 1))
    l, r, p, b = O.lbvh(x, sc, perm)
    tl, tr, tp, tb = radix_tree_top_down(sc, x[perm])
    assert np.array_equal(l, tl) and np.array_equal(r, tr) and np.array_equal(p, tp)
    assert np.array_equal(b.view(np.uint32), tb.view(np.uint32))
    # every internal node and every leaf but the root has exactly one parent
    n = len(x)
    kids = np.concatenate([l, r])
    assert sorted(kids.tolist()) == list(range(1, 2 * n - 1))
