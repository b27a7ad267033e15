"""GPU parity of the DFSPH substep: every field of every substep bit-exact
against the binary32 oracle, plus the reference's own solver unit tests
(tests/unit/test_dfsph_density.cu) and particle-store tests
(tests/unit/test_particle_store.cu) through the C ABI.
"""
from __future__ import annotations

import math

import numpy as np
import pytest

import water_simulation_amd as ws
from helpers import Pair, assert_bitexact, puddle_ref, random_box
from oracle import oracle as O

pytestmark = pytest.mark.gpu


# ---- reference unit tests (tests/unit/test_dfsph_density.cu) ----

def _single(cfg_over, pos=(0.5, 0.5, 0.5)):
    store = ws.ParticleStore(1)
    store.resize(1)
    store.set_positions(np.array([pos], np.float32))
    cfg = ws.DFSPHConfig(rest_density=1000.0, particle_radius=0.01, smoothing_length=0.04,
                         domain_min=(0, 0, 0), domain_max=(1, 1, 1), **cfg_over)
    return store, ws.DFSPHSolver(store, None, cfg)


def test_isolated_particle_density_is_mass_times_W0():
    # :16-36
    store, solver = _single({})
    solver.step(0.001)
    expected = solver.particle_mass() / (np.float32(math.pi) * 0.04 ** 3)
    assert solver.density_at(0) == pytest.approx(expected, rel=1e-4)
    assert solver.density_at(0) == pytest.approx(39.789, rel=1e-4)


def test_dense_pack_center_density_approaches_rest_density():
    # :38-79
    r = 0.01
    pts = np.array([[0.1 + x * 2 * r, 0.1 + y * 2 * r, 0.1 + z * 2 * r]
                    for z in range(10) for y in range(10) for x in range(10)], np.float32)
    store = ws.ParticleStore(len(pts))
    store.resize(len(pts))
    store.set_positions(pts)
    cfg = ws.DFSPHConfig(rest_density=1000.0, particle_radius=r, smoothing_length=4 * r)
    solver = ws.DFSPHSolver(store, None, cfg)
    solver.step(0.001)
    d = solver.density_at(555)
    assert d == pytest.approx(1000.0, rel=0.30)
    assert 500.0 < d < 1500.0
    a = solver.alpha_at(555)
    assert a >= 0.0 and math.isfinite(a)


def test_single_particle_falls_without_lateral_drift():
    # :81-109
    store, solver = _single({"viscosity": 0.0})
    store.set_velocities(np.zeros((1, 3), np.float32))
    for _ in range(50):
        solver.step(0.01)
    p = store.get_positions()[0]
    assert p[1] < 0.5
    assert p[0] == pytest.approx(0.5, rel=0.01)
    assert p[2] == pytest.approx(0.5, rel=0.01)


def test_density_at_out_of_range_raises():
    store, solver = _single({})
    solver.step(0.001)
    with pytest.raises(IndexError):
        solver.density_at(1)


# ---- particle store (tests/unit/test_particle_store.cu) ----

def test_particle_store_contract():
    s = ws.ParticleStore(100)
    assert s.capacity() == 100 and s.count() == 0
    assert s.has_attribute("position") and s.has_attribute("velocity")
    assert not s.has_attribute("density")
    s.resize(10)
    assert s.count() == 10
    s.resize(100)
    with pytest.raises(IndexError):
        s.resize(101)
    h = s.register_attribute("density", ws.AttribType.F32)
    assert s.has_attribute("density")
    with pytest.raises(RuntimeError):
        s.register_attribute("density", ws.AttribType.F32)
    with pytest.raises(IndexError):
        s.attribute_data(9999)
    vals = np.arange(100, dtype=np.float32)
    ws.memcpy_h2d(s.attribute_data(h), vals)
    assert np.array_equal(s.get_attribute(h), vals)
    assert np.all(s.get_attribute("velocity", width=3) == 0.0)  # zero-filled


# ---- bitwise parity with the oracle, substep by substep ----

def _run(pair: Pair, steps: int, dt_rule):
    pair.compare("initial")  # boundary masses etc. are set at construction
    for k in range(steps):
        dt = dt_rule(pair)
        pair.step(dt)
        pair.compare(f"step {k}")


def _puddle_dt(pair):
    # tests/regression/test_settled_puddle.cu:168-172
    v = max(pair.gpu.max_velocity(), 0.5)
    return max(min(0.001, 0.3 * 0.012 / v), 1e-5)


def test_parity_settled_puddle_case():
    x, v, xb, cfg = puddle_ref()
    pair = Pair(x, v, xb, cfg)
    assert_bitexact(pair.bnd.get_attribute("dfsph_boundary_mass"), pair.ref.get(O.FIELD_BMASS),
                    "boundary mass")
    for k in range(6):
        pair.step(_puddle_dt(pair))
        pair.compare(f"puddle step {k}")


@pytest.mark.parametrize("over", [
    {},
    {"pressure_form": 1},
    {"enable_divergence_solve": True, "max_divergence_iters": 8},
    {"viscosity": 0.0, "max_density_iters": 1},
    {"max_density_iters": 0},
])
def test_parity_random_box(over):
    x, v, xb, cfg = random_box(n=3375, seed=11, **over)
    pair = Pair(x, v, xb, cfg)
    for k in range(4):
        pair.step(0.002)
        pair.compare(f"box {over} step {k}")


def test_parity_without_boundary():
    x, v, _, cfg = random_box(n=1000, seed=5, boundary=False)
    pair = Pair(x, v, None, cfg)
    for k in range(4):
        pair.step(0.002)
        pair.compare(f"noboundary step {k}")


def test_parity_escaping_particles_and_clamps():
    # particles launched at the walls at > 5 m/s: velocity clamp + domain clamp
    x, v, xb, cfg = random_box(n=512, seed=3)
    v = (v * 20.0).astype(np.float32)
    pair = Pair(x, v, xb, cfg)
    for k in range(5):
        pair.step(0.004)
        pair.compare(f"clamp step {k}")


def test_zero_particles_step_is_a_no_op():
    store = ws.ParticleStore(16)
    solver = ws.DFSPHSolver(store, None, ws.DFSPHConfig())
    solver.step(0.001)  # dfsph.cu:121-122
    assert solver.max_velocity() == 0.0


def test_fluid_grid_view_matches_oracle_grid():
    x, v, xb, cfg = random_box(n=1000, seed=9)
    pair = Pair(x, v, xb, cfg)
    pair.step(0.002)
    g = pair.gpu.fluid_grid()
    origin, dims = O.make_grid(cfg["domain_min"], cfg["domain_max"], cfg["smoothing_length"])
    keys, perm, cs = O.grid_build(x, origin, dims, cfg["smoothing_length"])
    assert_bitexact(g.read_sorted_indices(), perm, "solver fluid grid perm")
    assert_bitexact(g.read_cell_start(), cs, "solver fluid grid cell_start")
    bg = pair.gpu.boundary_grid()
    bkeys, bperm, bcs = O.grid_build(xb, origin, dims, cfg["smoothing_length"])
    assert_bitexact(bg.read_sorted_indices(), bperm, "boundary grid perm")
    assert_bitexact(bg.read_cell_start(), bcs, "boundary grid cell_start")


@pytest.mark.parametrize("seed,scale", [(11, 1.0), (12, 20.0)])
def test_fluid_grid_partial_table_rebuild(seed, scale):
    # the solver's grid rewrites only the cells between the previous and the
    # current key span; over substeps whose spans grow, shrink and (scale 20,
    # particles thrown at the walls) reach the domain clamp, every table
    # equals the oracle's full table of the positions it was built from
    x, v, xb, cfg = random_box(n=1500, seed=seed)
    v = (v * scale).astype(np.float32)
    pair = Pair(x, v, xb, cfg)
    origin, dims = O.make_grid(cfg["domain_min"], cfg["domain_max"], cfg["smoothing_length"])
    g = pair.gpu.fluid_grid()
    for k in range(12):
        x0 = pair.fluid.get_positions()
        pair.step(0.004)
        keys, perm, cs = O.grid_build(x0, origin, dims, cfg["smoothing_length"])
        assert_bitexact(g.read_sorted_indices(), perm, f"substep {k} perm")
        assert_bitexact(g.read_cell_start(), cs, f"substep {k} cell_start")

