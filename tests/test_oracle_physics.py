"""Pins the oracle's DFSPH substep against closed forms and invariants.

Closed forms: isolated-particle density m W(0) (the reference's KAT,
tests/unit/test_dfsph_density.cu:16-36), interior density and alpha of an
ideal lattice, Akinci boundary mass of a 3-layer wall, free fall under
gravity with damping.  Invariants: momentum conservation of XSPH and of the
pressure solve, uniform-velocity and epsilon = 0 no-ops, the kappa/rho^2
versus kappa/rho contraction ratio, advection clamps,
divergence-free fields, and the binary64 error reduction.  The binary32
oracle's error against the binary64 oracle is stated for every field.
"""
from __future__ import annotations

import math

import numpy as np
import pytest

from oracle import oracle as O

PI_F = float(np.float32(math.pi))
S2, S3 = math.sqrt(2.0), math.sqrt(3.0)


def lattice(n, h, origin=(0.0, 0.0, 0.0), jitter=0.0, seed=0):
    """n^3 particles at cell centres of a grid with cell h (grid origin 0)."""
    i = np.arange(n, dtype=np.float64)
    g = np.stack(np.meshgrid(i, i, i, indexing="ij"), -1).reshape(-1, 3)
    x = np.asarray(origin) + (g + 0.5) * h
    if jitter:
        x += np.random.default_rng(seed).uniform(-jitter, jitter, x.shape) * h
    return x.astype(np.float32)


def cfg_box(h, extent, **kw):
    # domain [h, extent] so that the grid origin (domain_min - h) is 0
    base = dict(rest_density=1000.0, particle_radius=h / 2, smoothing_length=h, viscosity=0.0,
                damping=0.0, gravity=(0, 0, 0), domain_min=(h, h, h),
                domain_max=(extent, extent, extent), max_density_iters=0)
    base.update(kw)
    return O.config(**base)


def full_lattice_sum_W(sig):
    return sig * (1 + 6 * 0.25 + 12 * 0.25 * (2 - S2) ** 3 + 8 * 0.25 * (2 - S3) ** 3)


def test_isolated_particle_density_kat():
    # tests/unit/test_dfsph_density.cu:16-36: rho = m W(0) = 39.789 kg/m^3
    c = O.config(rest_density=1000, particle_radius=0.01, smoothing_length=0.04)
    for prec in (O.F32, O.F64):
        s = O.Solver(c, [[0.5, 0.5, 0.5]], precision=prec)
        s.step(0.001)
        rho = s.get(O.FIELD_RHO)[0]
        assert rho == pytest.approx(s.particle_mass() / (PI_F * 0.04 ** 3), rel=1e-6)
        assert rho == pytest.approx(39.789, rel=1e-4)


def test_ideal_lattice_interior_density_and_alpha():
    h = 1.0 / 64  # binary-exact lattice positions and radius
    x = lattice(9, h)
    s = O.Solver(cfg_box(h, 0.5), x, precision=O.F64)
    s.density_alpha()
    rho, alpha = s.get(O.FIELD_RHO), s.get(O.FIELD_ALPHA)
    m = 1000.0 * h ** 3  # rho0 (2r)^3 with 2r = h
    sig = 1.0 / (PI_F * h ** 3)
    centre = (4 * 9 + 4) * 9 + 4
    rho_c = m * full_lattice_sum_W(sig)
    assert rho[centre] == pytest.approx(rho_c, rel=1e-12)
    assert rho[centre] == pytest.approx(1000.0, rel=1e-3)   # h = spacing: ~rest density
    # alpha = rho / (|sum m gradW|^2 + sum |m gradW|^2); the first sum vanishes by symmetry
    g = lambda t: 0.75 * sig * t * t / h  # |grad W| at q = 2 - t
    q = m * m * (6 * g(1.0) ** 2 + 12 * g(2 - S2) ** 2 + 8 * g(2 - S3) ** 2)
    assert alpha[centre] == pytest.approx(rho_c / q, rel=1e-10)
    # binary32 error of the same fields, stated: below 1e-6 relative
    s32 = O.Solver(cfg_box(h, 0.5), x, precision=O.F32)
    s32.density_alpha()
    assert np.max(np.abs(s32.get(O.FIELD_RHO) / rho - 1)) < 1e-6
    assert np.max(np.abs(s32.get(O.FIELD_ALPHA) / alpha - 1)) < 1e-5


def test_reference_dense_pack_kat():
    # tests/unit/test_dfsph_density.cu:38-79 (h = 4r: truncated stencil, within 30 %)
    r = 0.01
    x = np.array([[0.1 + a * 2 * r, 0.1 + b * 2 * r, 0.1 + c * 2 * r]
                  for c in range(10) for b in range(10) for a in range(10)], np.float32)
    s = O.Solver(O.config(rest_density=1000, particle_radius=r, smoothing_length=4 * r), x)
    s.step(0.001)
    d = s.get(O.FIELD_RHO)[555]
    a = s.get(O.FIELD_ALPHA)[555]
    assert 700 <= d <= 1300 and 500 < d < 1500
    assert a >= 0 and math.isfinite(a)


def test_boundary_mass_of_three_layer_wall():
    h = 1.0 / 64
    i = np.arange(12, dtype=np.float64)
    g = np.stack(np.meshgrid(i, np.arange(3.0), i, indexing="ij"), -1).reshape(-1, 3)
    xb = ((g + 0.5) * h + np.array([0.0, 5 * h, 0.0])).astype(np.float32)  # layers at y cells 5..7
    s = O.Solver(cfg_box(h, 0.5), [[0.4, 0.4, 0.4]], xb=xb, precision=O.F64)
    mb = s.get(O.FIELD_BMASS)
    sig = 1.0 / (PI_F * h ** 3)
    idx = lambda a, layer, c: (a * 3 + layer) * 12 + c
    mid = idx(6, 1, 6)
    assert mb[mid] == pytest.approx(1000.0 / full_lattice_sum_W(sig), rel=1e-12)
    outer = idx(6, 0, 6)
    sw = sig * (1 + 5 * 0.25 + 8 * 0.25 * (2 - S2) ** 3 + 4 * 0.25 * (2 - S3) ** 3)
    assert mb[outer] == pytest.approx(1000.0 / sw, rel=1e-12)


def test_free_fall_with_damping_closed_form():
    # v_{k+1} = v_k e^{-d dt} + g dt ; x_{k+1} = x_k + v_{k+1} dt
    c = O.config(rest_density=1000, particle_radius=0.01, smoothing_length=0.04, viscosity=0.0,
                 damping=0.7, domain_min=(0, -10, 0), domain_max=(1, 10, 1))
    s = O.Solver(c, [[0.5, 0.5, 0.5]], precision=O.F64)
    dt, g = 0.01, float(np.float32(-9.81))
    dec = math.exp(-float(np.float32(0.7)) * float(np.float32(dt)))
    dtf = float(np.float32(dt))
    v, y = 0.0, float(np.float32(0.5))
    for _ in range(40):
        s.step(dt)
        v = v * dec + g * dtf
        y = y + v * dtf
    assert s.get(O.FIELD_V)[0, 1] == pytest.approx(v, rel=1e-12)
    assert s.get(O.FIELD_X)[0, 1] == pytest.approx(y, rel=1e-12)
    assert s.get(O.FIELD_X)[0, 0] == pytest.approx(0.5, abs=1e-15)


def test_xsph_conserves_momentum_and_keeps_uniform_fields():
    h = 0.02
    x = lattice(8, h, origin=(0.1, 0.1, 0.1), jitter=0.2, seed=1)
    rng = np.random.default_rng(2)
    v = rng.normal(0, 0.3, x.shape).astype(np.float32)
    s = O.Solver(cfg_box(h, 0.5, viscosity=0.05), x, v, precision=O.F64)
    s.step(1e-5)
    dv = s.get(O.FIELD_V) - v.astype(np.float64)
    assert np.abs(dv).max() > 1e-4                                   # XSPH acted
    assert np.abs(dv.sum(0)).max() < 1e-12 * np.abs(dv).sum()        # momentum conserved
    # a uniform field is a fixed point; epsilon = 0 is a no-op
    u = np.tile(np.float32([0.3, -0.2, 0.1]), (len(x), 1))
    s = O.Solver(cfg_box(h, 0.5, viscosity=0.05), x, u, precision=O.F64)
    s.step(1e-5)
    assert np.allclose(s.get(O.FIELD_V), u.astype(np.float64), rtol=0, atol=1e-15)
    s = O.Solver(cfg_box(h, 0.5, viscosity=0.0), x, v, precision=O.F64)
    s.step(1e-5)
    assert np.array_equal(s.get(O.FIELD_V), v.astype(np.float64))


def test_density_change_invariants():
    h = 0.02
    x = lattice(7, h, origin=(0.1, 0.1, 0.1), jitter=0.1, seed=3)
    u = np.tile(np.float32([0.2, -0.4, 0.1]), (len(x), 1))
    s = O.Solver(cfg_box(h, 0.5, max_density_iters=1), x, u, precision=O.F64)
    s.step(1e-4)
    # uniform velocity, no boundary: rho_dot = 0, so rho* = rho
    assert np.array_equal(s.get(O.FIELD_RHO_ADV), s.get(O.FIELD_RHO))
    # a particle moving toward a wall compresses (boundary term sign)
    xb = np.array([[a * h + h / 2, 0.1 + h / 2, c * h + h / 2] for a in range(8) for c in range(8)],
                  np.float32)
    xs = np.array([[4 * h + h / 2, 0.1 + 1.5 * h, 4 * h + h / 2]], np.float32)
    for vy, sign in ((-1.0, 1), (1.0, -1)):
        s = O.Solver(cfg_box(h, 0.5, max_density_iters=1), xs, [[0, vy, 0]], xb=xb,
                     precision=O.F64)
        s.step(1e-4)
        drho = s.get(O.FIELD_RHO_ADV)[0] - s.get(O.FIELD_RHO)[0]
        assert np.sign(drho) == sign


def test_pressure_solve_conserves_momentum():
    h = 0.02
    x = lattice(8, 0.9 * h, origin=(0.1, 0.1, 0.1), jitter=0.1, seed=4)  # compressed block
    s = O.Solver(cfg_box(h, 0.5, max_density_iters=3), x, precision=O.F64)
    s.step(1e-4)
    v = s.get(O.FIELD_V)
    assert np.abs(v).max() > 1e-6
    assert np.abs(v.sum(0)).max() < 1e-11 * np.abs(v).sum()


def test_kappa_scaling_contraction_ratio():
    # With alpha = rho/D, one sweep of the reference kappa/rho^2 form
    # removes 1/rho of what a kappa/rho sweep removes.  Cell-aligned lattice
    # (identical interior rho), compressed through the mass (r > h/2), with a
    # converging velocity field so that kappa varies between particles.
    h = 1.0 / 64
    x = lattice(12, h, origin=(0.1, 0.1, 0.1))
    centre = (6 * 12 + 6) * 12 + 6
    dx = (x - x[centre]).astype(np.float64)
    v = (-dx * (dx * dx).sum(1, keepdims=True) * 3e3).astype(np.float32)  # cubic: varying div v
    kw = dict(particle_radius=0.52 * h, eta_density=1e-12)

    def rho_adv(iters, form):
        s = O.Solver(cfg_box(h, 0.5, max_density_iters=iters, pressure_form=form, **kw), x, v,
                     precision=O.F64)
        s.step(1e-4)
        return s.get(O.FIELD_RHO_ADV), s.get(O.FIELD_RHO)

    first, rho = rho_adv(1, 0)
    d0 = rho_adv(2, 0)[0] - first
    d1 = rho_adv(2, 1)[0] - first
    ring = [centre + 1, centre - 12, centre + 144]  # interior particles with kappa gradients
    for i in [centre] + ring:
        assert d1[i] != 0.0
        assert d1[i] / d0[i] == pytest.approx(rho[i], rel=1e-9)


def test_advect_clamps_closed_form():
    c = O.config(rest_density=1000, particle_radius=0.01, smoothing_length=0.02, viscosity=0.0,
                 damping=0.0, gravity=(0, 0, 0), max_density_iters=0)
    x = np.float32([[0.5, 0.5, 0.5], [0.0005, 0.5, 0.5], [0.5, 0.999, 0.5], [0.2, 0.2, 0.2]])
    v = np.float32([[10, 0, 0], [-1, 0, 0], [0, 3, 0], [3, 4, 0]])
    s = O.Solver(c, x, v, precision=O.F64)
    s.step(0.01)
    X, V = s.get(O.FIELD_X), s.get(O.FIELD_V)
    assert V[0].tolist() == pytest.approx([5, 0, 0])                     # |v| clamp at 5 m/s
    assert X[0, 0] == pytest.approx(0.55)
    assert X[1, 0] == pytest.approx(float(np.float32(1e-3)))              # min + eps, v -> 0
    assert V[1, 0] == 0
    assert X[2, 1] == pytest.approx(1 - float(np.float32(1e-3)))          # max - eps, v -> 0
    assert V[2, 1] == 0
    assert X[3].tolist() == pytest.approx([0.23, 0.24, 0.2])              # |v| = 5, untouched
    st = s.stats()
    assert st["max_velocity"] == pytest.approx(5.0)


def test_density_error_and_max_velocity_reductions():
    h = 0.02
    x = lattice(8, 0.93 * h, origin=(0.1, 0.1, 0.1), jitter=0.2, seed=6)
    rng = np.random.default_rng(7)
    v = rng.normal(0, 0.5, x.shape).astype(np.float32)
    s = O.Solver(cfg_box(h, 0.5, max_density_iters=3, eta_density=1e-9), x, v, precision=O.F32)
    s.step(1e-3)
    st = s.stats()
    ra = s.get(O.FIELD_RHO_ADV)
    want = np.maximum(np.float32(0), ra - np.float32(1000)).astype(np.float64).mean()
    assert st["mean_density_error"] == pytest.approx(want, rel=1e-12)
    V = s.get(O.FIELD_V).astype(np.float64)
    assert st["max_velocity"] == pytest.approx(np.sqrt((V * V).sum(1)).max(), rel=1e-6)
    assert st["density_iters"] == 3


def test_density_solve_early_exit_needs_two_iterations():
    # dfsph.cu:254-258: the error is tested from the second iteration on
    h = 0.02
    x = lattice(6, 1.2 * h, origin=(0.1, 0.1, 0.1))  # expanded: rho < rho0, error 0
    s = O.Solver(cfg_box(h, 0.5, max_density_iters=50), x)
    s.step(1e-3)
    assert s.stats()["density_iters"] == 2
    assert s.stats()["mean_density_error"] == 0.0


def test_divergence_solve_invariants():
    h = 0.02
    x = lattice(8, h, origin=(0.1, 0.1, 0.1), jitter=0.05, seed=8)
    u = np.tile(np.float32([0.1, 0.0, -0.1]), (len(x), 1))
    kw = dict(enable_divergence_solve=1, max_divergence_iters=20, eta_divergence=1e-9)
    s = O.Solver(cfg_box(h, 0.5, **kw), x, u, precision=O.F64)
    s.step(1e-4)
    assert np.allclose(s.get(O.FIELD_V), u.astype(np.float64), atol=1e-14)  # uniform: no-op
    # converging field: the solve lowers the mean positive divergence
    centre = x.mean(0)
    conv = (-(x - centre) * 2.0).astype(np.float32)
    errs = []
    for iters in (1, 20):
        kw["max_divergence_iters"] = iters
        s = O.Solver(cfg_box(h, 0.5, **kw), x, conv, precision=O.F64)
        s.step(1e-4)
        errs.append(s.stats()["mean_divergence_error"])
    assert errs[1] < errs[0]


def test_binary32_error_against_binary64_per_field():
    """States the binary32 oracle's error against the binary64 oracle on one
    substep of a compressed, moving block (the GPU path equals the binary32
    oracle bit for bit, so these are the GPU's errors too)."""
    h = 0.02
    x = lattice(10, 0.97 * h, origin=(0.1, 0.1, 0.1), jitter=0.2, seed=9)
    v = np.random.default_rng(10).normal(0, 0.3, x.shape).astype(np.float32)
    kw = dict(viscosity=0.02, damping=0.5, gravity=(0, -9.81, 0), max_density_iters=5,
              eta_density=1e-9)
    a = O.Solver(cfg_box(h, 0.5, **kw), x, v, precision=O.F32)
    b = O.Solver(cfg_box(h, 0.5, **kw), x, v, precision=O.F64)
    a.step(5e-4)
    b.step(5e-4)
    rel = lambda f: np.max(np.abs(a.get(f).astype(np.float64) - b.get(f)) /
                           np.maximum(np.abs(b.get(f)), 1e-30))
    assert rel(O.FIELD_RHO) < 1e-6
    assert rel(O.FIELD_ALPHA) < 1e-5
    assert rel(O.FIELD_X) < 1e-6
    # kappa and v carry cancellation: bound them against the field scale
    for f in (O.FIELD_KAPPA, O.FIELD_V, O.FIELD_RHO_ADV):
        d = np.abs(a.get(f).astype(np.float64) - b.get(f)).max()
        assert d <= 1e-4 * np.abs(b.get(f)).max()


def test_regression_diagnostics_closed_forms():
    """Escapees, mass, momentum and kinetic energy after a substep: isolated
    particles (no neighbours, no gravity, no damping) keep their velocities,
    so the binary64 sums have closed forms; a particle that reaches the
    domain clamp counts as an escapee (and its outward velocity is zeroed,
    dfsph_kernels.cu:376-385); a NaN particle is neither summed nor an
    escapee.  A wrong factor 1/2, a dropped axis or a sum over the NaN
    particle fails it."""
    c = O.config(rest_density=1000, particle_radius=0.01, smoothing_length=0.02, viscosity=0.0,
                 damping=0.0, gravity=(0, 0, 0), max_density_iters=0)
    x = np.array([[0.2, 0.2, 0.2], [0.5, 0.5, 0.5], [0.8, 0.2, 0.5], [0.0015, 0.5, 0.5],
                  [0.3, 0.7, 0.3]], np.float32)
    v = np.array([[1.0, 0.0, 0.0], [0.0, -2.0, 0.5], [0.25, 0.25, -1.0], [-3.0, 0.0, 0.0],
                  [np.nan, 0.0, 0.0]], np.float32)
    s = O.Solver(c, x, v, None, O.F64)
    s.step(0.001)
    st = s.stats()
    m = 1000.0 * (2.0 * float(np.float32(0.01))) ** 3  # the binary64 instance's mass
    assert m == pytest.approx(s.particle_mass(), rel=1e-7)
    vs = v[:3].astype(np.float64)  # particle 3 is clamped at x = 0.001: v.x -> 0
    assert st["escapees"] == 1
    assert st["nan_count"] == 1
    assert st["total_mass"] == pytest.approx(4 * m, rel=1e-15)
    mom = m * vs.sum(0)
    assert st["momentum"] == pytest.approx(list(mom), rel=1e-14)
    assert st["kinetic_energy"] == pytest.approx(0.5 * m * (vs ** 2).sum(), rel=1e-14)
    out = s.get(O.FIELD_X)
    assert out[3, 0] == float(np.float32(1e-3))  # domain_min + 1e-3f


def dwdq(q, sig):
    """dW/dq of the cubic spline (sph_kernels.cuh:25-41), binary64."""
    return sig * (-3.0 * q + 2.25 * q * q) if q < 1.0 else -0.75 * sig * (2.0 - q) ** 2


def test_density_change_closed_form_for_a_linear_velocity_field():
    """rho-dot (dfsph_kernels.cu:196-224) of an interior particle of an ideal
    lattice (spacing = cell = l) under v(x) = A x + u: with grad W = r c(|r|),
    sum_j m (A r_j) . r_j c_j = m tr(A) S, S = l^2 (2 c(l) + 8 c(sqrt2 l) +
    8 c(sqrt3 l)) by the lattice's cubic symmetry, whatever the constant u.
    One density iteration (max_density_iters = 1) leaves rho* = rho +
    dt rho-dot in the predicted density.  A flipped (v_j - v_i), a dropped
    mass, W in place of grad W or a missing neighbour shell fails it."""
    h = 1.0 / 64
    x = lattice(9, h)
    A = np.array([[0.3, 0.1, -0.2], [0.05, -0.7, 0.4], [0.2, 0.15, 0.25]])
    u = np.array([0.4, -1.1, 0.25])
    v = x.astype(np.float64) @ A.T + u
    dt = 1e-3
    s = O.Solver(cfg_box(h, 0.5, max_density_iters=1), x, precision=O.F64)
    s.set_state_f64(x.astype(np.float64), v)
    s.step(dt)
    centre = (4 * 9 + 4) * 9 + 4
    rho, rho_adv = s.get(O.FIELD_RHO)[centre], s.get(O.FIELD_RHO_ADV)[centre]
    m = 1000.0 * h ** 3
    sig = 1.0 / (PI_F * h ** 3)
    c = lambda q: dwdq(q, sig) / (h * h * q)  # dW/dq / (l d), d = q l
    S = h * h * (2 * c(1.0) + 8 * c(S2) + 8 * c(S3))
    rho_dot = m * np.trace(A) * S
    assert (rho_adv - rho) / float(np.float32(dt)) == pytest.approx(rho_dot, rel=1e-9)
    assert abs(rho_dot) > 1.0  # a non-trivial rate


@pytest.mark.parametrize("pform", [0, 1])
def test_boundary_pressure_mirror_term_closed_form(pform):
    """Akinci boundary terms of density, alpha, kappa and the pressure
    velocity update (dfsph_kernels.cu:95-107, 145-163, 226-231, 281-296) for a
    lone fluid particle at height l above a one-layer wall lattice (spacing =
    cell = l), directly over a wall particle: its nine wall neighbours lie at
    l (1), sqrt2 l (4) and sqrt3 l (4), all at height l.  Boundary masses are
    rho0 / (sigma (2 + (2 - sqrt2)^3)) (interior of the layer); with v = 0 and
    g = 0 the first sweep applies
        dv = -dt m_b (k_i / rho^x + k_i / rho0^x) sum_b grad W_ib,   x = 2 (form 0), 1 (form 1),
    the second term being the mirror pressure of the wall.  A dropped mirror
    term, a wrong power of rho or rho0, or a sign fails it."""
    h = 1.0 / 64
    i = np.arange(11, dtype=np.float64)
    g = np.stack(np.meshgrid(i, [0.0], i, indexing="ij"), -1).reshape(-1, 3)
    xb = ((g + 0.5) * h + np.array([5 * h, 5 * h, 5 * h])).astype(np.float32)  # plane at cell y = 5
    xp = np.array([[10.5 * h, 6.5 * h, 10.5 * h]], np.float32)                 # over wall point (5, 5)
    dt = 1e-3
    # spacing 2r = 1.4 l: the particle's own mass puts rho above rho0
    cfg = cfg_box(h, 0.5, particle_radius=0.7 * h, max_density_iters=1, pressure_form=pform)
    s = O.Solver(cfg, xp, xb=xb, precision=O.F64)
    s.step(dt)
    sig = 1.0 / (PI_F * h ** 3)
    rho0 = 1000.0
    r32 = float(np.float32(0.7 * h))
    m = rho0 * (2 * r32) ** 3
    Wq = lambda q: sig * (1 - 1.5 * q * q + 0.75 * q ** 3) if q < 1 else 0.25 * sig * (2 - q) ** 3
    mb = rho0 / (Wq(0) + 4 * Wq(1.0) + 4 * Wq(S2))
    assert s.get(O.FIELD_BMASS)[5 * 11 + 5] == pytest.approx(mb, rel=1e-12)
    rho = m * Wq(0) + mb * (Wq(1.0) + 4 * Wq(S2) + 4 * Wq(S3))
    c = lambda q: dwdq(q, sig) / (h * h * q)
    G = mb * h * (c(1.0) + 4 * c(S2) + 4 * c(S3))  # y component; x and z cancel
    Q = mb * mb * (c(1.0) ** 2 * h * h + 4 * c(S2) ** 2 * 2 * h * h + 4 * c(S3) ** 2 * 3 * h * h)
    alpha = rho / (G * G + Q)
    dtf = float(np.float32(dt))  # the solver's time step is binary32
    kappa = max(0.0, rho - rho0) * alpha / (dtf * dtf)
    p = 2 if pform == 0 else 1
    dv = -dtf * (kappa / rho ** p + kappa / rho0 ** p) * G
    assert rho > rho0 * 1.05
    assert s.get(O.FIELD_RHO)[0] == pytest.approx(rho, rel=1e-12)
    assert s.get(O.FIELD_ALPHA)[0] == pytest.approx(alpha, rel=1e-10)
    assert s.get(O.FIELD_KAPPA)[0] == pytest.approx(kappa, rel=1e-10)
    vout = s.get(O.FIELD_V)[0]
    assert vout[1] == pytest.approx(dv, rel=1e-10)
    assert dv > 0  # the wall pushes the compressed particle up
    assert abs(vout[0]) < 1e-12 * abs(dv) and abs(vout[2]) < 1e-12 * abs(dv)
