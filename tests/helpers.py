"""Shared scenario builders and comparison helpers for the parity tests."""
from __future__ import annotations

import numpy as np

import water_simulation_amd as ws
from oracle import oracle as O


def bits(a) -> np.ndarray:
    return np.ascontiguousarray(np.asarray(a, dtype=np.float32)).view(np.uint32)


def assert_bitexact(got, want, what: str) -> None:
    g = np.ascontiguousarray(got)
    w = np.ascontiguousarray(want)
    assert g.shape == w.shape, f"{what}: shape {g.shape} != {w.shape}"
    if g.dtype == np.float32 or w.dtype == np.float32:
        gb, wb = bits(g), bits(w)
        bad = np.nonzero(gb.ravel() != wb.ravel())[0]
        if bad.size:
            d = np.abs(g.astype(np.float64) - w.astype(np.float64)).ravel()
            raise AssertionError(
                f"{what}: {bad.size}/{gb.size} values differ bitwise; first at flat index "
                f"{bad[0]} got {g.ravel()[bad[0]]!r} want {w.ravel()[bad[0]]!r}; "
                f"max |diff| {d.max():.3e}")
    else:
        bad = np.nonzero(g.ravel() != w.ravel())[0]
        if bad.size:
            raise AssertionError(
                f"{what}: {bad.size}/{g.size} entries differ; first at {bad[0]} got "
                f"{g.ravel()[bad[0]]} want {w.ravel()[bad[0]]}")


def puddle_ref():
    """Settled-puddle regression case (tests/regression/test_settled_puddle.cu:119-158)."""
    r = 0.012
    x = O.lattice_block([0.30, 0.30, 0.30], [0.70, 0.70, 0.70], 2 * r)
    xb = O.sample_aabb_boundary([0, 0, 0], [1, 1, 1], 2 * r, 3)
    cfg = dict(rest_density=1000.0, particle_radius=r, smoothing_length=2 * r, viscosity=5e-2,
               surface_tension=0.0, gravity=(0, -9.81, 0), domain_min=(0, 0, 0),
               domain_max=(1, 1, 1), max_density_iters=10, eta_density=5e-3, damping=3.0)
    return x, None, xb, cfg


def random_box(n=3000, seed=0, boundary=True, **over):
    """Jittered lattice with random velocities inside a walled 0.8 m box."""
    rng = np.random.default_rng(seed)
    r = 0.01
    side = int(round(n ** (1 / 3)))
    g = np.stack(np.meshgrid(*[np.arange(side)] * 3, indexing="ij"), -1).reshape(-1, 3)
    x = (0.2 + 2 * r * g + rng.uniform(-0.3 * r, 0.3 * r, g.shape)).astype(np.float32)
    v = rng.normal(0.0, 0.5, x.shape).astype(np.float32)
    xb = O.sample_aabb_boundary([0, 0, 0], [0.8, 0.8, 0.8], 2 * r, 3) if boundary else None
    cfg = dict(rest_density=1000.0, particle_radius=r, smoothing_length=2 * r, viscosity=0.01,
               damping=0.5, domain_min=(0, 0, 0), domain_max=(0.8, 0.8, 0.8),
               max_density_iters=20, eta_density=1e-3)
    cfg.update(over)
    return x, v, xb, cfg


class Pair:
    """The HIP solver and the binary32 oracle on identical inputs."""

    def __init__(self, x, v, xb, cfg):
        self.x, self.v, self.xb, self.cfg = x, v, xb, cfg
        n = len(x)
        self.fluid = ws.ParticleStore(n)
        self.fluid.resize(n)
        self.fluid.set_positions(x)
        if v is not None:
            self.fluid.set_velocities(v)
        self.bnd = None
        if xb is not None:
            self.bnd = ws.ParticleStore(len(xb))
            self.bnd.resize(len(xb))
            self.bnd.set_positions(xb)
        self.gpu = ws.DFSPHSolver(self.fluid, self.bnd, ws.DFSPHConfig(**cfg))
        self.ref = O.Solver(O.config(**cfg), x, v, xb, O.F32)

    def step(self, dt):
        self.gpu.step(dt)
        self.ref.step(dt)

    def compare(self, tag=""):
        f = self.fluid
        assert_bitexact(f.get_positions(), self.ref.get(O.FIELD_X), f"{tag} position")
        assert_bitexact(f.get_velocities(), self.ref.get(O.FIELD_V), f"{tag} velocity")
        assert_bitexact(f.get_attribute("dfsph_density"), self.ref.get(O.FIELD_RHO), f"{tag} density")
        assert_bitexact(f.get_attribute("dfsph_alpha"), self.ref.get(O.FIELD_ALPHA), f"{tag} alpha")
        if not self.cfg.get("enable_divergence_solve"):
            assert_bitexact(f.get_attribute("dfsph_kappa"), self.ref.get(O.FIELD_KAPPA),
                            f"{tag} kappa")
            assert_bitexact(f.get_attribute("dfsph_density_adv"), self.ref.get(O.FIELD_RHO_ADV),
                            f"{tag} density_adv")
        gs, rs = self.gpu.stats(), self.ref.stats()
        assert gs["density_iters"] == rs["density_iters"], (tag, gs, rs)
        assert gs["divergence_iters"] == rs["divergence_iters"], (tag, gs, rs)
        assert gs["nan_count"] == rs["nan_count"], (tag, gs, rs)
        assert bits(gs["max_velocity"]) == bits(rs["max_velocity"]), (tag, gs, rs)
        if rs["density_iters"]:
            assert abs(gs["mean_density_error"] - rs["mean_density_error"]) <= \
                1e-12 * max(1.0, abs(rs["mean_density_error"])), (tag, gs, rs)
        assert_diagnostics(gs, rs, tag)


def assert_diagnostics(gs, rs, tag=""):
    """Regression diagnostics of a substep: the escapee count exactly, and
    the binary64 mass, momentum and kinetic energy, whose sums run in a
    different order on the device (tile partials) than in the oracle (index
    order).  Reordering a binary64 sum of n terms moves it by at most about
    n * 2^-53 of the sum of the terms' magnitudes: 1e-10 of that scale
    covers 10^6 particles (the momentum's scale is mass * max |v|, since a
    settled fluid's momentum is a cancellation)."""
    rel = 1e-10
    assert gs["escapees"] == rs["escapees"], (tag, gs, rs)
    scale_m = max(abs(rs["total_mass"]), 1e-300)
    assert abs(gs["total_mass"] - rs["total_mass"]) <= rel * scale_m, (tag, gs, rs)
    vscale = max(abs(rs["max_velocity"]), 1e-6)
    for a in range(3):
        assert abs(gs["momentum"][a] - rs["momentum"][a]) <= rel * scale_m * vscale, \
            (tag, a, gs, rs)
    assert abs(gs["kinetic_energy"] - rs["kinetic_energy"]) <= rel * rs["kinetic_energy"] + 1e-300, \
        (tag, gs, rs)


def radix_tree_top_down(sorted_codes, points):
    """Binary radix tree over sorted Morton codes built top-down from its
    definition: a node's range [lo, hi] splits after the last leaf that shares
    the range's longest common prefix in the order (code, index); Karras's
    numbering puts an internal node at the end of its range that touches its
    parent's split (the root is node 0).  Independent of the per-node search
    of spatial_accel.cu:57-100.  -> (left, right, parent, boxes (2n-1, 2, 3))."""
    n = len(sorted_codes)
    ni = n - 1
    key = [(int(c) << 32) | k for k, c in enumerate(sorted_codes)]
    left = np.zeros(max(ni, 0), np.uint32)
    right = np.zeros(max(ni, 0), np.uint32)
    parent = np.zeros(n + max(ni, 0), np.uint32)
    parent[0] = 0xFFFFFFFF
    boxes = np.zeros((n + max(ni, 0), 2, 3), np.float32)
    pts = np.asarray(points, np.float32)
    boxes[ni:, 0] = pts
    boxes[ni:, 1] = pts
    stack = [(0, 0, n - 1)] if n > 1 else []
    while stack:
        node, lo, hi = stack.pop()
        top = (key[lo] ^ key[hi]).bit_length()  # bits below the common prefix
        g = lo
        while g + 1 <= hi and (key[lo] ^ key[g + 1]).bit_length() < top:
            g += 1
        a = ni + g if g == lo else g
        b = ni + g + 1 if g + 1 == hi else g + 1
        left[node], right[node] = a, b
        parent[a] = parent[b] = node
        boxes[node, 0] = pts[lo:hi + 1].min(0)
        boxes[node, 1] = pts[lo:hi + 1].max(0)
        if a < ni:
            stack.append((a, lo, g))
        if b < ni:
            stack.append((b, g + 1, hi))
    return left, right, parent, boxes
