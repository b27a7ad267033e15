"""Batched uneven scenes: one BatchSolver over many scenes against one plain
DFSPHSolver per scene (bit for bit, every field and statistic), against the
binary32 oracle, with and without graph replay, for every launch shape, with
the divergence solve, and the argument checks of the batch entry points.
"""
from __future__ import annotations

import ctypes as C

import numpy as np
import pytest

import water_simulation_amd as ws
from water_simulation_amd import scenes as SC
from helpers import assert_bitexact, bits
from oracle import oracle as O

pytestmark = pytest.mark.gpu

FIELDS = ("dfsph_density", "dfsph_alpha", "dfsph_kappa", "dfsph_density_adv")


def wall_positions():
    return ws.sample_aabb_boundary((0, 0, 0), (SC.BOX,) * 3, SC.WALL_SPACING, SC.WALL_LAYERS)


def boundary_store(xb):
    """A fresh wall store: each solver registers its boundary masses on its own store
    (a second registration is the reference's duplicate-attribute error)."""
    b = ws.ParticleStore(len(xb))
    b.resize(len(xb))
    b.set_positions(xb)
    return b


def fluid_store(x, v):
    n = len(x)
    f = ws.ParticleStore(max(n, 1))
    f.resize(n)
    if n:
        f.set_positions(x)
        f.set_velocities(v)
    return f


def scene_dt(vmax, ts=ws.TimeStepper()):
    return ts.next_dt(max(vmax, 0.5), SC.RADIUS, 1.0)


class Batch:
    def __init__(self, mix, xb, cfg, graphs=True):
        self.mix = mix
        self.counts = [len(s.positions) for s in mix]
        x = np.concatenate([s.positions for s in mix]).astype(np.float32)
        v = np.concatenate([s.velocities for s in mix]).astype(np.float32)
        self.fluid = fluid_store(x, v)
        self.bnd = boundary_store(xb)
        self.solver = ws.BatchSolver(self.fluid, self.bnd, ws.DFSPHConfig(**cfg), self.counts,
                                     [s.viscosity for s in mix])
        self.solver.set_graphs(graphs)
        self.vmax = [SC.initial_max_speed(s.velocities) for s in mix]
        self.dts = []

    def step(self):
        dt = np.array([scene_dt(v) for v in self.vmax], np.float32)
        self.dts.append(dt)
        self.solver.step(dt)
        self.vmax = [st["max_velocity"] for st in self.solver.all_stats()]

    def scene(self, s):
        a, b = self.solver.offsets[s], self.solver.offsets[s + 1]
        f = self.fluid
        out = {"x": f.get_positions()[a:b], "v": f.get_velocities()[a:b]}
        for name in FIELDS:
            out[name] = f.get_attribute(name)[a:b]
        return out


def run_plain(scene, xb, cfg, dts, graphs=False, shapes=None):
    f = fluid_store(scene.positions, scene.velocities)
    bnd = boundary_store(xb)
    c = dict(cfg, viscosity=scene.viscosity)
    s = ws.DFSPHSolver(f, bnd, ws.DFSPHConfig(**c))
    s.set_graphs(graphs)
    if shapes:
        s.set_shapes(**shapes)
    stats = []
    for dt in dts:
        s.step(float(dt))
        stats.append(s.stats())
    out = {"x": f.get_positions(), "v": f.get_velocities()}
    for name in FIELDS:
        out[name] = f.get_attribute(name)
    return out, stats


def compare_scene(got, want, tag):
    for k in want:
        assert_bitexact(got[k], want[k], f"{tag} {k}")


def compare_stats(gs, ws_, tag):
    for key in ("density_iters", "divergence_iters", "nan_count"):
        assert gs[key] == ws_[key], (tag, key, gs, ws_)
    assert bits(gs["max_velocity"]) == bits(ws_["max_velocity"]), (tag, gs, ws_)
    for key in ("mean_density_error", "mean_divergence_error"):
        assert np.float64(gs[key]).view(np.uint64) == np.float64(ws_[key]).view(np.uint64), \
            (tag, key, gs, ws_)


def small_mix():
    """Uneven sizes incl. an empty scene, a lone particle and a partial tile."""
    mix = [SC.box_scene(i, n_min=40, n_max=3000) for i in range(5)]
    lone = SC.BoxScene(np.array([[0.3, 0.3, 0.3]], np.float32),
                       np.array([[0.1, -0.2, 0.05]], np.float32), 0.02, (1, 1, 1))
    empty = SC.BoxScene(np.zeros((0, 3), np.float32), np.zeros((0, 3), np.float32), 0.01,
                        (0, 0, 0))
    no_visc = SC.box_scene(7, n_min=200, n_max=400)
    no_visc.viscosity = 0.0
    return [mix[0], empty, mix[1], lone, mix[2], no_visc, mix[3], mix[4]]


@pytest.fixture(scope="module")
def walls():
    return wall_positions()


def test_batch_matches_plain_per_scene(walls):
    xb = walls
    cfg = SC.box_config(max_density_iters=30)
    mix = small_mix()
    b = Batch(mix, xb, cfg)
    for _ in range(6):
        b.step()
    dts = np.array(b.dts)
    for s, scene in enumerate(mix):
        st = b.solver.scene_stats(s)
        if len(scene.positions) == 0:
            assert st["density_iters"] == 0 and st["max_velocity"] == 0.0
            assert st["nan_count"] == 0 and st["mean_density_error"] == 0.0
            continue
        want, wstats = run_plain(scene, xb, cfg, dts[:, s])
        compare_scene(b.scene(s), want, f"scene {s}")
        compare_stats(st, wstats[-1], f"scene {s}")


def test_batch_scenes_match_oracle(walls):
    xb = walls
    cfg = SC.box_config(max_density_iters=12)
    mix = small_mix()
    b = Batch(mix, xb, cfg)
    for _ in range(3):
        b.step()
    dts = np.array(b.dts)
    for s in (0, 3, 5):  # uneven block, lone particle, no viscosity
        sc = mix[s]
        ref = O.Solver(O.config(**dict(cfg, viscosity=sc.viscosity)), sc.positions,
                       sc.velocities, xb, O.F32)
        for dt in dts[:, s]:
            ref.step(float(dt))
        got = b.scene(s)
        assert_bitexact(got["x"], ref.get(O.FIELD_X), f"scene {s} position")
        assert_bitexact(got["v"], ref.get(O.FIELD_V), f"scene {s} velocity")
        assert_bitexact(got["dfsph_density"], ref.get(O.FIELD_RHO), f"scene {s} density")
        assert_bitexact(got["dfsph_alpha"], ref.get(O.FIELD_ALPHA), f"scene {s} alpha")
        assert_bitexact(got["dfsph_kappa"], ref.get(O.FIELD_KAPPA), f"scene {s} kappa")
        rs, gs = ref.stats(), b.solver.scene_stats(s)
        assert gs["density_iters"] == rs["density_iters"]
        assert bits(gs["max_velocity"]) == bits(rs["max_velocity"])
        assert abs(gs["mean_density_error"] - rs["mean_density_error"]) <= \
            1e-12 * max(1.0, abs(rs["mean_density_error"]))


def test_graph_replay_is_bitwise_eager(walls):
    xb = walls
    cfg = SC.box_config(max_density_iters=20)
    mix = small_mix()
    a = Batch(mix, xb, cfg, graphs=True)
    e = Batch(mix, xb, cfg, graphs=False)
    for _ in range(4):
        a.step()
        e.step()
    for s in range(len(mix)):
        compare_scene(a.scene(s), e.scene(s), f"graph scene {s}")
        compare_stats(a.solver.scene_stats(s), e.solver.scene_stats(s), f"graph scene {s}")
    # plain solver with graphs against the eager plain solver
    dts = np.array(a.dts)
    g, gst = run_plain(mix[2], xb, cfg, dts[:, 2], graphs=True)
    p, pst = run_plain(mix[2], xb, cfg, dts[:, 2], graphs=False)
    compare_scene(g, p, "plain graph")
    compare_stats(gst[-1], pst[-1], "plain graph")


SHAPES = [
    dict(density_alpha_nt=64, xsph_nt=512, change_nt=512, change_pu=2, apply_nt=128, apply_pu=8,
         cell_cpt=4, remap=0),
    dict(density_alpha_nt=512, xsph_nt=64, change_nt=1024, change_pu=8, apply_nt=512, apply_pu=2,
         cell_cpt=8, remap=1),
    dict(density_alpha_nt=128, xsph_nt=128, change_nt=256, change_pu=4, apply_nt=64, apply_pu=4,
         cell_cpt=16, remap=0),
]


@pytest.mark.parametrize("k", range(len(SHAPES)))
def test_launch_shapes_are_bitwise_identical(walls, k):
    xb = walls
    cfg = SC.box_config(max_density_iters=15)
    scene = SC.box_scene(11, n_min=5000, n_max=6000)
    dts = [0.002, 0.0015, 0.002]
    want, wst = run_plain(scene, xb, cfg, dts)
    got, gst = run_plain(scene, xb, cfg, dts, shapes=SHAPES[k])
    compare_scene(got, want, f"shape {k}")
    compare_stats(gst[-1], wst[-1], f"shape {k}")


def test_autotune_keeps_results(walls):
    xb = walls
    cfg = SC.box_config(max_density_iters=15)
    mix = small_mix()
    ref = Batch(mix, xb, cfg)
    tuned = Batch(mix, xb, cfg)
    log = tuned.solver.autotune(reps=2)
    kinds = {r[0] for r in log}
    assert {"grid_build", "density_alpha", "xsph", "density_change", "apply_kappa"} <= kinds
    assert all(r[5] > 0.0 for r in log)
    for _ in range(3):
        ref.step()
        tuned.step()
    for s in range(len(mix)):
        compare_scene(tuned.scene(s), ref.scene(s), f"tuned scene {s}")
        compare_stats(tuned.solver.scene_stats(s), ref.solver.scene_stats(s), f"tuned {s}")
    print("\nautotuned shapes:", tuned.solver.shapes())


def test_batch_divergence_solve_matches_plain(walls):
    xb = walls
    cfg = SC.box_config(max_density_iters=10, enable_divergence_solve=True,
                        max_divergence_iters=6, pressure_form=1)
    mix = small_mix()[:4]
    b = Batch(mix, xb, cfg)
    for _ in range(3):
        b.step()
    dts = np.array(b.dts)
    fields = FIELDS[:2] + ("dfsph_kappa_v",)
    for s, scene in enumerate(mix):
        if len(scene.positions) == 0:
            continue
        want, wst = run_plain(scene, xb, cfg, dts[:, s])
        got = b.scene(s)
        for k in ("x", "v", "dfsph_density", "dfsph_alpha"):
            assert_bitexact(got[k], want[k], f"div scene {s} {k}")
        compare_stats(b.solver.scene_stats(s), wst[-1], f"div scene {s}")
    del fields


def test_uneven_mix_batched_equals_sequential(walls):
    xb = walls
    cfg = SC.box_config()
    mix = SC.uneven_mix(64)
    b = Batch(mix, xb, cfg)
    for _ in range(2):
        b.step()
    dts = np.array(b.dts)
    for s, scene in enumerate(mix):
        want, wst = run_plain(scene, xb, cfg, dts[:, s])
        got = b.scene(s)
        assert_bitexact(got["x"], want["x"], f"mix scene {s} position")
        assert_bitexact(got["v"], want["v"], f"mix scene {s} velocity")
        assert_bitexact(got["dfsph_density"], want["dfsph_density"], f"mix scene {s} density")
        compare_stats(b.solver.scene_stats(s), wst[-1], f"mix scene {s}")


def test_batch_argument_checks(walls):
    xb = walls
    cfg = ws.DFSPHConfig(**SC.box_config())
    x = np.full((10, 3), 0.3, np.float32)
    f = fluid_store(x, np.zeros_like(x))
    with pytest.raises(ws.InvalidArgument):
        ws.BatchSolver(f, None, cfg, [4, 5])          # counts do not add up to the store
    with pytest.raises(ValueError):
        ws.BatchSolver(f, None, cfg, [4, 6], [0.01])  # one viscosity per scene
    b = ws.BatchSolver(f, boundary_store(xb), cfg, [4, 0, 6])
    assert list(b.offsets) == [0, 4, 4, 10]
    with pytest.raises(ws.OutOfRange):
        b.scene_stats(3)
    f.resize(9)
    with pytest.raises(ws.InvalidArgument):
        b.step(0.001)                                # the store count changed
    plain = ws.DFSPHSolver(fluid_store(x, np.zeros_like(x)), None, cfg)
    d = (C.c_float * 1)(0.001)
    assert ws._ffi.lib().wsim_batch_step(plain.handle, d) == 3  # not a batch
    with pytest.raises(ValueError):
        plain.set_shapes(change_nt=128)              # k_change needs whole tiles
