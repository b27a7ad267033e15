"""ctypes wrapper of the CPU oracle (oracle/liboracle.so) and of the compiled
reference host sources (oracle/_ref/libref_host.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py -- never by the product package.
"""
from __future__ import annotations

import ctypes as C
import os
from pathlib import Path
from typing import Optional

import numpy as np

HERE = Path(__file__).resolve().parent
# WSIM_ORACLE_LIB selects another build of the oracle (the mutation suite's
# planted-mistake builds); the default is the one build() makes.
ORACLE_LIB = Path(os.environ.get("WSIM_ORACLE_LIB", str(HERE / "liboracle.so")))
REF_LIB = HERE / "_ref" / "libref_host.so"

F32, F64, DUAL = 32, 64, 33
FIELD_X, FIELD_V, FIELD_RHO, FIELD_ALPHA, FIELD_KAPPA, FIELD_RHO_ADV, FIELD_BMASS, FIELD_XB = range(8)

F3 = C.c_float * 3
I3 = C.c_int32 * 3
D3 = C.c_double * 3


class OrcConfig(C.Structure):
    _fields_ = [
        ("rest_density", C.c_float),
        ("particle_radius", C.c_float),
        ("smoothing_length", C.c_float),
        ("viscosity", C.c_float),
        ("damping", C.c_float),
        ("surface_tension", C.c_float),
        ("gravity", C.c_float * 3),
        ("max_density_iters", C.c_uint32),
        ("max_divergence_iters", C.c_uint32),
        ("eta_density", C.c_float),
        ("eta_divergence", C.c_float),
        ("domain_min", C.c_float * 3),
        ("domain_max", C.c_float * 3),
        ("enable_divergence_solve", C.c_uint32),
        ("pressure_form", C.c_uint32),
        ("enable_surface_tension", C.c_uint32),
    ]


class OrcStats(C.Structure):
    _fields_ = [
        ("density_iters", C.c_uint32),
        ("divergence_iters", C.c_uint32),
        ("mean_density_error", C.c_double),
        ("mean_divergence_error", C.c_double),
        ("max_velocity", C.c_float),
        ("nan_count", C.c_uint32),
        ("escapees", C.c_uint32),
        ("total_mass", C.c_double),
        ("momentum", C.c_double * 3),
        ("kinetic_energy", C.c_double),
    ]


vp, sz = C.c_void_p, C.c_size_t
_SIGS = {
    "orc_set_threads": (None, [C.c_int]),
    "orc_solver_seed": (None, [vp, vp, vp, C.c_double]),
    "orc_solver_set_state_f64": (None, [vp, vp, vp]),
    "orc_solver_set_viscosity_f64": (None, [vp, C.c_double]),
    "orc_solver_get_tangent": (None, [vp, C.c_int, vp]),
    "orc_tape_create": (vp, []),
    "orc_tape_destroy": (None, [vp]),
    "orc_tape_builds": (sz, [vp]),
    "orc_solver_set_tape": (None, [vp, vp, C.c_int]),
    "orc_set_xsph_in_place": (None, [C.c_int]),
    "orc_set_analysis": (None, [C.c_int, C.c_int]),
    "orc_get_threads": (C.c_int, []),
    "orc_W_f32": (C.c_float, [C.c_float, C.c_float]),
    "orc_W_f64": (C.c_double, [C.c_double, C.c_double]),
    "orc_gradW_f32": (None, [F3, C.c_float, F3]),
    "orc_gradW_f64": (None, [D3, C.c_double, D3]),
    "orc_cell_keys": (None, [vp, sz, F3, I3, C.c_float, vp]),
    "orc_grid_build": (None, [vp, sz, F3, I3, C.c_float, vp, vp, vp]),
    "orc_make_grid": (None, [F3, F3, C.c_float, F3, I3]),
    "orc_neighbor_counts": (None, [vp, sz, vp, sz, F3, I3, C.c_float, C.c_float, vp, vp]),
    "orc_morton": (None, [vp, sz, F3, F3, vp, vp, vp]),
    "orc_lbvh": (None, [vp, vp, vp, sz, vp, vp, vp, vp]),
    "orc_sample_aabb_boundary": (sz, [F3, F3, C.c_float, C.c_int, vp, sz]),
    "orc_lattice_block": (sz, [F3, F3, C.c_float, vp, sz]),
    "orc_damping_decay": (C.c_float, [C.c_float, C.c_float]),
    "orc_next_dt": (C.c_float, [C.c_float, C.c_float, C.c_float, C.c_uint32, C.c_float,
                                C.c_float, C.c_float]),
    "orc_config_default": (None, [C.POINTER(OrcConfig)]),
    "orc_solver_create": (vp, [C.POINTER(OrcConfig), C.c_int, vp, vp, sz, vp, sz]),
    "orc_solver_destroy": (None, [vp]),
    "orc_solver_step": (None, [vp, C.c_float]),
    "orc_solver_density_alpha": (None, [vp]),
    "orc_solver_surface_tension": (None, [vp, C.c_float]),
    "orc_solver_set_state": (None, [vp, vp, vp]),
    "orc_solver_get": (None, [vp, C.c_int, vp]),
    "orc_solver_get_f32": (None, [vp, C.c_int, vp]),
    "orc_solver_stats": (None, [vp, C.POINTER(OrcStats)]),
    "orc_solver_particle_mass": (C.c_float, [vp]),
    "orc_solver_count": (sz, [vp]),
}

_lib = None
_ref = None


def lib() -> C.CDLL:
    global _lib
    if _lib is None:
        if not ORACLE_LIB.exists():
            raise RuntimeError(f"{ORACLE_LIB} missing: run __graft_entry__.build()")
        h = C.CDLL(str(ORACLE_LIB))
        for n, (r, a) in _SIGS.items():
            f = getattr(h, n)
            f.restype, f.argtypes = r, a
        _lib = h
    return _lib


def ref() -> Optional[C.CDLL]:
    """The compiled reference host sources, or None where they were not built."""
    global _ref
    if _ref is None and REF_LIB.exists():
        h = C.CDLL(str(REF_LIB))
        h.ref_sample_aabb_boundary.restype = sz
        h.ref_sample_aabb_boundary.argtypes = [F3, F3, C.c_float, C.c_int, vp, sz]
        h.ref_next_dt.restype = C.c_float
        h.ref_next_dt.argtypes = [C.c_float, C.c_float, C.c_float, C.c_uint32, C.c_float,
                                  C.c_float, C.c_float]
        _ref = h
    return _ref


def _f3(v):
    return F3(*[float(x) for x in v])


def _xyz(a) -> np.ndarray:
    return np.ascontiguousarray(np.asarray(a, dtype=np.float32).reshape(-1, 3))


def set_threads(n: int) -> None:
    lib().orc_set_threads(int(n))


def W(r: float, l: float, precision: int = F32) -> float:
    return lib().orc_W_f64(r, l) if precision == F64 else lib().orc_W_f32(r, l)


def gradW(r, l: float, precision: int = F32) -> np.ndarray:
    if precision == F64:
        out = D3()
        lib().orc_gradW_f64(D3(*[float(x) for x in r]), l, out)
        return np.array(list(out), dtype=np.float64)
    out = F3()
    lib().orc_gradW_f32(_f3(r), l, out)
    return np.array(list(out), dtype=np.float32)


def make_grid(dmin, dmax, h: float):
    o, d = F3(), I3()
    lib().orc_make_grid(_f3(dmin), _f3(dmax), h, o, d)
    return tuple(o), tuple(d)


def cell_keys(x, origin, dims, cell: float) -> np.ndarray:
    x = _xyz(x)
    out = np.empty(len(x), np.uint32)
    lib().orc_cell_keys(x.ctypes.data, len(x), _f3(origin), I3(*dims), cell, out.ctypes.data)
    return out


def grid_build(x, origin, dims, cell: float):
    """-> (sorted_keys, perm, cell_start)"""
    x = _xyz(x)
    n = len(x)
    C_ = int(np.prod(np.asarray(dims, dtype=np.int64)))
    keys = np.empty(n, np.uint32)
    perm = np.empty(n, np.uint32)
    cs = np.empty(C_ + 1, np.uint32)
    lib().orc_grid_build(x.ctypes.data, n, _f3(origin), I3(*dims), cell, keys.ctypes.data,
                         perm.ctypes.data, cs.ctypes.data)
    return keys, perm, cs


def neighbor_counts(xq, xs, origin, dims, cell: float, h: float):
    xq, xs = _xyz(xq), _xyz(xs)
    cand = np.empty(len(xq), np.uint32)
    nbr = np.empty(len(xq), np.uint32)
    lib().orc_neighbor_counts(xq.ctypes.data, len(xq), xs.ctypes.data, len(xs), _f3(origin),
                              I3(*dims), cell, h, cand.ctypes.data, nbr.ctypes.data)
    return cand, nbr


def morton(x, bmin, bmax):
    """-> (codes in input order, sorted codes, perm)"""
    x = _xyz(x)
    n = len(x)
    codes, sc, perm = (np.empty(n, np.uint32) for _ in range(3))
    lib().orc_morton(x.ctypes.data, n, _f3(bmin), _f3(bmax), codes.ctypes.data, sc.ctypes.data,
                     perm.ctypes.data)
    return codes, sc, perm


def lbvh(x, sorted_codes, perm):
    """Karras radix tree and node boxes over sorted Morton codes
    (spatial_accel.cu:44-113): -> (left, right, parent, boxes (2n-1, 2, 3))."""
    x = _xyz(x)
    n = len(sorted_codes)
    ni = max(n - 1, 0)
    left, right = np.zeros(ni, np.uint32), np.zeros(ni, np.uint32)
    parent = np.zeros(n + ni, np.uint32)
    boxes = np.zeros((n + ni, 2, 3), np.float32)
    sc = np.ascontiguousarray(sorted_codes, np.uint32)
    pm = np.ascontiguousarray(perm, np.uint32)
    lib().orc_lbvh(sc.ctypes.data, x.ctypes.data, pm.ctypes.data, n, left.ctypes.data,
                   right.ctypes.data, parent.ctypes.data, boxes.ctypes.data)
    return left, right, parent, boxes


def sample_aabb_boundary(lo, hi, spacing: float, n_layers: int = 3, use_reference=False):
    L = ref() if use_reference else lib()
    if L is None:
        raise RuntimeError("oracle/_ref not built")
    fn = L.ref_sample_aabb_boundary if use_reference else L.orc_sample_aabb_boundary
    n = fn(_f3(lo), _f3(hi), spacing, n_layers, None, 0)
    out = np.empty((n, 3), np.float32)
    fn(_f3(lo), _f3(hi), spacing, n_layers, out.ctypes.data, n)
    return out


def lattice_block(lo, hi, spacing: float) -> np.ndarray:
    n = lib().orc_lattice_block(_f3(lo), _f3(hi), spacing, None, 0)
    out = np.empty((n, 3), np.float32)
    lib().orc_lattice_block(_f3(lo), _f3(hi), spacing, out.ctypes.data, n)
    return out


def set_analysis(freeze_geometry: bool = False, free_advect: int = 0) -> None:
    """Analysis-only switches (0 everywhere else): dual-number kernel terms
    without position tangents; advection without the |v| clamp (bit 0) or the
    domain clamp (bit 1)."""
    lib().orc_set_analysis(int(freeze_geometry), int(free_advect))


def damping_decay(damping, dt) -> float:
    return lib().orc_damping_decay(damping, dt)


def next_dt(frame_dt, dt_max, cfl_lambda, max_substeps, max_velocity, particle_radius,
            remaining, use_reference=False) -> float:
    if use_reference:
        return ref().ref_next_dt(frame_dt, dt_max, cfl_lambda, max_substeps, max_velocity,
                                 particle_radius, remaining)
    return lib().orc_next_dt(frame_dt, dt_max, cfl_lambda, max_substeps, max_velocity,
                             particle_radius, remaining)


def config(**kw) -> OrcConfig:
    """The oracle evaluates the reference's expressions in IEEE arithmetic:
    the solver's tolerance-mode switch (fast_math) has no meaning here."""
    c = OrcConfig()
    lib().orc_config_default(C.byref(c))
    for k, v in kw.items():
        if k == "fast_math":
            continue
        if k in ("gravity", "domain_min", "domain_max"):
            setattr(c, k, (C.c_float * 3)(*[float(x) for x in v]))
        elif k in ("max_density_iters", "max_divergence_iters", "pressure_form",
                   "enable_divergence_solve", "enable_surface_tension"):
            setattr(c, k, int(v))
        else:
            setattr(c, k, float(v))
    return c


class Solver:
    """Stateful oracle DFSPH solver (dfsph.cu:71-133 restated)."""

    def __init__(self, cfg: OrcConfig, x, v=None, xb=None, precision: int = F32):
        self.x0 = _xyz(x)
        self.n = len(self.x0)
        v = None if v is None else _xyz(v)
        xb = None if xb is None else _xyz(xb)
        self.nb = 0 if xb is None else len(xb)
        self.precision = precision
        self.cfg = cfg
        self._h = lib().orc_solver_create(C.byref(cfg), precision, self.x0.ctypes.data,
                                          None if v is None else v.ctypes.data, self.n,
                                          None if xb is None else xb.ctypes.data, self.nb)

    def __del__(self):
        try:
            if self._h:
                lib().orc_solver_destroy(self._h)
                self._h = None
        except Exception:
            pass

    def step(self, dt: float) -> None:
        lib().orc_solver_step(self._h, float(dt))

    def density_alpha(self) -> None:
        lib().orc_solver_density_alpha(self._h)

    def surface_tension(self, dt: float) -> None:
        """The surface-tension phase alone, on the state after density_alpha()."""
        lib().orc_solver_surface_tension(self._h, float(dt))

    def set_state(self, x=None, v=None) -> None:
        xa = None if x is None else _xyz(x)
        va = None if v is None else _xyz(v)
        lib().orc_solver_set_state(self._h, None if xa is None else xa.ctypes.data,
                                   None if va is None else va.ctypes.data)

    def set_state_f64(self, x=None, v=None) -> None:
        xa = None if x is None else np.ascontiguousarray(x, np.float64).reshape(-1, 3)
        va = None if v is None else np.ascontiguousarray(v, np.float64).reshape(-1, 3)
        lib().orc_solver_set_state_f64(self._h, None if xa is None else xa.ctypes.data,
                                       None if va is None else va.ctypes.data)

    def set_viscosity_f64(self, viscosity: float) -> None:
        lib().orc_solver_set_viscosity_f64(self._h, float(viscosity))

    def _size(self, field: int) -> tuple:
        if field in (FIELD_X, FIELD_V):
            return (self.n, 3)
        if field == FIELD_XB:
            return (self.nb, 3)
        if field == FIELD_BMASS:
            return (self.nb,)
        return (self.n,)

    def get(self, field: int) -> np.ndarray:
        shape = self._size(field)
        if self.precision == F32:
            out = np.empty(shape, np.float32)
            lib().orc_solver_get_f32(self._h, field, out.ctypes.data)
            return out
        out = np.empty(shape, np.float64)
        lib().orc_solver_get(self._h, field, out.ctypes.data)
        return out

    def stats(self) -> dict:
        s = OrcStats()
        lib().orc_solver_stats(self._h, C.byref(s))
        return {f[0]: (list(getattr(s, f[0])) if f[0] == "momentum" else getattr(s, f[0]))
                for f in s._fields_}

    # gradient checks
    def seed(self, dx=None, dv=None, d_viscosity: float = 0.0) -> None:
        """Tangent seeds of a DUAL solver (initial positions, velocities, viscosity)."""
        assert self.precision == DUAL
        ax = None if dx is None else np.ascontiguousarray(dx, np.float64).reshape(-1, 3)
        av = None if dv is None else np.ascontiguousarray(dv, np.float64).reshape(-1, 3)
        lib().orc_solver_seed(self._h, None if ax is None else ax.ctypes.data,
                              None if av is None else av.ctypes.data, float(d_viscosity))

    def tangent(self, field: int) -> np.ndarray:
        assert self.precision == DUAL
        out = np.empty(self._size(field), np.float64)
        lib().orc_solver_get_tangent(self._h, field, out.ctypes.data)
        return out

    def set_tape(self, tape: "Tape", mode: int) -> None:
        """mode 1: record the neighbour structure and iteration counts; 2: replay."""
        self._tape = tape
        lib().orc_solver_set_tape(self._h, tape._h if tape else None, int(mode))

    def particle_mass(self) -> float:
        return lib().orc_solver_particle_mass(self._h)


class Tape:
    """Frozen structure (per grid build: home cells; per solve: iteration count)."""

    def __init__(self):
        self._h = lib().orc_tape_create()

    def builds(self) -> int:
        return lib().orc_tape_builds(self._h)

    def __del__(self):
        try:
            if self._h:
                lib().orc_tape_destroy(self._h)
                self._h = None
        except Exception:
            pass
