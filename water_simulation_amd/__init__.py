"""water_simulation_amd -- MI355X-native DFSPH neighbour search + pressure solve.

Python mirror of the reference's solver surface (Qervas/water_simulation),
calling the HIP library libwsim.so through its C ABI (include/wsim.h):

    ParticleStore      core/include/water/core/particle_store.h:48-107
    CellGrid           core/include/water/core/cell_grid.h:16-50
    SpatialAccel       core/include/water/core/spatial_accel.h:19-53 (Morton part)
    DFSPHSolver        solvers/dfsph/include/water/solvers/dfsph.h:17-84
    TimeStepper        core/include/water/core/timestep.h:8-34
    sample_aabb_boundary  solvers/dfsph/include/water/solvers/boundary_sampler.h:24-25
    load_scene         scene/include/water/scene/scene.h:45

Errors follow the reference's exception types: std::out_of_range ->
OutOfRange (an IndexError), std::runtime_error -> WsimError (a RuntimeError).
Device pointers are plain integers; helpers move numpy arrays in and out.
"""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass, field, fields
from typing import Optional, Sequence

import numpy as np

from . import _ffi
from ._ffi import DeviceError, InvalidArgument, OutOfRange, WsimError, check

__all__ = [
    "AttribType", "ParticleStore", "CellGrid", "SpatialAccel", "DFSPHConfig", "DFSPHSolver",
    "BatchSolver", "TimeStepper", "TimeStepperConfig", "Scene", "load_scene", "sample_aabb_boundary",
    "lattice_block", "write_frame", "read_frame", "DeviceBuffer", "memcpy_h2d", "memcpy_d2h",
    "device_count", "device_synchronize", "WsimError", "OutOfRange", "InvalidArgument",
    "DeviceError", "library_path", "pair_ops_check",
]


def library_path() -> str:
    return str(_ffi.LIB_PATH)


def _L():
    return _ffi.lib()


def _stream(s) -> Optional[int]:
    return None if s is None else int(s)


class AttribType:
    F32 = 0
    F32x3 = 1
    U32 = 2


def device_count() -> int:
    n = C.c_int(0)
    check(_L().wsim_device_count(C.byref(n)))
    return n.value


def device_synchronize() -> None:
    check(_L().wsim_device_synchronize())


def pair_ops_check(smoothing_length: float) -> dict:
    """Run every operand the solver kernels can produce for this smoothing
    length through the pair-path sqrt / quotients and the IEEE operations on
    the device; returns whether the kernels use the pair path and the counts
    of results that differ in any bit."""
    out = (C.c_uint64 * 4)()
    check(_L().wsim_pair_ops_check(float(smoothing_length), out))
    return {"fast": bool(out[0]), "sqrt": int(out[1]), "quotient": int(out[2]),
            "gradient": int(out[3])}


def memcpy_h2d(dst: int, array: np.ndarray, stream=None) -> None:
    a = np.ascontiguousarray(array)
    check(_L().wsim_memcpy_h2d(dst, a.ctypes.data, a.nbytes, _stream(stream)))


def memcpy_d2h(out: np.ndarray, src: int, stream=None) -> np.ndarray:
    assert out.flags["C_CONTIGUOUS"]
    check(_L().wsim_memcpy_d2h(out.ctypes.data, src, out.nbytes, _stream(stream)))
    return out


class DeviceBuffer:
    """Owned device allocation (device_buffer.h:21-89)."""

    def __init__(self, nbytes: int):
        p = C.c_void_p()
        check(_L().wsim_device_alloc(nbytes, C.byref(p)))
        self.ptr = p.value or 0
        self.nbytes = nbytes

    @classmethod
    def from_array(cls, array: np.ndarray) -> "DeviceBuffer":
        a = np.ascontiguousarray(array)
        b = cls(max(a.nbytes, 4))
        if a.nbytes:
            memcpy_h2d(b.ptr, a)
        return b

    def to_array(self, dtype, count: int) -> np.ndarray:
        out = np.empty(count, dtype=dtype)
        if out.nbytes:
            memcpy_d2h(out, self.ptr)
        return out

    def free(self) -> None:
        if self.ptr:
            _L().wsim_device_free(self.ptr)
            self.ptr = 0

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


def _f3(v) -> C.Array:
    return _ffi.F3(*(float(x) for x in v))


def _read_u32(ptr: int, count: int) -> np.ndarray:
    out = np.empty(count, dtype=np.uint32)
    if count:
        if not ptr:
            raise WsimError("buffer not allocated")
        memcpy_d2h(out, ptr)
    return out


# --------------------------------------------------------------------------
class ParticleStore:
    """Named per-particle device attributes (particle_store.h:48-107)."""

    def __init__(self, capacity: int, stream=None):
        h = C.c_void_p()
        check(_L().wsim_store_create(capacity, _stream(stream), C.byref(h)))
        self._h = h.value
        self.stream = stream

    @property
    def handle(self) -> int:
        return self._h

    def close(self) -> None:
        if self._h:
            _L().wsim_store_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def capacity(self) -> int:
        return _L().wsim_store_capacity(self._h)

    def count(self) -> int:
        return _L().wsim_store_count(self._h)

    def resize(self, count: int) -> None:
        check(_L().wsim_store_resize(self._h, count))

    def register_attribute(self, name: str, type: int) -> int:
        h = C.c_size_t()
        check(_L().wsim_store_register_attribute(self._h, name.encode(), type, C.byref(h)))
        return h.value

    def has_attribute(self, name: str) -> bool:
        return bool(_L().wsim_store_has_attribute(self._h, name.encode()))

    def find_attribute(self, name: str) -> int:
        h = C.c_size_t()
        check(_L().wsim_store_find_attribute(self._h, name.encode(), C.byref(h)))
        return h.value

    def attribute_data(self, handle: int) -> int:
        p = C.c_void_p()
        check(_L().wsim_store_attribute_data(self._h, handle, C.byref(p)))
        return p.value or 0

    def positions(self) -> int:
        p = C.c_void_p()
        check(_L().wsim_store_positions(self._h, C.byref(p)))
        return p.value or 0

    def velocities(self) -> int:
        p = C.c_void_p()
        check(_L().wsim_store_velocities(self._h, C.byref(p)))
        return p.value or 0

    def position_handle(self) -> int:
        return _L().wsim_store_position_handle(self._h)

    def velocity_handle(self) -> int:
        return _L().wsim_store_velocity_handle(self._h)

    # numpy helpers (N x 3 float32, original particle order)
    def set_positions(self, xyz: np.ndarray) -> None:
        memcpy_h2d(self.positions(), np.asarray(xyz, dtype=np.float32).reshape(-1, 3))

    def set_velocities(self, xyz: np.ndarray) -> None:
        memcpy_h2d(self.velocities(), np.asarray(xyz, dtype=np.float32).reshape(-1, 3))

    def get_positions(self, n: Optional[int] = None) -> np.ndarray:
        n = self.count() if n is None else n
        return memcpy_d2h(np.empty((n, 3), np.float32), self.positions())

    def get_velocities(self, n: Optional[int] = None) -> np.ndarray:
        n = self.count() if n is None else n
        return memcpy_d2h(np.empty((n, 3), np.float32), self.velocities())

    def get_attribute(self, handle, dtype=np.float32, width: int = 1,
                      n: Optional[int] = None) -> np.ndarray:
        if isinstance(handle, str):
            handle = self.find_attribute(handle)
        n = self.count() if n is None else n
        shape = (n,) if width == 1 else (n, width)
        return memcpy_d2h(np.empty(shape, dtype), self.attribute_data(handle))


# --------------------------------------------------------------------------
class CellGrid:
    """Uniform cell grid (cell_grid.h:16-50)."""

    def __init__(self, origin: Sequence[float], cells_per_axis: Sequence[int], cell_length: float,
                 stream=None, _view: Optional[int] = None, _owner=None):
        self._owner = _owner
        if _view is not None:
            self._h = _view
            self._owned = False
            return
        h = C.c_void_p()
        check(_L().wsim_grid_create(_f3(origin), _ffi.I3(*[int(d) for d in cells_per_axis]),
                                    float(cell_length), _stream(stream), C.byref(h)))
        self._h = h.value
        self._owned = True

    def close(self) -> None:
        if self._h and self._owned:
            _L().wsim_grid_destroy(self._h)
        self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def reserve(self, capacity: int) -> None:
        check(_L().wsim_grid_reserve(self._h, capacity))

    def build(self, positions: int, count: int) -> None:
        check(_L().wsim_grid_build(self._h, positions, count))

    def cell_start(self) -> int:
        return _L().wsim_grid_cell_start(self._h) or 0

    def sorted_indices(self) -> int:
        return _L().wsim_grid_sorted_indices(self._h) or 0

    def sorted_keys(self) -> int:
        return _L().wsim_grid_sorted_keys(self._h) or 0

    def origin(self):
        o = _ffi.F3()
        _L().wsim_grid_origin(self._h, o)
        return tuple(o)

    def cells_per_axis(self):
        d = _ffi.I3()
        _L().wsim_grid_cells_per_axis(self._h, d)
        return tuple(d)

    def cell_length(self) -> float:
        return _L().wsim_grid_cell_length(self._h)

    def total_cells(self) -> int:
        return _L().wsim_grid_total_cells(self._h)

    def count(self) -> int:
        return _L().wsim_grid_count(self._h)

    def read_cell_start(self) -> np.ndarray:
        return _read_u32(self.cell_start(), self.total_cells() + 1)

    def read_sorted_indices(self) -> np.ndarray:
        return _read_u32(self.sorted_indices(), self.count())

    def read_sorted_keys(self) -> np.ndarray:
        return _read_u32(self.sorted_keys(), self.count())


# --------------------------------------------------------------------------
class SpatialAccel:
    """Morton codes, their stable sort and the linear BVH over them
    (spatial_accel.h:19-53, spatial_accel.cu:44-199)."""

    def __init__(self):
        h = C.c_void_p()
        check(_L().wsim_accel_create(C.byref(h)))
        self._h = h.value

    def close(self) -> None:
        if self._h:
            _L().wsim_accel_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def build(self, positions: int, count: int, bounds_min, bounds_max, stream=None) -> None:
        check(_L().wsim_accel_build(self._h, positions, count, _f3(bounds_min), _f3(bounds_max),
                                    _stream(stream)))

    def leaf_count(self) -> int:
        return _L().wsim_accel_leaf_count(self._h)

    def read_sorted_indices(self) -> np.ndarray:
        return _read_u32(_L().wsim_accel_sorted_indices(self._h) or 0, self.leaf_count())

    def read_morton_codes(self) -> np.ndarray:
        return _read_u32(_L().wsim_accel_morton_codes(self._h) or 0, self.leaf_count())

    def read_internal_nodes(self):
        """(left, right, parent): children of the n - 1 internal nodes (a child
        >= n - 1 is leaf child - (n - 1)) and the parent of every node (root:
        0xFFFFFFFF)."""
        n = self.leaf_count()
        ni = max(n - 1, 0)
        return (_read_u32(_L().wsim_accel_internal_left(self._h) or 0, ni),
                _read_u32(_L().wsim_accel_internal_right(self._h) or 0, ni),
                _read_u32(_L().wsim_accel_internal_parent(self._h) or 0, n + ni))

    def read_node_aabbs(self) -> np.ndarray:
        """(2n - 1, 2, 3) float32: min and max of every node, internal nodes first."""
        n = self.leaf_count()
        m = 2 * n - 1 if n else 0
        out = np.empty((m, 2, 3), np.float32)
        if m:
            memcpy_d2h(out, _L().wsim_accel_node_aabbs(self._h))
        return out


# --------------------------------------------------------------------------
@dataclass
class DFSPHConfig:
    """DFSPHSolver::Config (dfsph.h:19-33) plus two extensions."""
    rest_density: float = 1000.0
    particle_radius: float = 0.0015
    smoothing_length: float = 0.003
    viscosity: float = 1e-3
    damping: float = 0.5
    surface_tension: float = 0.0
    gravity: tuple = (0.0, -9.81, 0.0)
    max_density_iters: int = 100
    max_divergence_iters: int = 100
    eta_density: float = 1e-3
    eta_divergence: float = 1e-3
    domain_min: tuple = (0.0, 0.0, 0.0)
    domain_max: tuple = (1.0, 1.0, 1.0)
    enable_divergence_solve: bool = False
    pressure_form: int = 0  # 0: reference kappa/rho^2, 1: kappa/rho
    enable_surface_tension: bool = False  # Akinci 2013, coefficient surface_tension
    fast_math: bool = False  # tolerance mode of the pair terms (hardware rsqrt/rcp, fma)

    def to_c(self) -> _ffi.DFSPHConfigC:
        c = _ffi.DFSPHConfigC()
        for f in fields(self):
            v = getattr(self, f.name)
            if f.name in ("gravity", "domain_min", "domain_max"):
                setattr(c, f.name, (C.c_float * 3)(*[float(x) for x in v]))
            elif f.name in ("max_density_iters", "max_divergence_iters", "pressure_form"):
                setattr(c, f.name, int(v))
            elif f.name in ("enable_divergence_solve", "enable_surface_tension", "fast_math"):
                setattr(c, f.name, 1 if v else 0)
            else:
                setattr(c, f.name, float(v))
        return c


class DFSPHSolver:
    """DFSPH with Akinci boundary particles and XSPH (dfsph.h:17-84)."""

    def __init__(self, fluid: ParticleStore, boundary: Optional[ParticleStore],
                 cfg: DFSPHConfig, stream=None):
        self._fluid = fluid
        self._boundary = boundary
        self.cfg = cfg
        h = C.c_void_p()
        c = cfg.to_c()
        check(_L().wsim_solver_create(fluid.handle, boundary.handle if boundary else None,
                                      C.byref(c), _stream(stream), C.byref(h)))
        self._h = h.value

    def close(self) -> None:
        if self._h:
            _L().wsim_solver_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    @property
    def handle(self) -> int:
        return self._h

    def step(self, dt: float) -> None:
        check(_L().wsim_solver_step(self._h, float(dt)))

    def advance(self, time: float, max_substeps: int, stepper: Optional["TimeStepper"] = None,
                v_floor: float = 0.0, with_stats: bool = False):
        """sim_cli's substep loop (main.cpp:171-178) in one call: substeps of
        dt = stepper.next_dt(max(max_velocity(), v_floor), r, remaining) until
        the remaining time is at most 1e-6 or max_substeps ran.  Returns the
        time steps taken (and, with_stats, each substep's statistics); each
        substep computes the same bits as step(dt)."""
        cfg = (stepper or TimeStepper()).cfg
        c = _ffi.TimeStepConfigC(cfg.frame_dt, cfg.dt_max, cfg.cfl_lambda, cfg.max_substeps)
        m = max(int(max_substeps), 1)
        dts = np.zeros(m, np.float32)
        recs = (_ffi.StepStatsC * m)() if with_stats else None
        n = C.c_uint32()
        check(_L().wsim_solver_advance(self._h, C.byref(c), float(v_floor), float(time),
                                       int(max_substeps),
                                       dts.ctypes.data_as(C.POINTER(C.c_float)), recs,
                                       C.byref(n)))
        dts = dts[: n.value].copy()
        if not with_stats:
            return dts
        return dts, [recs[k].as_dict() for k in range(n.value)]

    def max_velocity(self) -> float:
        return _L().wsim_solver_max_velocity(self._h)

    def particle_mass(self) -> float:
        return _L().wsim_solver_particle_mass(self._h)

    def last_density_iters(self) -> int:
        return _L().wsim_solver_last_density_iters(self._h)

    def _at(self, fn, i: int) -> float:
        v = C.c_float()
        check(fn(self._h, i, C.byref(v)))
        return v.value

    def density_at(self, i: int) -> float:
        return self._at(_L().wsim_solver_density_at, i)

    def alpha_at(self, i: int) -> float:
        return self._at(_L().wsim_solver_alpha_at, i)

    def boundary_mass_at(self, i: int) -> float:
        return self._at(_L().wsim_solver_boundary_mass_at, i)

    def stats(self) -> dict:
        s = _ffi.StepStatsC()
        check(_L().wsim_solver_stats(self._h, C.byref(s)))
        return s.as_dict()

    def neighbor_counts(self) -> dict:
        """Per-particle neighbour counts of the density/alpha walk on the
        store's current positions (original order): stencil candidates and
        in-support neighbours in the fluid and the boundary store
        (ITER_NEIGHBORS, dfsph_kernels.cu:13-35)."""
        n = self._fluid.count()
        bufs = [DeviceBuffer(max(4 * n, 4)) for _ in range(4)]
        check(_L().wsim_solver_neighbor_counts(self._h, *(b.ptr for b in bufs)))
        out = {}
        for k, b in zip(("cand_fluid", "nbr_fluid", "cand_boundary", "nbr_boundary"), bufs):
            a = np.empty(n, np.uint32)
            if n:
                memcpy_d2h(a, b.ptr)
            out[k] = a
        return out

    def fluid_grid(self) -> CellGrid:
        return CellGrid(None, None, None, _view=_L().wsim_solver_fluid_grid(self._h), _owner=self)

    def boundary_grid(self) -> Optional[CellGrid]:
        v = _L().wsim_solver_boundary_grid(self._h)
        return CellGrid(None, None, None, _view=v, _owner=self) if v else None

    def set_timing(self, on: bool) -> None:
        check(_L().wsim_solver_set_timing(self._h, 1 if on else 0))

    def reset_timing(self) -> None:
        check(_L().wsim_solver_reset_timing(self._h))

    def kernel_times(self) -> dict:
        out = {}
        for k in range(_L().wsim_solver_kernel_kinds()):
            n = C.c_uint64()
            ms = C.c_double()
            check(_L().wsim_solver_kernel_time(self._h, k, C.byref(n), C.byref(ms)))
            out[_L().wsim_solver_kernel_name(k).decode()] = (n.value, ms.value)
        return out

    # launch shapes, autotuner and graph replay (no counterpart in the reference)
    def shapes(self) -> dict:
        c = _ffi.KernelShapesC()
        check(_L().wsim_solver_get_shapes(self._h, C.byref(c)))
        return {f[0]: getattr(c, f[0]) for f in c._fields_}

    def set_shapes(self, **kw) -> None:
        cur = self.shapes()
        unknown = set(kw) - set(cur)
        if unknown:
            raise ValueError(f"unknown shape fields {sorted(unknown)}")
        cur.update(kw)
        c = _ffi.KernelShapesC(**cur)
        check(_L().wsim_solver_set_shapes(self._h, C.byref(c)))

    def autotune(self, reps: int = 5) -> list:
        """Times every launch-shape candidate on the current state and keeps
        the fastest; returns [(kernel, workgroup, pairs/trip, remap, cells/thread, ms)]."""
        check(_L().wsim_solver_autotune(self._h, int(reps)))
        out = []
        shape = (C.c_uint32 * 5)()
        ms = C.c_double()
        for i in range(_L().wsim_solver_tune_count(self._h)):
            check(_L().wsim_solver_tune_entry(self._h, i, shape, C.byref(ms)))
            out.append((_L().wsim_solver_kernel_name(shape[0]).decode(), shape[1], shape[2],
                        shape[3], shape[4], ms.value))
        return out

    def set_graphs(self, on: bool) -> None:
        check(_L().wsim_solver_set_graphs(self._h, 1 if on else 0))

    def set_viscosity(self, viscosity: float) -> None:
        """XSPH viscosity of a plain solver (the differentiable parameter)."""
        check(_L().wsim_solver_set_viscosity(self._h, float(viscosity)))
        self.cfg.viscosity = float(viscosity)


class BatchSolver(DFSPHSolver):
    """Many independent scenes of uneven sizes in one solver: one packed
    working set, one segmented grid sort and one launch per phase, substeps
    replayed as HIP graphs.  Scene s owns fluid-store indices
    [offsets[s], offsets[s+1]); it steps with its own dt, viscosity,
    convergence test and statistics, and matches a plain DFSPHSolver on that
    scene alone bit for bit.  The boundary store is shared by all scenes."""

    def __init__(self, fluid: ParticleStore, boundary: Optional[ParticleStore],
                 cfg: DFSPHConfig, counts, viscosity=None, stream=None):
        self._fluid = fluid
        self._boundary = boundary
        self.cfg = cfg
        counts = np.ascontiguousarray(counts, dtype=np.uint32)
        self.num_scenes = int(counts.size)
        visc = None
        if viscosity is not None:
            visc = np.ascontiguousarray(viscosity, dtype=np.float32)
            if visc.size != counts.size:
                raise ValueError("one viscosity per scene")
        h = C.c_void_p()
        c = cfg.to_c()
        self._h = None
        check(_L().wsim_batch_create(
            fluid.handle, boundary.handle if boundary else None, C.byref(c), self.num_scenes,
            counts.ctypes.data_as(C.POINTER(C.c_uint32)),
            visc.ctypes.data_as(C.POINTER(C.c_float)) if visc is not None else None,
            _stream(stream), C.byref(h)))
        self._h = h.value
        off = (C.c_uint32 * (self.num_scenes + 1))()
        check(_L().wsim_batch_scene_offsets(self._h, off))
        self.offsets = np.array(off[:], dtype=np.int64)

    def step(self, dt) -> None:
        d = np.ascontiguousarray(np.broadcast_to(np.asarray(dt, np.float32), (self.num_scenes,)),
                                 dtype=np.float32)
        check(_L().wsim_batch_step(self._h, d.ctypes.data_as(C.POINTER(C.c_float))))

    def scene_stats(self, s: int) -> dict:
        st = _ffi.StepStatsC()
        check(_L().wsim_batch_scene_stats(self._h, int(s), C.byref(st)))
        return st.as_dict()

    def all_stats(self) -> list:
        return [self.scene_stats(s) for s in range(self.num_scenes)]

    def max_velocities(self) -> np.ndarray:
        return np.array([self.scene_stats(s)["max_velocity"] for s in range(self.num_scenes)],
                        np.float32)


# --------------------------------------------------------------------------
@dataclass
class TimeStepperConfig:
    frame_dt: float = 1.0 / 24.0
    dt_max: float = 1.0 / 24.0
    cfl_lambda: float = 0.4
    max_substeps: int = 16


class TimeStepper:
    """CFL time step (timestep.h:8-34)."""

    def __init__(self, cfg: Optional[TimeStepperConfig] = None):
        self.cfg = cfg or TimeStepperConfig()

    def next_dt(self, max_velocity: float, particle_radius: float, remaining: float) -> float:
        c = _ffi.TimeStepConfigC(self.cfg.frame_dt, self.cfg.dt_max, self.cfg.cfl_lambda,
                                 self.cfg.max_substeps)
        return _L().wsim_timestep_next_dt(C.byref(c), max_velocity, particle_radius, remaining)


def sample_aabb_boundary(lo, hi, spacing: float, n_layers: int = 3) -> np.ndarray:
    """Akinci wall particles on the six faces of an AABB (boundary_sampler.cpp:6-32)."""
    n = C.c_size_t()
    check(_L().wsim_sample_aabb_boundary(_f3(lo), _f3(hi), spacing, n_layers, None, 0, C.byref(n)))
    out = np.empty((n.value, 3), np.float32)
    check(_L().wsim_sample_aabb_boundary(_f3(lo), _f3(hi), spacing, n_layers, out.ctypes.data,
                                         n.value, C.byref(n)))
    return out


def lattice_block(lo, hi, spacing: float) -> np.ndarray:
    """sim_cli's fluid block (main.cpp:65-74)."""
    n = C.c_size_t()
    check(_L().wsim_lattice_block(_f3(lo), _f3(hi), spacing, None, 0, C.byref(n)))
    out = np.empty((n.value, 3), np.float32)
    check(_L().wsim_lattice_block(_f3(lo), _f3(hi), spacing, out.ctypes.data, n.value,
                                  C.byref(n)))
    return out


def write_frame(path: str, positions: np.ndarray) -> None:
    a = np.ascontiguousarray(positions, dtype=np.float32).reshape(-1, 3)
    check(_L().wsim_write_frame(str(path).encode(), a.ctypes.data, a.shape[0]))


def read_frame(path: str) -> np.ndarray:
    """Reader of the frame format (tools/viz.py:42-46): u32 N + N x 3 float32 LE."""
    raw = open(path, "rb").read()
    n = int(np.frombuffer(raw[:4], "<u4")[0])
    return np.frombuffer(raw[4:4 + 12 * n], "<f4").reshape(n, 3).copy()


@dataclass
class Scene:
    name: str = ""
    width: int = 1920
    height: int = 1080
    fps: float = 24.0
    frame_start: int = 0
    frame_end: int = 120
    format: str = "exr"
    video: str = "mp4"
    cam_position: tuple = (0.0, 0.0, 1.0)
    cam_look_at: tuple = (0.0, 0.0, 0.0)
    cam_up: tuple = (0.0, 1.0, 0.0)
    cam_fov_y_deg: float = 30.0
    solver: str = "dfsph"
    particle_radius: float = 0.0015
    rest_density: float = 1000.0
    viscosity: float = 1e-3
    surface_tension: float = 0.0728
    gravity: tuple = (0.0, -9.81, 0.0)
    initial_block_min: tuple = (0.4, 0.05, 0.4)
    initial_block_max: tuple = (0.6, 0.25, 0.6)


def load_scene(path: str) -> Scene:
    """Scene JSON schema "1.0" (loader.cpp:58-83); WsimError on failure."""
    s = _ffi.SceneC()
    check(_L().wsim_scene_load(str(path).encode(), C.byref(s)))
    kw = {}
    for f in fields(Scene):
        v = getattr(s, f.name)
        if isinstance(v, bytes):
            v = v.decode()
        elif not isinstance(v, (int, float)):
            v = tuple(float(x) for x in v)
        kw[f.name] = v
    return Scene(**kw)
