"""Differentiable DFSPH rollouts.

`Rollout` runs a plain DFSPHSolver forward with a state checkpoint every
`checkpoint_every` substeps and back-propagates d(loss)/d(final state) to
d(loss)/d(initial positions, initial velocities, XSPH viscosity) with the
library's adjoint kernels (binary64 adjoints, pressure iterations recomputed
from the checkpoints).  `RolloutFunction` wraps it as a torch.autograd.Function
so a loss written in torch drives it.  The derivative is that of the substep
map with its discrete structure frozen as the forward pass took it (neighbour
sets, iteration counts, clamp and max(0, .) masks).

The inverse-problem scene: a 32^3 block (r = 0.005 m) resting on the floor of
a 0.6 m Akinci box (centred in x and z), fixed dt = 0.002 s, initial velocity
a seeded smooth random field (a sum of low-frequency Fourier modes scaled to
0.5 m/s), loss = squared distance of the final 16^3 cell density field
(cloud-in-cell deposit of particle mass, divided by the cell volume) to a
target.  The scene uses the kappa/rho pressure scaling: with the reference's
kappa/rho^2 scaling and this dt the block sinks through the three wall layers
within 80 substeps (mean density error 170-390 kg/m^3 at the 100-iteration
cap), while kappa/rho keeps it above the walls at eta = 0.1 % in 2-4
iterations.
"""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass

import numpy as np

from . import _ffi
from ._ffi import check
from . import DeviceBuffer, memcpy_d2h, memcpy_h2d


def _L():
    return _ffi.lib()


class Rollout:
    """Checkpointed forward / backward over a plain solver (numpy or device pointers)."""

    def __init__(self, solver, checkpoint_every: int = 20):
        self.solver = solver
        self.n = solver._fluid.count()
        h = C.c_void_p()
        check(_L().wsim_rollout_create(solver.handle, int(checkpoint_every), C.byref(h)))
        self._h = h.value
        self._bufs = {}

    def close(self) -> None:
        if getattr(self, "_h", None):
            _L().wsim_rollout_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _buf(self, key: str, nbytes: int) -> DeviceBuffer:
        b = self._bufs.get(key)
        if b is None or b.nbytes < nbytes:
            b = DeviceBuffer(nbytes)
            self._bufs[key] = b
        return b

    # device-pointer interface (torch tensors pass data_ptr())
    def forward_ptr(self, x0: int, v0: int, dts) -> None:
        d = np.ascontiguousarray(dts, np.float32)
        check(_L().wsim_rollout_forward(self._h, x0, v0, d.ctypes.data_as(C.POINTER(C.c_float)),
                                        int(d.size)))
        # the checkpoints of this forward pass replace the previous ones
        self.generation = getattr(self, "generation", 0) + 1

    def backward_ptr(self, gx: int, gv: int, gx0: int, gv0: int) -> float:
        g = C.c_double()
        check(_L().wsim_rollout_backward(self._h, gx, gv, gx0, gv0, C.byref(g)))
        return g.value

    # numpy interface
    def forward(self, x0: np.ndarray, v0: np.ndarray, dts) -> None:
        n3 = 3 * self.n
        bx = self._buf("x0", 4 * n3)
        bv = self._buf("v0", 4 * n3)
        memcpy_h2d(bx.ptr, np.ascontiguousarray(x0, np.float32).reshape(-1))
        memcpy_h2d(bv.ptr, np.ascontiguousarray(v0, np.float32).reshape(-1))
        self.forward_ptr(bx.ptr, bv.ptr, dts)

    def backward(self, gx: np.ndarray, gv: np.ndarray):
        n3 = 3 * self.n
        bufs = [self._buf(k, 8 * n3) for k in ("gx", "gv", "gx0", "gv0")]
        memcpy_h2d(bufs[0].ptr, np.ascontiguousarray(gx, np.float64).reshape(-1))
        memcpy_h2d(bufs[1].ptr, np.ascontiguousarray(gv, np.float64).reshape(-1))
        gvisc = self.backward_ptr(bufs[0].ptr, bufs[1].ptr, bufs[2].ptr, bufs[3].ptr)
        gx0 = np.empty(n3, np.float64)
        gv0 = np.empty(n3, np.float64)
        memcpy_d2h(gx0, bufs[2].ptr)
        memcpy_d2h(gv0, bufs[3].ptr)
        return gx0.reshape(-1, 3), gv0.reshape(-1, 3), gvisc


def adjoint_step(solver, dt: float, gx_out: np.ndarray, gv_out: np.ndarray):
    """One recorded substep from the solver's current state and its adjoint
    (the store is left at the input state).  Returns (gx_in, gv_in, g_viscosity)."""
    n3 = 3 * solver._fluid.count()
    a = [DeviceBuffer(8 * n3) for _ in range(4)]
    memcpy_h2d(a[0].ptr, np.ascontiguousarray(gx_out, np.float64).reshape(-1))
    memcpy_h2d(a[1].ptr, np.ascontiguousarray(gv_out, np.float64).reshape(-1))
    g = C.c_double()
    check(_L().wsim_adjoint_step(solver.handle, float(dt), a[0].ptr, a[1].ptr, a[2].ptr, a[3].ptr,
                                 C.byref(g)))
    gx = np.empty(n3, np.float64)
    gv = np.empty(n3, np.float64)
    memcpy_d2h(gx, a[2].ptr)
    memcpy_d2h(gv, a[3].ptr)
    return gx.reshape(-1, 3), gv.reshape(-1, 3), g.value


def rollout_function():
    """torch.autograd.Function over a Rollout: (x0, v0, viscosity) -> (x_T, v_T).

    torch and libwsim must share one HIP runtime: both runtimes answer to the
    soname libamdhip64.so.7 and the first one loaded serves the process, and
    torch only runs on its own.  Import torch before this package loads its
    library (the first solver, store or device call)."""
    import torch

    if not torch.cuda.is_available():
        raise RuntimeError(
            "torch sees no HIP device: import torch before water_simulation_amd loads "
            "libwsim, so that both use torch's HIP runtime (libamdhip64.so.7)")

    class RolloutFunction(torch.autograd.Function):
        @staticmethod
        def forward(ctx, x0, v0, viscosity, rollout, dts):
            if x0.dtype != torch.float32 or v0.dtype != torch.float32 or not x0.is_cuda:
                raise TypeError("x0 and v0 must be float32 tensors on the GPU")
            x0c = x0.detach().contiguous()
            v0c = v0.detach().contiguous()
            rollout.solver.set_viscosity(float(viscosity))
            torch.cuda.synchronize()
            rollout.forward_ptr(x0c.data_ptr(), v0c.data_ptr(), dts)
            store = rollout.solver._fluid
            n = rollout.n
            xT = torch.empty((n, 3), dtype=torch.float32, device=x0.device)
            vT = torch.empty((n, 3), dtype=torch.float32, device=x0.device)
            _device_copy(xT.data_ptr(), store.positions(), 12 * n)
            _device_copy(vT.data_ptr(), store.velocities(), 12 * n)
            ctx.rollout = rollout
            ctx.generation = rollout.generation
            ctx.visc_is_tensor = torch.is_tensor(viscosity)
            ctx.vdtype = viscosity.dtype if ctx.visc_is_tensor else torch.float64
            return xT, vT

        @staticmethod
        def backward(ctx, gxT, gvT):
            rollout = ctx.rollout
            if getattr(rollout, "generation", None) != ctx.generation:
                raise RuntimeError(
                    "RolloutFunction.backward: the Rollout ran another forward pass after this "
                    "one; its checkpoints belong to that pass.  Use one Rollout per "
                    "outstanding graph, or call backward before the next apply().")
            n = rollout.n
            dev = gxT.device if gxT is not None else gvT.device
            gx = (gxT if gxT is not None else torch.zeros((n, 3), device=dev)).double().contiguous()
            gv = (gvT if gvT is not None else torch.zeros((n, 3), device=dev)).double().contiguous()
            gx0 = torch.empty((n, 3), dtype=torch.float64, device=dev)
            gv0 = torch.empty((n, 3), dtype=torch.float64, device=dev)
            torch.cuda.synchronize()
            gvisc = rollout.backward_ptr(gx.data_ptr(), gv.data_ptr(), gx0.data_ptr(),
                                         gv0.data_ptr())
            gv_visc = (torch.tensor(gvisc, dtype=ctx.vdtype, device=dev) if ctx.visc_is_tensor
                       else None)  # a Python float viscosity takes no gradient
            return gx0.float(), gv0.float(), gv_visc, None, None

    return RolloutFunction


def _device_copy(dst: int, src: int, nbytes: int) -> None:
    check(_L().wsim_memcpy_d2d(dst, src, nbytes, None))
    check(_L().wsim_device_synchronize())


# ---- the inverse-problem scene and its loss ---------------------------------
BOX = 0.6
RADIUS = 0.005
DT = 0.002
FIELD_CELLS = 16


def inverse_config(**over) -> dict:
    cfg = dict(rest_density=1000.0, particle_radius=RADIUS, smoothing_length=2 * RADIUS,
               viscosity=0.01, damping=0.5, max_density_iters=100, eta_density=1e-3,
               domain_min=(0.0, 0.0, 0.0), domain_max=(BOX, BOX, BOX), pressure_form=1)
    cfg.update(over)
    return cfg


def block_positions(side: int = 32) -> np.ndarray:
    """side^3 lattice at spacing 2r = 0.01 m, centred in x and z, its lowest
    layer one spacing above the top wall layer (y = 0.03 m)."""
    i = np.arange(side, dtype=np.float32)
    g = np.stack(np.meshgrid(i, i, i, indexing="ij"), -1).reshape(-1, 3)
    s = np.float32(2 * RADIUS)
    lo = np.float32(BOX / 2 - side * RADIUS + RADIUS)
    origin = np.array([lo, np.float32(0.03), lo], np.float32)
    return (origin + s * g).astype(np.float32)


def smooth_velocity(x: np.ndarray, seed: int, amplitude: float = 0.5, modes: int = 4) -> np.ndarray:
    """Sum of low-frequency Fourier modes with seeded coefficients, scaled so the
    largest speed equals `amplitude`."""
    from .scenes import Stream

    rng = Stream(seed)
    v = np.zeros((len(x), 3), np.float64)
    xd = x.astype(np.float64) / BOX
    for _ in range(modes):
        k = 1.0 + np.floor(3.0 * rng.uniform(3))
        phase = 2.0 * np.pi * rng.uniform(3)
        amp = rng.normal(3)
        arg = 2.0 * np.pi * (xd * k).sum(1)
        for a in range(3):
            v[:, a] += amp[a] * np.sin(arg + phase[a])
    speed = np.sqrt((v * v).sum(1)).max()
    return (amplitude * v / speed).astype(np.float32)


def cic_weights(x: np.ndarray, cells: int = FIELD_CELLS, box: float = BOX):
    """Cloud-in-cell: for each particle the 8 (cell index, weight) pairs of the
    cells whose centres surround it, clamped to the grid."""
    h = box / cells
    u = x.astype(np.float64) / h - 0.5
    i0 = np.floor(u).astype(np.int64)
    f = u - i0
    idx, w = [], []
    for dx in (0, 1):
        for dy in (0, 1):
            for dz in (0, 1):
                ix = np.clip(i0[:, 0] + dx, 0, cells - 1)
                iy = np.clip(i0[:, 1] + dy, 0, cells - 1)
                iz = np.clip(i0[:, 2] + dz, 0, cells - 1)
                wx = f[:, 0] if dx else 1 - f[:, 0]
                wy = f[:, 1] if dy else 1 - f[:, 1]
                wz = f[:, 2] if dz else 1 - f[:, 2]
                idx.append((ix * cells + iy) * cells + iz)
                w.append(wx * wy * wz)
    return np.stack(idx, 1), np.stack(w, 1)


def density_field_np(x: np.ndarray, mass: float, cells: int = FIELD_CELLS, box: float = BOX):
    idx, w = cic_weights(x, cells, box)
    d = np.zeros(cells ** 3, np.float64)
    np.add.at(d, idx.ravel(), (mass * w).ravel())
    return d / (box / cells) ** 3


def density_loss_np(x: np.ndarray, target: np.ndarray, mass: float):
    """L = sum_c (D_c - T_c)^2 and dL/dx (binary64)."""
    cells = round(len(target) ** (1 / 3))
    h = BOX / cells
    d = density_field_np(x, mass, cells)
    r = d - target
    loss = float((r * r).sum())
    # dL/dx: weights are products of per-axis linear factors
    u = x.astype(np.float64) / h - 0.5
    i0 = np.floor(u).astype(np.int64)
    f = u - i0
    g = np.zeros((len(x), 3), np.float64)
    scale = 2.0 * mass / h ** 3
    for dx in (0, 1):
        for dy in (0, 1):
            for dz in (0, 1):
                ix = np.clip(i0[:, 0] + dx, 0, cells - 1)
                iy = np.clip(i0[:, 1] + dy, 0, cells - 1)
                iz = np.clip(i0[:, 2] + dz, 0, cells - 1)
                rc = r[(ix * cells + iy) * cells + iz]
                wx = f[:, 0] if dx else 1 - f[:, 0]
                wy = f[:, 1] if dy else 1 - f[:, 1]
                wz = f[:, 2] if dz else 1 - f[:, 2]
                sx = (1.0 if dx else -1.0) / h
                sy = (1.0 if dy else -1.0) / h
                sz = (1.0 if dz else -1.0) / h
                g[:, 0] += scale * rc * sx * wy * wz
                g[:, 1] += scale * rc * wx * sy * wz
                g[:, 2] += scale * rc * wx * wy * sz
    return loss, g


def density_field_torch(x, mass: float, cells: int = FIELD_CELLS, box: float = BOX):
    """Differentiable CIC deposit in torch (same definition as density_field_np)."""
    import torch

    h = box / cells
    u = x.double() / h - 0.5
    i0 = torch.floor(u).detach()
    f = u - i0
    i0 = i0.long()
    d = torch.zeros(cells ** 3, dtype=torch.float64, device=x.device)
    for dx in (0, 1):
        for dy in (0, 1):
            for dz in (0, 1):
                ix = (i0[:, 0] + dx).clamp(0, cells - 1)
                iy = (i0[:, 1] + dy).clamp(0, cells - 1)
                iz = (i0[:, 2] + dz).clamp(0, cells - 1)
                wx = f[:, 0] if dx else 1 - f[:, 0]
                wy = f[:, 1] if dy else 1 - f[:, 1]
                wz = f[:, 2] if dz else 1 - f[:, 2]
                d = d.index_add(0, (ix * cells + iy) * cells + iz, mass * wx * wy * wz)
    return d / h ** 3


@dataclass
class InverseProblem:
    x0: np.ndarray
    v_true: np.ndarray
    v_guess: np.ndarray
    xb: np.ndarray
    cfg: dict
    substeps: int
    dts: np.ndarray


def inverse_problem(side: int = 32, substeps: int = 200, seed_true: int = 0x1A7E,
                    seed_guess: int = 0x5EED) -> InverseProblem:
    from . import sample_aabb_boundary

    x0 = block_positions(side)
    return InverseProblem(
        x0=x0, v_true=smooth_velocity(x0, seed_true), v_guess=smooth_velocity(x0, seed_guess),
        xb=sample_aabb_boundary((0, 0, 0), (BOX,) * 3, 2 * RADIUS, 3), cfg=inverse_config(),
        substeps=substeps, dts=np.full(substeps, DT, np.float32))
