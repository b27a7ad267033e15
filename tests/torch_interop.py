"""torch.autograd.Function over a checkpointed rollout, checked against the
host reference gradient.  Run as a script in a fresh interpreter: torch is
imported before the library loads, so both share torch's HIP runtime.

Scene: 8^3 particles of the inverse-problem block in the 0.6 m box, 40
substeps of 0.002 s, cell-density loss (16^3 CIC) against the field of the
seeded true velocities; gradients w.r.t. the initial velocities and the XSPH
viscosity, compared along seeded directions with the DUAL oracle's tangents
dotted with the loss gradient.  Prints one JSON line.
"""
import json
import sys
from pathlib import Path

import torch  # first: owns the HIP runtime

HERE = Path(__file__).resolve().parent
sys.path.insert(0, str(HERE.parent))
sys.path.insert(0, str(HERE))

import numpy as np  # noqa: E402

import water_simulation_amd as ws  # noqa: E402
from water_simulation_amd import adjoint as A  # noqa: E402
from oracle import oracle as O  # noqa: E402


def main():
    p = A.inverse_problem(side=8, substeps=40)
    cfg = dict(p.cfg, max_density_iters=10)
    n = len(p.x0)
    f = ws.ParticleStore(n)
    f.resize(n)
    f.set_positions(p.x0)
    b = ws.ParticleStore(len(p.xb))
    b.resize(len(p.xb))
    b.set_positions(p.xb)
    solver = ws.DFSPHSolver(f, b, ws.DFSPHConfig(**cfg))
    mass = solver.particle_mass()
    ro = A.Rollout(solver, checkpoint_every=20)
    ro.forward(p.x0, p.v_true, p.dts)
    target = A.density_field_np(f.get_positions(), mass)

    Fn = A.rollout_function()
    dev = torch.device("cuda:0")
    x0 = torch.tensor(p.x0, device=dev)
    v0 = torch.tensor(p.v_guess, device=dev, requires_grad=True)
    eps = torch.tensor(cfg["viscosity"], dtype=torch.float64, device=dev, requires_grad=True)
    xT, vT = Fn.apply(x0, v0, eps, ro, p.dts)
    loss = ((A.density_field_torch(xT, mass) - torch.tensor(target, device=dev)) ** 2).sum()
    loss.backward()
    gv0 = v0.grad.double().cpu().numpy()
    ge = float(eps.grad)
    xTn = xT.detach().cpu().numpy()
    lnp, gx_loss = A.density_loss_np(xTn, target, mass)

    rng = np.random.default_rng(1)
    pairs, same = [], True
    for _ in range(2):
        uv = rng.normal(size=p.x0.shape)
        d = O.Solver(O.config(**cfg), p.x0, p.v_guess, p.xb, O.DUAL)
        d.seed(dv=uv, d_viscosity=0.5)
        for dt in p.dts:
            d.step(float(dt))
        same &= bool(np.array_equal(d.get(O.FIELD_X), xTn.astype(np.float64)))
        want = float((gx_loss * d.tangent(O.FIELD_X)).sum())
        got = float((gv0 * uv).sum() + ge * 0.5)
        pairs.append((got, want))
    # a Python float viscosity takes no gradient: backward returns None for it
    v1 = torch.tensor(p.v_guess, device=dev, requires_grad=True)
    xf, _ = Fn.apply(x0, v1, float(cfg["viscosity"]), ro, p.dts[:5])
    xf.sum().backward()
    float_visc_ok = bool(v1.grad is not None and torch.isfinite(v1.grad).all())
    # a second forward pass on the same Rollout before backward: the first
    # graph's backward must refuse the replaced checkpoints
    v2 = torch.tensor(p.v_guess, device=dev, requires_grad=True)
    xa, _ = Fn.apply(x0, v2, eps, ro, p.dts[:5])
    Fn.apply(x0, v2, eps, ro, p.dts[:5])
    try:
        xa.sum().backward()
        stale_refused = False
    except RuntimeError as e:
        stale_refused = "another forward pass" in str(e)
    print(json.dumps({"same_trajectory": same, "loss": lnp,
                      "loss_rel_err": abs(lnp - float(loss)) / lnp, "directional": pairs,
                      "float_viscosity_backward": float_visc_ok,
                      "stale_tape_refused": stale_refused}))


if __name__ == "__main__":
    main()
