"""Seeded synthetic scenes for the batched solver: many independent box
scenes of uneven particle counts.

Random numbers come from SplitMix64 (Steele, Lea and Flood, "Fast splittable
pseudorandom number generators", OOPSLA 2014): output k of a stream seeded
with s is mix(s + (k + 1) * 0x9E3779B97F4A7C15), so a whole stream is one
vectorised numpy expression.  Uniform doubles take the top 53 bits; normals
use Box-Muller.  The same seed gives the same scene on every machine.

The uneven mix: every scene is a 0.6 m box whose walls are Akinci particles
sampled at 0.01 m in 3 layers (one boundary store shared by all scenes);
the fluid is an axis-aligned lattice block at spacing 2r = 0.01 m with a
particle count drawn log-uniformly from [n_min, n_max], a random aspect ratio
and a random position inside the walls; velocities are i.i.d. N(0, sigma)
per component and the XSPH viscosity is uniform in [0.005, 0.05].
"""
from __future__ import annotations

from dataclasses import dataclass

import numpy as np

_GOLDEN = np.uint64(0x9E3779B97F4A7C15)
_M1 = np.uint64(0xBF58476D1CE4E5B9)
_M2 = np.uint64(0x94D049BB133111EB)
_MASK = (1 << 64) - 1


def splitmix64(seed: int, count: int, first: int = 0) -> np.ndarray:
    """Outputs first .. first+count-1 of the SplitMix64 stream of `seed`."""
    k = np.arange(first + 1, first + count + 1, dtype=np.uint64)
    with np.errstate(over="ignore"):
        z = np.uint64(seed & _MASK) + k * _GOLDEN
        z = (z ^ (z >> np.uint64(30))) * _M1
        z = (z ^ (z >> np.uint64(27))) * _M2
    return z ^ (z >> np.uint64(31))


class Stream:
    """Sequential draws from one SplitMix64 stream."""

    def __init__(self, seed: int):
        self.seed = seed
        self.pos = 0

    def u64(self, n: int) -> np.ndarray:
        out = splitmix64(self.seed, n, self.pos)
        self.pos += n
        return out

    def uniform(self, n: int = 1) -> np.ndarray:
        """U[0, 1) doubles from the top 53 bits."""
        return (self.u64(n) >> np.uint64(11)).astype(np.float64) * (1.0 / 9007199254740992.0)

    def normal(self, n: int = 1) -> np.ndarray:
        """N(0, 1) by Box-Muller, two uniforms per normal (cosine branch)."""
        u = self.uniform(2 * n).reshape(n, 2)
        r = np.sqrt(-2.0 * np.log1p(-u[:, 0]))  # 1 - u in (0, 1]
        return r * np.cos(2.0 * np.pi * u[:, 1])

    def integer(self, lo: int, hi: int) -> int:
        """Uniform integer in [lo, hi]."""
        return lo + int(self.uniform(1)[0] * (hi - lo + 1))


@dataclass
class BoxScene:
    positions: np.ndarray   # (n, 3) float32
    velocities: np.ndarray  # (n, 3) float32
    viscosity: float
    dims: tuple             # lattice particles per axis


BOX = 0.6          # box edge, m
RADIUS = 0.005     # particle radius, m (spacing 0.01 m)
WALL_SPACING = 0.01
WALL_LAYERS = 3
SEED_BASE = 0x5EED0000


def box_config(**over) -> dict:
    """Solver settings shared by all scenes of the mix (DFSPHConfig fields)."""
    cfg = dict(rest_density=1000.0, particle_radius=RADIUS, smoothing_length=2 * RADIUS,
               viscosity=0.01, damping=0.5, max_density_iters=100, max_divergence_iters=100,
               eta_density=1e-3, eta_divergence=1e-3, domain_min=(0.0, 0.0, 0.0),
               domain_max=(BOX, BOX, BOX))
    cfg.update(over)
    return cfg


def box_scene(index: int, n_min: int = 2000, n_max: int = 60000, sigma: float = 0.5,
              seed_base: int = SEED_BASE) -> BoxScene:
    rng = Stream(seed_base + index)
    spacing = 2 * RADIUS
    # lattice slots inside the three wall layers: 0.03 .. 0.57
    first = 0.03
    slots = int(round((BOX - 2 * first) / spacing)) + 1   # 55
    target = float(np.exp(np.log(n_min) + (np.log(n_max) - np.log(n_min)) * rng.uniform(1)[0]))
    aspect = np.exp(rng.uniform(3) - 0.5)
    base = (target / aspect.prod()) ** (1.0 / 3.0)
    nx = int(np.clip(round(base * aspect[0]), 2, slots))
    ny = int(np.clip(round(base * aspect[1]), 2, slots))
    nz = int(np.clip(round(target / (nx * ny)), 2, slots))
    o = [rng.integer(0, slots - d) for d in (nx, ny, nz)]
    ix, iy, iz = np.meshgrid(np.arange(nx), np.arange(ny), np.arange(nz), indexing="ij")
    idx = np.stack([ix, iy, iz], -1).reshape(-1, 3) + np.array(o)
    pos = (np.float32(first) + np.float32(spacing) * idx.astype(np.float32)).astype(np.float32)
    n = len(pos)
    vel = (sigma * rng.normal(3 * n)).astype(np.float32).reshape(n, 3)
    visc = float(np.float32(0.005 + 0.045 * rng.uniform(1)[0]))
    return BoxScene(pos, vel, visc, (nx, ny, nz))


def uneven_mix(count: int = 64, **kw) -> list:
    """The uneven batched mix: `count` box scenes (seeds SEED_BASE + index)."""
    return [box_scene(i, **kw) for i in range(count)]


def initial_max_speed(v: np.ndarray) -> float:
    """max |v| in binary32, the value a solver reports after a step."""
    if len(v) == 0:
        return 0.0
    v = v.astype(np.float32)
    return float(np.sqrt((v * v).sum(1, dtype=np.float32)).max())
