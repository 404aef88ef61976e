"""Synthetic scenes for the SPH step: seeded jittered particle lattices (numpy, host side).

The reference's own test fills positions with rand() (libCL/TEST/main.cpp:202-208); the workloads
here are portable and seeded instead:

* SplitMix64 stream per seed; a uniform float is (u64 >> 40) * 2^-24.
* Lattice order is x fastest, then y, then z; each particle draws three uniforms (x, y, z jitter),
  jitter = (2u - 1) * amplitude.
* State at t = 0: vel = 0, pos.w = vel.w = 0, force = (-0, -0, -9.8, -0), the value clGravity
  leaves between steps (libCL/phys/oclFluid3D.cl:294-298).
"""
from __future__ import annotations

import math
from dataclasses import dataclass, field

import numpy as np

_GOLDEN = np.uint64(0x9E3779B97F4A7C15)
_M1 = np.uint64(0xBF58476D1CE4E5B9)
_M2 = np.uint64(0x94D049BB133111EB)

GRAVITY_FORCE = np.array([-0.0, -0.0, -9.8, -0.0], np.float32)


def splitmix64(seed: int, count: int, offset: int = 0) -> np.ndarray:
    """Outputs offset+1 .. offset+count of the SplitMix64 generator started at `seed`."""
    with np.errstate(over="ignore"):
        i = np.arange(offset + 1, offset + count + 1, dtype=np.uint64)
        z = np.uint64(seed) + i * _GOLDEN
        z = (z ^ (z >> np.uint64(30))) * _M1
        z = (z ^ (z >> np.uint64(27))) * _M2
        return z ^ (z >> np.uint64(31))


def uniform(seed: int, count: int, offset: int = 0) -> np.ndarray:
    """float64 uniforms in [0, 1) with 24 random bits each."""
    return (splitmix64(seed, count, offset) >> np.uint64(40)).astype(np.float64) * (2.0 ** -24)


def normal(seed: int, count: int) -> np.ndarray:
    """Standard normals by Box-Muller over pairs of uniforms (float64)."""
    m = (count + 1) // 2
    u = uniform(seed, 2 * m)
    u1 = 1.0 - u[0::2]  # (0, 1]
    u2 = u[1::2]
    r = np.sqrt(-2.0 * np.log(u1))
    z = np.concatenate([r * np.cos(2 * np.pi * u2), r * np.sin(2 * np.pi * u2)])
    out = np.empty(2 * m)
    out[0::2] = z[:m]
    out[1::2] = z[m:]
    return out[:count]


def lattice(nx: int, ny: int, nz: int, spacing: float, origin, jitter: float, seed: int,
            count: int | None = None) -> np.ndarray:
    """float32 [count, 4] positions of an nx*ny*nz lattice (x fastest), truncated to `count`."""
    total = nx * ny * nz
    count = total if count is None else count
    iz, iy, ix = np.meshgrid(np.arange(nz), np.arange(ny), np.arange(nx), indexing="ij")
    grid = np.stack([ix.ravel(), iy.ravel(), iz.ravel()], axis=1)[:count].astype(np.float64)
    u = uniform(seed, 3 * count).reshape(count, 3)
    xyz = np.asarray(origin, np.float64)[None, :] + grid * spacing + (2.0 * u - 1.0) * jitter
    pos = np.zeros((count, 4), np.float32)
    pos[:, :3] = xyz.astype(np.float32)
    return pos


def sph_param_values(h: float = 0.02) -> dict:
    """Base Params of the seeded SPH workloads: h = cellSize, spacing 0.5h, radius 0.25h,
    dt 1e-4, rest density 1000, mass = rho0 * spacing^3, viscosity 6, stiffness 1, p0 0,
    velocity limit 500. The three kernel constants are derived from cellSize and viscosity
    exactly as clInitFluid does (libCL/phys/oclFluid3D.cl:145-147)."""
    spacing = 0.5 * h
    return dict(deltaTime=1e-4, particleRadius=0.25 * h, cellSize=h, mass=1000.0 * spacing ** 3, viscosity=6.0,
                pressure=0.0, density=1000.0, spacing=spacing, stiffness=1.0, velocitylimit=500.0)


@dataclass
class Scene:
    """One fluid scene: initial state, box and the base Params fields."""
    name: str
    pos: np.ndarray
    vel: np.ndarray
    force: np.ndarray
    box_min: tuple
    box_max: tuple
    params: dict = field(default_factory=sph_param_values)
    steps: int = 1

    @property
    def n(self) -> int:
        return self.pos.shape[0]


def _state(pos):
    n = pos.shape[0]
    return np.zeros((n, 4), np.float32), np.tile(GRAVITY_FORCE, (n, 1))


def dam_break(nx: int, ny: int, nz: int, box: float, seed: int, steps: int, h: float = 0.02, name: str = "dam") -> Scene:
    """A jittered block at 0.5h spacing, offset 0.25h from the origin corner of [0, box]^3."""
    pos = lattice(nx, ny, nz, 0.5 * h, (0.25 * h,) * 3, 0.01 * h, seed)
    vel, force = _state(pos)
    return Scene(name, pos, vel, force, (0.0, 0.0, 0.0, 0.0), (box, box, box, 0.0), sph_param_values(h), steps)


def config_small() -> Scene:
    """16,384 particles (32x16x32, seed 1) in the unit box, 1000 steps."""
    return dam_break(32, 16, 32, 1.0, 1, 1000, name="sph_16k_unit_box")


def config_1m() -> Scene:
    """1,048,576 particles (128x64x128, seed 2) in [0, 4]^3, 200 steps."""
    return dam_break(128, 64, 128, 4.0, 2, 200, name="sph_1m_box4")


def config_grad(n_side=(64, 32, 32), seed=6) -> tuple[Scene, np.ndarray]:
    """Differentiable case: a 64x32x32 block (65,536 particles), 200 steps; the target cloud is the
    initial block shifted 0.1 in x plus N(0, 0.01^2) noise per coordinate (seed 6)."""
    nx, ny, nz = n_side
    sc = dam_break(nx, ny, nz, 1.0, seed, 200, name=f"sph_grad_{nx * ny * nz}")
    noise = normal(seed, 3 * sc.n).reshape(sc.n, 3)
    target = sc.pos[:, :3].astype(np.float64) + np.array([0.1, 0.0, 0.0]) + 0.01 * noise
    return sc, target


def scene_counts(num: int = 256, seed: int = 5, lo: int = 2048, hi: int = 65536) -> np.ndarray:
    """Particle counts log-uniform in [lo, hi], rounded down to a multiple of 256."""
    u = uniform(seed, num)
    raw = np.exp(math.log(lo) + u * (math.log(hi) - math.log(lo)))
    return (256 * np.floor(raw / 256)).astype(np.int64)


def batched_scenes(num: int = 256, seed: int = 5, steps: int = 500, h: float = 0.02) -> list[Scene]:
    """`num` independent scenes, each a 32x32xceil(N/1024) lattice truncated to N points in its own
    unit box, jitter seed = seed*1000 + scene index."""
    out = []
    for s, n in enumerate(scene_counts(num, seed)):
        nz = -(-int(n) // 1024)
        pos = lattice(32, 32, nz, 0.5 * h, (0.25 * h,) * 3, 0.01 * h, seed * 1000 + s, count=int(n))
        vel, force = _state(pos)
        out.append(Scene(f"scene{s}", pos, vel, force, (0.0, 0.0, 0.0, 0.0), (1.0, 1.0, 1.0, 0.0),
                         sph_param_values(h), steps))
    return out


def sort_keys(n: int, seed: int, skew: bool = False, bits: int = 24) -> np.ndarray:
    """uint32 keys uniform over 2^bits cells; skew=True draws half the keys from a hot set holding
    10% of the cells (chosen with seed+1)."""
    cells = 1 << bits
    r = splitmix64(seed, n)
    keys = (r & np.uint64(cells - 1)).astype(np.uint32)
    if skew:
        hot = (splitmix64(seed + 1, cells // 10) & np.uint64(cells - 1)).astype(np.uint32)
        pick = (r >> np.uint64(40)) & np.uint64(1)
        which = ((r >> np.uint64(41)) % np.uint64(hot.shape[0])).astype(np.int64)
        keys = np.where(pick == 1, hot[which], keys)
    return keys
