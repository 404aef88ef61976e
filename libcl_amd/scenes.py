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


# ---- triangle meshes for the BVH and the particle-vs-mesh collision ---------------------------------

@dataclass
class Mesh:
    """A triangle mesh as the BVH program takes it: float4 vertices (w = 1) and 3 uint32 vertex
    indices per triangle."""
    name: str
    vertex: np.ndarray
    index: np.ndarray

    @property
    def triangles(self) -> int:
        return self.index.shape[0] // 3


def floor_mesh(nx: int, ny: int, lo=(0.0, 0.0), hi=(1.0, 1.0), z: float = 0.0, jitter: float = 0.0, seed: int = 0,
               name: str = "floor", slope=(0.0, 0.0)) -> Mesh:
    """An nx x ny grid of quads over [lo, hi] at height z + slope . (x, y) (each vertex moved by
    jitter * (2u - 1) in z, seed), two triangles per quad facing +z; vertices x fastest. (A triangle
    in a plane z = const has a box of zero thickness, which the reference's box test never enters.)"""
    xs = np.linspace(lo[0], hi[0], nx + 1)
    ys = np.linspace(lo[1], hi[1], ny + 1)
    gy, gx = np.meshgrid(ys, xs, indexing="ij")
    nv = (nx + 1) * (ny + 1)
    vertex = np.ones((nv, 4), np.float32)
    vertex[:, 0] = gx.ravel()
    vertex[:, 1] = gy.ravel()
    vertex[:, 2] = (z + slope[0] * gx.ravel() + slope[1] * gy.ravel() +
                    (jitter * (2.0 * uniform(seed, nv) - 1.0) if jitter else 0.0))
    j, i = np.meshgrid(np.arange(ny), np.arange(nx), indexing="ij")
    v00 = (j * (nx + 1) + i).ravel()
    v10, v01, v11 = v00 + 1, v00 + nx + 1, v00 + nx + 2
    index = np.stack([v00, v10, v01, v10, v11, v01], axis=1).reshape(-1).astype(np.uint32)
    return Mesh(name, vertex, index)


def random_triangles(nv: int = 1000, seed: int = 1, name: str = "random") -> Mesh:
    """The shape of libCL/TEST/main.cpp:349-383: nv vertices uniform in [-0.5, 0.5)^3 with w = 1 and
    indices 0..nv-1, so nv // 3 large, overlapping triangles (seeded here instead of rand())."""
    vertex = np.ones((nv, 4), np.float32)
    vertex[:, :3] = (uniform(seed, 3 * nv).reshape(nv, 3) - 0.5).astype(np.float32)
    return Mesh(name, vertex, np.arange(nv, dtype=np.uint32))


def sphere_mesh(n_lat: int, n_lon: int, center=(0.0, 0.0, 0.0), radius: float = 1.0, name: str = "sphere") -> Mesh:
    """A UV sphere with outward-facing triangles (two poles, n_lat - 1 rings of n_lon vertices)."""
    c = np.asarray(center, np.float64)
    rings = []
    for a in range(1, n_lat):
        th = math.pi * a / n_lat
        for b in range(n_lon):
            ph = 2.0 * math.pi * b / n_lon
            rings.append((math.sin(th) * math.cos(ph), math.sin(th) * math.sin(ph), math.cos(th)))
    pts = np.array([(0.0, 0.0, 1.0)] + rings + [(0.0, 0.0, -1.0)])
    south = len(pts) - 1
    tris = []
    ring = lambda a, b: 1 + (a - 1) * n_lon + (b % n_lon)  # noqa: E731
    for b in range(n_lon):
        tris.append((0, ring(1, b), ring(1, b + 1)))
        tris.append((south, ring(n_lat - 1, b + 1), ring(n_lat - 1, b)))
    for a in range(1, n_lat - 1):
        for b in range(n_lon):
            tris.append((ring(a, b), ring(a + 1, b), ring(a + 1, b + 1)))
            tris.append((ring(a, b), ring(a + 1, b + 1), ring(a, b + 1)))
    tri = np.array(tris, np.int64)
    p = pts * radius + c
    nrm = np.cross(p[tri[:, 1]] - p[tri[:, 0]], p[tri[:, 2]] - p[tri[:, 0]])
    out = (nrm * (p[tri].mean(1) - c)).sum(1) < 0
    tri[out] = tri[out][:, ::-1]
    vertex = np.ones((len(p), 4), np.float32)
    vertex[:, :3] = p.astype(np.float32)
    return Mesh(name, vertex, tri.reshape(-1).astype(np.uint32))


def config_collide_1m() -> tuple[Scene, Mesh]:
    """Particle-vs-mesh workload: the 1,048,576-particle block of config_1m falling at 1.5 units/s
    (vx, vy ~ N(0, 0.02^2), seed 3) onto a 96 x 48-quad floor (9,216 triangles) across the block at
    z = 0.6 with 0.002 height jitter (seed 4). The collision runs with clInitFluid's Params
    (dt = 0.005): its parallel-ray cut |n . v dt| < 1e-6 uses the unnormalised triangle normal, so a
    contact needs |2 area x v dt| >= 1e-6, which these triangles (legs ~0.013) and speed meet."""
    sc = config_1m()
    noise = normal(3, 2 * sc.n).reshape(sc.n, 2)
    sc.vel[:, 0] = (0.02 * noise[:, 0]).astype(np.float32)
    sc.vel[:, 1] = (0.02 * noise[:, 1]).astype(np.float32)
    sc.vel[:, 2] = -1.5
    sc.name = "collide_1m_floor9k"
    mesh = floor_mesh(96, 48, (0.0, 0.0), (1.28, 0.64), 0.6, 0.002, 4, name="floor9k")
    return sc, mesh


# ---- 2D states for oclFluid3Dnext ----------------------------------------------------------------------

def next_block(nx: int, ny: int, spacing: float = 0.7, origin=(20.0, 20.0), jitter: float = 0.1, seed: int = 0,
               vel=(0.0, 0.0)) -> np.ndarray:
    """float32 [nx * ny, 4] states (x, y, vx, vy) of a jittered nx x ny block (x fastest) inside the
    reference's 1024 x 768 domain (clamp of libCL/phys/oclFluid3Dnext.cl:258)."""
    iy, ix = np.meshgrid(np.arange(ny), np.arange(nx), indexing="ij")
    n = nx * ny
    u = uniform(seed, 2 * n).reshape(n, 2)
    st = np.zeros((n, 4), np.float32)
    st[:, 0] = (origin[0] + ix.ravel() * spacing + (2 * u[:, 0] - 1) * jitter).astype(np.float32)
    st[:, 1] = (origin[1] + iy.ravel() * spacing + (2 * u[:, 1] - 1) * jitter).astype(np.float32)
    st[:, 2] = vel[0]
    st[:, 3] = vel[1]
    return st


def config_next_1m() -> np.ndarray:
    """1,048,576 particles: a 1024 x 1024 block at spacing 0.7 from (20, 20), jitter 0.1 (seed 8), at rest."""
    return next_block(1024, 1024, 0.7, (20.0, 20.0), 0.1, 8)
