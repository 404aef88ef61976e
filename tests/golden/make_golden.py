"""Writes the golden fixtures under tests/golden/ (run: python tests/golden/make_golden.py).

The reference cannot execute here (its kernels need an OpenCL device; the image has none) and ships
no vectors, so the fixtures are derived independently of both the oracle and the HIP library:

* hash_kat.json: cell keys of libCL/phys/oclFluid3D.cl:44-52 evaluated with Python integers
  (exact mod-2^32 products), and grid coordinates (cl:34-42) evaluated with numpy's IEEE float32
  division and floor, at cell boundaries and negative coordinates.
* step_*.npz: one SPH step (libCL/phys/oclFluid3D.cpp:293-342) evaluated by an O(N^2) all-pairs
  float64 brute force over the kernel formulas of libCL/phys/oclFluid3D.cl:150-361. Each pair (i, j)
  is weighted by its visit multiplicity, the number of the 27 neighbour offsets d of i whose bucket
  hash(g_i + d) equals j's key: the reference does not de-duplicate buckets, and near the origin
  the XOR hash maps offsets such as (-1,-1,-1) and (1,1,1) to one bucket, so such neighbours count
  twice. With fresh bucket tables (no stale ranges) this weighting equals the 27-bucket traversal.
"""
from __future__ import annotations

import json
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))

from libcl_amd import scenes  # noqa: E402

BUCKETS = 262144
P1, P2, P3 = 73856093, 19349663, 83492791
MASK32 = (1 << 32) - 1


def hash_int(gx: int, gy: int, gz: int) -> int:
    """uint wraparound products of int grid coordinates, XOR, mod 2^18 (Python big ints)."""
    n = ((P1 * (gx & MASK32)) & MASK32) ^ ((P2 * (gy & MASK32)) & MASK32) ^ ((P3 * (gz & MASK32)) & MASK32)
    return n % BUCKETS


def hash_np(g: np.ndarray) -> np.ndarray:
    g = g.astype(np.int64) & MASK32
    n = ((P1 * g[:, 0]) & MASK32) ^ ((P2 * g[:, 1]) & MASK32) ^ ((P3 * g[:, 2]) & MASK32)
    return (n % BUCKETS).astype(np.uint32)


def hash_kat() -> dict:
    rng = np.random.default_rng(12345)
    cases = [(0, 0, 0), (1, 0, 0), (0, 1, 0), (0, 0, 1), (-1, 0, 0), (0, -1, 0), (0, 0, -1), (-1, -1, -1),
             (63, 63, 63), (64, 64, 64), (199, 0, 199), (2 ** 20, -2 ** 20, 12345), (-2 ** 31, 2 ** 31 - 1, 7)]
    cases += [tuple(int(v) for v in rng.integers(-5000, 5000, 3)) for _ in range(64)]
    keys = [[gx, gy, gz, hash_int(gx, gy, gz)] for gx, gy, gz in cases]
    h = np.float32(0.02)
    xs = [0.0, -0.0, 0.02, -0.02, 0.04, 0.0199999996, 0.0200000014, -1e-9, 1e-9, 0.5, 0.98, 1.0, -1.0, 3.99, 2.56,
          np.nextafter(np.float32(0.06), np.float32(0)), np.float32(0.06), np.nextafter(np.float32(0.06), np.float32(1))]
    xs = [np.float32(x) for x in xs] + [np.float32(v) for v in rng.uniform(-2, 5, 40)]
    grid = [[float(x), float(h), int(np.floor(np.float32(x) / h))] for x in xs]
    return {"keys": keys, "grid": grid, "cell_size": float(h)}


def brute_step(pos32, vel32, force32, prm: dict, box_min, box_max, gravity=(0, 0, 1, 0)):
    """One SPH step by all-pairs float64 evaluation of the reference formulas."""
    f64 = np.float64
    h32 = np.float32(prm["cellSize"])
    g = np.floor(pos32[:, :3] / h32).astype(np.int64)  # float32 IEEE division, as the kernel
    key = hash_np(g)
    perm = np.argsort(key, kind="stable").astype(np.uint32)
    skey = key[perm]
    uk, first = np.unique(skey, return_index=True)
    last = np.searchsorted(skey, uk, side="right")

    h = f64(h32)
    m = f64(np.float32(prm["mass"]))
    kc = f64(np.float32(prm["kernelConstant"]))
    kp = f64(np.float32(prm["pressureConstant"]))
    kv = f64(np.float32(prm["viscosityConstant"]))
    rho0 = f64(np.float32(prm["density"]))
    p0 = f64(np.float32(prm["pressure"]))
    kk = f64(np.float32(prm["stiffness"]))
    dt = f64(np.float32(prm["deltaTime"]))
    vmax = f64(np.float32(prm["velocitylimit"]))
    rad = f64(np.float32(prm["particleRadius"]))

    x = pos32[perm, :3].astype(f64)
    v = vel32[perm, :3].astype(f64)
    n = x.shape[0]
    # visit multiplicity of j by i: number of neighbour offsets of i whose bucket is j's key
    offs = np.array([(ox, oy, oz) for oz in (-1, 0, 1) for oy in (-1, 0, 1) for ox in (-1, 0, 1)], np.int64)
    gs = g[perm]
    nb_keys = np.stack([hash_np(gs + o) for o in offs], axis=1)  # [n, 27]
    mult = (nb_keys[:, :, None] == skey[None, None, :]).sum(1).astype(f64)  # [n, n]
    d = x[:, None, :] - x[None, :, :]
    r = np.sqrt((d * d).sum(-1))
    nbr = (r < h) & ~np.eye(n, dtype=bool) & (mult > 0)
    mult = np.where(nbr, mult, 0.0)
    count = mult.sum(1)
    w = mult * np.where(nbr, (h * h - r * r) ** 3, 0.0)
    acc = w.sum(1)
    rho = np.where(count > 2, np.maximum(1.0, m * kc * acc), rho0)
    prs = np.where(count > 2, p0 + kk * (rho - rho0), p0)

    rr = np.where(nbr, r, 1.0)
    hr = np.where(nbr, h - r, 0.0)
    rho2 = rho[:, None] * rho[None, :]
    q = (prs[:, None] + prs[None, :]) / rho2
    pres = ((mult * q * hr * hr / rr)[:, :, None] * d).sum(1)
    visc = ((mult * hr / rho2)[:, :, None] * (v[None, :, :] - v[:, None, :])).sum(1)
    f_in = force32[perm].astype(f64)
    f_sph = f_in.copy()
    f_sph[:, :3] += (kp * pres + kv * visc) * m
    f_sph[:, 3] = 0.0

    # clClipBox walls y-, y+, z-, z+, x-, x+ (libCL/phys/oclFluid3D.cl:315-361)
    f_clip = f_sph.copy()
    vel4 = np.concatenate([v, prs[:, None]], axis=1)
    mn = np.asarray(box_min, f64)
    mx = np.asarray(box_max, f64)
    walls = [(1, +1, lambda: rad - (x[:, 1] - mn[1])), (1, -1, lambda: rad - (mx[1] - x[:, 1])),
             (2, +1, lambda: rad - (x[:, 2] - mn[2])), (2, -1, lambda: rad - (mx[2] - x[:, 2])),
             (0, +1, lambda: rad - (x[:, 0] - mn[0])), (0, -1, lambda: rad - (mx[0] - x[:, 0]))]
    for axis, sign, diff_fn in walls:
        diff = diff_fn()
        hit = diff > 1e-5
        s = 2000.0 * diff - 256.0 * sign * vel4[:, axis]
        f_clip[:, axis] += np.where(hit, sign * s, 0.0)

    # back to the user (unsorted) order
    inv = np.empty(n, np.int64)
    inv[perm] = np.arange(n)
    f_user = f_clip[inv]
    speed = np.sqrt((f_user * f_user).sum(1))
    scale = np.where(speed > vmax, vmax / np.where(speed > 0, speed, 1.0), 1.0)
    f_lim = f_user * scale[:, None]
    vel_new = vel32.astype(f64) + f_lim * dt
    pos_new = pos32.astype(f64) + vel_new * dt
    force_new = np.tile(-9.8 * np.asarray(gravity, f64), (n, 1))
    return dict(key=skey, perm=perm, bucket=uk.astype(np.uint32), start=first.astype(np.uint32),
                end=last.astype(np.uint32), rho=rho, prs=prs, count=count.astype(np.int32),
                max_mult=np.int32(mult.max()),
                force_sph=f_sph[inv], force_clip=f_user, vel_new=vel_new, pos_new=pos_new, force_new=force_new)


def params_f32(prm: dict) -> np.ndarray:
    """13 float32 fields in the libCL/phys/oclFluid3D.h:48-63 order, kernel constants derived the way
    clInitFluid does (libCL/phys/oclFluid3D.cl:124-128,145-147) but computed here independently."""
    pif = np.float32(np.pi)
    h = np.float32(prm["cellSize"])
    h9 = np.float64(np.float32(np.float64(h) ** 9))
    h6 = np.float32(np.float64(h) ** 6)
    kc = np.float32(315.0 / (64.0 * np.float64(pif) * h9))
    visc_c = -45.0 / np.float64(np.float32(pif * h6))
    kv = np.float32(-np.float64(np.float32(prm["viscosity"])) * visc_c)
    kp = np.float32(-visc_c)
    order = ["deltaTime", "particleRadius", "cellSize", "mass", "viscosity", "pressure", "density", "spacing",
             "stiffness", "viscosityConstant", "pressureConstant", "kernelConstant", "velocitylimit"]
    full = dict(prm, viscosityConstant=kv, pressureConstant=kp, kernelConstant=kc)
    return np.array([np.float32(full[k]) for k in order], np.float32)


PARAM_ORDER = ["deltaTime", "particleRadius", "cellSize", "mass", "viscosity", "pressure", "density", "spacing",
               "stiffness", "viscosityConstant", "pressureConstant", "kernelConstant", "velocitylimit"]


def write_step_case(name, pos, vel, force, prm, box_min, box_max):
    p32 = params_f32(prm)
    pd = {k: float(v) for k, v in zip(PARAM_ORDER, p32)}
    out = brute_step(pos, vel, force, pd, box_min, box_max)
    np.savez_compressed(os.path.join(HERE, f"step_{name}.npz"), pos=pos, vel=vel, force=force, params=p32,
                        box_min=np.asarray(box_min, np.float32), box_max=np.asarray(box_max, np.float32), **out)
    return out


def main():
    with open(os.path.join(HERE, "hash_kat.json"), "w") as fh:
        json.dump(hash_kat(), fh, indent=0)

    h = 0.02
    prm = scenes.sph_param_values(h)
    # 1) jittered 8x8x8 block at 0.5h in the unit-box corner: walls active, near-uniform density
    pos = scenes.lattice(8, 8, 8, 0.5 * h, (0.25 * h,) * 3, 0.01 * h, seed=11)
    vel = np.zeros_like(pos)
    force = np.tile(scenes.GRAVITY_FORCE, (pos.shape[0], 1))
    write_step_case("lattice512", pos, vel, force, prm, (0, 0, 0, 0), (1, 1, 1, 0))

    # 2) uniform random cloud with random velocities: uneven densities, sparse (count <= 2) particles,
    #    viscosity active, negative coordinates, a tight velocity limit that clamps large forces
    u = scenes.uniform(12, 512 * 6).reshape(512, 6)
    pos = np.zeros((512, 4), np.float32)
    pos[:, :3] = (u[:, :3] * 0.16 - 0.03).astype(np.float32)
    vel = np.zeros((512, 4), np.float32)
    vel[:, :3] = ((u[:, 3:] - 0.5) * 2.0).astype(np.float32)
    force = np.tile(scenes.GRAVITY_FORCE, (512, 1))
    prm2 = dict(prm, velocitylimit=40.0, stiffness=3.0, pressure=5.0)
    write_step_case("random512", pos, vel, force, prm2, (-0.02, -0.02, -0.02, 0), (0.12, 0.12, 0.12, 0))

    # 3) compressed lattice (0.35h spacing) with a custom initial force field: large positive pressure
    pos = scenes.lattice(8, 8, 8, 0.35 * h, (0.3, 0.3, 0.3), 0.02 * h, seed=13)
    vel = np.zeros_like(pos)
    uf = scenes.uniform(14, 512 * 3).reshape(512, 3)
    force = np.zeros((512, 4), np.float32)
    force[:, :3] = ((uf - 0.5) * 20.0).astype(np.float32)
    write_step_case("compressed512", pos, vel, force, dict(prm, stiffness=50.0), (0, 0, 0, 0), (1, 1, 1, 0))


if __name__ == "__main__":
    main()
