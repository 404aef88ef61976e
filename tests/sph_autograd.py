"""Test-harness autograd glue for the SPH adjoint, and an independent FP64 CPU reference.

SPHRollout wraps libmfl's forward rollout + hand-written adjoint (mfl_adjoint_*) in a
torch.autograd.Function, so a loss written in torch differentiates through the GPU rollout.
cpu_rollout_fp64 restates the same step (libCL/phys/oclFluid3D.cl:150-361) in float64 torch ops on the
CPU, with pair multiplicities from the cell hash; torch's own autograd of it is the FP64 CPU adjoint
the hand-written one is checked against. It is exact for fresh bucket tables (a full bucket clear).
"""
from __future__ import annotations

import numpy as np
import torch

import libcl_amd as L
from libcl_amd import scenes

P1, P2, P3 = 73856093, 19349663, 83492791


class Simulation:
    """One program + adjoint for a scene; parameters are (viscosity, stiffness, rest density)."""

    def __init__(self, sc: scenes.Scene, steps: int, checkpoint_every: int = 20, list_cap: int = 192,
                 clear: int = 65536):
        self.sc = sc
        self.base = L.params_from(sc.params)
        self.prog = L.oclFluid3D()
        assert self.prog.compile() == 1, L.last_error()
        assert self.prog.setParticleCount(sc.n) == 1
        self.prog.setBox(sc.box_min, sc.box_max)
        self.prog.setBucketClear(clear)
        self.adj = L.Adjoint(self.prog, steps, checkpoint_every, list_cap)
        self.steps = steps
        self.tables0 = (self.prog.bfCellStart.read(np.uint32), self.prog.bfCellEnd.read(np.uint32))

    def params(self, viscosity, stiffness, density) -> L.Params:
        vals = self.base.as_dict()
        vals.update(viscosity=float(viscosity), stiffness=float(stiffness), density=float(density))
        for k in ("viscosityConstant", "pressureConstant", "kernelConstant"):
            vals.pop(k)
        return L.params_from(vals)

    def run(self, v0: np.ndarray, viscosity, stiffness, density) -> np.ndarray:
        """Forward rollout from the scene's initial positions/forces with velocities v0 [N, 3]."""
        self.prog.setParameters(self.params(viscosity, stiffness, density))
        vel = np.zeros((self.sc.n, 4), np.float32)
        vel[:, :3] = v0
        assert self.prog.upload(self.sc.pos, vel, self.sc.force) == 1
        assert self.prog.bfCellStart.write(self.tables0[0]) == 1 and self.prog.bfCellEnd.write(self.tables0[1]) == 1
        assert self.adj.forward() == 1, L.last_error()
        L.load().mfl_device_finish()
        return self.prog.state()["pos"]


class SPHRollout(torch.autograd.Function):
    """x_T = rollout(v0, viscosity, stiffness, rest density); backward calls mfl_adjoint_backward."""

    @staticmethod
    def forward(ctx, v0, viscosity, stiffness, density, sim: Simulation):
        pos = sim.run(v0.detach().cpu().numpy().astype(np.float32), viscosity.item(), stiffness.item(),
                      density.item())
        ctx.sim = sim
        return torch.from_numpy(pos[:, :3].copy()).to(v0.device)

    @staticmethod
    def backward(ctx, g_xt):
        sim = ctx.sim
        gb = torch.zeros(g_xt.shape[0], 4, dtype=torch.float32, device="cuda")
        gb[:, :3] = g_xt.float().to("cuda")
        assert sim.adj.backward(gb.contiguous()) == 1, L.last_error()
        gv = torch.from_numpy(sim.adj.velocity_grad()[:, :3].copy()).to(g_xt.device, g_xt.dtype)
        pg = sim.adj.param_grads()
        t = lambda x: torch.tensor(x, dtype=torch.float64)  # noqa: E731
        return gv, t(pg["viscosity"]), t(pg["stiffness"]), t(pg["density"]), None


# ---------------------------------------------------------------------------- FP64 CPU reference

def _hash(g: np.ndarray) -> np.ndarray:
    m = (1 << 32) - 1
    g = g.astype(np.int64) & m
    return ((((P1 * g[:, 0]) & m) ^ ((P2 * g[:, 1]) & m) ^ ((P3 * g[:, 2]) & m)) % 262144).astype(np.int64)


_OFFS = np.array([(x, y, z) for z in (-1, 0, 1) for y in (-1, 0, 1) for x in (-1, 0, 1)], np.int64)


def _pairs(x32: np.ndarray, h: float):
    """(i, j, multiplicity) of every accepted visit for fresh tables: j within h of i, weighted by
    the number of i's 27 neighbour buckets equal to j's key. All pairs for small n, a cell list above."""
    if x32.shape[0] > 2048:
        return _pairs_cells(x32, h)
    return _pairs_dense(x32, h)


def _candidates(x32: np.ndarray, radius: float):
    """Pairs (i, j), i != j, at distance < radius, by a cell list of cell size `radius`: such a j lies in
    one of the 27 cells around i's cell."""
    g = np.floor(x32.astype(np.float64) / radius).astype(np.int64)
    n = x32.shape[0]
    lo = g.min(0) - 1
    dims = g.max(0) - lo + 2
    lin = ((g[:, 2] - lo[2]) * dims[1] + (g[:, 1] - lo[1])) * dims[0] + (g[:, 0] - lo[0])
    order = np.argsort(lin, kind="stable")
    slin = lin[order]
    x64 = x32.astype(np.float64)
    ii_all, jj_all = [], []
    for o in _OFFS:
        t = lin + (o[2] * dims[1] + o[1]) * dims[0] + o[0]
        a = np.searchsorted(slin, t, side="left")
        b = np.searchsorted(slin, t, side="right")
        cnt = b - a
        ii = np.repeat(np.arange(n), cnt)
        first = np.repeat(a - np.concatenate([[0], np.cumsum(cnt)[:-1]]), cnt)
        jj = order[first + np.arange(ii.shape[0])]
        d = x64[ii] - x64[jj]
        keep = (np.sqrt((d * d).sum(1)) < radius) & (ii != jj)
        ii_all.append(ii[keep])
        jj_all.append(jj[keep])
    return np.concatenate(ii_all), np.concatenate(jj_all)


def _with_multiplicity(x32: np.ndarray, h: float, ii: np.ndarray, jj: np.ndarray):
    """Accepted visits among candidate pairs: distance < h (fp64 of the fp32 positions), weighted by the
    number of i's 27 neighbour buckets equal to j's key."""
    g = np.floor(x32 / np.float32(h)).astype(np.int64)  # float32 IEEE division, as the kernel
    key = _hash(g)
    x64 = x32.astype(np.float64)
    d = x64[ii] - x64[jj]
    near = np.sqrt((d * d).sum(1)) < h
    ii, jj = ii[near], jj[near]
    nbk = np.stack([_hash(g + o) for o in _OFFS], 1)  # [n, 27]
    mult = (nbk[ii] == key[jj][:, None]).sum(1)
    keep = mult > 0
    return ii[keep], jj[keep], mult[keep].astype(np.float64)


def _pairs_cells(x32: np.ndarray, h: float):
    """_pairs by a cell list (the same pairs as the all-pairs search, in another order)."""
    ii, jj = _candidates(x32, h)
    return _with_multiplicity(x32, h, ii, jj)


class PairCache:
    """_pairs over a rollout, with candidates within h + skin kept between steps: while no particle has
    moved skin / 2 from where the candidates were found, every pair within h is among them, so the
    accepted pairs are exactly those of a fresh search."""

    def __init__(self, h: float, skin: float = 0.2):
        self.h, self.skin = h, skin * h
        self.x0 = None

    def __call__(self, x32: np.ndarray):
        if x32.shape[0] <= 2048:
            return _pairs_dense(x32, self.h)
        moved = 0.0 if self.x0 is None else np.sqrt(((x32.astype(np.float64) - self.x0) ** 2).sum(1)).max()
        if self.x0 is None or moved >= 0.5 * self.skin:
            self.x0 = x32.astype(np.float64)
            self.cand = _candidates(x32, self.h + self.skin)
        return _with_multiplicity(x32, self.h, *self.cand)


def _pairs_dense(x32: np.ndarray, h: float):
    """All-pairs form of _pairs (n^2 distances)."""
    hf = np.float32(h)
    g = np.floor(x32 / hf).astype(np.int64)
    key = _hash(g)
    n = x32.shape[0]
    d = x32[:, None, :].astype(np.float64) - x32[None, :, :].astype(np.float64)
    r = np.sqrt((d * d).sum(-1))
    near = (r < h) & ~np.eye(n, dtype=bool)
    ii, jj = np.nonzero(near)
    nbk = np.stack([_hash(g + o) for o in _OFFS], 1)  # [n, 27]
    mult = (nbk[ii] == key[jj][:, None]).sum(1)
    keep = mult > 0
    return ii[keep], jj[keep], mult[keep].astype(np.float64)


def cpu_rollout_fp64(pos0, vel0, force0, prm: L.Params, box_min, box_max, steps, viscosity, stiffness, density):
    """Rollout in float64 torch ops (autograd-capable in vel0 and the three parameters)."""
    f64 = torch.float64
    h = float(prm.cellSize)
    m, kc, kp = float(prm.mass), float(prm.kernelConstant), float(prm.pressureConstant)
    kv0, mu0 = float(prm.viscosityConstant), float(prm.viscosity)
    kv = viscosity * (kv0 / mu0)
    p0 = float(prm.pressure)
    dt, vmax, rad = float(prm.deltaTime), float(prm.velocitylimit), float(prm.particleRadius)
    x = torch.tensor(pos0[:, :3], dtype=f64)
    v = vel0
    fin = torch.tensor(force0[:, :3], dtype=f64)
    grav = torch.tensor([-0.0, -0.0, -9.8], dtype=f64)
    mn, mx = np.asarray(box_min, np.float64), np.asarray(box_max, np.float64)
    n = x.shape[0]
    pairs = PairCache(h)
    for s in range(steps):
        ii, jj, mult = pairs(x.detach().numpy().astype(np.float32))
        ii_t, jj_t = torch.from_numpy(ii), torch.from_numpy(jj)
        mu_t = torch.from_numpy(mult)
        d = x[ii_t] - x[jj_t]
        r2 = (d * d).sum(1)
        w = h * h - r2
        acc = torch.zeros(n, dtype=f64).index_add(0, ii_t, mu_t * w ** 3)
        count = np.bincount(ii, weights=mult, minlength=n)
        dense = torch.from_numpy(count > 2)
        rho_raw = m * kc * acc
        rho = torch.where(dense, torch.clamp(rho_raw, min=1.0), density * torch.ones(n, dtype=f64))
        prs = torch.where(dense, p0 + stiffness * (rho - density), p0 * torch.ones(n, dtype=f64))
        r = torch.sqrt(r2)
        rr = rho[ii_t] * rho[jj_t]
        q = (prs[ii_t] + prs[jj_t]) / rr
        hr = h - r
        pres = torch.zeros(n, 3, dtype=f64).index_add(0, ii_t, (mu_t * q * hr * hr / r)[:, None] * d)
        visc = torch.zeros(n, 3, dtype=f64).index_add(0, ii_t, (mu_t * hr / rr)[:, None] * (v[jj_t] - v[ii_t]))
        F = (fin if s == 0 else grav) + (kp * pres + kv * visc) * m
        xs = x.detach().numpy()
        for axis, sign, dist in ((1, 1, xs[:, 1] - mn[1]), (1, -1, mx[1] - xs[:, 1]), (2, 1, xs[:, 2] - mn[2]),
                                 (2, -1, mx[2] - xs[:, 2]), (0, 1, xs[:, 0] - mn[0]), (0, -1, mx[0] - xs[:, 0])):
            hit = torch.from_numpy(rad - dist > 1e-5)
            dw = rad - sign * (x[:, axis] - (mn[axis] if sign > 0 else mx[axis]))
            add = sign * (2000.0 * dw - 256.0 * sign * v[:, axis])
            F = F.clone()
            F[:, axis] = F[:, axis] + torch.where(hit, add, torch.zeros_like(add))
        sp = torch.sqrt((F * F).sum(1))
        F = torch.where((sp > vmax)[:, None], F * (vmax / sp)[:, None], F)
        v = v + F * dt
        x = x + v * dt
    return x
