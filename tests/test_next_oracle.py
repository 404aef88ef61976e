"""CPU tests that pin the oclFluid3Dnext oracle (oracle/next_oracle.c) without the reference running.

The reference holds no test or fixture for oclFluid3Dnext (libCL/TEST/main.cpp does not use it). Each
check is independent of the oracle's code: the closed-form cubic spline and its normalisation, numpy
floor division for the hash, numpy first/last occurrence for the bounds, an fp64 all-pairs density
restricted to the 3 x 3 cell window, the Tait pressure in fp64, an independent search for the last
neighbour in visit order for the force, and the closed-form clamped Euler update.
"""
import numpy as np
import pytest

from libcl_amd import scenes


def _w64(h, r):
    q = np.asarray(r, np.float64) / h
    v = np.where(q < 1, 1 - 1.5 * q * q + 0.75 * q ** 3, np.where(q < 2, 0.25 * (2 - q) ** 3, 0.0))
    return 2.0 / (3.0 * h) * v


def test_spline_closed_form_and_normalisation(oracle):
    h = 2.0
    qs = np.linspace(0.0, 5.0, 2001)
    got = np.array([oracle.next_w(h, float(q)) for q in qs])
    np.testing.assert_allclose(got, _w64(h, qs), rtol=2e-6, atol=1e-9)
    # 1D cubic spline: the integral over the line is 1
    x = np.linspace(-2 * h, 2 * h, 40001)
    vals = np.array([oracle.next_w(h, float(abs(t))) for t in x[::10]])
    integral = np.trapezoid(vals, x[::10])
    assert abs(integral - 1.0) < 1e-4


def test_default_params(oracle):
    p = oracle.next_default_params()
    assert p.cellCountX == 512 and p.cellCountY == 384 and p.cellSize == 2.0
    assert p.deltaTime == np.float32(0.0333333333) and p.density == 10000.0
    assert p.particleRadius == np.float32(0.87 * (1e-4) ** float(np.float32(1 / 3)))


def _state(n_side=(64, 48), seed=3, spread=False):
    if spread:
        n = n_side[0] * n_side[1]
        u = scenes.uniform(seed, 4 * n).reshape(n, 4)
        st = np.zeros((n, 4), np.float32)
        st[:, 0] = 1 + 1023 * u[:, 0]
        st[:, 1] = 1 + 767 * u[:, 1]
        st[:, 2] = (u[:, 2] - 0.5) * 10
        st[:, 3] = (u[:, 3] - 0.5) * 10
        return st
    return scenes.next_block(*n_side, spacing=0.7, origin=(100.0, 200.0), jitter=0.1, seed=seed)


def test_hash_and_bounds(oracle):
    st = _state()
    st[:5, 0] = 1024.0  # gx = 512: aliases column 0 of the next row
    st[5:9, 1] = 768.0  # gy = 384: outside the grid, not recorded
    p = oracle.next_default_params()
    f = oracle.Next(st, p)
    f.hash()
    gx = np.floor(st[:, 0] / np.float32(2.0)).astype(np.int64)
    gy = np.floor(st[:, 1] / np.float32(2.0)).astype(np.int64)
    np.testing.assert_array_equal(f.cell, (gy * 512 + gx).astype(np.uint32))
    order = np.argsort(f.cell, kind="stable")
    cell = f.cell[order].copy()
    f.cell[...] = cell
    f.bounds()
    nc = 512 * 384
    want_s = np.full(nc, 0xFFFFFFFF, np.uint32)
    want_e = np.full(nc, 0xFFFFFFFF, np.uint32)
    u, first, counts = np.unique(cell, return_index=True, return_counts=True)
    ok = u < nc
    want_s[u[ok]] = first[ok]
    want_e[u[ok]] = (first + counts)[ok]
    np.testing.assert_array_equal(f.start, want_s)
    np.testing.assert_array_equal(f.end, want_e)


def _prepared(oracle, st):
    p = oracle.next_default_params()
    f = oracle.Next(st, p)
    f.hash()
    order = np.argsort(f.cell, kind="stable").astype(np.uint32)
    f.cell[...] = f.cell[order]
    f.index[...] = order
    f.sorted[...] = st[order]
    f.bounds()
    return f


@pytest.mark.parametrize("spread", [False, True])
def test_density_pressure_against_fp64_window_sums(oracle, spread):
    st = _state((48, 40), seed=5, spread=spread)
    f = _prepared(oracle, st)
    f.pressure_pass()
    p = f.params
    xy = f.sorted[:, :2].astype(np.float64)
    g = np.floor(f.sorted[:, :2] / np.float32(p.cellSize)).astype(np.int64)
    n = xy.shape[0]
    for i in range(0, n, 7):
        x0, x1 = max(g[i, 0] - 1, 0), min(g[i, 0] + 1, p.cellCountX - 1)
        y0, y1 = max(g[i, 1] - 1, 0), min(g[i, 1] + 1, p.cellCountY - 1)
        inwin = (g[:, 0] >= x0) & (g[:, 0] <= x1) & (g[:, 1] >= y0) & (g[:, 1] <= y1)
        inwin[i] = False
        r = np.sqrt(((xy[inwin] - xy[i]) ** 2).sum(1))
        rho = _w64(p.cellSize, r).sum()
        want = 0.0 if rho == 0 else 0.87 * ((rho / 100.0) ** 7 - 1.0) / (rho * rho)
        assert f.pressure[i, 0] == f.pressure[i, 1]
        assert abs(f.pressure[i, 0] - want) <= 2e-5 * abs(want) + 1e-12, (i, f.pressure[i, 0], want)


@pytest.mark.parametrize("spread", [False, True])
def test_force_is_the_last_neighbour_in_visit_order(oracle, spread):
    st = _state((48, 40), seed=6, spread=spread)
    f = _prepared(oracle, st)
    f.pressure_pass()
    f.force_pass()
    p = f.params
    cells = f.cell
    xy = f.sorted[:, :2].astype(np.float64)
    g = np.floor(f.sorted[:, :2] / np.float32(p.cellSize)).astype(np.int64)
    members = {}
    for j, c in enumerate(cells):
        members.setdefault(int(c), []).append(j)
    for i in range(0, xy.shape[0], 5):
        x0, x1 = max(g[i, 0] - 1, 0), min(g[i, 0] + 1, p.cellCountX - 1)
        y0, y1 = max(g[i, 1] - 1, 0), min(g[i, 1] + 1, p.cellCountY - 1)
        last = None
        for x in range(x0, x1 + 1):
            for y in range(y0, y1 + 1):
                for j in members.get(y * p.cellCountX + x, []):
                    if j != i:
                        last = j
        if last is None:
            want = np.array([0.0, -9.81])
        else:
            d = xy[last] - xy[i]
            r = np.sqrt((d * d).sum())
            w = _w64(p.cellSize, r)
            k = -(float(f.pressure[i, 0]) + float(f.pressure[last, 0]))
            want = k * d / r * w + 0.25 * d * w - np.array([0.0, 9.81])
        np.testing.assert_allclose(f.force[i], want, rtol=1e-5, atol=1e-6 * (1 + abs(want).max()))


def test_step_updates_state_by_clamped_euler(oracle):
    st = _state((40, 32), seed=7, spread=True)
    st[::11, 0] = 1023.999
    st[::11, 2] = 50.0  # pushed past the x clamp
    p = oracle.next_default_params()
    f = oracle.Next(st, p)
    f.step()
    dt = float(p.deltaTime)
    o = f.index
    v = st[o, 2:4].astype(np.float64) + f.force.astype(np.float64) * dt
    x = np.clip(st[o, 0:2].astype(np.float64) + v * dt, [1, 1], [1024, 768])
    np.testing.assert_allclose(f.state[o, 2:4], v, rtol=1e-6, atol=1e-6)
    np.testing.assert_allclose(f.state[o, 0:2], x, rtol=1e-6, atol=1e-5)
    assert (f.state[:, 0] <= 1024).all() and (f.state[:, 0] >= 1).all()
    # the index is the stable cell order of the hashed state
    gx = np.floor(st[:, 0] / np.float32(2.0)).astype(np.int64)
    gy = np.floor(st[:, 1] / np.float32(2.0)).astype(np.int64)
    np.testing.assert_array_equal(o, np.argsort(gy * 512 + gx, kind="stable"))
