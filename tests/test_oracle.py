"""CPU tests that pin the oracle (oracle/sph_oracle.c) without the reference running.

Each check is independent of the oracle's own code: Python-integer known answers for the hash,
numpy / std::stable_sort for the sort, an fp64 all-pairs brute force for the fields, closed forms,
textbook kernel normalisations and conservation laws.
"""
import json
import math
import os

import numpy as np
import pytest

from conftest import GOLDEN, STEP_CASES, force_errors, load_step_case
from libcl_amd import scenes


# ---------------------------------------------------------------------------------------- hash

def test_hash_known_answers(oracle):
    kat = json.load(open(os.path.join(GOLDEN, "hash_kat.json")))
    for gx, gy, gz, key in kat["keys"]:
        assert oracle.hash_key(gx, gy, gz) == key, (gx, gy, gz)
    h = np.float32(kat["cell_size"])
    xs = np.array([g[0] for g in kat["grid"]], np.float32)
    pos = np.zeros((xs.size, 4), np.float32)
    pos[:, 0] = xs
    pos[:, 1] = xs[::-1]
    pos[:, 2] = -xs
    cell, index = oracle.hash_cells(pos, float(h))
    gx = np.array([g[2] for g in kat["grid"]])
    gy = gx[::-1]
    gz = np.floor(-xs / h).astype(np.int64)
    for k in range(xs.size):
        assert cell[k] == oracle.hash_key(int(gx[k]), int(gy[k]), int(gz[k]))
    np.testing.assert_array_equal(index, np.arange(xs.size, dtype=np.uint32))


def test_hash_range_and_wrap(oracle):
    # products wrap mod 2^32 before the XOR; the key is the low 18 bits
    assert oracle.hash_key(1, 0, 0) == 73856093 % 262144
    assert oracle.hash_key(0, 0, 1) == 83492791 % 262144
    assert oracle.hash_key(-1, 0, 0) == ((-73856093) % 2 ** 32) % 262144
    assert oracle.hash_key(1, 1, 1) == (73856093 ^ 19349663 ^ 83492791) % 262144


# ---------------------------------------------------------------------------------------- sort

def _stable(keys, vals, start, end):
    d = (keys.astype(np.uint64) >> np.uint64(start)) & np.uint64((1 << (end - start)) - 1)
    o = np.argsort(d, kind="stable")
    return keys[o], vals[o]


@pytest.mark.parametrize("n", [1, 2, 255, 256, 257, 1000, 4096, 70001])
@pytest.mark.parametrize("bits", [(0, 24), (0, 32), (8, 16), (4, 8)])
def test_restated_radix_sort_is_stable_sort(oracle, n, bits):
    keys = (scenes.splitmix64(n, n) & np.uint64(0xFFFFFFFF)).astype(np.uint32)
    keys[: n // 3] &= np.uint32(0xF0F)  # many duplicates
    vals = np.arange(n, dtype=np.uint32)[::-1].copy()
    k, v = keys.copy(), vals.copy()
    assert oracle.radix_sort(k, v, *bits)
    ek, ev = _stable(keys, vals, *bits)
    np.testing.assert_array_equal(k, ek)
    np.testing.assert_array_equal(v, ev)
    k2, v2 = keys.copy(), vals.copy()
    oracle.std_stable_sort(k2, v2, *bits)
    np.testing.assert_array_equal(k2, ek)
    np.testing.assert_array_equal(v2, ev)
    k3, v3 = keys.copy(), vals.copy()
    oracle.merge_sort(k3, v3, *bits)
    np.testing.assert_array_equal(k3, ek)
    np.testing.assert_array_equal(v3, ev)


def test_restated_sort_rejections(oracle):
    k = np.zeros(16, np.uint32)
    v = np.zeros(16, np.uint32)
    assert not oracle.radix_sort(k, v, 0, 18)  # bit range not a multiple of 4 (oclRadixSort.cpp:76-80)
    big = np.zeros(oracle.SORT_MAX, np.uint32)
    assert not oracle.radix_sort(big, big.copy(), 0, 4)  # N >= 4,194,304 (cpp:70-74)
    assert oracle.radix_sort(big[:-1].copy(), big[:-1].copy(), 0, 4)


# ---------------------------------------------------------------------------------------- params

def test_kernel_constants_are_textbook_normalisations(oracle):
    """poly6: Kc * 4pi * int_0^h (h^2-r^2)^3 r^2 dr = 1; spiky: its gradient coefficient 45/(pi h^6)
    is the pressure constant and 15/(pi h^6) normalises (h-r)^3; viscosity constant = mu * 45/(pi h^6)."""
    for h in (0.02, 0.0316, 0.1):
        p = oracle.make_params(cellSize=h, viscosity=6.0)
        h64 = float(np.float32(h))
        # numerical quadrature, independent of the closed forms
        r = np.linspace(0.0, h64, 200001)
        poly6 = np.trapezoid((h64 ** 2 - r ** 2) ** 3 * 4 * np.pi * r ** 2, r)
        assert abs(p.kernelConstant * poly6 - 1.0) < 2e-6
        spiky = np.trapezoid((h64 - r) ** 3 * 4 * np.pi * r ** 2, r)
        assert abs(p.pressureConstant / 3.0 * spiky - 1.0) < 2e-6
        assert abs(p.viscosityConstant / (6.0 * p.pressureConstant) - 1.0) < 2e-7


def test_default_params_relations(oracle):
    """clInitFluid relations: spacing = 0.87 (m/rho0)^(1/3), radius = spacing/2, h = 2 spacing."""
    p = oracle.default_params()
    assert p.deltaTime == np.float32(0.005) and p.mass == np.float32(0.006) and p.density == 1000
    assert abs(p.spacing - 0.87 * (0.006 / 1000) ** (1 / 3)) < 1e-8
    assert p.particleRadius == np.float32(0.5 * p.spacing)
    assert p.cellSize == np.float32(2 * p.spacing)


# ---------------------------------------------------------------------------------------- step

@pytest.mark.parametrize("name", STEP_CASES)
def test_oracle_step_matches_fp64_brute_force(oracle, name):
    z = load_step_case(name)
    p = oracle.Params(*[float(x) for x in z["params"]])
    f = oracle.Fluid(z["pos"], z["vel"], z["force"], p, box_min=z["box_min"], box_max=z["box_max"])
    f.step(integrate=False)
    np.testing.assert_array_equal(f.cell, z["key"])
    np.testing.assert_array_equal(f.index, z["perm"])
    np.testing.assert_array_equal(f.cell_start[z["bucket"]], z["start"])
    np.testing.assert_array_equal(f.cell_end[z["bucket"]], z["end"])
    assert np.max(np.abs(f.spos[:, 3] - z["rho"]) / z["rho"]) < 5e-6
    nw, per = force_errors(f.force, z["force_sph"])
    assert nw < 5e-6 and per < 5e-6
    f2 = oracle.Fluid(z["pos"], z["vel"], z["force"], p, box_min=z["box_min"], box_max=z["box_max"])
    f2.step()
    assert np.max(np.abs(f2.vel[:, :3] - z["vel_new"][:, :3])) <= 1e-5 * (np.abs(z["vel_new"]).max() + 1e-3)
    assert np.max(np.abs(f2.pos[:, :3] - z["pos_new"][:, :3])) <= 1e-6


def test_golden_cases_cover_the_quirks():
    """The fixtures contain repeated buckets (multiplicity > 1), count <= 2 particles and overflowing
    neighbourhoods, so the checks above exercise each of those paths."""
    assert int(load_step_case("lattice512")["max_mult"]) >= 2
    assert (load_step_case("random512")["count"] <= 2).sum() > 0
    assert load_step_case("compressed512")["count"].max() > 64


def _params(**kw):
    from oracle import oracle as O
    base = dict(scenes.sph_param_values(0.02))
    base.update(kw)
    return O.make_params(**base)


def test_tetrahedron_closed_form(oracle):
    """Four particles on a regular tetrahedron of edge a < h, at rest, far from the origin: each has
    three neighbours, rho = m Kc 3 (h^2-a^2)^3, and the pressure force points away from the centroid
    with magnitude Kp m (2p/rho^2) (h-a)^2 / a * 4 |x_i - c|."""
    p = _params(stiffness=50.0)
    a = 0.012
    verts = np.array([[1, 1, 1], [1, -1, -1], [-1, 1, -1], [-1, -1, 1]], np.float64) * a / (2 * math.sqrt(2))
    c = np.array([0.5013, 0.4987, 0.5021])
    pos = np.zeros((4, 4), np.float32)
    pos[:, :3] = (verts + c).astype(np.float32)
    f = oracle.Fluid(pos, np.zeros((4, 4), np.float32), np.zeros((4, 4), np.float32), p,
                     box_min=(-5, -5, -5, 0), box_max=(5, 5, 5, 0))
    f.step(integrate=False)
    x = pos[:, :3].astype(np.float64)
    cc = x.mean(0)
    d = np.linalg.norm(x[:, None] - x[None], axis=-1)[0, 1]
    h, m = float(p.cellSize), float(p.mass)
    rho = m * float(p.kernelConstant) * 3 * (h * h - d * d) ** 3
    prs = float(p.pressure) + float(p.stiffness) * (rho - float(p.density))
    inv = np.argsort(f.index)
    np.testing.assert_allclose(f.spos[inv, 3], rho, rtol=2e-6)
    np.testing.assert_allclose(f.svel[inv, 3], prs, rtol=2e-6, atol=1e-3)
    q = 2 * prs / rho ** 2
    expect = float(p.pressureConstant) * m * q * (h - d) ** 2 / d * 4 * (x - cc)
    np.testing.assert_allclose(f.force[:, :3], expect, rtol=3e-5, atol=1e-6 * np.abs(expect).max())


def test_sparse_particles_fall_back_to_rest_density(oracle):
    """count <= 2 (cl:189-199): three particles on a line 0.6h apart: the ends have one neighbour,
    the middle two, so every density is rho0 and every pressure p0."""
    p = _params(pressure=3.0)
    pos = np.zeros((3, 4), np.float32)
    pos[:, 0] = np.array([0.70, 0.712, 0.724], np.float32)
    pos[:, 1:3] = 0.55
    f = oracle.Fluid(pos, np.zeros_like(pos), np.zeros_like(pos), p, box_min=(-5, -5, -5, 0), box_max=(5, 5, 5, 0))
    f.step(integrate=False)
    assert np.all(f.spos[:, 3] == np.float32(p.density))
    assert np.all(f.svel[:, 3] == np.float32(p.pressure))


def test_internal_forces_conserve_momentum(oracle):
    """Pressure and viscosity pair terms are antisymmetric, so with no repeated bucket and no stale
    range the internal forces sum to zero."""
    rng = np.random.default_rng(4)
    n = 768
    pos = np.zeros((n, 4), np.float32)
    pos[:, :3] = (0.6 + rng.uniform(0, 0.08, (n, 3))).astype(np.float32)
    vel = np.zeros((n, 4), np.float32)
    vel[:, :3] = rng.normal(0, 0.5, (n, 3)).astype(np.float32)
    p = _params(stiffness=20.0)
    f = oracle.Fluid(pos, vel, np.zeros_like(pos), p, box_min=(-5, -5, -5, 0), box_max=(5, 5, 5, 0))
    f.step(integrate=False)
    total = f.force[:, :3].astype(np.float64).sum(0)
    scale = np.abs(f.force[:, :3]).sum()
    assert np.all(np.abs(total) <= 2e-6 * scale), (total, scale)


def test_wall_repulsion_closed_form(oracle):
    """A lone particle (rho0 fallback, no SPH force) inside the y- wall band and moving towards it:
    f_y += 2000 d - 256 v_y with d = r - (y - ymin) (cl:324-328); nothing else changes."""
    p = _params()
    pos = np.array([[0.5, 0.003, 0.5, 0.0]], np.float32)
    vel = np.array([[0.1, -0.2, 0.3, 0.0]], np.float32)
    force = np.array([[1.0, 2.0, 3.0, 0.0]], np.float32)
    f = oracle.Fluid(pos, vel, force, p, box_min=(0, 0, 0, 0), box_max=(1, 1, 1, 0))
    f.step(integrate=False)
    oracle.clip_box(f.spos, f.svel, f.force, f.index, p, (0, 0, 0, 0), (1, 1, 1, 0))
    d = float(p.particleRadius) - 0.003
    assert f.force[0, 0] == 1.0 and f.force[0, 2] == 3.0
    assert abs(f.force[0, 1] - (2.0 + 2000 * d + 256 * 0.2)) < 1e-4
    # beyond the band (d <= 1e-5) nothing is added
    f2 = oracle.Fluid(np.array([[0.5, 0.3, 0.5, 0.0]], np.float32), vel, force, p, box_min=(0, 0, 0, 0),
                      box_max=(1, 1, 1, 0))
    f2.step(integrate=False)
    oracle.clip_box(f2.spos, f2.svel, f2.force, f2.index, p, (0, 0, 0, 0), (1, 1, 1, 0))
    np.testing.assert_array_equal(f2.force, force * np.array([1, 1, 1, 0], np.float32))


def test_integration_clamp_and_gravity(oracle):
    p = _params(velocitylimit=10.0, deltaTime=0.01)
    vel = np.array([[1.0, 0.0, 0.0, 0.5], [0.0, 0.0, 0.0, 0.0]], np.float32)
    force = np.array([[30.0, 40.0, 0.0, 0.0], [3.0, 4.0, 0.0, 0.0]], np.float32)
    v = vel.copy()
    oracle.integrate_force(v, force, p)
    # |f| = 50 > 10: f scaled to length 10 before v += f dt; the second is below the limit
    np.testing.assert_allclose(v[0], [1.0 + 6.0 * 0.01, 8.0 * 0.01, 0.0, 0.5], rtol=1e-6)
    np.testing.assert_allclose(v[1], [0.03, 0.04, 0.0, 0.0], rtol=1e-6)
    g = np.zeros((3, 4), np.float32)
    oracle.gravity(g)
    assert np.all(g[:, 2] == np.float32(-9.8))
    assert np.all(np.signbit(g[:, [0, 1, 3]]))  # -9.8f * 0 = -0.0
    pos = np.array([[1.0, 2.0, 3.0, 4.0]], np.float32)
    oracle.integrate_velocity(pos, np.array([[1.0, 1.0, 1.0, 1.0]], np.float32), p)
    np.testing.assert_allclose(pos, [[1.01, 2.01, 3.01, 4.01]], rtol=1e-7)


def test_find_bounds_and_partial_clear(oracle):
    """cl:85-113 against first/last+1 occurrences; the reference's byte-sized fill resets only the
    first 65,536 buckets, so entries above keep their previous values."""
    rng = np.random.default_rng(1)
    keys = np.sort(rng.integers(0, 262144, 5000).astype(np.uint32))
    cs = np.full(262144, 7, np.uint32)
    ce = np.full(262144, 9, np.uint32)
    oracle.clear_buckets(cs, ce, oracle.REF_CLEAR)
    oracle.find_bounds(keys, cs, ce)
    uk, first = np.unique(keys, return_index=True)
    last = np.searchsorted(keys, uk, side="right")
    np.testing.assert_array_equal(cs[uk], first)
    np.testing.assert_array_equal(ce[uk], last)
    empty = np.ones(262144, bool)
    empty[uk] = False
    low = np.arange(262144) < 65536
    assert np.all(cs[empty & low] == 0xFFFFFFFF) and np.all(ce[empty & low] == 0xFFFFFFFF)
    assert np.all(cs[empty & ~low] == 7) and np.all(ce[empty & ~low] == 9)


def test_stale_buckets_are_visited(oracle):
    """Two steps with the reference clear (65,536 entries): a cluster leaves a bucket >= 65,536, whose
    stale range then indexes other particles of the new sorted order. A particle whose neighbourhood
    contains that bucket gets the extra visits; the density difference to a full clear is exactly the
    poly6 sum over the stale range (computed here from the tables, independently of the traversal)."""
    p = _params()
    h = float(p.cellSize)
    rng = np.random.default_rng(9)
    found = False
    for trial in range(40):
        n = 256
        pos = np.zeros((n, 4), np.float32)
        pos[:, :3] = (0.5 + rng.uniform(0, 4 * h, (n, 3))).astype(np.float32)
        fa = oracle.Fluid(pos, np.zeros_like(pos), np.zeros_like(pos), p, box_min=(-5, -5, -5, 0),
                          box_max=(5, 5, 5, 0))
        fa.step(integrate=False)
        pos2 = pos.copy()
        pos2[:, 0] += np.float32(0.55 * h)  # shift: some buckets empty out
        runs = {}
        for clear in (65536, 262144):
            fb = oracle.Fluid(pos2, np.zeros_like(pos), np.zeros_like(pos), p, box_min=(-5, -5, -5, 0),
                              box_max=(5, 5, 5, 0), clear_count=clear, cell_start=fa.cell_start,
                              cell_end=fa.cell_end)
            fb.step(integrate=False)
            runs[clear] = fb
        part, full = runs[65536], runs[262144]
        np.testing.assert_array_equal(part.index, full.index)
        stale = np.nonzero((part.cell_start != full.cell_start) & (np.arange(262144) >= 65536))[0]
        if stale.size == 0:
            continue
        # expected extra poly6 sum per sorted particle from stale ranges in its 27-neighbourhood
        x = part.spos[:, :3].astype(np.float64)
        g = np.floor(part.spos[:, :3] / np.float32(h)).astype(np.int64)
        extra = np.zeros(n)
        extra_cnt = np.zeros(n, np.int64)
        stale_set = set(stale.tolist())
        for i in range(n):
            for dz in (-1, 0, 1):
                for dy in (-1, 0, 1):
                    for dx in (-1, 0, 1):
                        b = oracle.hash_key(int(g[i, 0] + dx), int(g[i, 1] + dy), int(g[i, 2] + dz))
                        if b in stale_set:
                            for j in range(part.cell_start[b], part.cell_end[b]):
                                r = np.linalg.norm(x[i] - x[j])
                                if j != i and r < h:
                                    extra[i] += (h * h - r * r) ** 3
                                    extra_cnt[i] += 1
        hit = (extra_cnt > 0) & (full.spos[:, 3] != p.density)
        if not hit.any():
            continue
        kc = float(p.kernelConstant) * float(p.mass)
        np.testing.assert_allclose(part.spos[hit, 3] - full.spos[hit, 3], kc * extra[hit], rtol=2e-3,
                                   atol=1e-3)
        found = True
        break
    assert found, "no configuration produced a stale visit"


@pytest.mark.parametrize("name", STEP_CASES)
def test_oracle_accepted_counts_match_brute_force(oracle, name):
    """Accepted visits per particle (the integer that drives the count > 2 switch, cl:179-199) against the
    fp64 all-pairs fixture, which weights each pair within h by its visit multiplicity. Particles with
    a pair whose distance lies within 1e-6 h of h are left out: there the fp32 decision of the restated
    fast_length may round either way."""
    z = load_step_case(name)
    p = oracle.Params(*[float(x) for x in z["params"]])
    f = oracle.Fluid(z["pos"], z["vel"], z["force"], p, box_min=z["box_min"], box_max=z["box_max"])
    f.step(integrate=False)
    x = f.spos[:, :3].astype(np.float64)
    h = float(np.float32(p.cellSize))
    r = np.linalg.norm(x[:, None] - x[None], axis=-1)
    ambiguous = (np.abs(r - h) <= 1e-6 * h).any(1)
    assert ambiguous.sum() <= 2
    np.testing.assert_array_equal(f.counts[~ambiguous], z["count"][~ambiguous].astype(np.uint32))
    assert f.counts.max() > 0
