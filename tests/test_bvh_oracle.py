"""CPU tests that pin the BVH / collision oracle (oracle/bvh_oracle.c) without the reference running.

The reference holds no golden vectors for oclBvhTrimesh or clCollideBVH (its test only logs the root,
libCL/TEST/main.cpp:349-404). Each check here is independent of the oracle's code: a bit-by-bit
interleave for code(), numpy min/max for the boxes and the global box, numpy's stable argsort for the
key order, the properties that define a binary radix tree for the links, closed-form single-triangle
contacts derived by hand for the collision, a floor mesh whose every particle must hit the triangle
under it, and an independent formulation of the breadth-first first-hit order.
"""
import math

import numpy as np
import pytest

from libcl_amd import scenes

INF = np.float32(np.finfo(np.float32).max)


def _spread(x: int) -> int:
    return sum(((x >> i) & 1) << (3 * i) for i in range(10))


def test_morton_code_is_the_bit_interleave(oracle):
    for x in list(range(1024)) + [0x3FF, 0x2AA, 0x155]:
        assert oracle.morton_code(x) == _spread(x), x


def test_global_box_is_min_max_of_all_vertices(oracle):
    m = scenes.random_triangles(1000, seed=3)
    v = m.vertex.copy()
    v[::7, 3] = -2.0  # w takes part (the empty box starts at w = 0)
    mn, mx = oracle.bvh_aabb(v)
    np.testing.assert_array_equal(mn[:3], v[:, :3].min(0))
    np.testing.assert_array_equal(mx[:3], v[:, :3].max(0))
    assert mn[3] == min(0.0, v[:, 3].min()) and mx[3] == max(0.0, v[:, 3].max())


def _meshes():
    dup = scenes.floor_mesh(8, 8, z=0.0, jitter=0.01, seed=2)
    dup.index = np.concatenate([dup.index, dup.index, dup.index[:30]])  # repeated triangles: equal keys
    flat = scenes.floor_mesh(16, 12, z=0.25)                             # zero z extent
    return {
        "random1000": scenes.random_triangles(1000, seed=1),
        "floor": scenes.floor_mesh(32, 16, (0.0, 0.0), (2.0, 1.0), 0.5, 0.01, 5),
        "flat": flat,
        "sphere": scenes.sphere_mesh(10, 16, (0.1, -0.2, 0.3), 0.7),
        "duplicates": dup,
        "two": scenes.Mesh("two", scenes.random_triangles(6, seed=9).vertex, np.arange(6, dtype=np.uint32)),
        "three": scenes.Mesh("three", scenes.random_triangles(9, seed=8).vertex, np.arange(9, dtype=np.uint32)),
    }


def _quantized(m: scenes.Mesh, mn, mx):
    """Centroid quantization in float64 and how far each value sits from an integer boundary."""
    tri = m.index[: 3 * m.triangles].reshape(-1, 3)
    c = m.vertex[tri][:, :, :3].astype(np.float64).mean(1)
    ext = mx[:3].astype(np.float64) - mn[:3]
    with np.errstate(invalid="ignore", divide="ignore"):
        q = (c - mn[:3]) / ext * 1023.0
    q = np.where(np.isfinite(q), q, 0.0)
    frac = np.abs(q - np.round(q))
    return np.clip(np.floor(q), 0, 1023).astype(np.int64), frac


def _unspread(k: int) -> int:
    return sum(((k >> (3 * i)) & 1) << i for i in range(10))


def _split(k0: int, k1: int) -> int:
    x = (int(k0) ^ int(k1)) & 0x3FFFFFFF
    return x.bit_length() - 1 if x else 0


def _inorder(nodes, root):
    out, stack, k = [], [], root
    while stack or k is not None:
        while k is not None:
            stack.append(k)
            lf = nodes[k]["left"] == nodes[k]["right"]
            k = None if lf else int(nodes[k]["left"])
        k = stack.pop()
        out.append(k)
        lf = nodes[k]["left"] == nodes[k]["right"]
        k = None if lf else int(nodes[k]["right"])
    return out


@pytest.mark.parametrize("name", list(_meshes().keys()))
def test_build_is_a_binary_radix_tree_over_the_sorted_morton_keys(oracle, name):
    m = _meshes()[name]
    T = m.triangles
    res = oracle.bvh_build(m.vertex, m.index)
    assert res is not None
    nodes, root, key, val = res
    assert nodes.shape[0] == 2 * T - 1 and root < T - 1
    # key order: numpy's stable argsort of the unsorted keys
    mn, mx = oracle.bvh_aabb(m.vertex)
    k0, v0 = oracle.bvh_morton(m.index, m.vertex, mn, mx)
    np.testing.assert_array_equal(v0, np.arange(T, dtype=np.uint32))
    order = np.argsort(k0, kind="stable")
    np.testing.assert_array_equal(key, k0[order])
    np.testing.assert_array_equal(val, order.astype(np.uint32))
    # keys interleave the quantized centroid (x at bit 0, y bit 1, z bit 2): every axis with an extent,
    # away from rounding edges
    q, frac = _quantized(m, mn, mx)
    ext = mx[:3] > mn[:3]
    for axis in range(3):
        got = np.array([_unspread(int(k) >> axis) for k in k0])
        sel = frac[:, axis] > 1e-3
        if ext[axis]:
            np.testing.assert_array_equal(got[sel], q[sel, axis])
    # internal nodes: split bits; leaves: triangle ids
    for i in range(T - 1):
        assert nodes[i]["bit"] == _split(key[i], key[i + 1])
        assert nodes[i]["trav"] == 2
    leaves = nodes[T - 1:]
    np.testing.assert_array_equal(leaves["bit"], val)
    assert (leaves["left"] == 0).all() and (leaves["right"] == 0).all() and (leaves["trav"] == 0).all()
    # in-order: leaf 0, node 0, leaf 1, node 1, ..., leaf T-1 (every node once)
    seq = _inorder(nodes, root)
    want = [T - 1]
    for i in range(T - 1):
        want += [i, T + i]
    assert seq == want
    # heap order of (bit desc, index asc), and key ranges for distinct keys
    lo = {T - 1 + g: g for g in range(T)}
    hi = dict(lo)

    def rng(k):
        if k >= T - 1:
            return lo[k], hi[k]
        a, _ = rng(int(nodes[k]["left"]))
        _, b = rng(int(nodes[k]["right"]))
        return a, b

    distinct = len(set(key.tolist())) == T
    for k in range(T - 1):
        for c in (int(nodes[k]["left"]), int(nodes[k]["right"])):
            if c < T - 1:
                assert nodes[k]["bit"] > nodes[c]["bit"] or (nodes[k]["bit"] == nodes[c]["bit"] and k < c)
        a, b = rng(k)
        assert a <= k < b
        if distinct:
            bit = int(nodes[k]["bit"])
            assert key[a] >> (bit + 1) == key[b] >> (bit + 1)
            assert (key[a] >> bit) & 1 == 0 and (key[b] >> bit) & 1 == 1
        # box = min / max over the triangles of the leaf range (w included, empty box has w = 0)
        tri = m.index[:3 * T].reshape(-1, 3)[val[a:b + 1]]
        pts = m.vertex[tri.reshape(-1)]
        np.testing.assert_array_equal(nodes[k]["bbMin"][:3], pts[:, :3].min(0))
        np.testing.assert_array_equal(nodes[k]["bbMax"][:3], pts[:, :3].max(0))
        assert nodes[k]["bbMin"][3] == min(0.0, pts[:, 3].min()) and nodes[k]["bbMax"][3] == max(0.0, pts[:, 3].max())


def test_build_rejects_fewer_than_two_triangles_and_bad_indices(oracle):
    m = scenes.random_triangles(6, seed=2)
    assert oracle.bvh_build(m.vertex, m.index[:5]) is None
    bad = m.index.copy()
    bad[4] = 6
    assert oracle.bvh_build(m.vertex, bad) is None


# ---------------------------------------------------------------------------------------- collision

def _params(oracle, radius=0.02, dt=0.005, h=0.04):
    return oracle.make_params(particleRadius=radius, deltaTime=dt, cellSize=h)


def _tri_mesh(v0, v1, v2):
    """The triangle of interest plus a tiny one far away (a BVH needs two)."""
    vertex = np.ones((6, 4), np.float32)
    vertex[:3, :3] = [v0, v1, v2]
    vertex[3:, :3] = [(50.0, 50.0, 50.0), (50.1, 50.0, 50.0), (50.0, 50.1, 50.0)]
    return vertex, np.arange(6, dtype=np.uint32)


def _contact_expected(p, pos, vel, force, v0, n_unit):
    """Closed form of a contact (derived by hand from cl:380-461): the displacement dt*v reaches the
    plane, the particle ends within the radius of it: x += (d/|d|)(n.d)(radius - dist), v = 0,
    F += (v - (n.v) n)/dt + (pc/rho0) m n (h - dist)^2, where dist = n.(x + d - v0)."""
    dt, R, h = float(p.deltaTime), float(p.particleRadius), float(p.cellSize)
    pc, rho, m = float(p.pressureConstant), float(p.density), float(p.mass)
    pos, vel, force = (np.asarray(a, np.float64) for a in (pos, vel, force))
    d = vel[:3] * dt
    dist = float(n_unit @ (pos[:3] + d - np.asarray(v0, np.float64)))
    depth = R - dist
    out_pos = pos.copy()
    out_pos[3] = 0.0
    out_pos[:3] += d / np.linalg.norm(d) * (n_unit @ d) * depth
    out_force = force.copy()
    out_force[:3] += (vel[:3] - (n_unit @ vel[:3]) * n_unit) / dt + pc / rho * m * n_unit * (h - dist) ** 2
    return out_pos, np.zeros(4), out_force


def _one(oracle, p, vertex, index, pos, vel, force=(0.0, 0.0, 0.0, 0.0)):
    nodes, root, _, _ = oracle.bvh_build(vertex, index)
    P = np.array([pos], np.float32)
    V = np.array([vel], np.float32)
    F = np.array([force], np.float32)
    return oracle.collide_bvh(P, V, F, p, nodes, root, index, vertex), nodes


# a tilted triangle (its box has thickness on every axis) in the plane z = 0.1 x + 0.05 y
TRI = ((0.0, 0.0, 0.0), (1.0, 0.0, 0.1), (0.0, 1.0, 0.05))
TRI_N = np.array([-0.1, -0.05, 1.0]) / np.linalg.norm([-0.1, -0.05, 1.0])


def _above(x, y, d):
    """The point d along the plane's unit normal from the plane point over (x, y)."""
    return np.array([x, y, 0.1 * x + 0.05 * y]) + d * TRI_N


@pytest.mark.parametrize("x,y,d,vel", [(0.2, 0.3, 0.01, -1.0 * TRI_N),
                                       (0.2, 0.3, 0.004, -0.5 * TRI_N + np.array([0.3, -0.2, 0.0])),
                                       (0.6, 0.1, 0.012, -2.0 * TRI_N + np.array([0.0, 0.4, 0.1]))])
def test_contact_with_a_triangle(oracle, x, y, d, vel):
    p = _params(oracle)
    vertex, index = _tri_mesh(*TRI)
    pos = (*_above(x, y, d), 0.5)
    vel = (*vel, 0.0)
    (P, V, F, hit), nodes = _one(oracle, p, vertex, index, pos, vel, (1.0, 2.0, 3.0, 0.0))
    assert hit[0] != 0xFFFFFFFF and nodes[hit[0]]["bit"] == 0
    ep, ev, ef = _contact_expected(p, pos, vel, (1.0, 2.0, 3.0, 0.0), TRI[0], TRI_N)
    np.testing.assert_allclose(P[0], ep, rtol=1e-5, atol=1e-6)
    np.testing.assert_array_equal(V[0], ev)
    np.testing.assert_allclose(F[0], ef, rtol=1e-4)


def test_a_zero_thickness_box_is_never_entered(oracle):
    """intersectBox (cl:464-484) accepts only smallest_tmax > largest_tmin: a box flat in z has
    tmin = tmax on that axis for every ray, so a triangle in a plane z = const is never hit."""
    p = _params(oracle)
    vertex, index = _tri_mesh((0.0, 0.0, 0.0), (1.0, 0.0, 0.0), (0.0, 1.0, 0.0))
    for vel in ((0.0, 0.0, -1.0, 0.0), (0.3, -0.2, -0.5, 0.0), (0.0, 0.0, 1.0, 0.0)):
        (P, V, F, hit), _ = _one(oracle, p, vertex, index, (0.2, 0.3, 0.01, 0.0), vel)
        assert hit[0] == 0xFFFFFFFF


@pytest.mark.parametrize("case", ["back_side", "moving_away", "parallel", "outside_s", "outside_t", "outside_st",
                                  "too_far", "at_rest"])
def test_no_contact(oracle, case):
    """Each early exit of cl:380-461 leaves the particle untouched."""
    p = _params(oracle)
    vertex, index = _tri_mesh(*TRI)
    tangent = np.array([1.0, 0.0, 0.1]) / np.linalg.norm([1.0, 0.0, 0.1])
    pos, vel = {
        "back_side": (_above(0.2, 0.2, -0.01), TRI_N),
        "moving_away": (_above(0.2, 0.2, 0.01), TRI_N),
        "parallel": (_above(0.2, 0.2, 0.01), tangent),
        "outside_s": (_above(-0.1, 0.3, 0.01), -TRI_N),
        "outside_t": (_above(0.3, -0.1, 0.01), -TRI_N),
        "outside_st": (_above(0.6, 0.6, 0.01), -TRI_N),
        "too_far": (_above(0.2, 0.2, 0.5), -TRI_N),
        "at_rest": (_above(0.2, 0.2, 0.01), np.zeros(3)),
    }[case]
    pos = np.float32((*pos, 0.0))
    vel = np.float32((*vel, 0.0))
    (P, V, F, hit), _ = _one(oracle, p, vertex, index, pos, vel, (1.0, 2.0, 3.0, 4.0))
    assert hit[0] == 0xFFFFFFFF
    np.testing.assert_array_equal(P[0], pos)
    np.testing.assert_array_equal(V[0], vel)
    np.testing.assert_array_equal(F[0], np.float32((1.0, 2.0, 3.0, 4.0)))


def test_floor_particles_hit_the_triangle_below(oracle):
    """Particles just above a tilted 16 x 16-quad floor falling with a small drift hit exactly the
    triangle containing their crossing point (picked away from the edges)."""
    p = _params(oracle)
    sx, sy = 0.1, 0.05
    m = scenes.floor_mesh(16, 16, (0.0, 0.0), (1.0, 1.0), 0.0, slope=(sx, sy))
    nrm = np.array([-sx, -sy, 1.0]) / np.linalg.norm([-sx, -sy, 1.0])
    nodes, root, _, _ = oracle.bvh_build(m.vertex, m.index)
    n = 2000
    u = scenes.uniform(11, 4 * n).reshape(n, 4)
    x = 0.02 + 0.96 * u[:, 0]
    y = 0.02 + 0.96 * u[:, 1]
    pos = np.zeros((n, 4), np.float32)
    vel = np.zeros((n, 4), np.float32)
    pos[:, 0] = x
    pos[:, 1] = y
    pos[:, 2] = sx * x + sy * y + 0.001 + 0.01 * u[:, 2]
    vel[:, 0] = 0.2 * (u[:, 3] - 0.5)
    vel[:, 2] = -1.0
    force = np.zeros((n, 4), np.float32)
    P, V, F, hit = oracle.collide_bvh(pos, vel, force, p, nodes, root, m.index, m.vertex)
    dt = float(p.deltaTime)
    p64, d = pos[:, :3].astype(np.float64), vel[:, :3].astype(np.float64) * dt
    r = -(p64 @ nrm) / (d @ nrm)
    ix, iy = p64[:, 0] + r * d[:, 0], p64[:, 1] + r * d[:, 1]
    fx, fy = ix * 16, iy * 16
    qx, qy = np.floor(fx).astype(int), np.floor(fy).astype(int)
    ex, ey = fx - qx, fy - qy
    clear = (np.minimum(np.minimum(ex, 1 - ex), np.minimum(ey, 1 - ey)) > 1e-3) & (np.abs(ex + ey - 1) > 1e-3)
    tri = 2 * (qy * 16 + qx) + (ex + ey > 1)
    dist = (p64 + d) @ nrm
    expect_hit = float(p.particleRadius) - dist > 0
    got = np.where(hit != 0xFFFFFFFF, nodes["bit"][np.minimum(hit, len(nodes) - 1)], -1)
    sel = clear & expect_hit
    assert sel.sum() > 1000
    np.testing.assert_array_equal(got[sel], tri[sel])
    for k in np.nonzero(sel)[0][:50]:
        ep, _, ef = _contact_expected(p, pos[k], vel[k], force[k], (0.0, 0.0, 0.0), nrm)
        np.testing.assert_allclose(P[k], ep, rtol=1e-5, atol=1e-6)
        np.testing.assert_allclose(F[k], ef, rtol=1e-4)
    assert (hit[~expect_hit] == 0xFFFFFFFF).all()


def _aabb_pass(nodes, pos, vel, radius):
    """collideAABB for every node at once (float32 numpy, the same definitions as cl:464-496)."""
    f = np.float32
    v = vel.astype(f)
    l = np.sqrt(f((v[0] * v[0] + v[1] * v[1]) + v[2] * v[2]) + v[3] * v[3])
    nd = v / l if l != 0 else v
    with np.errstate(divide="ignore", invalid="ignore", over="ignore"):
        inv = f(1.0) / nd[:3]
        tb = inv * (nodes["bbMin"][:, :3] - pos[:3].astype(f))
        tt = inv * (nodes["bbMax"][:, :3] - pos[:3].astype(f))
        tmin = np.where(tb < tt, tb, tt)
        tmax = np.where(tt < tb, tb, tt)
    mx = lambda a, b: np.where(a < b, b, a)  # noqa: E731
    mn = lambda a, b: np.where(b < a, b, a)  # noqa: E731
    near = mx(mx(tmin[:, 0], tmin[:, 1]), mx(tmin[:, 0], tmin[:, 2]))
    far = mn(mn(tmax[:, 0], tmax[:, 1]), mn(tmax[:, 0], tmax[:, 2]))
    lim = l if f(radius) < l else f(radius)
    return (far > near) & (near < lim)


def test_breadth_first_order_of_the_first_hit(oracle):
    """100 large overlapping triangles (199 nodes: the 256-entry ring never fills): the walk stops at the
    colliding leaf that comes first by (depth, in-order position) among the leaves whose every
    ancestor-to-child step passes the box test, the order a FIFO visit of a binary tree produces."""
    p = _params(oracle, radius=0.05, dt=0.05)
    m = scenes.random_triangles(300, seed=4)
    nodes, root, _, _ = oracle.bvh_build(m.vertex, m.index)
    T = m.triangles
    depth = {root: 0}
    parent = {}
    for k in range(T - 1):
        for c in (int(nodes[k]["left"]), int(nodes[k]["right"])):
            parent[c] = k
    pos_in = {k: i for i, k in enumerate(_inorder(nodes, root))}

    def dep(k):
        if k not in depth:
            depth[k] = dep(parent[k]) + 1
        return depth[k]

    n = 400
    u = scenes.uniform(21, 7 * n).reshape(n, 7)
    pos = np.zeros((n, 4), np.float32)
    vel = np.zeros((n, 4), np.float32)
    pos[:, :3] = (u[:, :3] - 0.5) * 1.2
    vel[:, :3] = (u[:, 3:6] - 0.5) * 4.0
    P, V, F, hit = oracle.collide_bvh(pos, vel, np.zeros_like(pos), p, nodes, root, m.index, m.vertex)
    hits = 0
    for q in range(n):
        ok = _aabb_pass(nodes, pos[q], vel[q], p.particleRadius)
        reach = []
        for g in range(T):
            leaf = T - 1 + g
            k, good = leaf, True
            while k != root:
                if not ok[k]:
                    good = False
                    break
                k = parent[k]
            if good:
                reach.append(leaf)
        reach.sort(key=lambda k: (dep(k), pos_in[k]))
        first = 0xFFFFFFFF
        for leaf in reach:
            tri = int(nodes[leaf]["bit"])
            sub = m.index.reshape(-1, 3)[tri]
            vtx, idx = _tri_mesh(*m.vertex[sub][:, :3])
            (_, _, _, h1), _ = _one(oracle, p, vtx, idx, pos[q], vel[q])
            if h1[0] != 0xFFFFFFFF:
                first = leaf
                break
        assert hit[q] == first, (q, hit[q], first)
        hits += first != 0xFFFFFFFF
    assert hits > 20
