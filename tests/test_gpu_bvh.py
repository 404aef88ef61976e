"""GPU parity of the triangle-mesh BVH (mfl_bvh_trimesh_*) and of clCollideBVH (mfl_collide_bvh,
mfl_fluid3d_collide_bvh) against the oracle: Morton order, nodes (links, split bits, boxes, trav) and
root bit-exact; collision outputs (position, velocity, force) bit-exact, including walks that fill
the reference's 256-entry ring."""
import numpy as np
import pytest
import torch

from libcl_amd import scenes

pytestmark = pytest.mark.gpu


def _meshes():
    dup = scenes.floor_mesh(8, 8, z=0.0, jitter=0.01, seed=2)
    dup.index = np.concatenate([dup.index, dup.index, dup.index[:30]])
    return {
        "random1000": scenes.random_triangles(1000, seed=1),
        "random30k": scenes.random_triangles(90000, seed=2),
        "floor": scenes.floor_mesh(32, 16, (0.0, 0.0), (2.0, 1.0), 0.5, 0.01, 5),
        "flat": scenes.floor_mesh(16, 12, z=0.25),
        "sphere": scenes.sphere_mesh(24, 48, (0.1, -0.2, 0.3), 0.7),
        "duplicates": dup,
        "two": scenes.Mesh("two", scenes.random_triangles(6, seed=9).vertex, np.arange(6, dtype=np.uint32)),
        "three": scenes.Mesh("three", scenes.random_triangles(10, seed=8).vertex, np.arange(10, dtype=np.uint32)),
        "floor9k": scenes.config_collide_1m()[1],
        "floor65k": scenes.floor_mesh(256, 128, (0.0, 0.0), (1.28, 0.64), 0.6, 0.002, 4),
        "floor1m": scenes.floor_mesh(1024, 512, (0.0, 0.0), (4.0, 2.0), 0.3, 0.001, 7),
    }


def _dev(a):
    return torch.from_numpy(np.ascontiguousarray(a)).cuda()


def _gpu_build(m):
    import libcl_amd as L
    b = L.oclBvhTrimesh()
    assert b.compile() == 1, L.last_error()
    v, i = _dev(m.vertex), _dev(m.index.view(np.int32))
    assert b.compute(v, i) == 1, L.last_error()
    return b, v, i


def _check_build(oracle, b, m):
    ref = oracle.bvh_build(m.vertex, m.index)
    assert ref is not None
    nodes, root, key, val = ref
    k, v = b.morton()
    np.testing.assert_array_equal(k, key)
    np.testing.assert_array_equal(v, val)
    assert b.getRootNode() == root and b.getNodeCount() == nodes.shape[0]
    got = b.nodes()
    for f in ("left", "right", "bit", "trav"):
        np.testing.assert_array_equal(got[f], nodes[f], err_msg=f)
    for f in ("bbMin", "bbMax"):
        np.testing.assert_array_equal(got[f].view(np.uint32), nodes[f].view(np.uint32), err_msg=f)
    return nodes, root


@pytest.mark.parametrize("name", list(_meshes().keys()))
def test_build_matches_oracle(oracle, name):
    m = _meshes()[name]
    print(f"bvh build {name}: {m.triangles} triangles", flush=True)
    b, _, _ = _gpu_build(m)
    _check_build(oracle, b, m)


def test_rebuild_with_other_meshes(oracle):
    import libcl_amd as L
    b = L.oclBvhTrimesh()
    assert b.compile() == 1
    ms = _meshes()
    for name in ("floor", "random1000", "two", "sphere", "floor"):
        m = ms[name]
        assert b.compute(_dev(m.vertex), _dev(m.index.view(np.int32))) == 1, L.last_error()
        _check_build(oracle, b, m)


def test_build_rejects_bad_meshes():
    import libcl_amd as L
    b = L.oclBvhTrimesh()
    assert b.compile() == 1
    m = scenes.random_triangles(9, seed=3)
    assert b.compute(_dev(m.vertex), _dev(m.index[:5].view(np.int32))) == 0
    assert "2 triangles" in L.last_error()
    bad = m.index.copy()
    bad[7] = 9
    assert b.compute(_dev(m.vertex), _dev(bad.view(np.int32))) == 0
    assert "index" in L.last_error()


def _particles_near(m, n, seed, height=0.012, speed=1.0, drift=0.2):
    """Particles above the triangles of m (a random triangle, barycentric point, offset along its
    normal), moving back towards it with a drift."""
    u = scenes.uniform(seed, 6 * n).reshape(n, 6)
    tri = m.index[: 3 * m.triangles].reshape(-1, 3)
    t = (u[:, 0] * m.triangles).astype(np.int64)
    a, b, c = (m.vertex[tri[t, k], :3].astype(np.float64) for k in range(3))
    nrm = np.cross(b - a, c - a)
    nrm /= np.linalg.norm(nrm, axis=1, keepdims=True) + 1e-30
    s, r = u[:, 1], u[:, 2]
    flip = s + r > 1
    s, r = np.where(flip, 1 - s, s), np.where(flip, 1 - r, r)
    p = a + s[:, None] * (b - a) + r[:, None] * (c - a) + (u[:, 3:4] * height) * nrm
    v = -speed * nrm + drift * (u[:, 4:6].mean(1, keepdims=True) - 0.5)
    pos = np.zeros((n, 4), np.float32)
    vel = np.zeros((n, 4), np.float32)
    pos[:, :3] = p
    vel[:, :3] = v
    pos[:, 3] = 0.25
    return pos, vel


def _collide_case(name):
    if name == "floor":
        m = scenes.floor_mesh(64, 64, (0.0, 0.0), (1.0, 1.0), 0.0, 0.0, 0, slope=(0.1, 0.05))
        pos, vel = _particles_near(m, 100000, 31)
    elif name == "sphere":
        m = scenes.sphere_mesh(32, 64, (0.0, 0.0, 0.0), 0.5)
        pos, vel = _particles_near(m, 100000, 32, height=0.02, speed=2.0)
    elif name == "random_ring":
        m = scenes.random_triangles(9000, seed=5)
        u = scenes.uniform(33, 6 * 20000).reshape(-1, 6)
        pos = np.zeros((20000, 4), np.float32)
        vel = np.zeros((20000, 4), np.float32)
        pos[:, :3] = (u[:, :3] - 0.5) * 1.2
        vel[:, :3] = (u[:, 3:] - 0.5) * 4.0
    elif name == "at_rest":
        m = scenes.floor_mesh(32, 32, (0.0, 0.0), (1.0, 1.0), 0.0, 0.003, 9)
        pos, vel = _particles_near(m, 50000, 34, speed=0.0, drift=0.0)
    elif name == "floor65k_u32":
        # more than 65,535 nodes: the 32-bit ring slots
        m = scenes.floor_mesh(256, 128, (0.0, 0.0), (1.28, 0.64), 0.6, 0.002, 4)
        pos, vel = _particles_near(m, 200000, 35, speed=10.0)
    elif name == "collide_1m":
        sc, m = scenes.config_collide_1m()
        pos, vel = sc.pos, sc.vel
    else:
        raise KeyError(name)
    force = np.zeros_like(pos)
    force[:, 2] = -9.8
    return m, pos, vel, force


@pytest.mark.parametrize("name", ["floor", "sphere", "random_ring", "at_rest", "floor65k_u32", "collide_1m"])
def test_collide_matches_oracle(oracle, name):
    import libcl_amd as L
    m, pos, vel, force = _collide_case(name)
    print(f"collide {name}: {pos.shape[0]} particles, {m.triangles} triangles", flush=True)
    prm = oracle.make_params(particleRadius=0.02, deltaTime=0.005, cellSize=0.04)
    b, vtx, idx = _gpu_build(m)
    nodes, root = _check_build(oracle, b, m)
    filled = np.zeros(pos.shape[0], np.uint32)
    ep, ev, ef, hit = oracle.collide_bvh(pos, vel, force, prm, nodes, root, m.index, m.vertex, filled)
    P, V, F = _dev(pos), _dev(vel), _dev(force)
    gp = L.params_from(prm.as_dict())
    assert L.collide_bvh(F, P, V, gp, b, idx, vtx) == 1, L.last_error()
    torch.cuda.synchronize()
    hits = int((hit != 0xFFFFFFFF).sum())
    print(f"  hits {hits}, ring filled {int(filled.sum())}", flush=True)
    if name != "at_rest":
        assert hits > 0
    if name == "random_ring":
        assert filled.sum() > 0
    np.testing.assert_array_equal(P.cpu().numpy().view(np.uint32), ep.view(np.uint32))
    np.testing.assert_array_equal(V.cpu().numpy().view(np.uint32), ev.view(np.uint32))
    np.testing.assert_array_equal(F.cpu().numpy().view(np.uint32), ef.view(np.uint32))


def test_collide_rejects_an_unbuilt_bvh():
    import libcl_amd as L
    b = L.oclBvhTrimesh()
    assert b.compile() == 1
    m = scenes.random_triangles(30, seed=6)
    P = torch.zeros(4, 4, device="cuda")
    assert L.collide_bvh(P, P, P, L.default_params(), b, _dev(m.index.view(np.int32)), _dev(m.vertex)) == 0
    assert "not been built" in L.last_error()


def test_integrate_handler_with_collision(oracle):
    """EVT_INTEGRATE (libCL/phys/oclFluid3D.cpp:325-328) with a handler that enqueues clIntegrateForce,
    clGravity, clCollideBVH and clIntegrateVelocity (libCL/TEST/main.cpp:270-301 plus the collision):
    a 16k-particle block falling at 1 unit/s onto a tilted 2-triangle plane across it, two steps,
    against the oracle's step without integration followed by the same four kernels. The fluid forces
    differ from the oracle's in FP32 summation order only (1e-5), so positions and velocities agree to
    that order except where a contact decision sits on a rounding edge."""
    import libcl_amd as L
    sc = scenes.dam_break(32, 16, 32, 1.0, 1, 2)
    sc.vel[:, 2] = -1.0
    prm = L.params_from(sc.params)
    mesh = scenes.floor_mesh(1, 1, (0.0, 0.0), (0.33, 0.17), 0.1, 0.0, 0, slope=(0.05, 0.03))
    f = L.oclFluid3D()
    assert f.compile() == 1
    assert f.setParticleCount(sc.n) == 1
    f.setParameters(prm)
    assert f.setBox(sc.box_min, sc.box_max) == 1
    assert f.upload(sc.pos, sc.vel, sc.force) == 1
    b, vtx, idx = _gpu_build(mesh)
    calls = []

    def handler(prog, stream):
        calls.append(1)
        assert prog.enqueueKernel("clIntegrateForce", stream) == 1
        assert prog.enqueueKernel("clGravity", stream) == 1
        assert prog.collideBVH(b, idx, vtx, stream) == 1
        assert prog.enqueueKernel("clIntegrateVelocity", stream) == 1

    assert f.addEventHandler(f.EVT_INTEGRATE, handler) == 1
    nodes, root = _check_build(oracle, b, mesh)
    op = oracle.make_params(**prm.as_dict())
    fo = oracle.Fluid(sc.pos, sc.vel, sc.force, op, box_min=sc.box_min, box_max=sc.box_max)
    total_hits = 0
    for step in range(2):
        assert f.compute() == 1, L.last_error()
        L.load().mfl_device_finish()
        fo.step(integrate=False)
        oracle.integrate_force(fo.vel, fo.force, op)
        oracle.gravity(fo.force)
        p2, v2, f2, hit = oracle.collide_bvh(fo.pos, fo.vel, fo.force, op, nodes, root, mesh.index, mesh.vertex)
        fo.pos[...], fo.vel[...], fo.force[...] = p2, v2, f2
        oracle.integrate_velocity(fo.pos, fo.vel, op)
        total_hits += int((hit != 0xFFFFFFFF).sum())
        st = f.state()
        dv = np.abs(st["vel"][:, :3] - fo.vel[:, :3]).max(1)
        dp = np.abs(st["pos"][:, :3] - fo.pos[:, :3]).max(1)
        bad = (dv > 1e-4) | (dp > 1e-6)
        print(f"step {step}: hits {int((hit != 0xFFFFFFFF).sum())}, off {int(bad.sum())}", flush=True)
        assert bad.sum() <= 3
        fo.pos[bad], fo.vel[bad], fo.force[bad] = st["pos"][bad], st["vel"][bad], st["force"][bad]
    assert len(calls) == 2 and total_hits > 100
