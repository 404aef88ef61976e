"""GPU parity of the SPH step (oclFluid3D::compute, libCL/phys/oclFluid3D.cpp:293-342).

Sort keys, permutation and bucket tables are compared bit-exact; density/pressure and force fields
within 1e-5 (relative; forces normwise and per particle floored by the RMS force, because interior
pressure forces are near-cancelling sums). Every comparison starts from an identical snapshot, so
chaotic divergence of trajectories never enters a tolerance.
"""
import numpy as np
import pytest

from conftest import STEP_CASES, force_errors, load_step_case
from libcl_amd import scenes

pytestmark = pytest.mark.gpu

TOL = 1e-5


# launch shapes of the neighbour passes: LDS tiles (tile_kept is the default), tiles so shallow that
# dense neighbourhoods are tested in several windows (or, past four windows, walked from global
# memory), and the recorded-visit-list passes
SHAPES = {"tile_list": {"neighbour_pass": "tile_list", "density_lanes": 4, "tile_depth": 192},
          "tile": {"neighbour_pass": "tile", "density_lanes": 8, "tile_depth": 256},
          "tile_list8": {"neighbour_pass": "tile_list", "density_lanes": 8, "tile_depth": 256, "force_chunks": 1},
          "tile64": {"neighbour_pass": "tile", "density_lanes": 8, "tile_depth": 64, "block_waves": 2},
          "tile_list64": {"neighbour_pass": "tile_list", "density_lanes": 8, "tile_depth": 64, "block_waves": 1},
          "tile_list4_64": {"neighbour_pass": "tile_list", "density_lanes": 4, "tile_depth": 64, "block_waves": 1},
          "tile_list4_ch2": {"neighbour_pass": "tile_list", "density_lanes": 4, "force_chunks": 2},
          "tile_kept": {},
          "tile_kept2": {"neighbour_pass": "tile_kept", "density_lanes": 2, "tile_depth": 96},
          "tile_kept2_64": {"neighbour_pass": "tile_kept", "density_lanes": 2, "tile_depth": 64, "block_waves": 1},
          "tile_list2": {"neighbour_pass": "tile_list", "density_lanes": 2, "tile_depth": 96},
          "tile_kept8": {"neighbour_pass": "tile_kept", "density_lanes": 8, "tile_depth": 256},
          "tile_kept64": {"neighbour_pass": "tile_kept", "density_lanes": 4, "tile_depth": 64, "block_waves": 1},
          "list": {"neighbour_pass": "list"}}


def make_fluid(pos, vel, force, params, box_min, box_max, clear=65536, tables=None, shape="tile_kept"):
    import libcl_amd as L
    f = L.oclFluid3D()
    assert f.compile() == 1, L.last_error()
    if SHAPES[shape]:
        assert f.setTuning(**SHAPES[shape]) == 1, L.last_error()
    assert f.setParticleCount(pos.shape[0]) == 1, L.last_error()
    assert f.setRecordCounts(True) == 1  # accepted visits per particle, compared with the oracle's
    f.setParameters(params)
    assert f.setBox(box_min, box_max) == 1
    assert f.setBucketClear(clear) == 1
    assert f.upload(pos, vel, force) == 1, L.last_error()
    if tables is not None:
        assert f.bfCellStart.write(tables[0]) == 1
        assert f.bfCellEnd.write(tables[1]) == 1
    return f


def params_of(values):
    import libcl_amd as L
    return L.Params(*[float(x) for x in values])


def run_forces_only(f):
    """compute() with an EVT_INTEGRATE handler that does nothing: leaves the force pass result."""
    import libcl_amd as L
    assert f.addEventHandler(f.EVT_INTEGRATE, lambda prog, stream: None) == 1
    assert f.compute() == 1, L.last_error()
    L.load().mfl_device_finish()
    f.removeEventHandler()


@pytest.mark.parametrize("shape", list(SHAPES))
@pytest.mark.parametrize("name", STEP_CASES)
def test_step_matches_golden(name, shape):
    """One step from the committed fp64 all-pairs fixture (tests/golden/make_golden.py)."""
    print(f"golden step {name} {shape}", flush=True)
    z = load_step_case(name)
    prm = params_of(z["params"])
    f = make_fluid(z["pos"], z["vel"], z["force"], prm, z["box_min"], z["box_max"], shape=shape)
    run_forces_only(f)
    it = f.internals()
    np.testing.assert_array_equal(it["cell"], z["key"])
    np.testing.assert_array_equal(it["index"], z["perm"])
    np.testing.assert_array_equal(it["cell_start"][z["bucket"]], z["start"])
    np.testing.assert_array_equal(it["cell_end"][z["bucket"]], z["end"])
    empty = np.ones(262144, bool)
    empty[z["bucket"]] = False
    assert np.all(it["cell_start"][empty] == 0xFFFFFFFF) and np.all(it["cell_end"][empty] == 0xFFFFFFFF)
    rel = np.abs(it["spos"][:, 3] - z["rho"]) / np.abs(z["rho"])
    assert rel.max() <= TOL
    prs_scale = np.abs(z["prs"]) + np.sqrt(np.mean(z["prs"] ** 2))
    assert (np.abs(it["svel"][:, 3] - z["prs"]) / prs_scale).max() <= TOL
    st = f.state()
    nw, per = force_errors(st["force"], z["force_sph"])
    assert nw <= TOL and per <= TOL, (nw, per)
    assert np.all(st["force"][:, 3] == 0)

    # full step (clip box, clamp, Euler, gravity) from the same snapshot
    f2 = make_fluid(z["pos"], z["vel"], z["force"], prm, z["box_min"], z["box_max"], shape=shape)
    assert f2.compute() == 1
    st2 = f2.state()
    np.testing.assert_array_equal(st2["force"], np.tile(scenes.GRAVITY_FORCE, (z["pos"].shape[0], 1)))
    dt = float(prm.deltaTime)
    fmag = np.linalg.norm(z["force_clip"][:, :3], axis=1)
    vtol = TOL * (fmag + np.sqrt(np.mean(fmag ** 2))) * dt + 4e-7 * (np.abs(z["vel_new"][:, :3]).max(1) + 1e-6)
    assert np.all(np.linalg.norm(st2["vel"][:, :3] - z["vel_new"][:, :3], axis=1) <= vtol)
    ptol = vtol * dt + 2.5e-7 * np.abs(z["pos_new"][:, :3]).max(1)
    assert np.all(np.linalg.norm(st2["pos"][:, :3] - z["pos_new"][:, :3], axis=1) <= ptol)


def _oracle_compare(f, fo, check_fields=True):
    """Keys, permutation, both bucket tables and every particle's accepted-visit count (the integer
    behind the count > 2 switch, cl:179-199; both sides form |d|^2 as fma(z, z, fma(y, y, x * x)))
    bit-exact; density and pressure within TOL."""
    it = f.internals()
    np.testing.assert_array_equal(it["cell"], fo.cell)
    np.testing.assert_array_equal(it["index"], fo.index)
    np.testing.assert_array_equal(it["cell_start"], fo.cell_start)
    np.testing.assert_array_equal(it["cell_end"], fo.cell_end)
    np.testing.assert_array_equal(f.bfNeighbourCount.read(np.uint32)[:fo.counts.shape[0]], fo.counts)
    if check_fields:
        rel = np.abs(it["spos"][:, 3] - fo.spos[:, 3]) / np.abs(fo.spos[:, 3])
        assert rel.max() <= TOL, rel.max()
        p = fo.svel[:, 3].astype(np.float64)
        assert (np.abs(it["svel"][:, 3] - p) / (np.abs(p) + np.sqrt(np.mean(p ** 2)))).max() <= TOL
    return it


_TRAJECTORIES = {}


def _oracle_snapshots(oracle, sc, prm, clear, targets):
    """States of the oracle's own trajectory at the given steps (positions, velocities, forces and
    bucket tables), computed once per scene and clear count."""
    key = (sc.name, clear, targets)
    if key not in _TRAJECTORIES:
        op = oracle.Params(*prm.values())
        fo = oracle.Fluid(sc.pos, sc.vel, sc.force, op, box_min=sc.box_min, box_max=sc.box_max, clear_count=clear)
        snaps, done = [], 0
        for target in targets:
            fo.steps(target - done)
            done = target
            snaps.append((target, fo.pos.copy(), fo.vel.copy(), fo.force.copy(), fo.cell_start.copy(),
                          fo.cell_end.copy()))
        _TRAJECTORIES[key] = snaps
    return _TRAJECTORIES[key]


def _step_from_snapshot(oracle, snap, prm, sc, clear, shape):
    """One step (forces only) from an oracle snapshot on both sides: keys, permutation and tables
    bit-exact, density/pressure and forces within TOL."""
    target, pos, vel, force, cs, ce = snap
    op = oracle.Params(*prm.values())
    f = make_fluid(pos, vel, force, prm, sc.box_min, sc.box_max, clear, tables=(cs, ce), shape=shape)
    run_forces_only(f)
    fo2 = oracle.Fluid(pos, vel, force, op, box_min=sc.box_min, box_max=sc.box_max, clear_count=clear,
                       cell_start=cs, cell_end=ce)
    fo2.step(integrate=False)
    _oracle_compare(f, fo2)
    nw, per = force_errors(f.state()["force"], fo2.force)
    assert nw <= TOL and per <= TOL, (target, nw, per)


@pytest.mark.parametrize("shape", list(SHAPES))
@pytest.mark.parametrize("clear", [65536, 262144])
def test_small_dam_break_vs_oracle(oracle, clear, shape):
    """16,384-particle dam break over its 1000 steps: snapshots along the oracle trajectory (steps 0, 1,
    10, 100, 1000) are stepped once on both sides, bucket tables included (stale ranges of the
    reference's partial clear are part of the state)."""
    sc = scenes.config_small()
    import libcl_amd as L
    prm = L.params_from(sc.params)
    for snap in _oracle_snapshots(oracle, sc, prm, clear, (0, 1, 10, 100, 1000)):
        print(f"16k snapshot step {snap[0]} clear {clear} {shape}", flush=True)
        _step_from_snapshot(oracle, snap, prm, sc, clear, shape)


def test_1m_trajectory_snapshots_vs_oracle(oracle):
    """1,048,576 particles over their 200 steps: snapshots of the oracle trajectory at steps 1, 10, 100
    and 200 (the block falling and spreading on the floor) stepped once by the default program."""
    sc = scenes.config_1m()
    import libcl_amd as L
    prm = L.params_from(sc.params)
    for snap in _oracle_snapshots(oracle, sc, prm, 65536, (1, 10, 100, 200)):
        print(f"1M snapshot step {snap[0]}", flush=True)
        _step_from_snapshot(oracle, snap, prm, sc, 65536, "tile_kept")


@pytest.mark.parametrize("spacing", [0.4, 0.3])
@pytest.mark.parametrize("shape", ["tile_kept", "tile_kept2", "tile_kept2_64", "tile_kept64",
                                   "tile_kept8", "tile_list2", "tile_list4_64", "tile64"])
def test_dense_blocks_vs_oracle(oracle, spacing, shape):
    """Blocks denser than the dam break's 0.5h lattice. At 0.4h spacing (about 16 particles per cell and
    190 kept entries per cell) the tile passes test the kept lists in several windows of `depth`
    entries; at 0.3h (about 37 per cell, more than 512 visits per neighbourhood) every particle walks
    its bucket ranges from global memory. One step against the oracle: keys, permutation, tables and
    accepted-visit counts bit-exact, fields within 1e-5."""
    import libcl_amd as L
    h = 0.02
    print(f"dense block spacing {spacing}h {shape}", flush=True)
    pos = scenes.lattice(32, 16, 16, spacing * h, (0.25 * h,) * 3, 0.01 * h, 43)
    vel = np.zeros_like(pos)
    force = np.tile(np.asarray(scenes.GRAVITY_FORCE, np.float32), (pos.shape[0], 1))
    prm = L.params_from(scenes.sph_param_values(h))
    box_min, box_max = (0.0, 0.0, 0.0, 0.0), (1.0, 1.0, 1.0, 0.0)
    f = make_fluid(pos, vel, force, prm, box_min, box_max, shape=shape)
    run_forces_only(f)
    fo = oracle.Fluid(pos, vel, force, oracle.Params(*prm.values()), box_min=box_min, box_max=box_max)
    fo.step(integrate=False)
    _oracle_compare(f, fo)
    nw, per = force_errors(f.state()["force"], fo.force)
    assert nw <= TOL and per <= TOL, (nw, per)


@pytest.mark.parametrize("total", [5120, 262144])
def test_bucket_run_longer_than_a_chunk(oracle, total):
    """5,120 particles packed into one cell (one bucket, so one high-digit run longer than the
    4,096 keys the per-run sort ranks at a time: counted first, then ranked in two chunks), alone and
    among 257,024 particles of a dam break (the run pass then writes no sorted state and the copy pass
    does). Keys, permutation, tables and counts bit-exact; density and pressure within 1e-5."""
    import libcl_amd as L
    h = 0.02
    print(f"long bucket run, {total} particles", flush=True)
    packed = scenes.lattice(16, 16, 20, 0.024 * h, (10.25 * h,) * 3, 0.002 * h, 47)
    if total > packed.shape[0]:
        sc = scenes.dam_break(64, 64, 64, 2.0, 7, 1)
        pos = np.concatenate([sc.pos[:total - packed.shape[0]], packed])
        box_min, box_max, prm = sc.box_min, sc.box_max, L.params_from(sc.params)
    else:
        pos = packed
        box_min, box_max = (0.0, 0.0, 0.0, 0.0), (1.0, 1.0, 1.0, 0.0)
        prm = L.params_from(scenes.sph_param_values(h))
    vel = np.zeros_like(pos)
    force = np.tile(np.asarray(scenes.GRAVITY_FORCE, np.float32), (pos.shape[0], 1))
    f = make_fluid(pos, vel, force, prm, box_min, box_max)
    run_forces_only(f)
    fo = oracle.Fluid(pos, vel, force, oracle.Params(*prm.values()), box_min=box_min, box_max=box_max)
    fo.step(integrate=False)
    _oracle_compare(f, fo)


@pytest.mark.parametrize("items", [4, 16])
def test_run_blocks_over_several_tiles(oracle, items):
    """The run pass's threads each take a contiguous range of hash tiles when there are more tiles than
    threads: 786,432 particles in tiles of 1,024 keys (two tiles per thread), and 1,048,576 in tiles of
    4,096 (one tile for every other thread). Keys, permutation, tables and counts bit-exact; density and
    pressure within 1e-5."""
    import libcl_amd as L
    sc = scenes.config_1m()
    count = 786432 if items == 4 else sc.n
    print(f"run pass over tiles of {256 * items} keys, {count} particles", flush=True)
    pos, vel, force = sc.pos[:count].copy(), sc.vel[:count].copy(), sc.force[:count].copy()
    prm = L.params_from(sc.params)
    f = make_fluid(pos, vel, force, prm, sc.box_min, sc.box_max)
    assert f.setTuning(sort_items=items) == 1, L.last_error()
    run_forces_only(f)
    fo = oracle.Fluid(pos, vel, force, oracle.Params(*prm.values()), box_min=sc.box_min, box_max=sc.box_max)
    fo.step(integrate=False)
    _oracle_compare(f, fo)


@pytest.mark.parametrize("shape", ["tile_list", "tile_list8", "tile", "tile_kept", "tile_kept2", "list"])
def test_1m_step_vs_oracle(oracle, shape):
    """1,048,576 particles in [0,4]^3: keys, permutation, bucket tables bit-exact; density, pressure
    and forces of one step within 1e-5."""
    print(f"1M step {shape}", flush=True)
    sc = scenes.config_1m()
    import libcl_amd as L
    prm = L.params_from(sc.params)
    op = oracle.Params(*prm.values())
    f = make_fluid(sc.pos, sc.vel, sc.force, prm, sc.box_min, sc.box_max, shape=shape)
    run_forces_only(f)
    fo = oracle.Fluid(sc.pos, sc.vel, sc.force, op, box_min=sc.box_min, box_max=sc.box_max)
    fo.step(integrate=False)
    _oracle_compare(f, fo)
    nw, per = force_errors(f.state()["force"], fo.force)
    assert nw <= TOL and per <= TOL, (nw, per)


@pytest.mark.parametrize("count", [262144 - 256, 262144])  # counts are multiples of cLocalSize = 256
def test_reorder_copy_forms_at_their_boundary(oracle, count):
    """The reorder stage gathers the sorted state inside its bounds pass below 2^18 particles and
    copies it from the caller's order through the inverse permutation from 2^18 up. On both sides of
    that boundary: keys, permutation and bucket tables bit-exact, sorted positions and velocities
    (a pure copy) bit-exact in xyz, density and pressure within 1e-5."""
    print(f"reorder boundary n={count}", flush=True)
    import libcl_amd as L
    sc = scenes.dam_break(64, 64, 64, 2.0, 7, 1)
    rng = np.random.default_rng(count)
    pos, force = sc.pos[:count].copy(), sc.force[:count].copy()
    vel = np.zeros_like(pos)
    vel[:, :3] = rng.uniform(-0.5, 0.5, (count, 3)).astype(np.float32)
    prm = L.params_from(sc.params)
    op = oracle.Params(*prm.values())
    f = make_fluid(pos, vel, force, prm, sc.box_min, sc.box_max)
    run_forces_only(f)
    fo = oracle.Fluid(pos, vel, force, op, box_min=sc.box_min, box_max=sc.box_max)
    fo.step(integrate=False)
    it = _oracle_compare(f, fo)
    np.testing.assert_array_equal(it["spos"][:, :3], fo.spos[:, :3])
    np.testing.assert_array_equal(it["svel"][:, :3], fo.svel[:, :3])


def test_multistep_tables_and_graph(oracle):
    """20 chained steps at 4,096 particles: eager and hipGraph replay give bit-identical state; the
    bucket tables (with their stale entries) match the oracle fed the same per-step positions."""
    import torch
    import libcl_amd as L
    sc = scenes.dam_break(16, 16, 16, 1.0, seed=31, steps=20)
    prm = L.params_from(sc.params)
    fa = make_fluid(sc.pos, sc.vel, sc.force, prm, sc.box_min, sc.box_max)
    fb = make_fluid(sc.pos, sc.vel, sc.force, prm, sc.box_min, sc.box_max)
    assert fb.setUseGraph(True) == 1
    stream = torch.cuda.Stream()
    op = oracle.Params(*prm.values())
    for step in range(20):
        before = fa.state()
        tables = fa.internals()
        assert fa.compute() == 1
        assert fb.compute(stream) == 1, L.last_error()
        stream.synchronize()
        L.load().mfl_device_finish()
        sa, sb = fa.state(), fb.state()
        for k in ("pos", "vel", "force"):
            np.testing.assert_array_equal(sa[k], sb[k])
        fo = oracle.Fluid(before["pos"], before["vel"], before["force"], op, box_min=sc.box_min,
                          box_max=sc.box_max, cell_start=tables["cell_start"], cell_end=tables["cell_end"])
        fo.step(integrate=False)
        _oracle_compare(fa, fo, check_fields=(step % 5 == 0))


def test_graph_recaptures_on_changes():
    """hipGraph replay against eager launches, step for step, while Params, box, gravity, bucket clear,
    launch shape change between steps: every change must reach the replayed
    graph (bit-identical state after every step)."""
    import torch
    import libcl_amd as L
    sc = scenes.dam_break(16, 16, 16, 1.0, seed=33, steps=12)
    prm = L.params_from(sc.params)
    fa = make_fluid(sc.pos, sc.vel, sc.force, prm, sc.box_min, sc.box_max)
    fb = make_fluid(sc.pos, sc.vel, sc.force, prm, sc.box_min, sc.box_max)
    assert fb.setUseGraph(True) == 1
    stream = torch.cuda.Stream()
    prm2 = L.params_from(dict(sc.params, viscosity=9.0, stiffness=2.0))
    changes = {
        2: lambda f: f.setParameters(prm2),
        4: lambda f: f.setBox((0.0, 0.0, 0.0, 0.0), (0.7, 1.0, 1.0, 0.0)),
        6: lambda f: f.setGravity((0.0, 0.5, 1.0, 0.0)),
        8: lambda f: f.setBucketClear(262144),
        10: lambda f: f.setTuning(neighbour_pass="tile_kept", density_lanes=8, tile_depth=256),
    }
    for step in range(12):
        if step in changes:
            for f in (fa, fb):
                assert changes[step](f) in (None, 1), L.last_error()
        assert fa.compute(stream) == 1, L.last_error()
        assert fb.compute(stream) == 1, L.last_error()
        stream.synchronize()
        sa, sb = fa.state(), fb.state()
        for k in ("pos", "vel", "force"):
            np.testing.assert_array_equal(sa[k], sb[k], err_msg=f"step {step} {k}")


def test_integrate_callback_path(oracle):
    """libCL/TEST/main.cpp:246-343: an EVT_INTEGRATE handler enqueues clIntegrateForce, clGravity,
    clIntegrateVelocity itself (no clip box)."""
    import libcl_amd as L
    z = load_step_case("random512")
    prm = params_of(z["params"])
    f = make_fluid(z["pos"], z["vel"], z["force"], prm, z["box_min"], z["box_max"])
    calls = []

    def handler(prog, stream):
        calls.append(1)
        for k in ("clIntegrateForce", "clGravity", "clIntegrateVelocity"):
            assert prog.enqueueKernel(k, stream) == 1

    assert f.addEventHandler(f.EVT_INTEGRATE, handler) == 1
    assert f.compute() == 1
    L.load().mfl_device_finish()
    assert calls == [1]
    fo = oracle.Fluid(z["pos"], z["vel"], z["force"], oracle.Params(*prm.values()), box_min=z["box_min"],
                      box_max=z["box_max"])
    fo.step(integrate=False)
    oracle.integrate_force(fo.vel, fo.force, fo.params)
    oracle.gravity(fo.force)
    oracle.integrate_velocity(fo.pos, fo.vel, fo.params)
    st = f.state()
    np.testing.assert_array_equal(st["force"], fo.force)
    assert np.max(np.abs(st["vel"] - fo.vel)) <= 1e-5 * (np.abs(fo.vel).max() + 1.0)
    assert np.max(np.abs(st["pos"] - fo.pos)) <= 1e-6


def test_stage_kernels_individually(oracle):
    """Each named kernel reachable through getKernel(name) matches its oracle restatement."""
    import libcl_amd as L
    z = load_step_case("compressed512")
    prm = params_of(z["params"])
    op = oracle.Params(*prm.values())
    f = make_fluid(z["pos"], z["vel"], z["force"], prm, z["box_min"], z["box_max"])
    for k in ("clHash",):
        assert f.enqueueKernel(k) == 1
    L.load().mfl_device_finish()
    it = f.internals()
    cell, index = oracle.hash_cells(z["pos"], float(prm.cellSize))
    np.testing.assert_array_equal(it["cell"], cell)
    np.testing.assert_array_equal(it["index"], index)
    assert f.enqueueKernel("clNoSuchKernel") == 0
    assert "clNoSuchKernel" in L.last_error()
    # clip box alone on a force field, in sorted order through the index buffer
    run_forces_only(f)
    it = f.internals()
    before = f.state()["force"]
    assert f.enqueueKernel("clClipBox") == 1
    L.load().mfl_device_finish()
    ref = before.copy()
    oracle.clip_box(it["spos"], it["svel"], ref, it["index"], op, z["box_min"], z["box_max"])
    np.testing.assert_allclose(f.state()["force"], ref, rtol=1e-6, atol=1e-6 * np.abs(ref).max())


def test_particle_count_rules():
    """N must be a non-zero multiple of cLocalSize = 256 (every NDRange uses local size 256)."""
    import libcl_amd as L
    f = L.oclFluid3D()
    assert f.compile() == 1
    assert f.getParticleCount() == 256
    assert f.setParticleCount(300) == 1
    assert f.compute() == 0 and "multiple of 256" in L.last_error()
    assert f.setParticleCount(512) == 1
    import torch
    small = torch.zeros(256, 4, dtype=torch.float32, device="cuda")
    assert f.setPositionBuffer(small) == 0
    big = torch.zeros(512, 4, dtype=torch.float32, device="cuda")
    assert f.setPositionBuffer(big) == 1
    assert f.compute() == 1
    L.load().mfl_device_finish()


def test_traversal_orders_bit_identical():
    """With the recorded-visit passes the visit order changes no result: sorted, caller and Morton
    traversal give bit-identical state over several steps (each particle's lane partial sums take the
    same visits in the same order whatever wave computes them)."""
    import libcl_amd as L
    sc = scenes.dam_break(32, 16, 16, 1.0, seed=41, steps=5)
    prm = L.params_from(sc.params)
    states = []
    for order in ("sorted", "caller", "morton"):
        f = make_fluid(sc.pos, sc.vel, sc.force, prm, sc.box_min, sc.box_max, shape="list")
        assert f.setTraversal(order, 2) == 1
        for _ in range(5):
            assert f.compute() == 1, L.last_error()
        L.load().mfl_device_finish()
        states.append((f.state(), f.internals()))
    for st, it in states[1:]:
        for k in ("pos", "vel", "force"):
            np.testing.assert_array_equal(st[k], states[0][0][k])
        for k in ("spos", "svel", "cell_start", "cell_end", "index"):
            np.testing.assert_array_equal(it[k], states[0][1][k])


def test_set_traversal_rebuilds_the_morton_order():
    """setTraversal makes the next step rebuild the Morton order: a program whose order was built from
    other positions (the same particles in a shuffled caller order) and then had its state rewritten
    steps bit-identically to a fresh program from that state (with the tile passes the FP32 summation
    order follows the traversal order, so a stale order shows)."""
    import libcl_amd as L
    sc = scenes.dam_break(32, 16, 16, 1.0, seed=41, steps=5)
    prm = L.params_from(sc.params)
    empty = np.full(262144, 0xFFFFFFFF, np.uint32)
    perm = np.random.default_rng(7).permutation(sc.n)
    used = make_fluid(sc.pos[perm], sc.vel[perm], sc.force[perm], prm, sc.box_min, sc.box_max, tables=(empty, empty))
    assert used.setTraversal("morton", 16) == 1
    for _ in range(3):
        assert used.compute() == 1, L.last_error()
    assert used.upload(sc.pos, sc.vel, sc.force) == 1
    assert used.bfCellStart.write(empty) == 1 and used.bfCellEnd.write(empty) == 1
    assert used.setTraversal("morton", 16) == 1
    fresh = make_fluid(sc.pos, sc.vel, sc.force, prm, sc.box_min, sc.box_max, tables=(empty, empty))
    assert fresh.setTraversal("morton", 16) == 1
    for f in (used, fresh):
        for _ in range(6):
            assert f.compute() == 1, L.last_error()
    L.load().mfl_device_finish()
    a, b = used.state(), fresh.state()
    for k in ("pos", "vel", "force"):
        np.testing.assert_array_equal(a[k], b[k])
    ia, ib = used.internals(), fresh.internals()
    for k in ("spos", "svel", "cell_start", "cell_end", "index"):
        np.testing.assert_array_equal(ia[k], ib[k])


@pytest.mark.parametrize("order", ["sorted", "caller", "morton"])
@pytest.mark.parametrize("shape", ["tile_list", "tile_list8", "tile", "tile64", "tile_list64", "tile_list4_64",
                                   "tile_list4_ch2", "tile_kept", "tile_kept8", "tile_kept64", "tile_kept2",
                                   "tile_kept2_64", "tile_list2"])
def test_tile_passes_match_list_passes(order, shape):
    """The LDS-tile passes take the same visits as the recorded-visit passes; only the FP32 summation
    order differs (a particle's kept entries are split among its lanes by position in the pruned
    tile, which depends on the wave's other particles). One step from identical state: identical keys,
    permutation and tables, densities and forces within 1e-6."""
    import libcl_amd as L
    sc = scenes.dam_break(32, 16, 16, 1.0, seed=41, steps=5)
    prm = L.params_from(sc.params)
    ref = make_fluid(sc.pos, sc.vel, sc.force, prm, sc.box_min, sc.box_max, shape="list")
    f = make_fluid(sc.pos, sc.vel, sc.force, prm, sc.box_min, sc.box_max, shape=shape)
    assert f.setTraversal(order, 1) == 1
    run_forces_only(ref)
    run_forces_only(f)
    a, b = ref.internals(), f.internals()
    for k in ("cell", "index", "cell_start", "cell_end"):
        np.testing.assert_array_equal(b[k], a[k])
    np.testing.assert_allclose(b["spos"][:, 3], a["spos"][:, 3], rtol=1e-6)
    np.testing.assert_array_equal(f.bfNeighbourCount.read(np.uint32), ref.bfNeighbourCount.read(np.uint32))
    nw, per = force_errors(f.state()["force"], ref.state()["force"])
    assert nw <= 1e-6 and per <= 1e-6, (nw, per)


def test_autotune_keeps_state_and_picks_a_valid_shape():
    import libcl_amd as L
    sc = scenes.config_small()
    prm = L.params_from(sc.params)
    f = make_fluid(sc.pos, sc.vel, sc.force, prm, sc.box_min, sc.box_max)
    for _ in range(3):
        assert f.compute() == 1
    L.load().mfl_device_finish()
    before, tables = f.state(), f.internals()
    res = f.autotune(steps=2)
    assert res is not None, L.last_error()
    best, ms = res
    print("autotune", best, ms, flush=True)
    assert ms > 0 and f.getTuning() == best
    after, tables2 = f.state(), f.internals()
    for k in ("pos", "vel", "force"):
        np.testing.assert_array_equal(after[k], before[k])
    for k in ("cell_start", "cell_end"):
        np.testing.assert_array_equal(tables2[k], tables[k])
    assert f.compute() == 1, L.last_error()
    L.load().mfl_device_finish()
    assert f.setTuning(tile_depth=100) == 0 and "tile_depth" in L.last_error()
    assert f.setTuning(neighbour_pass="tile", tile_depth=1024, block_waves=1) == 0 and "512" in L.last_error()


def test_graph_survives_autotune_at_1m():
    """hipGraph replay after an autotune at 1,048,576 particles. The autotuner's trials re-size the
    sort's temporaries (a 4-keys-per-thread trial doubles the tile count at this size), so a graph
    captured before it must not be replayed on the old buffers: compute, autotune, compute in graph
    mode equals the same steps launched eagerly, bit for bit (recorded-visit passes, whose results do
    not depend on the traversal order the autotuner invalidates)."""
    import torch
    import libcl_amd as L
    print("graph + autotune at 1M", flush=True)
    sc = scenes.config_1m()
    prm = L.params_from(sc.params)
    fa = make_fluid(sc.pos, sc.vel, sc.force, prm, sc.box_min, sc.box_max, shape="list")
    fb = make_fluid(sc.pos, sc.vel, sc.force, prm, sc.box_min, sc.box_max, shape="list")
    shape = fb.getTuning()
    assert fb.setUseGraph(True) == 1
    stream = torch.cuda.Stream()
    assert fa.compute(stream) == 1 and fb.compute(stream) == 1, L.last_error()
    stream.synchronize()
    assert fb.autotune(steps=1, stream=stream) is not None, L.last_error()
    assert fb.setTuning(**shape) == 1, L.last_error()  # the shape the first graph was captured with
    for _ in range(2):
        assert fa.compute(stream) == 1 and fb.compute(stream) == 1, L.last_error()
    stream.synchronize()
    sa, sb = fa.state(), fb.state()
    for k in ("pos", "vel", "force"):
        np.testing.assert_array_equal(sa[k], sb[k], err_msg=k)
    ia, ib = fa.internals(), fb.internals()
    for k in ("cell", "index", "cell_start", "cell_end", "spos", "svel"):
        np.testing.assert_array_equal(ia[k], ib[k], err_msg=k)
