"""GPU parity of the batched scenes: every scene of a batch matches the same scene stepped alone by
the single-scene program (oclFluid3D::compute, libCL/phys/oclFluid3D.cpp:293-342), bit for bit."""
import numpy as np
import pytest

from libcl_amd import scenes

pytestmark = pytest.mark.gpu


# batch launch shapes under test, and the single-scene shape each is compared with
BATCH_SHAPES = {
    "list": ({"neighbour_pass": "list"}, {"neighbour_pass": "list"}),
    "tile_list": ({"neighbour_pass": "tile_list"},
                  {"neighbour_pass": "tile_list", "density_lanes": 4, "tile_depth": 192, "block_waves": 2,
                   "force_chunks": 2}),
    "tile_kept": ({"neighbour_pass": "tile_kept"},
                  {"neighbour_pass": "tile_kept", "density_lanes": 4, "tile_depth": 192, "block_waves": 2,
                   "force_chunks": 2}),
    # 2 lanes per particle: four 16-lane segments of 8 particles per wave
    "tile_kept2": ({"neighbour_pass": "tile_kept", "density_lanes": 2, "tile_depth": 96},
                   {"neighbour_pass": "tile_kept", "density_lanes": 2, "tile_depth": 96, "block_waves": 2,
                    "force_chunks": 2}),
    # 2 lanes, 64-entry windows (dense cells take several)
    "tile_kept2_64": ({"neighbour_pass": "tile_kept", "density_lanes": 2, "tile_depth": 64},
                      {"neighbour_pass": "tile_kept", "density_lanes": 2, "tile_depth": 64, "block_waves": 2,
                       "force_chunks": 2}),
    # 8 lanes per particle: one 64-lane segment of 8 particles per wave
    "tile_kept8": ({"neighbour_pass": "tile_kept", "density_lanes": 8, "tile_depth": 256},
                   {"neighbour_pass": "tile_kept", "density_lanes": 8, "tile_depth": 256, "block_waves": 2,
                    "force_chunks": 2}),
    "default": (None, {"neighbour_pass": "tile_kept", "density_lanes": 4, "tile_depth": 192, "block_waves": 2,
                       "force_chunks": 2}),
}


def _plain(sc, prm, steps, tuning):
    import libcl_amd as L
    f = L.oclFluid3D()
    assert f.compile() == 1
    assert f.setTuning(**tuning) == 1, L.last_error()
    # the batch's traversal (the single-scene default for this size would be the sorted order, which
    # splits visits over a particle's lanes differently in the tile passes)
    assert f.setTraversal("morton", 16) == 1
    assert f.setParticleCount(sc.n) == 1
    f.setParameters(prm)
    f.setBox(sc.box_min, sc.box_max)
    assert f.upload(sc.pos, sc.vel, sc.force) == 1
    for _ in range(steps):
        assert f.compute() == 1, L.last_error()
    L.load().mfl_device_finish()
    return f


@pytest.mark.parametrize("graph", [False, True])
@pytest.mark.parametrize("shape", list(BATCH_SHAPES))
def test_batch_matches_single_scene_runs(graph, shape):
    """Every scene of the batch equals the same scene run alone by a single-scene program with the same
    launch shape (LIST: group traversal; TILE_LIST / TILE_KEPT: 4-lane tiles in a Morton order refreshed
    every 16 steps, recorded-visit or kept-entry forces), bit for bit."""
    import torch
    import libcl_amd as L
    from libcl_amd import _native as N
    scs = scenes.batched_scenes(num=12, seed=5, steps=6)
    prm = L.params_from(scs[0].params)
    counts = [s.n for s in scs]
    b = L.FluidBatch(counts)
    batch_tuning, single_tuning = BATCH_SHAPES[shape]
    if batch_tuning is not None:  # None: the batch's own default launch shape, never set
        assert b.setTuning(**batch_tuning) == 1, L.last_error()
    assert b.setParameters(prm) == 1
    assert b.setBox((0, 0, 0, 0), (1, 1, 1, 0)) == 1
    assert b.upload(np.concatenate([s.pos for s in scs]), np.concatenate([s.vel for s in scs]),
                    np.concatenate([s.force for s in scs])) == 1
    assert b.setUseGraph(graph) == 1
    stream = torch.cuda.Stream()
    steps = 6
    for _ in range(steps):
        assert b.compute(stream) == 1, L.last_error()
    stream.synchronize()
    st = b.state()
    index = b.read(N.MFL_INDEX, np.uint32)
    cell = b.read(N.MFL_CELL, np.uint32)
    spos = b.read(N.MFL_SORTED_POSITION, np.float32, (-1, 4))
    for s, sc in enumerate(scs):
        print(f"scene {s} n={sc.n} {shape}", flush=True)
        f = _plain(sc, prm, steps, single_tuning)
        ref = f.state()
        it = f.internals()
        lo, n = int(b.offsets[s]), sc.n
        for k in ("pos", "vel", "force"):
            np.testing.assert_array_equal(st[k][lo:lo + n], ref[k])
        np.testing.assert_array_equal(index[lo:lo + n] - lo, it["index"])
        np.testing.assert_array_equal(cell[lo:lo + n] & 0x3FFFF, it["cell"])
        assert np.all(cell[lo:lo + n] >> 18 == s)
        np.testing.assert_array_equal(spos[lo:lo + n], it["spos"])
        cs, ce = b.scene_tables(s)
        np.testing.assert_array_equal(cs, it["cell_start"])
        np.testing.assert_array_equal(ce, it["cell_end"])


def test_batch_per_scene_params_and_boxes():
    """Scenes with their own Params (cell size, viscosity, stiffness, time step) and clip boxes: each
    equals the scene run alone with those Params and that box, bit for bit (default launch shape)."""
    import torch
    import libcl_amd as L
    scs = scenes.batched_scenes(num=6, seed=9, steps=5)
    prms, boxes = [], []
    for s, sc in enumerate(scs):
        h = 0.02 * (1.0 + 0.15 * (s % 3))
        v = dict(scenes.sph_param_values(h), viscosity=4.0 + s, stiffness=1.0 + 0.5 * s, deltaTime=1e-4 * (1 + s % 2))
        prms.append(L.params_from(v))
        boxes.append(((0.0, 0.0, 0.0, 0.0), (0.6 + 0.1 * s, 1.0, 0.8, 0.0)))
    b = L.FluidBatch([sc.n for sc in scs])
    for s in range(len(scs)):
        assert b.setParameters(prms[s], scene=s) == 1, L.last_error()
        assert b.setBox(*boxes[s], scene=s) == 1, L.last_error()
    assert b.upload(np.concatenate([sc.pos for sc in scs]), np.concatenate([sc.vel for sc in scs]),
                    np.concatenate([sc.force for sc in scs])) == 1
    stream = torch.cuda.Stream()
    steps = 5
    for _ in range(steps):
        assert b.compute(stream) == 1, L.last_error()
    stream.synchronize()
    st = b.state()
    for s, sc in enumerate(scs):
        print(f"scene {s} n={sc.n} h={prms[s].cellSize}", flush=True)
        f = L.oclFluid3D()
        assert f.compile() == 1 and f.setTraversal("morton", 16) == 1 and f.setParticleCount(sc.n) == 1
        assert f.setTuning(neighbour_pass="tile_kept", density_lanes=4, tile_depth=192, block_waves=2,
                           force_chunks=2) == 1  # the batch's default shape and traversal
        f.setParameters(prms[s])
        assert f.setBox(*boxes[s]) == 1
        assert f.upload(sc.pos, sc.vel, sc.force) == 1
        for _ in range(steps):
            assert f.compute() == 1, L.last_error()
        L.load().mfl_device_finish()
        ref = f.state()
        lo = int(b.offsets[s])
        for k in ("pos", "vel", "force"):
            np.testing.assert_array_equal(st[k][lo:lo + sc.n], ref[k])


@pytest.mark.parametrize("sides", [[(8, 8, 4)], [(8, 8, 4), (16, 8, 4)], [(16, 16, 16)] * 3])
def test_batch_few_scenes(sides):
    """One, two and three scenes (256 to 4,096 particles; one scene has no scene bits in its keys):
    each equals its single-scene run bit for bit (default launch shape)."""
    import torch
    import libcl_amd as L
    scs = [scenes.dam_break(nx, ny, nz, 1.0, seed=40 + k, steps=4) for k, (nx, ny, nz) in enumerate(sides)]
    prm = L.params_from(scs[0].params)
    b = L.FluidBatch([sc.n for sc in scs])
    assert b.setParameters(prm) == 1 and b.setBox((0, 0, 0, 0), (1, 1, 1, 0)) == 1
    assert b.upload(np.concatenate([sc.pos for sc in scs]), np.concatenate([sc.vel for sc in scs]),
                    np.concatenate([sc.force for sc in scs])) == 1
    stream = torch.cuda.Stream()
    for _ in range(4):
        assert b.compute(stream) == 1, L.last_error()
    stream.synchronize()
    st = b.state()
    for s, sc in enumerate(scs):
        print(f"scene {s} of {len(scs)} n={sc.n}", flush=True)
        ref = _plain(sc, prm, 4, BATCH_SHAPES["default"][1]).state()
        lo = int(b.offsets[s])
        for k in ("pos", "vel", "force"):
            np.testing.assert_array_equal(st[k][lo:lo + sc.n], ref[k])


def test_batch_rejects_bad_counts():
    import libcl_amd as L
    with pytest.raises(RuntimeError, match="multiple of 256"):
        L.FluidBatch([256, 300])
    with pytest.raises(RuntimeError):
        L.FluidBatch([])


def test_batch_autotune_keeps_state():
    import torch
    import libcl_amd as L
    from libcl_amd import _native as N
    scs = scenes.batched_scenes(num=8, seed=7, steps=4)
    prm = L.params_from(scs[0].params)
    b = L.FluidBatch([s.n for s in scs])
    assert b.setParameters(prm) == 1 and b.setBox((0, 0, 0, 0), (1, 1, 1, 0)) == 1
    assert b.upload(np.concatenate([s.pos for s in scs]), np.concatenate([s.vel for s in scs]),
                    np.concatenate([s.force for s in scs])) == 1
    stream = torch.cuda.Stream()
    for _ in range(2):
        assert b.compute(stream) == 1
    stream.synchronize()
    before = b.state()
    tables = b.read(N.MFL_CELL_START, np.uint32)
    res = b.autotune(steps=2, stream=stream)
    assert res is not None, L.last_error()
    print("batch autotune", res, flush=True)
    assert b.getTuning() == res[0]
    after = b.state()
    for k in ("pos", "vel", "force"):
        np.testing.assert_array_equal(after[k], before[k])
    np.testing.assert_array_equal(b.read(N.MFL_CELL_START, np.uint32), tables)
    assert b.compute(stream) == 1, L.last_error()
    stream.synchronize()
    assert b.setTuning(neighbour_pass="tile") == 0 and "TILE_LIST" in L.last_error()


def test_batched_workload_vs_oracle():
    """All 256 scenes of the batched workload (2,048-65,536 particles each, 5.1 M in all) over their
    500 steps. From the batch's own state at steps 20 and 500 (positions, velocities, forces and each
    scene's bucket tables) one batched step and one oracle step per scene: keys, permutation and
    tables bit-exact, density and pressure within 1e-5, and the step's velocities and positions
    within the change a 1e-5 force error (floored by the scene's RMS force) makes in one step."""
    import torch
    import libcl_amd as L
    from libcl_amd import _native as N
    from oracle import oracle as O
    TOL = 1e-5
    scs = scenes.batched_scenes(num=256, seed=5, steps=500)
    prm = L.params_from(scs[0].params)
    op = O.Params(*prm.values())
    b = L.FluidBatch([s.n for s in scs])
    assert b.setParameters(prm) == 1 and b.setBox((0, 0, 0, 0), (1, 1, 1, 0)) == 1
    assert b.upload(np.concatenate([s.pos for s in scs]), np.concatenate([s.vel for s in scs]),
                    np.concatenate([s.force for s in scs])) == 1
    assert b.setUseGraph(True) == 1
    stream = torch.cuda.Stream()
    dt = float(prm.deltaTime)
    done = 0
    for target in (20, 500):
        for _ in range(target - done):
            assert b.compute(stream) == 1, L.last_error()
        stream.synchronize()
        done = target
        before = b.state()
        cs_all = b.read(N.MFL_CELL_START, np.uint32).reshape(len(scs), -1).copy()
        ce_all = b.read(N.MFL_CELL_END, np.uint32).reshape(len(scs), -1).copy()
        assert b.compute(stream) == 1, L.last_error()
        stream.synchronize()
        done += 1
        after = b.state()
        index = b.read(N.MFL_INDEX, np.uint32)
        cell = b.read(N.MFL_CELL, np.uint32)
        spos = b.read(N.MFL_SORTED_POSITION, np.float32, (-1, 4))
        svel = b.read(N.MFL_SORTED_VELOCITY, np.float32, (-1, 4))
        worst = 0.0
        for s, sc in enumerate(scs):
            lo, n = int(b.offsets[s]), sc.n
            sl = slice(lo, lo + n)
            fo = O.Fluid(before["pos"][sl], before["vel"][sl], before["force"][sl], op, box_min=(0, 0, 0, 0),
                         box_max=(1, 1, 1, 0), cell_start=cs_all[s], cell_end=ce_all[s])
            fo.step(integrate=False)
            np.testing.assert_array_equal(cell[sl] & 0x3FFFF, fo.cell, err_msg=f"step {target} scene {s}")
            np.testing.assert_array_equal(index[sl] - lo, fo.index, err_msg=f"step {target} scene {s}")
            cs, ce = b.scene_tables(s)
            np.testing.assert_array_equal(cs, fo.cell_start, err_msg=f"step {target} scene {s}")
            np.testing.assert_array_equal(ce, fo.cell_end, err_msg=f"step {target} scene {s}")
            rel = np.abs(spos[sl, 3] - fo.spos[:, 3]) / np.abs(fo.spos[:, 3])
            p = fo.svel[:, 3].astype(np.float64)
            prel = np.abs(svel[sl, 3] - p) / (np.abs(p) + np.sqrt(np.mean(p ** 2)))
            assert rel.max() <= TOL and prel.max() <= TOL, (target, s, rel.max(), prel.max())
            # the rest of the oracle step: clip box, integrate force, gravity, integrate velocity
            O.clip_box(fo.spos, fo.svel, fo.force, fo.index, op, (0, 0, 0, 0), (1, 1, 1, 0))
            fclip = fo.force.copy()
            O.integrate_force(fo.vel, fo.force, op)
            O.integrate_velocity(fo.pos, fo.vel, op)
            fmag = np.linalg.norm(fclip[:, :3].astype(np.float64), axis=1)
            vtol = TOL * (fmag + np.sqrt(np.mean(fmag ** 2))) * dt + 4e-7 * (np.abs(fo.vel[:, :3]).max(1) + 1e-6)
            dv = np.linalg.norm(after["vel"][sl, :3] - fo.vel[:, :3], axis=1)
            assert np.all(dv <= vtol), (target, s, (dv / vtol).max())
            ptol = vtol * dt + 2.5e-7 * np.abs(fo.pos[:, :3]).max(1)
            assert np.all(np.linalg.norm(after["pos"][sl, :3] - fo.pos[:, :3], axis=1) <= ptol), (target, s)
            np.testing.assert_array_equal(after["force"][sl], np.tile(scenes.GRAVITY_FORCE, (n, 1)))
            worst = max(worst, float(rel.max()), float((dv / vtol).max()) * TOL)
        print(f"C4 step {target}: 256 scenes match the oracle (worst {worst:.2e})", flush=True)
