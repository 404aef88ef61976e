"""Device memory of the program objects is returned when they are destroyed: create / use / destroy
loops of every program type leave the device's free memory where the first round left it."""
import gc

import numpy as np
import pytest

from libcl_amd import scenes

pytestmark = pytest.mark.gpu

SLACK = 16 << 20  # bytes: events, kernel-argument pools and other runtime-side allocations


def _free_bytes():
    import torch
    import libcl_amd as L
    L.load().mfl_device_finish()
    gc.collect()
    torch.cuda.synchronize()
    return torch.cuda.mem_get_info()[0]


def _fluid_round():
    import libcl_amd as L
    sc = scenes.dam_break(64, 64, 64, 2.0, 7, 1)
    f = L.oclFluid3D()
    assert f.compile() == 1 and f.setParticleCount(sc.n) == 1, L.last_error()
    f.setParameters(L.params_from(sc.params))
    assert f.setBox(sc.box_min, sc.box_max) == 1 and f.upload(sc.pos, sc.vel, sc.force) == 1
    assert f.setUseGraph(True) == 1
    import torch
    s = torch.cuda.Stream()
    for _ in range(2):
        assert f.compute(s) == 1, L.last_error()
    s.synchronize()
    del f


def _batch_round():
    import libcl_amd as L
    scs = scenes.batched_scenes(num=24, seed=5, steps=2)
    b = L.FluidBatch([s.n for s in scs])
    pos, vel, force = (np.concatenate([getattr(x, k) for x in scs]) for k in ("pos", "vel", "force"))
    assert b.setParameters(L.params_from(scs[0].params)) == 1 and b.upload(pos, vel, force) == 1
    for _ in range(2):
        assert b.compute() == 1, L.last_error()
    del b


def _next_round():
    import libcl_amd as L
    f = L.oclFluid3Dnext()
    assert f.compile() == 1 and f.setParticleCount(65536) == 1, L.last_error()
    rng = np.random.default_rng(3)
    st = np.zeros((65536, 4), np.float32)
    st[:, 0] = rng.uniform(0, 1000, 65536)
    st[:, 1] = rng.uniform(0, 700, 65536)
    assert f.upload(st) == 1
    assert f.compute() == 1, L.last_error()
    del f


def _bvh_round():
    import torch
    import libcl_amd as L
    rng = np.random.default_rng(4)
    vtx = torch.tensor(np.c_[rng.uniform(-1, 1, (30000, 3)), np.zeros(30000)].astype(np.float32), device="cuda")
    idx = torch.tensor(rng.integers(0, 30000, 30000).astype(np.int32), device="cuda")
    b = L.oclBvhTrimesh()
    assert b.compile() == 1 and b.compute(vtx, idx) == 1, L.last_error()
    del b


def _adjoint_round():
    import libcl_amd as L
    sc, _ = scenes.config_grad((16, 16, 8))
    f = L.oclFluid3D()
    assert f.compile() == 1 and f.setParticleCount(sc.n) == 1
    f.setParameters(L.params_from(sc.params))
    assert f.setBox(sc.box_min, sc.box_max) == 1 and f.upload(sc.pos, sc.vel, sc.force) == 1
    a = L.Adjoint(f, 8, 4)
    assert a.forward() == 1 and a.backward() == 1, L.last_error()
    del a, f


@pytest.mark.parametrize("make", [_fluid_round, _batch_round, _next_round, _bvh_round, _adjoint_round],
                         ids=["fluid3d", "batch", "fluid3dnext", "bvh", "adjoint"])
def test_destroy_returns_device_memory(make):
    print(f"memory round {make.__name__}", flush=True)
    make()
    first = _free_bytes()
    for _ in range(3):
        make()
    last = _free_bytes()
    assert first - last <= SLACK, f"{(first - last) / 2**20:.1f} MiB not returned after 3 create/destroy rounds"
