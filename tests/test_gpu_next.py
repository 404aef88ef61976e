"""GPU parity of oclFluid3Dnext (mfl_fluid3dnext_*) against the oracle: keys, permutation, sorted state
and both cell tables bit-exact; pressure, force and the updated state within 1e-5 (the GPU's powf and
the C library's differ in the last ulp); every step starts both from the same state."""
import numpy as np
import pytest

from libcl_amd import scenes

pytestmark = pytest.mark.gpu


def _make(n):
    import libcl_amd as L
    f = L.oclFluid3Dnext()
    assert f.compile() == 1, L.last_error()
    assert f.setParticleCount(n) == 1, L.last_error()
    return f


def _states():
    spread = np.zeros((65536, 4), np.float32)
    u = scenes.uniform(41, 4 * 65536).reshape(-1, 4)
    spread[:, 0] = 1 + 1023 * u[:, 0]
    spread[:, 1] = 1 + 767 * u[:, 1]
    spread[:, 2:] = (u[:, 2:] - 0.5) * 20
    edge = scenes.next_block(128, 128, 0.7, (930.0, 690.0), 0.1, 9, vel=(30.0, 30.0))  # runs into both clamps
    edge[::13, 0] = 1024.0
    edge[::17, 1] = 768.0
    return {"block16k": scenes.next_block(128, 128, 0.7, (100.0, 200.0), 0.1, 2),
            "spread64k": spread, "edges16k": edge, "block1m": scenes.config_next_1m()}


@pytest.mark.parametrize("name", ["block16k", "spread64k", "edges16k", "block1m"])
def test_steps_match_oracle(oracle, name):
    import libcl_amd as L
    st = _states()[name]
    n = st.shape[0]
    print(f"next {name}: {n} particles", flush=True)
    f = _make(n)
    op = oracle.next_default_params()
    fo = oracle.Next(st, op)
    for step in range(3):
        assert f.upload(fo.state) == 1
        assert f.compute() == 1, L.last_error()
        fo.step()
        a = f.internals()
        for k, ref in (("cell", fo.cell), ("index", fo.index), ("sorted", fo.sorted), ("cell_start", fo.start),
                       ("cell_end", fo.end)):
            np.testing.assert_array_equal(a[k], ref, err_msg=f"{k} step {step}")
        np.testing.assert_allclose(a["pressure"], fo.pressure, rtol=1e-5, atol=1e-9)
        np.testing.assert_allclose(a["force"], fo.force, rtol=1e-5, atol=1e-5)
        got = f.state()["state"]
        np.testing.assert_allclose(got[:, 2:], fo.state[:, 2:], rtol=1e-5, atol=1e-5)
        np.testing.assert_allclose(got[:, :2], fo.state[:, :2], rtol=1e-6, atol=1e-4)


def test_external_state_buffer_and_rejections():
    import torch

    import libcl_amd as L
    st = scenes.next_block(64, 32, 0.7, (300.0, 300.0), 0.1, 4)
    f = _make(st.shape[0])
    t = torch.from_numpy(st).cuda()
    assert f.setPositionBuffer(t) == 1
    assert f.compute() == 1, L.last_error()
    torch.cuda.synchronize()
    assert not torch.equal(t.cpu(), torch.from_numpy(st))  # the caller's buffer was stepped in place
    assert f.setParticleCount(st.shape[0] + 256) == 0      # larger than the bound buffer
    g = _make(300)
    assert g.compute() == 0 and "multiple of 256" in L.last_error()
