"""The C++ drop-in (include/oclmfl.hpp): a C++ program written like the reference's own test program
(libCL/TEST/main.cpp:133-343, tests/cpp/ref_smoke.cpp) compiles with g++ against the header and links
libmfl.so; on a GPU it runs testRadixSort, testFluid3D0 and testFluid3D1's sequences, and the state it
dumps around four of each fluid flow's steps is replayed by the oracle: sort result bit-exact, keys,
permutation and bucket tables bit-exact, density / pressure within 1e-5, velocities and positions
within the change a 1e-5 force error makes in one step."""
import os
import subprocess

import numpy as np
import pytest

from conftest import ROOT

SRC = os.path.join(ROOT, "tests", "cpp", "ref_smoke.cpp")
LIB_DIR = os.path.join(ROOT, "libcl_amd")
TOL = 1e-5


def _build(out_dir):
    exe = os.path.join(str(out_dir), "ref_smoke")
    cmd = ["g++", "-std=c++17", "-O2", "-Wall", "-Wextra", "-Werror", "-I", os.path.join(ROOT, "include"), SRC,
           "-L", LIB_DIR, "-lmfl", f"-Wl,-rpath,{LIB_DIR}", "-L/opt/rocm/lib", "-Wl,-rpath,/opt/rocm/lib", "-o", exe]
    r = subprocess.run(cmd, capture_output=True, text=True)
    assert r.returncode == 0, r.stderr
    return exe


def test_cpp_dropin_compiles_and_links(tmp_path):
    assert os.path.exists(_build(tmp_path))


def _load(d, name, dtype, cols=None):
    a = np.fromfile(os.path.join(d, name), dtype)
    return a.reshape(-1, cols) if cols else a


@pytest.mark.gpu
def test_cpp_dropin_reference_flows(tmp_path):
    from oracle import oracle as O
    import libcl_amd as L
    exe = _build(tmp_path)
    out = tmp_path / "dump"
    out.mkdir()
    print("running tests/cpp/ref_smoke.cpp", flush=True)
    r = subprocess.run([exe, str(out)], capture_output=True, text=True, timeout=120)
    print(r.stdout, flush=True)
    assert r.returncode == 0, r.stderr
    assert "ref_smoke ok" in r.stdout
    d = str(out)
    # testRadixSort: bits [0, 32) of 1024 keys, stable
    keys = _load(d, "sort_keys_in.bin", np.uint32)
    order = np.argsort(keys, kind="stable")
    np.testing.assert_array_equal(_load(d, "sort_keys_out.bin", np.uint32), keys[order])
    np.testing.assert_array_equal(_load(d, "sort_vals_out.bin", np.uint32), order.astype(np.uint32))
    # the fluid flows run with clInitFluid's Params and the reference box [-1, 1]^3
    prm = L.default_params()
    op = O.Params(*prm.values())
    dt = float(prm.deltaTime)
    for flow, clip in (("fluid0", True), ("fluid1", False)):
        for step in (0, 1, 10, 39):
            tag = f"{flow}_{step}"
            b = {k: _load(d, f"{tag}_before_{k}.bin", np.float32, 4) for k in ("pos", "vel", "force")}
            cs = _load(d, f"{tag}_before_cell_start.bin", np.uint32)
            ce = _load(d, f"{tag}_before_cell_end.bin", np.uint32)
            a = {k: _load(d, f"{tag}_after_{k}.bin", np.float32, 4) for k in ("pos", "vel", "force", "spos", "svel")}
            fo = O.Fluid(b["pos"], b["vel"], b["force"], op, cell_start=cs, cell_end=ce)
            fo.step(integrate=False)
            np.testing.assert_array_equal(_load(d, f"{tag}_after_cell.bin", np.uint32), fo.cell, err_msg=tag)
            np.testing.assert_array_equal(_load(d, f"{tag}_after_index.bin", np.uint32), fo.index, err_msg=tag)
            np.testing.assert_array_equal(_load(d, f"{tag}_after_cell_start.bin", np.uint32), fo.cell_start)
            np.testing.assert_array_equal(_load(d, f"{tag}_after_cell_end.bin", np.uint32), fo.cell_end)
            np.testing.assert_array_equal(a["spos"][:, :3], fo.spos[:, :3])
            rel = np.abs(a["spos"][:, 3] - fo.spos[:, 3]) / np.abs(fo.spos[:, 3])
            assert rel.max() <= TOL, (tag, rel.max())
            p = fo.svel[:, 3].astype(np.float64)
            assert (np.abs(a["svel"][:, 3] - p) / (np.abs(p) + np.sqrt(np.mean(p ** 2)) + 1e-30)).max() <= TOL
            # the rest of the step: the program's clip box + integration, or the handler's three kernels
            if clip:
                O.clip_box(fo.spos, fo.svel, fo.force, fo.index, op, (-1, -1, -1, 0), (1, 1, 1, 0))
            f_used = fo.force.copy()
            O.integrate_force(fo.vel, fo.force, op)
            O.gravity(fo.force)
            O.integrate_velocity(fo.pos, fo.vel, op)
            np.testing.assert_array_equal(a["force"], fo.force)
            fmag = np.linalg.norm(f_used[:, :3].astype(np.float64), axis=1)
            vtol = TOL * (fmag + np.sqrt(np.mean(fmag ** 2))) * dt + 4e-7 * (np.abs(fo.vel[:, :3]).max(1) + 1e-6)
            assert np.all(np.linalg.norm(a["vel"][:, :3] - fo.vel[:, :3], axis=1) <= vtol), tag
            ptol = vtol * dt + 2.5e-7 * np.abs(fo.pos[:, :3]).max(1)
            assert np.all(np.linalg.norm(a["pos"][:, :3] - fo.pos[:, :3], axis=1) <= ptol), tag
        print(f"{flow}: 4 steps replayed by the oracle", flush=True)
