"""CPU checks of the C ABI library: it loads, exports every symbol include/mfl.h declares, and its
host-side Params helpers agree bit for bit with the oracle and the independent fixture generator."""
import ctypes
import os
import re
import subprocess
import sys

import numpy as np

from conftest import ROOT, load_step_case


def _declared():
    text = open(os.path.join(ROOT, "include", "mfl.h")).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(mfl_[a-z0-9_]+)\s*\(", text)))


def test_library_exports_every_declared_symbol():
    from libcl_amd import _native as N
    lib = N.load()
    declared = _declared()
    assert len(declared) >= 30
    for name in declared:
        assert hasattr(lib, name), name
    assert set(declared) == set(N.EXPORTS)
    out = subprocess.run(["nm", "-D", "--defined-only", N.LIB_PATH], capture_output=True, text=True).stdout
    exported = set(re.findall(r" T (mfl_\w+)", out))
    assert set(declared) <= exported


def test_header_compiles_as_c(tmp_path):
    src = '#include "mfl.h"\nint main(void){ mfl_params p; mfl_params_default(&p); return (int)sizeof(p) - 52; }\n'
    r = subprocess.run(["gcc", "-std=c99", "-Wall", "-Werror", "-x", "c", "-", "-I", os.path.join(ROOT, "include"),
                        "-c", "-o", str(tmp_path / "hdr_check.o")], input=src, capture_output=True, text=True)
    assert r.returncode == 0, r.stderr


def test_sqrt_cut_is_the_exact_acceptance_threshold():
    """mfl_sqrt_cut(h) is the least float T with sqrtf(T) >= h, so r2 < T decides sqrtf(r2) < h for every
    float r2 >= 0 (numpy's float32 sqrt is IEEE, correctly rounded)."""
    from libcl_amd import _native as N
    from libcl_amd import scenes
    lib = N.load()
    hs = np.concatenate([np.float32([0.02, 0.0316, 0.04, 1.0, 2.0, 3.0]),
                         (1e-3 + 10 * scenes.uniform(17, 400)).astype(np.float32)])
    for h in hs:
        t = np.float32(lib.mfl_sqrt_cut(float(h)))
        below = np.nextafter(t, np.float32(0))
        assert np.sqrt(t) >= h and np.sqrt(below) < h, (h, t)
        # a window of floats around the threshold, decided both ways
        r2 = t + (np.arange(-64, 65, dtype=np.float32) * np.spacing(t)).astype(np.float32)
        np.testing.assert_array_equal(r2 < t, np.sqrt(r2) < h)
    assert lib.mfl_sqrt_cut(0.0) == 0.0 and lib.mfl_sqrt_cut(-1.0) == 0.0


def test_params_match_oracle_and_fixtures(oracle):
    import libcl_amd as L
    from libcl_amd import scenes
    assert ctypes.sizeof(L.Params) == 52
    d, o =L.default_params(), oracle.default_params()
    assert d.values() == [getattr(o, k) for k in L._native.PARAM_FIELDS]
    for h in (0.01, 0.02, 0.0316, 0.05):
        vals = scenes.sph_param_values(h)
        a = L.params_from(vals)
        b = oracle.make_params(**vals)
        assert a.values() == [getattr(b, k) for k in L._native.PARAM_FIELDS]
    z = load_step_case("lattice512")
    a = L.params_from(scenes.sph_param_values(0.02))
    np.testing.assert_array_equal(np.array(a.values(), np.float32), z["params"])


def test_program_creation_without_device_fails_cleanly():
    """Without a HIP device the program objects refuse creation with the HIP error text (no CPU
    fallback); with a device this is covered by the GPU tests."""
    code = ("import sys; sys.path.insert(0, %r)\n"
            "import torch\n"
            "if torch.cuda.is_available(): print('HAS_GPU'); sys.exit(0)\n"
            "import libcl_amd as L\n"
            "try:\n    L.oclFluid3D()\nexcept RuntimeError as e:\n    print('ERR', e)\n"
            "s = L.oclRadixSort()\nprint('COMPILE', s.compile(), L.last_error())\n") % ROOT
    r = subprocess.run([sys.executable, "-c", code], capture_output=True, text=True, timeout=300)
    if "HAS_GPU" in r.stdout:
        return
    assert "ERR" in r.stdout and "hip" in r.stdout.lower(), r.stdout + r.stderr
    assert "COMPILE 0" in r.stdout


def test_gpus_flag_rejected_above_one():
    r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "2"], capture_output=True,
                       text=True, timeout=120)
    assert r.returncode == 2
    assert "one GPU" in r.stderr


def test_committed_counter_profiles_match_the_timed_launch_shape():
    """The bench's roofline reads the committed counter summaries only when they were taken at the launch
    shape it times (the library's default shape at 1M particles, Morton traversal): they must be."""
    sys.path.insert(0, ROOT)
    import bench

    shape = dict(neighbour_pass="tile_kept", density_lanes=2, tile_depth=96, block_waves=2, sort_items=0,
                 force_chunks=2, traversal="morton")
    instr, fp32_share = bench.measured_valu("density", "sph_1m_box4", shape)
    assert instr > 0 and 0.0 < fp32_share < 1.0
    assert bench.measured_traffic("density", "sph_1m_box4", shape) > 0
    assert bench.measured_valu("density", "sph_1m_box4", dict(shape, tile_depth=64)) is None
