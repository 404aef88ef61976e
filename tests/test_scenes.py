"""CPU checks of the seeded workload generators (libcl_amd/scenes.py)."""
import numpy as np

from libcl_amd import scenes


def test_splitmix64_reference_values():
    # first outputs of SplitMix64 seeded with 0 (published test vector of the generator)
    v = scenes.splitmix64(0, 3)
    assert [int(x) for x in v] == [0xE220A8397B1DCDAF, 0x6E789E6AA1B965F4, 0x06C45D188009454F]


def test_configs_shapes_and_determinism():
    a = scenes.config_small()
    b = scenes.config_small()
    assert a.n == 16384 and a.steps == 1000
    np.testing.assert_array_equal(a.pos, b.pos)
    assert np.all(a.pos[:, 3] == 0) and np.all(a.vel == 0)
    assert np.all(a.force[:, 2] == np.float32(-9.8)) and np.all(np.signbit(a.force[:, [0, 1, 3]]))
    # lattice at 0.5h = 0.01 from 0.25h = 0.005, jitter within +-0.01h
    assert abs(a.pos[:, 0].min() - 0.005) <= 0.0002 + 1e-7
    assert abs(a.pos[:, 0].max() - (0.005 + 31 * 0.01)) <= 0.0002 + 1e-6
    c = scenes.config_1m()
    assert c.n == 1 << 20 and c.box_max[0] == 4.0


def test_scene_counts_are_multiples_of_256():
    counts = scenes.scene_counts()
    assert counts.shape == (256,)
    assert np.all(counts % 256 == 0) and counts.min() >= 2048 and counts.max() <= 65536
    assert 3.5e6 < counts.sum() < 6e6


def test_sort_keys_ranges():
    k = scenes.sort_keys(1 << 16, seed=3)
    assert k.max() < (1 << 24)
    ks = scenes.sort_keys(1 << 16, seed=3, skew=True)
    # half the keys come from 10% of the cells: the most common keys repeat far more often
    assert np.bincount(ks).max() >= np.bincount(k).max()
