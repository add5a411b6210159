import numpy as np

from mpm_amd import abi, scenes


def test_splitmix64_known_values():
    """SplitMix64 of seed 0, first outputs (Steele et al. 2014 reference sequence for state 0)."""
    z = scenes.splitmix64(0, np.arange(3, dtype=np.uint64))
    assert [int(v) for v in z] == [0xE220A8397B1DCDAF, 0x6E789E6AA1B965F4, 0x06C45D188009454F]


def test_jelly128_is_deterministic_and_in_box():
    a = scenes.jelly128()
    b = scenes.jelly128()
    assert np.array_equal(a.x, b.x)
    assert a.n == 32768
    assert a.x[:, 0].min() >= 0.25 and a.x[:, 0].max() <= 0.55
    assert a.x[:, 1].min() >= 0.45 and a.x[:, 1].max() <= 0.75
    assert np.all(a.v == 0) and np.all(a.material == abi.MAT_JELLY)


def test_mixed_scene_layout():
    s = scenes.mixed(n=30000)
    assert s.n == 30000
    assert set(np.unique(s.material)) == {abi.MAT_FLUID, abi.MAT_JELLY, abi.MAT_SNOW}
    assert np.all(np.abs(s.v) <= 1.0)
    for (lo, hi), m in zip(scenes.MIXED_BOXES, (0, 1, 2)):
        sel = s.x[s.material == m]
        assert sel[:, 0].min() >= lo[0] and sel[:, 0].max() <= hi[0]
        assert sel[:, 1].min() >= lo[1] and sel[:, 1].max() <= hi[1]
    # density kept at 8 particles per cell of the 1024-cell grid
    sel = s.x[s.material == 1]
    area = np.ptp(sel[:, 0]) * np.ptp(sel[:, 1]) * 1024 ** 2
    assert abs(len(sel) / area - 8.0) < 0.5


def test_batched_counts_log_uniform():
    s = scenes.batched()
    assert len(s.scene_counts) == 64
    assert s.scene_counts.min() >= 4096 and s.scene_counts.max() <= 524288
    assert s.n == int(s.scene_counts.sum())
    assert s.x.min() >= 0 and s.x.max() <= 1
