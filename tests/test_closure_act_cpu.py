"""CPU checks of the learned closure's inference activation: the piecewise-linear SiLU table the engine
builds on the host (mpm_closure_act_table) against the exact SiLU z / (1 + e^-z), evaluated the way the
device does (index clamp(rint(u), 0, 255) of u = 16 z + 159.5, one FMA)."""
import numpy as np

from mpm_amd import closure


def silu(z):
    return z / (1.0 + np.exp(-z))


def table_silu(z, table):
    u = 16.0 * np.asarray(z, dtype=np.float64) + 159.5
    return closure.act_silu(u, table)


def test_table_segments_and_error_bound():
    table = closure.act_table()
    slope, icpt = table
    assert slope.shape == (256,) and icpt.shape == (256,)
    assert np.all(np.isfinite(slope)) and np.all(np.isfinite(icpt))
    # interior segments: |error| <= 1.3e-4 (1/256 of a parabola with |silu''| <= 1/2, halved by the shift)
    # (open at the left end: at z = -10 + 1/16 exactly, u = 0.5 rounds to segment 0, the constant below)
    z = np.linspace(-10 + 1 / 16, 6, 400001)[1:-1]
    err = np.abs(table_silu(z, table) - silu(z))
    assert err.max() < 1.3e-4, err.max()
    # the shift halves the secant's error: without it the bound near z = 0 would be 2.4e-4
    assert err.max() > 1.0e-4
    # below z = -10 + 1/16 the constant half of silu there: |error| <= |silu(-9.9375)| / 2 ~ 2.4e-4
    zb = np.concatenate([np.linspace(-60, -10 + 1 / 16, 100001)[:-1], [-1e4]])
    errb = np.abs(table_silu(zb, table) - silu(zb))
    assert errb.max() <= 0.5 * abs(silu(-10 + 1 / 16)) * (1 + 1e-6), errb.max()
    assert slope[0] == 0.0
    # above z = 6 the last segment continues linearly: relative error below BF16's half ulp (2^-9) to z = 8
    zt = np.linspace(6, 8, 20001)
    rel = np.abs(table_silu(zt, table) - silu(zt)) / silu(zt)
    assert rel.max() < 2.0 ** -9, rel.max()


def test_table_index_units():
    """u = 16 z + 159.5: the segment boundaries are the multiples of 1/16 in z, and the lines of neighbouring
    interior segments meet within 1e-5 at the boundaries (the FP32 rounding of the intercepts; BF16 operands
    cannot resolve it)."""
    slope, icpt = closure.act_table()
    u = np.arange(2, 256, dtype=np.float64) - 0.5  # boundaries: rint switches segment at half-integers
    left = slope[1:255].astype(np.float64) * u + icpt[1:255]
    right = slope[2:256].astype(np.float64) * u + icpt[2:256]
    assert np.abs(left - right).max() < 1e-5
    z = (u - 159.5) / 16.0
    assert np.abs(right - silu(z)).max() < 1.3e-4
