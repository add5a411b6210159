"""CPU checks of the frame restatement (the reference renderer's background and particle blend) and of the
state-file format: closed forms for the background, the blend, the footprint and its clipping, order
independence of the blended frame, and the FNV-1a checksum against its published test vectors."""
import numpy as np
import pytest

from mpm_amd import abi, state


def blend(p, c):
    """drawKernel's alpha() for one channel, FP32 as written (cuda-render.cu:244-257)."""
    a = np.float32(100) * np.float32(np.float32(1) / np.float32(255))
    ia = np.float32(1) - a
    return int(np.float32(np.float32(np.float32(c) * a) + np.float32(np.float32(p) * ia)) + np.float32(0.5))


def background(x, y, w, h, chunk):
    if chunk and (x % chunk == 0 or y % chunk == 0):
        return (0, 0, 0)
    shade = int(np.float32(100) * np.float32(np.float32(y) / np.float32(h))) & 0xFF
    return ((255 - shade) & 0xFF, (160 - shade) & 0xFF, 80)


def test_background_closed_form(oracle_mod):
    w, h = 400, 330
    img = oracle_mod.render(np.zeros((0, 2), np.float32), w, h, chunk_px=160, scale=5.0)
    for (x, y) in [(0, 0), (1, 1), (159, 7), (160, 7), (7, 160), (399, 329), (200, 100), (321, 1)]:
        assert tuple(img[y, x, :3]) == background(x, y, w, h, 160), (x, y)
    assert (img[..., 3] == 255).all()
    # no grid lines when chunk_px = 0
    img0 = oracle_mod.render(np.zeros((0, 2), np.float32), w, h, chunk_px=0, scale=5.0)
    assert tuple(img0[0, 0, :3]) == background(0, 0, w, h, 0)


def test_single_particle_footprint_and_blend(oracle_mod):
    w, h = 64, 48
    bg = oracle_mod.render(np.zeros((0, 2), np.float32), w, h, chunk_px=0, scale=5.0)
    img = oracle_mod.render(np.array([[4.3, 5.1]], np.float32), w, h, chunk_px=0, scale=5.0)
    rx, ry = int(np.float32(4.3) * np.float32(5)), int(np.float32(5.1) * np.float32(5))
    changed = np.argwhere((img != bg).any(axis=2))
    expect = {(ry + i, rx + j) for i in range(-2, 3) for j in range(-2, 3) if not (abs(i) == 2 and abs(j) == 2)}
    assert {tuple(c) for c in changed} == expect
    for (yy, xx) in expect:
        assert tuple(img[yy, xx, :3]) == tuple(blend(bg[yy, xx, k], c) for k, c in enumerate((140, 210, 255)))


def test_footprint_clipping_and_truncation(oracle_mod):
    w, h = 32, 32
    bg = oracle_mod.render(np.zeros((0, 2), np.float32), w, h, chunk_px=0, scale=1.0)
    # (int)(-0.5) = 0: drawn at the corner, clipped to the 3x3 block minus its far corner
    img = oracle_mod.render(np.array([[-0.5, 0.2]], np.float32), w, h, chunk_px=0, scale=1.0)
    assert (img != bg).any(axis=2).sum() == 8
    # (int)(-1.5) = -1: off the image, nothing drawn
    img = oracle_mod.render(np.array([[-1.5, 3.0], [40.0, 3.0], [3.0, 32.0]], np.float32), w, h, chunk_px=0, scale=1.0)
    assert (img == bg).all()
    # the flip draws y as extent - y
    a = oracle_mod.render(np.array([[10.0, 7.0]], np.float32), w, h, chunk_px=0, scale=1.0, flip_extent=20.0)
    b = oracle_mod.render(np.array([[10.0, 13.0]], np.float32), w, h, chunk_px=0, scale=1.0)
    assert (a == b).all()


def test_blend_depends_only_on_coverage(oracle_mod):
    """k particles on one pixel give the k-fold blend whatever their order (the property the device
    rasteriser's coverage count relies on)."""
    rng = np.random.default_rng(0)
    w, h = 96, 64
    pos = np.concatenate([np.full((40, 2), 9.0, np.float32),
                          rng.uniform(0, 18, (300, 2)).astype(np.float32)])
    a = oracle_mod.render(pos, w, h, chunk_px=32, scale=5.0)
    b = oracle_mod.render(pos[rng.permutation(len(pos))], w, h, chunk_px=32, scale=5.0)
    assert (a == b).all()
    bg = oracle_mod.render(np.zeros((0, 2), np.float32), w, h, chunk_px=32, scale=5.0)
    px = bg[45, 45, :3].astype(int)
    for _ in range(40):
        px = [blend(px[k], c) for k, c in enumerate((140, 210, 255))]
    single = oracle_mod.render(np.full((40, 2), 9.0, np.float32), w, h, chunk_px=32, scale=5.0)
    assert tuple(single[45, 45, :3]) == tuple(px)


def test_fnv1a64_published_vectors():
    assert state.fnv1a64(b"") == 0xCBF29CE484222325
    assert state.fnv1a64(b"a") == 0xAF63DC4C8601EC8C
    assert state.fnv1a64(b"foobar") == 0x85944171F73967E8


def test_state_file_roundtrip(tmp_path):
    p = abi.preset(abi.PRESET_BATCH64X256)
    rng = np.random.default_rng(1)
    n = 37
    arrs = dict(x=rng.random((n, 2), np.float32), v=rng.random((n, 2), np.float32), C=rng.random((n, 4), np.float32),
                F=rng.random((n, 4), np.float32), J=rng.random(n, np.float32),
                material=rng.integers(0, 3, n).astype(np.uint8), scene=rng.integers(0, 64, n).astype(np.int32))
    path = tmp_path / "s.mpmstate"
    state.write_state(path, p, **arrs)
    assert path.stat().st_size == 200 + 57 * n
    st = state.read_state(path)
    assert st["n"] == n and st["num_scenes"] == 64 and st["params"].grid_w == p.grid_w
    for k, a in arrs.items():
        assert np.array_equal(st[k].reshape(a.shape), a), k
    raw = bytearray(path.read_bytes())
    raw[-1] ^= 1
    path.write_bytes(bytes(raw))
    with pytest.raises(ValueError):
        state.read_state(path)
