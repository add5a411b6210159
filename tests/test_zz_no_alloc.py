"""mpm_render_frame_noalloc (include/mpm_render.h) against mpm_render_frame, through the C-ABI.

The no-allocation entry point keeps its per-pixel coverage counters in a device array of the library instead
of a renderer's allocation; its frames must equal the renderer's byte for byte.  Two calls in a row, the
second with fewer particles and a smaller image (the counters keep the first call's values beyond the second
image's pixels), then an image over the counters' 2,097,152 pixels, which must be refused before any launch.
"""
import ctypes

import numpy as np
import pytest

from mpm_amd import abi, render
from mpm_amd.engine import Engine

pytestmark = pytest.mark.gpu

_hip = None


def hip():
    global _hip
    if _hip is None:
        _hip = ctypes.CDLL("libamdhip64.so")
        _hip.hipMalloc.argtypes = [ctypes.POINTER(ctypes.c_void_p), ctypes.c_size_t]
        _hip.hipMalloc.restype = ctypes.c_int
        _hip.hipFree.argtypes = [ctypes.c_void_p]
        _hip.hipFree.restype = ctypes.c_int
        _hip.hipMemcpy.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_int]
        _hip.hipMemcpy.restype = ctypes.c_int
        _hip.hipMemset.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_size_t]
        _hip.hipMemset.restype = ctypes.c_int
    return _hip


def scene(n, seed):
    """Particles over the reference domain (384 x 192 units) with a dense clump (high coverage counts)."""
    rng = np.random.default_rng(seed)
    return np.concatenate([rng.uniform([0, 0], [384, 192], (n, 2)),
                           rng.normal([100, 60], 1.5, (n // 4, 2)).clip(0, 192)]).astype(np.float32)


def noalloc_frame(lib, desc, view, pitch, height):
    """The frame of mpm_render_frame_noalloc in a device image the test allocates, copied to the host."""
    nbytes = pitch * height
    dev = ctypes.c_void_p()
    assert hip().hipMalloc(ctypes.byref(dev), nbytes) == 0
    try:
        assert hip().hipMemset(dev, 0x5A, nbytes) == 0
        rc = lib.mpm_render_frame_noalloc(ctypes.byref(desc), ctypes.byref(view), dev)
        assert rc == abi.MPM_OK, lib.mpm_render_message()
        host = np.zeros((height, pitch), np.uint8)
        assert hip().hipMemcpy(host.ctypes.data, dev, nbytes, 2) == 0
    finally:
        hip().hipFree(dev)
    return host


def test_noalloc_frame_equals_renderer_frame():
    lib = abi.load_library()
    p = abi.preset(abi.PRESET_REFERENCE)
    with Engine(p) as e:
        e.set_particles(scene(30000, 3))
        e.step(1)
        e.sync()
        view = e.read_view()
        assert view.count == 37500
        # call 1: the reference display, 1920 x 960, chunk lines every 160 px
        desc = abi.MpmRenderDesc(1920, 960, 0, 160, 5.0, 0.0)
        with render.Renderer(1920, 960, scale=5.0, chunk_px=160) as r:
            want = r.frame(view)
        got = noalloc_frame(lib, desc, view, 4 * 1920, 960).reshape(960, 1920, 4)
        assert np.array_equal(got, want)
        # call 2: fewer particles, a smaller image with padded rows and a y flip
        small = abi.MpmView(view.pos, view.vel, view.num_atoms, view.slot, 5000)
        desc2 = abi.MpmRenderDesc(640, 320, 4 * 640 + 64, 40, 5.0 / 3.0, 192.0)
        with render.Renderer(640, 320, scale=5.0 / 3.0, chunk_px=40, flip_extent=192.0, row_pitch=4 * 640 + 64) as r:
            want2 = r.frame(small)
        got2 = noalloc_frame(lib, desc2, small, 4 * 640 + 64, 320)
        assert np.array_equal(got2[:, :4 * 640].reshape(320, 640, 4), want2)
        # an image over the counters' capacity is refused without a launch
        big = abi.MpmRenderDesc(4096, 1024, 0, 0, 1.0, 0.0)
        assert lib.mpm_render_frame_noalloc(ctypes.byref(big), ctypes.byref(view), ctypes.c_void_p(16)) == \
            abi.MPM_ERR_CAPACITY
        assert b"2097152" in lib.mpm_render_message()
        bad = abi.MpmRenderDesc(0, 10, 0, 0, 1.0, 0.0)
        assert lib.mpm_render_frame_noalloc(ctypes.byref(bad), ctypes.byref(view), ctypes.c_void_p(16)) == \
            abi.MPM_ERR_INVALID_ARG
