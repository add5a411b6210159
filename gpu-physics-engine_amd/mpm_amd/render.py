"""Headless rasteriser over the C-ABI of include/mpm_render.h (the reference's render library without its
D3D12 interop: libs/render/source/cuda-render.cu).

  cuda_engine_init / bind_buffers    -> Renderer(width, height, ...)   (mpm_renderer_create)
  cuda_engine_render_frame(View)     -> Renderer.frame(view)           (mpm_render_frame, host image)
"""
import ctypes

import numpy as np

from . import abi


class RenderError(RuntimeError):
    def __init__(self, code, message):
        super().__init__(f"{abi.load_library().mpm_status_name(code).decode()} ({code}): {message}")
        self.code = code


def _check(lib, rc, what):
    if rc < 0:
        msg = lib.mpm_render_message()
        raise RenderError(rc, f"{what}: {msg.decode() if msg else ''}")
    return rc


# reference display: kDispWidth x kDispHeight = (12 * 32 * 5) x (6 * 32 * 5), chunk lines every 160 px
REFERENCE = dict(width=1920, height=960, scale=5.0, chunk_px=160, flip_extent=0.0)


class Renderer:
    def __init__(self, width, height, scale, chunk_px=0, flip_extent=0.0, row_pitch=0):
        self.lib = abi.load_library()
        self.desc = abi.MpmRenderDesc(int(width), int(height), int(row_pitch), int(chunk_px), float(scale),
                                      float(flip_extent))
        self.h = ctypes.c_void_p()
        _check(self.lib, self.lib.mpm_renderer_create(ctypes.byref(self.desc), ctypes.byref(self.h)),
               "mpm_renderer_create")
        self.pitch = row_pitch or 4 * width

    def close(self):
        if self.h:
            self.lib.mpm_renderer_destroy(self.h)
            self.h = ctypes.c_void_p()

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def frame(self, view):
        """Draw an MpmView (device pointers, e.g. Engine.read_view()) into a host uint8 [height, width, 4]."""
        d = self.desc
        buf = np.zeros(self.pitch * d.height, np.uint8)
        rc = self.lib.mpm_render_frame(self.h, ctypes.byref(view), buf.ctypes.data_as(ctypes.c_void_p), 0)
        _check(self.lib, rc, "mpm_render_frame")
        return buf.reshape(d.height, self.pitch)[:, :4 * d.width].reshape(d.height, d.width, 4)


def write_ppm(path, img):
    """Write uint8 [height, width, 4] as a binary PPM through the library's writer."""
    img = np.ascontiguousarray(img, dtype=np.uint8)
    h, w = img.shape[:2]
    lib = abi.load_library()
    _check(lib, lib.mpm_write_ppm(str(path).encode(), img.ctypes.data_as(ctypes.POINTER(ctypes.c_uint8)), w, h, 4 * w),
           "mpm_write_ppm")
