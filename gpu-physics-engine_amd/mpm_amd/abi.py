"""ctypes mirror of include/mpm_engine.h and the loader of the HIP engine library.

The engine is a C-ABI shared library (libmpm_engine.so) built from csrc/ by hipcc for gfx950.  This
module only describes its types and signatures; it raises ImportError-like RuntimeError when the library
is missing, so nothing silently falls back to another implementation.
"""
import ctypes
import os

_PKG_DIR = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
# MPM_ENGINE_LIB selects another build of the same sources (for example to compare two builds in one session).
LIB_PATH = os.environ.get("MPM_ENGINE_LIB") or os.path.join(_PKG_DIR, "lib", "libmpm_engine.so")

MPM_OK = 0
MPM_BUSY = 1
MPM_ERR_DEVICE = -1
MPM_ERR_NO_DEVICE = -2
MPM_ERR_STATE = -3
MPM_ERR_INVALID_ARG = -4
MPM_ERR_ALREADY = -5
MPM_ERR_CAPACITY = -6
MPM_ERR_NOT_BUILT = -7

MAT_FLUID = 0
MAT_JELLY = 1
MAT_SNOW = 2

MODE_REFERENCE = 0
MODE_TILED = 1
MODE_FUSED = 2

CLOSURE_ANALYTIC = 0
CLOSURE_LEARNED = 1

PRESET_REFERENCE = 0
PRESET_JELLY128 = 1
PRESET_MIXED1024 = 2
PRESET_BATCH64X256 = 3
PRESET_LEARNED1024 = 4
PRESET_LEARNED16X2048 = 5


class MpmParams(ctypes.Structure):
    """Field-for-field mirror of MpmParams in include/mpm_engine.h."""

    _fields_ = [
        ("grid_w", ctypes.c_int32),
        ("grid_h", ctypes.c_int32),
        ("halo", ctypes.c_int32),
        ("wall", ctypes.c_int32),
        ("dx", ctypes.c_float),
        ("inv_dx", ctypes.c_float),
        ("sim_w", ctypes.c_float),
        ("sim_h", ctypes.c_float),
        ("dt", ctypes.c_float),
        ("gravity_x", ctypes.c_float),
        ("gravity_y", ctypes.c_float),
        ("mass", ctypes.c_float),
        ("vol", ctypes.c_float),
        ("tait_k", ctypes.c_float),
        ("tait_gamma", ctypes.c_float),
        ("p_min", ctypes.c_float),
        ("p_max", ctypes.c_float),
        ("mu", ctypes.c_float),
        ("lam", ctypes.c_float),
        ("theta_c", ctypes.c_float),
        ("theta_s", ctypes.c_float),
        ("hardening", ctypes.c_float),
        ("j_min", ctypes.c_float),
        ("j_max", ctypes.c_float),
        ("substeps", ctypes.c_int32),
        ("capacity", ctypes.c_int32),
        ("num_scenes", ctypes.c_int32),
        ("sort_interval", ctypes.c_int32),
        ("mode", ctypes.c_int32),
        ("closure", ctypes.c_int32),
        ("block_x", ctypes.c_int32),
        ("block_y", ctypes.c_int32),
        ("chunk", ctypes.c_int32),
        ("out_chunk", ctypes.c_int32),
        ("reserved", ctypes.c_int32 * 7),
    ]

    def copy(self):
        c = MpmParams()
        ctypes.pointer(c)[0] = self
        return c

    def as_dict(self):
        return {f: getattr(self, f) for f, _ in self._fields_ if f != "reserved"}


class MpmView(ctypes.Structure):
    _fields_ = [
        ("pos", ctypes.c_void_p),
        ("vel", ctypes.c_void_p),
        ("num_atoms", ctypes.c_void_p),
        ("slot", ctypes.c_int32),
        ("count", ctypes.c_int32),
    ]


class MpmTuneResult(ctypes.Structure):
    _fields_ = [
        ("block_x", ctypes.c_int32),
        ("block_y", ctypes.c_int32),
        ("chunk", ctypes.c_int32),
        ("sort_interval", ctypes.c_int32),
        ("ms_per_substep", ctypes.c_float),
        ("candidates", ctypes.c_int32),
    ]


class MpmRenderDesc(ctypes.Structure):
    _fields_ = [
        ("width", ctypes.c_int32),
        ("height", ctypes.c_int32),
        ("row_pitch", ctypes.c_int32),
        ("chunk_px", ctypes.c_int32),
        ("scale", ctypes.c_float),
        ("flip_extent", ctypes.c_float),
    ]


_P = ctypes.POINTER
_f32p = _P(ctypes.c_float)
_i32p = _P(ctypes.c_int32)
_u8p = _P(ctypes.c_uint8)
_vp = ctypes.c_void_p

# name -> (restype, argtypes); every symbol the public headers declare.
SIGNATURES = {
    "mpm_params_preset": (ctypes.c_int32, [ctypes.c_int32, _P(MpmParams)]),
    "mpm_create": (ctypes.c_int32, [_P(MpmParams), _P(_vp)]),
    "mpm_destroy": (ctypes.c_int32, [_vp]),
    "mpm_set_particles": (
        ctypes.c_int32,
        [_vp, ctypes.c_int32, _f32p, _f32p, _f32p, _f32p, _f32p, _u8p, _i32p],
    ),
    "mpm_step": (ctypes.c_int32, [_vp, ctypes.c_int32]),
    "mpm_query": (ctypes.c_int32, [_vp]),
    "mpm_sync": (ctypes.c_int32, [_vp]),
    "mpm_last_step_ms": (ctypes.c_int32, [_vp, _f32p]),
    "mpm_read_view": (ctypes.c_int32, [_vp, _P(MpmView)]),
    "mpm_readback": (ctypes.c_int32, [_vp, _f32p, _f32p, _f32p, _f32p, _f32p, _u8p, _i32p]),
    "mpm_read_grid": (ctypes.c_int32, [_vp, _f32p, _f32p, _f32p]),
    "mpm_count": (ctypes.c_int32, [_vp, _i32p]),
    "mpm_save_state": (ctypes.c_int32, [_vp, ctypes.c_char_p]),
    "mpm_load_state": (ctypes.c_int32, [_vp, ctypes.c_char_p]),
    "mpm_hose_pipe": (
        ctypes.c_int32,
        [_vp, ctypes.c_float, ctypes.c_float, ctypes.c_float, ctypes.c_float, ctypes.c_float, ctypes.c_float,
         ctypes.c_float, ctypes.c_uint64, _i32p],
    ),
    "mpm_autotune": (ctypes.c_int32, [_P(MpmParams), ctypes.c_int32, _f32p, _f32p, _u8p, _i32p, ctypes.c_int32,
                                      _P(MpmTuneResult)]),
    "mpm_closure_samples": (ctypes.c_int32, [_P(MpmParams), ctypes.c_int32, _f32p, _f32p, _u8p, _f32p, _f32p, _i32p]),
    "mpm_closure_train": (ctypes.c_int32, [ctypes.c_int32, _f32p, _f32p, ctypes.c_int32, ctypes.c_int32, ctypes.c_float,
                                           ctypes.c_uint64, _f32p, _f32p]),
    "mpm_closure_grad": (ctypes.c_int32, [ctypes.c_int32, _f32p, _f32p, _f32p, _f32p, _f32p]),
    "mpm_closure_eval": (ctypes.c_int32, [ctypes.c_int32, _f32p, _f32p, _f32p]),
    "mpm_closure_act_table": (ctypes.c_int32, [_f32p]),
    "mpm_closure_message": (ctypes.c_char_p, []),
    "mpm_set_closure": (ctypes.c_int32, [_vp, _f32p]),
    "mpm_last_error": (ctypes.c_int32, [_vp]),
    "mpm_error_name": (ctypes.c_char_p, [ctypes.c_int32]),
    "mpm_status_name": (ctypes.c_char_p, [ctypes.c_int32]),
    "mpm_last_message": (ctypes.c_char_p, [_vp]),
    # include/mpm_render.h
    "mpm_renderer_create": (ctypes.c_int32, [_P(MpmRenderDesc), _P(_vp)]),
    "mpm_renderer_destroy": (ctypes.c_int32, [_vp]),
    "mpm_render_frame": (ctypes.c_int32, [_vp, _P(MpmView), _vp, ctypes.c_int32]),
    "mpm_render_frame_noalloc": (ctypes.c_int32, [_P(MpmRenderDesc), _P(MpmView), _vp]),
    "mpm_write_ppm": (ctypes.c_int32, [ctypes.c_char_p, _u8p, ctypes.c_int32, ctypes.c_int32, ctypes.c_int32]),
    "mpm_render_last_error": (ctypes.c_int32, []),
    "mpm_render_message": (ctypes.c_char_p, []),
    # include/mpm_diag.h
    "mpm_debug_bin": (ctypes.c_int32, [_vp, _P(ctypes.c_uint32), _i32p, _i32p]),
    "mpm_time_kernels": (ctypes.c_int32, [_vp, ctypes.c_int32, _f32p, _i32p]),
    "mpm_compute_stress": (ctypes.c_int32, [_vp, ctypes.c_int32, _u8p, _f32p, _f32p, _f32p]),
    "mpm_compute_snow_return": (ctypes.c_int32, [_vp, ctypes.c_int32, _f32p, _f32p, _f32p, _f32p, _f32p]),
}

CLOSURE_NUM_WEIGHTS = 4675

KT_BINNING = 0
KT_P2G = 1
KT_GRID = 2
KT_G2P = 3
KT_G2P2G = 4
KT_OUTPUT = 5
KT_COUNT = 6
KT_NAMES = ("binning", "p2g", "grid_update", "g2p", "g2p2g", "output")

_lib = None


def load_library(path=None):
    """Load libmpm_engine.so (raises RuntimeError when it has not been built)."""
    global _lib
    if _lib is not None and path is None:
        return _lib
    p = path or LIB_PATH
    if not os.path.exists(p):
        raise RuntimeError(
            f"HIP engine library not built: {p} is missing (run __graft_entry__.build() or make -C "
            f"gpu-physics-engine_amd/csrc)")
    lib = ctypes.CDLL(p, mode=ctypes.RTLD_GLOBAL)
    for name, (res, args) in SIGNATURES.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    if path is None:
        _lib = lib
    return lib


def preset(which):
    """MpmParams of a named preset (MPM_PRESET_*), filled by the engine library."""
    lib = load_library()
    p = MpmParams()
    rc = lib.mpm_params_preset(int(which), ctypes.byref(p))
    if rc != MPM_OK:
        raise ValueError(f"unknown preset {which}")
    return p
