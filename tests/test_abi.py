"""CPU checks of the C-ABI library: it loads, exports every symbol the public headers declare, fills
presets and rejects bad arguments with the documented status codes (no GPU needed)."""
import ctypes
import os
import re

import numpy as np
import pytest

from mpm_amd import abi

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_symbols():
    names = []
    for h in ("mpm_engine.h", "mpm_diag.h", "mpm_render.h"):
        txt = open(os.path.join(ROOT, "include", h)).read()
        txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
        names += re.findall(r"^\s*(?:const\s+)?[\w\*\s]+?\b(mpm_\w+)\s*\(", txt, flags=re.M)
    return sorted(set(names))


def test_library_exports_every_declared_symbol():
    lib = abi.load_library()
    names = declared_symbols()
    assert len(names) >= 19
    for n in names:
        assert hasattr(lib, n), n
    assert set(names) == set(abi.SIGNATURES), set(names) ^ set(abi.SIGNATURES)


def test_params_struct_layout_matches_header():
    txt = open(os.path.join(ROOT, "include", "mpm_engine.h")).read()
    body = txt[txt.index("typedef struct MpmParams {"):txt.index("} MpmParams;")]
    body = re.sub(r"/\*.*?\*/", "", body, flags=re.S)
    fields = []
    for line in body.splitlines()[1:]:
        line = line.strip().rstrip(";")
        if not line:
            continue
        typ, rest = line.split(None, 1)
        for name in rest.split(","):
            name = name.strip()
            m = re.match(r"(\w+)\[(\d+)\]", name)
            fields.append((m.group(1), int(m.group(2))) if m else (name, 1))
    mirror = [(n if n != "lam" else "lambda", getattr(t, "_length_", 1)) for n, t in abi.MpmParams._fields_]
    assert fields == mirror
    assert ctypes.sizeof(abi.MpmParams) == 4 * sum(c for _, c in fields)


def test_reference_preset_matches_constants():
    """constants.h:9-88 of the reference: 196 x 100 nodes, dx = 2, dt = 0.032 / 20, g = 700, Tait
    k = 8.2e4, gamma = 3.2, p in [-0, 8.2e5], m = 1, V0 = 1.2, 20 substeps, 400,000 particles."""
    p = abi.preset(abi.PRESET_REFERENCE)
    assert (p.grid_w, p.grid_h, p.halo, p.wall) == (196, 100, 2, 2)
    assert p.dx == 2.0 and p.inv_dx == 0.5
    assert (p.sim_w, p.sim_h) == (384.0, 192.0)
    assert p.dt == np.float32(np.float32(0.032) / np.float32(20))
    assert (p.gravity_x, p.gravity_y) == (0.0, 700.0)
    assert p.mass == 1.0 and p.vol == np.float32(1.2)
    assert p.tait_k == np.float32(8.2e4) and p.tait_gamma == np.float32(3.2)
    assert p.p_max == np.float32(8.2e5)
    assert p.substeps == 20 and p.capacity == 400000 and p.out_chunk == 1536


def test_unit_presets():
    p = abi.preset(abi.PRESET_JELLY128)
    assert p.grid_w == 132 and p.dx == np.float32(1 / 128) and p.dt == np.float32(1e-4)
    assert p.mass == np.float32(1 / 128 ** 2 / 4)
    assert p.mu == pytest.approx(1e3 / 2.4) and p.lam == pytest.approx(200 / 0.72)
    p = abi.preset(abi.PRESET_MIXED1024)
    assert p.grid_w == 1028 and p.capacity == 4194304 and p.dt == np.float32(2e-5)
    p = abi.preset(abi.PRESET_BATCH64X256)
    assert p.num_scenes == 64 and p.grid_w == 260
    with pytest.raises(ValueError):
        abi.preset(99)


def test_create_rejects_bad_params_before_touching_a_device():
    lib = abi.load_library()
    bad = abi.preset(abi.PRESET_JELLY128)
    bad.halo = 1
    h = ctypes.c_void_p()
    assert lib.mpm_create(ctypes.byref(bad), ctypes.byref(h)) == abi.MPM_ERR_INVALID_ARG
    assert b"halo" in lib.mpm_last_message(None)
    bad = abi.preset(abi.PRESET_JELLY128)
    bad.block_x, bad.block_y = 12, 12
    assert lib.mpm_create(ctypes.byref(bad), ctypes.byref(h)) == abi.MPM_ERR_INVALID_ARG
    bad = abi.preset(abi.PRESET_JELLY128)
    bad.sim_w = 2.0
    assert lib.mpm_create(ctypes.byref(bad), ctypes.byref(h)) == abi.MPM_ERR_INVALID_ARG
    bad = abi.preset(abi.PRESET_MIXED1024)
    bad.theta_c = 1.0  # singular values clamped to [0, ...]: Jp would divide by zero
    assert lib.mpm_create(ctypes.byref(bad), ctypes.byref(h)) == abi.MPM_ERR_INVALID_ARG
    assert b"theta_c" in lib.mpm_last_message(None)
    assert lib.mpm_create(None, ctypes.byref(h)) == abi.MPM_ERR_INVALID_ARG


def test_null_handle_and_error_names():
    lib = abi.load_library()
    assert lib.mpm_step(None, 1) == abi.MPM_ERR_INVALID_ARG
    assert lib.mpm_sync(None) == abi.MPM_ERR_INVALID_ARG
    assert lib.mpm_destroy(None) == abi.MPM_ERR_INVALID_ARG
    assert lib.mpm_status_name(abi.MPM_ERR_CAPACITY) == b"MPM_ERR_CAPACITY"
    assert lib.mpm_status_name(abi.MPM_OK) == b"MPM_OK"
    assert lib.mpm_status_name(-99) == b"MPM_UNKNOWN_STATUS"
    # mpm_error_name names runtime codes, like cuda_engine_error_name (cudaGetErrorName, cuda-render.cu:159)
    assert lib.mpm_error_name(0) == b"hipSuccess"
    assert lib.mpm_error_name(2) == b"hipErrorOutOfMemory"


def test_status_codes_keep_the_reference_meanings():
    """-1 .. -5 mean what the reference's C-ABI returns for them (libs/render/source/cuda-render.cu):
    record() turns a failed runtime call into -1 (:23-27), no matching adapter is -2 (:46), use before
    init / bind is -3 (:80, :171), a bad argument is -4 (:31, :81, :89), init / bind twice is -5 (:32, :82)."""
    import re
    hdr = open(os.path.join(ROOT, "include", "mpm_engine.h")).read()
    defs = {m.group(1): int(m.group(2)) for m in re.finditer(r"#define (MPM_[A-Z_]+) \(?(-?\d+)\)?", hdr)}
    assert defs["MPM_ERR_DEVICE"] == -1 == abi.MPM_ERR_DEVICE
    assert defs["MPM_ERR_NO_DEVICE"] == -2 == abi.MPM_ERR_NO_DEVICE
    assert defs["MPM_ERR_STATE"] == -3 == abi.MPM_ERR_STATE
    assert defs["MPM_ERR_INVALID_ARG"] == -4 == abi.MPM_ERR_INVALID_ARG
    assert defs["MPM_ERR_ALREADY"] == -5 == abi.MPM_ERR_ALREADY
    assert defs["MPM_ERR_CAPACITY"] == abi.MPM_ERR_CAPACITY and defs["MPM_ERR_NOT_BUILT"] == abi.MPM_ERR_NOT_BUILT
    statuses = [v for k, v in defs.items() if k.startswith("MPM_ERR_") or k in ("MPM_OK", "MPM_BUSY")]
    assert len(statuses) == len(set(statuses))
    # no handle, failure before any runtime call: the runtime code stays as it was (0: none failed yet)
    lib = abi.load_library()
    bad = abi.preset(abi.PRESET_JELLY128)
    bad.halo = 1
    h = ctypes.c_void_p()
    before = lib.mpm_last_error(None)
    assert lib.mpm_create(ctypes.byref(bad), ctypes.byref(h)) == abi.MPM_ERR_INVALID_ARG
    assert lib.mpm_last_error(None) == before


def test_engine_without_device_fails_loudly():
    """On a host without a GPU the engine refuses to run (no CPU fallback exists)."""
    import mpm_amd
    from mpm_amd.engine import MpmError
    try:
        import torch
        if torch.cuda.is_available():
            pytest.skip("a GPU is present")
    except ImportError:
        pass
    with pytest.raises(MpmError) as ei:
        mpm_amd.Engine(abi.preset(abi.PRESET_JELLY128))
    assert ei.value.code == abi.MPM_ERR_NO_DEVICE
