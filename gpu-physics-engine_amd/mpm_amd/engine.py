"""Python mirror of the reference engine surface for the MPM path, over the C-ABI of include/mpm_engine.h.

Reference interface (libs/cuda-engine/include/cuda-engine/cuda-engine.h) and its counterpart here:
  MPMArena() + instantiateMPMGraph()        -> Engine(params)            (mpm_create)
  hosePipe(stream, mpm, ...)                -> Engine.hose_pipe(...)      (mpm_hose_pipe)
  cudaGraphLaunch per tick                  -> Engine.step(n_ticks)       (mpm_step)
  cudaEventQuery(tick_done)                 -> Engine.query()             (mpm_query)
  cudaEventElapsedTime(t0, t1)              -> Engine.last_step_ms()      (mpm_last_step_ms)
  OutputBuffer::readTarget()                -> Engine.read_view()         (mpm_read_view)
plus host readbacks for tests (Engine.readback / Engine.read_grid) and state files (Engine.save_state /
Engine.load_state).  Every call raises MpmError on a
negative status; there is no fallback path.
"""
import ctypes
import json
import os

import numpy as np

from . import abi


class MpmError(RuntimeError):
    """A negative status of the C-ABI; `code` is the status, `runtime_error` the hipError_t that
    mpm_last_error holds when the status is MPM_ERR_DEVICE (else None)."""

    def __init__(self, code, message, runtime_error=None):
        lib = abi.load_library()
        text = f"{lib.mpm_status_name(code).decode()} ({code}): {message}"
        if runtime_error is not None:
            text += f" [{lib.mpm_error_name(runtime_error).decode()} ({runtime_error})]"
        super().__init__(text)
        self.code = code
        self.runtime_error = runtime_error


def _check(lib, h, rc, what):
    if rc < 0:
        msg = lib.mpm_last_message(h)
        rt = lib.mpm_last_error(h) if rc == abi.MPM_ERR_DEVICE else None
        raise MpmError(rc, f"{what}: {msg.decode() if msg else ''}", rt)
    return rc


def _fp(a):
    return None if a is None else a.ctypes.data_as(ctypes.POINTER(ctypes.c_float))


class Engine:
    """One engine handle: SoA particle store, grids, tick graphs and the output triple buffer."""

    def __init__(self, params):
        self.lib = abi.load_library()
        self.params = params.copy()
        h = ctypes.c_void_p()
        rc = self.lib.mpm_create(ctypes.byref(self.params), ctypes.byref(h))
        if rc < 0:
            msg = self.lib.mpm_last_message(None)
            rt = self.lib.mpm_last_error(None) if rc == abi.MPM_ERR_DEVICE else None
            raise MpmError(rc, f"mpm_create: {msg.decode() if msg else ''}", rt)
        self.h = h
        self.n = 0

    # -- lifetime --------------------------------------------------------------------------------------
    def close(self):
        if self.h is not None and self.h.value:
            self.lib.mpm_destroy(self.h)
        self.h = None

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    # -- population ------------------------------------------------------------------------------------
    def set_particles(self, x, v=None, C=None, F=None, J=None, material=None, scene_counts=None):
        x = np.ascontiguousarray(x, dtype=np.float32).reshape(-1, 2)
        n = x.shape[0]

        def arr(a, cols, dt=np.float32):
            if a is None:
                return None
            return np.ascontiguousarray(a, dtype=dt).reshape(n, cols) if cols > 1 else \
                np.ascontiguousarray(a, dtype=dt).reshape(n)
        v, C, F, J = arr(v, 2), arr(C, 4), arr(F, 4), arr(J, 1)
        m = arr(material, 1, np.uint8)
        sc = None if scene_counts is None else np.ascontiguousarray(scene_counts, dtype=np.int32).reshape(-1)
        if sc is not None and sc.size != self.params.num_scenes:
            # mpm_set_particles reads num_scenes counts
            raise ValueError(f"scene_counts has {sc.size} entries, the handle has {self.params.num_scenes} scenes")
        rc = self.lib.mpm_set_particles(
            self.h, n, _fp(x), _fp(v), _fp(C), _fp(F), _fp(J),
            None if m is None else m.ctypes.data_as(ctypes.POINTER(ctypes.c_uint8)),
            None if sc is None else sc.ctypes.data_as(ctypes.POINTER(ctypes.c_int32)))
        _check(self.lib, self.h, rc, "mpm_set_particles")
        self.n = n

    def set_scene(self, scene):
        self.set_particles(scene.x, scene.v, material=scene.material,
                           scene_counts=scene.scene_counts if self.params.num_scenes > 1 else None)

    def hose_pipe(self, origin, velocity, nozzle_width, spacing, jitter, tick):
        out = ctypes.c_int32()
        rc = self.lib.mpm_hose_pipe(self.h, float(origin[0]), float(origin[1]), float(velocity[0]),
                                    float(velocity[1]), float(nozzle_width), float(spacing), float(jitter),
                                    int(tick), ctypes.byref(out))
        _check(self.lib, self.h, rc, "mpm_hose_pipe")
        return out.value

    # -- stepping --------------------------------------------------------------------------------------
    def step(self, n_ticks=1):
        _check(self.lib, self.h, self.lib.mpm_step(self.h, int(n_ticks)), "mpm_step")

    def query(self):
        return _check(self.lib, self.h, self.lib.mpm_query(self.h), "mpm_query") == abi.MPM_OK

    def sync(self):
        _check(self.lib, self.h, self.lib.mpm_sync(self.h), "mpm_sync")

    def last_step_ms(self):
        ms = ctypes.c_float()
        _check(self.lib, self.h, self.lib.mpm_last_step_ms(self.h, ctypes.byref(ms)), "mpm_last_step_ms")
        return ms.value

    # -- readback --------------------------------------------------------------------------------------
    def count(self):
        n = ctypes.c_int32()
        _check(self.lib, self.h, self.lib.mpm_count(self.h, ctypes.byref(n)), "mpm_count")
        return n.value

    def readback(self, fields=("x", "v", "C", "F", "J", "material")):
        """Host copy of the particle state in caller order; `fields` names the arrays to copy (the others are
        passed to mpm_readback as NULL and left out of the result)."""
        n = self.count()
        shapes = {"x": ((n, 2), np.float32), "v": ((n, 2), np.float32), "C": ((n, 4), np.float32),
                  "F": ((n, 4), np.float32), "J": ((n,), np.float32), "material": ((n,), np.uint8)}
        unknown = set(fields) - set(shapes)
        if unknown:
            raise ValueError(f"unknown readback fields {sorted(unknown)}")
        out = {k: np.zeros(*shapes[k]) for k in fields}
        m = out.get("material")
        got = ctypes.c_int32()
        rc = self.lib.mpm_readback(self.h, _fp(out.get("x")), _fp(out.get("v")), _fp(out.get("C")),
                                   _fp(out.get("F")), _fp(out.get("J")),
                                   None if m is None else m.ctypes.data_as(ctypes.POINTER(ctypes.c_uint8)),
                                   ctypes.byref(got))
        _check(self.lib, self.h, rc, "mpm_readback")
        return out

    def read_grid(self):
        p = self.params
        nodes = p.num_scenes * p.grid_w * p.grid_h
        mass = np.zeros(nodes, np.float32)
        vx = np.zeros(nodes, np.float32)
        vy = np.zeros(nodes, np.float32)
        _check(self.lib, self.h, self.lib.mpm_read_grid(self.h, _fp(mass), _fp(vx), _fp(vy)), "mpm_read_grid")
        shape = (p.num_scenes, p.grid_h, p.grid_w)
        return mass.reshape(shape), vx.reshape(shape), vy.reshape(shape)

    def set_closure(self, weights):
        """Load learned-closure parameters (float[4675]) into the handle."""
        w = np.ascontiguousarray(weights, dtype=np.float32).reshape(-1)
        if w.size != abi.CLOSURE_NUM_WEIGHTS:
            raise ValueError(f"closure weights must have {abi.CLOSURE_NUM_WEIGHTS} entries")
        _check(self.lib, self.h, self.lib.mpm_set_closure(self.h, _fp(w)), "mpm_set_closure")

    def read_view(self):
        view = abi.MpmView()
        _check(self.lib, self.h, self.lib.mpm_read_view(self.h, ctypes.byref(view)), "mpm_read_view")
        return view

    def save_state(self, path):
        """Write the particle state to a state file (mpm_save_state)."""
        _check(self.lib, self.h, self.lib.mpm_save_state(self.h, str(path).encode()), "mpm_save_state")

    def load_state(self, path):
        """Replace the population with the one in a state file (mpm_load_state)."""
        _check(self.lib, self.h, self.lib.mpm_load_state(self.h, str(path).encode()), "mpm_load_state")

    # -- diagnostics (include/mpm_diag.h) --------------------------------------------------------------
    def debug_bin(self):
        """Bin the current state; returns (keys, ids) in binned order."""
        n = self.count()
        keys = np.zeros(n, np.uint32)
        ids = np.zeros(n, np.int32)
        got = ctypes.c_int32()
        rc = self.lib.mpm_debug_bin(self.h, keys.ctypes.data_as(ctypes.POINTER(ctypes.c_uint32)),
                                    ids.ctypes.data_as(ctypes.POINTER(ctypes.c_int32)), ctypes.byref(got))
        _check(self.lib, self.h, rc, "mpm_debug_bin")
        return keys, ids

    def time_kernels(self, substeps):
        """Advance `substeps` eagerly with HIP events around every launch; returns {class: (ms, launches)}."""
        ms = np.zeros(abi.KT_COUNT, np.float32)
        cnt = np.zeros(abi.KT_COUNT, np.int32)
        rc = self.lib.mpm_time_kernels(self.h, int(substeps), _fp(ms),
                                       cnt.ctypes.data_as(ctypes.POINTER(ctypes.c_int32)))
        _check(self.lib, self.h, rc, "mpm_time_kernels")
        return {abi.KT_NAMES[k]: (float(ms[k]), int(cnt[k])) for k in range(abi.KT_COUNT)}

    def compute_stress(self, material, F, J):
        """Kirchhoff stress (tau00, tau01, tau11) per particle from the engine's device code."""
        F = np.ascontiguousarray(F, dtype=np.float32).reshape(-1, 4)
        n = F.shape[0]
        J = np.ascontiguousarray(J, dtype=np.float32).reshape(n)
        m = np.ascontiguousarray(material, dtype=np.uint8).reshape(n)
        tau = np.zeros((n, 3), np.float32)
        rc = self.lib.mpm_compute_stress(self.h, n, m.ctypes.data_as(ctypes.POINTER(ctypes.c_uint8)), _fp(F), _fp(J),
                                         _fp(tau))
        _check(self.lib, self.h, rc, "mpm_compute_stress")
        return tau

    def snow_return(self, F_trial, Jp):
        """G2P return mapping of snow particles then their P2G stress with the mapping's rotation, by the
        engine's device code: (F float[n,4], Jp float[n], tau float[n,3])."""
        F = np.ascontiguousarray(F_trial, dtype=np.float32).reshape(-1, 4)
        n = F.shape[0]
        J = np.ascontiguousarray(Jp, dtype=np.float32).reshape(n)
        Fo = np.zeros((n, 4), np.float32)
        Jo = np.zeros(n, np.float32)
        tau = np.zeros((n, 3), np.float32)
        rc = self.lib.mpm_compute_snow_return(self.h, n, _fp(F), _fp(J), _fp(Fo), _fp(Jo), _fp(tau))
        _check(self.lib, self.h, rc, "mpm_compute_snow_return")
        return Fo, Jo, tau


def tune_key(params, scene, substeps=None):
    """What an autotuned launch shape depends on: grid, scenes, particle count (to the nearest power of two),
    the set of materials, mode, closure and substeps per timed tick.  A string, the key of the tune cache."""
    n = int(np.asarray(scene.x).reshape(-1, 2).shape[0])
    mats = sorted(int(m) for m in np.unique(np.asarray(scene.material)))
    return json.dumps({"grid": [int(params.grid_w), int(params.grid_h)], "scenes": int(params.num_scenes),
                       "particles_log2": int(round(np.log2(max(n, 1)))), "materials": mats,
                       "mode": int(params.mode), "closure": int(params.closure),
                       "substeps": int(substeps or params.substeps)}, sort_keys=True)


class TuneCache:
    """Autotuner choices kept in a small JSON file: {tune_key: {block_x, block_y, chunk, sort_interval,
    ms_per_substep}}.  Writes go to a temporary file that replaces the cache in one rename, so a reader
    never sees a half-written file; an unreadable file counts as empty."""

    FIELDS = ("block_x", "block_y", "chunk", "sort_interval", "ms_per_substep")

    def __init__(self, path):
        self.path = str(path)

    def _load(self):
        try:
            with open(self.path) as fh:
                data = json.load(fh)
            return data if isinstance(data, dict) else {}
        except (OSError, ValueError):
            return {}

    def get(self, key):
        entry = self._load().get(key)
        if not isinstance(entry, dict) or any(k not in entry for k in self.FIELDS):
            return None
        return entry

    def put(self, key, res):
        data = self._load()
        data[key] = {k: (float(getattr(res, k)) if k == "ms_per_substep" else int(getattr(res, k)))
                     for k in self.FIELDS}
        tmp = f"{self.path}.tmp{os.getpid()}"
        with open(tmp, "w") as fh:
            json.dump(data, fh, indent=1, sort_keys=True)
        os.replace(tmp, self.path)


def autotune(params, scene, substeps=None, cache=None):
    """Run the engine's autotuner (mpm_autotune) on a scene; returns a tuned copy of params and the
    MpmTuneResult.  With cache (a file path), a choice stored for the same tune_key is returned without
    tuning (the result's candidates field is then 0), and a new choice is stored."""
    p = params.copy()
    key = tune_key(params, scene, substeps) if cache else None
    if cache:
        hit = TuneCache(cache).get(key)
        if hit is not None:
            res = abi.MpmTuneResult(int(hit["block_x"]), int(hit["block_y"]), int(hit["chunk"]),
                                    int(hit["sort_interval"]), float(hit["ms_per_substep"]), 0)
            p.block_x, p.block_y, p.chunk, p.sort_interval = res.block_x, res.block_y, res.chunk, res.sort_interval
            return p, res
    lib = abi.load_library()
    x = np.ascontiguousarray(scene.x, dtype=np.float32)
    v = np.ascontiguousarray(scene.v, dtype=np.float32)
    m = np.ascontiguousarray(scene.material, dtype=np.uint8)
    sc = np.ascontiguousarray(scene.scene_counts, dtype=np.int32) if p.num_scenes > 1 else None
    res = abi.MpmTuneResult()
    rc = lib.mpm_autotune(ctypes.byref(p), x.shape[0], _fp(x), _fp(v), m.ctypes.data_as(ctypes.POINTER(ctypes.c_uint8)),
                          None if sc is None else sc.ctypes.data_as(ctypes.POINTER(ctypes.c_int32)),
                          int(substeps or p.substeps), ctypes.byref(res))
    if rc < 0:
        msg = lib.mpm_last_message(None)
        raise MpmError(rc, f"mpm_autotune: {msg.decode() if msg else ''}")
    p.block_x, p.block_y, p.chunk, p.sort_interval = res.block_x, res.block_y, res.chunk, res.sort_interval
    if cache:
        TuneCache(cache).put(key, res)
    return p, res
