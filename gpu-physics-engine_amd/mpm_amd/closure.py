"""Learned constitutive closure: dataset construction and GPU training through the engine's C-ABI.

The network (4-64-64-3, SiLU) maps the normalised SVD invariants of a particle's elastic deformation
gradient — singular values s1 >= s2, angle of U, and Jh (Jp for snow, 1 for jelly) — to tau / mu.  Its
training pairs are harvested from the plain (analytic) solver's own runs plus random deformation gradients
with singular values log-uniform in [0.6, 1.6]; targets are the analytic stress computed by the engine's
device code (mpm_closure_samples).  Training runs on the GPU with the engine's forward, backward and Adam
kernels (mpm_closure_train).
"""
import ctypes

import numpy as np

from . import abi, scenes
from .engine import Engine, MpmError

_f = ctypes.POINTER(ctypes.c_float)
_u8 = ctypes.POINTER(ctypes.c_uint8)
_i32 = ctypes.POINTER(ctypes.c_int32)

SHAPES = (("W1", (64, 4)), ("b1", (64,)), ("W2", (64, 64)), ("b2", (64,)), ("W3", (3, 64)), ("b3", (3,)))


def _p(a, t=_f):
    return a.ctypes.data_as(t)


def _lib_check(rc, what):
    lib = abi.load_library()
    if rc < 0:
        msg = lib.mpm_closure_message()
        raise MpmError(rc, f"{what}: {msg.decode() if msg else ''}")


def unflatten(w):
    """Flat float[4675] -> dict of named arrays (row = output unit)."""
    out, o = {}, 0
    for k, s in SHAPES:
        n = int(np.prod(s))
        out[k] = np.asarray(w[o:o + n], dtype=np.float64).reshape(s)
        o += n
    return out


def init_weights(seed=0):
    """He-style initialisation: N(0, 1/fan_in) weights, zero biases."""
    rng = np.random.default_rng(seed)
    parts = []
    for k, s in SHAPES:
        if k.startswith("W"):
            parts.append(rng.normal(0.0, 1.0 / np.sqrt(s[1]), s).reshape(-1))
        else:
            parts.append(np.zeros(int(np.prod(s))))
    return np.concatenate(parts).astype(np.float32)


def random_deformations(n, seed, sig_lo=0.6, sig_hi=1.6, jp_lo=0.7, jp_hi=1.3):
    """n deformation gradients F = U diag(s) V^T with s log-uniform in [sig_lo, sig_hi], U and V uniform
    rotations; half jelly (J slot 1), half snow (Jp uniform in [jp_lo, jp_hi])."""
    u = lambda stream, m: scenes.uniform(seed, stream, m)
    s1 = np.exp(np.log(sig_lo) + u(0, n) * (np.log(sig_hi) - np.log(sig_lo)))
    s2 = np.exp(np.log(sig_lo) + u(1, n) * (np.log(sig_hi) - np.log(sig_lo)))
    tu = (u(2, n) * 2 - 1) * np.pi
    tv = (u(3, n) * 2 - 1) * np.pi
    cu, su, cv, sv = np.cos(tu), np.sin(tu), np.cos(tv), np.sin(tv)
    # F = U diag(s) V^T, U = [[cu, -su], [su, cu]], V = [[cv, -sv], [sv, cv]]
    F = np.empty((n, 4))
    F[:, 0] = cu * s1 * cv + su * s2 * sv
    F[:, 1] = cu * s1 * sv - su * s2 * cv
    F[:, 2] = su * s1 * cv - cu * s2 * sv
    F[:, 3] = su * s1 * sv + cu * s2 * cv
    mat = np.where(u(4, n) < 0.5, abi.MAT_JELLY, abi.MAT_SNOW).astype(np.uint8)
    J = np.where(mat == abi.MAT_SNOW, jp_lo + u(5, n) * (jp_hi - jp_lo), 1.0)
    return F.astype(np.float32), J.astype(np.float32), mat


ACT_SCALE = 16.0  # hidden pre-activation z in the inference path's index units: u = 16 z + 159.5
ACT_SHIFT = 159.5


def act_table():
    """The inference path's SiLU segments: (slope[256], intercept[256]) float32 in index units u = 16 z + 159.5;
    silu(z) ~ slope[i] u + intercept[i] with i = clamp(rint(u), 0, 255), rint rounding half to even."""
    t = np.zeros(512, np.float32)
    _lib_check(abi.load_library().mpm_closure_act_table(_p(t)), "mpm_closure_act_table")
    return t[0::2].copy(), t[1::2].copy()


def act_silu(u, table=None):
    """Piecewise-linear SiLU of index-unit pre-activations u (float64 evaluation of the device's FMA)."""
    slope, icpt = act_table() if table is None else table
    u = np.asarray(u, dtype=np.float64)
    i = np.clip(np.rint(u), 0, 255).astype(np.int64)
    return slope[i].astype(np.float64) * u + icpt[i].astype(np.float64)


def samples(params, F, J, material):
    """Features [n,4], targets tau/mu [n,3] and validity mask from particle states (device computation)."""
    F = np.ascontiguousarray(F, dtype=np.float32).reshape(-1, 4)
    n = F.shape[0]
    J = np.ascontiguousarray(J, dtype=np.float32).reshape(n)
    m = np.ascontiguousarray(material, dtype=np.uint8).reshape(n)
    X = np.zeros((n, 4), np.float32)
    T = np.zeros((n, 3), np.float32)
    V = np.zeros(n, np.int32)
    rc = abi.load_library().mpm_closure_samples(ctypes.byref(params), n, _p(F), _p(J), _p(m, _u8), _p(X), _p(T),
                                                 _p(V, _i32))
    _lib_check(rc, "mpm_closure_samples")
    return X, T, V.astype(bool)


def harvest(params, scene, ticks):
    """(F, J, material) of every jelly / snow particle after each of `ticks` ticks of the analytic solver."""
    p = params.copy()
    p.closure = abi.CLOSURE_ANALYTIC
    p.capacity = max(p.capacity, scene.n)
    Fs, Js, Ms = [], [], []
    with Engine(p) as e:
        e.set_scene(scene)
        for _ in range(ticks):
            e.step(1)
            e.sync()
            o = e.readback()
            keep = o["material"] != abi.MAT_FLUID
            Fs.append(o["F"][keep])
            Js.append(o["J"][keep])
            Ms.append(o["material"][keep])
    return np.concatenate(Fs), np.concatenate(Js), np.concatenate(Ms)


def dataset(params, harvested, n_random, seed=1):
    """Training pairs: harvested states plus n_random random deformations; only in-range pairs kept."""
    F, J, M = harvested
    Fr, Jr, Mr = random_deformations(n_random, seed)
    F = np.concatenate([F, Fr])
    J = np.concatenate([J, Jr])
    M = np.concatenate([M, Mr])
    X, T, V = samples(params, F, J, M)
    return X[V], T[V]


def train(X, T, steps, batch=8192, lr=2e-3, seed=7, weights=None):
    """GPU training; returns (weights float32[4675], final minibatch loss)."""
    X = np.ascontiguousarray(X, dtype=np.float32)
    T = np.ascontiguousarray(T, dtype=np.float32)
    w = init_weights(seed) if weights is None else np.ascontiguousarray(weights, dtype=np.float32).copy()
    loss = ctypes.c_float()
    rc = abi.load_library().mpm_closure_train(X.shape[0], _p(X), _p(T), int(steps), int(batch), float(lr), int(seed),
                                               _p(w), ctypes.byref(loss))
    _lib_check(rc, "mpm_closure_train")
    return w, loss.value


def grad(X, T, weights):
    X = np.ascontiguousarray(X, dtype=np.float32)
    T = np.ascontiguousarray(T, dtype=np.float32)
    w = np.ascontiguousarray(weights, dtype=np.float32)
    g = np.zeros(abi.CLOSURE_NUM_WEIGHTS, np.float32)
    loss = ctypes.c_float()
    rc = abi.load_library().mpm_closure_grad(X.shape[0], _p(X), _p(T), _p(w), _p(g), ctypes.byref(loss))
    _lib_check(rc, "mpm_closure_grad")
    return loss.value, g


def evaluate(X, weights):
    """Network outputs through the engine's BF16 MFMA inference path."""
    X = np.ascontiguousarray(X, dtype=np.float32)
    w = np.ascontiguousarray(weights, dtype=np.float32)
    out = np.zeros((X.shape[0], 3), np.float32)
    rc = abi.load_library().mpm_closure_eval(X.shape[0], _p(X), _p(w), _p(out))
    _lib_check(rc, "mpm_closure_eval")
    return out
