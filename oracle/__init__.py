"""Python loader of the FP64 MLS-MPM oracle (oracle/mpm_oracle.cpp).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg,
as the checker or the timed CPU baseline.  The engine never imports it.
"""
import ctypes
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
# MPM_ORACLE_LIB selects another build of the same source (the sanitized one of `make sanitized`).
LIB_PATH = os.environ.get("MPM_ORACLE_LIB") or os.path.join(HERE, "liboracle_mpm.so")

_lib = None
_d = ctypes.POINTER(ctypes.c_double)
_u8 = ctypes.POINTER(ctypes.c_uint8)
_i32 = ctypes.POINTER(ctypes.c_int32)
_u32 = ctypes.POINTER(ctypes.c_uint32)
_f32 = ctypes.POINTER(ctypes.c_float)


def build(force=False):
    """Compile the oracle with its Makefile (g++, OpenMP)."""
    args = ["make", "-C", HERE]
    if force:
        args.append("-B")
    subprocess.run(args, check=True, stdout=subprocess.DEVNULL)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        L = ctypes.CDLL(LIB_PATH)
        vp = ctypes.c_void_p
        L.oracle_run.argtypes = [vp, ctypes.c_int32, _d, _d, _d, _d, _d, _u8, ctypes.c_int32, ctypes.c_int32, _d]
        L.oracle_run.restype = ctypes.c_int32
        L.oracle_p2g.argtypes = [vp, ctypes.c_int32, _d, _d, _d, _d, _d, _u8, _d]
        L.oracle_grid_update.argtypes = [vp, _d]
        L.oracle_g2p.argtypes = [vp, ctypes.c_int32, _d, _d, _d, _d, _d, _u8, _d]
        L.oracle_weights.argtypes = [vp, ctypes.c_double, _i32, _d, _d]
        L.oracle_tau.argtypes = [vp, ctypes.c_int32, _u8, _d, _d, _d]
        L.oracle_polar.argtypes = [ctypes.c_int32, _d, _d]
        L.oracle_svd.argtypes = [ctypes.c_int32, _d, _d, _d, _d]
        L.oracle_snow_project.argtypes = [vp, ctypes.c_int32, _d, _d]
        L.oracle_bin_keys_f32.argtypes = [vp, ctypes.c_int32, _f32, _i32, ctypes.c_int32, ctypes.c_int32, _u32]
        L.oracle_stable_order.argtypes = [ctypes.c_int32, _u32, _i32]
        L.oracle_mlp_forward.argtypes = [ctypes.c_int32, _d, _d, _d, _d, _d, _d, _d, _d]
        L.oracle_mlp_loss_grad.argtypes = [ctypes.c_int32, _d, _d, _d] + [_d] * 12
        L.oracle_mlp_loss_grad.restype = ctypes.c_double
        L.oracle_adam.argtypes = [ctypes.c_int32, _d, _d, _d, _d, ctypes.c_double, ctypes.c_double, ctypes.c_double,
                                  ctypes.c_double, ctypes.c_int32]
        L.oracle_bin_nodes_f32.argtypes = [vp, ctypes.c_int32, _f32, ctypes.c_int32, ctypes.c_int32, _u32]
        L.oracle_layered_order.argtypes = [ctypes.c_int32, _u32, _u32, ctypes.c_int32, _i32]
        L.oracle_checker_slots.argtypes = [ctypes.c_int32, _u32, ctypes.c_int32, ctypes.c_int32, _u32]
        L.oracle_render.argtypes = [ctypes.c_int32, _f32, ctypes.c_int32, ctypes.c_int32, ctypes.c_int32,
                                    ctypes.c_float, ctypes.c_float, _u8]
        _lib = L
    return _lib


def _p(a, t=_d):
    return a.ctypes.data_as(t)


class State:
    """FP64 particle state: x, v [n,2]; C, F [n,4] row-major; J [n]; material uint8 [n]."""

    def __init__(self, x, v=None, C=None, F=None, J=None, material=None):
        n = x.shape[0]
        self.x = np.ascontiguousarray(x, dtype=np.float64).reshape(n, 2).copy()
        self.v = np.zeros((n, 2)) if v is None else np.ascontiguousarray(v, dtype=np.float64).reshape(n, 2).copy()
        self.C = np.zeros((n, 4)) if C is None else np.ascontiguousarray(C, dtype=np.float64).reshape(n, 4).copy()
        if F is None:
            self.F = np.zeros((n, 4))
            self.F[:, 0] = 1.0
            self.F[:, 3] = 1.0
        else:
            self.F = np.ascontiguousarray(F, dtype=np.float64).reshape(n, 4).copy()
        self.J = np.ones(n) if J is None else np.ascontiguousarray(J, dtype=np.float64).reshape(n).copy()
        self.material = (np.zeros(n, dtype=np.uint8) if material is None
                         else np.ascontiguousarray(material, dtype=np.uint8).reshape(n).copy())

    @property
    def n(self):
        return self.x.shape[0]

    def copy(self):
        return State(self.x, self.v, self.C, self.F, self.J, self.material)


def run(params, state, substeps, threads=1, want_grid=False):
    """Advance state (in place) by substeps; returns the last grid (mass, vx, vy planes) if asked."""
    G = params.grid_w * params.grid_h
    grid = np.zeros(3 * G) if want_grid else None
    rc = lib().oracle_run(ctypes.byref(params), state.n, _p(state.x), _p(state.v), _p(state.C), _p(state.F),
                          _p(state.J), _p(state.material, _u8), int(substeps), int(threads),
                          _p(grid) if want_grid else None)
    if rc != 0:
        raise ValueError("oracle_run rejected its arguments")
    if want_grid:
        return grid[:G].reshape(params.grid_h, params.grid_w), grid[G:2 * G].reshape(params.grid_h, params.grid_w), \
            grid[2 * G:].reshape(params.grid_h, params.grid_w)
    return None


def p2g(params, state):
    """Grid (mom_x, mom_y, mass) [H, W, 3] after one P2G of state."""
    grid = np.zeros((params.grid_h, params.grid_w, 3))
    lib().oracle_p2g(ctypes.byref(params), state.n, _p(state.x), _p(state.v), _p(state.C), _p(state.F), _p(state.J),
                     _p(state.material, _u8), _p(grid))
    return grid


def grid_update(params, grid):
    g = np.ascontiguousarray(grid, dtype=np.float64).copy()
    lib().oracle_grid_update(ctypes.byref(params), _p(g))
    return g


def g2p(params, state, grid):
    g = np.ascontiguousarray(grid, dtype=np.float64)
    lib().oracle_g2p(ctypes.byref(params), state.n, _p(state.x), _p(state.v), _p(state.C), _p(state.F), _p(state.J),
                     _p(state.material, _u8), _p(g))


def weights(params, x):
    base = ctypes.c_int32()
    f = ctypes.c_double()
    w = np.zeros(3)
    lib().oracle_weights(ctypes.byref(params), float(x), ctypes.byref(base), ctypes.byref(f), _p(w))
    return base.value, f.value, w


def tau(params, material, F, J):
    n = F.shape[0]
    F = np.ascontiguousarray(F, dtype=np.float64).reshape(n, 4)
    J = np.ascontiguousarray(J, dtype=np.float64).reshape(n)
    m = np.ascontiguousarray(material, dtype=np.uint8).reshape(n)
    out = np.zeros((n, 4))
    lib().oracle_tau(ctypes.byref(params), n, _p(m, _u8), _p(F), _p(J), _p(out))
    return out


def polar(F):
    F = np.ascontiguousarray(F, dtype=np.float64).reshape(-1, 4)
    R = np.zeros_like(F)
    lib().oracle_polar(F.shape[0], _p(F), _p(R))
    return R


def svd(F):
    F = np.ascontiguousarray(F, dtype=np.float64).reshape(-1, 4)
    n = F.shape[0]
    U, s, V = np.zeros((n, 4)), np.zeros((n, 2)), np.zeros((n, 4))
    lib().oracle_svd(n, _p(F), _p(U), _p(s), _p(V))
    return U, s, V


def snow_project(params, F, Jp):
    F = np.ascontiguousarray(F, dtype=np.float64).reshape(-1, 4).copy()
    Jp = np.ascontiguousarray(Jp, dtype=np.float64).reshape(-1).copy()
    lib().oracle_snow_project(ctypes.byref(params), F.shape[0], _p(F), _p(Jp))
    return F, Jp


def bin_keys(params, x, block_x, block_y, scene_of=None):
    x = np.ascontiguousarray(x, dtype=np.float32).reshape(-1, 2)
    n = x.shape[0]
    keys = np.zeros(n, dtype=np.uint32)
    so = None if scene_of is None else np.ascontiguousarray(scene_of, dtype=np.int32)
    lib().oracle_bin_keys_f32(ctypes.byref(params), n, _p(x, _f32), None if so is None else _p(so, _i32),
                              int(block_x), int(block_y), _p(keys, _u32))
    return keys


def stable_order(keys):
    keys = np.ascontiguousarray(keys, dtype=np.uint32)
    perm = np.zeros(keys.shape[0], dtype=np.int32)
    lib().oracle_stable_order(keys.shape[0], _p(keys, _u32), _p(perm, _i32))
    return perm


def bin_nodes(params, x, block_x, block_y):
    """Stencil base node of each particle within its bin (row-major, x fastest)."""
    x = np.ascontiguousarray(x, dtype=np.float32).reshape(-1, 2)
    nodes = np.zeros(x.shape[0], dtype=np.uint32)
    lib().oracle_bin_nodes_f32(ctypes.byref(params), x.shape[0], _p(x, _f32), int(block_x), int(block_y),
                               _p(nodes, _u32))
    return nodes


def checker_slots(nodes, block_x, block_y):
    """Checkerboard slot of each in-bin node: nodes with lx + ly even first, then odd, each in node order."""
    nodes = np.ascontiguousarray(nodes, dtype=np.uint32)
    slots = np.zeros(nodes.shape[0], dtype=np.uint32)
    lib().oracle_checker_slots(nodes.shape[0], _p(nodes, _u32), int(block_x), int(block_y), _p(slots, _u32))
    return slots


def bin_slots(params, x, block_x, block_y):
    """Checkerboard slot of each particle's stencil base node within its bin: the node label of the engine's
    layered order."""
    return checker_slots(bin_nodes(params, x, block_x, block_y), block_x, block_y)


LAYERS = 24


def layered_order(keys, nodes, layers=LAYERS):
    """Binned order of the engine: bins ascending, inside a bin layers of one particle per stencil node."""
    keys = np.ascontiguousarray(keys, dtype=np.uint32)
    nodes = np.ascontiguousarray(nodes, dtype=np.uint32)
    perm = np.zeros(keys.shape[0], dtype=np.int32)
    lib().oracle_layered_order(keys.shape[0], _p(keys, _u32), _p(nodes, _u32), int(layers), _p(perm, _i32))
    return perm


class MLP:
    """FP64 4-64-64-3 SiLU network (row = output unit)."""

    SHAPES = (("W1", (64, 4)), ("b1", (64,)), ("W2", (64, 64)), ("b2", (64,)), ("W3", (3, 64)), ("b3", (3,)))

    def __init__(self, params=None):
        self.p = {k: np.zeros(s) for k, s in self.SHAPES}
        if params is not None:
            for k, _ in self.SHAPES:
                self.p[k] = np.ascontiguousarray(params[k], dtype=np.float64).copy()

    def forward(self, inp):
        inp = np.ascontiguousarray(inp, dtype=np.float64).reshape(-1, 4)
        out = np.zeros((inp.shape[0], 3))
        P = self.p
        lib().oracle_mlp_forward(inp.shape[0], _p(inp), _p(P["W1"]), _p(P["b1"]), _p(P["W2"]), _p(P["b2"]),
                                 _p(P["W3"]), _p(P["b3"]), _p(out))
        return out

    def loss_grad(self, inp, target, wgt=(1.0, 1.0, 1.0)):
        inp = np.ascontiguousarray(inp, dtype=np.float64).reshape(-1, 4)
        target = np.ascontiguousarray(target, dtype=np.float64).reshape(-1, 3)
        w = np.ascontiguousarray(wgt, dtype=np.float64)
        g = {k: np.zeros(s) for k, s in self.SHAPES}
        P = self.p
        loss = lib().oracle_mlp_loss_grad(inp.shape[0], _p(inp), _p(target), _p(w), _p(P["W1"]), _p(P["b1"]),
                                          _p(P["W2"]), _p(P["b2"]), _p(P["W3"]), _p(P["b3"]), _p(g["W1"]),
                                          _p(g["b1"]), _p(g["W2"]), _p(g["b2"]), _p(g["W3"]), _p(g["b3"]))
        return loss, g


def adam(w, g, m, v, lr, beta1, beta2, eps, t):
    """One Adam step in place on flat float64 arrays."""
    lib().oracle_adam(w.shape[0], _p(w), _p(g), _p(m), _p(v), lr, beta1, beta2, eps, int(t))


def render(pos, width, height, chunk_px=160, scale=5.0, flip_extent=0.0):
    """Serial frame of the reference renderer: background, then every particle blended in index order.
    pos: float32 [n,2]; returns uint8 [height, width, 4]."""
    pos = np.ascontiguousarray(pos, dtype=np.float32).reshape(-1, 2)
    img = np.zeros((height, width, 4), np.uint8)
    lib().oracle_render(pos.shape[0], _p(pos, _f32), width, height, chunk_px, scale, flip_extent, _p(img, _u8))
    return img


def philox4x32_10(counters, key):
    """Philox4x32-10 (Salmon, Moraes, Dror, Shaw 2011, "Parallel random numbers: as easy as 1, 2, 3"):
    ten rounds of two 32x32->64-bit products with the multipliers 0xD2511F53 and 0xCD9E8D57, the key bumped by
    the Weyl constants 0x9E3779B9 and 0xBB67AE85 after each round.  counters: uint32 [n, 4]; key: two uint32.
    Returns uint32 [n, 4].  The reference's emitter draws its jitter from cuRAND's Philox4_32_10
    (MPM.cu:378, 416-418); the engine's emitter uses this generator with its own counter layout."""
    c = np.ascontiguousarray(counters, dtype=np.uint64).reshape(-1, 4)
    k0, k1 = np.uint64(int(key[0]) & 0xFFFFFFFF), np.uint64(int(key[1]) & 0xFFFFFFFF)
    m0, m1, lo = np.uint64(0xD2511F53), np.uint64(0xCD9E8D57), np.uint64(0xFFFFFFFF)
    for _ in range(10):
        p0 = m0 * c[:, 0]
        p1 = m1 * c[:, 2]
        c = np.stack([(p1 >> np.uint64(32)) ^ c[:, 1] ^ k0, p1 & lo, (p0 >> np.uint64(32)) ^ c[:, 3] ^ k1, p0 & lo],
                     axis=1)
        k0 = (k0 + np.uint64(0x9E3779B9)) & lo
        k1 = (k1 + np.uint64(0xBB67AE85)) & lo
    return c.astype(np.uint32)


def hose_pipe_positions(origin, velocity, width, rows, spacing, jitter, tick):
    """Positions of one hose-pipe emission in float64 (MPMEmitterKernel, MPM.cu:380-438, with the engine's
    generator): particle i = row * width + col sits on a width x rows lattice trailing the nozzle,
    across = (col - (width - 1) / 2) spacing along the nozzle normal n = (-t_y, t_x) and
    along = -(row + 1/2) spacing along the stream direction t, plus a Gaussian jitter: Philox4x32-10 on the
    counter (i, 0, 0, 0) with the key (tick low word, tick high word), u1 = ((r0 >> 8) + 1) / 2^24,
    u2 = (r1 >> 8) / 2^24 and Box-Muller (sqrt(-2 ln u1) cos 2 pi u2, sqrt(-2 ln u1) sin 2 pi u2) * jitter.
    Returns (lattice, jitter offsets), both float64 [width * rows, 2]."""
    n = width * rows
    vx, vy = float(np.float32(velocity[0])), float(np.float32(velocity[1]))
    sp = np.hypot(vx, vy)
    tx, ty = (vx / sp, vy / sp) if sp > 0 else (0.0, 0.0)
    i = np.arange(n)
    across = (i % width - 0.5 * (width - 1)) * float(np.float32(spacing))
    along = -((i // width) + 0.5) * float(np.float32(spacing))
    lattice = np.stack([origin[0] - ty * across + tx * along, origin[1] + tx * across + ty * along], axis=1)
    off = np.zeros((n, 2))
    if jitter > 0:
        ctr = np.zeros((n, 4), np.uint32)
        ctr[:, 0] = i
        r = philox4x32_10(ctr, (tick & 0xFFFFFFFF, tick >> 32))
        u1 = ((r[:, 0] >> 8).astype(np.float64) + 1.0) / 16777216.0
        u2 = (r[:, 1] >> 8).astype(np.float64) / 16777216.0
        rad = np.sqrt(-2.0 * np.log(u1))
        off[:, 0] = rad * np.cos(2.0 * np.pi * u2) * float(np.float32(jitter))
        off[:, 1] = rad * np.sin(2.0 * np.pi * u2) * float(np.float32(jitter))
    return lattice, off
