"""Reader and writer of the engine's state files (mpm_save_state / mpm_load_state, include/mpm_engine.h).

Layout (little-endian): 200-byte header {char magic[8] = "MPMSTATE"; uint32 version = 1; uint32 header_bytes
= 200; int32 n; int32 num_scenes; uint64 checksum (64-bit FNV-1a of the payload); MpmParams (164 bytes);
uint32 reserved = 0}, then x f32[2n] | v f32[2n] | C f32[4n] | F f32[4n] | J f32[n] | scene i32[n] |
material u8[n].
"""
import ctypes
import struct

import numpy as np

from . import abi

MAGIC = b"MPMSTATE"
VERSION = 1
HEADER_BYTES = 200
_FIXED = struct.Struct("<8sIIiiQ")


def fnv1a64(data):
    """64-bit FNV-1a (offset basis 0xcbf29ce484222325, prime 0x100000001b3) of a bytes-like object."""
    h = 0xCBF29CE484222325
    prime = 0x100000001B3
    for b in memoryview(data).cast("B"):
        h = ((h ^ b) * prime) & 0xFFFFFFFFFFFFFFFF
    return h


def _payload(st):
    n = st["x"].shape[0]
    parts = [np.ascontiguousarray(st[k], np.float32).reshape(-1) for k in ("x", "v", "C", "F", "J")]
    parts.append(np.ascontiguousarray(st["scene"], np.int32).reshape(-1))
    out = b"".join(p.tobytes() for p in parts) + np.ascontiguousarray(st["material"], np.uint8).tobytes()
    assert len(out) == 57 * n
    return out


def read_state(path, verify=True):
    """Parse a state file; returns dict with n, num_scenes, checksum, params (MpmParams) and the arrays."""
    raw = open(path, "rb").read()
    magic, version, header_bytes, n, num_scenes, checksum = _FIXED.unpack_from(raw, 0)
    if magic != MAGIC or version != VERSION or header_bytes != HEADER_BYTES:
        raise ValueError("not a version-1 MPM state file")
    params = abi.MpmParams.from_buffer_copy(raw[32:32 + ctypes.sizeof(abi.MpmParams)])
    body = raw[HEADER_BYTES:]
    if len(body) != 57 * n:
        raise ValueError("payload size does not match the particle count")
    off = 0
    out = {"n": n, "num_scenes": num_scenes, "checksum": checksum, "params": params}
    for key, width, dt in (("x", 2, np.float32), ("v", 2, np.float32), ("C", 4, np.float32), ("F", 4, np.float32),
                           ("J", 1, np.float32), ("scene", 1, np.int32), ("material", 1, np.uint8)):
        cnt = n * width
        a = np.frombuffer(body, dt, cnt, off)
        off += cnt * np.dtype(dt).itemsize
        out[key] = a.reshape(n, width) if width > 1 else a.copy()
    if verify and fnv1a64(body) != checksum:
        raise ValueError("checksum mismatch")
    return out


def write_state(path, params, x, v, C, F, J, material, scene=None):
    """Write a state file from host arrays (same format as mpm_save_state)."""
    n = np.asarray(x).reshape(-1, 2).shape[0]
    st = {"x": x, "v": v, "C": C, "F": F, "J": J, "material": material,
          "scene": np.zeros(n, np.int32) if scene is None else scene}
    body = _payload(st)
    head = _FIXED.pack(MAGIC, VERSION, HEADER_BYTES, n, params.num_scenes, fnv1a64(body))
    head += bytes(params) + b"\0" * 4
    assert len(head) == HEADER_BYTES
    with open(path, "wb") as f:
        f.write(head + body)
