"""The output triple buffer on the GPU (OutputBuffer, cuda-engine.h:59-80 and cuda-engine.cpp:72-98;
MPMOutputKernel, MPM.cu:302-322), through the C-ABI: per-chunk particle counts of a population spanning
several output chunks, and a reader's slot that no tick overwrites while it is the read slot or the one
read before it."""
import ctypes

import numpy as np
import pytest

from mpm_amd import abi, scenes
from mpm_amd.engine import Engine

pytestmark = pytest.mark.gpu

_hip = None


def hip():
    global _hip
    if _hip is None:
        _hip = ctypes.CDLL("libamdhip64.so")
        _hip.hipMemcpy.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_int]
        _hip.hipMemcpy.restype = ctypes.c_int
    return _hip


def device_copy(ptr, shape, dtype):
    host = np.zeros(shape, dtype)
    assert hip().hipMemcpy(host.ctypes.data, ptr, host.nbytes, 2) == 0  # hipMemcpyDeviceToHost
    return host


def slot_bytes(view, n):
    return device_copy(view.pos, (n, 2), np.float32), device_copy(view.vel, (n, 2), np.float32)


def jelly_engine(n):
    p = abi.preset(abi.PRESET_JELLY128)
    p.substeps = 20
    p.sort_interval = 20
    p.capacity = n
    sc = scenes.jelly128(n=n)
    e = Engine(p)
    e.set_scene(sc)
    return e, p


def test_num_atoms_per_chunk_for_several_chunks():
    """5000 particles in chunks of 1536 (kMaxAtoms, constants.h:50): num_atoms = 1536, 1536, 1536, 392 —
    the full-chunk branch (i % 1536) + 1 == 1536 and the last-particle branch of MPM.cu:315-320 — and the
    slot holds the state in caller order."""
    e, p = jelly_engine(5000)
    with e:
        assert p.out_chunk == 1536
        e.step(1)
        e.sync()
        view = e.read_view()
        assert view.slot == 0 and view.count == 5000
        nums = device_copy(view.num_atoms, 4, np.int32)
        assert nums.tolist() == [1536, 1536, 1536, 392]
        pos, vel = slot_bytes(view, 5000)
        out = e.readback(("x", "v"))
        assert np.array_equal(pos, out["x"]) and np.array_equal(vel, out["v"])


def test_held_slot_is_never_overwritten():
    """A reader holds the published slot r.  mpm_step(h, 3), and three mpm_step(h, 1) calls without a
    publication, write slot r + 1 only; after the next publication (r + 1 is read) a tick writes r + 2, the
    cushion; only after one more publication may a tick write r again (the third slot of cuda-engine.h:59-61
    gives a reader one publication period)."""
    e, p = jelly_engine(3000)
    with e:
        e.step(1)
        e.sync()
        held = e.read_view()
        r = held.slot
        assert r == 0
        snap = slot_bytes(held, held.count)
        e.step(3)
        for _ in range(3):
            e.step(1)
        assert e.read_view().slot == r  # nothing published yet: the reader still sees r
        e.sync()
        now = e.read_view()
        assert now.slot == (r + 1) % 3
        latest = e.readback(("x", "v"))
        assert np.array_equal(slot_bytes(now, now.count)[0], latest["x"])
        for a, b in zip(snap, slot_bytes(held, held.count)):
            assert np.array_equal(a, b)
        # r + 1 is published, the writer moves to r + 2: r is still untouched
        e.step(2)
        e.sync()
        assert e.read_view().slot == (r + 2) % 3
        for a, b in zip(snap, slot_bytes(held, held.count)):
            assert np.array_equal(a, b)
        # two publications later the writer may reuse r, and the reader then sees a newer state there
        e.step(1)
        e.sync()
        assert e.read_view().slot == r
        assert not np.array_equal(snap[0], slot_bytes(held, held.count)[0])


def test_query_publishes_like_sync():
    """mpm_query returns MPM_BUSY or MPM_OK; once it returns MPM_OK the slot the ticks wrote is the read
    slot (the reference rolls after cudaEventQuery(tick_done) succeeds, MPM-master.cpp:83-89)."""
    e, p = jelly_engine(2000)
    with e:
        assert e.read_view().slot == -1
        e.step(2)
        while not e.query():
            pass
        view = e.read_view()
        assert view.slot == 0 and view.count == 2000
        pos, _ = slot_bytes(view, 2000)
        assert np.array_equal(pos, e.readback(("x",))["x"])
