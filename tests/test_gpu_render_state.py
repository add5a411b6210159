"""GPU tests of the headless rasteriser (against the serial frame restatement, byte for byte), of state files
(save -> parse -> load round trips, batched scenes, rejected files) and of the headless driver."""
import os
import subprocess
import sys

import numpy as np
import pytest

from mpm_amd import abi, render, scenes, state
from mpm_amd.engine import Engine, MpmError

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
DRIVER = os.path.join(ROOT, "gpu-physics-engine_amd", "lib", "mpm_run")


def reference_scene(n, seed):
    """Particles over the reference domain (384 x 192 units), a dense clump and a few on the borders."""
    rng = np.random.default_rng(seed)
    x = np.concatenate([rng.uniform([0, 0], [384, 192], (n, 2)),
                        rng.normal([100, 60], 1.5, (n // 4, 2)).clip(0, 192),
                        np.array([[0.0, 0.0], [384.0, 192.0], [0.1, 191.9], [383.9, 0.05]])]).astype(np.float32)
    return x


@pytest.mark.parametrize("ticks", [1, 3])
def test_render_matches_serial_frame(oracle_mod, ticks):
    p = abi.preset(abi.PRESET_REFERENCE)
    x = reference_scene(20000, 5 + ticks)
    with Engine(p) as e:
        e.set_particles(x)
        e.step(ticks)
        e.sync()
        view = e.read_view()
        pos = e.readback()["x"]
        assert view.count == pos.shape[0]
        with render.Renderer(**render.REFERENCE) as r:
            img = r.frame(view)
    ref = oracle_mod.render(pos, 1920, 960, chunk_px=160, scale=5.0)
    assert np.array_equal(img, ref), np.argwhere((img != ref).any(axis=2))[:5]


def test_render_flipped_unit_scene(oracle_mod):
    p = abi.preset(abi.PRESET_JELLY128)
    sc = scenes.jelly128()
    with Engine(p) as e:
        e.set_scene(sc)
        e.step(1)
        e.sync()
        pos = e.readback()["x"]
        with render.Renderer(256, 256, scale=256.0, chunk_px=0, flip_extent=1.0, row_pitch=4 * 256 + 64) as r:
            img = r.frame(e.read_view())
    ref = oracle_mod.render(pos, 256, 256, chunk_px=0, scale=256.0, flip_extent=1.0)
    assert np.array_equal(img, ref)


def test_renderer_rejects_bad_descriptions():
    for kw in (dict(width=0, height=10, scale=1.0), dict(width=10, height=10, scale=0.0),
               dict(width=10, height=10, scale=1.0, row_pitch=20), dict(width=10, height=10, scale=1.0, chunk_px=-1)):
        with pytest.raises(render.RenderError) as ei:
            render.Renderer(**kw)
        assert ei.value.code == abi.MPM_ERR_INVALID_ARG


def test_state_save_parse_load(tmp_path):
    p = abi.preset(abi.PRESET_MIXED1024)
    sc = scenes.for_preset(abi.PRESET_MIXED1024, scale=1 / 256)
    p.capacity = sc.n
    p.substeps = 40
    path = tmp_path / "mixed.mpmstate"
    with Engine(p) as e:
        e.set_scene(sc)
        e.step(2)
        e.sync()
        before = e.readback()
        e.save_state(path)
        e.step(1)
        e.sync()
        after = e.readback()
    st = state.read_state(path)
    assert st["n"] == sc.n and st["num_scenes"] == 1 and (st["scene"] == 0).all()
    for k in ("x", "v", "C", "F", "J", "material"):
        assert np.array_equal(st[k].reshape(before[k].shape), before[k]), k
    with Engine(p) as e2:
        e2.load_state(path)
        got = e2.readback()
        for k in ("x", "v", "C", "F", "J", "material"):
            assert np.array_equal(got[k], before[k]), k
        e2.step(1)
        e2.sync()
        cont = e2.readback()
    # float atomics make the grid sums order-dependent in the last bits; the continuation agrees closely
    rel = np.linalg.norm(cont["x"] - after["x"]) / np.linalg.norm(after["x"])
    assert rel < 1e-5


def test_state_keeps_batched_scenes(tmp_path):
    p = abi.preset(abi.PRESET_BATCH64X256)
    sc = scenes.for_preset(abi.PRESET_BATCH64X256, scale=1 / 64)
    p.capacity = sc.n
    p.substeps = 20
    path = tmp_path / "batch.mpmstate"
    with Engine(p) as e:
        e.set_scene(sc)
        e.step(1)
        e.sync()
        e.save_state(path)
        ref_mass = e.read_grid()[0].sum(axis=(1, 2))
    st = state.read_state(path)
    expect = np.repeat(np.arange(p.num_scenes), sc.scene_counts)
    assert np.array_equal(st["scene"], expect)
    with Engine(p) as e2:
        e2.load_state(path)
        e2.step(1)
        e2.sync()
        mass = e2.read_grid()[0].sum(axis=(1, 2))
    # every scene keeps its particles: per-scene grid mass is count * particle mass
    assert np.allclose(mass, np.asarray(sc.scene_counts) * np.float32(p.mass), rtol=1e-4)
    assert np.allclose(ref_mass, mass, rtol=1e-4)


def test_state_rejects_bad_files(tmp_path):
    p = abi.preset(abi.PRESET_JELLY128)
    sc = scenes.jelly128(n=4096)
    p.capacity = sc.n
    good = tmp_path / "good.mpmstate"
    with Engine(p) as e:
        e.set_scene(sc)
        e.save_state(good)
        raw = bytearray(good.read_bytes())
        cases = {"flip": raw[:-1] + bytes([raw[-1] ^ 1]), "trunc": raw[:-5], "magic": b"X" + raw[1:],
                 "extra": raw + b"\0"}
        for name, data in cases.items():
            bad = tmp_path / f"{name}.mpmstate"
            bad.write_bytes(bytes(data))
            with pytest.raises(MpmError) as ei:
                e.load_state(bad)
            assert ei.value.code == abi.MPM_ERR_INVALID_ARG, name
        with pytest.raises(MpmError):
            e.load_state(tmp_path / "missing.mpmstate")
    other = abi.preset(abi.PRESET_MIXED1024)
    other.capacity = sc.n
    with Engine(other) as e:
        with pytest.raises(MpmError) as ei:
            e.load_state(good)
        assert ei.value.code == abi.MPM_ERR_INVALID_ARG
    small = abi.preset(abi.PRESET_JELLY128)
    small.capacity = sc.n // 2
    with Engine(small) as e:
        with pytest.raises(MpmError) as ei:
            e.load_state(good)
        assert ei.value.code == abi.MPM_ERR_CAPACITY


def test_headless_driver(tmp_path, oracle_mod):
    """The reference tick loop: hose-pipe emission to 100,080 particles would take 417 ticks; 160 ticks with
    a frame every 80 and the final state saved."""
    assert os.path.exists(DRIVER), "driver not built"
    save = tmp_path / "end.mpmstate"
    print(f"[driver] {DRIVER} --ticks 160", flush=True)
    out = subprocess.run([DRIVER, "--ticks", "160", "--frames-every", "80", "--out", str(tmp_path), "--save", str(save)],
                         capture_output=True, text=True, timeout=120)
    sys.stdout.write(out.stdout[-2000:])
    assert out.returncode == 0, out.stderr
    assert "------- Tick 150 -------" in out.stdout and "Simulating 36000 atoms" in out.stdout
    last = out.stdout.strip().splitlines()[-1]
    assert last.startswith("mpm_run: ticks=160 particles=38400")
    st = state.read_state(save)
    assert st["n"] == 38400
    for t in (80, 160):
        f = tmp_path / f"frame_{t:05d}.ppm"
        data = f.read_bytes()
        assert data.startswith(b"P6\n1920 960\n255\n") and len(data) == 16 + 1920 * 960 * 3
    rgb = np.frombuffer((tmp_path / "frame_00160.ppm").read_bytes()[16:], np.uint8).reshape(960, 1920, 3)
    ref = oracle_mod.render(st["x"], 1920, 960, chunk_px=160, scale=5.0)
    assert np.array_equal(rgb, ref[..., :3])
