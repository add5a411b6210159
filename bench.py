"""Benchmark: MLS-MPM particle-substeps/s on one MI355X.

Workload (default): the 1024-cell mixed scene — 4,194,304 particles in three seeded boxes (fluid bulk 400,
fixed-corotated jelly E = 1e4 nu = 0.2, snow theta_c = 2.5e-2 theta_s = 7.5e-3 hardening 10), initial
velocity uniform in [-1,1]^2, dt = 2e-5.  One step = one tick of 100 substeps (the workload's run length),
launched as one hipGraph; particle state is resident in HBM before the timed region.

Prints one JSON line (rank 0): throughput, per-step time, the roofline of the dominant kernel measured
with HIP events on the engine's stream, and the FP64 oracle timed on this host's cores (bounded sample).
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.join(ROOT, "gpu-physics-engine_amd"))

from mpm_amd import abi, scenes  # noqa: E402
from mpm_amd.engine import Engine  # noqa: E402

METRIC = "MPM particle-substeps/s at 4M particles; learned+batched vs plain at ≤1e-3 err"
HBM_PEAK_GBS = 8000.0

# Algorithmic bytes per unit of the fused G2P->P2G kernel (k_g2p2g):
#   per particle: read x,y 8 + F 16 + J 4 + material 1 = 29 B, write x,y 8 + F 16 + J 4 = 28 B -> 57 B
#   per grid node: read the previous substep's mass/momentum 12 B (velocities are computed in the kernel: the
#   grid update is folded in) + accumulate mass/momentum 12 B + zero the set of the substep after next 12 B
#   -> 36 B
G2P2G_BYTES_PER_PARTICLE = 57
G2P2G_BYTES_PER_NODE = 36
# Whole-substep algorithmic bytes of the unfused three-kernel form: 132 B per particle, 40 B per node.
SUBSTEP_BYTES_PER_PARTICLE = 132
SUBSTEP_BYTES_PER_NODE = 40


def binning_bytes_per_particle(p):
    """Algorithmic bytes per particle of one binning: keys from x, y (read 8, write 4); P radix passes of 8
    bits over the bin bits (the histogram reads the key 4, the scatter reads and writes (key, id) 16); the
    bin ranges read the key 4; the layered order reads the key twice and the id and writes both (20); the
    full-state gather reads the id 4 and 57 B of state and writes 57 B."""
    bins = p.num_scenes * (-(-p.grid_w // p.block_x)) * (-(-p.grid_h // p.block_y))
    passes = -(-max(1, (bins - 1).bit_length()) // 8)
    return 12 + 20 * passes + 4 + 20 + 4 + 2 * 57


WORKLOADS = {
    # name: (preset, scene factory, description)
    "mixed1024": (abi.PRESET_MIXED1024, lambda scale: scenes.mixed(n=int(4194304 * scale)),
                  "mixed1024: 4,194,304 particles (fluid/jelly/snow boxes) on a 1024x1024-cell grid, dt=2e-5, "
                  "100 substeps per step"),
    "jelly128": (abi.PRESET_JELLY128, lambda scale: scenes.jelly128(n=int(32768 * scale)),
                 "jelly128: 32,768-particle jelly block on a 128x128-cell grid, dt=1e-4, 100 substeps per step"),
    "batch64x256": (abi.PRESET_BATCH64X256, lambda scale: scenes.batched(scale=scale),
                    "batch64x256: 64 scenes of 256x256 cells, particle counts log-uniform in [4096, 524288], "
                    "mixed materials, dt=2e-5, 200 substeps per step"),
    "learned1024": (abi.PRESET_LEARNED1024, lambda scale: scenes.for_preset(abi.PRESET_LEARNED1024, scale),
                    "learned1024: 4,194,304 jelly + snow particles on a 1024x1024-cell grid, BF16 MLP closure, "
                    "dt=2e-5, 100 substeps per step"),
    "learned16x2048": (abi.PRESET_LEARNED16X2048, lambda scale: scenes.for_preset(abi.PRESET_LEARNED16X2048, scale),
                       "learned16x2048: 16 scenes of 2048x2048 cells, 1,048,576-16,777,216 jelly + snow particles "
                       "each, BF16 MLP closure, dt=1e-5, 100 substeps per step"),
}
CLOSURE_WEIGHTS = os.path.join(ROOT, "gpu-physics-engine_amd", "mpm_amd", "data", "closure_learned1024.npy")


def traffic_from_profile(p, workload):
    """HBM bytes per launch of k_g2p2g measured by rocprofv3 PMC for this workload and launch shape
    (profiles/g2p2g_<workload>_traffic.json), or None."""
    path = os.path.join(ROOT, "profiles", f"g2p2g_{workload}_traffic.json")
    if not os.path.exists(path):
        return None
    d = json.load(open(path))
    if list(d.get("tile", [])) != [p.block_x, p.block_y] or d.get("chunk") != p.chunk:
        return None
    return d["traffic_bytes"]


def log(msg):
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def cpu_baseline(params, scene, budget_s=12.0):
    """FP64 oracle (OpenMP, per-thread private grids) on this host: particle-substeps/s on the full
    particle set for as many substeps as fit the budget (at least 1, at most 20)."""
    import oracle
    threads = min(16, len(os.sched_getaffinity(0)))
    if params.num_scenes > 1:
        # one scene at a time (each scene has its own grid), as many scenes as fit the budget
        done, t_all, k = 0, 0.0, 0
        for k, sl in enumerate(scene.scene_slices()):
            s = oracle.State(scene.x[sl], scene.v[sl], material=scene.material[sl])
            t0 = time.perf_counter()
            oracle.run(params, s, 1, threads=threads)
            t_all += time.perf_counter() - t0
            done += sl.stop - sl.start
            if t_all > budget_s:
                break
        return {"value": done / t_all, "unit": "particle-substeps/s", "cores": threads, "kind": "port",
                "sample": f"FP64 oracle, OpenMP {threads} threads, 1 substep of each of the first {k + 1} scenes "
                          f"({done} particles, {t_all:.1f} s)"}
    s = oracle.State(scene.x, scene.v, material=scene.material)
    t0 = time.perf_counter()
    oracle.run(params, s, 1, threads=threads)
    t1 = time.perf_counter() - t0
    k = int(max(1, min(20, budget_s / max(t1, 1e-3))))
    t0 = time.perf_counter()
    oracle.run(params, s, k, threads=threads)
    dt = time.perf_counter() - t0
    return {"value": scene.n * k / dt, "unit": "particle-substeps/s", "cores": threads, "kind": "port",
            "sample": f"FP64 oracle, OpenMP {threads} threads, full {scene.n}-particle scene, {k} substeps "
                      f"({dt:.1f} s) after 1 untimed substep"}


def copy_calibration(mib, reps=5):
    """Device-to-device copies of a known size (hipMemcpy), so profiler byte counters can be scaled to HBM
    bytes: each copy reads and writes exactly mib MiB."""
    import ctypes
    hip = ctypes.CDLL("libamdhip64.so")
    n = ctypes.c_size_t(mib << 20)
    a, b = ctypes.c_void_p(), ctypes.c_void_p()
    assert hip.hipMalloc(ctypes.byref(a), n) == 0 and hip.hipMalloc(ctypes.byref(b), n) == 0
    hip.hipMemset(a, 1, n)
    for _ in range(reps):
        assert hip.hipMemcpy(b, a, n, 3) == 0  # hipMemcpyDeviceToDevice
    hip.hipDeviceSynchronize()
    hip.hipFree(a)
    hip.hipFree(b)
    log(f"copy calibration: {reps} device copies of {mib} MiB")


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--mode", choices=["reference", "tiled", "fused"], default="fused")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--scale", type=float, default=1.0, help="particle count scale (1.0 = 4,194,304)")
    ap.add_argument("--autotune", action="store_true", help="pick tile shape, chunk and sort interval by mpm_autotune")
    ap.add_argument("--sort-interval", type=int, default=0, metavar="SUBSTEPS",
                    help="substeps between binnings (0 = the preset's; above the tick's substeps: every k-th tick)")
    ap.add_argument("--chunk", type=int, default=0, metavar="PARTICLES",
                    help="particles per work item (0 = the preset's)")
    ap.add_argument("--workload", choices=sorted(WORKLOADS), default="mixed1024",
                    help="mixed1024 is the headline; the others are the remaining configurations")
    ap.add_argument("--copy-calibration", type=int, default=0, metavar="MIB",
                    help="also run device copies of MIB MiB (profiler counter calibration)")
    args = ap.parse_args()
    if args.copy_calibration:
        copy_calibration(args.copy_calibration)
    if args.gpus != 1:
        print(f"bench.py: the engine runs on one GPU; --gpus {args.gpus} is not supported (sharding across GPUs "
              f"is designed, not implemented)", file=sys.stderr)
        sys.exit(2)

    preset, make_scene, description = WORKLOADS[args.workload]
    p = abi.preset(preset)
    p.mode = {"reference": abi.MODE_REFERENCE, "tiled": abi.MODE_TILED, "fused": abi.MODE_FUSED}[args.mode]
    if args.sort_interval > 0:
        p.sort_interval = args.sort_interval
    if args.chunk > 0:
        p.chunk = args.chunk
    scene = make_scene(args.scale)
    p.capacity = scene.n
    log(f"scene: {scene.description}")
    if args.autotune and p.mode != abi.MODE_REFERENCE:
        from mpm_amd.engine import autotune
        t0 = time.perf_counter()
        p, res = autotune(p, scene)
        log(f"autotune: {res.candidates} candidates in {time.perf_counter() - t0:.1f} s -> tile {res.block_x}x"
            f"{res.block_y}, chunk {res.chunk}, sort interval {res.sort_interval}, "
            f"{res.ms_per_substep * 1e3:.1f} us/substep")
    eng = Engine(p)
    eng.set_scene(scene)
    if p.closure == abi.CLOSURE_LEARNED:
        eng.set_closure(np.load(CLOSURE_WEIGHTS))
        log(f"learned closure loaded from {os.path.relpath(CLOSURE_WEIGHTS, ROOT)}")
    log("engine created, state resident in HBM")
    for _ in range(args.warmup):
        eng.step(1)
    eng.sync()
    log(f"warmup {args.warmup} ticks done")
    # Per-kernel HIP-event timing on the engine's stream: two binning cycles of eagerly launched ticks (a
    # cycle is one tick, or every tick up to the next binning when sort_interval spans several) right
    # before and two right after the timed region (events cannot sit inside the tick graph).  The scene
    # compresses under gravity and the fused kernel slows as it does, so they bracket the timed region and
    # their sum stands for its average (single ticks vary by a few per cent).
    cycle = 2 * p.substeps * (max(1, -(-p.sort_interval // p.substeps)) if p.mode != abi.MODE_REFERENCE else 1)
    kt_before = eng.time_kernels(cycle)
    t0 = time.perf_counter()
    eng.step(args.steps)
    eng.sync()
    wall = time.perf_counter() - t0
    last_tick_ms = eng.last_step_ms()
    ms_per_step = wall * 1e3 / args.steps
    value = scene.n * p.substeps * args.steps / wall
    log(f"timed {args.steps} ticks: {ms_per_step:.3f} ms/tick (last tick by HIP events {last_tick_ms:.3f} ms)")
    kt_after = eng.time_kernels(cycle)
    kt = {k: (kt_before[k][0] + kt_after[k][0], kt_before[k][1] + kt_after[k][1]) for k in kt_before}
    kernels = {k: {"ms_total": round(ms, 4), "launches": n, "us_avg": round(1e3 * ms / n, 2) if n else None}
               for k, (ms, n) in kt.items()}
    out = eng.readback()
    eng.close()
    finite = bool(np.all(np.isfinite(out["x"])))
    log(f"kernels: {json.dumps(kernels)}")

    nodes = p.grid_w * p.grid_h * p.num_scenes
    dominant = "g2p2g" if p.mode == abi.MODE_FUSED else ("p2g" if p.mode == abi.MODE_REFERENCE else "g2p")
    ms_dom, n_dom = kt[dominant]
    if dominant == "g2p2g":
        bytes_per_launch = scene.n * G2P2G_BYTES_PER_PARTICLE + nodes * G2P2G_BYTES_PER_NODE
    elif dominant == "p2g":
        bytes_per_launch = scene.n * 53 + nodes * 12
    else:
        bytes_per_launch = scene.n * 81 + nodes * 8
    achieved = bytes_per_launch / (ms_dom / n_dom * 1e-3) / 1e9 if n_dom else 0.0
    substep_s = wall / (args.steps * p.substeps)
    substep_bytes = scene.n * SUBSTEP_BYTES_PER_PARTICLE + nodes * SUBSTEP_BYTES_PER_NODE
    # the binning on its own line, amortised over the substeps that the event-timed cycles covered
    ms_bin, n_bin = kt["binning"]
    binning = None
    if n_bin:
        us_sort = 1e3 * ms_bin / n_bin
        bin_bytes = scene.n * binning_bytes_per_particle(p)
        binning = {"us_per_sort": us_sort, "bytes_per_sort": bin_bytes,
                   "achieved": bin_bytes / (us_sort * 1e-6) / 1e9, "unit": "GB/s",
                   "frac": bin_bytes / (us_sort * 1e-6) / 1e9 / HBM_PEAK_GBS,
                   "substeps_per_sort": 2 * cycle / n_bin,
                   "us_per_substep_amortised": 1e3 * ms_bin / (2 * cycle)}

    result = {
        "metric": METRIC,
        "value": value,
        "unit": "particle-substeps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_per_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {"workload": description if args.scale == 1.0 else
                               f"{args.workload} scaled x{args.scale}: {scene.n} particles",
                   "particles": scene.n, "scenes": p.num_scenes, "grid_cells": p.grid_w - 4,
                   "substeps_per_step": p.substeps,
                   "mode": args.mode, "tile": [p.block_x, p.block_y], "chunk": p.chunk,
                   "sort_interval": p.sort_interval},
        "roofline": {"bound": "hbm", "kernel": dominant, "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS,
                     "traffic": traffic_from_profile(p, args.workload) if dominant == "g2p2g" else None,
                     "bytes_per_launch": bytes_per_launch,
                     "us_per_launch": 1e3 * ms_dom / n_dom if n_dom else None},
        "substep": {"us": substep_s * 1e6, "algorithmic_bytes_unfused": substep_bytes,
                    "effective_GBs": substep_bytes / substep_s / 1e9},
        "binning": binning,
        "kernels": kernels,
        "state_finite": finite,
    }
    if p.closure == abi.CLOSURE_LEARNED:
        result["vs_plain"] = compare_plain(p, scene, args, value)
        result["roofline"]["mfma_tflops"] = scene.n * 2 * 64 * 64 / (ms_dom / n_dom * 1e-3) / 1e12 if n_dom else None
    if not args.no_cpu_baseline:
        log("cpu baseline (FP64 oracle) ...")
        result["cpu_baseline"] = cpu_baseline(p, scene)
    print(json.dumps(result), flush=True)


def compare_plain(p, scene, args, learned_value):
    """The same scene with the analytic closure: throughput over the same number of ticks, and the relative
    position error of the learned closure after one tick from the initial state."""
    pa = p.copy()
    pa.closure = abi.CLOSURE_ANALYTIC
    with Engine(pa) as e:
        e.set_scene(scene)
        e.step(1)
        e.sync()
        x_plain = e.readback()["x"].astype(np.float64)
        for _ in range(args.warmup):
            e.step(1)
        e.sync()
        t0 = time.perf_counter()
        e.step(args.steps)
        e.sync()
        plain_value = scene.n * pa.substeps * args.steps / (time.perf_counter() - t0)
    with Engine(p) as e:
        e.set_scene(scene)
        e.set_closure(np.load(CLOSURE_WEIGHTS))
        e.step(1)
        e.sync()
        x_learned = e.readback()["x"].astype(np.float64)
    err = float(np.linalg.norm(x_learned - x_plain) / np.linalg.norm(x_plain))
    log(f"plain (analytic) on the same scene: {plain_value:.4g} particle-substeps/s; learned / plain "
        f"{learned_value / plain_value:.3f}; relative position error after {p.substeps} substeps {err:.2e}")
    return {"plain_value": plain_value, "learned_over_plain": learned_value / plain_value,
            "rel_position_error_vs_plain": err, "substeps_compared": p.substeps}


if __name__ == "__main__":
    main()
