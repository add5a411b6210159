"""Benchmark: MLS-MPM particle-substeps/s on one MI355X.

Workload (default): the 1024-cell mixed scene — 4,194,304 particles in three seeded boxes (fluid bulk 400,
fixed-corotated jelly E = 1e4 nu = 0.2, snow theta_c = 2.5e-2 theta_s = 7.5e-3 hardening 10), initial
velocity uniform in [-1,1]^2, dt = 2e-5, particle density 16.  One step = one tick of 100 substeps (the
workload's run length), launched as one hipGraph; particle state is resident in HBM before the timed region.

Prints one JSON line: throughput, per-step time, the roofline of the dominant kernel timed with HIP events on
the engine's stream, the FP64 oracle timed on this host's cores (bounded sample), and three more workloads
timed the same way: the 64-scene batch (`batched`), the learned closure on the 4,194,304-particle jelly +
snow scene against the analytic engine on the same scene (`learned`), and the learned closure on 16 scenes of
2048^2 cells (91.5M particles) against the analytic engine on the same scenes (`learned_batched`).
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
# Kernel time per step (events, a replay of the timed window) may exceed the timed step by this much (run to
# run spread of the atomic-order noise and of the clocks) before the bench refuses its own roofline.
CONSISTENCY_SLACK = 1.02


def binning_bytes_per_particle(p):
    """Algorithmic bytes per particle of one steady-state binning: keys from x, y (read 8, write 4); P radix
    passes of 8 bits over the bin bits (the histogram reads the key 4, the scatter reads and writes (key, id)
    16); the bin ranges read the key 4; the layered order reads the key twice and the id and writes both
    (20); the gather reads the permutation 4 and moves the state: 57 B each way, or 33 B in the fused mode,
    whose ticks bin after their first P2G and leave v and C (24 B) behind."""
    bins = p.num_scenes * (-(-p.grid_w // p.block_x)) * (-(-p.grid_h // p.block_y))
    passes = -(-max(1, (bins - 1).bit_length()) // 8)
    state = 33 if p.mode == abi.MODE_FUSED else 57
    return 12 + 20 * passes + 4 + 20 + 4 + 2 * state


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
    """HBM bytes per launch of k_g2p2g measured by rocprofv3 PMC (FETCH_SIZE, WRITE_SIZE, each scaled by the
    same counter on a device copy of known size) for this workload and launch shape, as committed in
    profiles/g2p2g_<workload>_traffic.json; (bytes, source) or (None, None)."""
    name = f"g2p2g_{workload}_traffic.json"
    path = os.path.join(ROOT, "profiles", name)
    if not os.path.exists(path):
        return None, None
    d = json.load(open(path))
    if list(d.get("tile", [])) != [p.block_x, p.block_y] or d.get("chunk") != p.chunk:
        return None, None
    return d["traffic_bytes"], f"profiles/{name} (rocprofv3 --pmc FETCH_SIZE, WRITE_SIZE; replayed, not measured here)"


def log(msg):
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def scene_layout(workload, scene, p):
    """What the synthetic scene is: density, seed, boxes (or per-scene counts)."""
    rho = float(p.mass) / float(p.vol)
    out = {"rho": round(rho, 6), "seed": int(scene.seed), "dt": float(np.float32(p.dt))}
    if workload in ("mixed1024", "learned1024"):
        out["boxes"] = [[list(lo), list(hi)] for lo, hi in scenes.MIXED_BOXES]
        out["box_materials"] = ["fluid", "jelly", "snow"] if workload == "mixed1024" else ["jelly", "jelly", "snow"]
        out["v0"] = "uniform in [-1,1]^2"
    elif p.num_scenes > 1:
        out["scene_counts_min_max"] = [int(scene.scene_counts.min()), int(scene.scene_counts.max())]
    return out


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


def timed_run(p, scene, steps, warmup, weights=None, kernel_timing=True, label=""):
    """Warm-up ticks, then `steps` ticks timed by the host clock around mpm_step + mpm_sync (the engine's
    stream is idle before and after).  With kernel_timing, a second engine from the same initial state runs
    the same warm-up ticks and then the same window of ticks through mpm_time_kernels (each run of consecutive
    launches of one class captured into a graph and launched between HIP events on the engine's stream), so
    the per-class kernel times belong to exactly the ticks that were timed (the scene compresses under gravity
    and the fused kernel slows as it does).  Returns (result dict, readback of x after the timed ticks)."""
    def engine():
        e = Engine(p)
        e.set_scene(scene)
        if weights is not None:
            e.set_closure(weights)
        for _ in range(warmup):
            e.step(1)
        e.sync()
        return e

    with engine() as eng:
        t0 = time.perf_counter()
        eng.step(steps)
        eng.sync()
        wall = time.perf_counter() - t0
        last_tick_ms = eng.last_step_ms()
        out = eng.readback(("x",))
    ms_per_step = wall * 1e3 / steps
    value = scene.n * p.substeps * steps / wall
    log(f"{label}timed {steps} ticks: {ms_per_step:.3f} ms/tick (last tick by HIP events {last_tick_ms:.3f} ms), "
        f"{value:.4g} particle-substeps/s")
    res = {"value": value, "ms_per_step": ms_per_step, "steps": steps, "warmup": warmup,
           "last_tick_ms_events": last_tick_ms}
    if kernel_timing:
        with engine() as twin:
            res["kt"] = twin.time_kernels(steps * p.substeps)
        res["kt_substeps"] = steps * p.substeps
    return res, out["x"]


def roofline_of(p, scene, run, workload):
    """Roofline object of the dominant kernel from the graph-timed kernel classes of a timed run."""
    kt = run["kt"]
    nodes = p.grid_w * p.grid_h * p.num_scenes
    dominant = "g2p2g" if p.mode == abi.MODE_FUSED else ("p2g" if p.mode == abi.MODE_REFERENCE else "g2p")
    ms_dom, n_dom = kt[dominant]
    if dominant == "g2p2g":
        bytes_per_launch = scene.n * G2P2G_BYTES_PER_PARTICLE + nodes * G2P2G_BYTES_PER_NODE
    elif dominant == "p2g":
        bytes_per_launch = scene.n * 53 + nodes * 12
    else:
        bytes_per_launch = scene.n * 81 + nodes * 8
    us = 1e3 * ms_dom / n_dom if n_dom else None
    achieved = bytes_per_launch / (us * 1e-6) / 1e9 if n_dom else 0.0
    ticks = run["kt_substeps"] / p.substeps
    launches_per_step = n_dom / ticks
    kernel_ms_per_step = us * 1e-3 * launches_per_step if n_dom else 0.0
    traffic, source = (traffic_from_profile(p, workload) if dominant == "g2p2g" else (None, None))
    return {"bound": "hbm", "kernel": dominant, "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
            "frac": achieved / HBM_PEAK_GBS, "traffic": traffic, "traffic_source": source,
            "bytes_per_launch": bytes_per_launch, "us_per_launch": us, "launches_per_step": launches_per_step,
            "kernel_ms_per_step": kernel_ms_per_step, "kernel_share_of_step": kernel_ms_per_step / run["ms_per_step"],
            "timing": f"HIP events around graph launches of each run of {dominant} launches over the timed window "
                      f"of {run['kt_substeps']} substeps, replayed from the same initial state"}


def kernels_of(run):
    return {k: {"ms_total": round(ms, 4), "launches": n, "us_avg": round(1e3 * ms / n, 2) if n else None}
            for k, (ms, n) in run["kt"].items()}


def check_consistency(roof, ms_per_step, what):
    """The dominant kernel's time per step cannot exceed the step it belongs to.  Enforced for the headline
    workload; the batched and learned objects report the same ratio as kernel_share_of_step."""
    if roof["kernel_ms_per_step"] > CONSISTENCY_SLACK * ms_per_step:
        raise SystemExit(f"bench.py: {what}: {roof['kernel']} takes {roof['kernel_ms_per_step']:.3f} ms per step by "
                         f"its events, more than the timed step of {ms_per_step:.3f} ms; the roofline would not "
                         f"describe the timed ticks")


def batched_object(args):
    """The 64-scene batch (64 x 256^2 grids, 7,046,125 particles, 200 substeps per tick), timed like the
    headline."""
    preset, make_scene, description = WORKLOADS["batch64x256"]
    p = abi.preset(preset)
    scene = make_scene(1.0)
    p.capacity = scene.n
    steps = max(1, min(args.steps, 10))
    run, x = timed_run(p, scene, steps, min(args.warmup, 2), label="batch64x256: ")
    roof = roofline_of(p, scene, run, "batch64x256")
    return {"workload": description, "value": run["value"], "unit": "particle-substeps/s",
            "ms_per_step": run["ms_per_step"], "steps": steps, "warmup": min(args.warmup, 2), "particles": scene.n,
            "scenes": p.num_scenes, "substeps_per_step": p.substeps, "tile": [p.block_x, p.block_y], "chunk": p.chunk,
            "sort_interval": p.sort_interval, "layout": scene_layout("batch64x256", scene, p),
            "state_finite": bool(np.isfinite(x).all()),
            "roofline": {k: roof[k] for k in ("kernel", "achieved", "frac", "traffic", "traffic_source",
                                              "us_per_launch", "launches_per_step", "kernel_share_of_step")},
            "kernels": kernels_of(run)}


def learned_object(args, workload="learned1024", max_steps=10):
    """A learned-closure workload against the analytic engine on the same scene: throughput of both over the
    same ticks, and the relative position and velocity errors of the learned engine against the analytic one
    after one tick (100 substeps) from the initial state.  learned1024 is the 4,194,304-particle jelly + snow
    scene; learned16x2048 is the 16-scene batch of 2048^2 grids (91.5M particles)."""
    preset, make_scene, description = WORKLOADS[workload]
    p = abi.preset(preset)
    scene = make_scene(1.0)
    p.capacity = scene.n
    w = np.load(CLOSURE_WEIGHTS)
    steps = max(1, min(args.steps, max_steps))
    warm = min(args.warmup, 2)
    run, _ = timed_run(p, scene, steps, warm, weights=w, label=f"{workload}: ")
    roof = roofline_of(p, scene, run, workload)
    pa = p.copy()
    pa.closure = abi.CLOSURE_ANALYTIC
    plain, _ = timed_run(pa, scene, steps, warm, kernel_timing=False, label=f"{workload} scene, analytic: ")
    ref = {}
    for name, q, weights in (("learned", p, w), ("plain", pa, None)):
        with Engine(q) as e:
            e.set_scene(scene)
            if weights is not None:
                e.set_closure(weights)
            e.step(1)
            e.sync()
            ref[name] = e.readback(("x", "v"))

    def rel(a, b):
        a = np.asarray(a, np.float64)
        b = np.asarray(b, np.float64)
        return float(np.linalg.norm(a - b) / np.linalg.norm(b))
    ex, ev = rel(ref["learned"]["x"], ref["plain"]["x"]), rel(ref["learned"]["v"], ref["plain"]["v"])
    del ref
    ratio = run["value"] / plain["value"]
    log(f"{workload}: learned / plain {ratio:.3f}; after {p.substeps} substeps relative error vs plain: "
        f"x {ex:.2e}, v {ev:.2e}")
    out = {"workload": description, "value": run["value"], "unit": "particle-substeps/s",
           "ms_per_step": run["ms_per_step"], "steps": steps, "warmup": warm, "particles": scene.n,
           "scenes": p.num_scenes, "plain_value": plain["value"], "plain_ms_per_step": plain["ms_per_step"],
           "learned_over_plain": ratio, "rel_position_error_vs_plain": ex, "rel_velocity_error_vs_plain": ev,
           "substeps_compared": p.substeps, "layout": scene_layout(workload, scene, p),
           "roofline": {k: roof[k] for k in ("kernel", "achieved", "frac", "traffic", "traffic_source",
                                             "us_per_launch", "launches_per_step", "kernel_share_of_step")},
           "mfma_tflops": scene.n * 2 * 64 * 64 / (roof["us_per_launch"] * 1e-6) / 1e12,
           "kernels": kernels_of(run)}
    if p.num_scenes > 1:
        out.update({"tile": [p.block_x, p.block_y], "chunk": p.chunk, "sort_interval": p.sort_interval})
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--mode", choices=["reference", "tiled", "fused"], default="fused")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-extra", action="store_true",
                    help="skip the batched, learned and learned_batched objects")
    ap.add_argument("--scale", type=float, default=1.0, help="particle count scale (1.0 = 4,194,304)")
    ap.add_argument("--autotune", action="store_true", help="pick tile shape, chunk and sort interval by mpm_autotune")
    ap.add_argument("--sort-interval", type=int, default=0, metavar="SUBSTEPS",
                    help="substeps between binnings (0 = the preset's; above the tick's substeps: every k-th tick)")
    ap.add_argument("--chunk", type=int, default=0, metavar="PARTICLES",
                    help="particles per work item (0 = the preset's)")
    ap.add_argument("--tile", default="", metavar="BXxBY",
                    help="bin / tile shape in cells, one of 8x8, 16x8, 16x16, 32x16 (empty = the preset's)")
    ap.add_argument("--workload", choices=sorted(WORKLOADS), default="mixed1024",
                    help="mixed1024 is the headline; the others are the remaining configurations")
    ap.add_argument("--copy-calibration", type=int, default=0, metavar="MIB",
                    help="also run device copies of MIB MiB (profiler counter calibration)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py: the engine runs on one GPU; --gpus {args.gpus} is not supported (sharding across GPUs "
              f"is designed, not implemented)", file=sys.stderr)
        sys.exit(2)
    if args.copy_calibration:
        copy_calibration(args.copy_calibration)

    preset, make_scene, description = WORKLOADS[args.workload]
    p = abi.preset(preset)
    p.mode = {"reference": abi.MODE_REFERENCE, "tiled": abi.MODE_TILED, "fused": abi.MODE_FUSED}[args.mode]
    if args.sort_interval > 0:
        p.sort_interval = args.sort_interval
    if args.chunk > 0:
        p.chunk = args.chunk
    if args.tile:
        p.block_x, p.block_y = (int(v) for v in args.tile.lower().split("x"))
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
    weights = np.load(CLOSURE_WEIGHTS) if p.closure == abi.CLOSURE_LEARNED else None
    run, x = timed_run(p, scene, args.steps, args.warmup, weights=weights)
    roof = roofline_of(p, scene, run, args.workload)
    kernels = kernels_of(run)
    log(f"kernels: {json.dumps(kernels)}")
    log(f"{roof['kernel']}: {roof['us_per_launch']:.2f} us per launch x {roof['launches_per_step']:.1f} per step = "
        f"{roof['kernel_ms_per_step']:.3f} ms of the {run['ms_per_step']:.3f} ms step; frac {roof['frac']:.3f}")
    check_consistency(roof, run["ms_per_step"], args.workload)

    nodes = p.grid_w * p.grid_h * p.num_scenes
    substep_s = run["ms_per_step"] * 1e-3 / p.substeps
    substep_bytes = scene.n * SUBSTEP_BYTES_PER_PARTICLE + nodes * SUBSTEP_BYTES_PER_NODE
    # the binning on its own line, amortised over the substeps that the event-timed cycles covered
    ms_bin, n_bin = run["kt"]["binning"]
    binning = None
    if n_bin:
        us_sort = 1e3 * ms_bin / n_bin
        bin_bytes = scene.n * binning_bytes_per_particle(p)
        binning = {"us_per_sort": us_sort, "bytes_per_sort": bin_bytes,
                   "achieved": bin_bytes / (us_sort * 1e-6) / 1e9, "unit": "GB/s",
                   "frac": bin_bytes / (us_sort * 1e-6) / 1e9 / HBM_PEAK_GBS,
                   "substeps_per_sort": run["kt_substeps"] / n_bin,
                   "us_per_substep_amortised": 1e3 * ms_bin / run["kt_substeps"]}

    config = {"workload": description if args.scale == 1.0 else f"{args.workload} scaled x{args.scale}: {scene.n} particles",
              "particles": scene.n, "scenes": p.num_scenes, "grid_cells": p.grid_w - 4,
              "substeps_per_step": p.substeps, "mode": args.mode, "tile": [p.block_x, p.block_y], "chunk": p.chunk,
              "sort_interval": p.sort_interval}
    config.update(scene_layout(args.workload, scene, p))
    result = {
        "metric": METRIC,
        "value": run["value"],
        "unit": "particle-substeps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": run["ms_per_step"],
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": config,
        "roofline": roof,
        "substep": {"us": substep_s * 1e6, "algorithmic_bytes_unfused": substep_bytes,
                    "effective_GBs_unfused_model": substep_bytes / substep_s / 1e9},
        "binning": binning,
        "kernels": kernels,
        "state_finite": bool(np.isfinite(x).all()),
    }
    if args.workload == "mixed1024" and args.scale == 1.0 and not args.no_extra:
        result["batched"] = batched_object(args)
        result["learned"] = learned_object(args)
        result["learned_batched"] = learned_object(args, "learned16x2048", max_steps=3)
    if not args.no_cpu_baseline:
        log("cpu baseline (FP64 oracle) ...")
        result["cpu_baseline"] = cpu_baseline(p, scene)
    print(json.dumps(result), flush=True)


if __name__ == "__main__":
    main()
