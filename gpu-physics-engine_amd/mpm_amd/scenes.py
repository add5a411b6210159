"""Seeded synthetic scenes for every preset.

Initial states are generated once on the host; the engine and the oracle consume the same float32
arrays, so no random-number parity between them is needed.  The generator is counter based:
SplitMix64 (Steele, Lea, Flood 2014, "Fast splittable pseudorandom number generators") applied to
seed + (counter + 1) * 0x9E3779B97F4A7C15, counter = 8 * index + stream.  Uniform floats take the top
24 bits, so they are exact float32 values in [0, 1).
"""
import numpy as np

from . import abi

_GOLDEN = np.uint64(0x9E3779B97F4A7C15)
_M1 = np.uint64(0xBF58476D1CE4E5B9)
_M2 = np.uint64(0x94D049BB133111EB)

SEED_BASE = 0xC0FFEE


def splitmix64(seed, counters):
    """SplitMix64 output for each counter (uint64 array)."""
    with np.errstate(over="ignore"):
        z = np.uint64(seed) + (counters.astype(np.uint64) + np.uint64(1)) * _GOLDEN
        z = (z ^ (z >> np.uint64(30))) * _M1
        z = (z ^ (z >> np.uint64(27))) * _M2
        return z ^ (z >> np.uint64(31))


def uniform(seed, stream, n, offset=0):
    """n float64 values (exact float32) uniform in [0, 1) from one stream."""
    idx = np.arange(offset, offset + n, dtype=np.uint64)
    z = splitmix64(seed, idx * np.uint64(8) + np.uint64(stream))
    return (z >> np.uint64(40)).astype(np.float64) * (1.0 / (1 << 24))


class Scene:
    """Initial particle state: x, v float32[n,2], material uint8[n], scene_counts int32[num_scenes]."""

    def __init__(self, x, v, material, scene_counts, seed, description):
        self.x = np.ascontiguousarray(x, dtype=np.float32)
        self.v = np.ascontiguousarray(v, dtype=np.float32)
        self.material = np.ascontiguousarray(material, dtype=np.uint8)
        self.scene_counts = np.ascontiguousarray(scene_counts, dtype=np.int32)
        self.seed = seed
        self.description = description

    @property
    def n(self):
        return int(self.x.shape[0])

    def scene_slices(self):
        off = np.concatenate([[0], np.cumsum(self.scene_counts)])
        return [slice(int(off[k]), int(off[k + 1])) for k in range(len(self.scene_counts))]


def _box(seed, stream, n, lo, hi, offset=0):
    ux = uniform(seed, stream, n, offset)
    uy = uniform(seed, stream + 1, n, offset)
    x = np.empty((n, 2), dtype=np.float64)
    x[:, 0] = lo[0] + ux * (hi[0] - lo[0])
    x[:, 1] = lo[1] + uy * (hi[1] - lo[1])
    return x.astype(np.float32)


def jelly128(n=32768, seed=SEED_BASE + abi.PRESET_JELLY128):
    """Jelly block: n particles uniform in [0.25,0.55]x[0.45,0.75], v = 0."""
    x = _box(seed, 0, n, (0.25, 0.45), (0.55, 0.75))
    v = np.zeros((n, 2), dtype=np.float32)
    mat = np.full(n, abi.MAT_JELLY, dtype=np.uint8)
    return Scene(x, v, mat, [n], seed, f"jelly block, {n} particles in [0.25,0.55]x[0.45,0.75]")


# Three non-overlapping boxes of side 0.408 in the unit square (8 particles per cell at 4,194,304
# particles on a 1024-cell grid); materials fluid, jelly, snow.
MIXED_BOXES = (
    ((0.050, 0.050), (0.458, 0.458)),
    ((0.542, 0.050), (0.950, 0.458)),
    ((0.296, 0.520), (0.704, 0.928)),
)


def mixed(n=4194304, materials=(abi.MAT_FLUID, abi.MAT_JELLY, abi.MAT_SNOW), seed=SEED_BASE + abi.PRESET_MIXED1024,
          boxes=MIXED_BOXES):
    """n particles split evenly over three boxes, initial velocity uniform in [-1,1]^2.  For n below
    4,194,304 every box shrinks about its centre by sqrt(n / 4194304), keeping 8 particles per cell."""
    if n < 4194304:
        f = np.sqrt(n / 4194304.0)
        boxes = tuple((((lo[0] + hi[0]) / 2 - f * (hi[0] - lo[0]) / 2, (lo[1] + hi[1]) / 2 - f * (hi[1] - lo[1]) / 2),
                       ((lo[0] + hi[0]) / 2 + f * (hi[0] - lo[0]) / 2, (lo[1] + hi[1]) / 2 + f * (hi[1] - lo[1]) / 2))
                      for lo, hi in boxes)
    parts_x, parts_m = [], []
    counts = [n // 3, n // 3, n - 2 * (n // 3)]
    off = 0
    for k, ((lo, hi), cnt) in enumerate(zip(boxes, counts)):
        parts_x.append(_box(seed, 0, cnt, lo, hi, offset=off))
        parts_m.append(np.full(cnt, materials[k], dtype=np.uint8))
        off += cnt
    x = np.concatenate(parts_x)
    v = np.empty((n, 2), dtype=np.float64)
    v[:, 0] = uniform(seed, 2, n) * 2.0 - 1.0
    v[:, 1] = uniform(seed, 3, n) * 2.0 - 1.0
    return Scene(x, v.astype(np.float32), np.concatenate(parts_m), [n], seed,
                 f"three boxes ({', '.join(str(m) for m in materials)}), {n} particles, v in [-1,1]^2")


def log_uniform_counts(seed, num_scenes, lo, hi):
    u = uniform(seed, 4, num_scenes)
    return np.floor(np.exp(np.log(lo) + u * (np.log(hi) - np.log(lo)))).astype(np.int64).clip(lo, hi)


def batched(num_scenes=64, lo=4096, hi=524288, cells=256, materials=(abi.MAT_FLUID, abi.MAT_JELLY, abi.MAT_SNOW),
            ppc=16.0, seed=SEED_BASE + abi.PRESET_BATCH64X256, scale=1.0):
    """Independent scenes with log-uniform particle counts; one box per scene sized for ppc particles per
    cell (side capped at 0.8), material drawn uniformly per scene, v uniform in [-1,1]^2.
    scale < 1 shrinks every count (for CPU-sized parity runs)."""
    counts = log_uniform_counts(seed, num_scenes, lo, hi)
    counts = np.maximum((counts * scale).astype(np.int64), 16)
    mats = np.floor(uniform(seed, 5, num_scenes) * len(materials)).astype(np.int64)
    xs, ms = [], []
    off = 0
    for k in range(num_scenes):
        cnt = int(counts[k])
        side = min(0.8, np.sqrt(cnt / ppc) / cells)
        lo_xy = (0.5 - 0.5 * side, 0.1)
        hi_xy = (0.5 + 0.5 * side, 0.1 + side)
        xs.append(_box(seed, 0, cnt, lo_xy, hi_xy, offset=off))
        ms.append(np.full(cnt, materials[int(mats[k])], dtype=np.uint8))
        off += cnt
    n = off
    v = np.empty((n, 2), dtype=np.float64)
    v[:, 0] = uniform(seed, 2, n) * 2.0 - 1.0
    v[:, 1] = uniform(seed, 3, n) * 2.0 - 1.0
    return Scene(np.concatenate(xs), v.astype(np.float32), np.concatenate(ms), counts.astype(np.int32), seed,
                 f"{num_scenes} scenes of {cells} cells, counts log-uniform in [{lo},{hi}] x{scale}")


def for_preset(which, scale=1.0):
    """Initial scene of a named preset (scale < 1 shrinks particle counts for small runs)."""
    if which == abi.PRESET_JELLY128:
        return jelly128(n=max(16, int(32768 * scale)))
    if which == abi.PRESET_MIXED1024:
        return mixed(n=max(48, int(4194304 * scale)))
    if which == abi.PRESET_BATCH64X256:
        return batched(scale=scale)
    if which == abi.PRESET_LEARNED1024:
        return mixed(n=max(48, int(4194304 * scale)), materials=(abi.MAT_JELLY, abi.MAT_JELLY, abi.MAT_SNOW),
                     seed=SEED_BASE + abi.PRESET_LEARNED1024)
    if which == abi.PRESET_LEARNED16X2048:
        return batched(num_scenes=16, lo=1048576, hi=16777216, cells=2048, materials=(abi.MAT_JELLY, abi.MAT_SNOW),
                       seed=SEED_BASE + abi.PRESET_LEARNED16X2048, scale=scale)
    raise ValueError(f"no scene generator for preset {which}")
