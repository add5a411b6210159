"""Writes tests/golden/jelly128_1000.npz: the FP64 oracle's state of the 128-cell jelly block
(32,768 seeded particles, dx = 1/128, dt = 1e-4, E = 1e3, nu = 0.2) after 1000 substeps, every 32nd
particle.  The reference has no runnable CPU form and no fixtures, so the fixture is produced by the
repository's own oracle (serial loop, deterministic) and pins regressions of the engine.

    python tests/golden/make_golden.py
"""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.join(ROOT, "gpu-physics-engine_amd"))

import oracle  # noqa: E402
from mpm_amd import abi, scenes  # noqa: E402


def main():
    p = abi.preset(abi.PRESET_JELLY128)
    sc = scenes.jelly128()
    s = oracle.State(sc.x.astype(np.float64), sc.v.astype(np.float64), material=sc.material)
    oracle.run(p, s, 1000, threads=1)
    idx = np.arange(0, sc.n, 32, dtype=np.int32)
    np.savez_compressed(os.path.join(HERE, "jelly128_1000.npz"), index=idx, x=s.x[idx], v=s.v[idx], F=s.F[idx],
                        J=s.J[idx])
    print("wrote", len(idx), "particles; mean y", s.x[:, 1].mean())


if __name__ == "__main__":
    main()
