"""bench.py pieces that need no GPU: the one-GPU contract (--gpus above 1 exits with status 2 and says why, before
touching a device) and the binning's algorithmic byte count per particle."""
import os
import subprocess
import sys

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)

import bench  # noqa: E402
from mpm_amd import abi  # noqa: E402


def test_more_than_one_gpu_is_refused_with_a_message():
    r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "2"], capture_output=True, text=True,
                       timeout=300)
    assert r.returncode == 2
    assert "one GPU" in r.stderr and r.stdout == ""


def test_binning_bytes_follow_the_radix_pass_count():
    # keys 12 + P passes x 20 + bin ranges 4 + layered order 20 + gather (4 + 2 x 33 in the fused mode, whose
    # steady-state binning leaves v and C behind; 4 + 2 x 57 otherwise)
    for preset, bins, passes in ((abi.PRESET_MIXED1024, 65 * 65, 2), (abi.PRESET_BATCH64X256, 64 * 17 * 33, 2),
                                 (abi.PRESET_LEARNED16X2048, 16 * 129 * 129, 3)):
        p = abi.preset(preset)
        assert p.mode == abi.MODE_FUSED
        assert p.num_scenes * (-(-p.grid_w // p.block_x)) * (-(-p.grid_h // p.block_y)) == bins
        assert bench.binning_bytes_per_particle(p) == 12 + 20 * passes + 4 + 20 + 4 + 66
        p.mode = abi.MODE_TILED
        assert bench.binning_bytes_per_particle(p) == 12 + 20 * passes + 4 + 20 + 4 + 114
