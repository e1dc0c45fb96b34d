"""bench.py's CPU-side contract: the one-GPU rule, the counter table behind
roofline.traffic, and the flop count behind roofline.achieved."""
import os
import subprocess
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)
import bench  # noqa: E402


def test_more_than_one_gpu_exits_with_a_message():
    r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "2"],
                       capture_output=True, text=True, timeout=120)
    assert r.returncode == 2
    assert "one GPU" in r.stderr and r.stdout == ""


def test_traffic_table_covers_every_one_lane_shape():
    shapes = {(1, w, rc) for w in (1, 2, 4) for rc in (1, 2, 3, 4, 6, 8, 12, 16)}
    assert set(bench._SWEEP_PMC_KIB) == shapes
    for (lanes, w, rc), (fetch, write) in bench._SWEEP_PMC_KIB.items():
        # the launch writes one partial chi per omega and slot: 65536 omegas x
        # 2 species x 6 components x 16 B = 12288 KiB per slot, W RC slots
        assert write == pytest.approx(12288.0 * w * rc, rel=1e-9)
        # it reads at least the omegas (1 MiB), i.e. FETCH_SIZE >= 512 KiB at scale 2
        assert fetch * bench.FETCH_SCALE >= 1024.0
        assert bench.SWEEP_TRAFFIC_BYTES[(lanes, w, rc)] == pytest.approx(
            (fetch * bench.FETCH_SCALE + write * bench.WRITE_SCALE) * 1024.0)


def test_flops_per_eval_counts_the_sweep_terms():
    # 54 flops per (species, harmonic n in -nmax..nmax, interior-plus-end ppar node)
    assert bench.flops_per_eval([47, 20], 256) == 54.0 * 255 * (95 + 41)
    assert bench.flops_per_eval([0], 4) == 54.0 * 3
