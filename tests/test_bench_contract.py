"""bench.py pieces that run without a GPU: the measured-traffic lookup
behind `roofline.traffic`, and the single-GPU exit."""
import os
import subprocess
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)

import bench  # noqa: E402


@pytest.mark.parametrize("prec,w", [("fp32", 4), ("fp64", 8)])
def test_roofline_traffic_comes_from_the_committed_counters(prec, w):
    """The dominant kernel's HBM bytes per launch are found in
    profiles/hbm_traffic.json, for the link form the 96^3 workload resolves
    to (FP32: packed, its 67 MB of fields in the full form being above the
    32 MiB threshold), and are the algorithmic bytes give or take the
    measured excess (a few percent to 15%)."""
    from su2higgs_amd import abi
    p = bench.workload_params(abi.FP32 if prec == "fp32" else abi.FP64)
    packed = prec == "fp32"
    if packed:
        assert abi.volume(p) * 19 * 4 > 32 << 20
        p.link_format = abi.LINKS_PACKED
    t = bench.load_traffic("link_overrelax", "96cubed", prec, packed)
    assert t is not None
    alg = bench.algo_words(abi.K_LINK_OVERRELAX, p) * abi.volume(p) * w
    assert 0.95 < t / alg < 1.3, (t, alg)


def test_more_than_one_gpu_exits_with_a_message():
    r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "2"],
                       capture_output=True, text=True, timeout=120)
    assert r.returncode == 2
    assert "one GPU" in r.stderr
    assert r.stdout == ""
