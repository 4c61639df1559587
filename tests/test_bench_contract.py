"""bench.py's command-line contract, checked on the CPU (no GPU needed).

* `--gpus N` with N > 1 exits with status 2 and a message, before any GPU work;
* the per-frame flop tables cover the three timed disk times and the band
  pixels' flops are a small part of the frame's, so the roofline numerator
  (frame flops minus band flops) stays positive.
"""
import os
import subprocess
import sys

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_multi_gpu_request_exits_with_message():
    r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "2"],
                       capture_output=True, text=True, timeout=120, cwd=ROOT)
    assert r.returncode == 2
    assert "one GPU" in r.stderr
    assert r.stdout.strip() == ""


def test_flop_tables_consistent():
    sys.path.insert(0, ROOT)
    try:
        import bench
    finally:
        sys.path.pop(0)
    assert sorted(bench.FLOPS) == [0, 1, 2]
    assert sorted(bench.BAND_FLOPS) == [0, 1, 2]
    for k in bench.FLOPS:
        assert 0 < bench.BAND_FLOPS[k] < 0.05 * bench.FLOPS[k]
    assert bench.TRAFFIC_BYTES > 0
