"""Headless frame loop on the GPU: the bhrender tool (viewer layer ->
launchRenderPipeline -> PNG/PFM) against the same frames rendered in this
process through the C ABI, and the written files against the device
buffers, bit for bit."""
import json
import os
import subprocess

import numpy as np
import pytest

import bhdiff

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
CLI = os.path.join(ROOT, "black-hole_amd", "lib", "bhrender")


def say(*a):
    print("[gpu-test]", *a, flush=True)


def run_cli(args, tmp_path):
    if not os.path.exists(CLI):
        pytest.fail("bhrender not built: run __graft_entry__.build()")
    say("bhrender", " ".join(args))
    r = subprocess.run([CLI] + args, capture_output=True, text=True, timeout=120, cwd=tmp_path)
    assert r.returncode == 0, r.stderr
    return [json.loads(l) for l in r.stdout.splitlines() if l.startswith("{")]


@pytest.mark.parametrize("ext", ["png", "ppm"])
def test_cli_matches_in_process_frames(gpu, tmp_path, ext):
    W, H, n, dt = 320, 180, 4, 0.02
    args = ["--size", f"{W}x{H}", "--frames", str(n), "--dt", str(dt), "--model", "kerr", "--spin", "0.9",
            "--ori", "0.3,-0.2,0.1", "--quality", "2", "--out", f"f.{ext}", "--hdr", "f.pfm"]
    lines = run_cli(args, tmp_path)
    assert [l["frame"] for l in lines] == list(range(n))
    assert [l["restart"] for l in lines] == [1] + [0] * (n - 1)
    assert [l["sampleIndex"] for l in lines] == list(range(n))
    assert [l["accumMode"] for l in lines] == [0] + [2] * (n - 1)
    assert all(l["trace_ms"] > 0 for l in lines)
    # the same frames in-process
    v = gpu.Viewer(W, H)
    v.update(model=2, aStar=0.9, oriX=0.3, oriY=-0.2, oriZ=0.1)
    v.key("2")
    Ps = [v.frame(np.float32(dt))[0] for _ in range(n)]
    v.close()
    assert [P.diskTime for P in Ps] == pytest.approx([l["diskTime"] for l in lines], abs=1e-6)
    rgba, accum = gpu.render_sequence(Ps)
    img = bhdiff.read_image(str(tmp_path / f"f.{ext}"))
    want = rgba.copy()
    if ext == "ppm":
        want[..., 3] = 255
    assert np.array_equal(img, want)
    hdr = bhdiff.read_image(str(tmp_path / "f.pfm"))
    assert np.array_equal(hdr.view(np.uint32), np.ascontiguousarray(accum[..., :3]).view(np.uint32))
    say("frame ms", [l["trace_ms"] for l in lines])


def test_cli_workload_frame(gpu, tmp_path):
    lines = run_cli(["--workload", "schw_256_balanced", "--out", "w.png"], tmp_path)
    assert len(lines) == 1 and lines[0]["accumMode"] == 0
    P = gpu.workload("schw_256_balanced", 0)
    rgba, _ = gpu.render_pipeline(P)
    assert np.array_equal(bhdiff.read_image(str(tmp_path / "w.png")), rgba)


def test_viewer_progressive_frames_converge(gpu):
    """A still view (paused disk) accumulates the running mean of jittered
    samples: the HDR buffer after k frames equals the mean of the k
    single-sample frames rendered separately."""
    W, H, k = 160, 96, 4
    v = gpu.Viewer(W, H)
    v.key(" ")
    Ps = [v.frame(0.016)[0] for _ in range(k)]
    v.close()
    assert [P.accumMode for P in Ps] == [0] + [1] * (k - 1)
    _, acc = gpu.render_sequence(Ps)
    # the jittered sample of frame n alone: blend into a zero buffer, which
    # leaves sample / (n + 1)
    singles = []
    for P in Ps:
        _, a = gpu.render_sequence([P])
        singles.append(a.astype(np.float64) * (P.sampleIndex + 1))
    mean = np.mean(singles, axis=0)
    err = np.abs(acc[..., :3] - mean[..., :3]).max() / max(1e-6, np.abs(mean[..., :3]).max())
    say("running mean vs mean of singles: max rel err %.2e" % err)
    assert err < 1e-5
