"""Progressive accumulation (device sample index), hipGraph replay and the
autotuner.

Reference behaviour: the application's progressive anti-aliasing loop
(main.cpp:439-478) renders frame k of a steady view with sampleIndex = k
(R2 jitter, post_process.cuh:47-53) and blends it into accum with
accumBlend (running mean in mode 1, EMA in mode 2; post_process.cuh:56-68);
test_main.cu:596-616 (T6) checks accumulation determinism.

* bhAccumStep frames equal explicit launchRenderPipeline calls with
  sampleIndex = 0, 1, 2, ... bit for bit (accum and uchar4);
* a captured step replayed k times equals the eager sequence after every
  replay (host synchronizations and eager work between the replays), with
  the FP64 band pass on and off;
* the accumulated frames match the oracle's accumulation chain to 1 LSB;
* the autotuner returns a valid schedule, leaves the output bit-identical
  and answers a repeated query from its cache.
"""
import numpy as np
import pytest

from helpers import lsb_stats, scaled_workload

pytestmark = pytest.mark.gpu


def say(*a):
    print("[gpu-test]", *a, flush=True)


def explicit_sequence(gpu, P, n):
    """Frames 0..n-1 through launchRenderPipeline with explicit sample indices;
    returns [(rgba, accum)] after each frame."""
    import ctypes as C
    lib = gpu.load()
    W, H = P.width, P.height
    bw, bh = (W + 1) // 2, (H + 1) // 2
    out, acc = gpu.DeviceBuffer(W * H * 4), gpu.DeviceBuffer(W * H * 16)
    bA, bB = gpu.DeviceBuffer(bw * bh * 16), gpu.DeviceBuffer(bw * bh * 16)
    acc.fill(0)
    res = []
    for k in range(n):
        Q = P.copy()
        Q.sampleIndex = k
        gpu.check(lib.launchRenderPipeline(out.ptr, acc.ptr, bA.ptr, bB.ptr, C.byref(Q), None), "pipeline")
        gpu.synchronize()
        res.append((out.download(np.uint8).reshape(H, W, 4), acc.download(np.float32).reshape(H, W, 4)))
    return res


def same(a, b):
    return np.array_equal(a[0], b[0]) and np.array_equal(a[1].view(np.uint32), b[1].view(np.uint32))


@pytest.fixture(params=[0, 1], ids=["fp32", "fp64band"])
def band(gpu, request):
    old = gpu.get_precision()
    gpu.set_precision(band=request.param)
    yield request.param
    gpu.restore_precision(old)


@pytest.mark.parametrize("mode", [1, 2])
def test_accumulator_steps_and_graph_replays(gpu, band, mode):
    P = scaled_workload(gpu, "kerr09_1080p_high", 1, 320, 180)
    P.accumMode = mode
    n = 5
    ref = explicit_sequence(gpu, P, n)
    say("accumMode", mode, "band", band, "frames", n)
    A = gpu.Accumulator(P.width, P.height)
    try:
        A.reset()
        for k in range(n):
            A.step(P)
            assert same(A.readback(), ref[k]), ("eager step", k)
        assert A.sample_index() == n
        # graph: capture once, replay frame by frame with eager work between
        A.reset()
        A.capture(P)
        assert A.sample_index() == 0          # capture executes nothing
        for k in range(n):
            A.replay(1)
            got = A.readback()
            assert same(got, ref[k]), ("graph replay", k)
            gpu.render_pipeline(P)            # unrelated eager work between replays
        assert A.sample_index() == n
    finally:
        A.close()


def test_accumulation_vs_oracle(gpu, oracle):
    P = scaled_workload(gpu, "kerr09_1080p_high", 2, 256, 144)
    P.accumMode = 1
    ref = explicit_sequence(gpu, P, 3)
    acc = None
    for k in range(3):
        Q = P.copy()
        Q.sampleIndex = k
        o_rgba, acc = oracle.render_pipeline(Q, acc)
        st = lsb_stats(ref[k][0], o_rgba)
        say("frame", k, st)
        assert st["le1"] >= 0.999


def test_autotune(gpu):
    """Tuning is per workload class: the tuned schedule applies to frames of
    the tuned view's class only, the process default and other classes keep
    theirs, and no output changes."""
    P = scaled_workload(gpu, "kerr09_1080p_high", 0, 320, 180)
    B = scaled_workload(gpu, "kerr09_1080p_high", 0, 256, 144)      # another class (size)
    before = gpu.get_schedule()
    base, _ = gpu.render_pipeline(P)
    base_b, _ = gpu.render_pipeline(B)
    try:
        gpu.load().bhAutotuneClear()
        assert gpu.schedule_for(P).as_dict() == before.as_dict()
        s, ms = gpu.autotune(P, 1)
        say("autotune chose", s.as_dict(), "%.3f ms" % ms)
        assert s.wgSize in (64, 128, 256) and 1 <= s.raysPerRefill <= 64 and ms > 0
        assert gpu.schedule_for(P).as_dict() == s.as_dict()
        assert gpu.get_schedule().as_dict() == before.as_dict()
        assert gpu.schedule_for(B).as_dict() == before.as_dict()
        Q = P.copy()
        Q.sampleIndex, Q.diskTime = 3, 0.25                            # same class
        assert gpu.schedule_for(Q).as_dict() == s.as_dict()
        tuned, _ = gpu.render_pipeline(P)
        assert np.array_equal(tuned, base)
        tuned_b, _ = gpu.render_pipeline(B)
        assert np.array_equal(tuned_b, base_b)
        # a forced different schedule for the class gives the same frame too
        gpu.set_schedule(wg=256 if s.wgSize != 256 else 128)
        assert gpu.schedule_for(P).as_dict() == s.as_dict()
        s2, ms2 = gpu.autotune(P, 1)          # cached
        assert s2.as_dict() == s.as_dict() and ms2 == ms
    finally:
        gpu.load().bhAutotuneClear()
        gpu.restore_schedule(before)
