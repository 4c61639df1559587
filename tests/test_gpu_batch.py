"""Batched multi-view rendering (bhRenderBatch): views of different sizes,
models and disk settings must reproduce per-view launchRenderPipeline output
bit for bit, and match the oracle.

The persistent single-queue multi-view kernel hung on its first two GPU runs
(cause not yet shown); bhRenderBatch now runs the views as single-view
pipelines on worker streams forked from the caller's stream (views finish in
any order), and these tests exercise that path.  The persistent kernel's own
launch path returns hipErrorNotSupported.
"""
import numpy as np
import pytest

from helpers import lsb_stats

pytestmark = pytest.mark.gpu


def say(*a):
    print("[gpu-test]", *a, flush=True)


def small_views(bt, n=12):
    """The seeded uneven batch, shrunk to oracle-friendly sizes, plus a
    Schwarzschild and a Reissner-Nordstrom view."""
    views = bt.workload_views("uneven_views", 1, n)
    for k, P in enumerate(views):
        s = [(96, 54), (160, 90), (128, 72), (200, 112)][k % 4]
        P.width, P.height = s
        P.aspect = float(np.float32(s[0]) / np.float32(s[1]))
    P = views[n - 1]
    P.model, P.aSpin, P.Qc = 0, 0.0, 0.0
    d, *_ = bt.validate_and_derive(0, 0.5, 0.0, 0.0)
    P.rPlus, P.rPhoton, P.rErgo, P.diskInner = d.rPlus, d.rPhoton, d.rErgo, 3.0
    P = views[n - 2]
    d, *_ = bt.validate_and_derive(1, 0.5, 0.0, 0.5)
    P.model, P.aSpin, P.Qc = 1, 0.0, d.Q
    P.rPlus, P.rPhoton, P.rErgo = d.rPlus, d.rPhoton, d.rErgo
    P.diskInner = float(bt.load().bhIscoFP64(0.5, 0.0, d.Q))
    return views


SKIP_PERSISTENT = pytest.mark.skip(
    reason="persistent multi-view kernel hung on the GPU twice (cause not shown); path disabled in this build")


@SKIP_PERSISTENT
def test_batch_equals_single_view_launches(gpu, oracle):
    views = small_views(gpu)
    say("batch of", len(views), "small views")
    B = gpu.BatchBuffers(views)
    B.render()
    gpu.synchronize()
    lat = B.latency_ms()
    assert all(x >= 0 for x in lat), lat
    for v, P in enumerate(views):
        rgba, acc = B.readback(v)
        ref_rgba, ref_acc = gpu.render_pipeline(P)
        assert np.array_equal(rgba, ref_rgba), v
        assert np.array_equal(acc.view(np.uint32), ref_acc.view(np.uint32)), v
        o_rgba, _ = oracle.render_pipeline(P)
        st = lsb_stats(rgba, o_rgba)
        assert st["le1"] >= 0.999, (v, st)


@SKIP_PERSISTENT
def test_uneven_64_view_batch(gpu):
    views = gpu.workload_views("uneven_views", 1, 64)
    sizes = [(P.width, P.height) for P in views]
    say("64-view batch, total rays", sum(w * h for w, h in sizes))
    B = gpu.BatchBuffers(views)
    B.render()
    gpu.synchronize()
    lat = B.latency_ms()
    say("per-view trace latency ms: min %.2f median %.2f max %.2f" % (min(lat), float(np.median(lat)), max(lat)))
    assert all(x >= 0 for x in lat)
    # the smallest views of the batch against stand-alone launches
    order = np.argsort([w * h for w, h in sizes])[:4]
    for v in order:
        rgba, acc = B.readback(int(v))
        assert (rgba[..., 3] == 255).all() and np.isfinite(acc).all()
        ref, _ = gpu.render_pipeline(views[int(v)])
        assert np.array_equal(rgba, ref)


# ---------------------------------------------------------------- sequential batch path
def test_sequential_batch_matches_single_views(gpu, oracle):
    views = small_views(gpu)
    say("sequential batch of", len(views), "small views")
    B = gpu.BatchBuffers(views)
    B.render()
    gpu.synchronize()
    lat = B.latency_ms()
    assert all(x >= 0 for x in lat), lat
    for v, P in enumerate(views):
        rgba, acc = B.readback(v)
        ref_rgba, ref_acc = gpu.render_pipeline(P)
        assert np.array_equal(rgba, ref_rgba), v
        assert np.array_equal(acc.view(np.uint32), ref_acc.view(np.uint32)), v
        o_rgba, _ = oracle.render_pipeline(P)
        st = lsb_stats(rgba, o_rgba)
        assert st["le1"] >= 0.999, (v, st)
