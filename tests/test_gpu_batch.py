"""Batched multi-view rendering (bhRenderBatch): views of different sizes,
models and disk settings must reproduce per-view launchRenderPipeline output
bit for bit, and match the oracle.

bhRenderBatch runs the views as single-view pipelines on worker streams
forked from the caller's stream (views finish in any order).
bhRenderBatchPersistent traces all views in one persistent launch over a
device queue of (view, tile) items; with the FP64 band pass on it first
classifies every view and traces the band rays of all views in one launch
over a queue of (view, 64-ray chunk) items beside it.  It is compared with
launchRenderPipeline under the same precision, band on and band off.
"""
import ctypes as C

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


HIP_ERROR_INVALID_VALUE = 1


class _Band:
    """The FP64 band pass on or off (default widths) for the block."""

    def __init__(self, gpu, on):
        self.gpu, self.on = gpu, on

    def __enter__(self):
        self.p0 = self.gpu.get_precision()
        self.gpu.set_precision(band=self.on)

    def __exit__(self, *exc):
        self.gpu.restore_precision(self.p0)


@pytest.mark.parametrize("band", [0, 1])
def test_persistent_batch_equals_single_view_launches(gpu, oracle, band):
    views = small_views(gpu)
    say("persistent batch of", len(views), "small views, FP64 band", band)
    B = gpu.BatchBuffers(views)
    with _Band(gpu, band):
        assert B.render_persistent() == 0
        gpu.synchronize()
        lat = B.latency_ms()
        say("persistent batch done; per-view latency ms", ["%.3f" % x for x in lat])
        assert all(x > 0 for x in lat), lat
        for v, P in enumerate(views):
            rgba, acc = B.readback(v)
            ref_rgba, ref_acc = gpu.render_pipeline(P)
            assert np.array_equal(rgba, ref_rgba), v
            assert np.array_equal(acc.view(np.uint32), ref_acc.view(np.uint32)), v
    p = gpu.get_precision()
    oracle.set_band(band, p.critBand, p.polarBand)
    try:
        for v, P in enumerate(views):
            o_rgba, _ = oracle.render_pipeline(P)
            st = lsb_stats(B.readback(v)[0], o_rgba)
            assert st["le1"] >= 0.999, (v, st)
    finally:
        p = gpu.get_precision()
        oracle.set_band(p.fp64Band, p.critBand, p.polarBand)


@pytest.mark.parametrize("band", [0, 1])
def test_persistent_uneven_64_view_batch(gpu, band):
    """The seed-1 64-view batch in one persistent launch: every view equals
    its own launchRenderPipeline under the same precision bit for bit, again
    on a second launch into the same buffers."""
    views = gpu.workload_views("uneven_views", 1, 64)
    B = gpu.BatchBuffers(views)
    with _Band(gpu, band):
        for rep in range(2):
            assert B.render_persistent() == 0
            gpu.synchronize()
            lat = B.latency_ms()
            say("persistent 64-view batch %d: view latency ms min %.2f median %.2f max %.2f" %
                (rep, min(lat), float(np.median(lat)), max(lat)))
            assert all(x > 0 for x in lat)
        for v, P in enumerate(views):
            rgba, acc = B.readback(v)
            ref, ref_acc = gpu.render_pipeline(P)
            assert np.array_equal(rgba, ref), v
            assert np.array_equal(acc.view(np.uint32), ref_acc.view(np.uint32)), v


@pytest.mark.parametrize("band", [1, 0])
def test_persistent_batch_graph_replays(gpu, band):
    """The 64-view persistent launch captured as a graph (autotuned trace
    workgroup): three replays, with host synchronizations and unrelated eager
    frames between them, each give the eager launch's frames bit for bit.
    With the FP64 band pass on (the default), whose classification and band
    trace the graph holds beside the FP32 trace, and with it off."""
    views = gpu.workload_views("uneven_views", 1, 64)
    B = gpu.BatchBuffers(views)
    with _Band(gpu, band):
        assert B.render_persistent() == 0
        gpu.synchronize()
        ref = [B.readback(v) for v in range(len(views))]
        G = B.graph(autotune=1)
        wg, tuned = G.info()
        say("batch graph: workgroup %d, tuning ms %s" % (wg, ["%.2f" % t for t in tuned]))
        assert wg in (64, 128, 256) and all(t > 0 for t in tuned)
        try:
            for rep in range(3):
                for b in B.bufs:
                    b[0].fill(0)
                G.launch()
                gpu.synchronize()
                lat = G.latency_ms()
                assert all(x > 0 for x in lat), lat
                for v in range(len(views)):
                    rgba, acc = B.readback(v)
                    assert np.array_equal(rgba, ref[v][0]), (rep, v)
                    assert np.array_equal(acc.view(np.uint32), ref[v][1].view(np.uint32)), (rep, v)
                say("replay %d identical; trace latency max %.2f ms" % (rep, max(lat)))
                gpu.render_pipeline(views[rep])      # unrelated eager work between replays
        finally:
            G.close()


def test_uneven_64_view_batch(gpu, oracle):
    """The seed-1 batch of 64 uneven views (320x180 .. 1920x1080, Kerr-Newman,
    disk on/off) through bhRenderBatch: every view equals its own
    launchRenderPipeline bit for bit, a repeated batch (work ordered by each
    view's previous records) gives the same frames, and the smallest views
    match the oracle to 1 LSB on >= 99.9% of pixels."""
    views = gpu.workload_views("uneven_views", 1, 64)
    sizes = [(P.width, P.height) for P in views]
    say("64-view batch, total rays", sum(w * h for w, h in sizes))
    B = gpu.BatchBuffers(views)
    B.render()
    gpu.synchronize()
    lat = B.latency_ms()
    say("per-view latency ms: min %.2f median %.2f max %.2f" % (min(lat), float(np.median(lat)), max(lat)))
    assert all(x >= 0 for x in lat)
    first = [B.readback(v) for v in range(len(views))]
    for v, P in enumerate(views):
        rgba, acc = first[v]
        assert (rgba[..., 3] == 255).all() and np.isfinite(acc).all(), v
        ref, ref_acc = gpu.render_pipeline(P)
        assert np.array_equal(rgba, ref), v
        assert np.array_equal(acc.view(np.uint32), ref_acc.view(np.uint32)), v
    say("64 views equal their single-view launches")
    B.render()
    gpu.synchronize()
    for v in range(len(views)):
        rgba, acc = B.readback(v)
        assert np.array_equal(rgba, first[v][0]), v
        assert np.array_equal(acc.view(np.uint32), first[v][1].view(np.uint32)), v
    say("repeated batch identical")
    for v in np.argsort([w * h for w, h in sizes])[:3]:
        o_rgba, _ = oracle.render_pipeline(views[int(v)])
        st = lsb_stats(first[int(v)][0], o_rgba)
        say("view", int(v), sizes[int(v)], st)
        assert st["le1"] >= 0.999, (int(v), st)


def test_batch_rejects_bad_view_before_enqueueing(gpu):
    views = small_views(gpu, 4)
    B = gpu.BatchBuffers(views)
    B.outs[2] = None
    assert gpu.load().bhRenderBatch(B.n, B.views, B.outs, B.accs, B.bAs, B.bBs, None) == HIP_ERROR_INVALID_VALUE
    views[1].bloomEnabled = 1
    B2 = gpu.BatchBuffers(views)
    B2.bBs[1] = None
    assert gpu.load().bhRenderBatch(B2.n, B2.views, B2.outs, B2.accs, B2.bAs, B2.bBs, None) == HIP_ERROR_INVALID_VALUE


def test_stream_release_frees_batch_state(gpu):
    """A batch on a private stream, released, then again on the same stream:
    the library recreates its workers and the frames are unchanged."""
    views = small_views(gpu, 6)
    st = C.c_void_p()
    gpu.check(gpu.load().bhStreamCreate(C.byref(st)), "bhStreamCreate")
    try:
        B = gpu.BatchBuffers(views)
        B.render(st)
        gpu.check(gpu.load().bhStreamSynchronize(st), "sync")
        first = [B.readback(v) for v in range(len(views))]
        gpu.check(gpu.load().bhStreamRelease(st), "bhStreamRelease")
        B.render(st)
        gpu.check(gpu.load().bhStreamSynchronize(st), "sync")
        for v in range(len(views)):
            assert np.array_equal(B.readback(v)[0], first[v][0]), v
    finally:
        gpu.check(gpu.load().bhStreamDestroy(st), "bhStreamDestroy")


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
