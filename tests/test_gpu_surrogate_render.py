"""The learned surrogate in the render path, and its ablation.

A network is trained by the library's own HIP kernels on rays traced by the
plain FP32 kernel, attached with bhSetSurrogate, and compared with the same
frame rendered without it (Kerr a*=0.9 workload at reduced size):

* untrained network (ablation baseline): the router accepts nothing and the
  frame equals the plain frame bit for bit -- the router never substitutes
  an unconfident prediction;
* trained network: accepted rays are ones the plain trace shows without
  disk light (the disk-hit head gates them; at most 0.2% may slip), their
  exit direction is reported
  against the plain FP32 trace and against the FP64 oracle, and the number
  of rays the network takes over the plain method is reported with the
  frame time of both paths.
"""
import ctypes as C

import numpy as np
import pytest

from helpers import scaled_workload

pytestmark = pytest.mark.gpu

W, H = 480, 270


def say(*a):
    print("[gpu-test]", *a, flush=True)


@pytest.fixture(scope="module")
def trained(gpu):
    s = gpu.Surrogate(seed=2)
    s.dataset(1024, 4096, seed=2)
    loss = s.train(3000, batch=65536, lr=2e-3)
    say("trained 3000 steps, last loss", loss[-50:].mean(0))
    yield s
    gpu.set_surrogate(None)
    s.close()


def frame_ms(gpu, P, n=3):
    lib = gpu.load()
    fb = gpu.Framebuffer(P.width, P.height)
    fb.render(P)
    gpu.synchronize()
    lib.bhEnableKernelTimers(1)
    tr, ts = C.c_float(), []
    for _ in range(n):
        fb.render(P)
        lib.bhLastKernelTimes(C.byref(tr), None, None)
        ts.append(tr.value)
    lib.bhEnableKernelTimers(0)
    fb.close()
    return min(ts)


def accepted_mask(rec):
    return ((rec["flags"] >> 8) == 0) & ((rec["flags"] & 1) == 1)


def test_camera_in_disk_layer_skips_the_network(gpu):
    """The headline camera sits inside the disk's emitting layer (|dz| = 1.7
    scale heights at r = 15): every ray starts in the disk, so the network is
    not consulted at all and the frame is the plain frame."""
    P = scaled_workload(gpu, "kerr09_1080p_high", 0, W, H)
    base, _ = gpu.render_pipeline(P)
    s = gpu.Surrogate(seed=5)
    try:
        gpu.set_surrogate(s, 1.0, 1.0)
        img, _ = gpu.render_pipeline(P)
        cand, acc = gpu.last_surrogate_stats()
    finally:
        gpu.set_surrogate(None)
        s.close()
    say("camera in the disk layer: candidates", cand, "accepted", acc)
    assert cand == 0 and acc == 0
    assert np.array_equal(img, base)


def test_untrained_network_changes_nothing(gpu):
    P = scaled_workload(gpu, "kerr09_1080p_high", 0, W, H)
    P.diskEnabled = 0
    base, _ = gpu.render_pipeline(P)
    s = gpu.Surrogate(seed=5)
    try:
        gpu.set_surrogate(s, 1e-3, 1e-3)
        img, _ = gpu.render_pipeline(P)
        cand, acc = gpu.last_surrogate_stats()
    finally:
        gpu.set_surrogate(None)
        s.close()
    say("untrained: candidates", cand, "accepted", acc)
    assert cand > 0 and acc == 0
    assert np.array_equal(img, base)


@pytest.mark.parametrize("sigma", [1e-3, 3e-3])
def test_trained_network_router_contract(gpu, oracle, trained, sigma):
    P = scaled_workload(gpu, "kerr09_1080p_high", 0, W, H)
    plain = gpu.trace_records(P)
    t_plain = frame_ms(gpu, P)
    gpu.set_surrogate(trained, sigma, 1e-3)
    try:
        rec = gpu.trace_records(P)
        cand, acc = gpu.last_surrogate_stats()
        t_sur = frame_ms(gpu, P)
    finally:
        gpu.set_surrogate(None)
    m = accepted_mask(rec)
    say("sigma %.0e: candidates %d accepted %d (%.1f%% of pixels); trace %.3f ms vs plain %.3f ms" %
        (sigma, cand, acc, 100.0 * acc / (W * H), t_sur, t_plain))
    assert int(m.sum()) == acc
    # accepted rays never lose disk light
    leak = m & ((plain["trans"] < 1.0) | (np.abs(plain["accum"]).sum(-1) > 0))
    assert int(leak.sum()) <= max(2, acc // 500), int(leak.sum())
    if acc:
        a = rec["escDir"][m].astype(np.float64)
        b = plain["escDir"][m].astype(np.float64)
        a /= np.linalg.norm(a, axis=1, keepdims=True)
        b /= np.linalg.norm(b, axis=1, keepdims=True)
        ang = 2 * np.arcsin(np.clip(np.linalg.norm(a - b, axis=1) / 2, 0, 1))
        # FP64 oracle on a sample of the accepted rays
        idx = np.flatnonzero(m.ravel())[:: max(1, int(m.sum()) // 500)]
        ys, xs = np.divmod(idx, W)
        F = np.array([P.camForward.x, P.camForward.y, P.camForward.z])
        R = np.array([P.camRight.x, P.camRight.y, P.camRight.z])
        U = np.array([P.camUp.x, P.camUp.y, P.camUp.z])
        fx = (xs + 0.5) / W
        fy = (ys + 0.5) / H
        d = F + R * ((fx * 2 - 1) * P.tanHalfFov * P.aspect)[:, None] + U * ((1 - fy * 2) * P.tanHalfFov)[:, None]
        d /= np.linalg.norm(d, axis=1, keepdims=True)
        o = oracle.probe_dirs(P, d.astype(np.float32), mode=1)
        e64 = o[:, 1:4] / np.linalg.norm(o[:, 1:4], axis=1, keepdims=True)
        g = rec["escDir"].reshape(-1, 3)[idx].astype(np.float64)
        g /= np.linalg.norm(g, axis=1, keepdims=True)
        a64 = 2 * np.arcsin(np.clip(np.linalg.norm(g - e64, axis=1) / 2, 0, 1))
        say("accepted exit error vs plain FP32: max %.2e p99.9 %.2e p50 %.2e; vs FP64 (%d sampled): max %.2e" %
            (ang.max(), np.quantile(ang, 0.999), np.median(ang), len(idx), a64.max()))
        assert np.isfinite(ang).all() and ang.max() < 0.05
