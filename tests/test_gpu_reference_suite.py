"""The reference's verification battery (tests/test_main.cu T3-T6) run on the
HIP path, plus schedule invariance and probe parity against the oracle.

T1/T2 (host-side derivation and validation) live in test_host.py.
"""
import ctypes as C

import numpy as np
import pytest

from helpers import compare_grids, lsb_stats, make_params, scaled_workload

pytestmark = pytest.mark.gpu

NX, NY = 96, 54
SCHW, RN, KERR, KN = 0, 1, 2, 3


def say(*a):
    print("[gpu-test]", *a, flush=True)


@pytest.fixture(autouse=True)
def exact_schedule(gpu):
    s0 = gpu.get_schedule()
    gpu.set_schedule(math=0)
    yield
    gpu.restore_schedule(s0)


def grid(gpu, P, nx=NX, ny=NY):
    g = gpu.probe_grid(P, nx, ny)
    return g["outcome"], g["escDir"], g


# ----------------------------------------------------------------- T3 (test_main.cu:340-443)
def test_T3_model_consistency(gpu):
    say("T3 grids 96x54")
    S = grid(gpu, make_params(gpu, SCHW, 0.0, 0.0))
    RN0 = grid(gpu, make_params(gpu, RN, 0.0, 0.0))
    RN5 = grid(gpu, make_params(gpu, RN, 0.0, 0.5))
    K0 = grid(gpu, make_params(gpu, KERR, 0.0, 0.0))
    K6 = grid(gpu, make_params(gpu, KERR, 0.6, 0.0))
    KN60 = grid(gpu, make_params(gpu, KN, 0.6, 0.0))
    KN05 = grid(gpu, make_params(gpu, KN, 0.0, 0.5))
    KN00 = grid(gpu, make_params(gpu, KN, 0.0, 0.0))
    d = compare_grids(S[0], S[1], RN0[0], RN0[1])
    say("RN(q=0) vs Schw", d["mismatch"], d["max"])
    assert d["mismatch"] == 0 and d["max"] < 1e-4
    d = compare_grids(K6[0], K6[1], KN60[0], KN60[1])
    say("KN(.6,0) vs Kerr(.6)", d["mismatch"], d["max"])
    assert d["mismatch"] == 0 and d["max"] < 1e-4
    n = NX * NY
    for A, B, name in [(S, K0, "Kerr(0)~Schw"), (RN5, KN05, "KN(0,.5)~RN(.5)"), (S, KN00, "KN(0,0)~Schw")]:
        d = compare_grids(A[0], A[1], B[0], B[1])
        say(name, d["mismatch"], d["median"])
        assert d["mismatch"] <= int(0.025 * n) and d["median"] < 0.02
    assert int((RN5[0] == 0).sum()) < int((S[0] == 0).sum())
    Kp = grid(gpu, make_params(gpu, KERR, 0.7, 0.0))
    Km = grid(gpu, make_params(gpu, KERR, -0.7, 0.0))
    d = compare_grids(Kp[0], Kp[1], Km[0], Km[1])
    assert d["mismatch"] > 0 or d["median"] > 1e-3

    def bias(o):
        cap = (o == 0)
        return int(cap[:, : NX // 2].sum() - cap[:, NX // 2:].sum())

    bp, bm = bias(Kp[0]), bias(Km[0])
    if bp != 0 or bm != 0:
        assert bp * bm <= 0
    naked = gpu.probe_grid(make_params(gpu, KERR, 1.2, 0.0, 0.012, 4000, True), 48, 27)
    assert np.isin(naked["outcome"], [0, 1]).all()


# ----------------------------------------------------------------- T4 (test_main.cu:448-470)
def test_T4_constraint(gpu):
    for args in [(KERR, 0.9, 0.0), (KN, 0.7, 0.5)]:
        g = gpu.probe_grid(make_params(gpu, *args), NX, NY)
        v = np.sort(g["maxHviol"].ravel())
        med, mx = float(v[len(v) // 2]), float(v[-1])
        say("T4", args, "median %.2e max %.2e" % (med, mx))
        assert med < 5e-3 and mx < 5e-2


# ----------------------------------------------------------------- T5 (test_main.cu:475-522)
def test_T5_convergence_and_shadow(gpu):
    A = grid(gpu, make_params(gpu, KERR, 0.9, 0.0, 0.012, 4000))
    B = grid(gpu, make_params(gpu, KERR, 0.9, 0.0, 0.006, 8000))
    Cc = grid(gpu, make_params(gpu, KERR, 0.9, 0.0, 0.003, 16000))
    dAB = compare_grids(A[0], A[1], B[0], B[1])
    dBC = compare_grids(B[0], B[1], Cc[0], Cc[1])
    say("T5 refinement", dAB["median"], dBC["median"])
    assert dBC["median"] <= dAB["median"] * 1.2 + 1e-5 and dBC["median"] < 1e-2

    P = make_params(gpu, SCHW, 0.0, 0.0, 0.006, 8000)
    r0 = 30.0
    P.camPos.x, P.camPos.y, P.camPos.z = r0, 0.0, 0.0
    P.camForward.x, P.camForward.y, P.camForward.z = -1.0, 0.0, 0.0
    P.camRight.x, P.camRight.y, P.camRight.z = 0.0, 0.0, -1.0
    P.camUp.x, P.camUp.y, P.camUp.z = 0.0, 1.0, 0.0
    M = 0.5
    bc = 3.0 * np.sqrt(3.0) * M
    alpha_th = np.arcsin(bc * np.sqrt(1.0 - 2.0 * M / r0) / r0)
    lo, hi = 0.0, 0.2
    for _ in range(30):
        mid = 0.5 * (lo + hi)
        d = np.array([[-np.cos(mid), np.sin(mid), 0.0]], np.float32)
        d /= np.linalg.norm(d)
        pr = gpu.probe_dirs(P, d)
        if pr["outcome"][0] == 0:
            lo = mid
        else:
            hi = mid
    alpha = 0.5 * (lo + hi)
    rel = abs(alpha - alpha_th) / alpha_th
    say("shadow alpha %.6f theory %.6f rel %.2e" % (alpha, alpha_th, rel))
    assert rel < 0.01


# ----------------------------------------------------------------- T6 (test_main.cu:534-636)
def test_T6_invalid_inputs_render(gpu, oracle):
    W, H = 160, 90
    bad = []
    for k in range(4):
        P = make_params(gpu, KN, 0.7, 0.5)
        P.width, P.height, P.diskEnabled = W, H, 1
        bad.append(P)
    bad[0].dPhi = 0.0
    bad[1].maxSteps = 0
    bad[2].camPos.x, bad[2].camPos.y = float("nan"), float("nan")
    bad[3].aSpin, bad[3].Qc, bad[3].rPlus = 7.0, 9.0, -1.0
    for P in bad:
        say("T6 render with poisoned params")
        img = gpu.render_kernel(P)
        assert (img[..., 3] == 255).all()
        ref = oracle.render_kernel(P)
        st = lsb_stats(img, ref)
        assert st["le1"] >= 0.999, st


def test_T6_pipeline_determinism_and_accumulation(gpu, oracle):
    W, H = 160, 90
    P = make_params(gpu, KERR, 0.9, 0.0)
    P.width, P.height, P.diskEnabled, P.bloomEnabled = W, H, 1, 1
    P.sampleIndex, P.accumMode = 0, 0
    a, _ = gpu.render_pipeline(P)
    b, _ = gpu.render_pipeline(P)
    assert np.array_equal(a, b)
    # 8 progressive samples from a NaN-poisoned accumulation buffer
    poison = np.frombuffer(b"\xff" * (W * H * 16), np.float32).reshape(H, W, 4)
    fb_acc = poison.copy()
    o_acc = poison.copy()
    for s in range(8):
        P.sampleIndex, P.accumMode = s, (0 if s == 0 else 1)
        rgba, fb_acc = gpu.render_pipeline(P, accum_init=fb_acc)
        ref, o_acc = oracle.render_pipeline(P, accum=o_acc)
    assert (rgba[..., 3] == 255).all() and np.isfinite(fb_acc).all()
    st = lsb_stats(rgba, ref)
    say("accumulated 8 samples", st)
    assert st["le1"] >= 0.999
    # EMA mode and invalid accumulation fields
    P.sampleIndex, P.accumMode = 3, 2
    rgba, fb2 = gpu.render_pipeline(P, accum_init=fb_acc)
    ref, _ = oracle.render_pipeline(P, accum=o_acc)
    assert lsb_stats(rgba, ref)["le1"] >= 0.999
    P.sampleIndex, P.accumMode, P.bloomStrength = -7, 99, float("nan")
    rgba, _ = gpu.render_pipeline(P, accum_init=fb2)
    assert (rgba[..., 3] == 255).all()
    # hot spots at a fixed disk time: two frames match
    P = make_params(gpu, KERR, 0.9, 0.0)
    P.width, P.height, P.diskEnabled, P.bloomEnabled = W, H, 1, 1
    P.hotSpotsEnabled, P.hotSpotStrength, P.diskTime = 1, 1.0, 1.25
    a, _ = gpu.render_pipeline(P)
    b, _ = gpu.render_pipeline(P)
    assert np.array_equal(a, b)


# ----------------------------------------------------------------- schedule invariance
@pytest.mark.parametrize("math", [0, 1])
def test_schedule_invariance(gpu, math):
    """Workgroup size, steps per refill chunk, lane refill and cooperative
    disk sampling change only the order in which rays and sub-samples are
    traced, never a result bit."""
    P = scaled_workload(gpu, "kerr09_1080p_high", 0, 256, 144)
    outs = []
    before = gpu.get_schedule()
    try:
        for wg, chunk, refill, coop in [(256, 16, 64, 0), (64, 1, 64, 0), (128, 7, 64, 0), (256, 128, 64, 0),
                                        (64, 4096, 64, 1), (64, 4096, 16, 1), (64, 32, 16, 0), (64, 4096, 64, 2)]:
            gpu.set_schedule(wg=wg, chunk=chunk, math=math, refill=refill, coop=coop)
            say("schedule wg", wg, "chunk", chunk, "refill", refill, "coop", coop, "math", math)
            outs.append(gpu.render_pipeline(P))
    finally:
        gpu.restore_schedule(before)
    for rgba, acc in outs[1:]:
        assert np.array_equal(rgba, outs[0][0])
        assert np.array_equal(acc.view(np.uint32), outs[0][1].view(np.uint32))


# ----------------------------------------------------------------- probes vs oracle
@pytest.mark.parametrize("args", [(SCHW, 0.0, 0.0), (RN, 0.0, 0.5), (KERR, 0.9, 0.0), (KN, 0.7, 0.5),
                                  (KERR, -0.7, 0.0)])
def test_probe_grid_vs_oracle(gpu, oracle, args):
    P = make_params(gpu, *args)
    P.diskEnabled = 1
    say("probe grid vs oracle", args)
    g = gpu.probe_grid(P, NX, NY)
    o = oracle.probe_grid(P, NX, NY, 0)
    d = compare_grids(g["outcome"], g["escDir"], o[..., 0].astype(np.int32), o[..., 1:4])
    say("mismatch", d["mismatch"], "max chord %.2e median %.2e" % (d["max"], d["median"]))
    assert d["mismatch"] <= 2
    assert (d["angles"] <= 1e-3).mean() >= 0.999
    # the per-ray maximum |H| is a float32 round-off record of a chaotic orbit near
    # the photon region: compare its distribution, not ray by ray
    hv_g, hv_o = g["maxHviol"].ravel(), o[..., 4].ravel()
    for q in (0.5, 0.9, 0.99):
        qg, qo = float(np.quantile(hv_g, q)), float(np.quantile(hv_o, q))
        say("maxHviol q%.2f gpu %.3e oracle %.3e" % (q, qg, qo))
        assert qg <= 3.0 * qo + 1e-6 and qo <= 3.0 * qg + 1e-6, (q, qg, qo)
    assert (np.abs(hv_g - hv_o) < 1e-4).mean() >= 0.98
