#!/usr/bin/env python3
"""Throughput of the geodesic tracer on one MI355X.

A step is one full frame of the hot path -- trace, sky shading,
accumulation, bloom and tonemap through launchRenderPipeline -- on the
Kerr a*=0.9 1920x1080 'high' workload, cycling the disk time over
t = 0, 0.5, 1.  Inputs are device resident; timing brackets exactly
`--steps` frames, enqueued back to back, with device synchronizations,
after `--warmup` frames.  Per-kernel times (`kernel_ms`, the rooflines)
come from a second pass over the same frames with HIP-event timers.

Prints one JSON line.  `roofline` is the FP32 tile kernel's achieved
fraction of the FP32 vector peak: the exact algorithmic flops of the pixels
it traces (counted-float oracle, see FLOPS below) over its HIP-event time on
the stream it runs on.  `post_roofline` is bloom + tonemap against the HBM peak
(in the frame, and alone with the caches flushed before each run).
`fp64_error` is the exit-direction error of every pixel of each frame
against a full-frame FP64 trace.  `other_configs` times the other workloads
(Schwarzschild 256^2, Kerr-Newman 4K, the 64-view batch).  `surrogate` trains the MLP on 2^24 rays
with the library's HIP kernels and times MLP + RK4 fallback on the same
frames.  `cpu_baseline` is the CPU oracle (same algorithm, g++ -O3, OpenMP)
on a bounded sample of the same frame.
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "black-hole_amd"))

METRIC = "Mrays/s Kerr a*=0.9 1080p: plain FP32 RK4 vs MLP+fallback at <=1e-3 rad"
WORKLOAD = "kerr09_1080p_high"
FP32_PEAK_TFLOPS = 157.3
HBM_PEAK_GBS = 8000.0

# Exact algorithmic flops per frame of WORKLOAD (ray integration and disk
# radiative transfer of every pixel), and the share of the FP64 band pixels
# (the default band rule: 8,336 pixels, traced by the FP64 band kernel, not
# the FP32 tile kernel), from
#   python oracle/count_flops.py kerr09_1080p_high 0 1 2
# (counted-float instantiation of the oracle; 1 flop per +,-,*,/,sqrt,
# floor, transcendental).  Keyed by frame index (disk time 0, 0.5, 1).
FLOPS = {0: 506225668732, 1: 506225457204, 2: 506228100901}
BAND_FLOPS = {0: 4791139978, 1: 4793639269, 2: 4794055461}
# HBM bytes per launch of the FP32 tile kernel on WORKLOAD (FP64 band on):
# rocprofv3 --pmc FETCH_SIZE and WRITE_SIZE runs of this script, scaled by a
# plain-copy calibration (FETCH x2.0, WRITE x1.0); profiles/pmc_trace_kerr09_1080p_high.md
TRAFFIC_BYTES = (7611 * 2 + 65680) * 1024
FP32_NONPACKED_TFLOPS = 78.6
# The surrogate network: 224 -> 128 -> 128 -> 128 -> 128 -> 16 (padded)
# multiply-adds per row, 2 flops each; a training step is the forward pass
# and the two backward GEMMs per layer (3x) over a batch of 65536 rows.
MLP_MACS_PER_ROW = 224 * 128 + 3 * 128 * 128 + 128 * 16
MLP_FLOPS_PER_ROW = 2 * MLP_MACS_PER_ROW
MLP_TRAIN_FLOPS_PER_STEP = 3 * MLP_FLOPS_PER_ROW * 65536
MFMA_BF16_PEAK_TFLOPS = 2500.0
# The same counts for frame 0 of the other workloads: (every pixel, FP64 band
# pixels' share), from `python oracle/count_flops.py <workload> 0`.
OTHER_FLOPS = {"schw_256_balanced": (2202725945, 23447943),
               "kn_m07_q04_2160p_high": (1845561825607, 31190190725)}


_T0 = time.perf_counter()


def progress(msg):
    """One progress line on stderr (the JSON result is the only stdout line)."""
    print(f"[bench {time.perf_counter() - _T0:7.1f} s] {msg}", file=sys.stderr, flush=True)


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=9)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-other-configs", action="store_true",
                    help="skip the timing of the other workloads and the 64-view batch")
    ap.add_argument("--no-fp64-error", action="store_true",
                    help="skip the per-frame exit error against a full-frame FP64 trace")
    ap.add_argument("--cpu-row-step", type=int, default=4)
    ap.add_argument("--no-parity", action="store_true",
                    help="skip the full-frame LSB comparison of the headline frame with the CPU oracle")
    ap.add_argument("--math", type=int, default=None, help="0 exact, 1 fast (default: library default)")
    ap.add_argument("--band", type=int, default=None, help="FP64 critical-band pass 0/1 (default: library default)")
    ap.add_argument("--surrogate-steps", type=int, default=20000,
                    help="train the learned surrogate this many Adam steps and time MLP+fallback (0: skip)")
    ap.add_argument("--surrogate-sigma", type=float, default=3e-4, help="router: max predicted RMS exit error (rad)")
    ap.add_argument("--copy-calibration", type=int, default=0, metavar="MIB",
                    help="only run 3 device-to-device copies of MIB MiB (HBM counter calibration) and exit")
    return ap.parse_args()


def copy_calibration(mib):
    """Three plain device-to-device copies of `mib` MiB each: under
    rocprofv3 --pmc FETCH_SIZE / WRITE_SIZE this gives the counter reading of
    a known byte count (mib MiB read + mib MiB written per copy)."""
    import bhtrace as bt

    n = mib << 20
    src, dst = bt.DeviceBuffer(n), bt.DeviceBuffer(n)
    src.fill(1)
    bt.synchronize()
    for _ in range(3):
        bt.check(bt.load().bhCopyDevice(dst.ptr, src.ptr, n), "bhCopyDevice")
    bt.synchronize()
    print(json.dumps({"copy_calibration_bytes": n, "copies": 3}), flush=True)


def surrogate_leg(bt, lib, frames, steps, sigma, plain_ms):
    """The learned surrogate: trained by the library's HIP kernels on 2^24
    plain-kernel rays (4096 configurations x 4096 rays, seed 2), attached
    and evaluated
      * on the headline frames: the camera sits inside the disk's emitting
        layer, so the router does not consult the network (frame time vs
        plain reported);
      * on the headline camera with the disk off (the lensed sky), with the
        general network and with it adapted to the view (2^22 rays of that
        camera): candidates, accepted rays, exit error of the accepted rays
        against the library's full-frame FP64 trace (max, p99.9, over 1e-3
        rad), and trace time against the plain path."""
    import ctypes as C

    import numpy as np

    route_ms = []
    infer_roof = None

    def trace_ms(P, n=5):
        fb = bt.Framebuffer(P.width, P.height)
        fb.render(P)
        bt.synchronize()
        lib.bhEnableKernelTimers(1)
        tr, rk, ms = C.c_float(), C.c_float(), []
        for _ in range(n):
            fb.render(P)
            lib.bhLastKernelTimes(C.byref(tr), None, None)
            ms.append(tr.value)
            if lib.bhLastSurrogateKernelMs(C.byref(rk), None) == 0:
                route_ms.append(rk.value)
        lib.bhEnableKernelTimers(0)
        fb.close()
        return float(np.median(ms))

    def angles(a, b):
        a = a.astype(np.float64)
        b = b.astype(np.float64)
        a /= np.linalg.norm(a, axis=1, keepdims=True)
        b /= np.linalg.norm(b, axis=1, keepdims=True)
        return 2 * np.arcsin(np.clip(np.linalg.norm(a - b, axis=1) / 2, 0, 1))

    t0 = time.perf_counter()
    s = bt.Surrogate(seed=2)
    s.dataset(4096, 4096, seed=2)
    progress("surrogate: training set traced")
    t1 = time.perf_counter()
    loss = s.train(steps, batch=65536, lr=2e-3)
    bt.synchronize()
    step_ms = (time.perf_counter() - t1) * 1e3 / max(steps, 1)
    train_s = time.perf_counter() - t0
    progress(f"surrogate: {steps} Adam steps, {step_ms:.3f} ms/step")
    out = {"train_steps": steps, "train_s": round(train_s, 2), "adam_step_ms": round(step_ms, 4),
           "final_loss": [round(float(x), 5) for x in loss[-50:].mean(0)], "router_sigma": sigma}
    try:
        P = frames[0]
        bt.set_surrogate(s, sigma, 1e-3)
        bt.trace_records(P)
        cand, acc = bt.last_surrogate_stats()
        t_on = trace_ms(P)
        bt.set_surrogate(None)
        out["headline"] = {"candidates": int(cand), "accepted": int(acc), "trace_ms": round(t_on, 3),
                           "plain_trace_ms": round(plain_ms, 3), "speedup_vs_plain": round(plain_ms / t_on, 3),
                           "note": "camera inside the disk's emitting layer: network not consulted"}
        Q = P.copy()
        Q.diskEnabled = 0
        plain = bt.trace_records(Q)
        t_plain = trace_ms(Q)
        p0, s0 = bt.get_precision(), bt.get_schedule()
        try:
            bt.set_precision(band=1, crit=1e30, polar=1e30)
            bt.set_schedule(math=0)
            ref = bt.trace_records(Q)
        finally:
            bt.restore_precision(p0)
            bt.restore_schedule(s0)
        pe = (plain["flags"] & 1) == 1
        pa = angles(plain["escDir"][pe & ((ref["flags"] & 1) == 1)], ref["escDir"][pe & ((ref["flags"] & 1) == 1)])
        disk_off = {"plain_trace_ms": round(t_plain, 3), "plain_err_vs_fp64": {"max": float("%.3e" % pa.max()),
                                                                               "p999": float("%.3e" % np.quantile(pa, 0.999))}}
        for tag in ("general", "view_adapted"):
            if tag == "view_adapted":
                s.dataset_for([Q], 1 << 22, seed=11)
                s.train(10000, batch=65536, lr=5e-4)
                bt.synchronize()
            bt.set_surrogate(s, sigma, 1e-3)
            rec = bt.trace_records(Q)
            cand, acc = bt.last_surrogate_stats()
            route_ms.clear()
            t_s = trace_ms(Q)
            bt.set_surrogate(None)
            m = ((rec["flags"] >> 8) == 0) & ((rec["flags"] & 1) == 1) & ((ref["flags"] & 1) == 1)
            r = {"candidates": int(cand), "accepted": int(acc), "fallback_frac": round(1.0 - acc / rec.size, 5),
                 "trace_ms": round(t_s, 3), "speedup_vs_plain": round(t_plain / t_s, 3),
                 "mrays_trace": round(rec.size / t_s / 1e3, 3)}
            if m.any():
                e = angles(rec["escDir"][m], ref["escDir"][m])
                r["accepted_err_vs_fp64"] = {"max": float("%.3e" % e.max()), "p999": float("%.3e" % np.quantile(e, 0.999)),
                                             "p50": float("%.3e" % np.median(e)), "over_1e-3": int((e > 1e-3).sum())}
            if route_ms:
                rms = float(np.median(route_ms))
                r["route_kernel_ms"] = round(rms, 4)
                tf = cand * MLP_FLOPS_PER_ROW / (rms * 1e-3) / 1e12
                infer_roof = {"bound": "mfma_bf16", "kernel": "surrogateRouteKernel<8, 4>",
                              "kernels_timed": ["surrogateRouteKernel<8, 4>", "surrogateFinishKernel"], "rows": int(cand),
                              "flops_per_launch": int(cand * MLP_FLOPS_PER_ROW), "kernel_ms": round(rms, 4),
                              "achieved": round(tf, 2), "peak": MFMA_BF16_PEAK_TFLOPS, "unit": "TFLOP/s",
                              "frac": round(tf / MFMA_BF16_PEAK_TFLOPS, 4)}
            disk_off[tag] = r
            progress(f"surrogate ({tag}, disk off): accepted {acc} of {cand}, x{t_plain / t_s:.3f}")
        out["headline_camera_disk_off"] = disk_off
        tf = MLP_TRAIN_FLOPS_PER_STEP / (step_ms * 1e-3) / 1e12
        out["mlp_roofline"] = {
            "train": {"bound": "mfma_bf16", "kernels": ["gemmKernel<true, 64>", "reducePartsKernel", "featuresKernel",
                                                        "lossHeadKernel", "adamKernel", "zeroKernel"],
                      "flops_per_step": MLP_TRAIN_FLOPS_PER_STEP, "step_ms": round(step_ms, 4),
                      "achieved": round(tf, 2), "peak": MFMA_BF16_PEAK_TFLOPS, "unit": "TFLOP/s",
                      "frac": round(tf / MFMA_BF16_PEAK_TFLOPS, 4)},
            "infer": infer_roof}
    finally:
        bt.set_surrogate(None)
        s.close()
    return out


def fp64_error_leg(bt, frames):
    """Exit-direction error of the product path (the current precision and
    schedule) on every pixel of each frame, against the library's own FP64
    trace of the whole frame (every pixel in the FP64 band, exact math; that
    trace is checked against the oracle's FP64 trace in
    tests/test_gpu_fp64_band.py)."""
    import numpy as np

    p0, s0 = bt.get_precision(), bt.get_schedule()
    out = []
    for k, P in enumerate(frames):
        prod = bt.trace_records(P)
        try:
            bt.set_precision(band=1, crit=1e30, polar=1e30)
            bt.set_schedule(math=0)
            ref = bt.trace_records(P)
        finally:
            bt.restore_precision(p0)
            bt.restore_schedule(s0)
        _, st = bt.exit_angle_diff(prod, ref)
        out.append({"disk_time": P.diskTime, "rays": int(P.width * P.height), "escaped_both": st.bothEscaped,
                    "outcome_mismatch": st.outcomeMismatch, "max_rad": float("%.3e" % st.max),
                    "p999_rad": float("%.3e" % st.p999), "over_1e-3": st.over1mrad})
    return out


def post_roofline(W, H, post_ms):
    """Bloom + tonemap against the HBM peak.  Algorithmic bytes are the
    compulsory 20 B per pixel: the float4 accumulation read once and the
    uchar4 frame written once, which is also what the fused kernel moves
    (its half-resolution bloom image stays on chip; the reference's unfused
    chain moves 60 B per pixel)."""
    nbytes = 20 * W * H
    gbs = nbytes / (post_ms * 1e-3) / 1e9
    return {"bound": "hbm", "achieved": round(gbs, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
            "frac": round(gbs / HBM_PEAK_GBS, 4), "bytes_per_frame": nbytes, "ms": round(post_ms, 4),
            "kernels": ["postFusedKernel"]}


def post_flushed(bt, lib, fb, P, reps=5):
    """Bloom + tonemap alone (bhFramebufferPost) on the frame's HDR image,
    with the L2 and Infinity Cache flushed before each run by a 512 MiB
    device-to-device copy, and back to back without the flush."""
    import ctypes as C

    n = 512 << 20
    src, dst = bt.DeviceBuffer(n), bt.DeviceBuffer(n)
    po = C.c_float()
    hot, cold = [], []
    lib.bhEnableKernelTimers(1)
    for flush, out in ((False, hot), (True, cold)):
        for _ in range(reps):
            if flush:
                bt.check(lib.bhCopyDevice(dst.ptr, src.ptr, n), "bhCopyDevice")
            fb.post(P)
            bt.check(lib.bhLastKernelTimes(None, None, C.byref(po)), "bhLastKernelTimes")
            out.append(po.value)
    lib.bhEnableKernelTimers(0)
    bt.synchronize()
    src.close()
    dst.close()
    return sorted(hot)[reps // 2], sorted(cold)[reps // 2]


def other_configs(bt, lib, fp64):
    """The other workloads at their full sizes (not the headline metric):
    trace time and frame rate of one-frame launches, their exit error against
    a full-frame FP64 trace (when `fp64`), and the 64-view uneven batch
    (views rendered as single-view pipelines on the library's worker streams,
    and in one persistent launch with the FP64 band pass, with every view's
    exit error against its FP64 trace when `fp64`)."""
    import ctypes as C

    import numpy as np

    out = {}
    for name in ("schw_256_balanced", "kn_m07_q04_2160p_high"):
        P = bt.workload(name, 0)
        fb = bt.Framebuffer(P.width, P.height)
        fb.render(P)
        bt.synchronize()
        lib.bhEnableKernelTimers(1)
        tr, ms, tile = C.c_float(), [], []
        t0 = time.perf_counter()
        for _ in range(5):
            fb.render(P)
            lib.bhLastKernelTimes(C.byref(tr), None, None)
            ms.append(tr.value)
            tile.append(bt.last_trace_kernel_ms()[0])
        bt.synchronize()
        dt = (time.perf_counter() - t0) / 5
        lib.bhEnableKernelTimers(0)
        fb.close()
        out[name] = {"width": P.width, "height": P.height, "ms_per_frame": round(dt * 1e3, 3),
                     "trace_ms": round(float(np.median(ms)), 3), "mrays_s": round(P.width * P.height / dt / 1e6, 2)}
        prec = bt.get_precision()
        default_band = abs(prec.critBand - 0.04) < 1e-6 and abs(prec.polarBand - 1e-3) < 1e-9
        if prec.fp64Band and default_band:
            fl = OTHER_FLOPS[name][0] - OTHER_FLOPS[name][1]
            tms = float(np.mean(tile))
            ach = fl / (tms * 1e-3) / 1e12
            out[name]["roofline"] = {"bound": "valu_fp32", "kernel": "traceTileKernel", "flops_per_launch": fl,
                                     "kernel_ms": round(tms, 3), "achieved": round(ach, 3), "unit": "TFLOP/s",
                                     "frac": round(ach / FP32_PEAK_TFLOPS, 4),
                                     "frac_of_nonpacked_fp32": round(ach / FP32_NONPACKED_TFLOPS, 4)}
        if fp64:
            out[name]["fp64_error"] = fp64_error_leg(bt, [P])[0]
        progress(f"{name}: {out[name]['mrays_s']} Mrays/s")
    views = bt.workload_views("uneven_views", 1, 64)
    B = bt.BatchBuffers(views)
    B.render()
    bt.synchronize()
    dt, lat = None, None
    for _ in range(3):   # best of three batches
        t0 = time.perf_counter()
        B.render()
        bt.synchronize()
        t = time.perf_counter() - t0
        if dt is None or t < dt:
            dt, lat = t, B.latency_ms()
    rays = sum(P.width * P.height for P in views)
    out["uneven_views_64"] = {"rays": rays, "batch_ms": round(dt * 1e3, 2), "mrays_s": round(rays / dt / 1e6, 2),
                              "view_latency_ms_p50": round(float(np.median(lat)), 2),
                              "view_latency_ms_max": round(float(max(lat)), 2),
                              "launch": "bhRenderBatch: single-view pipelines (FP64 band on) on 6 worker streams"}
    # the same views in one persistent launch, eager and as a replayed
    # graph: with the FP64 band pass (the library default: every view
    # classified, the band rays of all views traced in one launch beside the
    # FP32 trace) and, for reference, every ray in FP32 (band off)
    # A second set of 64 views (seed 2) alternates with the first for the
    # 'fresh' figure: each batch then holds views not traced just before, so
    # its work order comes from the band statistics, not step counts.
    B2 = bt.BatchBuffers(bt.workload_views("uneven_views", 2, 64))

    def persistent(band):
        p0 = bt.get_precision()
        try:
            bt.set_precision(band=band)
            bt.check(B.render_persistent(), "bhRenderBatchPersistent")
            bt.synchronize()
            pe = None
            for _ in range(3):
                t0 = time.perf_counter()
                bt.check(B.render_persistent(), "bhRenderBatchPersistent")
                bt.synchronize()
                t = time.perf_counter() - t0
                pe = t if pe is None or t < pe else pe
            pn = None
            for _ in range(3):
                bt.check(B2.render_persistent(), "bhRenderBatchPersistent")
                bt.synchronize()
                t0 = time.perf_counter()
                bt.check(B.render_persistent(), "bhRenderBatchPersistent")
                bt.synchronize()
                t = time.perf_counter() - t0
                pn = t if pn is None or t < pn else pn
            G = B.graph(autotune=1)
            wg, tuned = G.info()
            G.launch()
            bt.synchronize()
            reps = 5
            t0 = time.perf_counter()
            for _ in range(reps):
                G.launch()
            bt.synchronize()
            tg = (time.perf_counter() - t0) / reps
            glat = G.latency_ms()
            G.close()
        finally:
            bt.restore_precision(p0)
        return {"eager_batch_ms": round(pe * 1e3, 2), "eager_mrays_s": round(rays / pe / 1e6, 2),
                "fresh_eager_batch_ms": round(pn * 1e3, 2), "fresh_eager_mrays_s": round(rays / pn / 1e6, 2),
                "graph_batch_ms": round(tg * 1e3, 2), "graph_mrays_s": round(rays / tg / 1e6, 2),
                "trace_latency_ms_p50": round(float(np.median(glat)), 2),
                "trace_latency_ms_max": round(float(max(glat)), 2),
                "trace_wg_size": wg, "autotune_ms_wg64_128_256": [round(x, 2) for x in tuned]}

    pb = persistent(1)
    progress(f"uneven_views_64 persistent (FP64 band): {pb['graph_mrays_s']} Mrays/s")
    pf = persistent(0)
    out["uneven_views_64_persistent"] = dict(
        rays=rays, **pb,
        launch="bhRenderBatchPersistent / bhBatchGraph: every view classified in one launch, then one persistent FP32 "
               "trace over a device queue of all views' tiles (costliest first) beside one FP64 trace of all views' "
               "band rays over a queue of (view, 64-ray chunk) items; batched shade/bloom/composite",
        fp32_only={k: pf[k] for k in ("eager_batch_ms", "eager_mrays_s", "fresh_eager_batch_ms", "graph_batch_ms",
                                      "graph_mrays_s")})
    del B2
    if fp64:
        # exit directions of every view against the library's FP64 trace of
        # the view (the batch's records equal the view's own band-on trace)
        p0, s0 = bt.get_precision(), bt.get_schedule()
        worst, p999, over, mism = 0.0, 0.0, 0, 0
        for V in views:
            Q, _ = bt.sanitized(V)
            prod = bt.trace_records(Q)
            try:
                bt.set_precision(band=1, crit=1e30, polar=1e30)
                bt.set_schedule(math=0)
                ref = bt.trace_records(Q)
            finally:
                bt.restore_precision(p0)
                bt.restore_schedule(s0)
            _, st = bt.exit_angle_diff(prod, ref)
            worst, p999 = max(worst, st.max), max(p999, st.p999)
            over += st.over1mrad
            mism += st.outcomeMismatch
        out["uneven_views_64_persistent"]["fp64_error"] = {
            "views": len(views), "max_rad": float("%.3e" % worst), "max_view_p999_rad": float("%.3e" % p999),
            "over_1e-3": int(over), "outcome_mismatch": int(mism)}
    return out


def parity_leg(bt, P):
    """The headline frame (t = 0) of the product against the CPU oracle:
    share of pixels whose channels all lie within 1 LSB, against the oracle
    with its FP64 band rule off (the reference's all-FP32 arithmetic) and
    with the library's band rule."""
    import numpy as np

    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import pyoracle

    O = pyoracle.Oracle()
    rgba, _ = bt.render_pipeline(P)
    prec = bt.get_precision()
    out = {}
    for key, band in (("vs_fp32_reference_semantics", 0), ("vs_oracle_with_band_rule", 1)):
        O.set_band(band and prec.fp64Band, prec.critBand, prec.polarBand)
        t0 = time.perf_counter()
        ref, _ = O.render_pipeline(P)
        progress(f"parity: oracle frame ({key}) in {time.perf_counter() - t0:.1f} s")
        d = np.abs(rgba[..., :3].astype(np.int32) - ref[..., :3].astype(np.int32)).max(-1)
        out[key] = {"le1_lsb": round(float((d <= 1).mean()), 6), "exact": round(float((d == 0).mean()), 6),
                    "pixels_over_1lsb": int((d > 1).sum()), "max_lsb": int(d.max())}
    O.set_band(0)
    out["frame"] = f"{WORKLOAD} t=0, {P.width}x{P.height}, product with FP64 band {'on' if prec.fp64Band else 'off'}"
    return out


def host_cpus():
    """(logical CPUs this process may run on, physical cores among them,
    total logical CPUs, CPU model name) from the affinity mask and sysfs /
    procfs topology."""
    cpus = sorted(os.sched_getaffinity(0))
    cores = set()
    for c in cpus:
        base = f"/sys/devices/system/cpu/cpu{c}/topology/"
        try:
            with open(base + "core_id") as f1, open(base + "physical_package_id") as f2:
                cores.add((f2.read().strip(), f1.read().strip()))
        except OSError:
            cores.add(("?", str(c)))
    model = "unknown"
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    return len(cpus), len(cores), os.cpu_count(), model


def cpu_baseline(bt, P, row_step, others, host):
    """CPU oracle (float, reference semantics: the FP32 integrators, disk, sky
    and post chain, g++ -O3 -ffp-contract=off, OpenMP dynamic over 16x8-pixel
    tiles) on every `row_step`-th row: one warm-up pass, then the median of
    three timed passes over the same rows, one thread per physical core of
    this host bound close (OMP_PROC_BIND=close, OMP_PLACES=cores, set at the
    start of main before the OpenMP runtime starts).  With `others`, the same for a row sample of
    each of the other workloads."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import pyoracle

    O = pyoracle.Oracle()
    logical, physical, nproc, model = host
    threads = physical
    progress(f"cpu baseline: {model}, nproc {nproc}, affinity {logical} CPUs, {physical} physical cores")

    def sample(Q, step, reps=1):
        O.trace_frame(Q, 0, 0, step, threads)            # warm-up
        times = []
        for _ in range(3):
            t = time.perf_counter()
            for _ in range(reps):
                hdr, _, _, _ = O.trace_frame(Q, 0, 0, step, threads)
            times.append(time.perf_counter() - t)
        dt = sorted(times)[1]
        rays = hdr.shape[0] * hdr.shape[1] * reps
        return rays / dt / 1e6, rays, dt

    v, rays, dt = sample(P, row_step)
    progress(f"cpu baseline: {v:.4f} Mrays/s on {threads} threads")
    res = {"value": round(v, 4), "unit": "Mrays/s", "cores": threads, "kind": "port",
           "sample": f"{WORKLOAD} t=0, every {row_step}th row ({rays} rays), median of 3 passes after a warm-up "
                     f"({dt:.2f} s each, {rays / dt / 1e6 / threads:.4f} Mrays/s per core)",
           "host": {"cpu_model": model, "nproc": nproc, "affinity_cpus": logical, "physical_cores": physical,
                    "omp_proc_bind": "close", "omp_places": "cores",
                    "flags": "g++ -O3 -ffp-contract=off -fno-fast-math -fopenmp (portable x86-64, no -march=native)"}}
    if others:
        res["other_configs"] = {}
        for name, step, reps in (("schw_256_balanced", 1, 8), ("kn_m07_q04_2160p_high", 16, 1)):
            v, rays, dt = sample(bt.workload(name, 0), step, reps)
            what = "every row" if step == 1 else f"every {step}th row"
            res["other_configs"][name] = {"value": round(v, 4), "unit": "Mrays/s",
                                          "sample": f"t=0, {what} x{reps} ({rays} rays), median of 3 after a "
                                                    f"warm-up ({dt:.2f} s)"}
    return res


def main():
    a = parse()
    # the host's CPUs, read before the oracle's OpenMP runtime starts (with
    # OMP_PROC_BIND it binds this thread to one core, which then is all the
    # affinity mask shows)
    host = host_cpus()
    # thread binding of the CPU oracle's OpenMP runtime (read when the
    # runtime starts, i.e. when the oracle library is first loaded)
    os.environ["OMP_PROC_BIND"] = "close"
    os.environ["OMP_PLACES"] = "cores"
    if a.gpus != 1:
        print(f"bench.py: this build runs on one GPU; --gpus {a.gpus} is not supported "
              "(a multi-GPU run would shard the pixels of a frame across GPUs by rows)", file=sys.stderr)
        sys.exit(2)
    import bhtrace as bt

    if a.copy_calibration > 0:
        copy_calibration(a.copy_calibration)
        return
    lib = bt.load()
    if lib.bhDeviceCount() < 1:
        print("bench.py: no HIP device", file=sys.stderr)
        sys.exit(1)
    if a.math is not None:
        bt.set_schedule(math=a.math)
    if a.band is not None:
        bt.set_precision(band=a.band)
    sched = bt.get_schedule()
    prec = bt.get_precision()
    frames = [bt.workload(WORKLOAD, f) for f in range(3)]
    W, H = frames[0].width, frames[0].height
    fb = bt.Framebuffer(W, H)
    for i in range(a.warmup):
        fb.render(frames[i % 3])
    bt.synchronize()

    import ctypes as C

    # the timed frames: enqueued back to back, no host synchronization
    # between the two device synchronizations that bracket them
    t0 = time.perf_counter()
    for i in range(a.steps):
        fb.render(frames[i % 3])
    bt.synchronize()
    total = time.perf_counter() - t0

    # kernel breakdown: the same frames again with HIP-event timers, each
    # frame's events read after it (this host wait is outside the timed loop)
    lib.bhEnableKernelTimers(1)
    tr, sh, po = C.c_float(), C.c_float(), C.c_float()
    trace_ms, shade_ms, post_ms, tile_ms, band_ms, per_frame = [], [], [], [], [], []
    for i in range(a.steps):
        f = i % 3
        fb.render(frames[f])
        bt.check(lib.bhLastKernelTimes(C.byref(tr), C.byref(sh), C.byref(po)), "bhLastKernelTimes")
        tk, bk = bt.last_trace_kernel_ms()
        trace_ms.append(tr.value)
        shade_ms.append(sh.value)
        post_ms.append(po.value)
        tile_ms.append(tk)
        band_ms.append(bk)
        per_frame.append(f)
    bt.synchronize()
    band_rays, _ = bt.last_band_stats()
    lib.bhEnableKernelTimers(0)

    ms_per_step = total * 1e3 / a.steps
    mrays = W * H * a.steps / total / 1e6
    avg_trace = sum(trace_ms) / len(trace_ms)
    avg_tile = sum(tile_ms) / len(tile_ms)
    roof = None
    default_band = abs(prec.critBand - 0.04) < 1e-6 and abs(prec.polarBand - 1e-3) < 1e-9
    if FLOPS and (default_band or not prec.fp64Band):
        # the tile kernel's algorithmic flops: every pixel's, less the band
        # pixels' when the band kernel traces those
        fl = sum(FLOPS[f] - (BAND_FLOPS[f] if prec.fp64Band else 0) for f in per_frame) / len(per_frame)
        achieved = fl / (avg_tile * 1e-3) / 1e12
        roof = {"bound": "valu_fp32", "achieved": round(achieved, 3), "peak": FP32_PEAK_TFLOPS, "unit": "TFLOP/s",
                "frac": round(achieved / FP32_PEAK_TFLOPS, 4),
                "traffic": TRAFFIC_BYTES if prec.fp64Band else None,
                "kernel": "traceTileKernel", "flops_per_launch": int(fl), "kernel_ms": round(avg_tile, 3),
                "frac_of_nonpacked_fp32": round(achieved / FP32_NONPACKED_TFLOPS, 4)}
    res = {
        "metric": METRIC, "value": round(mrays, 3), "unit": "Mrays/s", "n_gpus": 1, "steps": a.steps,
        "warmup": a.warmup, "ms_per_step": round(ms_per_step, 3), "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "fp32", "data": "synthetic",
        "config": {"workload": WORKLOAD, "width": W, "height": H, "disk_times": [0.0, 0.5, 1.0],
                   "path": "plain RK4: FP32 bulk + FP64 critical-band rays (no surrogate)" if prec.fp64Band
                   else "plain FP32 RK4 (no surrogate)",
                   "fp64_band": {"on": prec.fp64Band, "crit": prec.critBand, "polar": prec.polarBand,
                                 "band_rays": band_rays},
                   "math_mode": "fast" if sched.mathMode else "exact",
                   "wg_size": sched.wgSize, "steps_per_chunk": sched.stepsPerChunk},
        "kernel_ms": {"trace": round(avg_trace, 3), "tile_kernel": round(avg_tile, 3),
                      "band_kernel": round(sum(band_ms) / len(band_ms), 3),
                      "shade": round(sum(shade_ms) / len(shade_ms), 3),
                      "bloom_tonemap": round(sum(post_ms) / len(post_ms), 3)},
        "roofline": roof,
        "post_roofline": post_roofline(W, H, sum(post_ms) / len(post_ms)),
    }
    hot, cold = post_flushed(bt, lib, fb, frames[0])
    res["post_roofline"]["post_only_ms"] = round(hot, 4)
    res["post_roofline"]["cache_flushed"] = {k: v for k, v in post_roofline(W, H, cold).items()
                                             if k in ("achieved", "frac", "ms")}
    progress(f"headline: {mrays:.1f} Mrays/s, {ms_per_step:.3f} ms/frame")
    if not a.no_fp64_error:
        res["fp64_error"] = fp64_error_leg(bt, frames)
        progress("fp64 exit error done")
    if not a.no_other_configs:
        res["other_configs"] = other_configs(bt, lib, not a.no_fp64_error)
        progress("other configs done")
    if a.surrogate_steps > 0:
        res["surrogate"] = surrogate_leg(bt, lib, frames, a.surrogate_steps, a.surrogate_sigma, avg_trace)
        progress("surrogate done")
    if not a.no_parity:
        res["parity"] = parity_leg(bt, frames[0])
        progress("parity done")
    if not a.no_cpu_baseline:
        res["cpu_baseline"] = cpu_baseline(bt, frames[0], a.cpu_row_step, not a.no_other_configs, host)
        progress("cpu baseline done")
    print(json.dumps(res), flush=True)


if __name__ == "__main__":
    main()
