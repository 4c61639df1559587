"""Helpers shared by the tests: the reference test suite's parameter and
camera construction (test_main.cu:111-154), grid comparison
(test_main.cu:187-221) and pixel-parity statistics."""
from __future__ import annotations

import numpy as np

f32 = np.float32


def set_camera(P, az, el, dist):
    """Orbit camera of the reference test suite (test_main.cu:111-130)."""
    ca, sa = f32(np.cos(f32(az))), f32(np.sin(f32(az)))
    ce, se = f32(np.cos(f32(el))), f32(np.sin(f32(el)))
    d = f32(dist)
    p = [d * ce * ca, d * se, d * ce * sa]
    fl = f32(np.sqrt(p[0] * p[0] + p[1] * p[1] + p[2] * p[2]))
    f = [-p[0] / fl, -p[1] / fl, -p[2] / fl]
    r = [-f[2], f32(0), f[0]]
    rl = f32(np.sqrt(r[0] * r[0] + r[2] * r[2]))
    if rl < 1e-6:
        r, rl = [f32(1), f32(0), f32(0)], f32(1)
    r = [r[0] / rl, r[1], r[2] / rl]
    u = [r[1] * f[2] - r[2] * f[1], r[2] * f[0] - r[0] * f[2], r[0] * f[1] - r[1] * f[0]]
    P.camPos.x, P.camPos.y, P.camPos.z = [float(x) for x in p]
    P.camForward.x, P.camForward.y, P.camForward.z = [float(x) for x in f]
    P.camRight.x, P.camRight.y, P.camRight.z = [float(x) for x in r]
    P.camUp.x, P.camUp.y, P.camUp.z = [float(x) for x in u]
    P.tanHalfFov = float(f32(np.tan(f32(60.0 * 3.14159265 / 180.0 * 0.5))))
    P.aspect = float(f32(16.0) / f32(9.0))


def make_params(bt, model, a_star, q, dphi=0.012, max_steps=4000, allow_naked=False):
    """makeParams of the reference test suite (test_main.cu:132-154)."""
    d, M, _, _ = bt.validate_and_derive(model, 0.5, a_star, q, allow_naked)
    P = bt.defaults()
    P.model = model
    P.M = M
    P.aSpin, P.Qc = d.a, d.Q
    P.rPlus, P.rPhoton, P.rErgo = d.rPlus, d.rPhoton, d.rErgo
    P.allowNaked = 1 if allow_naked else 0
    P.diskInner = d.rIsco
    P.diskOuter = 8.0
    P.diskEnabled = 0
    P.dPhi = dphi
    P.maxSteps = max_steps
    set_camera(P, 0.6, 0.25, 16.0)
    return P


def scaled_workload(bt, name, frame, w, h):
    """A named workload at another resolution (same camera, field of view and
    aspect when w/h matches)."""
    P = bt.workload(name, frame)
    P.width, P.height = w, h
    P.aspect = float(f32(w) / f32(h))
    return P


def compare_grids(oa, ea, ob, eb):
    """compareGrids (test_main.cu:195-221): outcome mismatches and chord
    length between exit directions of rays escaping in both grids."""
    oa, ob = np.asarray(oa).ravel(), np.asarray(ob).ravel()
    ea, eb = np.asarray(ea, np.float64).reshape(-1, 3), np.asarray(eb, np.float64).reshape(-1, 3)
    mism = int((oa != ob).sum())
    both = (oa == 1) & (ob == 1)
    ang = np.sqrt(((ea[both] - eb[both]) ** 2).sum(-1))
    med = float(np.sort(ang)[len(ang) // 2]) if len(ang) else 0.0
    mx = float(ang.max()) if len(ang) else 0.0
    return dict(n=len(oa), mismatch=mism, median=med, max=mx, angles=ang)


def lsb_stats(a, b):
    """Per-pixel max channel difference of two RGBA8 images."""
    d = np.abs(a[..., :3].astype(np.int32) - b[..., :3].astype(np.int32)).max(-1)
    return dict(le1=float((d <= 1).mean()), exact=float((d == 0).mean()), max=int(d.max()),
                n_gt1=int((d > 1).sum()), n=int(d.size))


def angle_between(a, b):
    a = np.asarray(a, np.float64).reshape(-1, 3)
    b = np.asarray(b, np.float64).reshape(-1, 3)
    a = a / np.linalg.norm(a, axis=1, keepdims=True)
    b = b / np.linalg.norm(b, axis=1, keepdims=True)
    c = np.sqrt(((a - b) ** 2).sum(-1))
    return 2.0 * np.arcsin(np.clip(c / 2.0, 0.0, 1.0))


def pixel_dirs(P, xs, ys, jx=0.5, jy=0.5):
    """Pinhole directions of pixels (xs, ys) in float64 (cameraRayDir,
    trace.cuh:431-438)."""
    F = np.array([P.camForward.x, P.camForward.y, P.camForward.z])
    R = np.array([P.camRight.x, P.camRight.y, P.camRight.z])
    U = np.array([P.camUp.x, P.camUp.y, P.camUp.z])
    fx = (np.asarray(xs, np.float64) + jx) / P.width
    fy = (np.asarray(ys, np.float64) + jy) / P.height
    d = F + R * ((fx * 2 - 1) * P.tanHalfFov * P.aspect)[:, None] + U * ((1 - fy * 2) * P.tanHalfFov)[:, None]
    return d / np.linalg.norm(d, axis=1, keepdims=True)


def ill_conditioned(oracle, P, pix, eps=1e-6, amp=1e-3):
    """Rays (flat pixel indices) whose FP64 trace is ill-conditioned: moving
    the initial direction by eps rad along the camera's right or up axis
    changes the outcome or turns the exit direction by more than amp rad (an
    amplification of amp / eps).  For such rays two float traces that differ
    in the last bit of any intermediate cannot be expected to agree."""
    pix = np.asarray(pix, np.int64)
    if pix.size == 0:
        return np.zeros(0, bool)
    ys, xs = np.divmod(pix, P.width)
    d = pixel_dirs(P, xs, ys)
    base = oracle.probe_dirs(P, d.astype(np.float32), mode=1)
    ill = np.zeros(pix.size, bool)
    for ax in ((P.camRight.x, P.camRight.y, P.camRight.z), (P.camUp.x, P.camUp.y, P.camUp.z)):
        dp = d + eps * np.asarray(ax, np.float64)
        dp /= np.linalg.norm(dp, axis=1, keepdims=True)
        o = oracle.probe_dirs(P, dp.astype(np.float32), mode=1)
        ill |= o[:, 0] != base[:, 0]
        esc = (o[:, 0] == 1) & (base[:, 0] == 1)
        ill[esc] |= angle_between(o[esc, 1:4], base[esc, 1:4]) > amp
    return ill
