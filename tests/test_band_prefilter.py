"""The FP32 estimate of the band statistics that lets the classification skip
the FP64 evaluation far from the band (csrc/bh_band.h bandStatsF32).

A float32 restatement of its arithmetic (camera terms, invariants, 32-point
scan, 10-step golden-section search, polar turning point; the potential's
fused multiply-adds rounded once, from exact float64 products of the float32
operands) is compared with the
oracle's FP64 delta and sin^2 theta_min on Kerr-family views with a horizon.
The classification trusts the estimate only farther than 0.01 (delta) and
1e-3 (polar) from the band's edges; its error must stay an order of
magnitude below those margins, and no ray the FP64 rule puts in the band may
fall outside them.  (The GPU tests check that the device's band masks equal
the oracle's.)
"""
import numpy as np
import pytest

F = np.float32
MARGIN_DELTA, MARGIN_POLAR = 0.01, 1e-3


def camera_terms(P):
    a = float(P.aSpin)
    M, Q = float(P.M), float(P.Qc)
    ox, oy, oz = float(P.camPos.x), float(P.camPos.y), float(P.camPos.z)
    r0 = np.sqrt(ox * ox + oy * oy + oz * oz)
    cth0 = min(max(oy / r0, -1.0), 1.0)
    th0, ph0 = np.arccos(cth0), np.arctan2(-oz, ox)
    s0 = max(np.sin(th0), float(F(1e-4)))
    s02 = s0 * s0
    Sig0 = r0 * r0 + a * a * cth0 * cth0
    m2r0 = 2.0 * M * r0 - Q * Q
    gtt = min(-(1.0 - m2r0 / Sig0), float(F(-1e-5)))
    gtp = -a * m2r0 * s02 / Sig0
    gpp = (r0 * r0 + a * a + m2r0 * a * a * s02 / Sig0) * s02
    A = 1.0 / np.sqrt(max(gpp - gtp * gtp / gtt, float(F(1e-8))))
    return dict(a=a, M=M, Q=Q, s=np.sin(th0), c=np.cos(th0), sp=np.sin(ph0), cp=np.cos(ph0), cth0=cth0,
                s02=s02, gtt=gtt, gtp=gtp, gpp=gpp, A=A, pt0=1.0 / np.sqrt(-gtt), pthS=np.sqrt(Sig0))


def pixel_dirs64(P):
    W, H = P.width, P.height
    ys, xs = np.mgrid[0:H, 0:W]
    fx = (xs + 0.5) / float(F(W))
    fy = (ys + 0.5) / float(F(H))
    nx, ny = fx * 2.0 - 1.0, 1.0 - fy * 2.0
    Fw = np.array([P.camForward.x, P.camForward.y, P.camForward.z], np.float64)
    R = np.array([P.camRight.x, P.camRight.y, P.camRight.z], np.float64)
    U = np.array([P.camUp.x, P.camUp.y, P.camUp.z], np.float64)
    v = Fw + R * (nx * float(P.tanHalfFov) * float(P.aspect))[..., None] + U * (ny * float(P.tanHalfFov))[..., None]
    return v / np.sqrt(np.maximum((v * v).sum(-1), 1e-20))[..., None]


def estimate32(P, lo, hi):
    ct = camera_terms(P)
    d = pixel_dirs64(P).astype(F)
    s, c, sp, cp = F(ct["s"]), F(ct["c"]), F(ct["sp"]), F(ct["cp"])
    dtc = (d[..., 0] * (c * cp) + d[..., 1] * (-s)) + d[..., 2] * (-c * sp)
    dpc = d[..., 0] * (-sp) + d[..., 2] * (-cp)
    A, gtp, gtt, gpp = F(ct["A"]), F(ct["gtp"]), F(ct["gtt"]), F(ct["gpp"])
    ptUp = F(ct["pt0"]) - dpc * A * gtp * (F(1) / gtt)
    pfUp = dpc * A
    E = -(gtt * ptUp + gtp * pfUp)
    Lz = gtp * ptUp + gpp * pfUp
    a, cth0 = F(ct["a"]), F(ct["cth0"])
    pth = dtc * F(ct["pthS"])
    carter = pth * pth + cth0 * cth0 * (Lz * Lz * (F(1) / F(ct["s02"])) - a * a * E * E)
    iE = F(1) / E
    xi, eta = Lz * iE, carter * iE * iE
    M, Q = F(ct["M"]), F(ct["Q"])
    xa = xi - a
    etx, axi, a2Q2 = eta + xa * xa, a * xi, a * a + Q * Q

    a2s, twoM = a * a, F(2) * M

    def fma(x, y, z):
        # one rounding: the float64 product of two float32 values is exact
        x, y, z = np.asarray(x, F), np.asarray(y, F), np.asarray(z, F)
        return (x.astype(np.float64) * y.astype(np.float64) + z.astype(np.float64)).astype(F)

    def pot(r):
        r = np.asarray(r, F)
        ra = fma(r, r, a2s)
        P_ = ra - axi
        Del = fma(r, r - twoM, a2Q2)
        return fma(P_, P_, -(Del * etx)) * (F(1) / (ra * ra))

    lo, hi = F(lo), F(hi)
    dr = (hi - lo) * (F(1) / F(31))
    vb = pot(np.full_like(xi, lo))
    best = np.zeros(xi.shape, np.int32)
    for i in range(1, 32):
        v = pot(np.full_like(xi, lo + dr * F(i)))
        m = v < vb
        vb, best = np.where(m, v, vb), np.where(m, i, best)
    x0 = lo + dr * np.maximum(best - 1, 0).astype(F)
    x3 = lo + dr * np.minimum(best + 1, 31).astype(F)
    gr = F(0.618034)
    x1, x2 = x3 - gr * (x3 - x0), x0 + gr * (x3 - x0)
    f1, f2 = pot(x1), pot(x2)
    for _ in range(10):
        m = f1 < f2
        a3, a2, af2 = x2, x1, f1
        a1 = a3 - gr * (a3 - x0)
        b0, b1, bf1 = x1, x2, f2
        b2 = b0 + gr * (x3 - b0)
        fa1, fb2 = pot(a1), pot(b2)
        x0, x1 = np.where(m, x0, b0), np.where(m, a1, b1)
        x2, x3 = np.where(m, a2, b2), np.where(m, a3, x3)
        f1, f2 = np.where(m, fa1, bf1), np.where(m, af2, fb2)
    delta = np.minimum(vb, np.minimum(f1, f2))
    a2_ = a * a
    Bq = a2_ - eta - xi * xi
    D = np.sqrt(np.maximum(Bq * Bq + F(4) * a2_ * eta, F(0)))
    with np.errstate(all="ignore"):
        u = np.where(Bq <= 0, np.where(D - Bq > 0, F(2) * eta * (F(1) / (D - Bq)), F(1)),
                     np.where(a2_ > 0, (Bq + D) * (F(1) / (F(2) * a2_)), F(1)))
    ok = (E > F(1e-6)) & np.isfinite(E) & np.isfinite(Lz) & np.isfinite(delta)
    return delta, F(1) - u, ok


def views(bt):
    out = []
    for name, frame in (("kerr09_1080p_high", 0), ("kn_m07_q04_2160p_high", 0)):
        P = bt.workload(name, frame)
        P.width, P.height = 192, 108
        P.aspect = float(F(192) / F(108))
        out.append((name, P))
    for model, a_star, q in ((3, 0.3, 0.9), (2, -0.99, 0.0), (3, 0.9, 0.3)):
        P = bt.workload("kerr09_1080p_high", 0)
        d, M, a, qq = bt.validate_and_derive(model, 0.5, a_star, q, False)
        P.model, P.aSpin, P.Qc = model, a * 0.5, qq * 0.5
        P.rPlus, P.rPhoton, P.rErgo = d.rPlus, d.rPhoton, d.rErgo
        P.width, P.height = 192, 108
        P.aspect = float(F(192) / F(108))
        out.append(("model%d a*=%g q=%g" % (model, a_star, q), P))
    return out


def test_fp32_band_estimate_within_margins(bt, oracle):
    for label, P in views(bt):
        Q_, _ = bt.sanitized(P)
        assert Q_.rPlus > 0.0
        m, d64, p64 = oracle.band_mask(Q_, 0.04, 1e-3)
        rp = bt.load().bhHorizonOuterFP64(float(Q_.M), float(Q_.aSpin), float(Q_.Qc))
        lo = rp * 1.001
        hi = max(5.0 * float(Q_.M), lo * 1.5)
        d32, p32, ok = estimate32(Q_, lo, hi)
        fin = ok & np.isfinite(d64)
        assert fin.mean() > 0.99, label
        ed = np.abs(d32[fin].astype(np.float64) - d64[fin]).max()
        ep = np.abs(p32[fin].astype(np.float64) - p64[fin]).max()
        trusted = fin & ~((np.abs(d32) < 0.04 + MARGIN_DELTA) | (p32 < 1e-3 + MARGIN_POLAR))
        band64 = (np.abs(d64) < 0.04) | (p64 < 1e-3)
        print(label, "delta err %.2e polar err %.2e trusted %.3f" % (ed, ep, trusted.mean()))
        assert ed < MARGIN_DELTA / 10, (label, ed)
        assert ep < MARGIN_POLAR / 10, (label, ep)
        assert not (trusted & band64).any(), label
        assert trusted.mean() > 0.8, label
