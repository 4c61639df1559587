"""Pins of the oracle's FP64 critical-band test against closed forms (CPU).

The band test has no reference counterpart; its two statistics are checked
against what the mathematics fixes:

* on the critical curve of Kerr-Newman, parametrised by the radius r of the
  spherical photon orbit (double root of the radial potential R(r) = 0 =
  R'(r)), the normalised potential minimum `delta` is zero, and it moves off
  zero with the right sign (captured above the curve, turning point below);
* for a = 0 the scan reproduces the closed form 1 - b^2/b_c^2 (b_c^2 = 27 M^2
  for Schwarzschild) from a camera ray's own invariants;
* the polar statistic equals sin^2 of the theta turning point found by
  bisection on Theta(theta) = eta + a^2 cos^2 - xi^2 cot^2.
"""
import numpy as np
import pytest


def critical_curve(M, a, Q, r):
    """(xi, eta) of the spherical photon orbit at radius r: from R = 0 and
    R' = 0 with R = (r^2 + a^2 - a xi)^2 - Delta (eta + (xi - a)^2)."""
    D = r * r - 2 * M * r + a * a + Q * Q
    Dp = 2 * r - 2 * M
    P = 4 * r * D / Dp               # = r^2 + a^2 - a xi
    xi = (r * r + a * a - P) / a
    K = P * P / D                    # = eta + (xi - a)^2
    return xi, K - (xi - a) ** 2


def photon_shell(M, a, Q):
    """Radii where the critical curve has eta >= 0 (orbits reachable from
    the equatorial plane and above), by sampling."""
    rp = M + np.sqrt(M * M - a * a - Q * Q)
    rs = np.linspace(rp * 1.02, 4.5 * M, 4000)
    ok = []
    for r in rs:
        xi, eta = critical_curve(M, a, Q, r)
        if eta > 1e-3 and np.isfinite(xi) and r * r - 2 * M * r + a * a + Q * Q > 0 and 2 * r - 2 * M > 0:
            ok.append(r)
    return np.array(ok)


@pytest.mark.parametrize("M,a,Q", [(0.5, 0.45, 0.0), (0.5, -0.35, 0.2), (0.5, 0.2, 0.3), (0.5, 0.49, 0.0)])
def test_delta_vanishes_on_critical_curve(oracle, M, a, Q):
    Mf, af, Qf = (float(np.float32(x)) for x in (M, a, Q))
    rs = photon_shell(Mf, af, Qf)
    assert len(rs) > 10
    for r in rs[:: max(1, len(rs) // 12)]:
        xi, eta = critical_curve(Mf, af, Qf, r)
        d0 = oracle.L.bho_band_delta(Mf, af, Qf, xi, eta)
        assert abs(d0) < 1e-7, (r, xi, eta, d0)
        # above the curve (larger eta) R > 0 everywhere: captured, delta > 0;
        # below it a forbidden region opens: delta < 0
        up = oracle.L.bho_band_delta(Mf, af, Qf, xi, eta * 0.98)
        dn = oracle.L.bho_band_delta(Mf, af, Qf, xi, eta * 1.02)
        assert up * dn < 0, (r, up, dn)


def test_delta_schwarzschild_closed_form(bt, oracle):
    """Kerr model with a = 0: delta of every pixel equals 1 - b^2 / (27 M^2)."""
    P = bt.workload("kerr09_1080p_high", 0)
    P.aSpin = 0.0
    P.width, P.height = 64, 36
    m, d, pol = oracle.band_mask(P, 0.03, 0.0)
    # impact parameter of each pixel's ray from its direction (Schwarzschild)
    o = np.array([P.camPos.x, P.camPos.y, P.camPos.z], np.float64)
    F = np.array([P.camForward.x, P.camForward.y, P.camForward.z], np.float64)
    R = np.array([P.camRight.x, P.camRight.y, P.camRight.z], np.float64)
    U = np.array([P.camUp.x, P.camUp.y, P.camUp.z], np.float64)
    iy, ix = np.mgrid[0:P.height, 0:P.width]
    fx = (ix.astype(np.float64) + 0.5) / P.width
    fy = (iy.astype(np.float64) + 0.5) / P.height
    dirs = F + R * ((fx * 2 - 1) * P.tanHalfFov * P.aspect)[..., None] + U * ((1 - fy * 2) * P.tanHalfFov)[..., None]
    dirs /= np.linalg.norm(dirs, axis=-1, keepdims=True)
    r0 = np.linalg.norm(o)
    e1 = o / r0
    perp = dirs - (dirs @ e1)[..., None] * e1
    pl = np.linalg.norm(perp, axis=-1)
    f0 = 1 - 2 * P.M / r0
    b2 = (r0 * pl) ** 2 / f0
    expect = 1 - b2 / (27 * P.M * P.M)
    assert np.allclose(d, expect, atol=2e-9, rtol=0), np.abs(d - expect).max()


@pytest.mark.parametrize("a,xi,eta", [(0.45, 0.3, 5.0), (0.45, 0.02, 6.0), (-0.35, -1.2, 3.0), (0.3, 0.001, 1.5),
                                      (0.0, 0.5, 4.0)])
def test_polar_turning_point_by_bisection(oracle, a, xi, eta):
    def theta_pot(th):
        c, s = np.cos(th), np.sin(th)
        return eta + a * a * c * c - xi * xi * c * c / (s * s)

    lo, hi = 1e-9, np.pi / 2   # Theta < 0 near the pole (xi != 0), > 0 at the equator
    for _ in range(200):
        mid = 0.5 * (lo + hi)
        if theta_pot(mid) < 0:
            lo = mid
        else:
            hi = mid
    s2 = np.sin(0.5 * (lo + hi)) ** 2
    got = oracle.L.bho_band_polar(a, xi, eta)
    assert abs(got - s2) < 1e-10, (got, s2)


def test_band_fraction_small(bt, oracle):
    """The band stays a thin set of pixels on the configured workloads."""
    for name in ["kerr09_1080p_high", "kn_m07_q04_2160p_high", "schw_256_balanced"]:
        P = bt.workload(name, 0)
        P.width, P.height = 320, 180
        P.aspect = float(np.float32(320) / np.float32(180))
        m, _, _ = oracle.band_mask(P)
        assert 0 < m.mean() < 0.03, (name, m.mean())
