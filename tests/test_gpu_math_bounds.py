"""Error bounds of the library's own elementary-function approximations,
evaluated on the GPU exactly as the fast trace kernels evaluate them
(bhMathProbe), against long-double references on the host.

* atan2Poly (the fast policy's atan2: degree-7 polynomial in a^2 on the
  octant, then reflections): absolute error over a dense grid of angles and
  magnitudes;
* the fast policy's hardware sin/cos (of theta and of the winding phi), rcp,
  rsq, sqrt, and exp/log/pow from hardware exp2/log2;
* the fast policy's disk blackbody (the 2.2 power of Tanner Helland's fit
  taken in the log2 domain) and hot-spot azimuth wrap (floor instead of
  fmodf), against the published fit / the exact wrap in float64 and against
  the exact policy's own evaluation on the GPU;
* the FP64 band integrator's sin/cos (Cody-Waite reduction modulo pi/2 with
  a two-term constant, Taylor polynomials to degree 15/16) over |x| <= 300,
  its angle-addition step for |dl| < 1/16, and its reciprocal (hardware
  seed, two Newton steps).
Each bound below is the one the design relies on; the measured maxima are
printed.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

LD = np.longdouble
ATAN2, SINCOS, RCP, RSQ, SQRT, EXP, LOG, POW, D64SC, D64NEAR, D64RCP, BLACKBODY, WRAPPHI = range(13)


def say(*a):
    print("[gpu-test]", *a, flush=True)


def f32(x):
    return np.asarray(x, np.float32).astype(np.float64)


def test_atan2_poly_bound(gpu):
    ang = np.linspace(-np.pi, np.pi, 200001)
    rad = np.array([1e-3, 0.37, 1.0, 7.5, 1e3], np.float64)
    y = f32((rad[:, None] * np.sin(ang)[None, :]).ravel())
    x = f32((rad[:, None] * np.cos(ang)[None, :]).ravel())
    got = gpu.math_probe(ATAN2, y, x)[:, 0]
    ref = np.arctan2(y.astype(LD), x.astype(LD))
    err = np.abs(got.astype(LD) - ref).astype(np.float64)
    # the branch cut: atan2(+0, -x) = pi, atan2(-0, -x) = -pi
    err = np.minimum(err, np.abs(np.abs(err) - 2 * np.pi))
    say("atan2Poly: max abs error %.3e rad" % err.max())
    assert err.max() <= 3.5e-7


def test_fast_sin_cos_bounds(gpu):
    # theta in [0, pi] (polar angle, RK4 stages) and the winding azimuth phi
    for lo, hi, bound in ((0.0, np.pi, 1.5e-6), (-60.0, 60.0, 8e-6)):
        x = f32(np.linspace(lo, hi, 400001))
        r = gpu.math_probe(SINCOS, x)
        es = np.abs(r[:, 0].astype(LD) - np.sin(x.astype(LD))).astype(np.float64).max()
        ec = np.abs(r[:, 1].astype(LD) - np.cos(x.astype(LD))).astype(np.float64).max()
        say("hardware sin/cos on [%g, %g]: max abs error %.3e / %.3e" % (lo, hi, es, ec))
        assert es <= bound and ec <= bound


def test_fast_reciprocal_and_roots(gpu):
    x = f32(np.exp(np.linspace(np.log(1e-6), np.log(1e6), 200001)))
    for fn, name, ref in ((RCP, "rcp", 1 / x.astype(LD)), (RSQ, "rsq", 1 / np.sqrt(x.astype(LD))),
                          (SQRT, "sqrt", np.sqrt(x.astype(LD)))):
        got = gpu.math_probe(fn, x)[:, 0]
        rel = np.abs((got.astype(LD) - ref) / ref).astype(np.float64).max()
        say("%s: max relative error %.3e" % (name, rel))
        assert rel <= 2.5e-7


def test_fast_exp_log_pow(gpu):
    x = f32(np.linspace(-20.0, 20.0, 200001))
    got = gpu.math_probe(EXP, x)[:, 0]
    ref = np.exp(x.astype(LD))
    rel = np.abs((got.astype(LD) - ref) / ref).astype(np.float64)
    bound = 3e-7 + 8e-8 * np.abs(x)             # the argument's scaling by log2(e) in float
    say("exp: max relative error %.3e (bound at |x| = 20: %.1e)" % (rel.max(), bound.max()))
    assert np.all(rel <= bound)
    x = f32(np.exp(np.linspace(np.log(1e-6), np.log(1e6), 200001)))
    got = gpu.math_probe(LOG, x)[:, 0]
    ref = np.log(x.astype(LD))
    err = np.abs(got.astype(LD) - ref).astype(np.float64)
    lim = 2.5e-7 * (1.0 + np.abs(ref.astype(np.float64)))
    say("log: max absolute error %.3e, worst error / bound %.3f at x = %.4g" %
        (err.max(), (err / lim).max(), x[np.argmax(err / lim)]))
    assert np.all(err <= lim)
    # pow as the disk sampler and the blackbody use it: bases in (0, 40],
    # exponents 0.25 (flux -> temperature), 2.2 (gamma), 9, 50 (star magnitudes)
    for e in (0.25, 2.2, -0.1332047592, 9.0, 50.0):
        b = f32(np.linspace(1e-3, 40.0 if e < 9 else 1.0, 40001))
        got = gpu.math_probe(POW, b, np.float32(e))[:, 0]
        ref = b.astype(LD) ** LD(np.float32(e))
        fin = np.isfinite(ref.astype(np.float64)) & (ref > LD(1e-30))
        rel = np.abs((got.astype(LD)[fin] - ref[fin]) / ref[fin]).astype(np.float64)
        lim = 5e-7 + 2e-7 * abs(np.float32(e)) * (1.0 + np.abs(np.log2(b[fin])))
        say("pow(b, %g): max relative error %.3e" % (e, rel.max()))
        assert np.all(rel <= lim)


def test_d64_sincos_bound(gpu):
    x = np.linspace(-300.0, 300.0, 600001)
    r = gpu.math_probe(D64SC, x)
    es = np.abs(r[:, 0].astype(LD) - np.sin(x.astype(LD))).astype(np.float64).max()
    ec = np.abs(r[:, 1].astype(LD) - np.cos(x.astype(LD))).astype(np.float64).max()
    say("D64 sincos |x| <= 300: max abs error %.3e / %.3e" % (es, ec))
    assert es <= 2.5e-16 and ec <= 2.5e-16


def test_d64_sincos_near_bound(gpu):
    rng = np.random.default_rng(5)
    x0 = rng.uniform(0.0, np.pi, 200000)
    dl = rng.uniform(-0.0624, 0.0624, 200000)
    r = gpu.math_probe(D64NEAR, x0, dl)
    xs = x0.astype(LD) + dl.astype(LD)
    es = np.abs(r[:, 0].astype(LD) - np.sin(xs)).astype(np.float64).max()
    ec = np.abs(r[:, 1].astype(LD) - np.cos(xs)).astype(np.float64).max()
    say("D64 sincos by angle addition, |dl| < 1/16: max abs error %.3e / %.3e" % (es, ec))
    assert es <= 5e-16 and ec <= 5e-16


def test_d64_rcp_bound(gpu):
    x = np.exp(np.linspace(np.log(1e-8), np.log(1e8), 400001))
    x = np.concatenate([x, -x])
    got = gpu.math_probe(D64RCP, x)[:, 0]
    ref = 1 / x.astype(LD)
    rel = np.abs((got.astype(LD) - ref) / ref).astype(np.float64).max()
    say("D64 rcp: max relative error %.3e" % rel)
    assert rel <= 2.3e-16


def helland_linear(kelvin):
    """(Tanner Helland's fit / 255)^2.2 in float64: the disk's linear colour."""
    t = np.clip(kelvin, 1200.0, 40000.0) / 100.0
    with np.errstate(invalid="ignore", divide="ignore"):
        lo = t <= 66.0
        r = np.where(lo, 255.0, 329.698727446 * np.abs(t - 60.0) ** -0.1332047592)
        g = np.where(lo, 99.4708025861 * np.log(t) - 161.1195681661, 288.1221695283 * np.abs(t - 60.0) ** -0.0755148492)
        b = np.where(t >= 66.0, 255.0, np.where(t <= 19.0, 0.0, 138.5177312231 * np.log(np.maximum(t - 10.0, 1e-9)) - 305.0447927307))
    return (np.clip(np.stack([r, g, b]), 0.0, 255.0) / 255.0) ** 2.2


def test_fast_blackbody_bound(gpu):
    k = f32(np.concatenate([np.linspace(1000.0, 41000.0, 120001), [1900, 2000, 6599.99, 6600, 6600.01, 6700]]))
    ref = helland_linear(k)
    worst = []
    for c in range(3):
        r = gpu.math_probe(BLACKBODY, k, float(c))
        fast, exact = r[:, 0], r[:, 1]
        # absolute on the linear [0, 1] scale; the exact policy's float32
        # form of 138.52 ln(t - 10) - 305.04 cancels to ~1e-4 on the 0..255
        # scale near t = 20, which the 2.2 power then shrinks
        ef = np.abs(fast - ref[c]).max()
        ee = np.abs(exact - ref[c]).max()
        dfe = np.abs(fast - exact).max()
        worst.append((ef, ee, dfe))
        say("blackbody channel %d: fast %.3e, exact %.3e vs the fit; fast vs exact %.3e" % (c, ef, ee, dfe))
        assert ef <= 4e-6 and ee <= 4e-6 and dfe <= 4e-6


def test_fast_hot_spot_wrap_bound(gpu):
    d = f32(np.concatenate([np.linspace(-200.0, 200.0, 400001), [-np.pi, np.pi, 0.0, 2 * np.pi, -2 * np.pi]]))
    r = gpu.math_probe(WRAPPHI, d)
    # both forms wrap e = d + pi as rounded in float (d + kPi in the kernel)
    e = (d.astype(np.float32) + np.float32(np.pi)).astype(LD)
    ref = np.mod(e, 2 * LD(np.float32(np.pi))) - LD(np.float32(np.pi))
    e_fast = np.abs(r[:, 0].astype(LD) - ref).astype(np.float64)
    e_ex = np.abs(r[:, 1].astype(LD) - ref).astype(np.float64)
    # a value at the period's edge may land on either end: compare on the circle
    e_fast = np.minimum(e_fast, np.abs(e_fast - 2 * np.pi))
    e_ex = np.minimum(e_ex, np.abs(e_ex - 2 * np.pi))
    say("hot-spot wrap |d| <= 200: fast %.3e, fmodf %.3e rad" % (e_fast.max(), e_ex.max()))
    assert np.all((r[:, 0] >= -np.pi - 1e-6) & (r[:, 0] <= np.pi + 1e-6))
    assert e_fast.max() <= 1e-6 and e_ex.max() <= 1e-6
