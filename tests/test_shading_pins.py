"""Pins of the oracle's shading functions against published methods and the
invariants of radiative transfer (CPU).

What each test pins, and the plausible mistake it catches:
  * blackbodyRGB (trace.cuh:101-120) equals Tanner Helland's published
    kelvin -> sRGB fit (2012: "How to Convert Temperature (K) to RGB"), each
    channel clamped to [0, 255] / 255, then raised to 2.2, evaluated here in
    float64 from the published coefficients and branch rules (a wrong
    digit, branch threshold or exponent fails);
  * acesToneMap (post_process.cuh:34-40) equals Narkowicz's published
    rational fit x(2.51x + 0.03) / (x(2.43x + 0.59) + 0.14), clamped;
  * the finalizePixel dither (post_process.cuh:136-144) is triangular on
    (-1, 1) LSB and unbiased: the output histogram of a constant grey frame
    matches the rounding of value + TPDF noise, and its mean equals the
    undithered value (a one-hash rectangular dither, a missing "- 1" or a
    wrong scale fails);
  * sampleDiskSegment (trace.cuh:269-403) obeys the transfer invariants:
    no emission gives no light; no absorption keeps the transmittance at 1;
    light is linear in the emission scale; -ln(transmittance) is linear in
    the absorption scale; hot spots change nothing away from the inner disk;
    at a fixed point the radiance ratio of two photons is
    (g1/g2)^4 * B(T g1) / B(T g2) with g the closed-form emitter g-factor,
    T = 13500 K * F^(1/4) from the closed-form flux weight and B the pinned
    blackbody (a g^3 beaming law, a wrong T_inner or flux normalisation fails);
  * starLayer (trace.cuh:131-162): every star centre lies in its cell's
    [0.1, 0.9]^3 box; the radial profile is exp(-c d^2) + 0.08 exp(-0.1 c d^2)
    around it and exactly zero past the cut; brightness is linear in the
    layer intensity; the colour is a blackbody chromaticity of 2600-14600 K.
Still unpinned (restated from the source, no published form or golden
vector): the fbm octave map (x 2.03 + 11.7, x 2.01 - 5.2), the disk density
and turbulence constants (scale height 0.040 + 0.015 r/r_out, the 0.20/1.30,
0.40/1.00 turbulence mix, the m=1 spiral 0.38 cos(phi - 0.55 Omega t - 1.6 ln
r), hot-spot radii/widths/flicker), the star magnitude hash combination and
brightness law (0.05 m^9 + 9 m^50), and the Milky-Way band/airglow colours.
"""
import numpy as np
import pytest
from scipy import optimize

from helpers import make_params

f32 = np.float32


# ------------------------------------------------------------------ blackbody (Tanner Helland)
def helland_rgb(kelvin):
    """Tanner Helland's published fit, float64, channels in [0, 255]."""
    t = min(max(kelvin, 1200.0), 40000.0) / 100.0
    if t <= 66.0:
        r = 255.0
        g = 99.4708025861 * np.log(t) - 161.1195681661
    else:
        r = 329.698727446 * (t - 60.0) ** -0.1332047592
        g = 288.1221695283 * (t - 60.0) ** -0.0755148492
    if t >= 66.0:
        b = 255.0
    elif t <= 19.0:
        b = 0.0
    else:
        b = 138.5177312231 * np.log(t - 10.0) - 305.0447927307
    return np.clip([r, g, b], 0.0, 255.0)


def test_blackbody_is_tanner_helland_fit(oracle):
    """Compared on the fit's own 0..255 scale (the oracle's float32 form of
    e.g. 138.52 ln(t - 10) - 305.04 cancels to ~1e-4 absolute near t = 20)."""
    ks = np.concatenate([np.linspace(800.0, 42000.0, 1500), [1200, 1900, 2000, 6600, 6700, 40000]])
    worst = 0.0
    for k in ks:
        k32 = float(f32(k))
        ref = helland_rgb(k32)
        got = 255.0 * oracle.blackbody(k32).astype(np.float64) ** (1.0 / 2.2)
        err = np.abs(got - ref)
        assert np.all(err <= 2e-4), (k, got, ref)
        worst = max(worst, float(err.max()))
    assert worst < 2e-4


# ------------------------------------------------------------------ ACES (Narkowicz 2015)
def test_aces_is_narkowicz_fit(oracle):
    xs = np.concatenate([[0.0], np.logspace(-5, 2.5, 600)])
    for x in xs:
        x32 = float(f32(x))
        ref = x32 * (2.51 * x32 + 0.03) / (x32 * (2.43 * x32 + 0.59) + 0.14)
        ref = min(max(ref, 0.0), 1.0)
        got = float(oracle.aces([x32, x32, x32])[0])
        assert abs(got - ref) <= 4e-7 * max(ref, 1e-6) + 1e-12, (x, got, ref)
    # channels are independent
    v = oracle.aces([0.1, 1.0, 10.0])
    assert np.allclose(v, [oracle.aces([c] * 3)[0] for c in (0.1, 1.0, 10.0)], rtol=0, atol=0)


# ------------------------------------------------------------------ dither
def _tpdf_cdf(x):
    x = np.asarray(x, np.float64)
    return np.where(x <= -1, 0.0, np.where(x <= 0, (x + 1) ** 2 / 2, np.where(x <= 1, 1 - (1 - x) ** 2 / 2, 1.0)))


@pytest.mark.parametrize("grey", [0.02, 0.11, 0.37])
def test_dither_is_unbiased_triangular(bt, oracle, grey):
    P = bt.defaults()
    W = H = 1000
    P.width, P.height = W, H
    P.bloomEnabled = 0
    P.exposure = 1.0
    P.swapRB = 0
    acc = np.zeros((H, W, 4), np.float32)
    acc[..., :3] = grey
    out = oracle.post(P, acc)[..., 0].astype(np.float64)
    # undithered channel value in LSB: ACES (pinned above), no vibrance
    # change on grey, gamma 1/2.2
    a = grey * (2.51 * grey + 0.03) / (grey * (2.43 * grey + 0.59) + 0.14)
    v = 255.0 * a ** (1.0 / 2.2)
    k = np.floor(v)
    f = v - k
    js = np.arange(-1, 3)
    expect = _tpdf_cdf(js + 0.5 - f) - _tpdf_cdf(js - 0.5 - f)
    got = np.array([(out == k + j).mean() for j in js])
    assert got.sum() == 1.0, "outputs outside value -1 .. +2 LSB"
    assert np.abs(got - expect).max() < 4e-3, (got, expect)
    assert abs(out.mean() - v) < 4e-3, (out.mean(), v)


# ------------------------------------------------------------------ disk transfer
def _disk_params(bt, spots=1):
    P = make_params(bt, 2, 0.9, 0.0)
    P.diskEnabled = 1
    P.diskInner = float(bt.load().bhIscoFP64(0.5, float(P.aSpin), 0.0))
    P.diskOuter = 20.0
    P.hotSpotsEnabled = spots
    P.hotSpotStrength = 0.65
    P.diskEmisScale = 1.5
    P.diskAbsScale = 2.2
    return P


def _chord(r, phi, y0, y1, dr=0.3):
    a = np.array([r * np.cos(phi), y0, -r * np.sin(phi)], np.float32)
    b = np.array([(r + dr) * np.cos(phi + 0.02), y1, -(r + dr) * np.sin(phi + 0.02)], np.float32)
    return a, b


CHORDS = [(2.2, 0.3, 0.03, -0.02), (3.0, 1.7, 0.01, 0.05), (4.5, -2.0, -0.04, 0.02), (7.0, 2.9, 0.10, -0.12)]
PHOTON = (1.0, 2.1)   # E, Lz


def test_disk_no_emission_no_light_no_absorption_full_transmittance(bt, oracle):
    P = _disk_params(bt)
    lit = 0
    for (r, phi, y0, y1) in CHORDS:
        a, b = _chord(r, phi, y0, y1)
        acc, tr = oracle.disk_segment(P, *PHOTON, P.aSpin, P.Qc, a, b, 0.5)
        lit += acc.max() > 0
        Q = P.copy(); Q.diskEmisScale = 0.0
        acc0, tr0 = oracle.disk_segment(Q, *PHOTON, P.aSpin, P.Qc, a, b, 0.5)
        assert np.all(acc0 == 0.0) and tr0 == tr
        Q = P.copy(); Q.diskAbsScale = 0.0
        acc1, tr1 = oracle.disk_segment(Q, *PHOTON, P.aSpin, P.Qc, a, b, 0.5)
        assert tr1 == 1.0 and acc1.max() >= acc.max()
    assert lit == len(CHORDS), "every test chord must carry light"


def test_disk_light_linear_in_emission_and_optical_depth_in_absorption(bt, oracle):
    P = _disk_params(bt)
    for (r, phi, y0, y1) in CHORDS:
        a, b = _chord(r, phi, y0, y1)
        P.diskAbsScale = 0.3
        acc, tr = oracle.disk_segment(P, *PHOTON, P.aSpin, P.Qc, a, b, 0.5)
        Q = P.copy(); Q.diskEmisScale = P.diskEmisScale * 2.0
        acc2, tr2 = oracle.disk_segment(Q, *PHOTON, P.aSpin, P.Qc, a, b, 0.5)
        assert np.array_equal(acc2, 2.0 * acc) and tr2 == tr
        Q = P.copy(); Q.diskAbsScale = 0.6
        _, trA = oracle.disk_segment(Q, *PHOTON, P.aSpin, P.Qc, a, b, 0.5)
        assert tr < 1.0
        assert abs(np.log(trA) / np.log(tr) - 2.0) < 2e-4, (np.log(trA), np.log(tr))


def test_hot_spots_change_only_the_inner_disk(bt, oracle):
    P = _disk_params(bt, spots=1)
    Q = P.copy(); Q.hotSpotsEnabled = 0
    rin = P.diskInner
    for (r, phi, y0, y1) in [(1.7 * rin, 0.4, 0.02, -0.01), (2.5 * rin, 2.0, 0.0, 0.03), (5.0, -1.0, 0.05, 0.0)]:
        a, b = _chord(r, phi, y0, y1, dr=0.25)
        assert np.hypot(a[0], a[2]) > 1.6 * rin and np.hypot(b[0], b[2]) > 1.6 * rin
        on = oracle.disk_segment(P, *PHOTON, P.aSpin, P.Qc, a, b, 0.5)
        off = oracle.disk_segment(Q, *PHOTON, P.aSpin, P.Qc, a, b, 0.5)
        assert np.array_equal(on[0], off[0]) and on[1] == off[1]
    # the spots are live: some inner chord over a range of azimuths differs
    diffs = 0
    for phi in np.linspace(-np.pi, np.pi, 24, endpoint=False):
        a, b = _chord(1.06 * rin, phi, 0.0, 0.0, dr=0.02)
        on = oracle.disk_segment(P, *PHOTON, P.aSpin, P.Qc, a, b, 0.5)
        off = oracle.disk_segment(Q, *PHOTON, P.aSpin, P.Qc, a, b, 0.5)
        diffs += not np.array_equal(on[0], off[0])
    assert diffs > 0


def test_disk_beaming_is_g4_times_blackbody_at_T_g(bt, oracle):
    """Single sub-sample chord (length < 1/6.5) at fixed position, no spots,
    no absorption: only the emitter g-factor differs between photons."""
    P = _disk_params(bt, spots=0)
    P.diskAbsScale = 0.0
    rin = float(P.diskInner)
    for (r, phi) in [(2.0, 0.5), (3.3, -1.2), (6.0, 2.4)]:
        a, b = _chord(r, phi, 0.01, -0.01, dr=0.05)
        assert np.linalg.norm(b - a) < 1.0 / 6.5
        mid = 0.5 * (a + b)
        rc = float(np.hypot(mid[0], mid[2]))
        Fr = float(oracle.flux_weight(f32(rc), f32(rin))) / 0.075
        T = min(max(13500.0 * Fr ** 0.25, 2000.0), 26000.0)
        ph1, ph2 = (1.0, 3.0), (1.0, -3.0)   # co- and counter-rotating photons
        g1 = oracle.emitter_g(rc, P.M, P.aSpin, P.Qc, *ph1)
        g2 = oracle.emitter_g(rc, P.M, P.aSpin, P.Qc, *ph2)
        assert abs(g1 / g2 - 1.0) > 0.1
        acc1, _ = oracle.disk_segment(P, *ph1, P.aSpin, P.Qc, a, b, 0.0)
        acc2, _ = oracle.disk_segment(P, *ph2, P.aSpin, P.Qc, a, b, 0.0)
        B1 = (helland_rgb(min(max(T * g1, 1200.0), 40000.0)) / 255.0) ** 2.2
        B2 = (helland_rgb(min(max(T * g2, 1200.0), 40000.0)) / 255.0) ** 2.2
        expect = (g1 / g2) ** 4 * B1 / B2
        got = acc1.astype(np.float64) / acc2
        assert np.allclose(got, expect, rtol=2e-4), (r, got, expect)


# ------------------------------------------------------------------ star lattice
def _lum(c):
    return 0.2126 * c[0] + 0.7152 * c[1] + 0.0722 * c[2]


LAYERS = [(80.0, 1.00, 320.0), (190.0, 0.55, 520.0), (420.0, 0.30, 820.0), (900.0, 0.14, 1500.0)]


def _star_value(oracle, cell, f, scale, inten, sharp):
    p = (np.asarray(cell, np.float64) + np.asarray(f, np.float64)) / scale
    return oracle.star_layer(p.astype(np.float32), scale, inten, sharp)


def _find_star(oracle, cell, layer, grid):
    scale, inten, sharp = layer
    vals = np.array([_lum(_star_value(oracle, cell, f, scale, inten, sharp)) for f in grid])
    if vals.max() <= 0.0:
        return None
    x0 = grid[int(vals.argmax())]
    res = optimize.minimize(lambda f: -_lum(_star_value(oracle, cell, f, scale, inten, sharp)), x0,
                            method="Nelder-Mead", options=dict(xatol=1e-6, fatol=1e-12, maxiter=4000))
    return res.x


def test_star_centres_fill_the_inner_box(oracle):
    """Star centres over many cells lie in [0.1, 0.9]^3 and spread over it
    (a narrower or wider placement box fails)."""
    rng = np.random.default_rng(11)
    g = (np.arange(12) + 0.5) / 12.0
    grid = np.stack(np.meshgrid(g, g, g, indexing="ij"), -1).reshape(-1, 3)
    centres = []
    for k in range(160):
        layer = LAYERS[k % 4]
        c = _find_star(oracle, rng.integers(-500, 500, 3).astype(np.float64), layer, grid)
        if c is not None:
            centres.append(c)
    C = np.array(centres)
    assert len(C) >= 120
    assert C.min() > 0.1 - 1e-4 and C.max() < 0.9 + 1e-4, (C.min(0), C.max(0))
    assert np.all(C.min(0) < 0.13) and np.all(C.max(0) > 0.87), (C.min(0), C.max(0))


def test_star_profile_cut_intensity_and_colour(oracle):
    rng = np.random.default_rng(7)
    g = (np.arange(21) + 0.5) / 21.0
    grid = np.stack(np.meshgrid(g, g, g, indexing="ij"), -1).reshape(-1, 3)
    checked = 0
    for scale, inten, sharp in LAYERS:
        found = 0
        for _ in range(40):
            cell = rng.integers(-60, 60, 3).astype(np.float64)
            c = _find_star(oracle, cell, (scale, inten, sharp), grid)
            if c is None:
                continue
            peak = _lum(_star_value(oracle, cell, c, scale, inten, sharp))
            # radial profile core + 0.08 glow around the centre
            u = rng.normal(size=3)
            u /= np.linalg.norm(u)
            for d in [0.2, 0.6, 1.2, 2.0, 3.0]:
                d = d / np.sqrt(sharp)
                v = _lum(_star_value(oracle, cell, c + d * u, scale, inten, sharp))
                prof = (np.exp(-sharp * d * d) + 0.08 * np.exp(-0.1 * sharp * d * d)) / 1.08
                assert abs(v / peak - prof) < 2e-3, (d, v / peak, prof)
            # zero past the largest cut (d^2 c > 60), inside the cell
            dcut = np.sqrt(61.0 / sharp)
            for uu in np.eye(3):
                q = c + dcut * uu
                if np.all(q > 0) and np.all(q < 1):
                    assert np.all(_star_value(oracle, cell, q, scale, inten, sharp) == 0.0)
            # linear in the layer intensity
            s1 = _star_value(oracle, cell, c, scale, inten, sharp)
            s2 = _star_value(oracle, cell, c, scale, 2.0 * inten, sharp)
            assert np.array_equal(s2, 2.0 * s1)
            # blackbody chromaticity between 2600 and 14600 K
            col = s1.astype(np.float64) / _lum(s1)

            def bb_norm(T):
                b = (helland_rgb(T) / 255.0) ** 2.2
                return b / _lum(b)

            Ts = np.linspace(2600.0, 14600.0, 2401)
            errs = [np.abs(bb_norm(T) - col).max() for T in Ts]
            k = int(np.argmin(errs))
            T = optimize.minimize_scalar(lambda t: np.abs(bb_norm(t) - col).max(),
                                         bounds=(Ts[max(k - 1, 0)], Ts[min(k + 1, len(Ts) - 1)]),
                                         method="bounded").x
            assert np.abs(bb_norm(T) - col).max() < 1e-4, (T, col)
            checked += 1
            found += 1
            if found == 3:
                break
    assert checked >= 8
