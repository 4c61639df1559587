"""CPU tests of the C ABI library's host side: exported symbols, the public
headers, parameter validation/derivation (reference T1/T2,
test_main.cu:226-335), launch-time sanitation, FP64 roots, camera and
workload presets.  No GPU call is made."""
import ctypes as C
import math
import os
import re
import subprocess

import numpy as np
import pytest

from helpers import make_params

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HDR = os.path.join(ROOT, "include", "bh", "bhtrace.h")
SCHW, RN, KERR, KN = 0, 1, 2, 3


def declared_functions():
    txt = open(HDR).read()
    txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
    names = re.findall(r"^[A-Za-z_][\w\s\*]*?\b(\w+)\s*\(", txt, flags=re.M)
    return sorted(set(n for n in names if n not in ("defined",)))


def test_library_exports_every_declared_symbol(bt):
    lib = bt.load()
    names = declared_functions()
    assert "launchRenderPipeline" in names and "launchRenderKernel" in names
    bound = {n for n, _, _ in bt.SIGNATURES}
    for n in names:
        assert hasattr(lib, n), f"{n} declared in bhtrace.h but not exported"
        assert n in bound, f"{n} has no ctypes signature in bhtrace.py"


def test_public_headers_compile_as_c99_and_cxx(tmp_path):
    src = tmp_path / "t.c"
    src.write_text('#include "bh/bhtrace.h"\nint main(void){RenderParams p; bhDefaultRenderParams(&p); '
                   'return (int)sizeof(p) - 160;}\n')
    for cc, flags in [("gcc", ["-std=c99", "-x", "c"]), ("g++", ["-std=c++17", "-x", "c++"])]:
        r = subprocess.run([cc, *flags, "-fsyntax-only", "-I", os.path.join(ROOT, "include"), str(src)],
                           capture_output=True, text=True)
        assert r.returncode == 0, r.stderr


def test_render_params_layout(bt):
    RP = bt.RenderParams
    expect = dict(width=0, height=4, swapRB=8, camPos=12, camForward=24, camRight=36, camUp=48, tanHalfFov=60,
                  aspect=64, model=68, M=72, aSpin=76, Qc=80, rPlus=84, rPhoton=88, rErgo=92, allowNaked=96,
                  dPhi=100, maxSteps=104, diskEnabled=108, diskInner=112, diskOuter=116, diskTime=120,
                  diskEmisScale=124, diskAbsScale=128, hotSpotsEnabled=132, hotSpotStrength=136, exposure=140,
                  sampleIndex=144, accumMode=148, bloomEnabled=152, bloomStrength=156)
    for k, off in expect.items():
        assert getattr(RP, k).offset == off, k
    assert C.sizeof(RP) == 160


def test_defaults_match_reference(bt):
    P = bt.defaults()
    assert (P.width, P.height, P.model, P.maxSteps, P.diskEnabled, P.bloomEnabled) == (1280, 720, 0, 2000, 1, 1)
    assert math.isclose(P.dPhi, 0.012, rel_tol=1e-6) and math.isclose(P.diskOuter, 8.0)
    assert math.isclose(P.hotSpotStrength, 0.65, rel_tol=1e-6) and math.isclose(P.bloomStrength, 0.35, rel_tol=1e-6)


# ----------------------------------------------------------------- T1 (test_main.cu:226-278)
@pytest.mark.parametrize("model,a,q", [(SCHW, 0.0, 0.0), (RN, 0.0, 0.6), (KERR, 0.8, 0.0), (KN, 0.6, 0.5),
                                       (KN, 0.9, 0.3)])
def test_T1_horizons(bt, model, a, q):
    d, M, _, _ = bt.validate_and_derive(model, 0.5, a, q)
    aa = a * 0.5 if model in (KERR, KN) else 0.0
    QQ = q * 0.5 if model in (RN, KN) else 0.0
    rp = 0.5 + math.sqrt(0.25 - aa * aa - QQ * QQ)
    assert abs(d.rPlus - rp) < 2e-6
    assert abs(bt.load().bhHorizonOuterFP64(0.5, aa, QQ) - rp) < 1e-15


def test_T1_special_cases(bt):
    s, M, _, _ = bt.validate_and_derive(SCHW, 0.5, 0.0, 0.0)
    assert abs(s.rPlus - 1.0) < 1e-6 and abs(s.rIsco - 3.0) < 0.01 and abs(s.rPhoton - 1.5) < 1e-4
    p, *_ = bt.validate_and_derive(KERR, 0.5, 0.9, 0.0)
    m, *_ = bt.validate_and_derive(KERR, 0.5, -0.9, 0.0)
    assert abs(p.rPlus - m.rPlus) < 1e-5 and m.rIsco > p.rIsco + 0.05
    assert abs(m.rPlus - (0.5 + math.sqrt(0.25 - 0.45 ** 2))) < 2e-6


# ----------------------------------------------------------------- T2 (test_main.cu:283-335)
def test_T2_validation(bt):
    d, M, a, q = bt.validate_and_derive(RN, 0.5, 0.0, 2.0)
    assert d.clamped and q <= 0.9951 and math.isfinite(d.rPlus) and d.rPlus > 0
    d, M, a, q = bt.validate_and_derive(KERR, 0.5, 5.0, 0.0)
    assert d.clamped and a <= 0.9951 and d.rPlus > d.rMinus - 1e-6
    d, M, a, q = bt.validate_and_derive(KERR, 0.5, -5.0, 0.0)
    assert d.clamped and a >= -0.9951 and d.rPlus > 0
    d, M, a, q = bt.validate_and_derive(KN, 0.5, 0.9, 0.9)
    assert d.clamped and a * a + q * q <= 0.995 + 1e-4 and math.isfinite(d.rPlus)
    d, M, a, q = bt.validate_and_derive(KERR, 0.5, 1.2, 0.3, True)
    assert d.naked and d.rPlus == 0.0 and math.isfinite(d.rIsco)
    d, M, a, q = bt.validate_and_derive(KN, float("nan"), float("nan"), -3.0)
    assert math.isfinite(M) and math.isfinite(a) and q >= 0 and math.isfinite(d.rIsco)
    P = bt.defaults()
    P.dPhi, P.maxSteps = 0.0, 0
    P.camPos.x = float("nan")
    P.exposure, P.model = -5.0, 99
    Q, fixed = bt.sanitized(P)
    assert fixed and Q.dPhi >= np.float32(1e-4) and Q.maxSteps >= 16 and all(
        math.isfinite(v) for v in (Q.camPos.x, Q.camPos.y, Q.camPos.z)) and Q.exposure > 0 and 0 <= Q.model <= 3


# ----------------------------------------------------------------- parity of host logic with the oracle
def _random_params(bt, rng):
    P = bt.defaults()
    P.model = int(rng.integers(-1, 5))
    P.M = float(rng.choice([0.5, 0.5, 0.0, -1.0, np.nan, 2.0]))
    P.aSpin = float(rng.choice([0.0, 0.45, -0.35, 3.0, np.nan, np.inf]))
    P.Qc = float(rng.choice([0.0, 0.2, -0.1, 0.6, np.nan]))
    P.rPlus = float(rng.choice([1.0, 0.7, 0.0, -1.0, np.nan]))
    P.allowNaked = int(rng.integers(0, 2))
    P.dPhi = float(rng.choice([0.012, 0.0, -1.0, 0.5, np.nan]))
    P.maxSteps = int(rng.choice([2000, 0, -5, 10 ** 7]))
    P.exposure = float(rng.choice([1.0, -5.0, 1e9, np.nan]))
    P.width, P.height = int(rng.choice([320, 0, -3])), int(rng.choice([180, 0]))
    P.tanHalfFov = float(rng.choice([0.5, 0.0, 99.0, np.nan]))
    P.aspect = float(rng.choice([1.77, 0.0, np.nan]))
    for v in (P.camPos, P.camForward):
        v.x, v.y, v.z = [float(x) for x in rng.choice([0.0, 1.0, -3.0, 20.0, np.nan], 3)]
    P.diskInner = float(rng.choice([3.0, 0.1, -1.0, np.nan, 5000.0]))
    P.diskOuter = float(rng.choice([8.0, 1.0, np.nan, 1e5]))
    P.sampleIndex = int(rng.choice([0, -7, 10 ** 6, 5]))
    P.accumMode = int(rng.choice([0, 1, 2, 99, -1]))
    P.bloomStrength = float(rng.choice([0.35, np.nan, 9.0, -1.0]))
    P.hotSpotsEnabled = int(rng.choice([0, 1, 7]))
    P.hotSpotStrength = float(rng.choice([0.65, 100.0, np.nan]))
    P.diskEmisScale = float(rng.choice([1.5, np.nan, 99.0]))
    P.diskAbsScale = float(rng.choice([2.2, -1.0, np.nan]))
    return P


def test_sanitize_matches_oracle(bt, oracle):
    rng = np.random.default_rng(11)
    for _ in range(3000):
        P = _random_params(bt, rng)
        a, fa = bt.sanitized(P)
        b, fb = oracle.sanitize(P)
        assert fa == fb
        assert bytes(a) == bytes(b)


@pytest.mark.parametrize("model", [SCHW, RN, KERR, KN])
def test_validate_derive_matches_oracle(bt, oracle, model):
    for a in [0.0, 0.3, -0.7, 0.9, 0.995, 1.2, -5.0]:
        for q in [0.0, 0.4, 0.6, 1.1, -1.0]:
            for naked in (False, True):
                d, M, a1, q1 = bt.validate_and_derive(model, 0.5, a, q, naked)
                o, M2, a2, q2 = oracle.validate_derive(model, 0.5, a, q, naked)
                got = [d.a, d.Q, d.rPlus, d.rMinus, d.rPhoton, d.rErgo, d.rIsco, d.clamped, d.naked]
                exp = [o[k] for k in ["a", "Q", "rPlus", "rMinus", "rPhoton", "rErgo", "rIsco", "clamped", "naked"]]
                assert np.array_equal(np.float32(got), np.float32(exp)), (model, a, q, naked)
                assert (M, a1, q1) == (M2, a2, q2)


# ----------------------------------------------------------------- FP64 roots
@pytest.mark.parametrize("a_star", [-0.99, -0.7, -0.3, 0.0, 0.3, 0.7, 0.9, 0.99])
def test_isco_fp64_kerr_satisfies_bpt_condition(bt, a_star):
    """Kerr ISCO (orbit along +phi, signed spin) is the root of
    r^2 - 6 M r + 8 a sqrt(M r) - 3 a^2 = 0 (Bardeen, Press & Teukolsky 1972)."""
    M, a = 0.5, 0.5 * a_star
    r = bt.load().bhIscoFP64(M, a, 0.0)
    assert abs(r * r - 6 * M * r + 8 * a * math.sqrt(M * r) - 3 * a * a) < 1e-12
    if a_star == 0:
        assert abs(r - 6 * M) < 1e-14


@pytest.mark.parametrize("q", [0.2, 0.4, 0.6, 0.8, 0.95])
def test_isco_fp64_reissner_nordstrom_cubic(bt, q):
    """RN ISCO: the largest root of M r^3 - 6 M^2 r^2 + 9 M Q^2 r - 4 Q^4 = 0."""
    M, Q = 0.5, 0.5 * q
    r = bt.load().bhIscoFP64(M, 0.0, Q)
    assert abs(M * r ** 3 - 6 * M * M * r * r + 9 * M * Q * Q * r - 4 * Q ** 4) < 1e-11


@pytest.mark.parametrize("a_star,q", [(-0.7, 0.4), (0.6, 0.5), (0.9, 0.3)])
def test_isco_fp64_kerr_newman_vs_scan(bt, a_star, q):
    M = 0.5
    r = bt.load().bhIscoFP64(M, a_star * M, q * M)
    d, *_ = bt.validate_and_derive(KN, M, a_star, q)
    assert abs(r - d.rIsco) < 0.002 * M   # the float scan steps 0.001 M
    # energy is minimal at r: compare with a dense FP64 scan of the circular-orbit energy
    a, Q = a_star * M, q * M
    rr = np.linspace(r * 0.98, r * 1.02, 4001)
    w = np.sqrt(M * rr - Q * Q)
    Om = w / (rr * rr + a * w)
    m2r = 2 * M * rr - Q * Q
    gtt, gtp, gpp = -(1 - m2r / rr ** 2), -a * m2r / rr ** 2, rr ** 2 + a * a + m2r * a * a / rr ** 2
    E = -(gtt + Om * gtp) / np.sqrt(-(gtt + 2 * Om * gtp + Om * Om * gpp))
    assert abs(rr[np.argmin(E)] - r) < 2 * (rr[1] - rr[0])


# ----------------------------------------------------------------- camera and presets
def test_fill_camera_orthonormal(bt):
    P = bt.defaults()
    P.width, P.height = 1920, 1080
    bt.fill_camera(P, 0.6, math.radians(5), 15.0, 50.0)
    pos = np.array([P.camPos.x, P.camPos.y, P.camPos.z])
    F = np.array([P.camForward.x, P.camForward.y, P.camForward.z])
    R = np.array([P.camRight.x, P.camRight.y, P.camRight.z])
    U = np.array([P.camUp.x, P.camUp.y, P.camUp.z])
    assert abs(np.linalg.norm(pos) - 15.0) < 1e-5
    assert np.allclose(F, -pos / np.linalg.norm(pos), atol=1e-6)
    for v in (F, R, U):
        assert abs(np.linalg.norm(v) - 1) < 1e-6
    assert abs(F @ R) < 1e-6 and abs(F @ U) < 1e-6 and abs(R @ U) < 1e-6
    assert U[1] > 0 and abs(R[1]) < 1e-7
    assert abs(P.tanHalfFov - math.tan(math.radians(25))) < 1e-6 and abs(P.aspect - 16 / 9) < 1e-6


def test_workload_presets(bt):
    P = bt.workload("schw_256_balanced")
    assert (P.width, P.height, P.model, P.bloomEnabled, P.hotSpotsEnabled) == (256, 256, 0, 0, 0)
    assert (P.diskInner, P.diskOuter) == (3.0, 12.0)
    for f, t in enumerate([0.0, 0.5, 1.0]):
        P = bt.workload("kerr09_1080p_high", f)
        assert (P.width, P.height, P.model, P.maxSteps) == (1920, 1080, 2, 3500) and P.diskTime == t
        assert abs(P.diskInner - 1.160442) < 2e-6 and P.diskOuter == 20.0
        assert abs(math.hypot(P.camPos.x, P.camPos.y, P.camPos.z) - 15.0) < 1e-5
    P = bt.workload("kn_m07_q04_2160p_high")
    assert (P.width, P.height, P.model) == (3840, 2160, 3) and abs(P.diskInner - 3.964624) < 2e-6
    assert abs(P.aSpin + 0.35) < 1e-7 and abs(P.Qc - 0.2) < 1e-7
    with pytest.raises(ValueError):
        bt.workload("nope")
    for name, fr in [("schw_256_balanced", 0), ("kerr09_1080p_high", 2), ("kn_m07_q04_2160p_high", 0)]:
        Q, fixed = bt.sanitized(bt.workload(name, fr))
        assert not fixed


def test_schedule_validation(bt):
    s0 = bt.get_schedule()
    lib = bt.load()
    for bad in [dict(wgSize=96), dict(stepsPerChunk=0), dict(raysPerRefill=65), dict(mathMode=2), dict(diskCoop=3)]:
        s = bt.BhSchedule(**s0.as_dict())
        for k, v in bad.items():
            setattr(s, k, v)
        assert lib.bhSetSchedule(C.byref(s)) != 0
    assert lib.bhSetSchedule(C.byref(s0)) == 0


def test_null_and_invalid_arguments_rejected_without_gpu(bt):
    lib = bt.load()
    P = bt.workload("schw_256_balanced")
    assert lib.launchRenderKernel(None, C.byref(P), None) != 0
    assert lib.launchRenderPipeline(None, None, None, None, C.byref(P), None) != 0
    assert lib.bhProbeGrid(None, C.byref(P), 4, 4, None) != 0
    assert lib.bhFramebufferCreate(0, 10, C.byref(C.c_void_p())) != 0
    # the post chain alone: output and accumulation buffers are required, and
    # the bloom scratch too when bloom is on (as in launchRenderPipeline)
    dummy = C.c_void_p(0x1000)
    assert lib.bhPostChain(None, dummy, dummy, dummy, C.byref(P), None) != 0
    assert lib.bhPostChain(dummy, None, dummy, dummy, C.byref(P), None) != 0
    Q = P.copy()
    Q.bloomEnabled = 1
    assert lib.bhPostChain(dummy, dummy, None, None, C.byref(Q), None) != 0
    assert lib.bhFramebufferPost(None, C.byref(P), None) != 0
