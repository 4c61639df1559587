"""Viewer layer (host code of libbhtrace.so, no GPU): the caller-side
producer of RenderParams that follows the reference application loop
(main.cpp:48-228, 230-337, 418-478).

Pinned by the reference's stated defaults and knob constants, by scipy's
extrinsic Euler rotation for the body orientation, by orthonormality of the
camera basis, and by the FP64 ISCO root for the disk inner edge.
"""
import ctypes as C
import math

import numpy as np
import pytest
from scipy.spatial.transform import Rotation

import bhtrace as bt


@pytest.fixture
def viewer():
    v = bt.Viewer(640, 360)
    yield v
    v.close()


def vec(f3):
    return np.array([f3.x, f3.y, f3.z], np.float64)


def test_defaults_follow_reference(viewer):
    s = viewer.state
    # main.cpp:45-88
    assert (s.azimuth, s.elevation, s.distance) == (np.float32(0.6), np.float32(0.2), 16.0)
    assert (s.oriX, s.oriY, s.oriZ) == (0, 0, 0)
    assert s.model == 0 and s.aStar == np.float32(0.7) and s.q == np.float32(0.4) and s.allowNaked == 0
    assert s.exposure == np.float32(0.85) and s.diskOn and s.bloomOn and s.hotSpotsOn and not s.animPaused
    assert s.dPhi == np.float32(0.012) and s.maxSteps == 2000 and s.fovY == 60.0
    P, restarted = viewer.frame(0.016)
    assert restarted and P.accumMode == 0 and P.sampleIndex == 0
    assert (P.width, P.height) == (640, 360) and P.aspect == np.float32(640 / 360)
    assert P.tanHalfFov == pytest.approx(math.tan(math.radians(30)), rel=1e-6)
    assert P.M == 0.5 and P.diskOuter == 8.0 and P.hotSpotStrength == np.float32(0.65)
    assert P.bloomEnabled == 1 and P.diskEnabled == 1 and P.hotSpotsEnabled == 1
    Q, changed = bt.sanitized(P)
    assert not changed


def test_zero_orientation_equals_fill_camera():
    lib = bt.load()
    for az, el, dist, fov in [(0.6, 0.2, 16.0, 60.0), (-2.0, 1.4, 3.0, 50.0), (3.0, -0.7, 40.0, 70.0)]:
        a, b = bt.defaults(), bt.defaults()
        a.width = b.width = 1920
        a.height = b.height = 1080
        lib.bhFillCamera(C.byref(a), az, el, dist, fov)
        lib.bhFillCameraOriented(C.byref(b), az, el, dist, fov, 0.0, 0.0, 0.0)
        assert bytes(a) == bytes(b)


@pytest.mark.parametrize("ori", [(0.3, 0.0, 0.0), (0.0, -1.1, 0.0), (0.0, 0.0, 2.0), (0.4, -0.9, 1.7),
                                 (3.0, 2.5, -2.2)])
def test_body_orientation_is_extrinsic_xyz(ori):
    lib = bt.load()
    az, el, dist = 0.6, 0.25, 12.0
    body, world = bt.defaults(), bt.defaults()
    lib.bhFillCamera(C.byref(body), az, el, dist, 60.0)
    lib.bhFillCameraOriented(C.byref(world), az, el, dist, 60.0, *ori)
    # R = Rz Ry Rx about fixed axes == scipy's extrinsic "xyz" sequence
    R = Rotation.from_euler("xyz", ori).as_matrix()
    for name in ("camPos", "camForward", "camRight"):
        want = R @ vec(getattr(body, name))
        assert np.allclose(vec(getattr(world, name)), want, atol=3e-6 * max(1.0, np.linalg.norm(want))), name
    F, Rt, U = vec(world.camForward), vec(world.camRight), vec(world.camUp)
    B = np.stack([F, Rt, U])
    assert np.allclose(B @ B.T, np.eye(3), atol=2e-6)
    assert np.allclose(np.cross(Rt, F), U, atol=2e-6)
    pos = vec(world.camPos)
    assert np.allclose(F, -pos / np.linalg.norm(pos), atol=2e-6)


def test_clamps(viewer):
    viewer.orbit(0.0, 10000.0)
    assert viewer.state.elevation == np.float32(1.45)
    viewer.orbit(0.0, -100000.0)
    assert viewer.state.elevation == np.float32(-1.45)
    viewer.zoom(200.0)
    assert viewer.state.distance == np.float32(2.2)
    viewer.key("N")
    viewer.zoom(200.0)
    assert viewer.state.distance == np.float32(1.5)
    viewer.zoom(-500.0)
    assert viewer.state.distance == 55.0
    viewer.update(exposure=1e-4)
    assert viewer.state.exposure == np.float32(0.05)
    viewer.update(exposure=1e4)
    assert viewer.state.exposure == 20.0


def test_mouse_and_held_keys(viewer):
    s0 = viewer.state
    viewer.orbit(10.0, -4.0)
    s = viewer.state
    assert s.azimuth == pytest.approx(s0.azimuth + 0.05, abs=1e-6)
    assert s.elevation == pytest.approx(s0.elevation - 0.02, abs=1e-6)
    viewer.zoom(2.0)
    assert viewer.state.distance == pytest.approx(s0.distance * 0.92 ** 2, rel=1e-6)
    d0 = viewer.state.distance
    viewer.frame(0.05, held=["W"])
    assert viewer.state.distance == pytest.approx(d0 * 0.5 ** 0.05, rel=1e-6)
    viewer.frame(0.05, held=["X", "Z"])
    viewer.frame(0.05, held=["X", "SHIFT"])
    s = viewer.state
    assert s.oriX == pytest.approx(0.0, abs=1e-7) and s.oriZ == pytest.approx(1.1 * 0.05, rel=1e-6)
    viewer.frame(1.0, held=["A"])                      # dt is clamped to 0.1
    assert viewer.state.azimuth == pytest.approx(s.azimuth - 0.12, abs=1e-6)
    e0 = viewer.state.exposure
    viewer.frame(0.1, held=["="])
    assert viewer.state.exposure == pytest.approx(e0 * 2.5 ** 0.1, rel=1e-6)
    viewer.key("R")
    s = viewer.state
    assert (s.azimuth, s.elevation, s.distance, s.oriX, s.oriZ) == (np.float32(0.6), np.float32(0.2), 16.0, 0, 0)


def test_model_knobs(viewer):
    lib = bt.load()
    assert viewer.key(bt.KEY_F4) == 1
    P, _ = viewer.frame(0.0)
    assert P.model == 2 and P.aSpin == pytest.approx(0.7 * 0.5, rel=1e-6) and P.Qc == 0.0
    assert P.diskInner == np.float32(lib.bhIscoFP64(0.5, P.aSpin, 0.0))
    d = bt.validate_and_derive(2, 0.5, 0.7, 0.4)[0]
    # the reference's float scan steps 0.001 M
    assert abs(P.diskInner - d.rIsco) <= 1.1e-3 * 0.5 + 1e-6
    for _ in range(10):
        viewer.key("]", bt.KEY_REPEAT)
    P, _ = viewer.frame(0.0)
    assert viewer.state.aStar == np.float32(0.995) and P.aSpin == pytest.approx(0.995 * 0.5, rel=1e-6)
    for _ in range(60):
        viewer.key("[", bt.KEY_REPEAT)
    viewer.frame(0.0)
    assert viewer.state.aStar == np.float32(-0.995)
    viewer.key(bt.KEY_F5)
    for _ in range(30):
        viewer.key(",", bt.KEY_REPEAT)
    P, _ = viewer.frame(0.0)
    assert viewer.state.q == 0.0 and P.model == 3
    viewer.key(bt.KEY_F3)
    viewer.update(q=0.5)
    P, _ = viewer.frame(0.0)
    assert P.model == 1 and P.aSpin == 0.0 and P.Qc == pytest.approx(0.25, rel=1e-6)
    # RN ISCO from the FP64 root
    assert P.diskInner == np.float32(lib.bhIscoFP64(0.5, 0.0, P.Qc))
    # presses that are not repeatable ignore repeats; presets
    assert viewer.key("3", bt.KEY_REPEAT) == 0
    viewer.key("3")
    P, _ = viewer.frame(0.0)
    assert P.dPhi == np.float32(0.007) and P.maxSteps == 3500
    viewer.key("1")
    P, _ = viewer.frame(0.0)
    assert P.dPhi == np.float32(0.02) and P.maxSteps == 1200


def test_naked_mode_knob_range(viewer):
    viewer.key(bt.KEY_F4)
    viewer.key("N")
    for _ in range(40):
        viewer.key("]", bt.KEY_REPEAT)
    P, _ = viewer.frame(0.0)
    assert viewer.state.aStar == np.float32(1.5) and P.allowNaked == 1
    assert P.rPlus == 0.0
    Q, changed = bt.sanitized(P)
    assert not changed


def test_accumulation_rules(viewer):
    P, r = viewer.frame(0.016)
    assert r and P.sampleIndex == 0 and P.accumMode == 0
    P, r = viewer.frame(0.016)
    # the disk animates: EMA blend, index advances
    assert not r and P.sampleIndex == 1 and P.accumMode == 2
    t = viewer.state.diskTime
    viewer.key(" ")                                     # pause the disk clock
    P, r = viewer.frame(0.016)
    assert not r and P.accumMode == 1 and P.sampleIndex == 2 and viewer.state.diskTime == t
    viewer.update(exposure=2.0)                         # post-only change keeps accumulating
    viewer.key("B")
    P, r = viewer.frame(0.016)
    assert not r and P.sampleIndex == 3 and P.bloomEnabled == 0
    viewer.key("H")                                     # hot spots change the geometry
    P, r = viewer.frame(0.016)
    assert r and P.sampleIndex == 0 and P.accumMode == 0 and P.hotSpotsEnabled == 0
    viewer.key(" ")
    viewer.key(bt.KEY_F1)                               # disk off: running mean
    P, r = viewer.frame(0.016)
    assert r and P.diskEnabled == 0
    P, r = viewer.frame(0.016)
    assert not r and P.accumMode == 1
    viewer.orbit(1.0, 0.0)
    P, r = viewer.frame(0.016)
    assert r and P.sampleIndex == 0
    viewer.resize(320, 200)
    P, r = viewer.frame(0.016)
    assert r and (P.width, P.height) == (320, 200)
    for _ in range(5000):
        P, r = viewer.frame(0.0)
    assert P.sampleIndex == 4096 and viewer.state.sampleIndex == 4096


def test_disk_clock(viewer):
    viewer.frame(0.05)
    viewer.frame(0.5)                                   # clamped to 0.1
    P, _ = viewer.frame(0.02)
    assert P.diskTime == pytest.approx(0.17, abs=1e-6)


def test_bad_arguments():
    lib = bt.load()
    h = C.c_void_p()
    assert lib.bhViewerCreate(0, 10, C.byref(h)) == -1
    assert lib.bhViewerFrame(None, 0.0, 0, None) == -1
