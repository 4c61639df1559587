"""Exactly radial camera rays (trace.cuh:447-459, 500-505).

For the spherical models the orbital-plane basis of a radial ray is
degenerate and the reference's radialRay takes over: an inbound ray is
captured and an outbound one escapes along its own direction, in both cases
with no disk sampling and no integration steps -- even when the camera sits
inside the disk's layer.  The Kerr models have no such branch: their radial
rays go through the Hamiltonian integrator like any other.  Every case is
traced through bhProbeDirs in both math modes and compared with the oracle.
"""
import numpy as np
import pytest

from helpers import angle_between, make_params

pytestmark = pytest.mark.gpu

SCHW, RN, KERR, KN = 0, 1, 2, 3
MODELS = [("schw", SCHW, 0.0, 0.0), ("rn_q0.5", RN, 0.0, 0.5), ("kerr_a0.9", KERR, 0.9, 0.0),
          ("kn_a-0.7_q0.4", KN, -0.7, 0.4)]
CAMERAS = [(15.0, 0.0, 0.0), (0.0, 0.0, -12.0), (8.0, 8.0, 0.0), (0.0, 9.0, 0.0), (-6.0, 0.5, 6.0)]


def say(*a):
    print("[gpu-test]", *a, flush=True)


@pytest.fixture(autouse=True)
def default_schedule(gpu):
    s0 = gpu.get_schedule()
    yield
    gpu.restore_schedule(s0)


def params(gpu, model, a_star, q, cam):
    P = make_params(gpu, model, a_star, q, 0.012, 2000)
    P.diskEnabled, P.hotSpotsEnabled = 1, 1
    P.diskInner = float(gpu.load().bhIscoFP64(0.5, float(P.aSpin), float(P.Qc)))
    P.diskOuter = 20.0
    P.diskTime = 0.4
    P.camPos.x, P.camPos.y, P.camPos.z = cam
    return P


@pytest.mark.parametrize("math", [0, 1])
@pytest.mark.parametrize("name,model,a_star,q", MODELS, ids=[m[0] for m in MODELS])
def test_radial_rays(gpu, oracle, math, name, model, a_star, q):
    gpu.set_schedule(math=math)
    for cam in CAMERAS:
        P = params(gpu, model, a_star, q, cam)
        o = np.array(cam, np.float32)
        inward = (-o / np.float32(np.linalg.norm(o))).astype(np.float32)
        dirs = np.stack([inward, -inward]).astype(np.float32)
        got = gpu.probe_dirs(P, dirs)
        ref = oracle.probe_dirs(P, dirs, mode=0)
        say(name, "math", math, "camera", cam, "outcomes", got["outcome"].tolist(), "steps", got["steps"].tolist())
        assert got["outcome"].tolist() == ref[:, 0].astype(int).tolist()
        if model in (SCHW, RN):
            # radialRay: inbound captured, outbound escapes along itself, no
            # steps and no disk light
            assert got["outcome"].tolist() == [0, 1]
            assert got["steps"].tolist() == [0, 0]
            assert np.all(got["lum"] == 0.0) and np.all(got["trans"] == 1.0)
            assert np.array_equal(got["escDir"][1], dirs[1])
            assert np.array_equal(ref[:, 6], [0, 0]) and np.all(ref[:, 5] == 0.0)
        else:
            # the integrator: outbound escapes, inbound falls in (the camera
            # is outside the horizon); exit direction and disk light as the
            # oracle's
            assert got["outcome"].tolist() == [0, 1]
            ang = angle_between(got["escDir"][1:], ref[1:, 1:4])
            assert ang.max() < 1e-4, ang
            assert np.allclose(got["lum"], ref[:, 5], rtol=1e-2, atol=1e-4)
            assert np.allclose(got["trans"], ref[:, 7], rtol=1e-2, atol=1e-4)
