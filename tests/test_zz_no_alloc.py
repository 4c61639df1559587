"""launchRenderPipelineNoAlloc: the frame of launchRenderPipeline with every
library buffer taken from a static per-GPU array (no allocation).

Checked through the C ABI against launchRenderPipeline bit for bit (uchar4
frame and float4 accumulation) on two calls in a row, the second with a
smaller frame, and on a frame with the FP64 band pass off; frames whose
scratch would not fit the array are rejected before anything is enqueued.
"""
import ctypes as C

import numpy as np
import pytest

from helpers import scaled_workload

pytestmark = pytest.mark.gpu

HIP_ERROR_INVALID_VALUE = 1


def say(*a):
    print("[gpu-test]", *a, flush=True)


def buffers(bt, P):
    W, H = P.width, P.height
    nh = ((W + 1) // 2) * ((H + 1) // 2)
    b = [bt.DeviceBuffer(W * H * 4), bt.DeviceBuffer(W * H * 16), bt.DeviceBuffer(nh * 16),
         bt.DeviceBuffer(nh * 16)]
    b[1].fill(0)
    return b


def no_alloc_frame(bt, P):
    b = buffers(bt, P)
    err = bt.load().launchRenderPipelineNoAlloc(b[0].ptr, b[1].ptr, b[2].ptr, b[3].ptr, C.byref(P), None)
    bt.check(err, "launchRenderPipelineNoAlloc")
    bt.synchronize()
    W, H = P.width, P.height
    rgba = b[0].download(np.uint8).reshape(H, W, 4)
    acc = b[1].download(np.float32).reshape(H, W, 4)
    for x in b:
        x.close()
    return rgba, acc


def check_same(bt, P, what):
    say(what, P.width, "x", P.height)
    rgba, acc = no_alloc_frame(bt, P)
    ref, ref_acc = bt.render_pipeline(P)
    assert (rgba[..., 3] == 255).all()
    assert np.array_equal(rgba, ref), what
    assert np.array_equal(acc.view(np.uint32), ref_acc.view(np.uint32)), what


def test_no_alloc_equals_pipeline_two_calls(gpu):
    """1080p Kerr (the largest frame the array takes), then a smaller view."""
    check_same(gpu, gpu.workload("kerr09_1080p_high", 0), "kerr09 full frame")
    check_same(gpu, scaled_workload(gpu, "kn_m07_q04_2160p_high", 0, 640, 360), "kn reduced")


def test_no_alloc_band_off_and_spherical(gpu):
    p0 = gpu.get_precision()
    try:
        gpu.set_precision(band=0)
        check_same(gpu, scaled_workload(gpu, "kerr09_1080p_high", 1, 480, 270), "kerr09 band off")
    finally:
        gpu.restore_precision(p0)
    check_same(gpu, gpu.workload("schw_256_balanced"), "schwarzschild")


def test_no_alloc_rejects_frames_beyond_its_array(gpu):
    P = scaled_workload(gpu, "kerr09_1080p_high", 0, 2048, 1536)
    b = [C.c_void_p(16), C.c_void_p(16), C.c_void_p(16), C.c_void_p(16)]   # never dereferenced
    err = gpu.load().launchRenderPipelineNoAlloc(b[0], b[1], b[2], b[3], C.byref(P), None)
    assert err == HIP_ERROR_INVALID_VALUE
