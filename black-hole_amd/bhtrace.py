"""ctypes binding of libbhtrace.so, the gfx950 geodesic tracer's C ABI.

Python is plumbing here: buffers are device allocations made through the
library, every kernel runs through the C entry points declared in
include/bh/bhtrace.h.  Loading fails loudly when the library is missing.

    import bhtrace as bt
    lib = bt.load()
    P = bt.workload("kerr09_1080p_high", 0)
    fb = bt.Framebuffer(P.width, P.height)
    fb.render(P); rgba, hdr = fb.readback()
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "lib", "libbhtrace.so")


class F3(C.Structure):
    _fields_ = [("x", C.c_float), ("y", C.c_float), ("z", C.c_float)]


class RenderParams(C.Structure):
    """Mirror of RenderParams (include/bh/render_params.h; reference render_params.h:21-87)."""

    _fields_ = [
        ("width", C.c_int), ("height", C.c_int), ("swapRB", C.c_int),
        ("camPos", F3), ("camForward", F3), ("camRight", F3), ("camUp", F3),
        ("tanHalfFov", C.c_float), ("aspect", C.c_float),
        ("model", C.c_int), ("M", C.c_float), ("aSpin", C.c_float), ("Qc", C.c_float),
        ("rPlus", C.c_float), ("rPhoton", C.c_float), ("rErgo", C.c_float), ("allowNaked", C.c_int),
        ("dPhi", C.c_float), ("maxSteps", C.c_int),
        ("diskEnabled", C.c_int), ("diskInner", C.c_float), ("diskOuter", C.c_float), ("diskTime", C.c_float),
        ("diskEmisScale", C.c_float), ("diskAbsScale", C.c_float),
        ("hotSpotsEnabled", C.c_int), ("hotSpotStrength", C.c_float),
        ("exposure", C.c_float), ("sampleIndex", C.c_int), ("accumMode", C.c_int),
        ("bloomEnabled", C.c_int), ("bloomStrength", C.c_float),
    ]

    def copy(self) -> "RenderParams":
        c = RenderParams()
        C.memmove(C.byref(c), C.byref(self), C.sizeof(RenderParams))
        return c


assert C.sizeof(RenderParams) == 160


class BHDerived(C.Structure):
    _fields_ = [("a", C.c_float), ("Q", C.c_float), ("rPlus", C.c_float), ("rMinus", C.c_float),
                ("rPhoton", C.c_float), ("rErgo", C.c_float), ("rIsco", C.c_float),
                ("clamped", C.c_int), ("naked", C.c_int)]


class BhProbe(C.Structure):
    _fields_ = [("outcome", C.c_int), ("escDir", C.c_float * 3), ("maxHviol", C.c_float),
                ("lum", C.c_float), ("steps", C.c_int), ("trans", C.c_float)]


class BhSchedule(C.Structure):
    _fields_ = [("wgSize", C.c_int), ("stepsPerChunk", C.c_int), ("raysPerRefill", C.c_int),
                ("mathMode", C.c_int), ("diskCoop", C.c_int)]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


class BhPrecision(C.Structure):
    _fields_ = [("fp64Band", C.c_int), ("critBand", C.c_float), ("polarBand", C.c_float)]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


class BhViewState(C.Structure):
    _fields_ = [("azimuth", C.c_float), ("elevation", C.c_float), ("distance", C.c_float),
                ("oriX", C.c_float), ("oriY", C.c_float), ("oriZ", C.c_float),
                ("model", C.c_int), ("aStar", C.c_float), ("q", C.c_float), ("allowNaked", C.c_int),
                ("exposure", C.c_float), ("diskOn", C.c_int), ("bloomOn", C.c_int), ("hotSpotsOn", C.c_int),
                ("animPaused", C.c_int), ("dPhi", C.c_float), ("maxSteps", C.c_int), ("fovY", C.c_float),
                ("diskTime", C.c_float), ("sampleIndex", C.c_int)]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


KEY_PRESS, KEY_REPEAT = 1, 2
KEY_F1, KEY_F2, KEY_F3, KEY_F4, KEY_F5 = 0x101, 0x102, 0x103, 0x104, 0x105
HELD = {"W": 1 << 0, "S": 1 << 1, "A": 1 << 2, "D": 1 << 3, "Q": 1 << 4, "E": 1 << 5, "-": 1 << 6, "=": 1 << 7,
        "X": 1 << 8, "Y": 1 << 9, "Z": 1 << 10, "SHIFT": 1 << 11}


class BhDiffStats(C.Structure):
    _fields_ = [("pixels", C.c_longlong), ("maxDiff", C.c_int * 3), ("pad", C.c_int),
                ("hist", C.c_longlong * 8), ("fracWithin1", C.c_double)]


class BhAngleStats(C.Structure):
    _fields_ = [("bothEscaped", C.c_longlong), ("outcomeMismatch", C.c_longlong), ("over1mrad", C.c_longlong),
                ("median", C.c_double), ("p999", C.c_double), ("max", C.c_double)]


MLP_TARGET_DTYPE = np.dtype([("d0", "<f4"), ("d1", "<f4"), ("disk", "<f4"), ("hitR", "<f4"), ("escaped", "<f4"),
                             ("band", "<f4"), ("pad", "<f4", (2,))])

PROBE_DTYPE = np.dtype([("outcome", "<i4"), ("escDir", "<f4", (3,)), ("maxHviol", "<f4"),
                        ("lum", "<f4"), ("steps", "<i4"), ("trans", "<f4")])
RECORD_DTYPE = np.dtype([("escDir", "<f4", (3,)), ("flags", "<i4"), ("accum", "<f4", (3,)),
                         ("trans", "<f4")])

_lib = None

# (name, restype, argtypes) of every exported symbol declared in bh/bhtrace.h
_P = C.POINTER(RenderParams)
_vp = C.c_void_p
SIGNATURES = [
    ("launchRenderPipeline", C.c_int, [_vp, _vp, _vp, _vp, _P, _vp]),
    ("launchRenderKernel", C.c_int, [_vp, _P, _vp]),
    ("launchRenderPipelineNoAlloc", C.c_int, [_vp, _vp, _vp, _vp, _P, _vp]),
    ("bhDefaultRenderParams", None, [_P]),
    ("bhSanitizeRenderParams", C.c_int, [_P]),
    ("bhValidateAndDerive", None, [C.c_int, C.POINTER(C.c_float), C.POINTER(C.c_float),
                                   C.POINTER(C.c_float), C.c_int, C.POINTER(BHDerived)]),
    ("bhHorizonOuterFP64", C.c_double, [C.c_double, C.c_double, C.c_double]),
    ("bhIscoFP64", C.c_double, [C.c_double, C.c_double, C.c_double]),
    ("bhFillCamera", None, [_P, C.c_float, C.c_float, C.c_float, C.c_float]),
    ("bhFillCameraOriented", None, [_P, C.c_float, C.c_float, C.c_float, C.c_float, C.c_float, C.c_float,
                                     C.c_float]),
    ("bhViewStateDefaults", None, [C.POINTER(BhViewState)]),
    ("bhViewerCreate", C.c_int, [C.c_int, C.c_int, C.POINTER(C.c_void_p)]),
    ("bhViewerDestroy", None, [_vp]),
    ("bhViewerGetState", C.c_int, [_vp, C.POINTER(BhViewState)]),
    ("bhViewerSetState", C.c_int, [_vp, C.POINTER(BhViewState)]),
    ("bhViewerResize", C.c_int, [_vp, C.c_int, C.c_int, C.c_int]),
    ("bhViewerReset", None, [_vp]),
    ("bhViewerOrbit", None, [_vp, C.c_float, C.c_float]),
    ("bhViewerZoom", None, [_vp, C.c_float]),
    ("bhViewerKey", C.c_int, [_vp, C.c_int, C.c_int]),
    ("bhViewerFrame", C.c_int, [_vp, C.c_float, C.c_uint, _P]),
    ("bhWorkloadParams", C.c_int, [C.c_char_p, C.c_int, _P]),
    ("bhWorkloadViews", C.c_int, [C.c_char_p, C.c_uint64, C.c_int, _P]),
    ("bhRenderBatch", C.c_int, [C.c_int, _P, _vp, _vp, _vp, _vp, _vp]),
    ("bhRenderBatchPersistent", C.c_int, [C.c_int, _P, _vp, _vp, _vp, _vp, _vp]),
    ("bhBatchViewLatency", C.c_int, [C.POINTER(C.c_float), C.c_int, _vp]),
    ("bhBatchGraphCreate", C.c_int, [C.c_int, _P, _vp, _vp, _vp, _vp, C.c_int, C.POINTER(C.c_void_p)]),
    ("bhBatchGraphLaunch", C.c_int, [_vp, _vp]),
    ("bhBatchGraphInfo", C.c_int, [_vp, C.POINTER(C.c_int), C.POINTER(C.c_float)]),
    ("bhBatchGraphLatency", C.c_int, [_vp, C.POINTER(C.c_float), C.c_int, _vp]),
    ("bhBatchGraphDestroy", None, [_vp]),
    ("bhProbeGrid", C.c_int, [_vp, _P, C.c_int, C.c_int, _vp]),
    ("bhProbeDirs", C.c_int, [_vp, _vp, C.c_int, _P, _vp]),
    ("bhTraceRecords", C.c_int, [_vp, _P, _vp]),
    ("bhMathProbe", C.c_int, [C.c_int, _vp, C.c_int, _vp, _vp]),
    ("bhFramebufferCreate", C.c_int, [C.c_int, C.c_int, C.POINTER(C.c_void_p)]),
    ("bhFramebufferDestroy", None, [_vp]),
    ("bhFramebufferColor", _vp, [_vp]),
    ("bhFramebufferHDR", _vp, [_vp]),
    ("bhFramebufferRender", C.c_int, [_vp, _P, _vp]),
    ("bhFramebufferReadback", C.c_int, [_vp, _vp, _vp, _vp]),
    ("bhPostChain", C.c_int, [_vp, _vp, _vp, _vp, _P, _vp]),
    ("bhFramebufferPost", C.c_int, [_vp, _P, _vp]),
    ("bhGetSchedule", None, [C.POINTER(BhSchedule)]),
    ("bhSetSchedule", C.c_int, [C.POINTER(BhSchedule)]),
    ("bhAutotune", C.c_int, [_P, C.c_int, C.POINTER(BhSchedule), C.POINTER(C.c_float), _vp]),
    ("bhAutotuneClear", None, []),
    ("bhScheduleFor", None, [_P, C.POINTER(BhSchedule)]),
    ("bhAccumCreate", C.c_int, [C.c_int, C.c_int, C.POINTER(C.c_void_p)]),
    ("bhAccumDestroy", None, [_vp]),
    ("bhAccumColor", _vp, [_vp]),
    ("bhAccumHDR", _vp, [_vp]),
    ("bhAccumReset", C.c_int, [_vp, _vp]),
    ("bhAccumStep", C.c_int, [_vp, _P, _vp]),
    ("bhAccumCapture", C.c_int, [_vp, _P, _vp]),
    ("bhAccumReplay", C.c_int, [_vp, C.c_int, _vp]),
    ("bhAccumSampleIndex", C.c_int, [_vp, C.POINTER(C.c_int), _vp]),
    ("bhGetPrecision", None, [C.POINTER(BhPrecision)]),
    ("bhSetPrecision", C.c_int, [C.POINTER(BhPrecision)]),
    ("bhLastBandStats", C.c_int, [C.POINTER(C.c_uint), C.POINTER(C.c_float), _vp]),
    ("bhLastTileOrder", C.c_int, [_vp, _vp, C.c_int, C.POINTER(C.c_int), _vp]),
    ("bhLastTraceKernelMs", C.c_int, [C.POINTER(C.c_float), C.POINTER(C.c_float), _vp]),
    ("bhEnableKernelTimers", None, [C.c_int]),
    ("bhLastKernelTimes", C.c_int, [C.POINTER(C.c_float), C.POINTER(C.c_float), C.POINTER(C.c_float)]),
    ("bhSurrogateParamCount", C.c_int, []),
    ("bhSurrogateCreate", C.c_int, [C.c_uint64, C.POINTER(C.c_void_p)]),
    ("bhSurrogateDestroy", None, [_vp]),
    ("bhSurrogateSetParams", C.c_int, [_vp, _vp, C.c_int]),
    ("bhSurrogateGetParams", C.c_int, [_vp, _vp, C.c_int]),
    ("bhSurrogateConfigs", C.c_int, [C.c_uint64, C.c_int, _P]),
    ("bhSurrogateDataset", C.c_int, [_vp, C.c_int, C.c_int, C.c_uint64, _vp]),
    ("bhSurrogateDatasetFor", C.c_int, [_vp, _P, C.c_int, C.c_int, C.c_uint64, _vp]),
    ("bhSurrogateDatasetView", C.c_int, [_vp, C.POINTER(C.c_void_p), C.POINTER(C.c_void_p), C.POINTER(C.c_void_p),
                                         C.POINTER(C.c_int)]),
    ("bhSurrogateTrain", C.c_int, [_vp, C.c_int, C.c_int, C.c_float, _vp, _vp]),
    ("bhSurrogateForward", C.c_int, [_vp, _vp, C.c_int, _vp, _vp]),
    ("bhSurrogateForwardFused", C.c_int, [_vp, _vp, C.c_int, _vp, _vp]),
    ("bhSurrogateLossGrad", C.c_int, [_vp, _vp, _vp, C.c_int, _vp, _vp, _vp]),
    ("bhSetSurrogate", C.c_int, [_vp, C.c_float, C.c_float]),
    ("bhLastSurrogateStats", C.c_int, [C.POINTER(C.c_uint), C.POINTER(C.c_uint), _vp]),
    ("bhLastSurrogateKernelMs", C.c_int, [C.POINTER(C.c_float), _vp]),
    ("bhMlpGemmTest", C.c_int, [_vp, C.c_int, C.c_int, _vp, C.c_int, C.c_int, _vp, C.c_int, C.c_int, C.c_int,
                                C.c_int, C.c_int, _vp]),
    ("bhWritePPM", C.c_int, [C.c_char_p, _vp, C.c_int, C.c_int]),
    ("bhWritePNG", C.c_int, [C.c_char_p, _vp, C.c_int, C.c_int]),
    ("bhWritePFM", C.c_int, [C.c_char_p, _vp, C.c_int, C.c_int]),
    ("bhImageDiff", C.c_int, [_vp, _vp, C.c_int, C.c_int, C.POINTER(BhDiffStats)]),
    ("bhExitAngleDiff", C.c_int, [_vp, _vp, C.c_int, _vp, C.POINTER(BhAngleStats)]),
    ("bhDeviceAlloc", C.c_int, [C.POINTER(C.c_void_p), C.c_size_t]),
    ("bhDeviceFree", C.c_int, [_vp]),
    ("bhCopyToDevice", C.c_int, [_vp, _vp, C.c_size_t]),
    ("bhCopyToHost", C.c_int, [_vp, _vp, C.c_size_t]),
    ("bhDeviceFill", C.c_int, [_vp, C.c_ubyte, C.c_size_t]),
    ("bhCopyDevice", C.c_int, [_vp, _vp, C.c_size_t]),
    ("bhDeviceSynchronize", C.c_int, []),
    ("bhStreamCreate", C.c_int, [C.POINTER(C.c_void_p)]),
    ("bhStreamRelease", C.c_int, [_vp]),
    ("bhStreamDestroy", C.c_int, [_vp]),
    ("bhStreamSynchronize", C.c_int, [_vp]),
    ("bhErrorString", C.c_char_p, [C.c_int]),
    ("bhDeviceCount", C.c_int, []),
]


class BhError(RuntimeError):
    pass


def load(path: str | None = None):
    """Load libbhtrace.so (raises if it has not been built)."""
    global _lib
    if _lib is not None:
        return _lib
    p = path or LIB_PATH
    if not os.path.exists(p):
        raise BhError(f"libbhtrace.so not built: {p} missing (run __graft_entry__.build())")
    lib = C.CDLL(p)
    for name, res, args in SIGNATURES:
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    _lib = lib
    return lib


def check(err: int, what: str = "") -> None:
    if err != 0:
        msg = load().bhErrorString(err)
        raise BhError(f"{what}: HIP error {err} ({msg.decode() if msg else '?'})")


# ------------------------------------------------------------------ parameters
def workload(name: str, frame: int = 0) -> RenderParams:
    P = RenderParams()
    if load().bhWorkloadParams(name.encode(), frame, C.byref(P)) != 0:
        raise ValueError(f"unknown workload {name!r} frame {frame}")
    return P


def workload_views(name: str = "uneven_views", seed: int = 1, n: int = 64):
    arr = (RenderParams * n)()
    if load().bhWorkloadViews(name.encode(), seed, n, arr) != n:
        raise ValueError(f"unknown view batch {name!r}")
    return arr


def defaults() -> RenderParams:
    P = RenderParams()
    load().bhDefaultRenderParams(C.byref(P))
    return P


def sanitized(P: RenderParams) -> tuple[RenderParams, bool]:
    Q = P.copy()
    fixed = load().bhSanitizeRenderParams(C.byref(Q))
    return Q, bool(fixed)


def validate_and_derive(model: int, M: float, a_star: float, q: float, allow_naked: bool = False):
    m, a, qq = C.c_float(M), C.c_float(a_star), C.c_float(q)
    d = BHDerived()
    load().bhValidateAndDerive(model, C.byref(m), C.byref(a), C.byref(qq), int(allow_naked), C.byref(d))
    return d, m.value, a.value, qq.value


def fill_camera(P: RenderParams, az: float, el: float, dist: float, fov_deg: float) -> None:
    load().bhFillCamera(C.byref(P), az, el, dist, fov_deg)


# ------------------------------------------------------------------ device memory
class DeviceBuffer:
    """A device allocation owned by Python (freed on close/GC)."""

    def __init__(self, nbytes: int):
        self.nbytes = int(nbytes)
        p = C.c_void_p()
        check(load().bhDeviceAlloc(C.byref(p), max(self.nbytes, 1)), "bhDeviceAlloc")
        self.ptr = p

    def upload(self, arr: np.ndarray) -> None:
        a = np.ascontiguousarray(arr)
        assert a.nbytes <= self.nbytes
        check(load().bhCopyToDevice(self.ptr, a.ctypes.data_as(C.c_void_p), a.nbytes), "upload")

    def download(self, dtype, count=None) -> np.ndarray:
        dt = np.dtype(dtype)
        n = self.nbytes // dt.itemsize if count is None else count
        out = np.empty(n, dtype=dt)
        check(load().bhCopyToHost(out.ctypes.data_as(C.c_void_p), self.ptr, out.nbytes), "download")
        return out

    def fill(self, value: int) -> None:
        check(load().bhDeviceFill(self.ptr, value, self.nbytes), "fill")

    def close(self) -> None:
        if self.ptr:
            load().bhDeviceFree(self.ptr)
            self.ptr = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def synchronize() -> None:
    check(load().bhDeviceSynchronize(), "bhDeviceSynchronize")


class Framebuffer:
    """Headless device framebuffer (uchar4 colour + float4 HDR accumulation)."""

    def __init__(self, w: int, h: int):
        self.w, self.h = w, h
        fb = C.c_void_p()
        check(load().bhFramebufferCreate(w, h, C.byref(fb)), "bhFramebufferCreate")
        self.fb = fb

    def render(self, P: RenderParams, stream=None) -> None:
        check(load().bhFramebufferRender(self.fb, C.byref(P), stream), "bhFramebufferRender")

    def post(self, P: RenderParams, stream=None) -> None:
        """Bloom + tonemap of the framebuffer's HDR image only (no trace)."""
        check(load().bhFramebufferPost(self.fb, C.byref(P), stream), "bhFramebufferPost")

    def hdr_ptr(self) -> C.c_void_p:
        return C.c_void_p(load().bhFramebufferHDR(self.fb))

    def readback(self, stream=None):
        rgba = np.empty((self.h, self.w, 4), np.uint8)
        hdr = np.empty((self.h, self.w, 4), np.float32)
        check(load().bhFramebufferReadback(self.fb, rgba.ctypes.data_as(C.c_void_p),
                                           hdr.ctypes.data_as(C.c_void_p), stream), "readback")
        return rgba, hdr

    def close(self) -> None:
        if self.fb:
            load().bhFramebufferDestroy(self.fb)
            self.fb = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


# ------------------------------------------------------------------ convenience wrappers
def render_pipeline(P: RenderParams, accum_init: np.ndarray | None = None, frames: int = 1):
    """Run launchRenderPipeline `frames` times on fresh device buffers; returns
    (rgba uint8 [H,W,4], accum float32 [H,W,4])."""
    lib = load()
    W, H = P.width, P.height
    bw, bh = (W + 1) // 2, (H + 1) // 2
    out = DeviceBuffer(W * H * 4)
    acc = DeviceBuffer(W * H * 16)
    bA = DeviceBuffer(bw * bh * 16)
    bB = DeviceBuffer(bw * bh * 16)
    if accum_init is None:
        acc.fill(0)
    else:
        acc.upload(np.ascontiguousarray(accum_init, np.float32))
    for _ in range(frames):
        check(lib.launchRenderPipeline(out.ptr, acc.ptr, bA.ptr, bB.ptr, C.byref(P), None), "launchRenderPipeline")
    synchronize()
    rgba = out.download(np.uint8).reshape(H, W, 4)
    accum = acc.download(np.float32).reshape(H, W, 4)
    return rgba, accum


def render_sequence(params) -> tuple[np.ndarray, np.ndarray]:
    """launchRenderPipeline over a sequence of frames of one size sharing one
    accumulation buffer (initially zero); returns the last frame's
    (rgba uint8 [H,W,4], accum float32 [H,W,4])."""
    lib = load()
    params = list(params)
    W, H = params[0].width, params[0].height
    if any((P.width, P.height) != (W, H) for P in params):
        raise ValueError("all frames of a sequence must have one size")
    bw, bh = (W + 1) // 2, (H + 1) // 2
    out, acc = DeviceBuffer(W * H * 4), DeviceBuffer(W * H * 16)
    bA, bB = DeviceBuffer(bw * bh * 16), DeviceBuffer(bw * bh * 16)
    acc.fill(0)
    for P in params:
        check(lib.launchRenderPipeline(out.ptr, acc.ptr, bA.ptr, bB.ptr, C.byref(P), None), "launchRenderPipeline")
    synchronize()
    return out.download(np.uint8).reshape(H, W, 4), acc.download(np.float32).reshape(H, W, 4)


def render_kernel(P: RenderParams) -> np.ndarray:
    lib = load()
    W, H = max(P.width, 1), max(P.height, 1)
    Q, _ = sanitized(P)
    out = DeviceBuffer(Q.width * Q.height * 4)
    check(lib.launchRenderKernel(out.ptr, C.byref(P), None), "launchRenderKernel")
    synchronize()
    return out.download(np.uint8).reshape(Q.height, Q.width, 4)


def trace_records(P: RenderParams, buf: "DeviceBuffer | None" = None) -> np.ndarray:
    """Records of one trace of P (into `buf` when given, else a new buffer)."""
    lib = load()
    Q, _ = sanitized(P)
    if buf is None:
        buf = DeviceBuffer(Q.width * Q.height * RECORD_DTYPE.itemsize)
    check(lib.bhTraceRecords(buf.ptr, C.byref(P), None), "bhTraceRecords")
    synchronize()
    return buf.download(RECORD_DTYPE).reshape(Q.height, Q.width)


# ------------------------------------------------------------------ viewer
class Viewer:
    """Caller-side producer of RenderParams (orbit camera, body orientation,
    model knobs, disk clock, accumulation restarts); wraps bhViewer*."""

    def __init__(self, width: int = 1280, height: int = 720):
        self.lib = load()
        h = C.c_void_p()
        if self.lib.bhViewerCreate(width, height, C.byref(h)) != 0:
            raise BhError("bhViewerCreate failed")
        self.h = h

    def close(self):
        if self.h:
            self.lib.bhViewerDestroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    @property
    def state(self) -> BhViewState:
        s = BhViewState()
        self.lib.bhViewerGetState(self.h, C.byref(s))
        return s

    def set_state(self, s: BhViewState) -> None:
        if self.lib.bhViewerSetState(self.h, C.byref(s)) != 0:
            raise BhError("bhViewerSetState failed")

    def update(self, **kw) -> None:
        s = self.state
        for k, v in kw.items():
            setattr(s, k, v)
        self.set_state(s)

    def resize(self, width: int, height: int, swap_rb: bool = False) -> None:
        if self.lib.bhViewerResize(self.h, width, height, int(swap_rb)) != 0:
            raise BhError("bhViewerResize failed")

    def reset(self) -> None:
        self.lib.bhViewerReset(self.h)

    def orbit(self, dx: float, dy: float) -> None:
        self.lib.bhViewerOrbit(self.h, dx, dy)

    def zoom(self, steps: float) -> None:
        self.lib.bhViewerZoom(self.h, steps)

    def key(self, key, action: int = KEY_PRESS) -> int:
        k = ord(key) if isinstance(key, str) else int(key)
        return self.lib.bhViewerKey(self.h, k, action)

    def frame(self, dt: float, held=()):
        """Advance by dt with the named held keys; returns (RenderParams, restarted)."""
        mask = 0
        for k in held:
            mask |= HELD[k]
        P = RenderParams()
        r = self.lib.bhViewerFrame(self.h, dt, mask, C.byref(P))
        if r < 0:
            raise BhError("bhViewerFrame failed")
        return P, bool(r)


# ------------------------------------------------------------ headless output
def _u8_image(rgba: np.ndarray) -> np.ndarray:
    a = np.ascontiguousarray(rgba, np.uint8)
    if a.ndim != 3 or a.shape[2] != 4:
        raise ValueError("expected an RGBA8 image of shape (H, W, 4)")
    return a


def write_image(path: str, rgba: np.ndarray) -> None:
    """Write an (H, W, 4) uint8 frame as PPM (.ppm) or PNG (.png)."""
    a = _u8_image(rgba)
    fn = load().bhWritePNG if path.lower().endswith(".png") else load().bhWritePPM
    if fn(path.encode(), a.ctypes.data, a.shape[1], a.shape[0]) != 0:
        raise BhError(f"writing {path} failed")


def write_pfm(path: str, hdr: np.ndarray) -> None:
    """Write an (H, W, 4) float32 HDR buffer as PFM (RGB, little-endian)."""
    a = np.ascontiguousarray(hdr, np.float32)
    if a.ndim != 3 or a.shape[2] != 4:
        raise ValueError("expected a float4 image of shape (H, W, 4)")
    if load().bhWritePFM(path.encode(), a.ctypes.data, a.shape[1], a.shape[0]) != 0:
        raise BhError(f"writing {path} failed")


def image_diff(a: np.ndarray, b: np.ndarray) -> BhDiffStats:
    """Per-channel LSB statistics of two RGBA8 frames of one size."""
    a, b = _u8_image(a), _u8_image(b)
    if a.shape != b.shape:
        raise ValueError(f"image sizes differ: {a.shape} vs {b.shape}")
    st = BhDiffStats()
    if load().bhImageDiff(a.ctypes.data, b.ctypes.data, a.shape[1], a.shape[0], C.byref(st)) != 0:
        raise BhError("bhImageDiff failed")
    return st


def exit_angle_diff(a: np.ndarray, b: np.ndarray):
    """Exit-direction angles (rad, -1 where a ray did not escape in both) of
    two ray-record arrays, and their BhAngleStats."""
    a = np.ascontiguousarray(a, RECORD_DTYPE).ravel()
    b = np.ascontiguousarray(b, RECORD_DTYPE).ravel()
    if a.shape != b.shape:
        raise ValueError("record counts differ")
    ang = np.empty(a.size, np.float32)
    st = BhAngleStats()
    if load().bhExitAngleDiff(a.ctypes.data, b.ctypes.data, a.size, ang.ctypes.data, C.byref(st)) != 0:
        raise BhError("bhExitAngleDiff failed")
    return ang, st


def probe_grid(P: RenderParams, nx: int, ny: int) -> np.ndarray:
    lib = load()
    buf = DeviceBuffer(nx * ny * PROBE_DTYPE.itemsize)
    check(lib.bhProbeGrid(buf.ptr, C.byref(P), nx, ny, None), "bhProbeGrid")
    synchronize()
    return buf.download(PROBE_DTYPE).reshape(ny, nx)


def math_probe(fn: int, a, b=None) -> np.ndarray:
    """bhMathProbe: results [n, 2] (float64) of library function `fn` on the
    argument pairs (a, b)."""
    a = np.asarray(a, np.float64).ravel()
    b = np.zeros_like(a) if b is None else np.broadcast_to(np.asarray(b, np.float64), a.shape)
    x = np.ascontiguousarray(np.stack([a, b], 1))
    n = a.size
    di, do = DeviceBuffer(max(x.nbytes, 16)), DeviceBuffer(max(x.nbytes, 16))
    di.upload(x)
    check(load().bhMathProbe(fn, di.ptr, n, do.ptr, None), "bhMathProbe")
    synchronize()
    return do.download(np.float64, 2 * n).reshape(n, 2)


def probe_dirs(P: RenderParams, dirs: np.ndarray) -> np.ndarray:
    lib = load()
    d = np.ascontiguousarray(dirs, np.float32).reshape(-1, 3)
    dbuf = DeviceBuffer(d.nbytes)
    dbuf.upload(d)
    buf = DeviceBuffer(len(d) * PROBE_DTYPE.itemsize)
    check(lib.bhProbeDirs(buf.ptr, dbuf.ptr, len(d), C.byref(P), None), "bhProbeDirs")
    synchronize()
    return buf.download(PROBE_DTYPE)


def get_schedule() -> BhSchedule:
    s = BhSchedule()
    load().bhGetSchedule(C.byref(s))
    return s


def set_schedule(wg: int | None = None, chunk: int | None = None, math: int | None = None,
                 refill: int | None = None, coop: int | None = None) -> BhSchedule:
    s = get_schedule()
    if wg is not None:
        s.wgSize = wg
    if chunk is not None:
        s.stepsPerChunk = chunk
    if math is not None:
        s.mathMode = math
    if refill is not None:
        s.raysPerRefill = refill
    if coop is not None:
        s.diskCoop = coop
    check(load().bhSetSchedule(C.byref(s)), "bhSetSchedule")
    return s


def restore_schedule(s: BhSchedule) -> None:
    check(load().bhSetSchedule(C.byref(s)), "bhSetSchedule")


def schedule_for(P: RenderParams) -> BhSchedule:
    """The schedule a frame of P runs under (tuned for its class, or the default)."""
    s = BhSchedule()
    load().bhScheduleFor(C.byref(P), C.byref(s))
    return s


def autotune(P: RenderParams, reps: int = 2):
    """Run the library's autotuner on P; returns (chosen BhSchedule, best trace ms)."""
    s, ms = BhSchedule(), C.c_float()
    check(load().bhAutotune(C.byref(P), reps, C.byref(s), C.byref(ms), None), "bhAutotune")
    return s, ms.value


class Accumulator:
    """Progressive accumulation of one view (device sample index, graph replay)."""

    def __init__(self, w: int, h: int):
        self.w, self.h = w, h
        p = C.c_void_p()
        check(load().bhAccumCreate(w, h, C.byref(p)), "bhAccumCreate")
        self.ptr = p

    def reset(self):
        check(load().bhAccumReset(self.ptr, None), "bhAccumReset")

    def step(self, P: RenderParams):
        check(load().bhAccumStep(self.ptr, C.byref(P), None), "bhAccumStep")

    def capture(self, P: RenderParams):
        check(load().bhAccumCapture(self.ptr, C.byref(P), None), "bhAccumCapture")

    def replay(self, frames: int):
        check(load().bhAccumReplay(self.ptr, frames, None), "bhAccumReplay")

    def sample_index(self) -> int:
        i = C.c_int()
        check(load().bhAccumSampleIndex(self.ptr, C.byref(i), None), "bhAccumSampleIndex")
        return i.value

    def readback(self):
        lib = load()
        n = self.w * self.h
        rgba = np.empty((self.h, self.w, 4), np.uint8)
        hdr = np.empty((self.h, self.w, 4), np.float32)
        synchronize()
        check(lib.bhCopyToHost(rgba.ctypes.data_as(C.c_void_p), lib.bhAccumColor(self.ptr), n * 4), "color")
        check(lib.bhCopyToHost(hdr.ctypes.data_as(C.c_void_p), lib.bhAccumHDR(self.ptr), n * 16), "hdr")
        return rgba, hdr

    def close(self):
        if self.ptr:
            load().bhAccumDestroy(self.ptr)
            self.ptr = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Surrogate:
    """The learned surrogate (network parameters, training set, training)."""

    def __init__(self, seed: int = 2):
        p = C.c_void_p()
        check(load().bhSurrogateCreate(seed, C.byref(p)), "bhSurrogateCreate")
        self.ptr = p
        self.nparams = load().bhSurrogateParamCount()

    def params(self) -> np.ndarray:
        out = np.empty(self.nparams, np.float32)
        check(load().bhSurrogateGetParams(self.ptr, out.ctypes.data_as(C.c_void_p), self.nparams), "get params")
        return out

    def set_params(self, p: np.ndarray) -> None:
        a = np.ascontiguousarray(p, np.float32)
        check(load().bhSurrogateSetParams(self.ptr, a.ctypes.data_as(C.c_void_p), a.size), "set params")

    def dataset(self, n_cfg: int, rays_per_cfg: int, seed: int = 2) -> None:
        check(load().bhSurrogateDataset(self.ptr, n_cfg, rays_per_cfg, seed, None), "bhSurrogateDataset")

    def dataset_for(self, views, rays_per_view: int, seed: int = 2) -> None:
        """Training set from the cameras of the given RenderParams views."""
        arr = (RenderParams * len(views))(*views)
        check(load().bhSurrogateDatasetFor(self.ptr, arr, len(views), rays_per_view, seed, None), "dataset_for")

    def dataset_view(self):
        """(inputs [n,6], targets [n] MLP_TARGET_DTYPE, escDir [n,3]) copied to host."""
        x, t, e, n = C.c_void_p(), C.c_void_p(), C.c_void_p(), C.c_int()
        check(load().bhSurrogateDatasetView(self.ptr, C.byref(x), C.byref(t), C.byref(e), C.byref(n)), "view")
        n = n.value
        lib = load()
        X = np.empty((n, 6), np.float32)
        T = np.empty(n, MLP_TARGET_DTYPE)
        E = np.empty((n, 3), np.float32)
        synchronize()
        for dst, src in ((X, x), (T, t), (E, e)):
            if n:
                check(lib.bhCopyToHost(dst.ctypes.data_as(C.c_void_p), src, dst.nbytes), "download")
        return X, T, E

    def dataset_device(self):
        x, t, e, n = C.c_void_p(), C.c_void_p(), C.c_void_p(), C.c_int()
        check(load().bhSurrogateDatasetView(self.ptr, C.byref(x), C.byref(t), C.byref(e), C.byref(n)), "view")
        return x, t, e, n.value

    def train(self, steps: int, batch: int = 65536, lr: float = 1e-3) -> np.ndarray:
        loss = np.zeros((steps, 3), np.float32)
        check(load().bhSurrogateTrain(self.ptr, steps, batch, lr, loss.ctypes.data_as(C.c_void_p), None), "train")
        return loss

    def forward(self, x: np.ndarray, fused: bool = False) -> np.ndarray:
        """Head outputs [n x 5] from the training path's GEMMs, or (fused)
        from the render path's fused inference kernel."""
        a = np.ascontiguousarray(x, np.float32).reshape(-1, 6)
        n = a.shape[0]
        xi, yo = DeviceBuffer(a.nbytes), DeviceBuffer(n * 5 * 4)
        xi.upload(a)
        fn = load().bhSurrogateForwardFused if fused else load().bhSurrogateForward
        check(fn(self.ptr, xi.ptr, n, yo.ptr, None), "forward")
        synchronize()
        return yo.download(np.float32).reshape(n, 5)

    def loss_grad(self, x: np.ndarray, t: np.ndarray):
        a = np.ascontiguousarray(x, np.float32).reshape(-1, 6)
        tt = np.ascontiguousarray(t, MLP_TARGET_DTYPE)
        n = a.shape[0]
        xi, ti = DeviceBuffer(a.nbytes), DeviceBuffer(tt.nbytes)
        xi.upload(a)
        ti.upload(tt.view(np.uint8))
        loss = np.zeros(3, np.float32)
        grad = np.zeros(self.nparams, np.float32)
        check(load().bhSurrogateLossGrad(self.ptr, xi.ptr, ti.ptr, n, loss.ctypes.data_as(C.c_void_p),
                                         grad.ctypes.data_as(C.c_void_p), None), "loss_grad")
        return loss, grad

    def close(self):
        if self.ptr:
            load().bhSurrogateDestroy(self.ptr)
            self.ptr = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def set_surrogate(s, sigma_max: float = 1e-4, disk_prob_max: float = 1e-3) -> None:
    """Attach a Surrogate to the render path (None detaches)."""
    check(load().bhSetSurrogate(s.ptr if s is not None else None, sigma_max, disk_prob_max), "bhSetSurrogate")


def last_surrogate_stats(stream=None):
    c, a = C.c_uint(), C.c_uint()
    check(load().bhLastSurrogateStats(C.byref(c), C.byref(a), stream), "bhLastSurrogateStats")
    return c.value, a.value


def surrogate_configs(seed: int, n: int):
    arr = (RenderParams * n)()
    check(load().bhSurrogateConfigs(seed, n, arr), "bhSurrogateConfigs")
    return [arr[i] for i in range(n)]


def get_precision() -> BhPrecision:
    p = BhPrecision()
    load().bhGetPrecision(C.byref(p))
    return p


def set_precision(band: int | None = None, crit: float | None = None, polar: float | None = None) -> BhPrecision:
    """FP64 critical-band pass on/off and its thresholds (returns the new setting)."""
    p = get_precision()
    if band is not None:
        p.fp64Band = band
    if crit is not None:
        p.critBand = crit
    if polar is not None:
        p.polarBand = polar
    check(load().bhSetPrecision(C.byref(p)), "bhSetPrecision")
    return p


def last_trace_kernel_ms(stream=None):
    """(tile kernel ms, band kernel ms) of the last timed frame."""
    t, b = C.c_float(), C.c_float()
    check(load().bhLastTraceKernelMs(C.byref(t), C.byref(b), stream), "bhLastTraceKernelMs")
    return t.value, b.value


def last_tile_order(n_max: int, stream=None):
    """(keys, order) of the last frame's tile queue, or None when it was not ordered."""
    keys = np.zeros(n_max, np.uint32)
    order = np.zeros(n_max, np.uint32)
    n = C.c_int()
    e = load().bhLastTileOrder(keys.ctypes.data, order.ctypes.data, n_max, C.byref(n), stream)
    if n.value == 0:
        return None
    check(e, "bhLastTileOrder")
    return keys[:n.value], order[:n.value]


def restore_precision(p: BhPrecision) -> None:
    check(load().bhSetPrecision(C.byref(p)), "bhSetPrecision")


def last_band_stats(stream=None):
    """(band ray count, FP64 band kernel ms or 0) of the last frame on `stream`."""
    n, ms = C.c_uint(), C.c_float()
    check(load().bhLastBandStats(C.byref(n), C.byref(ms), stream), "bhLastBandStats")
    return n.value, ms.value


class BatchBuffers:
    """Device buffers for a batch of views (out, accum, bloomA, bloomB per view)."""

    def __init__(self, views):
        self.views = views
        self.n = len(views)
        self.bufs = []
        for P in views:
            W, H = P.width, P.height
            nh = ((W + 1) // 2) * ((H + 1) // 2)
            b = [DeviceBuffer(W * H * 4), DeviceBuffer(W * H * 16), DeviceBuffer(nh * 16), DeviceBuffer(nh * 16)]
            b[1].fill(0)
            self.bufs.append(b)
        ptrs = lambda k: (C.c_void_p * self.n)(*[b[k].ptr.value for b in self.bufs])
        self.outs, self.accs, self.bAs, self.bBs = ptrs(0), ptrs(1), ptrs(2), ptrs(3)

    def render(self, stream=None):
        check(load().bhRenderBatch(self.n, self.views, self.outs, self.accs, self.bAs, self.bBs, stream),
              "bhRenderBatch")

    def render_persistent(self, stream=None) -> int:
        """bhRenderBatchPersistent; returns its HIP error code."""
        return load().bhRenderBatchPersistent(self.n, self.views, self.outs, self.accs, self.bAs, self.bBs, stream)

    def latency_ms(self, stream=None):
        out = (C.c_float * self.n)()
        check(load().bhBatchViewLatency(out, self.n, stream), "bhBatchViewLatency")
        return list(out)

    def graph(self, autotune: int = 1) -> "BatchGraph":
        """The persistent launch of this batch captured as a graph."""
        return BatchGraph(self, autotune)

    def readback(self, v):
        P = self.views[v]
        rgba = self.bufs[v][0].download(np.uint8).reshape(P.height, P.width, 4)
        acc = self.bufs[v][1].download(np.float32).reshape(P.height, P.width, 4)
        return rgba, acc


class BatchGraph:
    """bhBatchGraph* over a BatchBuffers' views and buffers."""

    def __init__(self, buffers: "BatchBuffers", autotune: int = 1):
        self.b = buffers
        p = C.c_void_p()
        check(load().bhBatchGraphCreate(buffers.n, buffers.views, buffers.outs, buffers.accs, buffers.bAs, buffers.bBs,
                                        autotune, C.byref(p)), "bhBatchGraphCreate")
        self.ptr = p

    def launch(self, stream=None):
        check(load().bhBatchGraphLaunch(self.ptr, stream), "bhBatchGraphLaunch")

    def info(self):
        wg, ms = C.c_int(), (C.c_float * 3)()
        check(load().bhBatchGraphInfo(self.ptr, C.byref(wg), ms), "bhBatchGraphInfo")
        return wg.value, list(ms)

    def latency_ms(self, stream=None):
        out = (C.c_float * self.b.n)()
        check(load().bhBatchGraphLatency(self.ptr, out, self.b.n, stream), "bhBatchGraphLatency")
        return list(out)

    def close(self):
        if self.ptr:
            load().bhBatchGraphDestroy(self.ptr)
            self.ptr = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
