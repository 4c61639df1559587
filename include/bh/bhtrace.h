/*
 * bh/bhtrace.h -- C ABI of the MI355X (gfx950) null-geodesic tracer library
 * libbhtrace.so.
 *
 * Every entry point is extern "C", takes plain pointers and sizes, and never
 * throws.  Device pointers are HIP device allocations (hipMalloc or
 * bhDeviceAlloc below).  Error codes are HIP error codes (0 = success).
 *
 * The first two entry points are the drop-in replacements for the
 * reference's launch API; the rest are new (headless framebuffer, probes,
 * batched multi-view launch, scheduling, timing, learned surrogate).
 *
 * C++ callers get the reference's `const RenderParams&` form; C callers pass
 * `const RenderParams*` (ABI-identical).
 */
#ifndef BH_BHTRACE_H
#define BH_BHTRACE_H

#include <stddef.h>
#include <stdint.h>
#include "render_params.h"

#ifdef __cplusplus
extern "C" {
#define BH_PARAMS_IN const RenderParams&
#else
#define BH_PARAMS_IN const RenderParams*
#endif

typedef int   bh_error_t;    /* == hipError_t (int-sized enum, 0 = hipSuccess) */
typedef void* bh_stream_t;   /* == hipStream_t (opaque; 0 = null stream)      */

/* ------------------------------------------------------------------------ */
/* Drop-in launch API                                                       */
/* ------------------------------------------------------------------------ */

/* Full-quality frame: one sample per pixel blended into `accum`
 * (float4[W*H], read-modify-write), optional bloom, then
 * exposure/ACES/gamma/dither into `out` (uchar4[W*H]).  bloomA/bloomB
 * (float4[ceil(W/2)*ceil(H/2)] each) are required with bloom, as in the
 * reference, but serve as scratch only: the bloom image is rebuilt per
 * output tile on chip and their contents are left undefined.  `p` is copied
 * and sanitized; the caller's copy is untouched.  Asynchronous on `stream`.
 * Replaces: extern "C" cudaError_t launchRenderPipeline(uchar4*, float4*,
 * float4*, float4*, const RenderParams&, cudaStream_t)
 * (black_hole_kernel.cu:108-134). */
bh_error_t launchRenderPipeline(bh_uchar4* out, bh_float4* accum,
                                bh_float4* bloomA, bh_float4* bloomB,
                                BH_PARAMS_IN p, bh_stream_t stream);

/* Single-pass frame: centered sample, no accumulation, bloom forced off,
 * straight to `out`.
 * Replaces: extern "C" cudaError_t launchRenderKernel(uchar4*, const
 * RenderParams&, cudaStream_t) (black_hole_kernel.cu:151-163). */
bh_error_t launchRenderKernel(bh_uchar4* out, BH_PARAMS_IN p, bh_stream_t stream);

/* launchRenderPipeline that allocates nothing.  No counterpart in the
 * reference: for callers that cannot allocate device memory (code under HIP
 * graph capture, tight serving loops, harnesses that forbid hipMalloc).
 * Same arguments, semantics and frames as launchRenderPipeline, except:
 *   - the library's scratch (trace records, counters, FP64 band lists, tile
 *     order) comes from one static 96 MiB device array per GPU, held until
 *     the process ends; frames whose scratch does not fit -- 45 bytes per
 *     pixel plus 8 per 8x8 tile and ~5 KiB, i.e. up to 1920x1080 or about
 *     2.2 Mpixels of any shape -- are rejected with hipErrorInvalidValue
 *     before anything is enqueued;
 *   - no allocation, stream or event is created (the first call on a device
 *     looks up the array's address); the FP64 band and FP32 tile traces run
 *     one after the other on `stream`; kernel timers are not recorded; an
 *     attached learned surrogate is not applied;
 *   - every call on a GPU shares that GPU's array: calls that overlap on one
 *     GPU (several streams or host threads) race. */
bh_error_t launchRenderPipelineNoAlloc(bh_uchar4* out, bh_float4* accum,
                                       bh_float4* bloomA, bh_float4* bloomB,
                                       BH_PARAMS_IN p, bh_stream_t stream);

/* ------------------------------------------------------------------------ */
/* Host-side parameter helpers                                              */
/* ------------------------------------------------------------------------ */

/* Reference defaults (render_params.h:24-86). */
void bhDefaultRenderParams(RenderParams* p);

/* In-place sanitation; returns 1 if any field was changed.
 * Follows sanitizeRenderParams (metric.cuh:344-419). */
int bhSanitizeRenderParams(RenderParams* p);

typedef struct BHDerived
{
    float a, Q;             /* code units */
    float rPlus, rMinus;    /* horizons (0 when naked) */
    float rPhoton, rErgo;
    float rIsco;            /* float scan, as the reference */
    int   clamped, naked;
} BHDerived;

/* Clamp (M, a*, q) in place and derive horizons, photon orbit, ergosphere and
 * ISCO.  Follows bhValidateAndDerive (metric.cuh:257-329). */
void bhValidateAndDerive(int model, float* M, float* aStar, float* q,
                         int allowNaked, BHDerived* out);

/* FP64 roots (no reference counterpart; the reference scans in float). */
double bhHorizonOuterFP64(double M, double a, double Q);
double bhIscoFP64(double M, double a, double Q);   /* prograde equatorial ISCO */

/* Orbit camera looking at the origin: follows fillCamera (main.cpp:186-228)
 * with no body rotation, plus an explicit vertical field of view. */
void bhFillCamera(RenderParams* p, float azimuthRad, float elevationRad,
                  float distance, float fovYDegrees);

/* Orbit camera with the black hole's body orientation R = Rz(oriZ) Ry(oriY)
 * Rx(oriX) about fixed world axes: follows fillCamera and
 * applyBodyOrientation (main.cpp:147-167, 186-228).  With zero orientation
 * it equals bhFillCamera. */
void bhFillCameraOriented(RenderParams* p, float azimuthRad, float elevationRad, float distance,
                          float fovYDegrees, float oriX, float oriY, float oriZ);

/* ------------------------------------------------------------------------ */
/* Viewer layer: the caller-side producer of RenderParams                   */
/* Follows the reference application loop (main.cpp:48-228, 230-337,        */
/* 418-478): orbit camera and its clamps, body orientation, model knobs,     */
/* quality presets, disk clock, and the progressive-accumulation rules       */
/* (restart on any geometry change; accumMode 1 when the disk is still,      */
/* 2 when it animates; sample index capped at 4096).  Host only.            */
/* ------------------------------------------------------------------------ */
typedef struct BhViewState
{
    float azimuth, elevation, distance;   /* orbit camera: rad, rad, code units (rs = 1) */
    float oriX, oriY, oriZ;               /* body orientation (rad) */
    int   model;                          /* BH_SCHWARZSCHILD .. BH_KERR_NEWMAN */
    float aStar, q;                       /* knobs; each model uses the ones that apply */
    int   allowNaked;
    float exposure;
    int   diskOn, bloomOn, hotSpotsOn, animPaused;
    float dPhi;
    int   maxSteps;
    float fovY;                           /* degrees (60 in the reference) */
    float diskTime;                       /* disk clock (advanced by dt unless paused) */
    int   sampleIndex;                    /* sample index of the next frame */
} BhViewState;

typedef struct BhViewer BhViewer;
enum { BH_KEY_PRESS = 1, BH_KEY_REPEAT = 2 };
/* Keys: printable keys by their ASCII code ('[', ']', ',', '.', 'B', 'N',
 * 'H', 'R', ' ', '1', '2', '3'); function keys below. */
enum { BH_KEY_F1 = 0x101, BH_KEY_F2, BH_KEY_F3, BH_KEY_F4, BH_KEY_F5 };
/* Held-key mask for bhViewerFrame (W/S zoom, A/D azimuth, Q/E elevation,
 * -/= exposure, X/Y/Z body rotation, SHIFT reverses the rotation). */
enum
{
    BH_HELD_W = 1u << 0, BH_HELD_S = 1u << 1, BH_HELD_A = 1u << 2, BH_HELD_D = 1u << 3,
    BH_HELD_Q = 1u << 4, BH_HELD_E = 1u << 5, BH_HELD_EXPOSURE_DOWN = 1u << 6, BH_HELD_EXPOSURE_UP = 1u << 7,
    BH_HELD_X = 1u << 8, BH_HELD_Y = 1u << 9, BH_HELD_Z = 1u << 10, BH_HELD_SHIFT = 1u << 11
};
/* Reference defaults (main.cpp:48-88). */
void bhViewStateDefaults(BhViewState* s);
int  bhViewerCreate(int width, int height, BhViewer** v);
void bhViewerDestroy(BhViewer* v);
int  bhViewerGetState(const BhViewer* v, BhViewState* s);
/* Replaces the state (clamped); model knob changes are re-derived at the
 * next frame. */
int  bhViewerSetState(BhViewer* v, const BhViewState* s);
int  bhViewerResize(BhViewer* v, int width, int height, int swapRB);
/* Camera and orientation back to the defaults (key R). */
void bhViewerReset(BhViewer* v);
/* Mouse drag (0.005 rad per pixel) and wheel (distance x 0.92^steps). */
void bhViewerOrbit(BhViewer* v, float dxPixels, float dyPixels);
void bhViewerZoom(BhViewer* v, float wheelSteps);
/* One key event; returns 1 if it changed the state. */
int  bhViewerKey(BhViewer* v, int key, int action);
/* Advance by dt seconds (clamped to 0.1) with the held keys, and produce the
 * frame's parameters.  Returns 1 when the accumulation restarts (first
 * frame or changed geometry), 0 when it continues, -1 on bad arguments. */
int  bhViewerFrame(BhViewer* v, float dt, unsigned held, RenderParams* out);

/* Named workload presets (camera, model, disk and post settings).  `frame`
 * selects the disk time where a workload has several frames.  Returns 0 on
 * success, -1 for an unknown name.  Names:
 *   "schw_256_balanced"      Schwarzschild 256x256, disk [3,12], bloom off
 *   "kerr09_1080p_high"      Kerr a*=0.9 1920x1080, hot spots, bloom; frame 0,1,2 -> t=0,0.5,1
 *   "kn_m07_q04_2160p_high"  Kerr-Newman a*=-0.7 q=0.4 3840x2160, bloom */
int bhWorkloadParams(const char* name, int frame, RenderParams* out);

/* Seeded batch of views, "uneven_views": resolution uniform in {320x180,
 * 640x360, 1280x720, 1920x1080}, Kerr-Newman with a* ~ U[-0.99, 0.99] and
 * q ~ U[0, sqrt(0.98 - a*^2)], camera distance log-uniform in [6, 40],
 * elevation U[-80, 80] deg, vertical FOV U[40, 90] deg, disk on with
 * probability 1/2 from the FP64 ISCO to 8, disk time U[0, 10], balanced
 * preset, bloom on.  SplitMix64 stream from `seed`.  Writes `maxViews` views
 * and returns that count, or -1 for an unknown name. */
int bhWorkloadViews(const char* name, uint64_t seed, int maxViews, RenderParams* out);

/* ------------------------------------------------------------------------ */
/* Probes (verification; follow probeGridKernel/probeOneKernel,             */
/* test_main.cu:71-98)                                                      */
/* ------------------------------------------------------------------------ */
typedef struct BhProbe
{
    int   outcome;      /* 0 captured, 1 escaped */
    float escDir[3];
    float maxHviol;     /* max |K|/(E^2 (r^2+a^2)) along the ray (rotating models) */
    float lum;          /* accum.x + accum.y + accum.z */
    int   steps;
    float trans;
} BhProbe;

/* nx*ny rays through pixel centres of an nx x ny grid; devOut: BhProbe[nx*ny]. */
bh_error_t bhProbeGrid(BhProbe* devOut, BH_PARAMS_IN p, int nx, int ny, bh_stream_t stream);
/* n rays with explicit world directions devDirs[n]; devOut: BhProbe[n]. */
bh_error_t bhProbeDirs(BhProbe* devOut, const bh_float3* devDirs, int n,
                       BH_PARAMS_IN p, bh_stream_t stream);

/* The library's own approximations of elementary functions, evaluated on
 * the GPU as the fast trace kernels evaluate them (verification of their
 * error bounds; no reference counterpart).  devIn: double[2n] argument
 * pairs (a, b); devOut: double[2n] result pairs.  Functions: */
enum
{
    BH_MATH_ATAN2_POLY = 0,       /* atan2(y = a, x = b) in float, polynomial (fast policy) */
    BH_MATH_FAST_SINCOS = 1,      /* hardware sin(a), cos(a) in float */
    BH_MATH_FAST_RCP = 2,         /* hardware 1/a in float */
    BH_MATH_FAST_RSQ = 3,         /* hardware 1/sqrt(a) */
    BH_MATH_FAST_SQRT = 4,        /* hardware sqrt(a) */
    BH_MATH_FAST_EXP = 5,         /* exp(a) via hardware exp2 */
    BH_MATH_FAST_LOG = 6,         /* log(a) via hardware log2 */
    BH_MATH_FAST_POW = 7,         /* a^b via hardware exp2/log2 */
    BH_MATH_D64_SINCOS = 8,       /* sin(a), cos(a) in double: Cody-Waite reduction and Taylor polynomials */
    BH_MATH_D64_SINCOS_NEAR = 9,  /* sin, cos of a + b from those of a by angle addition (|b| < 1/16) */
    BH_MATH_D64_RCP = 10,         /* 1/a in double: hardware seed and two Newton steps */
    BH_MATH_BLACKBODY = 11,       /* channel (int)b of the disk's linear blackbody colour at a kelvin:
                                     [0] fast policy (2.2 power in the log2 domain), [1] exact policy */
    BH_MATH_WRAP_PHI = 12         /* a wrapped to (-pi, pi]: [0] fast policy (floor), [1] exact (fmodf) */
};
bh_error_t bhMathProbe(int fn, const double* devIn, int n, double* devOut, bh_stream_t stream);

/* Per-pixel trace record of one frame (no shading): for checking exit
 * directions against an FP64 trace.  devOut: BhRayRecord[W*H]. */
typedef struct BhRayRecord
{
    float escDir[3];
    int   flags;        /* bit0 escaped, bit1 captured, bits 8.. step count */
    float accum[3];     /* disk radiance */
    float trans;        /* remaining transmittance */
} BhRayRecord;
bh_error_t bhTraceRecords(BhRayRecord* devOut, BH_PARAMS_IN p, bh_stream_t stream);

/* ------------------------------------------------------------------------ */
/* Batched views                                                            */
/* ------------------------------------------------------------------------ */
/* Renders `nViews` views of any sizes and models, each with the full
 * pipeline of launchRenderPipeline (accumulation, bloom, tonemap) into its
 * own buffers: outs[v] uchar4[W*H], accums[v] float4[W*H], bloomAs[v] /
 * bloomBs[v] float4[ceil(W/2)*ceil(H/2)] (the arrays may be NULL when no view
 * has bloom).  `views` is a host array; each view is copied and sanitized,
 * and every view is checked before any work is enqueued
 * (hipErrorInvalidValue).  Results equal per-view launchRenderPipeline calls
 * bit for bit.  The views run as single-view pipelines, each under the
 * schedule bhScheduleFor gives it, on six library worker streams forked
 * from `stream` (largest views first, each on the least-loaded worker), and
 * `stream` continues when all have finished (also after an error).
 * Library state per caller stream: the six workers and the auxiliary streams
 * of their pipelines are created with a CU mask of every CU, which gives each
 * its own hardware queue (12 queues; HIP makes such streams blocking, so the
 * batch also orders against the legacy null stream); where CU-masked streams
 * cannot be created, pooled non-blocking streams are used.  Each view index
 * keeps its own trace records (32 bytes per pixel) between batches, so a view
 * repeated at the same index orders its work by its previous step counts.
 * bhStreamRelease frees all of it. */
bh_error_t bhRenderBatch(int nViews, const RenderParams* views, bh_uchar4* const* outs,
                         bh_float4* const* accums, bh_float4* const* bloomAs,
                         bh_float4* const* bloomBs, bh_stream_t stream);
/* The same batch as one persistent launch: one device queue of (view, tile)
 * items over all views, traced by one kernel (wavefronts take the next item
 * as they finish; the view index is uniform per wavefront), then batched
 * shade, bloom and composite kernels over a device table of the views, all
 * on `stream`.  With the FP64 band pass on (bhSetPrecision), every view is
 * first classified and the band rays of all views are traced in double by
 * one launch over a queue of (view, 64-ray chunk) items, beside the FP32
 * trace (which skips them) on a library stream forked from `stream` and
 * joined back; with it off every ray is traced in FP32 (the reference's
 * arithmetic).  Either way each view's frame equals launchRenderPipeline's
 * under the same precision, bit for bit.  Per-view completion times (the
 * later of a view's FP32 tiles and its band rays) are read with
 * bhBatchViewLatency.  No reference
 * counterpart (the reference renders one view per launch,
 * black_hole_kernel.cu:108-134). */
bh_error_t bhRenderBatchPersistent(int nViews, const RenderParams* views, bh_uchar4* const* outs,
                                   bh_float4* const* accums, bh_float4* const* bloomAs,
                                   bh_float4* const* bloomBs, bh_stream_t stream);
/* A steady batch captured once and replayed: the persistent launch of
 * bhRenderBatchPersistent for fixed views and buffers, as a hipGraph of its
 * kernels (counters reset by a kernel; no memset or copy nodes), with the
 * band pass when the precision at creation has it on.  Create uploads the
 * view table once, gives the batch its own records, band buffers, auxiliary
 * stream and counters, and with `autotune` first times one eager launch per trace
 * workgroup size (64, 128, 256; the frames do not depend on it) and
 * captures the fastest.  Launch replays the whole batch on `stream`;
 * Latency reads the per-view trace completion times of the last replay
 * (synchronizes `stream`).  Frames equal bhRenderBatchPersistent's. */
typedef struct BhBatchGraph BhBatchGraph;
bh_error_t bhBatchGraphCreate(int nViews, const RenderParams* views, bh_uchar4* const* outs,
                              bh_float4* const* accums, bh_float4* const* bloomAs,
                              bh_float4* const* bloomBs, int autotune, BhBatchGraph** g);
bh_error_t bhBatchGraphLaunch(BhBatchGraph* g, bh_stream_t stream);
/* Workgroup size the graph traces with, and the tuning times (ms, 0 when not tuned). */
bh_error_t bhBatchGraphInfo(BhBatchGraph* g, int* wgSize, float* tunedMs3);
bh_error_t bhBatchGraphLatency(BhBatchGraph* g, float* msOut, int nViews, bh_stream_t stream);
void       bhBatchGraphDestroy(BhBatchGraph* g);

/* Per-view completion time (ms after the batch began on the GPU) of the
 * last bhRenderBatch (the view's whole pipeline) or bhRenderBatchPersistent
 * (the view's last traced tile, from device wall-clock stamps) on `stream`;
 * synchronizes the stream. */
bh_error_t bhBatchViewLatency(float* msOut, int nViews, bh_stream_t stream);

/* ------------------------------------------------------------------------ */
/* Headless framebuffer (replaces the Vulkan interop image)                 */
/* ------------------------------------------------------------------------ */
typedef struct BhFramebuffer BhFramebuffer;
bh_error_t bhFramebufferCreate(int width, int height, BhFramebuffer** fb);
void       bhFramebufferDestroy(BhFramebuffer* fb);
bh_uchar4* bhFramebufferColor(BhFramebuffer* fb);   /* device uchar4[W*H] */
bh_float4* bhFramebufferHDR(BhFramebuffer* fb);     /* device float4[W*H] (accum) */
/* launchRenderPipeline into the framebuffer's own buffers (p.width/height
 * must match the framebuffer). */
bh_error_t bhFramebufferRender(BhFramebuffer* fb, BH_PARAMS_IN p, bh_stream_t stream);
/* Synchronous copy to host; either pointer may be NULL. */
bh_error_t bhFramebufferReadback(BhFramebuffer* fb, bh_uchar4* hostColor,
                                 bh_float4* hostHdr, bh_stream_t stream);

/* The post chain of launchRenderPipeline alone -- bloom (when
 * p.bloomEnabled) and the composite/tonemap -- on an existing accumulation
 * buffer (black_hole_kernel.cu:120-132): re-tonemaps a stored HDR image,
 * e.g. after an exposure or bloom-strength change, without tracing.  Reads
 * accum (not written), writes out; bloomA/bloomB as in launchRenderPipeline
 * (required with bloom, contents left undefined).  p is sanitized as in launchRenderPipeline.
 * With kernel timers on, the chain's time is reported as bhLastKernelTimes'
 * postMs (traceMs and shadeMs read 0). */
bh_error_t bhPostChain(bh_uchar4* out, const bh_float4* accum, bh_float4* bloomA,
                       bh_float4* bloomB, BH_PARAMS_IN p, bh_stream_t stream);
/* bhPostChain on the framebuffer's own buffers: its HDR image re-tonemapped
 * into its colour buffer (p.width/height must match the framebuffer). */
bh_error_t bhFramebufferPost(BhFramebuffer* fb, BH_PARAMS_IN p, bh_stream_t stream);

/* ------------------------------------------------------------------------ */
/* Progressive accumulation, graph-replayed                                 */
/* ------------------------------------------------------------------------ */
/* Frames of one steady view accumulated into one HDR buffer, as in the
 * reference's progressive anti-aliasing loop (main.cpp:439-478): frame k
 * takes sample index k (R2 sub-pixel jitter, running mean in accumMode 1,
 * EMA in accumMode 2; p.sampleIndex is ignored).  The index lives in device
 * memory and the step increments it, so one captured step (a hipGraph of
 * trace, shade, bloom and composite) replays for every later frame. */
typedef struct BhAccumulator BhAccumulator;
bh_error_t bhAccumCreate(int width, int height, BhAccumulator** acc);
void       bhAccumDestroy(BhAccumulator* acc);
bh_uchar4* bhAccumColor(BhAccumulator* acc);   /* device uchar4[W*H] */
bh_float4* bhAccumHDR(BhAccumulator* acc);     /* device float4[W*H] */
bh_error_t bhAccumReset(BhAccumulator* acc, bh_stream_t stream);   /* next frame is sample 0 */
/* One frame with eager launches on `stream`. */
bh_error_t bhAccumStep(BhAccumulator* acc, BH_PARAMS_IN p, bh_stream_t stream);
/* Capture one step for `p` as a graph (synchronizes `stream` first; nothing
 * is executed).  Replays run the captured step `frames` times on `stream`. */
bh_error_t bhAccumCapture(BhAccumulator* acc, BH_PARAMS_IN p, bh_stream_t stream);
bh_error_t bhAccumReplay(BhAccumulator* acc, int frames, bh_stream_t stream);
/* Sample index of the next frame (synchronizes `stream`). */
bh_error_t bhAccumSampleIndex(BhAccumulator* acc, int* index, bh_stream_t stream);

/* ------------------------------------------------------------------------ */
/* Scheduling                                                               */
/* ------------------------------------------------------------------------ */
typedef struct BhSchedule
{
    int wgSize;         /* threads per workgroup of the trace kernel: 64, 128, 256 */
    int stepsPerChunk;  /* max RK4 steps a wavefront runs between refill checks: 1..4096 */
    int raysPerRefill;  /* rays a wavefront takes from the queue at once: it refills
                           when this many of its 64 lanes are idle (1..63); 64
                           takes whole 8x8 tiles, costliest first */
    int mathMode;       /* 0 exact IEEE arithmetic, 1 hardware fast intrinsics */
    int diskCoop;       /* 0: each lane samples its own ray's disk chords;
                           1: chords sampled cooperatively by the wavefront;
                           2: cooperatively when a whole-tile frame has no
                           more 8x8 tiles than the tile grid has wavefronts */
} BhSchedule;
/* Process default: {64, 4096, 64, 1, 2} (one-wavefront workgroups, whole
 * tiles, fast math, automatic cooperative sampling). */
void bhGetSchedule(BhSchedule* s);
bh_error_t bhSetSchedule(const BhSchedule* s);   /* -1 (hipErrorInvalidValue-like) on bad values */
/* Times the trace kernel of `p` under 18 candidate schedules (workgroup
 * size 64/128/256 x whole tiles / lane refill of 32 rays every 256 steps /
 * 16 rays every 64 steps x cooperative disk sampling off/on; `reps` timed
 * frames each after one warm-up) on `stream` and caches the fastest for
 * p's workload class (model class, size, disk, step budget, band, math
 * mode).  Every later frame of that class (launchRenderPipeline,
 * launchRenderKernel, bhRenderBatch views, bhTraceRecords) runs under it;
 * other classes keep the process default of bhSetSchedule, which tuning
 * does not change.  Output does not depend on the schedule. */
bh_error_t bhAutotune(BH_PARAMS_IN p, int reps, BhSchedule* chosen, float* bestMs, bh_stream_t stream);
/* The schedule a frame of `p` runs under: its class's tuned schedule, else
 * the process default. */
void bhScheduleFor(BH_PARAMS_IN p, BhSchedule* s);
void bhAutotuneClear(void);

/* ------------------------------------------------------------------------ */
/* Precision: FP64 critical-band rays                                        */
/* ------------------------------------------------------------------------ */
/* With fp64Band set, every frame first classifies each pixel's ray from its
 * FP64 invariants: rays near the critical (photon-ring) curve, -critBand <
 * delta < critBand / 4 with delta = min_r R(r)/(r^2+a^2)^2 of the radial
 * potential (1 - b^2/b_c^2 for the spherical models; delta > 0: the ray
 * falls in and has no exit direction; escaping Kerr rays with sin^2 of the
 * polar turning angle >= 0.1 s^2 or 0.3 s^2 use critBand / 2 or / 4 on the
 * escaping side), and Kerr rays whose polar
 * turning point has sin^2(theta) < polarBand (near the critical curve;
 * 0.3, 0.1 and 0.03 of it from |delta| = 0.2 s, 0.5 s and s), are traced
 * in double precision on a second stream while the FP32 kernel traces the
 * rest.  The widths apply to steps dPhi <= 0.007; a frame with a coarser
 * step uses critBand * s and polarBand * s^2, s = dPhi / 0.007.
 * No counterpart in the reference (which is FP32 throughout). */
typedef struct BhPrecision
{
    int   fp64Band;     /* 0 off (pure FP32, the reference's arithmetic), 1 on */
    float critBand;     /* default 4e-2 */
    float polarBand;    /* default 1e-3 */
} BhPrecision;
void bhGetPrecision(BhPrecision* p);
bh_error_t bhSetPrecision(const BhPrecision* p);   /* rejects negative or non-finite bands */
/* Band rays of the last frame traced on `stream` and, when kernel timers
 * are on, the FP64 band kernel's time in ms (synchronizes the stream). */
bh_error_t bhLastBandStats(unsigned* count, float* bandMs, bh_stream_t stream);
/* HIP-event times (ms) of the FP32 tile kernel, on the stream it runs on,
 * and of the FP64 band kernel (0 when the band pass is off) of the last frame
 * traced on `stream` with kernel timers on.  hipErrorNotReady when that
 * frame was not timed.  Synchronizes the stream. */
bh_error_t bhLastTraceKernelMs(float* tileMs, float* bandMs, bh_stream_t stream);
/* Tile order of the last frame on `stream` (whole-tile schedule, where the
 * queue hands out the costliest tiles first): per-tile cost keys, the
 * smallest over the tile's FP32 rays of the float bits of 64 / (steps + 1)
 * from the previous frame's records when the frame repeats the previous
 * view into the same records, else (band pass on) of |delta| (all ones when
 * the tile has no FP32 ray), and the queue order, `capacity` entries each
 * (either may be NULL).  *nTiles receives the tile count; hipErrorNotReady
 * when the frame was not ordered (band pass off and a view not traced just
 * before).  Synchronizes the stream. */
bh_error_t bhLastTileOrder(unsigned* keys, unsigned* order, int capacity, int* nTiles, bh_stream_t stream);

/* ------------------------------------------------------------------------ */
/* Timing (HIP events recorded on the launch stream)                        */
/* ------------------------------------------------------------------------ */
/* When enabled, launchRenderPipeline/launchRenderKernel record events around
 * the trace kernel and around the post kernels of each frame. */
void bhEnableKernelTimers(int enable);
/* Elapsed ms of the last timed frame: trace, sky shading, bloom+composite.
 * Synchronizes on the recorded events. */
bh_error_t bhLastKernelTimes(float* traceMs, float* shadeMs, float* postMs);

/* ------------------------------------------------------------------------ */
/* Learned surrogate (no counterpart in the reference)                      */
/* ------------------------------------------------------------------------ */
/* A small MLP predicts, for a camera ray, its exit deflection, whether the
 * plain trace would sample the disk, the radius of that first chord and
 * the log-variance of its own exit prediction.  Inputs (normalised): screen
 * impact parameter b, image-plane angle alpha, a*, q, camera radius,
 * camera inclination -> 16-frequency Fourier features -> 4 x 128 SiLU ->
 * 5 outputs.  BF16 weights on MFMA with FP32 accumulation; FP32 master
 * weights trained with Adam by the library's own HIP kernels on rays traced
 * by the plain FP32 kernel. */
typedef struct BhSurrogate BhSurrogate;
typedef struct BhMlpTarget
{
    float d0, d1;     /* exit deflection in the ray frame, minus the weak-field term */
    float disk;       /* 1 when the plain trace sampled the disk */
    float hitR;       /* cylindrical radius of the first disk chord */
    float escaped;    /* 1 escaped, 0 captured */
    float band;       /* 1 critical / polar band ray: no regression target */
    float pad[2];
} BhMlpTarget;
int        bhSurrogateParamCount(void);
bh_error_t bhSurrogateCreate(uint64_t seed, BhSurrogate** s);
void       bhSurrogateDestroy(BhSurrogate* s);
bh_error_t bhSurrogateSetParams(BhSurrogate* s, const float* host, int count);   /* resets Adam */
bh_error_t bhSurrogateGetParams(BhSurrogate* s, float* host, int count);
/* Training configurations: KN black holes a* ~ U[-0.99, 0.99], q ~ U[0, 0.6]
 * with a*^2 + q^2 <= 0.995, camera radius log-uniform in [6, 40] rs,
 * elevation uniform in [-80, 80] deg, disk ISCO..20 rs, 'high' steps. */
int        bhSurrogateConfigs(uint64_t seed, int n, RenderParams* out);
/* Traces nCfg x raysPerCfg rays (raysPerCfg a multiple of 64) with the plain
 * FP32 kernel into the training set (replacing any previous one). */
bh_error_t bhSurrogateDataset(BhSurrogate* s, int nCfg, int raysPerCfg, uint64_t seed, bh_stream_t stream);
/* Training set from the cameras of given views (for adapting the network to
 * a steady view): raysPerView rays (a multiple of 64) from each view's
 * camera position, psi up to 65 degrees, traced like bhSurrogateDataset. */
bh_error_t bhSurrogateDatasetFor(BhSurrogate* s, const RenderParams* views, int nViews, int raysPerView,
                                 uint64_t seed, bh_stream_t stream);
/* Device pointers of the training set: inputs float[n*6], targets
 * BhMlpTarget[n], plain-trace exit directions float[n*3]. */
bh_error_t bhSurrogateDatasetView(BhSurrogate* s, float** inputs, void** targets, float** escDir, int* n);
/* `steps` Adam steps on consecutive batches of the training set (batch a
 * multiple of 64), cosine learning-rate decay from lr to lr/20.  lossHost
 * (may be NULL) receives per-step [regression, disk BCE, hit radius] means. */
bh_error_t bhSurrogateTrain(BhSurrogate* s, int steps, int batch, float lr, float* lossHost, bh_stream_t stream);
/* Network outputs for n rows of normalised inputs (device float[n*6] ->
 * device float[n*5]): exit deflection minus the weak-field term (two
 * columns, milliradians), disk-hit logit, hit radius / 10, and s, the log of
 * the expected squared exit error in mrad^2. */
bh_error_t bhSurrogateForward(BhSurrogate* s, const float* devInputs, int n, float* devOut, bh_stream_t stream);
/* The same outputs from the render path's fused inference kernel (the
 * network held in LDS, activations in registers, as the router runs it). */
bh_error_t bhSurrogateForwardFused(BhSurrogate* s, const float* devInputs, int n, float* devOut,
                                   bh_stream_t stream);
/* Loss terms (sums over rows) and the gradient of the mean loss with respect
 * to every parameter for n rows (host outputs; synchronizes). */
bh_error_t bhSurrogateLossGrad(BhSurrogate* s, const float* devInputs, const void* devTargets, int n,
                               float* lossHost, float* gradHost, bh_stream_t stream);
/* Attach a trained surrogate to the render path (NULL detaches).  With the
 * FP64 band pass on, every escaping non-band camera ray inside the training
 * domain (psi <= 65 deg, camera radius 6..40 rs, q <= 0.6) is evaluated by
 * the network; rays whose predicted disk-hit probability is below
 * diskProbMax (not consulted when the disk is off) and predicted exit
 * spread (the RMS exit error the network expects) below sigmaMax (rad) take
 * the predicted exit direction and skip RK4; all others are traced by RK4.
 * Frames whose camera sits inside the disk's emitting layer (every ray
 * starts in the disk) do not consult the network.  The surrogate must
 * outlive its attachment. */
bh_error_t bhSetSurrogate(BhSurrogate* s, float sigmaMax, float diskProbMax);
/* Candidates and accepted rays of the last frame on `stream` (synchronizes). */
bh_error_t bhLastSurrogateStats(unsigned* candidates, unsigned* accepted, bh_stream_t stream);
/* HIP-event time (ms) of the network inference and routing kernel of the
 * last frame on `stream` traced with kernel timers on and a surrogate
 * consulted; hipErrorNotReady otherwise.  Synchronizes the stream. */
bh_error_t bhLastSurrogateKernelMs(float* ms, bh_stream_t stream);
/* The MFMA GEMM used by the network, exposed for tests: C (+)= op(A) op(B)
 * with bf16 operands (device) and f32 C (device; zero it when splitK > 1).
 * With splitK > 1 and ldc == N the K splits go through the weight-gradient
 * path (per-split partials reduced into C; synchronizes the stream). */
bh_error_t bhMlpGemmTest(const void* A, int lda, int transA, const void* B, int ldb, int transB, float* C, int ldc,
                         int M, int N, int K, int splitK, bh_stream_t stream);

/* ------------------------------------------------------------------------ */
/* Headless output formats and image comparison (host memory only)          */
/* Replace the reference's presentation path: the Vulkan interop buffer and */
/* per-swapchain-image copy (vulkan_context.cpp:446-484, 529-588;           */
/* cuda_interop.cpp:61-74).  Return 0 on success, -1 on a bad argument or   */
/* an I/O failure.                                                          */
/* ------------------------------------------------------------------------ */
/* Binary PPM (P6): 8-bit RGB, alpha dropped, rows top to bottom. */
int bhWritePPM(const char* path, const bh_uchar4* rgba, int width, int height);
/* PNG: 8-bit RGBA, filter 0, zlib stream of stored (uncompressed) blocks. */
int bhWritePNG(const char* path, const bh_uchar4* rgba, int width, int height);
/* PFM ("PF"): float RGB of the HDR buffer, little-endian (scale -1), rows
 * bottom to top as the format defines; w dropped. */
int bhWritePFM(const char* path, const bh_float4* hdr, int width, int height);

typedef struct BhDiffStats
{
    long long pixels;
    int       maxDiff[3];   /* per channel, in LSB */
    int       pad;
    long long hist[8];      /* per-pixel max channel diff: 0,1,2,3,4-7,8-15,16-63,64+ */
    double    fracWithin1;  /* share of pixels whose channels all differ by <= 1 LSB */
} BhDiffStats;
/* LSB statistics of two RGBA8 images of the same size (alpha ignored). */
int bhImageDiff(const bh_uchar4* a, const bh_uchar4* b, int width, int height, BhDiffStats* out);

typedef struct BhAngleStats
{
    long long bothEscaped;      /* rays escaped in both records */
    long long outcomeMismatch;  /* escaped in one record only */
    long long over1mrad;        /* escaped in both, angle > 1e-3 rad */
    double    median, p999, max;   /* exit-direction angle (rad) over bothEscaped */
} BhAngleStats;
/* Exit-direction angle between two ray-record buffers of n rays (host
 * copies of bhTraceRecords output).  anglesOut (may be NULL): float[n], the
 * angle per ray, -1 where the ray did not escape in both. */
int bhExitAngleDiff(const BhRayRecord* a, const BhRayRecord* b, int n, float* anglesOut, BhAngleStats* out);

/* ------------------------------------------------------------------------ */
/* Device memory and stream utilities (plumbing for callers without HIP)    */
/* ------------------------------------------------------------------------ */
bh_error_t  bhDeviceAlloc(void** ptr, size_t bytes);
bh_error_t  bhDeviceFree(void* ptr);
bh_error_t  bhCopyToDevice(void* dst, const void* src, size_t bytes);
bh_error_t  bhCopyToHost(void* dst, const void* src, size_t bytes);
bh_error_t  bhDeviceFill(void* dst, unsigned char value, size_t bytes);
bh_error_t  bhCopyDevice(void* dst, const void* src, size_t bytes);   /* device to device */
bh_error_t  bhDeviceSynchronize(void);
bh_error_t  bhStreamCreate(bh_stream_t* s);
/* Frees what the library holds for a stream (per-stream scratch, batch
 * workers and their queues, batch view records) after waiting for the
 * stream's work; the stream itself stays valid.  bhStreamDestroy does this
 * before destroying the stream. */
bh_error_t  bhStreamRelease(bh_stream_t s);
bh_error_t  bhStreamDestroy(bh_stream_t s);
bh_error_t  bhStreamSynchronize(bh_stream_t s);
const char* bhErrorString(bh_error_t e);
int         bhDeviceCount(void);

#ifdef __cplusplus
}
#endif
#endif /* BH_BHTRACE_H */
