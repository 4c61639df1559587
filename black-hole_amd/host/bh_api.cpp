// ---------------------------------------------------------------------------
// bh_api.cpp -- C ABI entry points of libbhtrace.so.
//
// launchRenderPipeline / launchRenderKernel keep the reference's contract
// (black_hole_kernel.cu:108-163): the parameter block is copied and
// sanitized, work is enqueued asynchronously on the caller's stream, and the
// return value is the launch error state (no synchronization, no throw).
//
// Library-owned device state (documented in the public header):
//   * per-(device, stream) scratch: the per-pixel trace records and the ray
//     queue counter, grown on demand and kept for the life of the process;
//     guarded by a mutex, so concurrent streams never share scratch;
//   * the schedule (workgroup size, RK4 steps per refill chunk, math mode),
//     process-wide, read once per launch;
//   * optional per-stream timing events.
// ---------------------------------------------------------------------------
#include <hip/hip_runtime.h>
#include <atomic>
#include <algorithm>
#include <cstring>
#include <map>
#include <tuple>
#include <mutex>
#include <utility>
#include <vector>
#include "../csrc/bh_internal.h"

namespace bhk {

// FP64 frame constants: RenderParams floats widened, derived values in
// double with the reference's float literals widened (trace.cuh:466-890).
BandK makeBandK(const RenderParams& P, float critBand, float polarBand)
{
    BandK B{};
    const bool kerr = (P.model == BH_KERR || P.model == BH_KERR_NEWMAN);
    B.kerr = kerr ? 1 : 0;
    B.M = P.M;
    B.a = kerr ? (double)P.aSpin : 0.0;
    B.Q = P.Qc;
    B.rPlus = P.rPlus;
    B.rPhoton = P.rPhoton;
    B.dPhi = P.dPhi;
    B.diskOuter = P.diskOuter;
    B.diskTime = P.diskTime;
    B.rCap = (B.rPlus > (double)1e-6f) ? B.rPlus : (double)kSingCutOverM * B.M;
    B.rFloor = std::max(B.rPlus, B.rCap);
    B.escR = (double)120.f * B.M;
    B.capR = B.rCap * (double)1.002f + (double)1e-3f;
    B.rPhotonRefine = (double)1.5f * B.rPhoton;
    B.diskOuterCap = B.diskOuter * (double)1.3f;
    B.sixM = 6.0 * B.M;
    B.kRad = (double)8.0f * B.dPhi;
    B.kAng = (double)2.2f * B.dPhi;
    B.uHor = 1.0 / std::max(B.rCap, (double)1e-6f);
    B.uEscape = 1.0 / B.escR;
    B.invNineM = 1.0 / (9.0 * B.M);
    B.uBudget = 1.0 / std::max((double)1.5f * B.rPhoton, 3.0 * B.M);
    B.hFloor = B.dPhi * (double)0.03f;
    B.maxSteps = P.maxSteps;
    B.disk = P.diskEnabled != 0;
    // spherical models: b_c^2 = r_ph^2 / f(r_ph), r_ph = (3M + sqrt(9M^2 - 8Q^2)) / 2
    const double Q2 = B.Q * B.Q, disc = 9.0 * B.M * B.M - 8.0 * Q2;
    B.bCrit2 = 0.0;
    if (!kerr && disc >= 0.0)
    {
        const double rph = 0.5 * (3.0 * B.M + std::sqrt(disc));
        const double f = 1.0 - 2.0 * B.M / rph + Q2 / (rph * rph);
        if (f > 0.0) B.bCrit2 = rph * rph / f;
    }
    // Kerr models: scan the radial potential over [r+ (1 + 1e-3), 5M]
    const double rp = bhHorizonOuterFP64(B.M, B.a, B.Q);
    B.rScanLo = rp > 0.0 ? rp * 1.001 : (double)kSingCutOverM * B.M;
    B.rScanHi = std::max(5.0 * B.M, B.rScanLo * 1.5);
    // The widths are those that hold every ray within 1e-3 rad of FP64 at
    // steps dPhi <= 0.007 (the 'high' preset).  Coarser steps move FP32
    // traces further from FP64 (more of their adaptive-step decisions
    // differ), so the bands widen with the step: the critical band by
    // s = dPhi / 0.007, the polar band by s^2 (64 uneven 'balanced' views,
    // dPhi 0.012: the unscaled bands leave 7 views over 1e-3 rad).
    const double s = std::max(1.0, (double)P.dPhi / 0.007);
    B.critBand = (double)critBand * s;
    B.polarBand = (double)polarBand * s * s;
    // Away from the critical curve the polar band narrows (measured on the
    // 64 uneven views and the 1080p and 4K frames: near-pole rays with
    // |delta| >= 0.2 s and sin^2 theta_min >= 0.3 polarBand, >= 0.5 s and
    // >= 0.1 polarBand, >= s and >= 0.03 polarBand trace in FP32 within
    // 3.5e-4 rad of FP64).
    // Rays that fall in (delta > 0) carry no exit direction; in FP32 they
    // fall in like in FP64 from delta = 0.005 s on (no outcome differs on
    // those frames), so their critical band is a quarter as wide.
    B.critBandIn = 0.25 * B.critBand;
    // Kerr rays that escape with their polar turning point far from the axis
    // (sin^2 theta_min >= 0.1 s^2, >= 0.3 s^2) trace in FP32 within 2.9e-4
    // rad of FP64 from |delta| = critBand / 2, critBand / 4 on.
    B.critPolarTier[0] = 0.1 * s * s;
    B.critPolarTier[1] = 0.3 * s * s;
    B.polarTier[0] = 0.2 * s;
    B.polarTier[1] = 0.5 * s;
    B.polarTier[2] = 1.0 * s;
    // camera terms of the ray invariants (trace.cuh:213-247, 704-739), in
    // double from the float camera position, with contraction off (this
    // file is compiled with -ffp-contract=off)
    {
        const double a = B.a, M = B.M, Q = B.Q;
        const double ox = P.camPos.x, oy = P.camPos.y, oz = P.camPos.z;
        const double r0 = std::sqrt(ox * ox + oy * oy + oz * oz);
        const double cth0 = std::fmin(std::fmax(oy / r0, -1.0), 1.0);
        const double th0 = std::acos(cth0);
        const double ph0 = std::atan2(-oz, ox);
        double s0 = std::sin(th0);
        if (s0 < (double)1e-4f) s0 = (double)1e-4f;
        const double s02 = s0 * s0;
        const double Sig0 = r0 * r0 + a * a * cth0 * cth0;
        const double Del0 = r0 * r0 - 2.0 * M * r0 + a * a + Q * Q;
        const double m2r0 = 2.0 * M * r0 - Q * Q;
        double gtt = -(1.0 - m2r0 / Sig0);
        if (gtt > (double)-1e-5f) gtt = (double)-1e-5f;
        const double gtp = -a * m2r0 * s02 / Sig0;
        const double gpp = (r0 * r0 + a * a + m2r0 * a * a * s02 / Sig0) * s02;
        B.camR0 = r0;
        B.camCth0 = cth0;
        B.camTh0 = th0;
        B.camPh0 = ph0;
        B.camS = std::sin(th0);
        B.camC = std::cos(th0);
        B.camSp = std::sin(ph0);
        B.camCp = std::cos(ph0);
        B.camS02 = s02;
        B.camSig0 = Sig0;
        B.camDel0 = Del0;
        B.camGtt = gtt;
        B.camGtp = gtp;
        B.camGpp = gpp;
        B.camA = 1.0 / std::sqrt(std::fmax(gpp - gtp * gtp / gtt, (double)1e-8f));
        B.camPt0 = 1.0 / std::sqrt(-gtt);
        B.camPrScale = std::sqrt(Sig0 / std::fmax(Del0, (double)1e-6f));
        B.camPthScale = std::sqrt(Sig0);
    }
    return B;
}

}  // namespace bhk

namespace {

// Batches run their views as independent single-view pipelines on this many
// worker streams, so small views share the machine.  Each worker and the
// auxiliary stream of its pipelines hold a hardware queue of their own (12
// queues): 64 uneven views take 123 ms on 6 workers, 140 ms on 4, 165 ms on 2,
// and 165 ms again on 12, whose 24 queues the command processor time-slices.
constexpr int kBatchStreams = 6;

// FP64 band wavefronts per CU of a persistent batch's band trace.  Each
// holds two FP32 trace wavefronts' registers while it runs, so fewer, busier
// band wavefronts cost the FP32 trace less, as long as the band kernel still
// ends before it.  64 uneven views with sorted band lists and the uniform
// band widths: 91.5 ms at 3 per CU, 89.9 ms at 2, 88.0 ms at 1, 151 ms at
// 1/2 (the band kernel then sets the time).  With the narrowed polar and
// infall bands (a third of the band rays): 87.2 ms at 2, 85.7 ms at 1, the
// band kernel ending at about half of the FP32 trace.
constexpr int kBatchBandWavesPerCU = 1;

// The last view traced into a record buffer (RenderParams less the
// per-frame sample index and disk time) and that buffer: a frame of the same
// view orders its work by those records' step counts.
struct ViewMemo
{
    RenderParams lastView{};
    const BhRayRecord* lastRec = nullptr;
};

// Records of one view of a batch, kept from batch to batch so that a
// repeated view orders its work by its own previous records.
struct BatchSlot
{
    BhRayRecord* rec = nullptr;
    size_t cap = 0;
    ViewMemo memo;
};

// FP64 band pass of a persistent batch: the per-view band descriptors
// (device table, and its pinned host staging for the eager launch), every
// view's band mask and list (slices of one buffer each, in view order), and
// the pass's counters: list lengths [n], chunk offsets [n + 1], per-view
// chunk completions [n], queue head [1].
struct BatchBandBufs
{
    bhk::BandArgs* table = nullptr;
    bhk::BandArgs* hostTable = nullptr;
    unsigned char* mask = nullptr;
    unsigned* list = nullptr;
    unsigned* key = nullptr;      // cost key of each listed ray
    unsigned* sorted = nullptr;   // each view's list, longest rays first
    unsigned* hist = nullptr;     // [viewCap x 512] sort histograms
    unsigned* words = nullptr;
    size_t pixCap = 0;
    int viewCap = 0;
};

// The tile order of a persistent batch: every tile's cost key (all views'
// tiles in the FP32 trace's tile numbering), the order, costliest first,
// and the sort's bucket counts and cursors.
struct BatchTileBufs
{
    unsigned* key = nullptr;      // [cap]
    unsigned* order = nullptr;    // [cap]
    unsigned* hist = nullptr;     // [512]
    size_t cap = 0;
};

void freeBatchTiles(BatchTileBufs& b)
{
    (void)hipFree(b.key);
    (void)hipFree(b.order);
    (void)hipFree(b.hist);
    b = BatchTileBufs{};
}

// Grows the tile buffers to `tiles` tiles (after `st` has finished with the
// old ones).
hipError_t growBatchTiles(BatchTileBufs& b, size_t tiles, hipStream_t st)
{
    if (tiles <= b.cap && b.key) return hipSuccess;
    hipError_t e = st ? hipStreamSynchronize(st) : hipSuccess;
    if (e != hipSuccess) return e;
    const size_t nt = std::max(std::max(tiles, b.cap), (size_t)1);
    freeBatchTiles(b);
    e = hipMalloc(&b.key, sizeof(unsigned) * nt);
    if (e == hipSuccess) e = hipMalloc(&b.order, sizeof(unsigned) * nt);
    if (e == hipSuccess) e = hipMalloc(&b.hist, sizeof(unsigned) * 512);
    if (e != hipSuccess) { freeBatchTiles(b); return e; }
    b.cap = nt;
    return hipSuccess;
}

void freeBatchBand(BatchBandBufs& b)
{
    (void)hipFree(b.table);
    (void)hipHostFree(b.hostTable);
    (void)hipFree(b.mask);
    (void)hipFree(b.list);
    (void)hipFree(b.key);
    (void)hipFree(b.sorted);
    (void)hipFree(b.hist);
    (void)hipFree(b.words);
    b = BatchBandBufs{};
}

// Grows the buffers to n views and `pix` pixels (after `st` has finished
// with the old ones); `pinned`: with a host staging table.
hipError_t growBatchBand(BatchBandBufs& b, int n, size_t pix, bool pinned, hipStream_t st)
{
    if (n <= b.viewCap && pix <= b.pixCap && (!pinned || b.hostTable)) return hipSuccess;
    hipError_t e = st ? hipStreamSynchronize(st) : hipSuccess;
    if (e != hipSuccess) return e;
    const int nv = std::max(n, b.viewCap);
    const size_t np = std::max(std::max(pix, b.pixCap), (size_t)1);
    freeBatchBand(b);
    e = hipMalloc(&b.table, sizeof(bhk::BandArgs) * nv);
    if (e == hipSuccess && pinned) e = hipHostMalloc(&b.hostTable, sizeof(bhk::BandArgs) * nv, hipHostMallocDefault);
    if (e == hipSuccess) e = hipMalloc(&b.mask, np);
    if (e == hipSuccess) e = hipMalloc(&b.list, np * sizeof(unsigned));
    if (e == hipSuccess) e = hipMalloc(&b.key, np * sizeof(unsigned));
    if (e == hipSuccess) e = hipMalloc(&b.sorted, np * sizeof(unsigned));
    if (e == hipSuccess) e = hipMalloc(&b.hist, sizeof(unsigned) * 512 * (size_t)nv);
    if (e == hipSuccess) e = hipMalloc(&b.words, sizeof(unsigned) * (3 * (size_t)nv + 2));
    if (e != hipSuccess) { freeBatchBand(b); return e; }
    b.viewCap = nv;
    b.pixCap = np;
    return hipSuccess;
}

struct Scratch
{
    BhRayRecord* rec = nullptr;
    size_t recCap = 0;     // records
    unsigned* counter = nullptr;
    hipEvent_t ev[4] = {nullptr, nullptr, nullptr, nullptr};
    bool evValid = false;
    // batched launches: view table (device + pinned host staging), per-view
    // counters and stamps; `staged` marks the completion of the last upload
    // out of `hostViews`, which must not be rewritten before it.
    bhk::ViewDesc* views = nullptr;
    bhk::ViewDesc* hostViews = nullptr;
    hipEvent_t staged = nullptr;
    unsigned* viewDone = nullptr;
    unsigned long long* stamps = nullptr;
    int viewCap = 0;
    int lastViews = 0;
    bool lastPersistent = false;          // the last batch was a persistent launch (latency from stamps)
    std::vector<hipEvent_t> viewEvents;   // [0] batch start, [v+1] view v done
    // batched views run on library worker streams forked from this stream
    hipStream_t batchStreams[kBatchStreams] = {};
    hipEvent_t batchDone[kBatchStreams] = {};
    // FP64 band pass: per-pixel mask, band pixel list (count in counter[16])
    unsigned char* bandMask = nullptr;
    unsigned* bandList = nullptr;
    unsigned* bandKey = nullptr;          // [W*H] cost key per listed band ray
    unsigned* bandSorted = nullptr;       // [W*H] band list, longest rays first
    unsigned* bandHist = nullptr;         // [512] band sort: bucket counts, then cursors
    ViewMemo memo;                        // last view traced into `rec`
    std::vector<BatchSlot> slots;         // batched views: records per view of the last batch
    size_t bandCap = 0;
    // tile order of the FP32 tile tracer (longest tiles first): per-tile
    // keys, the order, and 512 counters of the counting sort
    unsigned* tileKey = nullptr;
    unsigned* tileOrder = nullptr;
    unsigned* tileHist = nullptr;
    size_t tileCap = 0;
    int lastOrderTiles = 0;      // tiles ordered in the last frame (0: row-major)
    // surrogate candidates (inputs, pixel list); counts in counter[32] and
    // counter[48] (accepted)
    float* candInputs = nullptr;
    unsigned* candList = nullptr;
    float4* candAcc = nullptr;                   // rays the surrogate answered (pixel, deflection outputs)
    size_t candCap = 0;
    hipEvent_t bandEv[2] = {nullptr, nullptr};   // timing: band pass start / end
    hipEvent_t tileEv[2] = {nullptr, nullptr};   // timing: FP32 tile kernel start / end (its own stream)
    hipEvent_t routeEv[2] = {nullptr, nullptr};  // timing: surrogate inference + routing start / end
    bool routeTimed = false;                     // routeEv hold the last timed frame's route kernel
    bool tileTimed = false;                      // tileEv hold the last timed frame
    // the FP32 tile trace runs on `aux` concurrently with the FP64 band
    // kernel on the caller's stream: fork after classification, join before
    // shading
    hipStream_t aux = nullptr;
    hipEvent_t fork = nullptr, join = nullptr;
    BatchBandBufs bband;         // band pass of persistent batches on this stream
    BatchTileBufs btiles;        // their tile order
    std::vector<RenderParams> lastBatchViews;   // views of the last persistent batch (sample index, disk time 0)
    const BhRayRecord* lastBatchRec = nullptr;  // and the records it traced into
};

std::mutex g_mu;
std::map<std::pair<int, hipStream_t>, Scratch> g_scratch;
// wgSize, stepsPerChunk, raysPerRefill, mathMode, diskCoop.  One-wavefront
// workgroups: a resident FP64 band wavefront then takes exactly one
// wavefront slot from the tile grid (a 4-wavefront workgroup cannot start
// on a CU whose free slots do not take all four).
BhSchedule g_sched = {64, 4096, 64, 1, 2};
BhPrecision g_prec = {1, 4e-2f, 1e-3f};
constexpr int kBandCountSlot = 16;            // counter[16]: band list length (own 64-byte line)
constexpr int kCandCountSlot = 32;            // counter[32]: surrogate candidates
constexpr int kAcceptSlot = 48;               // counter[48]: rays answered by the surrogate
constexpr int kBandQueueSlot = 64;            // counter[64]: band-trace queue head
constexpr size_t kCounterBytes = bhk::kFrameCounterWords * sizeof(unsigned);
// learned surrogate attached to the render path (null: off)
struct SurrogateRoute
{
    BhSurrogate* s = nullptr;
    float sigmaMax = 1e-4f;
    float diskProbMax = 1e-3f;
} g_sur;
std::atomic<int> g_timers{0};
std::pair<int, hipStream_t> g_lastTimed{-1, nullptr};
int g_numCU[64] = {0};

int numCUs(int dev)
{
    if (dev < 0 || dev >= 64) return 256;
    if (g_numCU[dev] == 0)
    {
        hipDeviceProp_t prop;
        if (hipGetDeviceProperties(&prop, dev) == hipSuccess) g_numCU[dev] = prop.multiProcessorCount;
        else g_numCU[dev] = 256;
    }
    return g_numCU[dev];
}

// A stream with a hardware queue of its own.  HIP pools a few hardware
// queues per process and shares them among streams in an order set by every
// stream created before, so two batch streams can land on one queue and run
// one after the other.  A stream created with a CU mask gets a queue outside
// that pool; the mask enables every CU, so dispatch is unchanged.  HIP creates
// it as a blocking stream: it also orders against the legacy null stream.
hipError_t createQueueStream(int dev, hipStream_t* s)
{
    const int n = numCUs(dev);
    std::vector<uint32_t> mask((size_t)(n + 31) / 32, 0xffffffffu);
    if (n % 32) mask.back() = (1u << (n % 32)) - 1u;
    if (hipExtStreamCreateWithCUMask(s, (uint32_t)mask.size(), mask.data()) == hipSuccess) return hipSuccess;
    // without CU-mask support: a pooled non-blocking stream (same results;
    // the runtime may then share its hardware queue with other streams)
    (void)hipGetLastError();
    *s = nullptr;
    return hipStreamCreateWithFlags(s, hipStreamNonBlocking);
}

// Scratch for the current device and `st`, holding at least `nRec` records.
hipError_t getScratch(hipStream_t st, size_t nRec, Scratch** out)
{
    int dev = 0;
    hipError_t e = hipGetDevice(&dev);
    if (e != hipSuccess) return e;
    std::lock_guard<std::mutex> lk(g_mu);
    Scratch& S = g_scratch[{dev, st}];
    if (!S.counter)
    {
        e = hipMalloc(&S.counter, kCounterBytes);
        if (e != hipSuccess) { S.counter = nullptr; return e; }
        e = hipMemset(S.counter, 0, kCounterBytes);
        if (e != hipSuccess) return e;
    }
    if (nRec > S.recCap)
    {
        if (S.rec)
        {
            // the old buffer may still be read by work queued on this stream
            e = hipStreamSynchronize(st);
            if (e != hipSuccess) return e;
            (void)hipFree(S.rec);
            S.rec = nullptr;
            S.recCap = 0;
        }
        e = hipMalloc(&S.rec, nRec * sizeof(BhRayRecord));
        if (e != hipSuccess) { S.rec = nullptr; return e; }
        S.recCap = nRec;
    }
    *out = &S;
    return hipSuccess;
}

hipError_t getBatchScratch(Scratch* S, int nViews, hipStream_t st)
{
    std::lock_guard<std::mutex> lk(g_mu);
    hipError_t e = hipSuccess;
    if (!S->staged)
    {
        e = hipEventCreateWithFlags(&S->staged, hipEventDisableTiming);
        if (e != hipSuccess) { S->staged = nullptr; return e; }
    }
    if (nViews <= S->viewCap) return hipSuccess;
    // the old tables may still be in use by queued work
    e = hipStreamSynchronize(st);
    if (e != hipSuccess) return e;
    (void)hipFree(S->views);
    (void)hipHostFree(S->hostViews);
    (void)hipFree(S->viewDone);
    (void)hipFree(S->stamps);
    S->views = nullptr; S->hostViews = nullptr; S->viewDone = nullptr; S->stamps = nullptr; S->viewCap = 0;
    e = hipMalloc(&S->views, sizeof(bhk::ViewDesc) * nViews);
    if (e == hipSuccess) e = hipHostMalloc(&S->hostViews, sizeof(bhk::ViewDesc) * nViews, hipHostMallocDefault);
    if (e == hipSuccess) e = hipMalloc(&S->viewDone, sizeof(unsigned) * nViews);
    if (e == hipSuccess) e = hipMalloc(&S->stamps, sizeof(unsigned long long) * (nViews + 1));
    if (e == hipSuccess) S->viewCap = nViews;
    return e;
}

BhSchedule currentSchedule()
{
    std::lock_guard<std::mutex> lk(g_mu);
    return g_sched;
}

BhPrecision currentPrecision()
{
    std::lock_guard<std::mutex> lk(g_mu);
    return g_prec;
}

hipError_t getCandScratch(Scratch* S, size_t n, hipStream_t st)
{
    std::lock_guard<std::mutex> lk(g_mu);
    if (n <= S->candCap) return hipSuccess;
    hipError_t e = hipStreamSynchronize(st);
    if (e != hipSuccess) return e;
    (void)hipFree(S->candInputs);
    (void)hipFree(S->candList);
    (void)hipFree(S->candAcc);
    S->candInputs = nullptr; S->candList = nullptr; S->candAcc = nullptr; S->candCap = 0;
    e = hipMalloc(&S->candInputs, n * 6 * sizeof(float));
    if (e == hipSuccess) e = hipMalloc(&S->candList, n * sizeof(unsigned));
    if (e == hipSuccess) e = hipMalloc(&S->candAcc, n * sizeof(float4));
    if (e != hipSuccess) return e;
    S->candCap = n;
    return hipSuccess;
}

SurrogateRoute currentSurrogate()
{
    std::lock_guard<std::mutex> lk(g_mu);
    return g_sur;
}

hipError_t getTileScratch(Scratch* S, size_t nTiles, hipStream_t st)
{
    std::lock_guard<std::mutex> lk(g_mu);
    if (nTiles <= S->tileCap) return hipSuccess;
    hipError_t e = hipSuccess;
    if (S->tileKey)
    {
        e = hipStreamSynchronize(st);
        if (e != hipSuccess) return e;
        (void)hipFree(S->tileKey);
        (void)hipFree(S->tileOrder);
        (void)hipFree(S->tileHist);
        S->tileKey = nullptr; S->tileOrder = nullptr; S->tileHist = nullptr; S->tileCap = 0;
    }
    e = hipMalloc(&S->tileKey, nTiles * sizeof(unsigned));
    if (e == hipSuccess) e = hipMalloc(&S->tileOrder, nTiles * sizeof(unsigned));
    if (e == hipSuccess) e = hipMalloc(&S->tileHist, 512 * sizeof(unsigned));
    if (e != hipSuccess) return e;
    S->tileCap = nTiles;
    return hipSuccess;
}

hipError_t getBandScratch(Scratch* S, size_t n, hipStream_t st)
{
    std::lock_guard<std::mutex> lk(g_mu);
    if (n <= S->bandCap) return hipSuccess;
    hipError_t e = hipSuccess;
    if (S->bandMask)
    {
        e = hipStreamSynchronize(st);
        if (e != hipSuccess) return e;
        (void)hipFree(S->bandMask);
        (void)hipFree(S->bandList);
        (void)hipFree(S->bandKey);
        (void)hipFree(S->bandSorted);
        (void)hipFree(S->bandHist);
        S->bandMask = nullptr; S->bandList = nullptr; S->bandKey = nullptr; S->bandSorted = nullptr;
        S->bandHist = nullptr;
        S->bandCap = 0;
    }
    e = hipMalloc(&S->bandMask, n);
    if (e == hipSuccess) e = hipMalloc(&S->bandList, n * sizeof(unsigned));
    if (e == hipSuccess) e = hipMalloc(&S->bandKey, n * sizeof(unsigned));
    if (e == hipSuccess) e = hipMalloc(&S->bandSorted, n * sizeof(unsigned));
    if (e == hipSuccess) e = hipMalloc(&S->bandHist, 512 * sizeof(unsigned));
    if (e != hipSuccess) return e;
    S->bandCap = n;
    if (!S->aux)
    {
        e = hipStreamCreateWithFlags(&S->aux, hipStreamNonBlocking);
        if (e == hipSuccess) e = hipEventCreateWithFlags(&S->fork, hipEventDisableTiming);
        if (e == hipSuccess) e = hipEventCreateWithFlags(&S->join, hipEventDisableTiming);
        if (e != hipSuccess) return e;
    }
    return hipSuccess;
}


inline bool isKerr(int model) { return model == BH_KERR || model == BH_KERR_NEWMAN; }

// A camera inside the disk's emitting layer: its own position passes the
// volumetric sampler's geometric gates (r_in <= rc <= r_out and |y| <= 2.6 H
// with H = (0.040 + 0.015 rc / r_out) rc; trace.cuh:298-310).  Every camera
// ray then starts inside the disk, so none can be answered without disk
// light, and the learned surrogate is not consulted for the frame.
bool cameraInDiskLayer(const RenderParams& P)
{
    if (!P.diskEnabled) return false;
    const float x = P.camPos.x, y = P.camPos.y, z = P.camPos.z;
    const float rc = std::sqrt(x * x + z * z);
    const float rin = std::fmax(P.diskInner, 1e-4f);
    if (rc < rin || rc > P.diskOuter) return false;
    const float H = (0.040f + 0.015f * (rc * (1.f / std::fmax(P.diskOuter, rin)))) * rc;
    return std::fabs(y / std::fmax(H, 1e-4f)) <= 2.6f;
}

float2 jitterOf(int accumMode, int sampleIndex)
{
    // R2 sequence (post_process.cuh:47-53), evaluated without contraction
    if (accumMode == 0 || sampleIndex <= 0) return make_float2(0.5f, 0.5f);
    const float n = (float)sampleIndex;
    const float x = 0.7548776662f * n + 0.5f, y = 0.5698402909f * n + 0.5f;
    return make_float2(x - std::floor(x), y - std::floor(y));
}

// Memo of the views traced by the fixed-buffer path, one per device (that
// path keeps its records in the library's static scratch of that device).
ViewMemo g_fixedMemo[64];

// Enqueues the trace of one frame into `rec`.  Buffers come from the
// stream's scratch `Sc`, or, with `fixed`, from caller-owned memory: then
// nothing is allocated or created, the surrogate is not applied, no timers
// are recorded, and the FP64 band and FP32 tile traces run one after the
// other on `st` (no auxiliary stream).  `memo` holds the last view traced
// into `rec` (null: the stream's own).
hipError_t enqueueTrace(const RenderParams& P, BhRayRecord* rec, Scratch* Sc, const BhSchedule& scIn,
                        hipStream_t st, const int* dynIndex = nullptr, bool allowTimers = true,
                        ViewMemo* memo = nullptr, const bhk::FrameBufs* fixed = nullptr)
{
    unsigned* counter = fixed ? fixed->counter : Sc->counter;
    const BhPrecision pr = currentPrecision();
    const bool band = pr.fp64Band != 0 && P.width > 0 && P.height > 0;
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess) dev = 0;
    if (!memo) memo = fixed ? &g_fixedMemo[(dev >= 0 && dev < 64) ? dev : 0] : &Sc->memo;
    if (Sc) Sc->routeTimed = false;
    const int kerr = isKerr(P.model) ? 1 : 0;
    BhSchedule sc = scIn;
    const int tiles = sc.raysPerRefill >= 64 ? 1 : 0;
    if (sc.diskCoop == 2)
    {
        // automatic: cooperative disk sampling when the whole-tile frame has
        // no more tiles than the tile grid has wavefronts.  Each wavefront
        // then traces one tile, the frame is as long as its slowest tile,
        // and lanes whose rays have ended take over the disk sub-samples of
        // the rays still running (the same operations in the same order).
        const int occ0 = sc.mathMode ? bhk::traceOccupancyFast(kerr, sc.wgSize, 0, tiles)
                                     : bhk::traceOccupancyExact(kerr, sc.wgSize, 0, tiles);
        const long long waves = (long long)occ0 * numCUs(dev) * (sc.wgSize / 64);
        const long long nTiles = (long long)((P.width + 7) / 8) * ((P.height + 7) / 8);
        sc.diskCoop = (tiles && nTiles <= waves) ? 1 : 0;
    }
    const int occ = sc.mathMode ? bhk::traceOccupancyFast(kerr, sc.wgSize, sc.diskCoop, tiles)
                                : bhk::traceOccupancyExact(kerr, sc.wgSize, sc.diskCoop, tiles);
    const long long want = (long long)occ * numCUs(dev);
    // Same view as the last frame traced into `rec` (RenderParams apart from
    // the sample index and disk time)?  Then its records predict this
    // frame's work.  Under graph capture always: replays repeat this view.
    // The prediction only orders the work; any prediction gives the same frame.
    bool repeatView, capturing;
    {
        RenderParams v = P;
        v.sampleIndex = 0;
        v.diskTime = 0.f;
        hipStreamCaptureStatus cs = hipStreamCaptureStatusNone;
        capturing = hipStreamIsCapturing(st, &cs) == hipSuccess && cs == hipStreamCaptureStatusActive;
        repeatView = capturing || (memo->lastRec == rec && std::memcmp(&v, &memo->lastView, sizeof(v)) == 0);
        memo->lastView = v;
        memo->lastRec = rec;
    }
    const SurrogateRoute sur = (fixed || cameraInDiskLayer(P)) ? SurrogateRoute{} : currentSurrogate();
    bhk::BandArgs BA;
    BA.tileKey = nullptr;
    unsigned char* bandMask = nullptr;
    unsigned* tileOrder = nullptr;
    if (band)
    {
        const size_t n = (size_t)P.width * (size_t)P.height;
        const long long nT = (long long)((P.width + 7) / 8) * ((P.height + 7) / 8);
        hipError_t e = hipSuccess;
        unsigned* tileHist = nullptr;
        if (fixed)
        {
            bandMask = fixed->bandMask;
            BA.list = fixed->bandList;
            BA.listKey = fixed->bandKey;
            BA.sorted = fixed->bandSorted;
            BA.listHist = fixed->bandHist;
            if (tiles) { BA.tileKey = fixed->tileKey; tileOrder = fixed->tileOrder; tileHist = fixed->tileHist; }
        }
        else
        {
            e = getBandScratch(Sc, n, st);
            if (e != hipSuccess) return e;
            bandMask = Sc->bandMask;
            BA.list = Sc->bandList;
            BA.listKey = Sc->bandKey;
            BA.sorted = Sc->bandSorted;
            BA.listHist = Sc->bandHist;
            Sc->lastOrderTiles = 0;
            if (tiles)
            {
                // whole-tile schedule: the queue hands out the costliest tiles
                // first, by the classification's cost keys
                e = getTileScratch(Sc, (size_t)nT, st);
                if (e != hipSuccess) return e;
                BA.tileKey = Sc->tileKey;
                tileOrder = Sc->tileOrder;
                tileHist = Sc->tileHist;
                Sc->lastOrderTiles = (int)nT;
            }
        }
        BA.P = P;
        BA.B = bhk::makeBandK(P, pr.critBand, pr.polarBand);
        BA.D = bhk::makeDisk(P, BA.B.kerr ? P.aSpin : 0.f, P.Qc);
        BA.rec = rec;
        BA.mask = bandMask;
        BA.prevRec = repeatView ? rec : nullptr;
        BA.count = counter + kBandCountSlot;
        BA.queue = counter + kBandQueueSlot;
        BA.tilesX = (P.width + 7) / 8;
        BA.jitter = jitterOf(P.accumMode, P.sampleIndex);
        BA.dynIndex = dynIndex;
        BA.candInputs = nullptr;
        BA.candList = nullptr;
        BA.candCount = counter + kCandCountSlot;
        BA.psiMax = (float)(65.0 * 3.14159265358979323846 / 180.0);
        if (sur.s)
        {
            e = getCandScratch(Sc, n, st);
            if (e != hipSuccess) return e;
            BA.candInputs = Sc->candInputs;
            BA.candList = Sc->candList;
        }
        const int nTiles = (int)nT;
        // every counter word of the frame (the tile queue head included) and
        // the tile keys in one launch
        e = bhk::launchFrameReset(counter, bhk::kFrameCounterWords, BA.tileKey, nTiles, tileHist, BA.listHist, st);
        if (e == hipSuccess) e = bhk::launchBandClassify(BA, st);
        if (e == hipSuccess) e = bhk::launchBandSort(BA, st);
        if (e == hipSuccess && BA.tileKey) e = bhk::launchTileOrder(BA.tileKey, nTiles, tileHist, tileOrder, st);
        if (e != hipSuccess) return e;
        if (sur.s)
        {
            // network inference + routing: accepted rays get their record and
            // mask 2 before the RK4 tracers start
            const bhk::SurrogateNet net = bhk::surrogateNet(sur.s);
            bhk::RouteArgs R;
            R.P = P;
            R.jitter = BA.jitter;
            R.dynIndex = dynIndex;
            R.inputs = Sc->candInputs;
            R.list = Sc->candList;
            R.count = counter + kCandCountSlot;
            R.rows = 0;
            R.outRows = nullptr;
            R.W = net.W;
            R.Pf = net.P;
            for (int l = 0; l < bhk::kMlpLayers; ++l) { R.wOff[l] = net.wOff[l]; R.bOff[l] = net.bOff[l]; }
            R.sigmaMax = sur.sigmaMax;
            // without a disk no ray can gather disk light: the disk-hit head
            // (trained with the disk on) is not consulted
            R.diskProbMax = P.diskEnabled ? sur.diskProbMax : 2.f;
            R.rec = rec;
            R.mask = bandMask;
            R.accepted = counter + kAcceptSlot;
            R.accList = Sc->candAcc;
            const bool rt = allowTimers && g_timers.load() != 0;
            if (rt && !Sc->routeEv[0])
            {
                if (hipEventCreate(&Sc->routeEv[0]) != hipSuccess || hipEventCreate(&Sc->routeEv[1]) != hipSuccess)
                    return hipErrorOutOfMemory;
            }
            Sc->routeTimed = rt;
            if (rt) (void)hipEventRecord(Sc->routeEv[0], st);
            e = bhk::launchSurrogateRoute(R, (int)n, numCUs(dev), st);
            if (rt && e == hipSuccess) (void)hipEventRecord(Sc->routeEv[1], st);
            if (e != hipSuccess) return e;
        }
    }
    bool orderedNoBand = false;
    const long long nTilesFrame = (long long)((P.width + 7) / 8) * ((P.height + 7) / 8);
    // (under graph capture nothing may be allocated: without the buffers the
    // captured frame keeps the row-major order)
    if (!band && !fixed && tiles && repeatView && P.width > 0 && P.height > 0 &&
        (!capturing || Sc->tileCap >= (size_t)nTilesFrame))
    {
        // band pass off (the reference's arithmetic): no classification, but
        // a view traced again still orders its tiles longest first, by the
        // step counts its previous frame left in the records
        const long long nT = nTilesFrame;
        hipError_t e = getTileScratch(Sc, (size_t)nT, st);
        if (e == hipSuccess) e = bhk::launchFrameReset(counter, bhk::kFrameCounterWords, Sc->tileKey, (int)nT, Sc->tileHist,
                                                      nullptr, st);
        if (e == hipSuccess) e = bhk::launchTileKeys(rec, P.width, P.height, (P.width + 7) / 8, Sc->tileKey, st);
        if (e == hipSuccess) e = bhk::launchTileOrder(Sc->tileKey, (int)nT, Sc->tileHist, Sc->tileOrder, st);
        if (e != hipSuccess) return e;
        tileOrder = Sc->tileOrder;
        Sc->lastOrderTiles = (int)nT;
        orderedNoBand = true;
    }
    else if (!band && Sc)
        Sc->lastOrderTiles = 0;
    bhk::TraceArgs A;
    const BhSchedule& scT = sc;
    A.bandMask = bandMask;
    A.tileOrder = (BA.tileKey || orderedNoBand) ? tileOrder : nullptr;
    A.P = P;
    A.rec = rec;
    A.counter = counter;
    A.tilesX = (P.width + 7) / 8;
    const int tilesY = (P.height + 7) / 8;
    A.totalIds = A.tilesX * tilesY * 64;
    A.chunk = scT.stepsPerChunk;
    A.refillMin = scT.raysPerRefill;
    A.coop = scT.diskCoop;
    A.jitter = jitterOf(P.accumMode, P.sampleIndex);
    A.dynIndex = dynIndex;
    long long maxUseful = ((long long)A.totalIds + scT.wgSize - 1) / scT.wgSize;
    const long long wantT = want;
    bhk::LaunchShape S;
    S.wgSize = scT.wgSize;
    S.blocks = (int)(wantT < maxUseful ? wantT : maxUseful);
    if (S.blocks < 1) S.blocks = 1;
    // with the band, or the tiles ordered without it: counters reset above
    hipError_t e = (band || orderedNoBand) ? hipSuccess : bhk::launchResetCounter(counter, st);
    if (e != hipSuccess) return e;
    const int bandBlocks = 4 * numCUs(dev);
    if (fixed)
    {
        if (band)
        {
            e = sc.mathMode ? bhk::launchBandTraceFast(BA, bandBlocks, st)
                            : bhk::launchBandTraceExact(BA, bandBlocks, st);
            if (e != hipSuccess) return e;
        }
        return sc.mathMode ? bhk::launchTraceFast(A, S, st) : bhk::launchTraceExact(A, S, st);
    }
    const bool timed = allowTimers && g_timers.load() != 0;
    if (timed && !Sc->tileEv[0])
    {
        if (hipEventCreate(&Sc->tileEv[0]) != hipSuccess || hipEventCreate(&Sc->tileEv[1]) != hipSuccess)
            return hipErrorOutOfMemory;
    }
    Sc->tileTimed = timed;
    if (!band)
    {
        if (timed) (void)hipEventRecord(Sc->tileEv[0], st);
        e = sc.mathMode ? bhk::launchTraceFast(A, S, st) : bhk::launchTraceExact(A, S, st);
        if (timed && e == hipSuccess) (void)hipEventRecord(Sc->tileEv[1], st);
        return e;
    }
    // The band kernel goes first on the caller's stream so its few, long
    // FP64 wavefronts are resident before the persistent FP32 tile grid
    // fills the machine; the tile kernel follows on the auxiliary stream.
    if (timed && !Sc->bandEv[0])
    {
        if (hipEventCreate(&Sc->bandEv[0]) != hipSuccess || hipEventCreate(&Sc->bandEv[1]) != hipSuccess)
            return hipErrorOutOfMemory;
    }
    e = hipEventRecord(Sc->fork, st);
    if (e != hipSuccess) return e;
    if (timed) (void)hipEventRecord(Sc->bandEv[0], st);
    e = sc.mathMode ? bhk::launchBandTraceFast(BA, bandBlocks, st) : bhk::launchBandTraceExact(BA, bandBlocks, st);
    if (e != hipSuccess) return e;
    if (timed) (void)hipEventRecord(Sc->bandEv[1], st);
    e = hipStreamWaitEvent(Sc->aux, Sc->fork, 0);
    if (e != hipSuccess) return e;
    if (timed) (void)hipEventRecord(Sc->tileEv[0], Sc->aux);
    e = sc.mathMode ? bhk::launchTraceFast(A, S, Sc->aux) : bhk::launchTraceExact(A, S, Sc->aux);
    if (timed && e == hipSuccess) (void)hipEventRecord(Sc->tileEv[1], Sc->aux);
    // the forked auxiliary stream is joined back also when the launch failed
    hipError_t j = hipEventRecord(Sc->join, Sc->aux);
    if (j == hipSuccess) j = hipStreamWaitEvent(st, Sc->join, 0);
    return e != hipSuccess ? e : j;
}

void recordEvent(Scratch* S, int i, hipStream_t st)
{
    if (!S->evValid)
    {
        for (auto& ev : S->ev)
            if (hipEventCreate(&ev) != hipSuccess) return;
        S->evValid = true;
    }
    (void)hipEventRecord(S->ev[i], st);
}

// Autotune cache: schedule chosen per workload class (renderPipeline looks
// a frame's class up here; frames of untuned classes use g_sched).
struct TuneKey
{
    int kerr, w, h, disk, maxSteps, band, math;
    bool operator<(const TuneKey& o) const
    {
        return std::tie(kerr, w, h, disk, maxSteps, band, math) <
               std::tie(o.kerr, o.w, o.h, o.disk, o.maxSteps, o.band, o.math);
    }
};
std::map<TuneKey, std::pair<BhSchedule, float>> g_tuned;

TuneKey tuneKeyOf(const RenderParams& P, int band, int math)
{
    return TuneKey{isKerr(P.model) ? 1 : 0, P.width, P.height, P.diskEnabled, P.maxSteps, band, math};
}

// Schedule of a (sanitized) frame: the autotuner's choice for its class
// under the current precision and math mode, else the process default.
BhSchedule scheduleFor(const RenderParams& P)
{
    std::lock_guard<std::mutex> lk(g_mu);
    auto it = g_tuned.find(tuneKeyOf(P, g_prec.fp64Band, g_sched.mathMode));
    return it != g_tuned.end() ? it->second.first : g_sched;
}

// Records of one frame: the stream's scratch, or a batch view's own slot
// (grown on demand; the stream must be idle on the old buffer first).
hipError_t slotRecords(BatchSlot* slot, size_t n, hipStream_t st)
{
    if (n <= slot->cap) return hipSuccess;
    if (slot->rec)
    {
        hipError_t e = hipStreamSynchronize(st);
        if (e != hipSuccess) return e;
        (void)hipFree(slot->rec);
        slot->rec = nullptr;
        slot->cap = 0;
    }
    hipError_t e = hipMalloc(&slot->rec, n * sizeof(BhRayRecord));
    if (e != hipSuccess) { slot->rec = nullptr; return e; }
    slot->cap = n;
    slot->memo = ViewMemo{};
    return hipSuccess;
}

hipError_t renderPipeline(uchar4* out, float4* accum, float4* bloomA, float4* bloomB, const RenderParams& p,
                          hipStream_t st, bool singlePass, const int* dynIndex = nullptr, bool allowTimers = true,
                          BatchSlot* slot = nullptr)
{
    RenderParams P = p;
    bhSanitizeRenderParams(&P);
    if (singlePass) P.bloomEnabled = 0;
    if (!out || (!singlePass && !accum)) return hipErrorInvalidValue;
    if (!singlePass && P.bloomEnabled && (!bloomA || !bloomB)) return hipErrorInvalidValue;
    const size_t n = (size_t)P.width * (size_t)P.height;
    Scratch* S = nullptr;
    hipError_t e = getScratch(st, slot ? 0 : n, &S);
    if (e != hipSuccess) return e;
    BhRayRecord* rec = S->rec;
    ViewMemo* memo = nullptr;
    if (slot)
    {
        e = slotRecords(slot, n, st);
        if (e != hipSuccess) return e;
        rec = slot->rec;
        memo = &slot->memo;
    }
    const BhSchedule sc = scheduleFor(P);
    const bool timed = allowTimers && g_timers.load() != 0;
    if (timed) recordEvent(S, 0, st);
    e = enqueueTrace(P, rec, S, sc, st, dynIndex, allowTimers, memo);
    if (e != hipSuccess) return e;
    if (timed) recordEvent(S, 1, st);
    if (singlePass)
    {
        e = bhk::launchShadeFinalize(P, rec, out, st);
        if (timed) { recordEvent(S, 2, st); recordEvent(S, 3, st); }
    }
    else
    {
        e = bhk::launchShadeBlend(P, rec, accum, dynIndex, st);
        if (e != hipSuccess) return e;
        if (timed) recordEvent(S, 2, st);
        // bloom + composite in one launch; bloomA/bloomB stay the caller's
        // scratch (validated above, as the reference requires them)
        e = bhk::launchPostFused(P, accum, out, st);
        if (timed) recordEvent(S, 3, st);
    }
    if (timed)
    {
        int dev = 0;
        if (hipGetDevice(&dev) != hipSuccess) dev = 0;
        std::lock_guard<std::mutex> lk(g_mu);
        g_lastTimed = {dev, st};
    }
    if (e != hipSuccess) return e;
    return hipGetLastError();
}

}  // namespace

namespace bhk {

hipError_t renderPipelineFixed(uchar4* out, float4* accum, float4* bloomA, float4* bloomB, const RenderParams& p,
                               hipStream_t st, const FrameBufs& fb)
{
    RenderParams P = p;
    bhSanitizeRenderParams(&P);
    if (!out || !accum) return hipErrorInvalidValue;
    if (P.bloomEnabled && (!bloomA || !bloomB)) return hipErrorInvalidValue;
    hipError_t e = enqueueTrace(P, fb.rec, nullptr, scheduleFor(P), st, nullptr, false, nullptr, &fb);
    if (e == hipSuccess) e = launchShadeBlend(P, fb.rec, accum, nullptr, st);
    if (e == hipSuccess) e = launchPostFused(P, accum, out, st);
    if (e != hipSuccess) return e;
    return hipGetLastError();
}

}  // namespace bhk

namespace {

// The persistent multi-view launch: one trace kernel over a device queue of
// (view, tile) items of every view (traceBatchKernel), then batched shade,
// bloom and composite kernels over a device table of the views.  With the
// FP64 band pass on, every view is classified first and one kernel traces
// the band rays of all views (traceBandBatchKernel) beside the FP32 trace;
// frames equal launchRenderPipeline's under the same precision.  (Its first GPU runs
// hung: the queue loop ended with a lane-0 block that the compiler merged
// with the next iteration's queue atomic, see traceBatchKernel; the loop was
// restructured.)
constexpr bool kPersistentBatchEnabled = true;

// The view table of a persistent batch (records not yet assigned): per view
// the sanitized params, frame constants, buffers and its range of queue
// items; totals for the launch shapes.
struct BatchTable
{
    std::vector<bhk::ViewDesc> D;
    size_t totalPix = 0;
    int tiles = 0, maxW = 1, maxH = 1;
};

hipError_t buildBatchTable(int n, const RenderParams* views, uchar4* const* outs, float4* const* accums,
                           float4* const* bloomAs, float4* const* bloomBs, BatchTable& T)
{
    if (n < 1 || !views || !outs || !accums) return hipErrorInvalidValue;
    T.D.assign((size_t)n, bhk::ViewDesc{});
    for (int v = 0; v < n; ++v)
    {
        RenderParams P = views[v];
        bhSanitizeRenderParams(&P);
        if (!outs[v] || !accums[v]) return hipErrorInvalidValue;
        if (P.bloomEnabled && (!bloomAs || !bloomBs || !bloomAs[v] || !bloomBs[v])) return hipErrorInvalidValue;
        bhk::ViewDesc& d = T.D[(size_t)v];
        std::memset(&d, 0, sizeof(d));
        d.P = P;
        d.K = bhk::makeFrame(P);
        d.out = outs[v];
        d.accum = accums[v];
        d.bloomA = bloomAs ? bloomAs[v] : nullptr;
        d.bloomB = bloomBs ? bloomBs[v] : nullptr;
        d.tilesX = (P.width + 7) / 8;
        d.tileBase = T.tiles;
        d.tileCount = d.tilesX * ((P.height + 7) / 8);
        d.jitter = jitterOf(P.accumMode, P.sampleIndex);
        T.tiles += d.tileCount;
        T.totalPix += (size_t)P.width * P.height;
        T.maxW = P.width > T.maxW ? P.width : T.maxW;
        T.maxH = P.height > T.maxH ? P.height : T.maxH;
    }
    return hipSuccess;
}

void assignBatchRecords(BatchTable& T, BhRayRecord* rec)
{
    size_t off = 0;
    for (auto& d : T.D)
    {
        d.rec = rec + off;
        off += (size_t)d.P.width * d.P.height;
    }
}

// The FP64 band pass of a persistent batch, ready to enqueue: each view's
// classification arguments (by value, as for a single view), the batched
// band trace's arguments, and the auxiliary stream the FP32 batch trace
// forks onto.
struct BatchBandPlan
{
    std::vector<bhk::BandArgs> views;
    bhk::BatchBandArgs A{};
    hipStream_t aux = nullptr;
    hipEvent_t fork = nullptr, join = nullptr;
    int blocks = 1;
    unsigned* order = nullptr;  // the batch's tile order, written after the classification
};

// Band descriptors of every view of T (records assigned) in the buffers b,
// with the band widths of `pr` (makeBandK, as for a single view); sets each
// view's band mask in T.  Each view's band list is sorted longest rays
// first, by keys from the step counts its records hold when the batch
// repeats (`repeat`), else from the band statistics.  The device table
// b.table is not uploaded here.
void planBatchBand(BatchTable& T, const BhPrecision& pr, const BatchBandBufs& b, const BatchTileBufs& tb, bool repeat,
                   BatchBandPlan& plan)
{
    const int n = (int)T.D.size();
    plan.views.assign((size_t)n, bhk::BandArgs{});
    unsigned* w = b.words;
    size_t off = 0;
    unsigned blocks = 0;
    for (int v = 0; v < n; ++v)
    {
        bhk::ViewDesc& d = T.D[(size_t)v];
        bhk::BandArgs& a = plan.views[(size_t)v];
        std::memset(&a, 0, sizeof(a));
        a.P = d.P;
        a.B = bhk::makeBandK(d.P, pr.critBand, pr.polarBand);
        a.D = bhk::makeDisk(d.P, a.B.kerr ? d.P.aSpin : 0.f, d.P.Qc);
        a.rec = d.rec;
        a.mask = b.mask + off;
        a.list = b.list + off;
        a.listKey = b.key + off;
        a.sorted = b.sorted + off;
        a.listHist = b.hist + (size_t)512 * v;
        a.prevRec = repeat ? d.rec : nullptr;
        a.count = w + v;
        a.tilesX = d.tilesX;
        a.jitter = d.jitter;
        a.classifyBase = blocks;
        a.tileKey = tb.key + d.tileBase;
        d.bandMask = a.mask;
        const size_t px = (size_t)d.P.width * d.P.height;
        blocks += (unsigned)((px + 255) / 256);
        off += px;
    }
    plan.A.classifyBlocks = blocks;
    plan.A.tileKey = tb.key;
    plan.A.nTiles = T.tiles;
    plan.A.tileHist = tb.hist;
    plan.order = tb.order;
    plan.A.views = b.table;
    plan.A.nViews = n;
    plan.A.counts = w;
    plan.A.chunkBase = w + n;
    plan.A.bandDone = w + 2 * n + 1;
    plan.A.queue = w + 3 * n + 1;
    plan.A.hist = b.hist;
}

// The kernels of a persistent batch on `st`: counters reset, the trace over
// one queue of (view, tile) items, shade + accumulation, bloom + composite.
// With `band`: every view classified, then the FP64 band rays of all views
// traced on `st` beside the FP32 batch trace on the auxiliary stream (forked
// after the classifications, joined before shading), as the single-view
// pipeline does.  Without it, `stepOrder` (a batch that repeats the
// previous one into the same records): the tiles ordered costliest first by
// the step counts those records hold.  No memset and no copy: capturable as
// a graph.
hipError_t enqueueBatchKernels(const bhk::BatchArgs& A, int maxW, int maxH, int wgSize, int math, hipStream_t st,
                               Scratch* timers, const BatchBandPlan* band = nullptr,
                               const BatchTileBufs* stepOrder = nullptr)
{
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess) dev = 0;
    const int occ = math ? bhk::traceBatchOccupancyFast(wgSize) : bhk::traceBatchOccupancyExact(wgSize);
    bhk::LaunchShape L;
    L.wgSize = wgSize;
    const long long want = (long long)occ * numCUs(dev), useful = ((long long)A.totalTiles * 64 + wgSize - 1) / wgSize;
    L.blocks = (int)(want < useful ? want : useful);
    if (L.blocks < 1) L.blocks = 1;
    if (timers) recordEvent(timers, 0, st);
    hipError_t e = bhk::launchResetBatch(A, st);
    if (e != hipSuccess) return e;
    if (band)
    {
        bhk::BatchBandArgs B = band->A;
        B.stamps = A.stamps;
        e = bhk::launchBatchBandReset(B, st);
        if (e == hipSuccess) e = bhk::launchBandClassifyBatch(B, st);
        if (e == hipSuccess) e = bhk::launchTileOrder(B.tileKey, B.nTiles, B.tileHist, band->order, st);
        if (e == hipSuccess) e = bhk::launchBatchBandSort(B, st);
        if (e == hipSuccess) e = bhk::launchBatchBandScan(B, st);
        if (e == hipSuccess) e = hipEventRecord(band->fork, st);
        if (e == hipSuccess)
            e = math ? bhk::launchBandTraceBatchFast(B, band->blocks, st) : bhk::launchBandTraceBatchExact(B, band->blocks, st);
        if (e == hipSuccess) e = hipStreamWaitEvent(band->aux, band->fork, 0);
        if (e == hipSuccess)
        {
            // once the auxiliary stream is forked it is joined back, also when
            // the trace launch fails, so no work of this batch outlives the
            // caller's stream order
            // the FP32 trace takes the batch's tiles costliest first, by the
            // classification's tile keys
            bhk::BatchArgs AO = A;
            AO.order = band->order;
            e = math ? bhk::launchTraceBatchFast(AO, L, band->aux) : bhk::launchTraceBatchExact(AO, L, band->aux);
            hipError_t j = hipEventRecord(band->join, band->aux);
            if (j == hipSuccess) j = hipStreamWaitEvent(st, band->join, 0);
            if (e == hipSuccess) e = j;
        }
    }
    else
    {
        bhk::BatchArgs AO = A;
        if (stepOrder)
        {
            e = bhk::launchFrameReset(nullptr, 0, stepOrder->key, A.totalTiles, stepOrder->hist, nullptr, st);
            if (e == hipSuccess) e = bhk::launchTileKeysBatch(A.views, A.nViews, maxW * maxH, stepOrder->key, st);
            if (e == hipSuccess) e = bhk::launchTileOrder(stepOrder->key, A.totalTiles, stepOrder->hist, stepOrder->order, st);
            if (e != hipSuccess) return e;
            AO.order = stepOrder->order;
        }
        e = math ? bhk::launchTraceBatchFast(AO, L, st) : bhk::launchTraceBatchExact(AO, L, st);
    }
    if (e != hipSuccess) return e;
    if (timers) recordEvent(timers, 1, st);
    e = bhk::launchShadeBlendBatch(A.views, A.nViews, maxW, maxH, st);
    if (e != hipSuccess) return e;
    if (timers) recordEvent(timers, 2, st);
    e = bhk::launchBloomBatch(A.views, A.nViews, maxW, maxH, st);
    if (e != hipSuccess) return e;
    if (timers) recordEvent(timers, 3, st);
    return hipSuccess;
}

hipError_t renderBatchPersistent(int n, const RenderParams* views, uchar4* const* outs, float4* const* accums,
                                 float4* const* bloomAs, float4* const* bloomBs, hipStream_t st)
{
    if (!kPersistentBatchEnabled) return hipErrorNotSupported;
    BatchTable T;
    hipError_t e = buildBatchTable(n, views, outs, accums, bloomAs, bloomBs, T);
    if (e != hipSuccess) return e;
    Scratch* S = nullptr;
    e = getScratch(st, T.totalPix, &S);
    if (e != hipSuccess) return e;
    e = getBatchScratch(S, n, st);
    if (e != hipSuccess) return e;
    assignBatchRecords(T, S->rec);
    const BhPrecision pr = currentPrecision();
    const bool band = pr.fp64Band != 0;
    // the same views as the last persistent batch on this stream (apart from
    // sample index and disk time), into the same records: their step counts
    // order the tiles and band rays
    bool repeat;
    {
        std::vector<RenderParams> key(T.D.size());
        for (size_t v = 0; v < T.D.size(); ++v)
        {
            key[v] = T.D[v].P;
            key[v].sampleIndex = 0;
            key[v].diskTime = 0.f;
        }
        repeat = S->lastBatchRec == S->rec && S->lastBatchViews.size() == key.size() &&
                 std::memcmp(S->lastBatchViews.data(), key.data(), sizeof(RenderParams) * key.size()) == 0;
        S->lastBatchViews = key;
        S->lastBatchRec = S->rec;
    }
    if (band || repeat)
    {
        e = growBatchTiles(S->btiles, (size_t)T.tiles, st);
        if (e != hipSuccess) return e;
    }
    BatchBandPlan plan;
    if (band)
    {
        e = growBatchBand(S->bband, n, T.totalPix, true, st);
        if (e == hipSuccess && !S->aux)
        {
            e = hipStreamCreateWithFlags(&S->aux, hipStreamNonBlocking);
            if (e == hipSuccess) e = hipEventCreateWithFlags(&S->fork, hipEventDisableTiming);
            if (e == hipSuccess) e = hipEventCreateWithFlags(&S->join, hipEventDisableTiming);
        }
        if (e != hipSuccess) return e;
        planBatchBand(T, pr, S->bband, S->btiles, repeat, plan);
        plan.aux = S->aux;
        plan.fork = S->fork;
        plan.join = S->join;
        int dev = 0;
        if (hipGetDevice(&dev) != hipSuccess) dev = 0;
        plan.blocks = kBatchBandWavesPerCU * numCUs(dev);
    }
    // stage the tables in pinned memory (after the previous uploads from
    // them have completed) and copy them asynchronously on the stream
    e = hipEventSynchronize(S->staged);
    if (e != hipSuccess) return e;
    std::memcpy(S->hostViews, T.D.data(), sizeof(bhk::ViewDesc) * n);
    e = hipMemcpyAsync(S->views, S->hostViews, sizeof(bhk::ViewDesc) * n, hipMemcpyHostToDevice, st);
    if (e != hipSuccess) return e;
    if (band)
    {
        std::memcpy(S->bband.hostTable, plan.views.data(), sizeof(bhk::BandArgs) * n);
        e = hipMemcpyAsync(S->bband.table, S->bband.hostTable, sizeof(bhk::BandArgs) * n, hipMemcpyHostToDevice, st);
        if (e != hipSuccess) return e;
    }
    e = hipEventRecord(S->staged, st);
    if (e != hipSuccess) return e;
    bhk::BatchArgs A;
    A.views = S->views;
    A.nViews = n;
    A.totalTiles = T.tiles;
    A.counter = S->counter;
    A.viewDone = S->viewDone;
    A.stamps = S->stamps;
    A.order = nullptr;
    S->lastViews = n;
    S->lastPersistent = true;
    const BhSchedule sc = currentSchedule();
    const bool timed = g_timers.load() != 0;
    e = enqueueBatchKernels(A, T.maxW, T.maxH, sc.wgSize, sc.mathMode, st, timed ? S : nullptr, band ? &plan : nullptr,
                            (!band && repeat) ? &S->btiles : nullptr);
    if (e != hipSuccess) return e;
    if (timed)
    {
        int d2 = 0;
        if (hipGetDevice(&d2) != hipSuccess) d2 = 0;
        std::lock_guard<std::mutex> lk(g_mu);
        g_lastTimed = {d2, st};
    }
    return hipGetLastError();
}

// Per-view times (ms after the batch start) from device wall-clock stamps:
// stamps[v] the view's last traced tile, stamps[n] the start.
hipError_t stampsToMs(const unsigned long long* dStamps, int n, float* msOut, int nOut)
{
    std::vector<unsigned long long> h((size_t)n + 1);
    int dev = 0, khz = 0;
    hipError_t e = hipMemcpy(h.data(), dStamps, h.size() * sizeof(unsigned long long), hipMemcpyDeviceToHost);
    if (e == hipSuccess) e = hipGetDevice(&dev);
    if (e == hipSuccess) e = hipDeviceGetAttribute(&khz, hipDeviceAttributeWallClockRate, dev);
    if (e != hipSuccess) return e;
    if (khz <= 0) return hipErrorNotSupported;
    for (int v = 0; v < nOut; ++v)
        msOut[v] = h[(size_t)v] >= h[(size_t)n] ? (float)((double)(h[(size_t)v] - h[(size_t)n]) / khz) : -1.f;
    return hipSuccess;
}


// Checks every view of a batch before any work is enqueued (the arguments
// launchRenderPipeline checks, on each view's sanitized copy).
hipError_t validateBatch(int n, const RenderParams* views, uchar4* const* outs, float4* const* accums,
                         float4* const* bloomAs, float4* const* bloomBs)
{
    if (n < 1 || !views || !outs || !accums) return hipErrorInvalidValue;
    for (int v = 0; v < n; ++v)
    {
        RenderParams P = views[v];
        bhSanitizeRenderParams(&P);
        if (!outs[v] || !accums[v]) return hipErrorInvalidValue;
        if (P.bloomEnabled && (!bloomAs || !bloomBs || !bloomAs[v] || !bloomBs[v])) return hipErrorInvalidValue;
    }
    return hipSuccess;
}

hipError_t renderBatch(int n, const RenderParams* views, uchar4* const* outs, float4* const* accums,
                       float4* const* bloomAs, float4* const* bloomBs, hipStream_t st)
{
    hipError_t e = validateBatch(n, views, outs, accums, bloomAs, bloomBs);
    if (e != hipSuccess) return e;
    Scratch* S = nullptr;
    e = getScratch(st, 0, &S);
    if (e != hipSuccess) return e;
    {
        std::lock_guard<std::mutex> lk(g_mu);
        if ((int)S->viewEvents.size() < n + 1)
        {
            for (size_t k = S->viewEvents.size(); k < (size_t)n + 1; ++k)
            {
                hipEvent_t ev = nullptr;
                e = hipEventCreate(&ev);
                if (e != hipSuccess) return e;
                S->viewEvents.push_back(ev);
            }
        }
        if ((int)S->slots.size() < n) S->slots.resize((size_t)n);
        S->lastViews = n;
        S->lastPersistent = false;
    }
    // worker streams (default priority) forked from the caller's stream; they
    // and the auxiliary streams of their pipelines each get a hardware queue
    // of their own where the runtime allows it
    int dev = 0;
    e = hipGetDevice(&dev);
    if (e != hipSuccess) return e;
    {
        std::lock_guard<std::mutex> lk(g_mu);
        for (int k = 0; k < kBatchStreams; ++k)
        {
            if (!S->batchStreams[k])
            {
                e = createQueueStream(dev, &S->batchStreams[k]);
                if (e != hipSuccess) { S->batchStreams[k] = nullptr; return e; }
            }
            if (!S->batchDone[k])
            {
                e = hipEventCreateWithFlags(&S->batchDone[k], hipEventDisableTiming);
                if (e != hipSuccess) { S->batchDone[k] = nullptr; return e; }
            }
        }
    }
    for (int k = 0; k < kBatchStreams; ++k)
    {
        Scratch* W = nullptr;
        e = getScratch(S->batchStreams[k], 0, &W);
        if (e != hipSuccess) return e;
        std::lock_guard<std::mutex> lk(g_mu);
        if (!W->aux)
        {
            e = createQueueStream(dev, &W->aux);
            if (e != hipSuccess) { W->aux = nullptr; return e; }
            e = hipEventCreateWithFlags(&W->fork, hipEventDisableTiming);
            if (e == hipSuccess) e = hipEventCreateWithFlags(&W->join, hipEventDisableTiming);
            if (e != hipSuccess) return e;
        }
    }
    e = hipEventRecord(S->viewEvents[0], st);
    if (e != hipSuccess) return e;
    for (int k = 0; k < kBatchStreams; ++k)
    {
        e = hipStreamWaitEvent(S->batchStreams[k], S->viewEvents[0], 0);
        if (e != hipSuccess) return e;
    }
    // largest views first, each on the worker stream with the fewest pixels
    // so far; each view traces into its own records, so a view repeated in
    // the next batch orders its work by its previous step counts
    std::vector<int> ord((size_t)n);
    for (int v = 0; v < n; ++v) ord[(size_t)v] = v;
    std::stable_sort(ord.begin(), ord.end(), [&](int x, int y) {
        return (long long)views[x].width * views[x].height > (long long)views[y].width * views[y].height;
    });
    long long load[kBatchStreams] = {};
    hipError_t err = hipSuccess;
    for (int v : ord)
    {
        int k = 0;
        for (int j = 1; j < kBatchStreams; ++j)
            if (load[j] < load[k]) k = j;
        load[k] += (long long)std::max(views[v].width, 1) * std::max(views[v].height, 1);
        hipStream_t ws = S->batchStreams[k];
        err = renderPipeline(outs[v], accums[v], bloomAs ? bloomAs[v] : nullptr, bloomBs ? bloomBs[v] : nullptr,
                             views[v], ws, false, nullptr, true, &S->slots[(size_t)v]);
        if (err == hipSuccess) err = hipEventRecord(S->viewEvents[(size_t)v + 1], ws);
        if (err != hipSuccess) break;
    }
    // the caller's stream continues once every worker stream has finished
    // (also after an error: work already enqueued stays ordered before the
    // caller's later work)
    for (int k = 0; k < kBatchStreams; ++k)
    {
        e = hipEventRecord(S->batchDone[k], S->batchStreams[k]);
        if (e == hipSuccess) e = hipStreamWaitEvent(st, S->batchDone[k], 0);
        if (e != hipSuccess && err == hipSuccess) err = e;
    }
    return err;
}


// Frees every buffer, event and stream a Scratch holds (its streams must be
// idle).
void freeScratch(Scratch& S)
{
    (void)hipFree(S.rec);
    (void)hipFree(S.counter);
    if (S.evValid)
        for (auto& ev : S.ev) (void)hipEventDestroy(ev);
    (void)hipFree(S.views);
    (void)hipHostFree(S.hostViews);
    (void)hipFree(S.viewDone);
    (void)hipFree(S.stamps);
    if (S.staged) (void)hipEventDestroy(S.staged);
    for (hipEvent_t ev : S.viewEvents) (void)hipEventDestroy(ev);
    for (int k = 0; k < kBatchStreams; ++k)
        if (S.batchDone[k]) (void)hipEventDestroy(S.batchDone[k]);
    (void)hipFree(S.bandMask);
    (void)hipFree(S.bandList);
    (void)hipFree(S.bandKey);
    (void)hipFree(S.bandSorted);
    (void)hipFree(S.bandHist);
    (void)hipFree(S.tileKey);
    (void)hipFree(S.tileOrder);
    (void)hipFree(S.tileHist);
    (void)hipFree(S.candInputs);
    (void)hipFree(S.candList);
    (void)hipFree(S.candAcc);
    for (auto& ev : S.bandEv)
        if (ev) (void)hipEventDestroy(ev);
    for (auto& ev : S.tileEv)
        if (ev) (void)hipEventDestroy(ev);
    for (auto& ev : S.routeEv)
        if (ev) (void)hipEventDestroy(ev);
    for (auto& sl : S.slots) (void)hipFree(sl.rec);
    if (S.aux)
    {
        (void)hipStreamSynchronize(S.aux);
        (void)hipStreamDestroy(S.aux);
    }
    if (S.fork) (void)hipEventDestroy(S.fork);
    if (S.join) (void)hipEventDestroy(S.join);
    freeBatchBand(S.bband);
    freeBatchTiles(S.btiles);
}

// Releases what the library holds for `st` on the current device: its
// scratch, and the worker streams of its batches with their scratch and
// auxiliary streams.  Waits for the stream's work first.
hipError_t releaseStream(hipStream_t st)
{
    int dev = 0;
    hipError_t e = hipGetDevice(&dev);
    if (e != hipSuccess) return e;
    e = hipStreamSynchronize(st);
    if (e != hipSuccess) return e;
    std::vector<Scratch> gone;
    std::vector<hipStream_t> workers;
    {
        std::lock_guard<std::mutex> lk(g_mu);
        auto it = g_scratch.find({dev, st});
        if (it == g_scratch.end()) return hipSuccess;
        gone.push_back(it->second);
        g_scratch.erase(it);
        for (int k = 0; k < kBatchStreams; ++k)
        {
            hipStream_t ws = gone[0].batchStreams[k];
            if (!ws) continue;
            workers.push_back(ws);
            auto w = g_scratch.find({dev, ws});
            if (w != g_scratch.end())
            {
                gone.push_back(w->second);
                g_scratch.erase(w);
            }
        }
        if (g_lastTimed.first == dev && (g_lastTimed.second == st ||
                                         std::find(workers.begin(), workers.end(), g_lastTimed.second) != workers.end()))
            g_lastTimed = {-1, nullptr};
    }
    for (hipStream_t ws : workers) (void)hipStreamSynchronize(ws);
    for (auto& S : gone) freeScratch(S);
    for (hipStream_t ws : workers) (void)hipStreamDestroy(ws);
    return hipSuccess;
}

}  // namespace

struct BhFramebuffer
{
    int w, h;
    uchar4* color;
    float4* hdr;
    float4* bloomA;
    float4* bloomB;
};

// A steady batch of views as one captured graph of the persistent launch.
struct BhBatchGraph
{
    int n = 0, maxW = 1, maxH = 1, tiles = 0;
    bhk::ViewDesc* table = nullptr;     // device view table (uploaded once)
    BhRayRecord* rec = nullptr;         // records of every view
    unsigned* counter = nullptr;        // queue head
    unsigned* viewDone = nullptr;
    unsigned long long* stamps = nullptr;
    hipStream_t cs = nullptr;           // private capture stream
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    int wgSize = 64;
    float tunedMs[3] = {0.f, 0.f, 0.f}; // trace+shade+post of one launch at workgroup 64, 128, 256 (0: not timed)
    bool band = false;                  // FP64 band pass (the precision at creation)
    BatchBandBufs bband;
    BatchTileBufs btiles;               // tile order (keys from the previous replay's step counts)
    BatchBandPlan plan;                 // its aux stream and fork/join events are the graph's own
};

struct BhAccumulator
{
    BhFramebuffer* fb = nullptr;
    int* index = nullptr;              // device sample index of the next frame
    hipStream_t capStream = nullptr;   // private stream the step graph is captured on
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
};


extern "C" {

bh_error_t launchRenderPipeline(bh_uchar4* out, bh_float4* accum, bh_float4* bloomA, bh_float4* bloomB,
                                const RenderParams& p, bh_stream_t stream)
{
    return (bh_error_t)renderPipeline(reinterpret_cast<uchar4*>(out), reinterpret_cast<float4*>(accum),
                                      reinterpret_cast<float4*>(bloomA), reinterpret_cast<float4*>(bloomB), p,
                                      (hipStream_t)stream, false);
}

bh_error_t launchRenderKernel(bh_uchar4* out, const RenderParams& p, bh_stream_t stream)
{
    return (bh_error_t)renderPipeline(reinterpret_cast<uchar4*>(out), nullptr, nullptr, nullptr, p,
                                      (hipStream_t)stream, true);
}

bh_error_t bhTraceRecords(BhRayRecord* devOut, const RenderParams& p, bh_stream_t stream)
{
    if (!devOut) return hipErrorInvalidValue;
    RenderParams P = p;
    bhSanitizeRenderParams(&P);
    hipStream_t st = (hipStream_t)stream;
    Scratch* S = nullptr;
    hipError_t e = getScratch(st, 0, &S);
    if (e != hipSuccess) return e;
    e = enqueueTrace(P, devOut, S, scheduleFor(P), st);
    if (e != hipSuccess) return e;
    return hipGetLastError();
}

static bh_error_t probeCommon(BhProbe* devOut, const bh_float3* dirs, int n, int nx, int ny, const RenderParams& p,
                              bh_stream_t stream)
{
    if (!devOut || n < 0) return hipErrorInvalidValue;
    bhk::ProbeArgs A;
    A.P = p;
    A.out = devOut;
    A.dirs = dirs;
    A.n = n;
    A.nx = nx;
    A.ny = ny;
    const BhSchedule sc = currentSchedule();
    hipError_t e = sc.mathMode ? bhk::launchProbeFast(A, (hipStream_t)stream) : bhk::launchProbeExact(A, (hipStream_t)stream);
    return (bh_error_t)e;
}

bh_error_t bhProbeGrid(BhProbe* devOut, const RenderParams& p, int nx, int ny, bh_stream_t stream)
{
    if (nx <= 0 || ny <= 0) return hipErrorInvalidValue;
    return probeCommon(devOut, nullptr, nx * ny, nx, ny, p, stream);
}

bh_error_t bhProbeDirs(BhProbe* devOut, const bh_float3* devDirs, int n, const RenderParams& p, bh_stream_t stream)
{
    if (!devDirs) return hipErrorInvalidValue;
    return probeCommon(devOut, devDirs, n, 0, 0, p, stream);
}

// ---------------------------------------------------------------- batched views
bh_error_t bhRenderBatch(int nViews, const RenderParams* views, bh_uchar4* const* outs, bh_float4* const* accums,
                         bh_float4* const* bloomAs, bh_float4* const* bloomBs, bh_stream_t stream)
{
    return (bh_error_t)renderBatch(nViews, views, reinterpret_cast<uchar4* const*>(outs),
                                   reinterpret_cast<float4* const*>(accums), reinterpret_cast<float4* const*>(bloomAs),
                                   reinterpret_cast<float4* const*>(bloomBs), (hipStream_t)stream);
}

bh_error_t bhRenderBatchPersistent(int nViews, const RenderParams* views, bh_uchar4* const* outs,
                                   bh_float4* const* accums, bh_float4* const* bloomAs, bh_float4* const* bloomBs,
                                   bh_stream_t stream)
{
    return (bh_error_t)renderBatchPersistent(nViews, views, reinterpret_cast<uchar4* const*>(outs),
                                             reinterpret_cast<float4* const*>(accums),
                                             reinterpret_cast<float4* const*>(bloomAs),
                                             reinterpret_cast<float4* const*>(bloomBs), (hipStream_t)stream);
}

bh_error_t bhBatchViewLatency(float* msOut, int nViews, bh_stream_t stream)
{
    hipStream_t st = (hipStream_t)stream;
    Scratch* S = nullptr;
    hipError_t e = getScratch(st, 0, &S);
    if (e != hipSuccess) return e;
    if (!msOut || nViews < 1 || nViews > S->lastViews) return hipErrorInvalidValue;
    if (S->lastPersistent)
    {
        e = hipStreamSynchronize(st);
        if (e != hipSuccess) return e;
        return stampsToMs(S->stamps, S->lastViews, msOut, nViews);
    }
    if ((int)S->viewEvents.size() < nViews + 1) return hipErrorInvalidValue;
    e = hipStreamSynchronize(st);
    if (e != hipSuccess) return e;
    for (int v = 0; v < nViews; ++v)
    {
        e = hipEventElapsedTime(&msOut[v], S->viewEvents[0], S->viewEvents[(size_t)v + 1]);
        if (e != hipSuccess) return e;
    }
    return hipSuccess;
}

// ---------------------------------------------------------------- graph-captured persistent batch
void bhBatchGraphDestroy(BhBatchGraph* g)
{
    if (!g) return;
    if (g->cs) (void)hipStreamSynchronize(g->cs);
    if (g->exec) (void)hipGraphExecDestroy(g->exec);
    if (g->graph) (void)hipGraphDestroy(g->graph);
    if (g->cs) (void)hipStreamDestroy(g->cs);
    if (g->plan.aux)
    {
        (void)hipStreamSynchronize(g->plan.aux);
        (void)hipStreamDestroy(g->plan.aux);
    }
    if (g->plan.fork) (void)hipEventDestroy(g->plan.fork);
    if (g->plan.join) (void)hipEventDestroy(g->plan.join);
    freeBatchBand(g->bband);
    freeBatchTiles(g->btiles);
    (void)hipFree(g->table);
    (void)hipFree(g->rec);
    (void)hipFree(g->counter);
    (void)hipFree(g->viewDone);
    (void)hipFree(g->stamps);
    delete g;
}

bh_error_t bhBatchGraphCreate(int nViews, const RenderParams* views, bh_uchar4* const* outs, bh_float4* const* accums,
                              bh_float4* const* bloomAs, bh_float4* const* bloomBs, int autotune, BhBatchGraph** out)
{
    if (!out) return hipErrorInvalidValue;
    *out = nullptr;
    BatchTable T;
    hipError_t e = buildBatchTable(nViews, views, reinterpret_cast<uchar4* const*>(outs),
                                   reinterpret_cast<float4* const*>(accums), reinterpret_cast<float4* const*>(bloomAs),
                                   reinterpret_cast<float4* const*>(bloomBs), T);
    if (e != hipSuccess) return e;
    BhBatchGraph* g = new (std::nothrow) BhBatchGraph;
    if (!g) return hipErrorOutOfMemory;
    g->n = nViews;
    g->maxW = T.maxW;
    g->maxH = T.maxH;
    g->tiles = T.tiles;
    e = hipMalloc(&g->rec, std::max<size_t>(T.totalPix, 1) * sizeof(BhRayRecord));
    if (e == hipSuccess) e = hipMalloc(&g->table, sizeof(bhk::ViewDesc) * nViews);
    if (e == hipSuccess) e = hipMalloc(&g->counter, kCounterBytes);
    if (e == hipSuccess) e = hipMalloc(&g->viewDone, sizeof(unsigned) * nViews);
    if (e == hipSuccess) e = hipMalloc(&g->stamps, sizeof(unsigned long long) * (nViews + 1));
    if (e == hipSuccess) e = hipStreamCreateWithFlags(&g->cs, hipStreamNonBlocking);
    const BhPrecision pr = currentPrecision();
    g->band = pr.fp64Band != 0;
    if (e == hipSuccess)
    {
        assignBatchRecords(T, g->rec);
        e = growBatchTiles(g->btiles, (size_t)T.tiles, nullptr);
        if (e == hipSuccess && g->band)
        {
            e = growBatchBand(g->bband, nViews, T.totalPix, false, nullptr);
            if (e == hipSuccess) e = hipStreamCreateWithFlags(&g->plan.aux, hipStreamNonBlocking);
            if (e == hipSuccess) e = hipEventCreateWithFlags(&g->plan.fork, hipEventDisableTiming);
            if (e == hipSuccess) e = hipEventCreateWithFlags(&g->plan.join, hipEventDisableTiming);
            if (e == hipSuccess)
            {
                // replays repeat the batch: band rays ordered by the step
                // counts of the previous launch's records
                planBatchBand(T, pr, g->bband, g->btiles, true, g->plan);
                int dev = 0;
                if (hipGetDevice(&dev) != hipSuccess) dev = 0;
                g->plan.blocks = kBatchBandWavesPerCU * numCUs(dev);
                e = hipMemcpy(g->bband.table, g->plan.views.data(), sizeof(bhk::BandArgs) * nViews,
                              hipMemcpyHostToDevice);
            }
        }
        if (e == hipSuccess) e = hipMemcpy(g->table, T.D.data(), sizeof(bhk::ViewDesc) * nViews, hipMemcpyHostToDevice);
    }
    if (e != hipSuccess) { bhBatchGraphDestroy(g); return e; }
    bhk::BatchArgs A;
    A.views = g->table;
    A.nViews = nViews;
    A.totalTiles = T.tiles;
    A.counter = g->counter;
    A.viewDone = g->viewDone;
    A.stamps = g->stamps;
    A.order = nullptr;
    const BhSchedule sc = currentSchedule();
    g->wgSize = sc.wgSize;
    if (autotune)
    {
        // one warm-up and one timed eager launch per workgroup size; the
        // frames do not depend on it
        const int wgs[3] = {64, 128, 256};
        hipEvent_t t0 = nullptr, t1 = nullptr;
        e = hipEventCreate(&t0);
        if (e == hipSuccess) e = hipEventCreate(&t1);
        float best = 1e30f;
        for (int w = 0; w < 3 && e == hipSuccess; ++w)
        {
            for (int r = 0; r < 2 && e == hipSuccess; ++r)
            {
                e = hipEventRecord(t0, g->cs);
                if (e == hipSuccess)
                    e = enqueueBatchKernels(A, g->maxW, g->maxH, wgs[w], sc.mathMode, g->cs, nullptr,
                                            g->band ? &g->plan : nullptr, g->band ? nullptr : &g->btiles);
                if (e == hipSuccess) e = hipEventRecord(t1, g->cs);
                if (e == hipSuccess) e = hipEventSynchronize(t1);
            }
            float ms = 0.f;
            if (e == hipSuccess) e = hipEventElapsedTime(&ms, t0, t1);
            if (e == hipSuccess)
            {
                g->tunedMs[w] = ms;
                if (ms < best) { best = ms; g->wgSize = wgs[w]; }
            }
        }
        if (t0) (void)hipEventDestroy(t0);
        if (t1) (void)hipEventDestroy(t1);
        if (e != hipSuccess) { bhBatchGraphDestroy(g); return e; }
    }
    e = hipStreamBeginCapture(g->cs, hipStreamCaptureModeThreadLocal);
    if (e != hipSuccess) { bhBatchGraphDestroy(g); return e; }
    // replays repeat the batch: tiles (and band rays) ordered by the step
    // counts of the previous replay's records
    hipError_t le = enqueueBatchKernels(A, g->maxW, g->maxH, g->wgSize, sc.mathMode, g->cs, nullptr,
                                        g->band ? &g->plan : nullptr, g->band ? nullptr : &g->btiles);
    hipGraph_t graph = nullptr;
    e = hipStreamEndCapture(g->cs, &graph);
    if (le != hipSuccess || e != hipSuccess)
    {
        if (graph) (void)hipGraphDestroy(graph);
        bhBatchGraphDestroy(g);
        return le != hipSuccess ? le : e;
    }
    g->graph = graph;
    e = hipGraphInstantiate(&g->exec, graph, nullptr, nullptr, 0);
    if (e != hipSuccess) { g->exec = nullptr; bhBatchGraphDestroy(g); return e; }
    *out = g;
    return hipSuccess;
}

bh_error_t bhBatchGraphLaunch(BhBatchGraph* g, bh_stream_t stream)
{
    if (!g || !g->exec) return hipErrorInvalidValue;
    return hipGraphLaunch(g->exec, (hipStream_t)stream);
}

bh_error_t bhBatchGraphInfo(BhBatchGraph* g, int* wgSize, float* tunedMs3)
{
    if (!g) return hipErrorInvalidValue;
    if (wgSize) *wgSize = g->wgSize;
    if (tunedMs3)
        for (int k = 0; k < 3; ++k) tunedMs3[k] = g->tunedMs[k];
    return hipSuccess;
}

bh_error_t bhBatchGraphLatency(BhBatchGraph* g, float* msOut, int nViews, bh_stream_t stream)
{
    if (!g || !msOut || nViews < 1 || nViews > g->n) return hipErrorInvalidValue;
    hipError_t e = hipStreamSynchronize((hipStream_t)stream);
    if (e != hipSuccess) return e;
    return stampsToMs(g->stamps, g->n, msOut, nViews);
}

// ---------------------------------------------------------------- framebuffer
bh_error_t bhFramebufferCreate(int width, int height, BhFramebuffer** fb)
{
    if (!fb || width < 1 || height < 1) return hipErrorInvalidValue;
    *fb = nullptr;
    BhFramebuffer* F = new (std::nothrow) BhFramebuffer{width, height, nullptr, nullptr, nullptr, nullptr};
    if (!F) return hipErrorOutOfMemory;
    const size_t n = (size_t)width * height, nh = (size_t)((width + 1) / 2) * ((height + 1) / 2);
    hipError_t e = hipMalloc(&F->color, n * sizeof(uchar4));
    if (e == hipSuccess) e = hipMalloc(&F->hdr, n * sizeof(float4));
    if (e == hipSuccess) e = hipMalloc(&F->bloomA, nh * sizeof(float4));
    if (e == hipSuccess) e = hipMalloc(&F->bloomB, nh * sizeof(float4));
    if (e == hipSuccess) e = hipMemset(F->hdr, 0, n * sizeof(float4));
    if (e != hipSuccess)
    {
        bhFramebufferDestroy(F);
        return e;
    }
    *fb = F;
    return hipSuccess;
}

void bhFramebufferDestroy(BhFramebuffer* fb)
{
    if (!fb) return;
    (void)hipFree(fb->color);
    (void)hipFree(fb->hdr);
    (void)hipFree(fb->bloomA);
    (void)hipFree(fb->bloomB);
    delete fb;
}

bh_uchar4* bhFramebufferColor(BhFramebuffer* fb) { return fb ? reinterpret_cast<bh_uchar4*>(fb->color) : nullptr; }
bh_float4* bhFramebufferHDR(BhFramebuffer* fb) { return fb ? reinterpret_cast<bh_float4*>(fb->hdr) : nullptr; }

bh_error_t bhFramebufferRender(BhFramebuffer* fb, const RenderParams& p, bh_stream_t stream)
{
    if (!fb || p.width != fb->w || p.height != fb->h) return hipErrorInvalidValue;
    return (bh_error_t)renderPipeline(fb->color, fb->hdr, fb->bloomA, fb->bloomB, p, (hipStream_t)stream, false);
}

bh_error_t bhPostChain(bh_uchar4* out, const bh_float4* accum, bh_float4* bloomA, bh_float4* bloomB,
                       const RenderParams& p, bh_stream_t stream)
{
    RenderParams P = p;
    bhSanitizeRenderParams(&P);
    if (!out || !accum) return hipErrorInvalidValue;
    if (P.bloomEnabled && (!bloomA || !bloomB)) return hipErrorInvalidValue;
    hipStream_t st = (hipStream_t)stream;
    Scratch* S = nullptr;
    hipError_t e = getScratch(st, 0, &S);
    if (e != hipSuccess) return e;
    const float4* acc = reinterpret_cast<const float4*>(accum);
    uchar4* o = reinterpret_cast<uchar4*>(out);
    const bool timed = g_timers.load() != 0;
    if (timed) { recordEvent(S, 0, st); recordEvent(S, 1, st); recordEvent(S, 2, st); }
    e = bhk::launchPostFused(P, acc, o, st);
    if (e != hipSuccess) return e;
    if (timed)
    {
        recordEvent(S, 3, st);
        int dev = 0;
        if (hipGetDevice(&dev) != hipSuccess) dev = 0;
        std::lock_guard<std::mutex> lk(g_mu);
        g_lastTimed = {dev, st};
    }
    return hipGetLastError();
}

bh_error_t bhFramebufferPost(BhFramebuffer* fb, const RenderParams& p, bh_stream_t stream)
{
    if (!fb || p.width != fb->w || p.height != fb->h) return hipErrorInvalidValue;
    return bhPostChain(reinterpret_cast<bh_uchar4*>(fb->color), reinterpret_cast<const bh_float4*>(fb->hdr),
                       reinterpret_cast<bh_float4*>(fb->bloomA), reinterpret_cast<bh_float4*>(fb->bloomB), p, stream);
}

bh_error_t bhFramebufferReadback(BhFramebuffer* fb, bh_uchar4* hostColor, bh_float4* hostHdr, bh_stream_t stream)
{
    if (!fb) return hipErrorInvalidValue;
    hipStream_t st = (hipStream_t)stream;
    const size_t n = (size_t)fb->w * fb->h;
    hipError_t e = hipStreamSynchronize(st);
    if (e == hipSuccess && hostColor) e = hipMemcpy(hostColor, fb->color, n * sizeof(uchar4), hipMemcpyDeviceToHost);
    if (e == hipSuccess && hostHdr) e = hipMemcpy(hostHdr, fb->hdr, n * sizeof(float4), hipMemcpyDeviceToHost);
    return e;
}

// ---------------------------------------------------------------- progressive accumulation
bh_error_t bhAccumCreate(int width, int height, BhAccumulator** acc)
{
    if (!acc) return hipErrorInvalidValue;
    *acc = nullptr;
    BhAccumulator* A = new (std::nothrow) BhAccumulator;
    if (!A) return hipErrorOutOfMemory;
    hipError_t e = (hipError_t)bhFramebufferCreate(width, height, &A->fb);
    if (e == hipSuccess) e = hipMalloc(&A->index, sizeof(int) * 64);
    if (e == hipSuccess) e = hipMemset(A->index, 0, sizeof(int) * 64);
    if (e == hipSuccess) e = hipStreamCreateWithFlags(&A->capStream, hipStreamNonBlocking);
    if (e != hipSuccess)
    {
        bhAccumDestroy(A);
        return e;
    }
    *acc = A;
    return hipSuccess;
}

void bhAccumDestroy(BhAccumulator* acc)
{
    if (!acc) return;
    if (acc->exec) (void)hipGraphExecDestroy(acc->exec);
    if (acc->graph) (void)hipGraphDestroy(acc->graph);
    if (acc->capStream)
    {
        (void)releaseStream(acc->capStream);
        (void)hipStreamDestroy(acc->capStream);
    }
    (void)hipFree(acc->index);
    bhFramebufferDestroy(acc->fb);
    delete acc;
}

bh_uchar4* bhAccumColor(BhAccumulator* acc) { return acc ? bhFramebufferColor(acc->fb) : nullptr; }
bh_float4* bhAccumHDR(BhAccumulator* acc) { return acc ? bhFramebufferHDR(acc->fb) : nullptr; }

bh_error_t bhAccumReset(BhAccumulator* acc, bh_stream_t stream)
{
    if (!acc) return hipErrorInvalidValue;
    return bhk::launchSetIndex(acc->index, 0, (hipStream_t)stream);
}

static hipError_t accumStepOn(BhAccumulator* acc, const RenderParams& p, hipStream_t st, bool allowTimers)
{
    BhFramebuffer* F = acc->fb;
    if (p.width != F->w || p.height != F->h) return hipErrorInvalidValue;
    hipError_t e = renderPipeline(F->color, F->hdr, F->bloomA, F->bloomB, p, st, false, acc->index, allowTimers);
    if (e != hipSuccess) return e;
    return bhk::launchBumpIndex(acc->index, st);
}

bh_error_t bhAccumStep(BhAccumulator* acc, const RenderParams& p, bh_stream_t stream)
{
    if (!acc) return hipErrorInvalidValue;
    return accumStepOn(acc, p, (hipStream_t)stream, true);
}

bh_error_t bhAccumCapture(BhAccumulator* acc, const RenderParams& p, bh_stream_t stream)
{
    if (!acc) return hipErrorInvalidValue;
    BhFramebuffer* F = acc->fb;
    if (p.width != F->w || p.height != F->h) return hipErrorInvalidValue;
    hipStream_t cs = acc->capStream;
    // work already queued by the caller comes first
    hipError_t e = hipStreamSynchronize((hipStream_t)stream);
    if (e != hipSuccess) return e;
    // scratch of the capture stream is allocated before capture (no
    // allocation may happen inside it)
    RenderParams P = p;
    bhSanitizeRenderParams(&P);
    const size_t n = (size_t)P.width * (size_t)P.height;
    Scratch* S = nullptr;
    e = getScratch(cs, n, &S);
    // the tile order's buffers with the band pass on or off (a captured step
    // repeats its view, so its tiles are ordered either way)
    if (e == hipSuccess) e = getTileScratch(S, (size_t)((P.width + 7) / 8) * (size_t)((P.height + 7) / 8), cs);
    if (e == hipSuccess && currentPrecision().fp64Band)
    {
        e = getBandScratch(S, n, cs);
        if (e == hipSuccess && currentSurrogate().s) e = getCandScratch(S, n, cs);
    }
    if (e != hipSuccess) return e;
    if (acc->exec) { (void)hipGraphExecDestroy(acc->exec); acc->exec = nullptr; }
    if (acc->graph) { (void)hipGraphDestroy(acc->graph); acc->graph = nullptr; }
    e = hipStreamBeginCapture(cs, hipStreamCaptureModeThreadLocal);
    if (e != hipSuccess) return e;
    hipError_t le = accumStepOn(acc, p, cs, false);
    hipGraph_t g = nullptr;
    e = hipStreamEndCapture(cs, &g);
    if (le != hipSuccess) { if (g) (void)hipGraphDestroy(g); return le; }
    if (e != hipSuccess) return e;
    acc->graph = g;
    e = hipGraphInstantiate(&acc->exec, g, nullptr, nullptr, 0);
    if (e != hipSuccess) { acc->exec = nullptr; return e; }
    return hipSuccess;
}

bh_error_t bhAccumReplay(BhAccumulator* acc, int frames, bh_stream_t stream)
{
    if (!acc || frames < 0) return hipErrorInvalidValue;
    if (!acc->exec) return hipErrorNotReady;
    for (int i = 0; i < frames; ++i)
    {
        hipError_t e = hipGraphLaunch(acc->exec, (hipStream_t)stream);
        if (e != hipSuccess) return e;
    }
    return hipSuccess;
}

bh_error_t bhAccumSampleIndex(BhAccumulator* acc, int* index, bh_stream_t stream)
{
    if (!acc || !index) return hipErrorInvalidValue;
    hipError_t e = hipStreamSynchronize((hipStream_t)stream);
    if (e != hipSuccess) return e;
    return hipMemcpy(index, acc->index, sizeof(int), hipMemcpyDeviceToHost);
}

// ---------------------------------------------------------------- autotuner
bh_error_t bhAutotune(const RenderParams& p, int reps, BhSchedule* chosen, float* bestMs, bh_stream_t stream)
{
    if (reps < 1 || reps > 64) return hipErrorInvalidValue;
    RenderParams P = p;
    bhSanitizeRenderParams(&P);
    hipStream_t st = (hipStream_t)stream;
    const BhSchedule cur = currentSchedule();
    const BhPrecision pr = currentPrecision();
    const TuneKey key = tuneKeyOf(P, pr.fp64Band, cur.mathMode);
    {
        std::lock_guard<std::mutex> lk(g_mu);
        auto it = g_tuned.find(key);
        if (it != g_tuned.end())
        {
            if (chosen) *chosen = it->second.first;
            if (bestMs) *bestMs = it->second.second;
            return hipSuccess;
        }
    }
    const size_t n = (size_t)P.width * (size_t)P.height;
    Scratch* S = nullptr;
    hipError_t e = getScratch(st, n, &S);
    if (e != hipSuccess) return e;
    hipEvent_t t0 = nullptr, t1 = nullptr;
    e = hipEventCreate(&t0);
    if (e == hipSuccess) e = hipEventCreate(&t1);
    if (e != hipSuccess) return e;
    // candidates: workgroup size x (whole tiles | lane refill at two chunk
    // lengths) x cooperative disk sampling
    const int wgs[3] = {256, 128, 64};
    const int modes[3][2] = {{64, 4096}, {32, 256}, {16, 64}};
    BhSchedule best = cur;
    float bestT = 1e30f;
    for (int w = 0; w < 3 && e == hipSuccess; ++w)
        for (int m = 0; m < 3 && e == hipSuccess; ++m)
            for (int coop = 0; coop < 2 && e == hipSuccess; ++coop)
            {
                BhSchedule c = cur;
                c.wgSize = wgs[w];
                c.raysPerRefill = modes[m][0];
                c.stepsPerChunk = modes[m][1];
                c.diskCoop = coop;
                float tMin = 1e30f;
                for (int r = 0; r <= reps && e == hipSuccess; ++r)   // first run warms up
                {
                    e = hipEventRecord(t0, st);
                    if (e == hipSuccess) e = enqueueTrace(P, S->rec, S, c, st, nullptr, false);
                    if (e == hipSuccess) e = hipEventRecord(t1, st);
                    if (e == hipSuccess) e = hipEventSynchronize(t1);
                    float ms = 0.f;
                    if (e == hipSuccess) e = hipEventElapsedTime(&ms, t0, t1);
                    if (e == hipSuccess && r > 0) tMin = std::min(tMin, ms);
                }
                if (e == hipSuccess && tMin < bestT) { bestT = tMin; best = c; }
            }
    (void)hipEventDestroy(t0);
    (void)hipEventDestroy(t1);
    if (e != hipSuccess) return e;
    {
        std::lock_guard<std::mutex> lk(g_mu);
        g_tuned[key] = {best, bestT};
    }
    if (chosen) *chosen = best;
    if (bestMs) *bestMs = bestT;
    return hipSuccess;
}

void bhScheduleFor(const RenderParams& p, BhSchedule* s)
{
    if (!s) return;
    RenderParams P = p;
    bhSanitizeRenderParams(&P);
    *s = scheduleFor(P);
}

void bhAutotuneClear(void)
{
    std::lock_guard<std::mutex> lk(g_mu);
    g_tuned.clear();
}

// ---------------------------------------------------------------- schedule / timers
void bhGetSchedule(BhSchedule* s)
{
    if (s) *s = currentSchedule();
}

bh_error_t bhSetSchedule(const BhSchedule* s)
{
    if (!s) return hipErrorInvalidValue;
    if (s->wgSize != 64 && s->wgSize != 128 && s->wgSize != 256) return hipErrorInvalidValue;
    if (s->stepsPerChunk < 1 || s->stepsPerChunk > 4096) return hipErrorInvalidValue;
    if (s->mathMode != 0 && s->mathMode != 1) return hipErrorInvalidValue;
    if (s->raysPerRefill < 1 || s->raysPerRefill > 64) return hipErrorInvalidValue;
    if (s->diskCoop < 0 || s->diskCoop > 2) return hipErrorInvalidValue;
    std::lock_guard<std::mutex> lk(g_mu);
    g_sched = *s;
    return hipSuccess;
}

void bhEnableKernelTimers(int enable) { g_timers.store(enable ? 1 : 0); }

void bhGetPrecision(BhPrecision* p)
{
    if (p) *p = currentPrecision();
}
bh_error_t bhSetPrecision(const BhPrecision* p)
{
    if (!p) return hipErrorInvalidValue;
    if (p->fp64Band != 0 && p->fp64Band != 1) return hipErrorInvalidValue;
    if (!(p->critBand >= 0.f) || !(p->polarBand >= 0.f) || !std::isfinite(p->critBand) ||
        !std::isfinite(p->polarBand))
        return hipErrorInvalidValue;
    std::lock_guard<std::mutex> lk(g_mu);
    g_prec = *p;
    return hipSuccess;
}
bh_error_t bhSetSurrogate(BhSurrogate* s, float sigmaMax, float diskProbMax)
{
    if (s && (!(sigmaMax > 0.f) || !(diskProbMax > 0.f) || diskProbMax > 1.f)) return hipErrorInvalidValue;
    std::lock_guard<std::mutex> lk(g_mu);
    g_sur.s = s;
    if (s)
    {
        g_sur.sigmaMax = sigmaMax;
        g_sur.diskProbMax = diskProbMax;
    }
    return hipSuccess;
}

bh_error_t bhLastSurrogateStats(unsigned* candidates, unsigned* accepted, bh_stream_t stream)
{
    hipStream_t st = (hipStream_t)stream;
    int dev = 0;
    hipError_t e = hipGetDevice(&dev);
    if (e != hipSuccess) return e;
    Scratch* S = nullptr;
    {
        std::lock_guard<std::mutex> lk(g_mu);
        auto it = g_scratch.find({dev, st});
        if (it == g_scratch.end() || !it->second.counter) return hipErrorNotReady;
        S = &it->second;
    }
    e = hipStreamSynchronize(st);
    if (e == hipSuccess && candidates)
        e = hipMemcpy(candidates, S->counter + kCandCountSlot, sizeof(unsigned), hipMemcpyDeviceToHost);
    if (e == hipSuccess && accepted)
        e = hipMemcpy(accepted, S->counter + kAcceptSlot, sizeof(unsigned), hipMemcpyDeviceToHost);
    return e;
}

bh_error_t bhLastSurrogateKernelMs(float* ms, bh_stream_t stream)
{
    hipStream_t st = (hipStream_t)stream;
    int dev = 0;
    hipError_t e = hipGetDevice(&dev);
    if (e != hipSuccess) return e;
    Scratch* S = nullptr;
    {
        std::lock_guard<std::mutex> lk(g_mu);
        auto it = g_scratch.find({dev, st});
        if (it == g_scratch.end() || !it->second.routeTimed || !it->second.routeEv[0]) return hipErrorNotReady;
        S = &it->second;
    }
    e = hipStreamSynchronize(st);
    if (e == hipSuccess && ms) e = hipEventElapsedTime(ms, S->routeEv[0], S->routeEv[1]);
    return e;
}

bh_error_t bhLastBandStats(unsigned* count, float* bandMs, bh_stream_t stream)
{
    hipStream_t st = (hipStream_t)stream;
    int dev = 0;
    hipError_t e = hipGetDevice(&dev);
    if (e != hipSuccess) return e;
    Scratch* S = nullptr;
    {
        std::lock_guard<std::mutex> lk(g_mu);
        auto it = g_scratch.find({dev, st});
        if (it == g_scratch.end() || !it->second.counter) return hipErrorNotReady;
        S = &it->second;
    }
    e = hipStreamSynchronize(st);
    if (e != hipSuccess) return e;
    if (count)
    {
        e = hipMemcpy(count, S->counter + kBandCountSlot, sizeof(unsigned), hipMemcpyDeviceToHost);
        if (e != hipSuccess) return e;
    }
    if (bandMs)
    {
        *bandMs = 0.f;
        if (S->bandEv[0]) e = hipEventElapsedTime(bandMs, S->bandEv[0], S->bandEv[1]);
    }
    return e;
}
bh_error_t bhLastTraceKernelMs(float* tileMs, float* bandMs, bh_stream_t stream)
{
    hipStream_t st = (hipStream_t)stream;
    int dev = 0;
    hipError_t e = hipGetDevice(&dev);
    if (e != hipSuccess) return e;
    Scratch* S = nullptr;
    {
        std::lock_guard<std::mutex> lk(g_mu);
        auto it = g_scratch.find({dev, st});
        if (it == g_scratch.end() || !it->second.tileTimed || !it->second.tileEv[0]) return hipErrorNotReady;
        S = &it->second;
    }
    e = hipStreamSynchronize(st);
    if (e != hipSuccess) return e;
    if (tileMs) e = hipEventElapsedTime(tileMs, S->tileEv[0], S->tileEv[1]);
    if (e == hipSuccess && bandMs)
    {
        *bandMs = 0.f;
        if (S->bandEv[0]) e = hipEventElapsedTime(bandMs, S->bandEv[0], S->bandEv[1]);
    }
    return e;
}

bh_error_t bhLastTileOrder(unsigned* keys, unsigned* order, int capacity, int* nTiles, bh_stream_t stream)
{
    hipStream_t st = (hipStream_t)stream;
    int dev = 0;
    hipError_t e = hipGetDevice(&dev);
    if (e != hipSuccess) return e;
    Scratch* S = nullptr;
    {
        std::lock_guard<std::mutex> lk(g_mu);
        auto it = g_scratch.find({dev, st});
        if (it == g_scratch.end()) return hipErrorNotReady;
        S = &it->second;
    }
    e = hipStreamSynchronize(st);
    if (e != hipSuccess) return e;
    const int n = S->lastOrderTiles;
    if (nTiles) *nTiles = n;
    if (n == 0) return hipErrorNotReady;
    if (capacity < n) return hipErrorInvalidValue;
    if (keys) e = hipMemcpy(keys, S->tileKey, (size_t)n * sizeof(unsigned), hipMemcpyDeviceToHost);
    if (e == hipSuccess && order) e = hipMemcpy(order, S->tileOrder, (size_t)n * sizeof(unsigned), hipMemcpyDeviceToHost);
    return e;
}

bh_error_t bhLastKernelTimes(float* traceMs, float* shadeMs, float* postMs)
{
    Scratch* S = nullptr;
    {
        std::lock_guard<std::mutex> lk(g_mu);
        auto it = g_scratch.find(g_lastTimed);
        if (it == g_scratch.end() || !it->second.evValid) return hipErrorNotReady;
        S = &it->second;
    }
    hipError_t e = hipEventSynchronize(S->ev[3]);
    if (e != hipSuccess) return e;
    float a = 0, b = 0, c = 0;
    e = hipEventElapsedTime(&a, S->ev[0], S->ev[1]);
    if (e == hipSuccess) e = hipEventElapsedTime(&b, S->ev[1], S->ev[2]);
    if (e == hipSuccess) e = hipEventElapsedTime(&c, S->ev[2], S->ev[3]);
    if (traceMs) *traceMs = a;
    if (shadeMs) *shadeMs = b;
    if (postMs) *postMs = c;
    return e;
}

// ---------------------------------------------------------------- plumbing
bh_error_t bhDeviceAlloc(void** ptr, size_t bytes) { return hipMalloc(ptr, bytes); }
bh_error_t bhDeviceFree(void* ptr) { return hipFree(ptr); }
bh_error_t bhCopyToDevice(void* dst, const void* src, size_t bytes) { return hipMemcpy(dst, src, bytes, hipMemcpyHostToDevice); }
bh_error_t bhCopyToHost(void* dst, const void* src, size_t bytes) { return hipMemcpy(dst, src, bytes, hipMemcpyDeviceToHost); }
bh_error_t bhDeviceFill(void* dst, unsigned char value, size_t bytes) { return hipMemset(dst, value, bytes); }
bh_error_t bhCopyDevice(void* dst, const void* src, size_t bytes) { return hipMemcpy(dst, src, bytes, hipMemcpyDeviceToDevice); }
bh_error_t bhDeviceSynchronize(void) { return hipDeviceSynchronize(); }
bh_error_t bhStreamCreate(bh_stream_t* s)
{
    hipStream_t st = nullptr;
    hipError_t e = hipStreamCreateWithFlags(&st, hipStreamNonBlocking);
    if (s) *s = (bh_stream_t)st;
    return e;
}
bh_error_t bhStreamRelease(bh_stream_t s) { return releaseStream((hipStream_t)s); }
bh_error_t bhStreamDestroy(bh_stream_t s)
{
    hipError_t e = releaseStream((hipStream_t)s);
    hipError_t d = hipStreamDestroy((hipStream_t)s);
    return e != hipSuccess ? e : d;
}
bh_error_t bhStreamSynchronize(bh_stream_t s) { return hipStreamSynchronize((hipStream_t)s); }
const char* bhErrorString(bh_error_t e) { return hipGetErrorString((hipError_t)e); }
int bhDeviceCount(void)
{
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess) return 0;
    return n;
}

}  // extern "C"
