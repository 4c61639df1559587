// ---------------------------------------------------------------------------
// bh_trace.hip -- geodesic trace kernels for gfx950.
//
// Compiled twice: BHK_FAST=0 (MathExact, -ffp-contract=off) and BHK_FAST=1
// (MathFast, -ffp-contract=on).  Each object exports its own launchers.
//
// traceKernel: persistent wavefront-refill tracer.  Every lane owns one ray;
// a lane whose ray has terminated writes the ray's record (exit direction,
// outcome, disk radiance, transmittance, step count) and later takes the
// next ray id from a device-wide queue counter -- one atomic per wavefront
// per refill.  Ray ids walk the
// image in 8x8 pixel tiles so a wavefront's rays start spatially coherent.
// Sky shading is deferred to a separate pixel-order kernel (bh_post.hip):
// it runs once per ray and would otherwise execute with most lanes idle.
// Disk chords of a step are sampled cooperatively by the whole wavefront
// (diskChordWave) when kCoop is set.  Lanes whose ray ends write its record
// at once; the wavefront refills only when `refillMin` lanes are idle.
//
// probeKernel: one ray per thread with constraint statistics (tests).
// ---------------------------------------------------------------------------
#include "bh_rays.h"
#include "bh_internal.h"
#include "bh_surrogate_geom.h"

#if BHK_FAST
#define BHK_POLICY MathFast
#define BHK_SUFFIX(name) name##Fast
#else
#define BHK_POLICY MathExact
#define BHK_SUFFIX(name) name##Exact
#endif

namespace bhk {

using Pol = BHK_POLICY;

BHK_INL void pixelOfId(unsigned id, int tilesX, int& px, int& py)
{
    unsigned tile = id >> 6, inT = id & 63u;
    px = (int)((tile % (unsigned)tilesX) * 8u + (inT & 7u));
    py = (int)((tile / (unsigned)tilesX) * 8u + (inT >> 3));
}

template <class Ray>
BHK_INL void writeRecord(BhRayRecord* rec, int idx, const Ray& R)
{
    float4* o = reinterpret_cast<float4*>(rec + idx);
    int flags = (R.status == RAY_ESCAPED ? 1 : 2) | (R.step << 8);
    o[0] = make_float4(R.escDir.x, R.escDir.y, R.escDir.z, __int_as_float(flags));
    o[1] = make_float4(R.acc.x, R.acc.y, R.acc.z, R.trans);
}

template <class Ray, bool kCoop>
__global__ __launch_bounds__(256) void traceKernel(TraceArgs A)
{
    const RenderParams& P = A.P;
    const FrameK& K = A.K;
    const v3 cam = mk3(P.camPos.x, P.camPos.y, P.camPos.z);
    const unsigned lane = __lane_id();
    const unsigned total = (unsigned)A.totalIds;
    const int chunk = A.chunk;
    const int refillMin = A.refillMin;
    const float2 jit = launchJitter(P, A.jitter, A.dynIndex);

    Ray R;
    Chord ch;
    ch.n = 0;
    bool active = false, drained = false;
    int pix = 0;
    for (;;)
    {
        // ---- refill: one queue atomic per wavefront for all idle lanes
        const bool need = !active && !drained;
        const unsigned long long needMask = __ballot(need);
        if (needMask)
        {
            const int leader = __ffsll((long long)needMask) - 1;
            unsigned base = 0;
            if ((int)lane == leader) base = atomicAdd(A.counter, (unsigned)__popcll(needMask));
            base = __shfl(base, leader);
            if (need)
            {
                const unsigned id = base + (unsigned)__popcll(needMask & ((1ull << lane) - 1ull));
                if (id >= total) drained = true;
                else
                {
                    int px, py;
                    pixelOfId(id, A.tilesX, px, py);
                    if (px < P.width && py < P.height && !(A.bandMask && A.bandMask[py * P.width + px]))
                    {
                        const float fx = Pol::div((float)px + jit.x, (float)P.width);
                        const float fy = Pol::div((float)py + jit.y, (float)P.height);
                        R.init(K, cam, cameraRayDir<Pol>(P, fx, fy));
                        pix = py * P.width + px;
                        if (R.status == RAY_RUNNING) active = true;
                        else writeRecord(A.rec, pix, R);    // exactly radial ray
                    }
                }
            }
        }
        if (__ballot(active) == 0ull)
        {
            if (__ballot(!drained) == 0ull) break;
            continue;
        }
        // ---- up to `chunk` RK4 steps; leave early once enough lanes idle
        for (int k = 0; k < chunk; ++k)
        {
            bool done = false;
            if (active) done = R.advance(K, ch);
            else ch.n = 0;
            if (kCoop) diskChordWave<Pol>(K.D, active && ch.n > 0, ch, R.E, R.Lz, R.acc, R.trans);
            else if (active && ch.n > 0) diskChordSerial<Pol>(K.D, R.E, R.Lz, ch, R.acc, R.trans);
            if (done)
            {
                writeRecord(A.rec, pix, R);
                active = false;
            }
            const unsigned long long idle = __ballot(!active && !drained);
            if (__popcll(idle) >= refillMin || __ballot(active) == 0ull) break;
        }
    }
}

// ---------------------------------------------------------------- whole tiles
// One wavefront traces one 8x8 tile to completion, then takes the next tile
// from the queue (no mid-flight refill: the rays of a tile cross the disk
// during the same steps, which keeps the wavefront coherent).  Used by the
// single-view launch when raysPerRefill == 64 and by the batched launch.
template <class Ray, bool kCoop>
BHK_INL void traceTile(const RenderParams& P, const FrameK& K, BhRayRecord* rec, int tilesX, float2 jit,
                       unsigned tile, unsigned lane, const unsigned char* bandMask = nullptr)
{
    const int px = (int)((tile % (unsigned)tilesX) * 8u + (lane & 7u));
    const int py = (int)((tile / (unsigned)tilesX) * 8u + (lane >> 3));
    Ray R;
    Chord ch;
    ch.n = 0;
    bool active = false;
    int pix = 0;
    if (px < P.width && py < P.height && !(bandMask && bandMask[py * P.width + px]))
    {
        const float fx = Pol::div((float)px + jit.x, (float)P.width);
        const float fy = Pol::div((float)py + jit.y, (float)P.height);
        R.init(K, mk3(P.camPos.x, P.camPos.y, P.camPos.z), cameraRayDir<Pol>(P, fx, fy));
        pix = py * P.width + px;
        if (R.status == RAY_RUNNING) active = true;
        else writeRecord(rec, pix, R);
    }
    while (__ballot(active) != 0ull)
    {
        bool done = false;
        if (active) done = R.advance(K, ch);
        else ch.n = 0;
        if (kCoop) diskChordWave<Pol>(K.D, active && ch.n > 0, ch, R.E, R.Lz, R.acc, R.trans);
        else if (active && ch.n > 0) diskChordSerial<Pol>(K.D, R.E, R.Lz, ch, R.acc, R.trans);
        if (done)
        {
            writeRecord(rec, pix, R);
            active = false;
        }
    }
}

BHK_INL unsigned nextTile(unsigned* counter, unsigned lane)
{
    unsigned t = 0;
    if (lane == 0) t = atomicAdd(counter, 1u);
    return (unsigned)__builtin_amdgcn_readfirstlane((int)__shfl(t, 0));
}

template <class Ray, bool kCoop>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(Pol::fast ? 4 : 3))) void traceTileKernel(
    TraceArgs A)
{
    const unsigned lane = __lane_id();
    const unsigned nTiles = (unsigned)A.totalIds / 64u;
    const float2 jit = launchJitter(A.P, A.jitter, A.dynIndex);
    for (;;)
    {
        const unsigned t = nextTile(A.counter, lane);
        if (t >= nTiles) break;
        const unsigned tile = A.tileOrder ? A.tileOrder[t] : t;
        traceTile<Ray, kCoop>(A.P, A.K, A.rec, A.tilesX, jit, tile, lane, A.bandMask);
    }
}

// ---------------------------------------------------------------- batched views
// Wavefronts take (view, tile) items from one queue over all views.  The
// view index is wavefront-uniform, so its descriptor is read through scalar
// loads; per-view completion times are stamped for latency reporting.
//
// The loop has one lane-0 region per iteration, at its top and before the
// cross-lane read of the queue item: the completion count of the tile the
// wavefront finished last, then the next queue item.  (A lane-0 region at
// the end of the body -- the completion count after the trace -- let the
// compiler merge it with the next iteration's queue atomic on the loop's
// exit path: lanes 1..63 re-entered the loop while lane 0 was parked there,
// read the next item from the inactive lane 0 as 0, and traced tile 0 of
// view 0 forever.)  Every wavefront also leaves after totalTiles + 1 items
// at most, an exit condition each lane reaches on its own count.
template <class SRay, class KRay>
__global__ __launch_bounds__(256) void traceBatchKernel(const ViewDesc* __restrict__ views, BatchArgs A)
{
    const unsigned lane = __lane_id();
    int prev = -1;      // view of the tile this wavefront finished last (wavefront-uniform)
    for (int it = 0; it <= A.totalTiles; ++it)
    {
        unsigned t = 0;
        if (lane == 0)
        {
            if (prev >= 0)
            {
                // all of this wavefront's record stores are issued before the count
                __threadfence();
                const unsigned d = atomicAdd(&A.viewDone[prev], 1u) + 1u;
                // the view's FP64 band rays may finish later: the later stamp stays
                if (d == (unsigned)views[prev].tileCount) atomicMax(&A.stamps[prev], (unsigned long long)wall_clock64());
            }
            t = atomicAdd(A.counter, 1u);
        }
        t = (unsigned)__builtin_amdgcn_readfirstlane((int)__shfl(t, 0));
        if (t >= (unsigned)A.totalTiles) break;
        // the batch tile at this queue position (costliest first when the
        // batch's tiles were ordered by their cost keys)
        if (A.order) t = (unsigned)__builtin_amdgcn_readfirstlane((int)A.order[t]);
        int lo = 0, hi = A.nViews - 1;
        while (lo < hi)
        {
            const int mid = (lo + hi + 1) >> 1;
            if ((unsigned)views[mid].tileBase <= t) lo = mid;
            else hi = mid - 1;
        }
        const int v = __builtin_amdgcn_readfirstlane(lo);
        const ViewDesc& V = views[v];
        const unsigned tile = t - (unsigned)V.tileBase;
        if (V.P.model == BH_KERR || V.P.model == BH_KERR_NEWMAN)
            traceTile<KRay, false>(V.P, V.K, V.rec, V.tilesX, V.jitter, tile, lane, V.bandMask);
        else traceTile<SRay, false>(V.P, V.K, V.rec, V.tilesX, V.jitter, tile, lane, V.bandMask);
        prev = v;
    }
}

// ---------------------------------------------------------------- FP64 band rays
// Band rays come from the classification list through a queue, one ray per
// lane.  Band rays differ widely in length, so a lane that finishes takes
// the next ray instead of idling until the wavefront's longest ray ends:
// every kBandChunk steps the wavefront refills once at least kBandRefill of
// its lanes are idle (one queue atomic for all of them).  The geometry is
// double precision; the disk chords are sampled in float with the object's
// policy, exactly as for the FP32 rays.
constexpr int kBandChunk = 32;
constexpr int kBandRefill = 16;
// band lists at least this long (64 full wavefronts) are spread over fewer
// wavefronts than they have 64-ray chunks
constexpr unsigned kBandLongList = 4096u;

// At most 256 VGPRs (two wavefronts per SIMD): a band wavefront shares its
// SIMD with the FP32 tile grid, whose wavefronts hold 128 VGPRs each, so the
// band wavefront's allocation sets how many tile wavefronts it displaces --
// two at 256 (a few FP64 values spill to scratch), three at the 328 the
// compiler picks unbounded.
template <class Ray>
__global__ __launch_bounds__(64) __attribute__((amdgpu_waves_per_eu(2))) void traceBandKernel(BandArgs A)
{
    const unsigned n = *A.count;
    {
        // Wavefronts beyond what the list needs leave at once.  A long list
        // is spread at 1.25 rays per lane: its last chunks hold its shortest
        // rays, which lanes of the first wavefronts take as they free up, so
        // fewer band wavefronts hold tile slots (1080p Kerr frame: trace
        // 9.22 -> 9.12 ms, identical records).  A short list keeps one ray per
        // lane: its rays are all long and the band kernel is the frame's
        // critical path (256^2 Schwarzschild frame: 1.2 -> 1.8 ms at 1.25).
        const unsigned want = n >= kBandLongList ? (n + 79u) / 80u : (n + 63u) / 64u;
        if (blockIdx.x >= (want > 0u ? want : 1u)) return;
    }
    // rays sorted longest first: the wavefronts that take the first chunks
    // run longest, the later ones carry short rays and soon give their
    // SIMD slots back to the FP32 tile grid
    const unsigned* list = A.sorted ? A.sorted : A.list;
    // band rays are the longest of the frame and run beside the FP32 tile
    // kernel: give their wavefronts issue priority so the frame is not
    // waiting on them
    __builtin_amdgcn_s_setprio(3);
    const unsigned lane = __lane_id();
    const d3 o = dk3(A.P.camPos.x, A.P.camPos.y, A.P.camPos.z);
    const float2 jit = launchJitter(A.P, A.jitter, A.dynIndex);
    Ray R;
    Chord ch;
    ch.n = 0;
    bool active = false, drained = false;
    unsigned pix = 0;
    for (;;)
    {
        const bool need = !active && !drained;
        const unsigned long long needMask = __ballot(need);
        if (needMask)
        {
            const int leader = __ffsll((long long)needMask) - 1;
            unsigned base = 0;
            if ((int)lane == leader) base = atomicAdd(A.queue, (unsigned)__popcll(needMask));
            base = __shfl(base, leader);
            if (need)
            {
                const unsigned id = base + (unsigned)__popcll(needMask & ((1ull << lane) - 1ull));
                if (id >= n) drained = true;
                else
                {
                    pix = list[id];
                    const int px = (int)(pix % (unsigned)A.P.width), py = (int)(pix / (unsigned)A.P.width);
                    R.init(A.B, o, cameraDir64(A.P, px, py, jit));
                    active = true;
                }
            }
        }
        if (__ballot(active) == 0ull)
        {
            if (__ballot(!drained) == 0ull) break;
            continue;
        }
        for (int k = 0; k < kBandChunk; ++k)
        {
            if (active)
            {
                bool done = (R.status != RAY_RUNNING);
                if (!done)
                {
                    done = R.advance(A.B, A.D, ch);
                    if (ch.n > 0) diskChordSerial<Pol>(A.D, R.diskE(), R.diskLz(), ch, R.acc, R.trans);
                }
                if (done)
                {
                    float4* out = reinterpret_cast<float4*>(A.rec + pix);
                    const int flags = (R.status == RAY_ESCAPED ? 1 : 2) | (R.step << 8);
                    out[0] = make_float4((float)R.escDir.x, (float)R.escDir.y, (float)R.escDir.z,
                                         __int_as_float(flags));
                    out[1] = make_float4(R.acc.x, R.acc.y, R.acc.z, R.trans);
                    active = false;
                }
            }
            const unsigned long long idle = __ballot(!active && !drained);
            if (__popcll(idle) >= kBandRefill || __ballot(active) == 0ull) break;
        }
    }
}

// ---------------------------------------------------------------- FP64 band rays of a batch
// One ray of a view's band list (entry j of the sorted list when there is
// one, when the list has that many), traced to completion: the same
// integration and record as traceBandKernel's.
template <class Ray> BHK_INL void traceBandListRay(const BandArgs& V, unsigned j)
{
    if (j >= *V.count) return;
    const unsigned pix = (V.sorted ? V.sorted : V.list)[j];
    const int px = (int)(pix % (unsigned)V.P.width), py = (int)(pix / (unsigned)V.P.width);
    const d3 o = dk3(V.P.camPos.x, V.P.camPos.y, V.P.camPos.z);
    const float2 jit = launchJitter(V.P, V.jitter, V.dynIndex);
    Ray R;
    R.init(V.B, o, cameraDir64(V.P, px, py, jit));
    Chord ch;
    ch.n = 0;
    bool done = (R.status != RAY_RUNNING);
    while (!done)
    {
        done = R.advance(V.B, V.D, ch);
        if (ch.n > 0) diskChordSerial<Pol>(V.D, R.diskE(), R.diskLz(), ch, R.acc, R.trans);
    }
    float4* out = reinterpret_cast<float4*>(V.rec + pix);
    const int flags = (R.status == RAY_ESCAPED ? 1 : 2) | (R.step << 8);
    out[0] = make_float4((float)R.escDir.x, (float)R.escDir.y, (float)R.escDir.z, __int_as_float(flags));
    out[1] = make_float4(R.acc.x, R.acc.y, R.acc.z, R.trans);
}

// Wavefronts take 64-ray chunks of the views' band lists from one queue; a
// chunk's view is wavefront-uniform, so its constants come through scalar
// loads, and each lane traces one ray of the chunk.  The loop has the shape
// of traceBatchKernel's: one lane-0 region per iteration, at its top and
// before the cross-lane read of the queue item (the completion count of the
// chunk finished last, then the next chunk), and every wavefront leaves
// after chunks + 1 items at most, a bound each lane counts itself.  Band
// wavefronts run beside the FP32 batch trace at raised issue priority, as
// traceBandKernel's do, and at most 256 VGPRs.
template <class SRay, class KRay>
__global__ __launch_bounds__(64) __attribute__((amdgpu_waves_per_eu(2))) void traceBandBatchKernel(BatchBandArgs A)
{
    __builtin_amdgcn_s_setprio(3);
    const unsigned lane = __lane_id();
    const unsigned total = A.chunkBase[A.nViews];
    int prev = -1;      // view of the chunk this wavefront finished last (wavefront-uniform)
    for (unsigned it = 0; it <= total; ++it)
    {
        unsigned c = 0;
        if (lane == 0)
        {
            if (prev >= 0)
            {
                // all of this wavefront's record stores are issued before the count
                __threadfence();
                const unsigned d = atomicAdd(&A.bandDone[prev], 1u) + 1u;
                if (d == A.chunkBase[prev + 1] - A.chunkBase[prev])
                    atomicMax(&A.stamps[prev], (unsigned long long)wall_clock64());
            }
            c = atomicAdd(A.queue, 1u);
        }
        c = (unsigned)__builtin_amdgcn_readfirstlane(__shfl((int)c, 0));
        if (c >= total) break;
        // the chunk's view: the last with chunkBase <= c (views without band
        // rays share the next view's base and come before it)
        int lo = 0, hi = A.nViews - 1;
        while (lo < hi)
        {
            const int mid = (lo + hi + 1) >> 1;
            if (A.chunkBase[mid] <= c) lo = mid;
            else hi = mid - 1;
        }
        const int v = __builtin_amdgcn_readfirstlane(lo);
        const BandArgs& V = A.views[v];
        const unsigned j = (c - A.chunkBase[v]) * (unsigned)kBandChunkRays + lane;
        if (V.B.kerr) traceBandListRay<KRay>(V, j);
        else traceBandListRay<SRay>(V, j);
        prev = v;
    }
}

template <class Ray>
__global__ __launch_bounds__(256) void probeKernel(ProbeArgs A)
{
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= A.n) return;
    const RenderParams& P = A.P;
    const FrameK& K = A.K;
    v3 dir;
    if (A.dirs)
    {
        bh_float3 d = A.dirs[i];
        dir = mk3(d.x, d.y, d.z);
    }
    else
    {
        int ix = i % A.nx, iy = i / A.nx;
        float fx = Pol::div((float)ix + 0.5f, (float)A.nx);
        float fy = Pol::div((float)iy + 0.5f, (float)A.ny);
        dir = cameraRayDir<Pol>(P, fx, fy);
    }
    Ray R;
    R.init(K, mk3(P.camPos.x, P.camPos.y, P.camPos.z), dir);
    Chord ch;
    bool done = (R.status != RAY_RUNNING);
    while (!done)
    {
        done = R.advance(K, ch);
        if (ch.n > 0) diskChordSerial<Pol>(K.D, R.E, R.Lz, ch, R.acc, R.trans);
    }
    BhProbe o;
    o.outcome = (R.status == RAY_ESCAPED) ? 1 : 0;
    o.escDir[0] = R.escDir.x; o.escDir[1] = R.escDir.y; o.escDir[2] = R.escDir.z;
    o.maxHviol = rayMaxHviol(R);
    o.lum = R.acc.x + R.acc.y + R.acc.z;
    o.steps = R.step;
    o.trans = R.trans;
    A.out[i] = o;
}

template <class Ray> static hipError_t launchTraceT(const TraceArgs& Ain, const LaunchShape& S, hipStream_t st)
{
    TraceArgs A = Ain;
    A.K = makeFrame(A.P);
    dim3 block(S.wgSize), grid(S.blocks);
    const bool tiles = A.refillMin >= 64;
    if (tiles && A.coop) hipLaunchKernelGGL((traceTileKernel<Ray, true>), grid, block, 0, st, A);
    else if (tiles) hipLaunchKernelGGL((traceTileKernel<Ray, false>), grid, block, 0, st, A);
    else if (A.coop) hipLaunchKernelGGL((traceKernel<Ray, true>), grid, block, 0, st, A);
    else hipLaunchKernelGGL((traceKernel<Ray, false>), grid, block, 0, st, A);
    return hipGetLastError();
}

template <class Ray> static int occupancyT(int wgSize, int coop, int tiles)
{
    int nb = 0;
    hipError_t e;
    if (tiles && coop) e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, traceTileKernel<Ray, true>, wgSize, 0);
    else if (tiles) e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, traceTileKernel<Ray, false>, wgSize, 0);
    else if (coop) e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, traceKernel<Ray, true>, wgSize, 0);
    else e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, traceKernel<Ray, false>, wgSize, 0);
    if (e != hipSuccess) return 1;
    return nb > 0 ? nb : 1;
}

hipError_t BHK_SUFFIX(launchTrace)(const TraceArgs& A, const LaunchShape& S, hipStream_t st)
{
    const bool kerr = (A.P.model == BH_KERR || A.P.model == BH_KERR_NEWMAN);
    return kerr ? launchTraceT<KerrRay<Pol, false>>(A, S, st) : launchTraceT<SphRay<Pol>>(A, S, st);
}

int BHK_SUFFIX(traceOccupancy)(int kerr, int wgSize, int coop, int tiles)
{
    return kerr ? occupancyT<KerrRay<Pol, false>>(wgSize, coop, tiles) : occupancyT<SphRay<Pol>>(wgSize, coop, tiles);
}

hipError_t BHK_SUFFIX(launchTraceBatch)(const BatchArgs& A, const LaunchShape& S, hipStream_t st)
{
    using SR = SphRay<Pol>;
    using KR = KerrRay<Pol, false>;
    hipLaunchKernelGGL((traceBatchKernel<SR, KR>), dim3(S.blocks), dim3(S.wgSize), 0, st, A.views, A);
    return hipGetLastError();
}

int BHK_SUFFIX(traceBatchOccupancy)(int wgSize)
{
    int nb = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, traceBatchKernel<SphRay<Pol>, KerrRay<Pol, false>>, wgSize,
                                                     0) != hipSuccess)
        return 1;
    return nb > 0 ? nb : 1;
}

#if BHK_FAST
// ---------------------------------------------------------------- surrogate training set
// One ray per thread; rays of a configuration are contiguous (raysPerCfg is
// a multiple of 64, so a wavefront shares one configuration).  psi is
// uniform in (0, psiMax], alpha uniform in (-pi, pi], from a counter-based
// hash of (seed, ray index).  The ray is traced by the plain FP32 Kerr
// integrator with the configuration's disk geometry; the disk-hit label is
// set when a sub-sample of one of its chords passes the geometric gates of
// the volumetric sampler (the radiance itself is not evaluated).
BHK_INL float uhash(unsigned long long seed, unsigned long long i, unsigned k)
{
    unsigned long long z = seed + 0x9E3779B97F4A7C15ull * (i * 4ull + k + 1ull);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    z ^= z >> 31;
    return (float)((z >> 40) + 0.5) * (1.0f / 16777216.0f);
}

__global__ __launch_bounds__(256) void surrogateDatasetKernel(DatasetArgs A)
{
    const unsigned long long i = (unsigned long long)blockIdx.x * blockDim.x + threadIdx.x;
    const unsigned long long n = (unsigned long long)A.nCfg * (unsigned long long)A.raysPerCfg;
    if (i >= n) return;
    const int cfg = __builtin_amdgcn_readfirstlane((int)(i / (unsigned long long)A.raysPerCfg));
    const FrameK& K = A.frames[cfg];
    const RenderParams& P = A.params[cfg];
    const v3 o = mk3(P.camPos.x, P.camPos.y, P.camPos.z);
    v3 er, eth, eph;
    float rc, thc;
    camBasis(o, er, eth, eph, rc, thc);
    const float psi = A.psiMax * uhash(A.seed, i, 0);
    const float alpha = 3.14159265358979f * (2.f * uhash(A.seed, i, 1) - 1.f);
    const float sp = sinf(psi), cpsi = cosf(psi), sa = sinf(alpha), ca = cosf(alpha);
    const v3 dir = unit3(add3(scale3(er, -cpsi), add3(scale3(eth, -sp * ca), scale3(eph, sp * sa))));
    const RayFrame F = rayFrame(dir, er, eth, eph, rc);
    float x[kMlpInputs];
    surrogateInputs(F, P.aSpin / P.M, P.Qc / P.M, rc, thc, x);
    for (int k = 0; k < kMlpInputs; ++k) A.inputs[i * kMlpInputs + k] = x[k];

    KerrRay<MathFast, false> R;
    R.init(K, o, dir);
    Chord ch;
    bool done = (R.status != RAY_RUNNING);
    float disk = 0.f, hitR = 0.f;
    while (!done)
    {
        done = R.advance(K, ch);
        if (ch.n > 0 && disk == 0.f)
        {
            // disk-hit label: some sub-sample of the chord passes the
            // geometric gates (without it the disk cannot touch the ray)
            for (int j = 0; j < ch.n; ++j)
            {
                const v3 p = lerp3(ch.a, ch.b, MathFast::div((float)j + 0.5f, (float)ch.n));
                if (diskSubSampleCanContribute<MathFast>(K.D, p))
                {
                    disk = 1.f;
                    hitR = sqrtf(p.x * p.x + p.z * p.z);
                    break;
                }
            }
        }
    }
    MlpTarget t;
    {
        // band rays (critical curve, polar turning point) carry no regression target
        double delta, polar;
        const BandK& Bk = A.bands[cfg];
        const bool ok = bandStats(Bk, dk3(o.x, o.y, o.z), dk3(dir.x, dir.y, dir.z), delta, polar);
        t.band = bandDecision(Bk, ok, delta, polar) ? 1.f : 0.f;
    }
    t.escaped = (R.status == RAY_ESCAPED) ? 1.f : 0.f;
    t.disk = disk;
    t.hitR = hitR;
    t.d0 = 0.f;
    t.d1 = 0.f;
    t.pad[0] = t.pad[1] = 0.f;
    if (t.escaped > 0.5f)
    {
        float dA, dB;
        exitDeflection(F, R.escDir, dA, dB);
        t.d0 = dA - sphDeflection(P.M, P.Qc, rc, F.psi, 120.f * P.M);
        t.d1 = dB;
    }
    A.targets[i] = t;
    if (A.escDir)
    {
        A.escDir[i * 3 + 0] = R.escDir.x;
        A.escDir[i * 3 + 1] = R.escDir.y;
        A.escDir[i * 3 + 2] = R.escDir.z;
    }
}

hipError_t launchSurrogateDataset(const DatasetArgs& A, hipStream_t st)
{
    const long long n = (long long)A.nCfg * A.raysPerCfg;
    if (n <= 0) return hipSuccess;
    if (A.raysPerCfg % 64) return hipErrorInvalidValue;
    hipLaunchKernelGGL(surrogateDatasetKernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, A);
    return hipGetLastError();
}
#endif

hipError_t BHK_SUFFIX(launchBandTrace)(const BandArgs& A, int blocks, hipStream_t st)
{
    using D = typename std::conditional<Pol::fast, D64Fast, D64Exact>::type;
    using KR = KerrRay64<Pol, D>;
    using SR = SphRay64<Pol, D>;
    if (A.B.kerr) hipLaunchKernelGGL(traceBandKernel<KR>, dim3(blocks), dim3(64), 0, st, A);
    else hipLaunchKernelGGL(traceBandKernel<SR>, dim3(blocks), dim3(64), 0, st, A);
    return hipGetLastError();
}

hipError_t BHK_SUFFIX(launchBandTraceBatch)(const BatchBandArgs& A, int blocks, hipStream_t st)
{
    using D = typename std::conditional<Pol::fast, D64Fast, D64Exact>::type;
    hipLaunchKernelGGL((traceBandBatchKernel<SphRay64<Pol, D>, KerrRay64<Pol, D>>), dim3(blocks), dim3(64), 0, st, A);
    return hipGetLastError();
}

hipError_t BHK_SUFFIX(launchProbe)(const ProbeArgs& Ain, hipStream_t st)
{
    ProbeArgs A = Ain;
    A.K = makeFrame(A.P);
    const bool kerr = (A.P.model == BH_KERR || A.P.model == BH_KERR_NEWMAN);
    dim3 block(128), grid((A.n + 127) / 128);
    if (A.n <= 0) return hipSuccess;
    using KerrProbe = KerrRay<Pol, true>;
    using SphProbe = SphRay<Pol>;
    if (kerr) hipLaunchKernelGGL(probeKernel<KerrProbe>, grid, block, 0, st, A);
    else hipLaunchKernelGGL(probeKernel<SphProbe>, grid, block, 0, st, A);
    return hipGetLastError();
}

}  // namespace bhk
