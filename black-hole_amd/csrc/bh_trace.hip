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
    const float2 jit = A.jitter;

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
                    if (px < P.width && py < P.height)
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
    if (A.coop) hipLaunchKernelGGL((traceKernel<Ray, true>), grid, block, 0, st, A);
    else hipLaunchKernelGGL((traceKernel<Ray, false>), grid, block, 0, st, A);
    return hipGetLastError();
}

template <class Ray> static int occupancyT(int wgSize, int coop)
{
    int nb = 0;
    hipError_t e = coop ? hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, traceKernel<Ray, true>, wgSize, 0)
                        : hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, traceKernel<Ray, false>, wgSize, 0);
    if (e != hipSuccess) return 1;
    return nb > 0 ? nb : 1;
}

hipError_t BHK_SUFFIX(launchTrace)(const TraceArgs& A, const LaunchShape& S, hipStream_t st)
{
    const bool kerr = (A.P.model == BH_KERR || A.P.model == BH_KERR_NEWMAN);
    return kerr ? launchTraceT<KerrRay<Pol, false>>(A, S, st) : launchTraceT<SphRay<Pol>>(A, S, st);
}

int BHK_SUFFIX(traceOccupancy)(int kerr, int wgSize, int coop)
{
    return kerr ? occupancyT<KerrRay<Pol, false>>(wgSize, coop) : occupancyT<SphRay<Pol>>(wgSize, coop);
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
