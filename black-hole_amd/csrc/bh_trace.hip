// ---------------------------------------------------------------------------
// bh_trace.hip -- geodesic trace kernels for gfx950.
//
// Compiled twice: BHK_FAST=0 (MathExact, -ffp-contract=off) and BHK_FAST=1
// (MathFast, -ffp-contract=on).  Each object exports its own launchers.
//
// traceKernel: persistent wavefront-refill tracer.  Every lane owns one ray;
// between chunks of `chunk` RK4 steps the lanes whose ray has terminated
// write the ray's record (exit direction, outcome, disk radiance,
// transmittance, step count) and take the next ray id from a device-wide
// queue counter -- one atomic per wavefront per refill.  Ray ids walk the
// image in 8x8 pixel tiles so a wavefront's rays start spatially coherent.
// Sky shading is deferred to a separate pixel-order kernel (bh_post.hip):
// it runs once per ray and would otherwise execute with most lanes idle.
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

template <class Ray>
__global__ __launch_bounds__(256) void traceKernel(TraceArgs A)
{
    const RenderParams& P = A.P;
    const FrameK K = makeFrame(P);
    const v3 cam = mk3(P.camPos.x, P.camPos.y, P.camPos.z);
    const unsigned lane = __lane_id();
    const unsigned total = (unsigned)A.totalIds;
    const int chunk = A.chunk;
    const float2 jit = A.jitter;

    Ray R;
    bool active = false, drained = false, done = false;
    int pix = 0;
    for (;;)
    {
        bool need = !active && !drained;
        unsigned long long needMask = __ballot(need);
        if (needMask)
        {
            int leader = __ffsll((long long)needMask) - 1;
            unsigned n = (unsigned)__popcll(needMask);
            unsigned base = 0;
            if ((int)lane == leader) base = atomicAdd(A.counter, n);
            base = __shfl(base, leader);
            if (need)
            {
                unsigned rank = (unsigned)__popcll(needMask & ((1ull << lane) - 1ull));
                unsigned id = base + rank;
                if (id >= total) drained = true;
                else
                {
                    int px, py;
                    pixelOfId(id, A.tilesX, px, py);
                    if (px < P.width && py < P.height)
                    {
                        float fx = Pol::div((float)px + jit.x, (float)P.width);
                        float fy = Pol::div((float)py + jit.y, (float)P.height);
                        v3 dir = cameraRayDir<Pol>(P, fx, fy);
                        R.init(K, cam, dir);
                        pix = py * P.width + px;
                        active = true;
                        done = (R.status != RAY_RUNNING);
                    }
                }
            }
        }
        if (__ballot(active) == 0)
        {
            if (__ballot(!drained) == 0) break;
            continue;
        }
        for (int k = 0; k < chunk; ++k)
        {
            if (active && !done) done = R.advance(K);
            if (__ballot(active && !done) == 0) break;
        }
        if (active && done)
        {
            writeRecord(A.rec, pix, R);
            active = false;
            done = false;
        }
    }
}

template <class Ray>
__global__ __launch_bounds__(256) void probeKernel(ProbeArgs A)
{
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= A.n) return;
    const RenderParams& P = A.P;
    const FrameK K = makeFrame(P);
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
    if (R.status == RAY_RUNNING)
        while (!R.advance(K)) {}
    BhProbe o;
    o.outcome = (R.status == RAY_ESCAPED) ? 1 : 0;
    o.escDir[0] = R.escDir.x; o.escDir[1] = R.escDir.y; o.escDir[2] = R.escDir.z;
    o.maxHviol = rayMaxHviol(R);
    o.lum = R.acc.x + R.acc.y + R.acc.z;
    o.steps = R.step;
    o.trans = R.trans;
    A.out[i] = o;
}

template <class Ray> static hipError_t launchTraceT(const TraceArgs& A, const LaunchShape& S, hipStream_t st)
{
    dim3 block(S.wgSize), grid(S.blocks);
    hipLaunchKernelGGL(traceKernel<Ray>, grid, block, 0, st, A);
    return hipGetLastError();
}

template <class Ray> static int occupancyT(int wgSize)
{
    int nb = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, traceKernel<Ray>, wgSize, 0) != hipSuccess) return 1;
    return nb > 0 ? nb : 1;
}

hipError_t BHK_SUFFIX(launchTrace)(const TraceArgs& A, const LaunchShape& S, hipStream_t st)
{
    const bool kerr = (A.P.model == BH_KERR || A.P.model == BH_KERR_NEWMAN);
    return kerr ? launchTraceT<KerrRay<Pol, false>>(A, S, st) : launchTraceT<SphRay<Pol>>(A, S, st);
}

int BHK_SUFFIX(traceOccupancy)(int kerr, int wgSize)
{
    return kerr ? occupancyT<KerrRay<Pol, false>>(wgSize) : occupancyT<SphRay<Pol>>(wgSize);
}

hipError_t BHK_SUFFIX(launchProbe)(const ProbeArgs& A, hipStream_t st)
{
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
