// ---------------------------------------------------------------------------
// bh_band.hip -- FP64 critical-band classification of every pixel.
//
// One thread per pixel evaluates the ray's FP64 invariants and band tests
// (bh_band.h), writes the pixel's mask byte (read by the FP32 tile tracer,
// which skips band pixels) and appends band pixels to a list with one
// atomic per wavefront.  Pure FP64 arithmetic without contraction; compiled
// once.
// ---------------------------------------------------------------------------
#include "bh_internal.h"

namespace bhk {

__global__ __launch_bounds__(256) void bandClassifyKernel(BandArgs A)
{
    const int W = A.P.width, H = A.P.height;
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    bool band = false;
    if (i < W * H)
    {
        const int px = i % W, py = i / W;
        const d3 o = dk3(A.P.camPos.x, A.P.camPos.y, A.P.camPos.z);
        band = inCriticalBand(A.B, o, cameraDir64(A.P, px, py, launchJitter(A.P, A.jitter, A.dynIndex)));
        A.mask[i] = band ? 1 : 0;
    }
    const unsigned long long m = __ballot(band);
    if (m == 0ull) return;
    const unsigned lane = __lane_id();
    unsigned base = 0;
    if (lane == 0) base = atomicAdd(A.count, (unsigned)__popcll(m));
    base = (unsigned)__shfl((int)base, 0);
    if (band)
    {
        const unsigned below = (unsigned)__popcll(m & ((1ull << lane) - 1ull));
        A.list[base + below] = (unsigned)i;
    }
}

hipError_t launchBandClassify(const BandArgs& A, hipStream_t st)
{
    const int n = A.P.width * A.P.height;
    if (n <= 0) return hipSuccess;
    hipLaunchKernelGGL(bandClassifyKernel, dim3((n + 255) / 256), dim3(256), 0, st, A);
    return hipGetLastError();
}

}  // namespace bhk
