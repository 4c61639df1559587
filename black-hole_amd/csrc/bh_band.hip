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
#include "bh_surrogate_geom.h"

namespace bhk {

__global__ __launch_bounds__(256) void bandClassifyKernel(BandArgs A)
{
    const int W = A.P.width, H = A.P.height;
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    bool band = false, cand = false;
    float x[kMlpInputs];
    if (i < W * H)
    {
        const int px = i % W, py = i / W;
        const float2 jit = launchJitter(A.P, A.jitter, A.dynIndex);
        const d3 o = dk3(A.P.camPos.x, A.P.camPos.y, A.P.camPos.z);
        double delta, polar;
        const bool ok = bandStats(A.B, o, cameraDir64(A.P, px, py, jit), delta, polar);
        band = !ok || (!(A.B.kerr == 0 && delta == 1.0) && (fabs(delta) < A.B.critBand || polar < A.B.polarBand));
        A.mask[i] = band ? 1 : 0;
        if (A.candList && !band && delta < -A.B.critBand)
        {
            // candidate of the surrogate: escaping ingoing ray inside the training domain
            const float fx = MathExact::div((float)px + jit.x, (float)W);
            const float fy = MathExact::div((float)py + jit.y, (float)H);
            const v3 d = cameraRayDir<MathExact>(A.P, fx, fy);
            v3 er, eth, eph;
            float rc, thc;
            camBasis(mk3(A.P.camPos.x, A.P.camPos.y, A.P.camPos.z), er, eth, eph, rc, thc);
            const RayFrame F = rayFrame(d, er, eth, eph, rc);
            const float aStar = A.B.kerr ? A.P.aSpin / A.P.M : 0.f, q = A.P.Qc / A.P.M;
            cand = F.psi <= A.psiMax && rc >= kSurRMin && rc <= kSurRMax && q >= 0.f && q <= 0.6f &&
                   fabsf(aStar) <= 0.99f;
            if (cand) surrogateInputs(F, aStar, q, rc, thc, x);
        }
    }
    const unsigned lane = __lane_id();
    const unsigned long long m = __ballot(band);
    if (m != 0ull)
    {
        unsigned base = 0;
        if (lane == 0) base = atomicAdd(A.count, (unsigned)__popcll(m));
        base = (unsigned)__shfl((int)base, 0);
        if (band) A.list[base + (unsigned)__popcll(m & ((1ull << lane) - 1ull))] = (unsigned)i;
    }
    const unsigned long long mc = __ballot(cand);
    if (mc != 0ull)
    {
        unsigned base = 0;
        if (lane == 0) base = atomicAdd(A.candCount, (unsigned)__popcll(mc));
        base = (unsigned)__shfl((int)base, 0);
        if (cand)
        {
            const unsigned slot = base + (unsigned)__popcll(mc & ((1ull << lane) - 1ull));
            A.candList[slot] = (unsigned)i;
            for (int k = 0; k < kMlpInputs; ++k) A.candInputs[(size_t)slot * kMlpInputs + k] = x[k];
        }
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
