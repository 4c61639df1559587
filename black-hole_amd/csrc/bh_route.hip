// ---------------------------------------------------------------------------
// bh_route.hip -- surrogate inference fused with the router.
//
// A workgroup of four wavefronts runs the whole network for 128 candidate
// rays at a time (32 per wavefront): Fourier features into LDS, then the
// five layers as 16x16x32 BF16 MFMA tiles.  Each layer's weights are staged
// once per pass into LDS by the whole workgroup (16-byte copies, rows padded
// by 8 elements) and read from there as B operands by all four wavefronts;
// activations round-trip through each wavefront's own LDS rows in the
// A-operand layout, with FP32 accumulation, bias and SiLU in the epilogue.
// Finally the routing decision per ray: a ray is
// accepted when its predicted disk-hit probability and exit-direction spread
// are both below the thresholds; it then gets its record (exit direction
// from the predicted deflection, no disk light) and mask value 2, so the RK4
// tracers skip it.  Rejected rays keep mask 0 and are traced by RK4.
// ---------------------------------------------------------------------------
#include <algorithm>

#include "bh_internal.h"
#include "bh_surrogate_geom.h"

namespace bhk {

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

constexpr int kRouteRows = 32;                 // rays per wavefront pass
constexpr int kStrideA = kMlpIn + 8;           // LDS row strides (bf16), 16-byte multiples
constexpr int kStrideB = kMlpHidden + 8;
constexpr int kRouteWaves = 4;                 // wavefronts per workgroup

struct alignas(16) RouteLds
{
    __bf16 a[kRouteRows * kStrideA];
    __bf16 b[kRouteRows * kStrideB];
    float out[kRouteRows * 8];
};
// staged weights of one layer: out rows of in + 8 elements (layer 0 is the largest)
constexpr int kRouteWElems = kMlpHidden * kStrideA;

__global__ __launch_bounds__(64 * kRouteWaves) void surrogateRouteKernel(RouteArgs A)
{
    __shared__ RouteLds lds[kRouteWaves];
    __shared__ __attribute__((aligned(16))) __bf16 Wlds[kRouteWElems];
    const unsigned n = *A.count;
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    RouteLds& L = lds[wave];
    const int r16 = lane & 15, kq = (lane >> 4) * 8;
    const float2 jit = launchJitter(A.P, A.jitter, A.dynIndex);
    const unsigned step = gridDim.x * (kRouteRows * kRouteWaves);
    for (unsigned base0 = blockIdx.x * (kRouteRows * kRouteWaves); base0 < n; base0 += step)
    {
        const unsigned base = base0 + wave * kRouteRows;
        // ---- Fourier features: two lanes per row, three inputs each
        {
            const int row = lane >> 1, half = lane & 1;
            const bool valid = base + row < n;
            __bf16* o = L.a + row * kStrideA;
            for (int dd = 0; dd < 3; ++dd)
            {
                const int d = half * 3 + dd;
                const float v = valid ? A.inputs[(size_t)(base + row) * kMlpInputs + d] : 0.f;
                for (int f = 0; f < kMlpFreqs; ++f)
                {
                    float s, c;
                    __sincosf(mlpFreq(f) * v, &s, &c);
                    o[(d * kMlpFreqs + f) * 2 + 0] = (__bf16)(valid ? s : 0.f);
                    o[(d * kMlpFreqs + f) * 2 + 1] = (__bf16)(valid ? c : 0.f);
                }
                o[kMlpFF + d] = (__bf16)v;
            }
            if (half)
                for (int c = kMlpFF + kMlpInputs; c < kMlpIn; ++c) o[c] = (__bf16)0.f;
        }
        // ---- layers
        for (int l = 0; l < kMlpLayers; ++l)
        {
            const int in = mlpLayerIn(l), out = mlpLayerOut(l);
            const int ws = in + 8;
            const __bf16* src = (l & 1) ? L.b : L.a;
            const int ss = (l & 1) ? kStrideB : kStrideA;
            __bf16* dst = (l & 1) ? L.a : L.b;
            const int ds = (l & 1) ? kStrideA : kStrideB;
            const float* bl = A.Pf + A.bOff[l];
            // stage the layer's weights (every wavefront is past the previous
            // layer's reads of Wlds)
            __syncthreads();
            {
                const __bf16* Wl = A.W + A.wOff[l];
                const int vecPerRow = in / 8;
                for (int q = threadIdx.x; q < out * vecPerRow; q += 64 * kRouteWaves)
                {
                    const int o = q / vecPerRow, c = (q - o * vecPerRow) * 8;
                    *reinterpret_cast<bf16x8*>(Wlds + o * ws + c) =
                        *reinterpret_cast<const bf16x8*>(Wl + (size_t)o * in + c);
                }
            }
            __syncthreads();
            // all output tiles at once: 2 x nOt independent accumulators, the
            // A fragments of a k step read once
            const int nOt = out / 16;
            f32x4 acc[kMlpHidden / 16][2];
#pragma unroll
            for (int ot = 0; ot < kMlpHidden / 16; ++ot)
                acc[ot][0] = acc[ot][1] = f32x4{0.f, 0.f, 0.f, 0.f};
            for (int kb = 0; kb < in; kb += 32)
            {
                const bf16x8 a0 = *reinterpret_cast<const bf16x8*>(src + r16 * ss + kb + kq);
                const bf16x8 a1 = *reinterpret_cast<const bf16x8*>(src + (16 + r16) * ss + kb + kq);
#pragma unroll
                for (int ot = 0; ot < kMlpHidden / 16; ++ot)
                {
                    if (ot < nOt)
                    {
                        const bf16x8 bw = *reinterpret_cast<const bf16x8*>(Wlds + (ot * 16 + r16) * ws + kb + kq);
                        acc[ot][0] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a0, bw, acc[ot][0], 0, 0, 0);
                        acc[ot][1] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a1, bw, acc[ot][1], 0, 0, 0);
                    }
                }
            }
#pragma unroll
            for (int ot = 0; ot < kMlpHidden / 16; ++ot)
            {
                if (ot >= nOt) continue;
                const int col = ot * 16 + r16;
                const float bias = bl[col];
#pragma unroll
                for (int v = 0; v < 4; ++v)
                {
                    const int row0 = (lane >> 4) * 4 + v;
                    const float z0 = acc[ot][0][v] + bias, z1 = acc[ot][1][v] + bias;
                    if (l < kMlpLayers - 1)
                    {
                        dst[row0 * ds + col] = (__bf16)(z0 / (1.f + __expf(-z0)));
                        dst[(16 + row0) * ds + col] = (__bf16)(z1 / (1.f + __expf(-z1)));
                    }
                    else if (col < 8)
                    {
                        L.out[row0 * 8 + col] = z0;
                        L.out[(16 + row0) * 8 + col] = z1;
                    }
                }
            }
        }
        __syncthreads();   // the head outputs of every lane are in L.out
        // ---- routing: one lane per ray
        bool acc = false;
        if (lane < kRouteRows && base + lane < n)
        {
            const float* y = L.out + lane * 8;
            const float pDisk = 1.f / (1.f + __expf(-y[2]));
            const float sigma = kMlpSigma0 * __expf(0.5f * fminf(fmaxf(y[4], -20.f), 20.f));
            if (pDisk < A.diskProbMax && sigma < A.sigmaMax)
            {
                const unsigned pix = A.list[base + lane];
                const int W = A.P.width;
                const int px = (int)(pix % (unsigned)W), py = (int)(pix / (unsigned)W);
                const float fx = MathExact::div((float)px + jit.x, (float)W);
                const float fy = MathExact::div((float)py + jit.y, (float)A.P.height);
                const v3 d = cameraRayDir<MathExact>(A.P, fx, fy);
                v3 er, eth, eph;
                float rc, thc;
                camBasis(mk3(A.P.camPos.x, A.P.camPos.y, A.P.camPos.z), er, eth, eph, rc, thc);
                const RayFrame F = rayFrame(d, er, eth, eph, rc);
                const v3 e = exitFromDeflection(
                    F, y[0] * (1.f / kMlpOutPerRad) + sphDeflection(A.P.M, A.P.Qc, rc, F.psi, 120.f * A.P.M),
                    y[1] * (1.f / kMlpOutPerRad));
                float4* r = reinterpret_cast<float4*>(A.rec + pix);
                r[0] = make_float4(e.x, e.y, e.z, __int_as_float(1));
                r[1] = make_float4(0.f, 0.f, 0.f, 1.f);
                A.mask[pix] = 2;
                acc = true;
            }
        }
        const unsigned long long m = __ballot(acc);
        if (lane == 0 && m) atomicAdd(A.accepted, (unsigned)__popcll(m));
        __syncthreads();
    }
}

hipError_t launchSurrogateRoute(const RouteArgs& A, int maxRows, int numCU, hipStream_t st)
{
    if (maxRows <= 0) return hipSuccess;
    const int per = kRouteRows * kRouteWaves;
    const int want = (maxRows + per - 1) / per;
    const int blocks = std::max(1, std::min(want, numCU * 8));
    hipLaunchKernelGGL(surrogateRouteKernel, dim3(blocks), dim3(64 * kRouteWaves), 0, st, A);
    return hipGetLastError();
}

}  // namespace bhk
