// ---------------------------------------------------------------------------
// bh_route.hip -- surrogate inference fused with the router.
//
// One persistent workgroup per CU holds the whole network in LDS for the
// kernel's lifetime: the five layers' BF16 weights (156 KB) are staged once,
// in MFMA-fragment order (each 16x32 A operand of v_mfma_f32_16x16x32_bf16 is
// one 1 KB block that the 64 lanes read with one conflict-free 16-byte load),
// together with the hidden layers' biases.  Each wavefront then pushes 64
// candidate rays at a time through the network with no further LDS writes
// and no barriers:
//
//   * the weights are the MFMA A operand (M = output units, K = input units)
//     and the rays the B operand (N = rays), so a layer's output tile D[unit][ray]
//     lands in the lanes exactly where the next layer's B operand needs it:
//     lane (g, r) holds units 4g..4g+3 of two consecutive 16-unit tiles for
//     ray r.  The next layer's K order is permuted to match when its weights
//     are staged, so activations never leave the registers;
//   * the Fourier features of the first layer are computed per K chunk
//     straight into the B operand registers (chunk c < 6 is input c, lane
//     group g its frequencies 4g..4g+3, sine and cosine);
//   * the accumulators start from the layer's bias, and SiLU is
//     z / (1 + 2^(-log2(e) z)): one exp2, one reciprocal and three ALU ops
//     per unit, rounded to BF16 where the training path rounds.
//
// Finally the routing decision per ray (one lane per ray): a ray is accepted
// when its predicted disk-hit probability and exit-direction spread are both
// below the thresholds, and is appended to a compact list.  A second kernel
// gives each listed ray its record (exit direction from the predicted
// deflection, no disk light) and mask value 2, so the RK4 tracers skip it;
// the FP64 quadrature of the deflection's spherical base runs there, on
// full lanes, and its registers stay out of the network kernel (which
// then holds 64 rays per wavefront at two wavefronts per SIMD without
// spilling).  Rejected rays keep mask 0 and are traced by RK4.  With
// `outRows` set the kernel only evaluates the network and writes the five
// head outputs of every row (the layout of bhSurrogateForward).
// ---------------------------------------------------------------------------
#include <algorithm>

#include "bh_internal.h"
#include "bh_surrogate_geom.h"

namespace bhk {

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef __bf16 bf16x4 __attribute__((ext_vector_type(4)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

constexpr int kHidTiles = kMlpHidden / 16;     // 8 output tiles of a hidden layer
constexpr int kHidChunks = kMlpHidden / 32;    // 4 K chunks of a hidden layer's input
constexpr int kInChunks = kMlpIn / 32;         // 7 K chunks of the first layer's input
constexpr int kFragElems = 64 * 8;             // one A operand: 64 lanes x 8 bf16
constexpr int kFragsL0 = kHidTiles * kInChunks;        // 56
constexpr int kFragsHid = kHidTiles * kHidChunks;      // 32 per hidden layer
constexpr int kFragsOut = kHidChunks;                  // 4 (16 padded outputs)
constexpr int kFrags = kFragsL0 + 3 * kFragsHid + kFragsOut;   // 156
constexpr float kLog2e = 1.4426950408889634f;

static_assert(kMlpIn % 32 == 0 && kMlpHidden % 32 == 0 && kMlpOutPad == 16, "fragment shapes");
static_assert(kMlpFF == 6 * 32 && kMlpInputs <= 8, "feature chunks: one input per 32-column chunk, raw inputs in chunk 6");

__device__ __forceinline__ int fragBase(int l)
{
    return l == 0 ? 0 : (l < kMlpLayers - 1 ? kFragsL0 + (l - 1) * kFragsHid : kFragsL0 + 3 * kFragsHid);
}

// silu(z) = z / (1 + 2^(-log2(e) z)), rounded to bf16 as the training
// path's epilogue rounds it.
__device__ __forceinline__ __bf16 silu(float z)
{
    return (__bf16)(z * __builtin_amdgcn_rcpf(1.f + __builtin_amdgcn_exp2f(z * -kLog2e)));
}

// Next layer's B operand (chunk c) from this layer's accumulators: units
// 32c + 4g + v (tile 2c) then 32c + 16 + 4g + v (tile 2c + 1), v = 0..3.
__device__ __forceinline__ bf16x8 packChunk(const f32x4& lo, const f32x4& hi)
{
    bf16x8 b;
#pragma unroll
    for (int v = 0; v < 4; ++v)
    {
        b[v] = silu(lo[v]);
        b[4 + v] = silu(hi[v]);
    }
    return b;
}

// kRouteWaves wavefronts per workgroup; kRayTiles 16-ray B tiles per
// wavefront pass (one lane per ray when routing).
template <int kRouteWaves, int kRayTiles>
__global__ __launch_bounds__(64 * kRouteWaves) void surrogateRouteKernel(RouteArgs A)
{
    __shared__ __attribute__((aligned(16))) __bf16 Wf[kFrags * kFragElems];
    __shared__ __attribute__((aligned(16))) float Bs[(kMlpLayers - 1) * kMlpHidden];

    // ---- stage the network once: weights in fragment order, K permuted for
    //      layers fed by registers; hidden biases
    for (int q = threadIdx.x; q < kFrags * 64; q += 64 * kRouteWaves)
    {
        const int frag = q >> 6, ln = q & 63;
        int l, mt, kc;
        if (frag < kFragsL0)
        {
            l = 0;
            mt = frag / kInChunks;
            kc = frag - mt * kInChunks;
        }
        else if (frag < kFragsL0 + 3 * kFragsHid)
        {
            const int f = frag - kFragsL0;
            l = 1 + f / kFragsHid;
            mt = (f % kFragsHid) / kHidChunks;
            kc = f % kHidChunks;
        }
        else
        {
            l = kMlpLayers - 1;
            mt = 0;
            kc = frag - (kFragsL0 + 3 * kFragsHid);
        }
        const int in = mlpLayerIn(l), g = ln >> 4;
        const __bf16* src = A.W + A.wOff[l] + (size_t)(mt * 16 + (ln & 15)) * in + 32 * kc;
        bf16x8 v;
        if (l == 0)
            v = *reinterpret_cast<const bf16x8*>(src + 8 * g);
        else
        {
            const bf16x4 lo = *reinterpret_cast<const bf16x4*>(src + 4 * g);
            const bf16x4 hi = *reinterpret_cast<const bf16x4*>(src + 16 + 4 * g);
            v = bf16x8{lo[0], lo[1], lo[2], lo[3], hi[0], hi[1], hi[2], hi[3]};
        }
        *reinterpret_cast<bf16x8*>(Wf + (size_t)q * 8) = v;
    }
    for (int q = threadIdx.x; q < (kMlpLayers - 1) * kMlpHidden; q += 64 * kRouteWaves)
    {
        const int l = q / kMlpHidden;
        Bs[q] = A.Pf[A.bOff[l] + (q - l * kMlpHidden)];
    }
    __syncthreads();

    const unsigned n = A.count ? *A.count : (unsigned)A.rows;
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int g = lane >> 4, r16 = lane & 15;
    // this lane's four feature frequencies 1.25^(4g+i) pi/2 (the same float
    // values the training features use)
    float fq[4];
#pragma unroll
    for (int i = 0; i < 4; ++i)
        fq[i] = g == 0 ? mlpFreq(i) : g == 1 ? mlpFreq(4 + i) : g == 2 ? mlpFreq(8 + i) : mlpFreq(12 + i);
    float bOut[kMlpOut];
#pragma unroll
    for (int c = 0; c < kMlpOut; ++c) bOut[c] = A.Pf[A.bOff[kMlpLayers - 1] + c];

    constexpr unsigned kPass = 16u * kRayTiles;
    const unsigned step = gridDim.x * kRouteWaves * kPass;
    for (unsigned base = (blockIdx.x * kRouteWaves + wave) * kPass; base < n; base += step)
    {
        f32x4 acc[kHidTiles][kRayTiles];
        bf16x8 hb[kHidChunks][kRayTiles];
        // ---- layer 0: features computed per K chunk into the B registers
        {
            float x[kRayTiles][kMlpInputs];
#pragma unroll
            for (int t = 0; t < kRayTiles; ++t)
            {
                const unsigned r = base + 16 * t + r16;
#pragma unroll
                for (int d = 0; d < kMlpInputs; ++d)
                    x[t][d] = r < n ? A.inputs[(size_t)r * kMlpInputs + d] : 0.f;
            }
#pragma unroll
            for (int mt = 0; mt < kHidTiles; ++mt)
            {
                const f32x4 b = *reinterpret_cast<const f32x4*>(Bs + mt * 16 + 4 * g);
#pragma unroll
                for (int t = 0; t < kRayTiles; ++t) acc[mt][t] = b;
            }
#pragma unroll
            for (int kc = 0; kc < kInChunks; ++kc)
            {
                bf16x8 bx[kRayTiles];
#pragma unroll
                for (int t = 0; t < kRayTiles; ++t)
                {
                    if (kc < kMlpInputs)
                    {
#pragma unroll
                        for (int i = 0; i < 4; ++i)
                        {
                            float s, c;
                            __sincosf(fq[i] * x[t][kc], &s, &c);
                            bx[t][2 * i] = (__bf16)s;
                            bx[t][2 * i + 1] = (__bf16)c;
                        }
                    }
                    else
                    {
#pragma unroll
                        for (int j = 0; j < 8; ++j)
                            bx[t][j] = (__bf16)((g == 0 && j < kMlpInputs) ? x[t][j < kMlpInputs ? j : 0] : 0.f);
                    }
                }
#pragma unroll
                for (int mt = 0; mt < kHidTiles; ++mt)
                {
                    const bf16x8 a =
                        *reinterpret_cast<const bf16x8*>(Wf + ((size_t)(mt * kInChunks + kc) * 64 + lane) * 8);
#pragma unroll
                    for (int t = 0; t < kRayTiles; ++t)
                        acc[mt][t] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, bx[t], acc[mt][t], 0, 0, 0);
                }
                // keep the scheduler from hoisting every chunk's LDS loads
                // (56 fragments) to the top of the layer
                __builtin_amdgcn_sched_barrier(0);
            }
#pragma unroll
            for (int c = 0; c < kHidChunks; ++c)
#pragma unroll
                for (int t = 0; t < kRayTiles; ++t)
                    hb[c][t] = packChunk(acc[2 * c][t], acc[2 * c + 1][t]);
        }
        // ---- hidden layers 1..3: registers in, registers out
#pragma unroll 1
        for (int l = 1; l < kMlpLayers - 1; ++l)
        {
            const __bf16* Wl = Wf + (size_t)fragBase(l) * kFragElems;
#pragma unroll
            for (int mt = 0; mt < kHidTiles; ++mt)
            {
                const f32x4 b = *reinterpret_cast<const f32x4*>(Bs + l * kMlpHidden + mt * 16 + 4 * g);
#pragma unroll
                for (int t = 0; t < kRayTiles; ++t) acc[mt][t] = b;
            }
#pragma unroll
            for (int kc = 0; kc < kHidChunks; ++kc)
            {
#pragma unroll
                for (int mt = 0; mt < kHidTiles; ++mt)
                {
                    const bf16x8 a = *reinterpret_cast<const bf16x8*>(Wl + ((size_t)(mt * kHidChunks + kc) * 64 + lane) * 8);
#pragma unroll
                    for (int t = 0; t < kRayTiles; ++t)
                        acc[mt][t] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, hb[kc][t], acc[mt][t], 0, 0, 0);
                }
                __builtin_amdgcn_sched_barrier(0);
            }
#pragma unroll
            for (int c = 0; c < kHidChunks; ++c)
#pragma unroll
                for (int t = 0; t < kRayTiles; ++t) hb[c][t] = packChunk(acc[2 * c][t], acc[2 * c + 1][t]);
        }
        // ---- output layer: 16 padded units, lane (g, r) holds units 4g..4g+3 of ray r
        f32x4 y[kRayTiles];
#pragma unroll
        for (int t = 0; t < kRayTiles; ++t) y[t] = f32x4{0.f, 0.f, 0.f, 0.f};
        {
            const __bf16* Wl = Wf + (size_t)fragBase(kMlpLayers - 1) * kFragElems;
#pragma unroll
            for (int kc = 0; kc < kHidChunks; ++kc)
            {
                const bf16x8 a = *reinterpret_cast<const bf16x8*>(Wl + ((size_t)kc * 64 + lane) * 8);
#pragma unroll
                for (int t = 0; t < kRayTiles; ++t)
                    y[t] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, hb[kc][t], y[t], 0, 0, 0);
            }
        }
        // ---- one lane per ray: lane L < kPass takes ray base + L, i.e. tile
        //      L / 16 (= g) at column r16; units 0..3 sit in lane r16, unit 4
        //      in lane 16 + r16
        float o[kMlpOut];
#pragma unroll
        for (int c = 0; c < kMlpOut; ++c)
        {
            const int src = c < 4 ? r16 : 16 + r16;
            const int v = c & 3;
            float raw = __shfl(y[0][v], src);
#pragma unroll
            for (int t = 1; t < kRayTiles; ++t)
            {
                const float st = __shfl(y[t][v], src);
                raw = g == t ? st : raw;
            }
            o[c] = raw + bOut[c];
        }
        const unsigned ray = lane < (int)kPass ? base + lane : n;
        if (A.outRows)
        {
            if (ray < n)
#pragma unroll
                for (int c = 0; c < kMlpOut; ++c) A.outRows[(size_t)ray * kMlpOut + c] = o[c];
            continue;
        }
        // routing decision; accepted rays are appended (wavefront-aggregated)
        // to the list the finishing kernel turns into records
        bool accepted = false;
        if (ray < n)
        {
            const float pDisk = 1.f / (1.f + __expf(-o[2]));
            const float sigma = kMlpSigma0 * __expf(0.5f * fminf(fmaxf(o[4], -20.f), 20.f));
            accepted = pDisk < A.diskProbMax && sigma < A.sigmaMax;
        }
        const unsigned long long m = __ballot(accepted);
        if (m)
        {
            unsigned first = 0;
            if (lane == 0) first = atomicAdd(A.accepted, (unsigned)__popcll(m));
            first = __shfl(first, 0);
            if (accepted)
            {
                const unsigned k = first + (unsigned)__builtin_amdgcn_mbcnt_hi(
                                               (unsigned)(m >> 32), __builtin_amdgcn_mbcnt_lo((unsigned)m, 0u));
                A.accList[k] = make_float4(__uint_as_float(A.list[ray]), o[0], o[1], 0.f);
            }
        }
    }
}

// Records of the accepted rays: exit direction from the predicted deflection
// (network output + the spherical-metric base, FP64 quadrature), no disk
// light, mask value 2 (the RK4 tracers skip the pixel).
__global__ __launch_bounds__(256) void surrogateFinishKernel(RouteArgs A)
{
    const unsigned n = *A.accepted;
    if (blockIdx.x * 256u >= n) return;
    const float2 jit = launchJitter(A.P, A.jitter, A.dynIndex);
    v3 er, eth, eph;
    float rc, thc;
    camBasis(mk3(A.P.camPos.x, A.P.camPos.y, A.P.camPos.z), er, eth, eph, rc, thc);
    const int W = A.P.width;
    for (unsigned k = blockIdx.x * 256u + threadIdx.x; k < n; k += gridDim.x * 256u)
    {
        const float4 q = A.accList[k];
        const unsigned pix = __float_as_uint(q.x);
        const int px = (int)(pix % (unsigned)W), py = (int)(pix / (unsigned)W);
        const float fx = MathExact::div((float)px + jit.x, (float)W);
        const float fy = MathExact::div((float)py + jit.y, (float)A.P.height);
        const v3 d = cameraRayDir<MathExact>(A.P, fx, fy);
        const RayFrame F = rayFrame(d, er, eth, eph, rc);
        const v3 e = exitFromDeflection(
            F, q.y * (1.f / kMlpOutPerRad) + sphDeflection(A.P.M, A.P.Qc, rc, F.psi, 120.f * A.P.M),
            q.z * (1.f / kMlpOutPerRad));
        float4* rp = reinterpret_cast<float4*>(A.rec + pix);
        rp[0] = make_float4(e.x, e.y, e.z, __int_as_float(1));
        rp[1] = make_float4(0.f, 0.f, 0.f, 1.f);
        A.mask[pix] = 2;
    }
}

template <int kRouteWaves, int kRayTiles>
hipError_t launchRouteShape(const RouteArgs& A, int maxRows, int numCU, hipStream_t st)
{
    const int per = 16 * kRayTiles * kRouteWaves;
    const int want = (maxRows + per - 1) / per;
    // the network fills a CU's LDS: one persistent workgroup per CU at most
    const int blocks = std::max(1, std::min(want, numCU));
    hipLaunchKernelGGL((surrogateRouteKernel<kRouteWaves, kRayTiles>), dim3(blocks), dim3(64 * kRouteWaves), 0, st, A);
    return hipGetLastError();
}

hipError_t launchSurrogateRoute(const RouteArgs& A, int maxRows, int numCU, hipStream_t st)
{
    if (maxRows <= 0) return hipSuccess;
    // 8 wavefronts of 64 rays (2 per SIMD, 256 VGPRs, no spills); timed on
    // 1.94 M rows against 4x4, 8x2, 12x2 and 16x1: 0.444 against 0.453-0.531 ms
    const hipError_t e = launchRouteShape<8, 4>(A, maxRows, numCU, st);
    if (e != hipSuccess || A.outRows) return e;
    const int blocks = std::max(1, std::min((maxRows + 255) / 256, numCU * 4));
    hipLaunchKernelGGL(surrogateFinishKernel, dim3(blocks), dim3(256), 0, st, A);
    return hipGetLastError();
}

}  // namespace bhk
