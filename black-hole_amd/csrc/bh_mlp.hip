// ---------------------------------------------------------------------------
// bh_mlp.hip -- the learned surrogate's network on gfx950: BF16 operands on
// MFMA (v_mfma_f32_16x16x32_bf16, FP32 accumulation), plus the elementwise
// kernels of the forward pass, the loss, the backward pass and Adam.
//
// Network: 6 inputs -> Fourier features (16 frequencies x 6 inputs, sin and
// cos, plus the 6 raw inputs; 198 columns zero-padded to 224) -> 4 hidden
// layers of width 128 with SiLU -> 5 outputs (padded to 16 columns).
// Weights are kept as FP32 masters (Adam) with BF16 copies for the MFMA
// operands, stored out x in (row o holds the weights of output o).
//
// MFMA 16x16x32 operand layout (one wavefront, lane l):
//   A (16 x 32): row l % 16, k = 8 (l / 16) + j, j = 0..7  (8 bf16 per lane)
//   B (32 x 16): col l % 16, k = 8 (l / 16) + j
//   D (16 x 16): col l % 16, row = 4 (l / 16) + v, v = 0..3 (4 f32 per lane)
// ---------------------------------------------------------------------------
#include <hip/hip_runtime.h>
#include "bh_mlp.h"

namespace bhk {

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

BHK_MLP_INL __bf16 tobf(float x) { return (__bf16)x; }

// ------------------------------------------------------------------ GEMM
// C[M x N] (+)= op(A)[M x K] . op(B)[K x N].  op(A) = A (row-major, lda) or
// A^T (A stored K x M); op(B) = B (row-major K x N) or B^T (B stored N x K).
// One wavefront per 16 x 16 output tile; gridDim.z splits K (accumulated
// with atomics when split > 1 or `accumulate`).
__global__ __launch_bounds__(64) void gemmKernel(GemmArgs g)
{
    const int lane = threadIdx.x;
    const int tn = blockIdx.x, tm = blockIdx.y;
    const int r16 = lane & 15, kq = (lane >> 4) * 8;
    const int row = tm * 16 + r16;         // A row of this lane
    const int col = tn * 16 + r16;         // B column of this lane
    const int kPer = ((g.K + 31) / 32 + gridDim.z - 1) / gridDim.z * 32;
    const int k0 = blockIdx.z * kPer, k1 = min(g.K, k0 + kPer);
    f32x4 acc = {0.f, 0.f, 0.f, 0.f};
    for (int kb = k0; kb < k1; kb += 32)
    {
        bf16x8 a, b;
#pragma unroll
        for (int j = 0; j < 8; ++j)
        {
            const int k = kb + kq + j;
            __bf16 av = tobf(0.f), bv = tobf(0.f);
            if (k < k1)
            {
                if (row < g.M) av = g.transA ? g.A[(size_t)k * g.lda + row] : g.A[(size_t)row * g.lda + k];
                if (col < g.N) bv = g.transB ? g.B[(size_t)col * g.ldb + k] : g.B[(size_t)k * g.ldb + col];
            }
            a[j] = av;
            b[j] = bv;
        }
        acc = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, b, acc, 0, 0, 0);
    }
    const int oc = tn * 16 + r16;
#pragma unroll
    for (int v = 0; v < 4; ++v)
    {
        const int orow = tm * 16 + (lane >> 4) * 4 + v;
        if (orow < g.M && oc < g.N)
        {
            float* dst = g.C + (size_t)orow * g.ldc + oc;
            if (g.accumulate || gridDim.z > 1) atomicAdd(dst, acc[v]);
            else *dst = acc[v];
        }
    }
}

hipError_t launchGemm(const GemmArgs& g, int splitK, hipStream_t st)
{
    if (g.M <= 0 || g.N <= 0 || g.K <= 0) return hipSuccess;
    dim3 grid((g.N + 15) / 16, (g.M + 15) / 16, splitK < 1 ? 1 : splitK);
    hipLaunchKernelGGL(gemmKernel, grid, dim3(64), 0, st, g);
    return hipGetLastError();
}

// ------------------------------------------------------------------ features
// x: [M x 6] normalised inputs -> X0 bf16 [M x kIn] (sin, cos of each input
// at 16 frequencies, then the raw inputs, then zeros).
__global__ __launch_bounds__(256) void featuresKernel(const float* x, int M, __bf16* X0)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= M) return;
    __bf16* o = X0 + (size_t)i * kMlpIn;
    for (int d = 0; d < kMlpInputs; ++d)
    {
        const float v = x[(size_t)i * kMlpInputs + d];
        for (int f = 0; f < kMlpFreqs; ++f)
        {
            float s, c;
            __sincosf(mlpFreq(f) * v, &s, &c);
            o[(d * kMlpFreqs + f) * 2 + 0] = tobf(s);
            o[(d * kMlpFreqs + f) * 2 + 1] = tobf(c);
        }
    }
    for (int d = 0; d < kMlpInputs; ++d) o[kMlpFF + d] = tobf(x[(size_t)i * kMlpInputs + d]);
    for (int c = kMlpFF + kMlpInputs; c < kMlpIn; ++c) o[c] = tobf(0.f);
}

hipError_t launchFeatures(const float* x, int M, __bf16* X0, hipStream_t st)
{
    if (M <= 0) return hipSuccess;
    hipLaunchKernelGGL(featuresKernel, dim3((M + 255) / 256), dim3(256), 0, st, x, M, X0);
    return hipGetLastError();
}

// ------------------------------------------------------------------ bias + SiLU
// Z = Y + b (kept in FP32 for the backward pass when Z != null), A = silu(Z)
// (or Z itself when silu == 0) in bf16.  Y may alias Z.
__global__ __launch_bounds__(256) void biasActKernel(const float* Y, const float* bias, int M, int N, float* Z,
                                                     __bf16* A, int silu)
{
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= (size_t)M * N) return;
    const float z = Y[i] + bias[i % N];
    if (Z) Z[i] = z;
    const float a = silu ? z / (1.f + __expf(-z)) : z;
    A[i] = tobf(a);
}

hipError_t launchBiasAct(const float* Y, const float* bias, int M, int N, float* Z, __bf16* A, int silu,
                         hipStream_t st)
{
    const size_t n = (size_t)M * N;
    if (!n) return hipSuccess;
    hipLaunchKernelGGL(biasActKernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, Y, bias, M, N, Z, A, silu);
    return hipGetLastError();
}

// ------------------------------------------------------------------ SiLU backward
// dZ = dA * silu'(Z) in bf16 (GEMM operand) and f32 column sums into db.
__global__ __launch_bounds__(256) void siluBackKernel(const float* dA, const float* Z, int M, int N, __bf16* dZ,
                                                      float* db)
{
    __shared__ float colSum[kMlpHidden];
    for (int c = threadIdx.x; c < N; c += blockDim.x) colSum[c] = 0.f;
    __syncthreads();
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < (size_t)M * N)
    {
        const float z = Z[i];
        const float sg = 1.f / (1.f + __expf(-z));
        const float g = dA[i] * (sg * (1.f + z * (1.f - sg)));
        dZ[i] = tobf(g);
        atomicAdd(&colSum[i % N], g);
    }
    __syncthreads();
    for (int c = threadIdx.x; c < N; c += blockDim.x)
        if (colSum[c] != 0.f) atomicAdd(&db[c], colSum[c]);
}

hipError_t launchSiluBack(const float* dA, const float* Z, int M, int N, __bf16* dZ, float* db, hipStream_t st)
{
    const size_t n = (size_t)M * N;
    if (!n || N > kMlpHidden) return n ? hipErrorInvalidValue : hipSuccess;
    hipLaunchKernelGGL(siluBackKernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, dA, Z, M, N, dZ, db);
    return hipGetLastError();
}

// ------------------------------------------------------------------ loss head
// Output columns: 0, 1 exit deflection minus the spherical base (rad), 2
// disk-hit logit, 3 hit radius / 10, 4 s = log of the expected squared exit
// error over sigma0^2.  Per row (mean over the M rows of the batch):
//   all rows:               BCE(disk logit, disk)
//   escaped, not band:      kRegW x Huber(|e|; kClip) (quadratic inside the
//                           clip radius, linear beyond), and
//                           0.1 (s - log(min(|e|, kClip)^2 / sigma0^2 + 1e-6))^2
//                           with the error term held fixed (s learns the
//                           error it will make)
//   escaped, not band, disk: (hitR_pred - hitR / 10)^2
// dOut gets dL/dY for the head; dOutTrunk the same with the s column zero,
// so the log-variance head does not shape the hidden layers.
constexpr float kMlpClip = 0.02f;
constexpr float kMlpRegW = 1e3f;

__global__ __launch_bounds__(256) void lossHeadKernel(const float* Y, const float* bias, const MlpTarget* T, int M,
                                                      __bf16* dOut, __bf16* dOutTrunk, float* lossOut)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    float lReg = 0.f, lBce = 0.f, lHit = 0.f;
    if (i < M)
    {
        float y[kMlpOutPad];
#pragma unroll
        for (int c = 0; c < kMlpOutPad; ++c) y[c] = c < kMlpOut ? Y[(size_t)i * kMlpOutPad + c] + bias[c] : 0.f;
        float g[kMlpOutPad];
#pragma unroll
        for (int c = 0; c < kMlpOutPad; ++c) g[c] = 0.f;
        const MlpTarget t = T[i];
        const float invM = 1.f / (float)M;
        const float p = 1.f / (1.f + __expf(-y[2]));
        lBce = fmaxf(y[2], 0.f) - y[2] * t.disk + __logf(1.f + __expf(-fabsf(y[2])));
        g[2] = (p - t.disk) * invM;
        if (t.escaped > 0.5f && t.band < 0.5f)
        {
            const float e0 = y[0] - t.d0, e1 = y[1] - t.d1;
            const float en2 = e0 * e0 + e1 * e1;
            const float c2 = fminf(en2, kMlpClip * kMlpClip);
            if (en2 <= kMlpClip * kMlpClip)
            {
                lReg = kMlpRegW * en2;
                g[0] = 2.f * kMlpRegW * e0 * invM;
                g[1] = 2.f * kMlpRegW * e1 * invM;
            }
            else
            {
                // Huber: linear beyond the clip radius
                const float en = sqrtf(en2);
                lReg = kMlpRegW * (2.f * kMlpClip * en - kMlpClip * kMlpClip);
                g[0] = 2.f * kMlpRegW * kMlpClip * e0 / en * invM;
                g[1] = 2.f * kMlpRegW * kMlpClip * e1 / en * invM;
            }
            const float tgt = __logf(c2 / (kMlpSigma0 * kMlpSigma0) + 1e-6f);
            const float ds = y[4] - tgt;
            lReg += 0.1f * ds * ds;
            g[4] = 0.2f * ds * invM;
            if (t.disk > 0.5f)
            {
                const float eh = y[3] - t.hitR * 0.1f;
                lHit = eh * eh;
                g[3] = 2.f * eh * invM;
            }
        }
#pragma unroll
        for (int c = 0; c < kMlpOutPad; ++c)
        {
            dOut[(size_t)i * kMlpOutPad + c] = tobf(g[c]);
            dOutTrunk[(size_t)i * kMlpOutPad + c] = tobf(c == 4 ? 0.f : g[c]);
        }
    }
    // wavefront reduction of the three loss terms
    for (int off = 32; off > 0; off >>= 1)
    {
        lReg += __shfl_down(lReg, off);
        lBce += __shfl_down(lBce, off);
        lHit += __shfl_down(lHit, off);
    }
    if ((threadIdx.x & 63) == 0)
    {
        atomicAdd(&lossOut[0], lReg);
        atomicAdd(&lossOut[1], lBce);
        atomicAdd(&lossOut[2], lHit);
    }
}

// head bias gradient: db[c] = sum_i dOut[i][c] (dOut already scaled by 1/M)
__global__ __launch_bounds__(256) void headBiasGradKernel(const __bf16* dOut, int M, float* db)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    float v[kMlpOut];
#pragma unroll
    for (int c = 0; c < kMlpOut; ++c) v[c] = i < M ? (float)dOut[(size_t)i * kMlpOutPad + c] : 0.f;
    for (int off = 32; off > 0; off >>= 1)
#pragma unroll
        for (int c = 0; c < kMlpOut; ++c) v[c] += __shfl_down(v[c], off);
    if ((threadIdx.x & 63) == 0)
#pragma unroll
        for (int c = 0; c < kMlpOut; ++c) atomicAdd(&db[c], v[c]);
}

hipError_t launchLossHead(const float* Y, const float* bias, const MlpTarget* T, int M, __bf16* dOut,
                          __bf16* dOutTrunk, float* lossOut, float* dbHead, hipStream_t st)
{
    if (M <= 0) return hipSuccess;
    hipLaunchKernelGGL(lossHeadKernel, dim3((M + 255) / 256), dim3(256), 0, st, Y, bias, T, M, dOut, dOutTrunk,
                       lossOut);
    hipLaunchKernelGGL(headBiasGradKernel, dim3((M + 255) / 256), dim3(256), 0, st, dOut, M, dbHead);
    return hipGetLastError();
}

// network outputs: out[i][c] = Y[i][c] + b[c], c < kMlpOut
__global__ __launch_bounds__(256) void headOutKernel(const float* Y, const float* bias, int M, float* out)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= M) return;
#pragma unroll
    for (int c = 0; c < kMlpOut; ++c) out[(size_t)i * kMlpOut + c] = Y[(size_t)i * kMlpOutPad + c] + bias[c];
}

hipError_t launchHeadOut(const float* Y, const float* bias, int M, float* out, hipStream_t st)
{
    if (M <= 0) return hipSuccess;
    hipLaunchKernelGGL(headOutKernel, dim3((M + 255) / 256), dim3(256), 0, st, Y, bias, M, out);
    return hipGetLastError();
}

// ------------------------------------------------------------------ Adam
// p -= lr * mhat / (sqrt(vhat) + eps); BF16 copy of p written alongside.
// Bias-corrected with the step count t (>= 1).
__global__ __launch_bounds__(256) void adamKernel(float* p, const float* g, float* m, float* v, __bf16* pb, int n,
                                                  float lr, float b1, float b2, float eps, float c1, float c2)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const float gi = g[i];
    const float mi = b1 * m[i] + (1.f - b1) * gi;
    const float vi = b2 * v[i] + (1.f - b2) * gi * gi;
    m[i] = mi;
    v[i] = vi;
    const float pi = p[i] - lr * (mi * c1) / (sqrtf(vi * c2) + eps);
    p[i] = pi;
    if (pb) pb[i] = tobf(pi);
}

hipError_t launchAdam(float* p, const float* g, float* m, float* v, __bf16* pb, int n, float lr, int t,
                      hipStream_t st)
{
    if (n <= 0) return hipSuccess;
    const float b1 = 0.9f, b2 = 0.999f;
    const float c1 = 1.f / (1.f - powf(b1, (float)t)), c2 = 1.f / (1.f - powf(b2, (float)t));
    hipLaunchKernelGGL(adamKernel, dim3((n + 255) / 256), dim3(256), 0, st, p, g, m, v, pb, n, lr, b1, b2, 1e-8f,
                       c1, c2);
    return hipGetLastError();
}

// f32 -> bf16 copy (weights after upload)
__global__ __launch_bounds__(256) void toBf16Kernel(const float* x, __bf16* y, int n)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) y[i] = tobf(x[i]);
}

hipError_t launchToBf16(const float* x, __bf16* y, int n, hipStream_t st)
{
    if (n <= 0) return hipSuccess;
    hipLaunchKernelGGL(toBf16Kernel, dim3((n + 255) / 256), dim3(256), 0, st, x, y, n);
    return hipGetLastError();
}

// fill with zeros (graph-safe, no memset nodes)
__global__ __launch_bounds__(256) void zeroKernel(float* x, size_t n)
{
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) x[i] = 0.f;
}

hipError_t launchZero(float* x, size_t n, hipStream_t st)
{
    if (!n) return hipSuccess;
    hipLaunchKernelGGL(zeroKernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, x, n);
    return hipGetLastError();
}

}  // namespace bhk
