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
#include <algorithm>

#include "bh_mlp.h"

namespace bhk {

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

BHK_MLP_INL __bf16 tobf(float x) { return (__bf16)x; }

// ------------------------------------------------------------------ GEMM
// C[M x N] (+)= op(A)[M x K] . op(B)[K x N].  op(A) = A (row-major, lda) or
// A^T (A stored K x M); op(B) = B (row-major K x N) or B^T (B stored N x K).
//
// Workgroups of 4 wavefronts own a 64 x kBN output tile (each wavefront a
// 32 x kBN/2 quarter: 2 x kBN/32 MFMA tiles; kBN = 64) and walk K in chunks
// of 64: both
// operand tiles are staged in LDS with k contiguous ([row][k], rows padded
// to 72 elements), so every MFMA fragment is one 16-byte LDS read whatever
// the operand's storage order.  An operand stored k-contiguous is copied
// with 16-byte loads and stores; one stored k-major is read 8 elements of a
// row at a time (16-byte loads) and transposed on its way into LDS.  The
// vector path needs 8-element-aligned rows and extents (all the network's
// shapes); other shapes load element by element with bounds checks.
// gridDim.z splits K: each split's partial tile is stored to `part` for a
// reduction, or accumulated into C with atomics.
//
// Fused epilogues (kEpiBiasAct, kEpiSiluBack) replace the separate
// bias/SiLU and SiLU-backward passes: same expressions, same rounding
// points, no f32 round trip through HBM.
constexpr int kGBM = 64, kGBN = 64, kGBK = 64;
constexpr int kGLds = kGBK + 8;    // LDS row stride (bf16): 144 bytes, 16-byte aligned

template <bool kVec, int kRows = 64>
BHK_MLP_INL void stageTile(const __bf16* __restrict__ X, int ld, int trans, int rows0, int nRows, int k0, int K,
                           __bf16* __restrict__ T, int tid)
{
    // T[r][k] = op(X)[rows0 + r][k0 + k], r < kRows, k < 64: 8 kRows groups of 8
#pragma unroll
    for (int it = 0; it < kRows / 32; ++it)
    {
        const int q = tid + 256 * it;
        if (!trans)
        {
            // X[row][k]: 8 consecutive k of one row
            const int r = q >> 3, kg = (q & 7) * 8;
            const int row = rows0 + r, k = k0 + kg;
            bf16x8 v;
            if (kVec)
            {
                if (row < nRows && k < K) v = *reinterpret_cast<const bf16x8*>(X + (size_t)row * ld + k);
                else
                {
#pragma unroll
                    for (int j = 0; j < 8; ++j) v[j] = tobf(0.f);
                }
            }
            else
            {
#pragma unroll
                for (int j = 0; j < 8; ++j) v[j] = (row < nRows && k + j < K) ? X[(size_t)row * ld + k + j] : tobf(0.f);
            }
            *reinterpret_cast<bf16x8*>(T + r * kGLds + kg) = v;
        }
        else
        {
            // X[k][row]: 8 consecutive rows of one k, transposed into T
            constexpr int kGroups = kRows / 8;
            const int kr = q / kGroups, rg = (q % kGroups) * 8;
            const int k = k0 + kr, row = rows0 + rg;
            bf16x8 v;
            if (kVec)
            {
                if (k < K && row < nRows) v = *reinterpret_cast<const bf16x8*>(X + (size_t)k * ld + row);
                else
                {
#pragma unroll
                    for (int j = 0; j < 8; ++j) v[j] = tobf(0.f);
                }
            }
            else
            {
#pragma unroll
                for (int j = 0; j < 8; ++j) v[j] = (k < K && row + j < nRows) ? X[(size_t)k * ld + row + j] : tobf(0.f);
            }
#pragma unroll
            for (int j = 0; j < 8; ++j) T[(rg + j) * kGLds + kr] = v[j];
        }
    }
}

template <bool kVec, int kBN>
__global__ __launch_bounds__(256) void gemmKernel(GemmArgs g)
{
    constexpr int kNT = kBN / 32;   // 16-column MFMA tiles per wavefront
    __shared__ __attribute__((aligned(16))) __bf16 As[kGBM * kGLds];
    __shared__ __attribute__((aligned(16))) __bf16 Bs[kBN * kGLds];
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int m0 = blockIdx.y * kGBM, n0 = blockIdx.x * kBN;
    const int wm = (wave >> 1) * 32, wn = (wave & 1) * (kBN / 2);
    const int r16 = lane & 15, kq = (lane >> 4) * 8;
    const int chunks = (g.K + kGBK - 1) / kGBK;
    const int per = (chunks + (int)gridDim.z - 1) / (int)gridDim.z;
    const int c0 = (int)blockIdx.z * per, c1 = min(chunks, c0 + per);
    f32x4 acc[2][kNT];
#pragma unroll
    for (int i = 0; i < 2; ++i)
#pragma unroll
        for (int j = 0; j < kNT; ++j) acc[i][j] = f32x4{0.f, 0.f, 0.f, 0.f};
    for (int c = c0; c < c1; ++c)
    {
        const int k0 = c * kGBK;
        stageTile<kVec>(g.A, g.lda, g.transA, m0, g.M, k0, g.K, As, tid);
        stageTile<kVec, kBN>(g.B, g.ldb, !g.transB, n0, g.N, k0, g.K, Bs, tid);
        __syncthreads();
#pragma unroll
        for (int ks = 0; ks < kGBK; ks += 32)
        {
            bf16x8 a[2], b[kNT];
#pragma unroll
            for (int i = 0; i < 2; ++i)
                a[i] = *reinterpret_cast<const bf16x8*>(As + (wm + i * 16 + r16) * kGLds + ks + kq);
#pragma unroll
            for (int j = 0; j < kNT; ++j)
                b[j] = *reinterpret_cast<const bf16x8*>(Bs + (wn + j * 16 + r16) * kGLds + ks + kq);
#pragma unroll
            for (int i = 0; i < 2; ++i)
#pragma unroll
                for (int j = 0; j < kNT; ++j)
                    acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a[i], b[j], acc[i][j], 0, 0, 0);
        }
        __syncthreads();
    }
    // D layout: column r16 of each 16 x 16 tile, rows 4 (lane / 16) + v
    const bool atom = g.accumulate || gridDim.z > 1;
#pragma unroll
    for (int j = 0; j < kNT; ++j)
    {
        const int col = n0 + wn + j * 16 + r16;
        const bool cok = col < g.N;
        float colSum = 0.f;
        const float bias = (g.epi == kEpiBiasAct && cok) ? g.bias[col] : 0.f;
#pragma unroll
        for (int i = 0; i < 2; ++i)
#pragma unroll
            for (int v = 0; v < 4; ++v)
            {
                const int row = m0 + wm + i * 16 + (lane >> 4) * 4 + v;
                if (!cok || row >= g.M) continue;
                const size_t o = (size_t)row * g.ldc + col;
                const float x = acc[i][j][v];
                if (g.epi == kEpiStore)
                {
                    if (g.part) g.part[(size_t)blockIdx.z * g.M * g.N + (size_t)row * g.N + col] = x;
                    else if (atom) atomicAdd(g.C + o, x);
                    else g.C[o] = x;
                }
                else if (g.epi == kEpiBiasAct)
                {
                    const float z = x + bias;
                    if (g.Zout) g.Zout[o] = z;
                    const float y = g.act ? z / (1.f + __expf(-z)) : z;
                    g.Aout[o] = tobf(y);
                }
                else
                {
                    const float z = g.Zin[o];
                    const float sg = 1.f / (1.f + __expf(-z));
                    const float gr = x * (sg * (1.f + z * (1.f - sg)));
                    g.dZout[o] = tobf(gr);
                    colSum += gr;
                }
            }
        if (g.epi == kEpiSiluBack)
        {
            // the four lanes of a column hold its 32 rows of this wavefront
            colSum += __shfl_xor(colSum, 16);
            colSum += __shfl_xor(colSum, 32);
            if (g.dbPart)
            {
                if (lane < 16 && cok) g.dbPart[(size_t)(blockIdx.y * 2 + (wave >> 1)) * g.N + col] = colSum;
            }
            else if (lane < 16 && cok && colSum != 0.f)
                atomicAdd(g.db + col, colSum);
        }
    }
}

static bool aligned16(const void* p) { return ((uintptr_t)p & 15u) == 0; }

// out[y * L + i] (+)= sum over p in segment y of part[p * L + i]: a workgroup
// owns 64 consecutive i and one segment of p (blockIdx.y); its 16 wavefronts
// sum interleaved slices of the segment (coalesced 64-lane rows), then the
// slices are added in LDS in a fixed order.  `add`: accumulate into out.
__global__ __launch_bounds__(1024) void reducePartsKernel(const float* part, int P, int L, float* out, int add)
{
    __shared__ float sl[16][64];
    const int c = threadIdx.x & 63, w = threadIdx.x >> 6;
    const int i = blockIdx.x * 64 + c;
    const int seg = (P + (int)gridDim.y - 1) / (int)gridDim.y;
    const int p0 = (int)blockIdx.y * seg, p1 = min(P, p0 + seg);
    float s = 0.f;
    if (i < L)
        for (int p = p0 + w; p < p1; p += 16) s += part[(size_t)p * L + i];
    sl[w][c] = s;
    __syncthreads();
    if (w == 0 && i < L)
    {
        float t = 0.f;
#pragma unroll
        for (int k = 0; k < 16; ++k) t += sl[k][c];
        float* o = out + (size_t)blockIdx.y * L + i;
        *o = add ? *o + t : t;
    }
}

hipError_t launchReduceParts(const float* part, int P, int L, float* dst, float* tmp, hipStream_t st)
{
    if (P <= 0 || L <= 0) return hipSuccess;
    const unsigned bx = (unsigned)((L + 63) / 64);
    if (P > 256 && tmp)
    {
        // few long columns: segments of 64 partials in parallel, then their sums
        const int S = (P + 63) / 64;
        hipLaunchKernelGGL(reducePartsKernel, dim3(bx, S), dim3(1024), 0, st, part, P, L, tmp, 0);
        hipError_t e = hipGetLastError();
        if (e != hipSuccess) return e;
        part = tmp;
        P = S;
    }
    hipLaunchKernelGGL(reducePartsKernel, dim3(bx, 1), dim3(1024), 0, st, part, P, L, dst, 1);
    return hipGetLastError();
}

hipError_t launchGemm(const GemmArgs& g, int splitK, hipStream_t st)
{
    if (g.M <= 0 || g.N <= 0 || g.K <= 0) return hipSuccess;
    if (splitK < 1) splitK = 1;
    if (g.epi != kEpiStore && (splitK > 1 || g.accumulate)) return hipErrorInvalidValue;
    // vector staging: every operand row starts 16-byte aligned and its
    // contiguous extent is a multiple of 8 elements
    const int contA = g.transA ? g.M : g.K, contB = g.transB ? g.K : g.N;
    const bool vec = aligned16(g.A) && aligned16(g.B) && g.lda % 8 == 0 && g.ldb % 8 == 0 && contA % 8 == 0 &&
                     contB % 8 == 0;
    // 64 x 64 tiles: 64 x 128 tiles (each A row block staged once for the
    // full width of a layer) halve the workgroup count and were slower on
    // every training shape (step 436 -> 513 us at batch 65536)
    dim3 grid((g.N + kGBN - 1) / kGBN, (g.M + kGBM - 1) / kGBM, splitK);
    if (vec) hipLaunchKernelGGL((gemmKernel<true, kGBN>), grid, dim3(256), 0, st, g);
    else hipLaunchKernelGGL((gemmKernel<false, kGBN>), grid, dim3(256), 0, st, g);
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

// ------------------------------------------------------------------ loss head
// Output columns: 0, 1 exit deflection minus the spherical base (mrad), 2
// disk-hit logit, 3 hit radius / 10, 4 s = log of the expected squared exit
// error over sigma0^2.  Per row (mean over the M rows of the batch):
//   all rows:               BCE(disk logit, disk)
//   escaped, not band:      kRegW x Huber(|e|; kClip) with e in mrad
//                           (quadratic inside the clip radius, linear
//                           beyond), and the negative log-likelihood
//                           0.1 (s + q e^-s) of an exponential law for
//                           q = min(|e|, kClip)^2 + 1e-6 (|e| / sigma0 with
//                           sigma0 = 1 mrad, the output unit),
//                           held fixed: minimal at e^s = q, so s learns
//                           log E[q], the mean squared error it will make
//                           (e^-s clamped at |s| <= 20)
//   escaped, not band, disk: (hitR_pred - hitR / 10)^2
// dOut gets dL/dY for the head; dOutTrunk the same with the s column zero,
// so the log-variance head does not shape the hidden layers.
constexpr float kMlpClip = 20.f;      // mrad
constexpr float kMlpRegW = 1e-3f;   // (mrad^2: the balance against the disk BCE of a 1e3 weight on rad^2)

__global__ __launch_bounds__(256) void lossHeadKernel(const float* Y, const float* bias, const MlpTarget* T, int M,
                                                      __bf16* dOut, __bf16* dOutTrunk, float* lossOut, float* dbHead)
{
    // grid-stride over rows; per-thread sums of the three loss terms and of
    // the head-bias gradient db[c] = sum_i dOut[i][c] (the bf16 values the
    // backward GEMMs read), reduced per workgroup before one atomic each
    float part[3 + kMlpOut];
#pragma unroll
    for (int c = 0; c < 3 + kMlpOut; ++c) part[c] = 0.f;
    const float invM = 1.f / (float)M;
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < M; i += gridDim.x * blockDim.x)
    {
        float y[kMlpOutPad];
#pragma unroll
        for (int c = 0; c < kMlpOutPad; ++c) y[c] = c < kMlpOut ? Y[(size_t)i * kMlpOutPad + c] + bias[c] : 0.f;
        float g[kMlpOutPad];
#pragma unroll
        for (int c = 0; c < kMlpOutPad; ++c) g[c] = 0.f;
        float lReg = 0.f, lHit = 0.f;
        const MlpTarget t = T[i];
        const float p = 1.f / (1.f + __expf(-y[2]));
        const float lBce = fmaxf(y[2], 0.f) - y[2] * t.disk + __logf(1.f + __expf(-fabsf(y[2])));
        g[2] = (p - t.disk) * invM;
        if (t.escaped > 0.5f && t.band < 0.5f)
        {
            const float e0 = y[0] - t.d0 * kMlpOutPerRad, e1 = y[1] - t.d1 * kMlpOutPerRad;
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
            const float q = c2 + 1e-6f;
            const float qe = q * __expf(-fminf(fmaxf(y[4], -20.f), 20.f));
            lReg += 0.1f * (y[4] + qe);
            g[4] = 0.1f * (1.f - (fabsf(y[4]) < 20.f ? qe : 0.f)) * invM;
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
            const __bf16 gb = tobf(g[c]);
            dOut[(size_t)i * kMlpOutPad + c] = gb;
            dOutTrunk[(size_t)i * kMlpOutPad + c] = tobf(c == 4 ? 0.f : g[c]);
            if (c < kMlpOut) part[3 + c] += (float)gb;
        }
        part[0] += lReg;
        part[1] += lBce;
        part[2] += lHit;
    }
    __shared__ float red[4][3 + kMlpOut];
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
#pragma unroll
    for (int c = 0; c < 3 + kMlpOut; ++c)
    {
        float v = part[c];
        for (int off = 32; off > 0; off >>= 1) v += __shfl_down(v, off);
        if (lane == 0) red[wave][c] = v;
    }
    __syncthreads();
    if (threadIdx.x < 3 + kMlpOut)
    {
        const int c = threadIdx.x;
        const float v = red[0][c] + red[1][c] + red[2][c] + red[3][c];
        atomicAdd(c < 3 ? &lossOut[c] : &dbHead[c - 3], v);
    }
}

hipError_t launchLossHead(const float* Y, const float* bias, const MlpTarget* T, int M, __bf16* dOut,
                          __bf16* dOutTrunk, float* lossOut, float* dbHead, hipStream_t st)
{
    if (M <= 0) return hipSuccess;
    // at most 128 workgroups: 1024 atomics in all, on 8 addresses
    const int blocks = std::max(1, std::min((M + 255) / 256, 128));
    hipLaunchKernelGGL(lossHeadKernel, dim3(blocks), dim3(256), 0, st, Y, bias, T, M, dOut, dOutTrunk, lossOut,
                       dbHead);
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
