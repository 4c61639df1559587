// ---------------------------------------------------------------------------
// bh_mlp.h -- shapes and launch descriptors of the surrogate network
// (kernels in bh_mlp.hip, training and routing in host/bh_surrogate.cpp).
// ---------------------------------------------------------------------------
#pragma once
#include <hip/hip_runtime.h>
#include <cstddef>

namespace bhk {

#define BHK_MLP_INL __device__ __forceinline__

constexpr int kMlpInputs = 6;                         // b, alpha, a*, q, r_cam, inclination (normalised)
constexpr int kMlpFreqs = 16;                         // Fourier frequencies per input
constexpr int kMlpFF = kMlpInputs * kMlpFreqs * 2;    // 192 sin/cos features
constexpr int kMlpIn = 224;                           // 192 + 6 raw inputs, zero-padded to 7 x 32
constexpr int kMlpHidden = 128;                       // width of the 4 hidden layers
constexpr int kMlpLayers = 5;                         // 4 hidden + output
constexpr int kMlpOut = 5;                            // deflection x2, disk logit, hit radius, log-variance
constexpr int kMlpOutPad = 16;
constexpr float kMlpSigma0 = 1e-3f;                   // scale of the predicted exit-direction spread (rad)
// The exit outputs (columns 0, 1) are in milliradians: targets of a few
// mrad are then O(1) outputs, whose last-layer weights Adam moves by steps
// of the learning rate without swamping them (in radians the late steps of
// a 1e-4 learning rate alone moved the outputs by ~1e-3 rad).
constexpr float kMlpOutPerRad = 1000.f;

// Frequencies 1.25^f * pi / 2, f = 0..15 (pi/2 .. 44.6 per unit input).
__host__ __device__ inline float mlpFreq(int f)
{
    float w = 1.5707963267948966f;
    for (int i = 0; i < f; ++i) w *= 1.25f;
    return w;
}

// Layer l: out x in weights (row-major), in = kMlpIn for l = 0 else kMlpHidden;
// out = kMlpHidden for l < 4 else kMlpOutPad.
__host__ __device__ inline int mlpLayerIn(int l) { return l == 0 ? kMlpIn : kMlpHidden; }
__host__ __device__ inline int mlpLayerOut(int l) { return l == kMlpLayers - 1 ? kMlpOutPad : kMlpHidden; }

struct MlpTarget
{
    float d0, d1;       // exit deflection components (radians) minus the analytic base
    float disk;         // 1 when the plain trace sampled the disk
    float hitR;         // cylindrical radius of the first disk chord (M units of the params)
    float escaped;      // 1 escaped, 0 captured
    float band;         // 1: critical / polar band ray (no regression target)
    float pad[2];
};

// Epilogue of a GEMM launch.
enum GemmEpi
{
    kEpiStore = 0,      // C (+)= acc (f32)
    kEpiBiasAct = 1,    // z = acc + bias[col]; Zout = z (f32, when set); Aout = bf16(act ? silu(z) : z)
    kEpiSiluBack = 2,   // g = acc * silu'(Zin); dZout = bf16(g); db[col] += sum over rows of g
};

struct GemmArgs
{
    const __bf16* A;
    const __bf16* B;
    float* C;
    int M, N, K;
    int lda, ldb, ldc;
    int transA, transB;
    int accumulate;
    int epi;                // GemmEpi; the fused epilogues need splitK == 1 and use ldc for every row stride
    int act;                // kEpiBiasAct: 1 SiLU, 0 identity
    const float* bias;      // kEpiBiasAct: [N]
    float* Zout;            // kEpiBiasAct: [M x N] pre-activations (or null)
    __bf16* Aout;           // kEpiBiasAct: [M x N] activations
    const float* Zin;       // kEpiSiluBack: [M x N] pre-activations of the layer
    __bf16* dZout;          // kEpiSiluBack: [M x N]
    float* db;              // kEpiSiluBack: [N] bias gradient (accumulated; unused when dbPart is set)
    float* part;            // kEpiStore: [splitK x M x N] per-split partial products, stored (not added)
                            // instead of atomics into C; reduce them with launchReduceParts (or null)
    float* dbPart;          // kEpiSiluBack: [2 * ceil(M / 64) x N] column sums of 32-row slabs, stored
                            // instead of atomics into db; reduce with launchReduceParts (or null)
};

hipError_t launchGemm(const GemmArgs& g, int splitK, hipStream_t st);
// dst[i] += sum over p < P of part[p * L + i], i < L (fixed summation order).
// tmp (or null): [ceil(P / 64) x L] floats, used when P > 256 for a first
// pass over segments of 64 partials.
hipError_t launchReduceParts(const float* part, int P, int L, float* dst, float* tmp, hipStream_t st);
inline int gemmDbParts(int M) { return 2 * ((M + 63) / 64); }
hipError_t launchFeatures(const float* x, int M, __bf16* X0, hipStream_t st);
hipError_t launchLossHead(const float* Y, const float* bias, const MlpTarget* T, int M, __bf16* dOut,
                          __bf16* dOutTrunk, float* lossOut, float* dbHead, hipStream_t st);
hipError_t launchHeadOut(const float* Y, const float* bias, int M, float* out, hipStream_t st);
hipError_t launchAdam(float* p, const float* g, float* m, float* v, __bf16* pb, int n, float lr, int t,
                      hipStream_t st);
hipError_t launchToBf16(const float* x, __bf16* y, int n, hipStream_t st);
hipError_t launchZero(float* x, size_t n, hipStream_t st);

}  // namespace bhk
