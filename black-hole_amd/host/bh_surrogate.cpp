// ---------------------------------------------------------------------------
// bh_surrogate.cpp -- the learned surrogate: training set generation by the
// plain FP32 tracer, training with the project's own HIP forward / backward
// / Adam kernels (bh_mlp.hip), inference, and parameter access for tests.
// No counterpart in the reference.
// ---------------------------------------------------------------------------
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <mutex>
#include <new>
#include <vector>

#include "../../include/bh/bhtrace.h"
#include "../csrc/bh_internal.h"

using bhk::kMlpHidden;
using bhk::kMlpIn;
using bhk::kMlpInputs;
using bhk::kMlpLayers;
using bhk::kMlpOut;
using bhk::kMlpOutPad;

namespace {

struct Rng
{
    uint64_t s;
    uint64_t next()
    {
        uint64_t z = (s += 0x9E3779B97F4A7C15ull);
        z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
        z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
        return z ^ (z >> 31);
    }
    double uniform() { return (double)(next() >> 11) * (1.0 / 9007199254740992.0); }
    double uniform(double lo, double hi) { return lo + (hi - lo) * uniform(); }
};

// Parameter layout: for each layer, W (out x in, row-major) then b (out).
struct Layout
{
    size_t w[kMlpLayers], b[kMlpLayers], total;
    Layout()
    {
        size_t o = 0;
        for (int l = 0; l < kMlpLayers; ++l)
        {
            w[l] = o;
            o += (size_t)bhk::mlpLayerOut(l) * bhk::mlpLayerIn(l);
            b[l] = o;
            o += (size_t)bhk::mlpLayerOut(l);
        }
        total = o;
    }
};
const Layout kL;
constexpr int kMaxSplit = 128;   // batch chunks of the weight-gradient GEMMs

template <class T> hipError_t dalloc(T** p, size_t n)
{
    return hipMalloc((void**)p, std::max<size_t>(n, 1) * sizeof(T));
}

}  // namespace

struct BhSurrogate
{
    float* p = nullptr;      // FP32 master parameters
    __bf16* pb = nullptr;    // BF16 copies (MFMA operands; biases unused)
    float* g = nullptr;      // gradients
    float* m = nullptr;      // Adam moments
    float* v = nullptr;
    int step = 0;
    // workspace for `cap` rows
    int cap = 0;
    float* x = nullptr;          // inputs copy (training batches are views into the set)
    __bf16* X0 = nullptr;        // [cap x 224]
    float* Z[4] = {};            // pre-activations [cap x 128]
    __bf16* A[4] = {};           // activations [cap x 128]
    float* Y5 = nullptr;         // [cap x 16]
    __bf16* dOut = nullptr;      // [cap x 16]
    __bf16* dOutT = nullptr;     // [cap x 16], log-variance column zero
    __bf16* dZ[2] = {};          // [cap x 128] each: gradients at the pre-activations (ping-pong)
    float* part = nullptr;       // split-K partials of the weight gradients [kMaxSplit x 128 x 224]
    float* dbPart = nullptr;     // column-sum partials of the bias gradients [gemmDbParts(cap) x 128]
    float* loss = nullptr;       // per-step loss terms [lossCap x 3]
    int lossCap = 0;
    // training set
    float* setX = nullptr;
    bhk::MlpTarget* setT = nullptr;
    float* setEsc = nullptr;
    int setN = 0;
    std::mutex mu;
};

namespace {

void freeWorkspace(BhSurrogate* s)
{
    (void)hipFree(s->X0);
    for (int l = 0; l < 4; ++l) { (void)hipFree(s->Z[l]); (void)hipFree(s->A[l]); s->Z[l] = nullptr; s->A[l] = nullptr; }
    (void)hipFree(s->Y5);
    (void)hipFree(s->dOut);
    (void)hipFree(s->dOutT);
    s->dOutT = nullptr;
    for (int k = 0; k < 2; ++k) { (void)hipFree(s->dZ[k]); s->dZ[k] = nullptr; }
    (void)hipFree(s->part);
    (void)hipFree(s->dbPart);
    s->part = nullptr; s->dbPart = nullptr;
    s->X0 = nullptr; s->Y5 = nullptr; s->dOut = nullptr;
    s->cap = 0;
}

hipError_t ensureWorkspace(BhSurrogate* s, int rows, hipStream_t st)
{
    if (rows <= s->cap) return hipSuccess;
    hipError_t e = hipStreamSynchronize(st);
    if (e != hipSuccess) return e;
    freeWorkspace(s);
    const size_t n = (size_t)rows;
    e = dalloc(&s->X0, n * kMlpIn);
    for (int l = 0; l < 4 && e == hipSuccess; ++l)
    {
        e = dalloc(&s->Z[l], n * kMlpHidden);
        if (e == hipSuccess) e = dalloc(&s->A[l], n * kMlpHidden);
    }
    if (e == hipSuccess) e = dalloc(&s->Y5, n * kMlpOutPad);
    if (e == hipSuccess) e = dalloc(&s->dOut, n * kMlpOutPad);
    if (e == hipSuccess) e = dalloc(&s->dOutT, n * kMlpOutPad);
    if (e == hipSuccess) e = dalloc(&s->dZ[0], n * kMlpHidden);
    if (e == hipSuccess) e = dalloc(&s->dZ[1], n * kMlpHidden);
    if (e == hipSuccess) e = dalloc(&s->part, (size_t)kMaxSplit * kMlpHidden * kMlpIn);
    if (e == hipSuccess) e = dalloc(&s->dbPart, (size_t)bhk::gemmDbParts(rows) * kMlpHidden);
    if (e != hipSuccess) { freeWorkspace(s); return e; }
    s->cap = rows;
    return hipSuccess;
}

bhk::GemmArgs gemm(const __bf16* A, int lda, int transA, const __bf16* B, int ldb, int transB, float* C, int ldc,
                   int M, int N, int K, int acc)
{
    bhk::GemmArgs g;
    g.A = A; g.lda = lda; g.transA = transA;
    g.B = B; g.ldb = ldb; g.transB = transB;
    g.C = C; g.ldc = ldc;
    g.M = M; g.N = N; g.K = K;
    g.accumulate = acc;
    g.epi = bhk::kEpiStore;
    g.act = 0;
    g.bias = nullptr; g.Zout = nullptr; g.Aout = nullptr; g.Zin = nullptr; g.dZout = nullptr; g.db = nullptr;
    g.part = nullptr; g.dbPart = nullptr;
    return g;
}

// Forward pass of `rows` rows of inputs x (device [rows x 6]).  Keeps the
// pre-activations and activations in the workspace; Y5 = head before bias.
// Hidden layers: GEMM with the bias + SiLU epilogue fused.
hipError_t forward(BhSurrogate* s, const float* x, int rows, hipStream_t st)
{
    hipError_t e = bhk::launchFeatures(x, rows, s->X0, st);
    const __bf16* X = s->X0;
    for (int l = 0; l < kMlpLayers && e == hipSuccess; ++l)
    {
        const int in = bhk::mlpLayerIn(l), out = bhk::mlpLayerOut(l);
        bhk::GemmArgs g = gemm(X, in, 0, s->pb + kL.w[l], in, 1, (l < 4) ? nullptr : s->Y5, out, rows, out, in, 0);
        if (l < 4)
        {
            g.epi = bhk::kEpiBiasAct;
            g.act = 1;
            g.bias = s->p + kL.b[l];
            g.Zout = s->Z[l];
            g.Aout = s->A[l];
        }
        e = bhk::launchGemm(g, 1, st);
        if (l < 4) X = s->A[l];
    }
    return e;
}

// dZ_{l-1} = (dZ_l . W_l) * silu'(Z_{l-1}) in one GEMM; its epilogue stores
// the column sums of db_{l-1} per 32-row slab in dbPart (reduced after)
bhk::GemmArgs backData(const __bf16* dY, int ldy, const __bf16* W, int in, int outN, const float* Zprev,
                       __bf16* dZout, float* dbPart, int rows)
{
    bhk::GemmArgs g = gemm(dY, ldy, 0, W, in, 0, nullptr, in, rows, in, outN, 0);
    g.epi = bhk::kEpiSiluBack;
    g.Zin = Zprev;
    g.dZout = dZout;
    g.dbPart = dbPart;
    return g;
}

// dW = dY^T X over the batch (K = rows), split over the batch into partial
// products that a reduction adds into the gradient (no atomics; the
// summation order is fixed)
hipError_t weightGrad(BhSurrogate* s, const __bf16* dY, int ldy, int outM, const __bf16* X, int in, float* gW,
                      int rows, hipStream_t st)
{
    const int split = std::max(1, std::min(kMaxSplit, rows / 512));
    bhk::GemmArgs g = gemm(dY, ldy, 1, X, in, 0, gW, in, outM, in, rows, 0);
    g.part = s->part;
    hipError_t e = bhk::launchGemm(g, split, st);
    if (e == hipSuccess) e = bhk::launchReduceParts(s->part, split, outM * in, gW, nullptr, st);
    return e;
}

hipError_t dataGrad(BhSurrogate* s, const __bf16* dY, int ldy, const __bf16* W, int in, int outN, const float* Zprev,
                    __bf16* dZout, float* gb, int rows, hipStream_t st)
{
    hipError_t e = bhk::launchGemm(backData(dY, ldy, W, in, outN, Zprev, dZout, s->dbPart, rows), 1, st);
    // the weight-gradient partials are consumed by now: their buffer holds
    // the first pass of this reduction
    if (e == hipSuccess) e = bhk::launchReduceParts(s->dbPart, bhk::gemmDbParts(rows), in, gb, s->part, st);
    return e;
}

// Backward pass after forward(): gradients accumulate into s->g (zeroed by
// the caller); loss terms into lossSlot[0..2].
hipError_t backward(BhSurrogate* s, const bhk::MlpTarget* T, int rows, float* lossSlot, hipStream_t st)
{
    hipError_t e =
        bhk::launchLossHead(s->Y5, s->p + kL.b[4], T, rows, s->dOut, s->dOutT, lossSlot, s->g + kL.b[4], st);
    // head: dW4 = dOut^T A3; dZ3 = (dOut W4) * silu'(Z3), db3
    if (e == hipSuccess) e = weightGrad(s, s->dOut, kMlpOutPad, kMlpOutPad, s->A[3], kMlpHidden, s->g + kL.w[4], rows, st);
    if (e == hipSuccess)
        e = dataGrad(s, s->dOutT, kMlpOutPad, s->pb + kL.w[4], kMlpHidden, kMlpOutPad, s->Z[3], s->dZ[1],
                     s->g + kL.b[3], rows, st);
    for (int l = 3, cur = 1; l >= 0 && e == hipSuccess; --l, cur ^= 1)
    {
        const int in = bhk::mlpLayerIn(l);
        const __bf16* X = (l == 0) ? s->X0 : s->A[l - 1];
        e = weightGrad(s, s->dZ[cur], kMlpHidden, kMlpHidden, X, in, s->g + kL.w[l], rows, st);
        if (e == hipSuccess && l > 0)
            e = dataGrad(s, s->dZ[cur], kMlpHidden, s->pb + kL.w[l], in, kMlpHidden, s->Z[l - 1], s->dZ[cur ^ 1],
                         s->g + kL.b[l - 1], rows, st);
    }
    return e;
}

hipError_t refreshBf16(BhSurrogate* s, hipStream_t st)
{
    return bhk::launchToBf16(s->p, s->pb, (int)kL.total, st);
}

}  // namespace

namespace bhk {
SurrogateNet surrogateNet(const BhSurrogate* s)
{
    SurrogateNet n;
    n.W = s->pb;
    n.P = s->p;
    for (int l = 0; l < kMlpLayers; ++l)
    {
        n.wOff[l] = (int)kL.w[l];
        n.bOff[l] = (int)kL.b[l];
    }
    return n;
}
}  // namespace bhk

extern "C" {

int bhSurrogateParamCount(void) { return (int)kL.total; }

bh_error_t bhSurrogateCreate(uint64_t seed, BhSurrogate** out)
{
    if (!out) return hipErrorInvalidValue;
    *out = nullptr;
    BhSurrogate* s = new (std::nothrow) BhSurrogate;
    if (!s) return hipErrorOutOfMemory;
    const size_t n = kL.total;
    hipError_t e = dalloc(&s->p, n);
    if (e == hipSuccess) e = dalloc(&s->pb, n);
    if (e == hipSuccess) e = dalloc(&s->g, n);
    if (e == hipSuccess) e = dalloc(&s->m, n);
    if (e == hipSuccess) e = dalloc(&s->v, n);
    if (e != hipSuccess) { bhSurrogateDestroy(s); return e; }
    // He-uniform hidden layers, small output layer; padded output rows and
    // all biases zero
    std::vector<float> h(n, 0.f);
    Rng r{seed};
    for (int l = 0; l < kMlpLayers; ++l)
    {
        const int in = bhk::mlpLayerIn(l), outN = bhk::mlpLayerOut(l);
        const int realOut = (l == kMlpLayers - 1) ? kMlpOut : outN;
        // output layer starts small: the exit heads predict the residual
        // over the analytic base in mrad (typically a few)
        const double lim = (l == kMlpLayers - 1) ? 0.05 * std::sqrt(3.0 / in) : std::sqrt(6.0 / in);
        for (int o = 0; o < realOut; ++o)
            for (int i = 0; i < in; ++i) h[kL.w[l] + (size_t)o * in + i] = (float)r.uniform(-lim, lim);
    }
    bh_error_t be = bhSurrogateSetParams(s, h.data(), (int)n);
    if (be != 0) { bhSurrogateDestroy(s); return be; }
    *out = s;
    return hipSuccess;
}

void bhSurrogateDestroy(BhSurrogate* s)
{
    if (!s) return;
    (void)hipDeviceSynchronize();
    freeWorkspace(s);
    (void)hipFree(s->p); (void)hipFree(s->pb); (void)hipFree(s->g); (void)hipFree(s->m); (void)hipFree(s->v);
    (void)hipFree(s->loss);
    (void)hipFree(s->setX); (void)hipFree(s->setT); (void)hipFree(s->setEsc);
    delete s;
}

bh_error_t bhSurrogateSetParams(BhSurrogate* s, const float* host, int count)
{
    if (!s || !host || count != (int)kL.total) return hipErrorInvalidValue;
    std::lock_guard<std::mutex> lk(s->mu);
    hipError_t e = hipMemcpy(s->p, host, kL.total * sizeof(float), hipMemcpyHostToDevice);
    if (e == hipSuccess) e = hipMemset(s->m, 0, kL.total * sizeof(float));
    if (e == hipSuccess) e = hipMemset(s->v, 0, kL.total * sizeof(float));
    if (e == hipSuccess) e = refreshBf16(s, nullptr);
    if (e == hipSuccess) e = hipDeviceSynchronize();
    s->step = 0;
    return e;
}

bh_error_t bhSurrogateGetParams(BhSurrogate* s, float* host, int count)
{
    if (!s || !host || count != (int)kL.total) return hipErrorInvalidValue;
    hipError_t e = hipDeviceSynchronize();
    if (e == hipSuccess) e = hipMemcpy(host, s->p, kL.total * sizeof(float), hipMemcpyDeviceToHost);
    return e;
}

namespace {

// Training set of raysPerCfg rays from the camera of each configuration.
hipError_t buildDataset(BhSurrogate* s, const std::vector<RenderParams>& P, int raysPerCfg, uint64_t seed,
                        hipStream_t st)
{
    const int nCfg = (int)P.size();
    std::vector<bhk::FrameK> K((size_t)nCfg);
    std::vector<bhk::BandK> Bk((size_t)nCfg);
    BhPrecision prec;
    bhGetPrecision(&prec);
    for (int c = 0; c < nCfg; ++c)
    {
        K[(size_t)c] = bhk::makeFrame(P[(size_t)c]);
        Bk[(size_t)c] = bhk::makeBandK(P[(size_t)c], prec.critBand, prec.polarBand);
    }
    const size_t n = (size_t)nCfg * (size_t)raysPerCfg;
    std::lock_guard<std::mutex> lk(s->mu);
    hipError_t e = hipStreamSynchronize(st);
    if (e != hipSuccess) return e;
    (void)hipFree(s->setX); (void)hipFree(s->setT); (void)hipFree(s->setEsc);
    s->setX = nullptr; s->setT = nullptr; s->setEsc = nullptr; s->setN = 0;
    e = dalloc(&s->setX, n * kMlpInputs);
    if (e == hipSuccess) e = dalloc(&s->setT, n);
    if (e == hipSuccess) e = dalloc(&s->setEsc, n * 3);
    bhk::FrameK* dK = nullptr;
    RenderParams* dP = nullptr;
    bhk::BandK* dB = nullptr;
    if (e == hipSuccess) e = dalloc(&dB, Bk.size());
    if (e == hipSuccess) e = hipMemcpy(dB, Bk.data(), Bk.size() * sizeof(bhk::BandK), hipMemcpyHostToDevice);
    if (e == hipSuccess) e = dalloc(&dK, K.size());
    if (e == hipSuccess) e = dalloc(&dP, P.size());
    if (e == hipSuccess) e = hipMemcpy(dK, K.data(), K.size() * sizeof(bhk::FrameK), hipMemcpyHostToDevice);
    if (e == hipSuccess) e = hipMemcpy(dP, P.data(), P.size() * sizeof(RenderParams), hipMemcpyHostToDevice);
    if (e == hipSuccess)
    {
        bhk::DatasetArgs A;
        A.frames = dK;
        A.params = dP;
        A.nCfg = nCfg;
        A.raysPerCfg = raysPerCfg;
        A.seed = seed * 0x2545F4914F6CDD1Dull + 1ull;
        A.psiMax = (float)(65.0 * 3.14159265358979323846 / 180.0);
        A.bands = dB;
        A.inputs = s->setX;
        A.targets = s->setT;
        A.escDir = s->setEsc;
        e = bhk::launchSurrogateDataset(A, st);
    }
    if (e == hipSuccess) e = hipStreamSynchronize(st);
    (void)hipFree(dK);
    (void)hipFree(dP);
    (void)hipFree(dB);
    if (e == hipSuccess) s->setN = (int)n;
    return e;
}

}  // namespace

bh_error_t bhSurrogateDataset(BhSurrogate* s, int nCfg, int raysPerCfg, uint64_t seed, bh_stream_t stream)
{
    if (!s || nCfg < 1 || raysPerCfg < 64 || raysPerCfg % 64) return hipErrorInvalidValue;
    std::vector<RenderParams> P((size_t)nCfg);
    bhSurrogateConfigs(seed, nCfg, P.data());
    return buildDataset(s, P, raysPerCfg, seed, (hipStream_t)stream);
}

bh_error_t bhSurrogateDatasetFor(BhSurrogate* s, const RenderParams* views, int nViews, int raysPerView,
                                 uint64_t seed, bh_stream_t stream)
{
    if (!s || !views || nViews < 1 || raysPerView < 64 || raysPerView % 64) return hipErrorInvalidValue;
    std::vector<RenderParams> P(views, views + nViews);
    for (auto& p : P) bhSanitizeRenderParams(&p);
    return buildDataset(s, P, raysPerView, seed, (hipStream_t)stream);
}

bh_error_t bhSurrogateDatasetView(BhSurrogate* s, float** inputs, void** targets, float** escDir, int* n)
{
    if (!s) return hipErrorInvalidValue;
    if (inputs) *inputs = s->setX;
    if (targets) *targets = s->setT;
    if (escDir) *escDir = s->setEsc;
    if (n) *n = s->setN;
    return hipSuccess;
}

bh_error_t bhSurrogateTrain(BhSurrogate* s, int steps, int batch, float lr, float* lossHost, bh_stream_t stream)
{
    if (!s || steps < 0 || batch < 64 || batch % 64 || !(lr > 0.f)) return hipErrorInvalidValue;
    if (s->setN < batch) return hipErrorNotReady;
    hipStream_t st = (hipStream_t)stream;
    std::lock_guard<std::mutex> lk(s->mu);
    hipError_t e = ensureWorkspace(s, batch, st);
    if (e != hipSuccess) return e;
    if (steps > s->lossCap)
    {
        (void)hipFree(s->loss);
        s->loss = nullptr;
        s->lossCap = 0;
        e = dalloc(&s->loss, (size_t)steps * 3);
        if (e != hipSuccess) return e;
        s->lossCap = steps;
    }
    e = bhk::launchZero(s->loss, (size_t)steps * 3, st);
    const int nb = s->setN / batch;
    for (int k = 0; k < steps && e == hipSuccess; ++k)
    {
        const int b0 = (s->step % nb) * batch;
        // cosine learning-rate decay over this call's steps, to lr / 20
        const float c = 0.5f * (1.f + std::cos(3.14159265f * (float)k / (float)std::max(steps, 1)));
        const float lrk = lr * (0.05f + 0.95f * c);
        e = bhk::launchZero(s->g, kL.total, st);
        if (e == hipSuccess) e = forward(s, s->setX + (size_t)b0 * kMlpInputs, batch, st);
        if (e == hipSuccess) e = backward(s, s->setT + b0, batch, s->loss + (size_t)k * 3, st);
        ++s->step;
        if (e == hipSuccess) e = bhk::launchAdam(s->p, s->g, s->m, s->v, s->pb, (int)kL.total, lrk, s->step, st);
    }
    if (e == hipSuccess && lossHost)
    {
        std::vector<float> tmp((size_t)steps * 3);
        e = hipMemcpyAsync(tmp.data(), s->loss, tmp.size() * sizeof(float), hipMemcpyDeviceToHost, st);
        if (e == hipSuccess) e = hipStreamSynchronize(st);
        for (int k = 0; k < steps; ++k)
            for (int j = 0; j < 3; ++j) lossHost[k * 3 + j] = tmp[(size_t)k * 3 + j] / (float)batch;
    }
    return e;
}

bh_error_t bhSurrogateForward(BhSurrogate* s, const float* devInputs, int n, float* devOut, bh_stream_t stream)
{
    if (!s || n < 0 || (n > 0 && (!devInputs || !devOut))) return hipErrorInvalidValue;
    hipStream_t st = (hipStream_t)stream;
    std::lock_guard<std::mutex> lk(s->mu);
    const int chunk = 1 << 18;
    hipError_t e = ensureWorkspace(s, std::min(n, chunk), st);
    for (int i0 = 0; i0 < n && e == hipSuccess; i0 += chunk)
    {
        const int rows = std::min(chunk, n - i0);
        e = forward(s, devInputs + (size_t)i0 * kMlpInputs, rows, st);
        if (e == hipSuccess) e = bhk::launchHeadOut(s->Y5, s->p + kL.b[4], rows, devOut + (size_t)i0 * kMlpOut, st);
    }
    return e;
}

bh_error_t bhSurrogateForwardFused(BhSurrogate* s, const float* devInputs, int n, float* devOut, bh_stream_t stream)
{
    if (!s || n < 0 || (n > 0 && (!devInputs || !devOut))) return hipErrorInvalidValue;
    if (n == 0) return hipSuccess;
    hipStream_t st = (hipStream_t)stream;
    std::lock_guard<std::mutex> lk(s->mu);
    int dev = 0, cus = 0;
    hipError_t e = hipGetDevice(&dev);
    if (e == hipSuccess) e = hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev);
    if (e != hipSuccess) return e;
    const bhk::SurrogateNet net = bhk::surrogateNet(s);
    bhk::RouteArgs R{};
    R.inputs = devInputs;
    R.count = nullptr;
    R.rows = n;
    R.outRows = devOut;
    R.W = net.W;
    R.Pf = net.P;
    for (int l = 0; l < kMlpLayers; ++l) { R.wOff[l] = net.wOff[l]; R.bOff[l] = net.bOff[l]; }
    return bhk::launchSurrogateRoute(R, n, cus, st);
}

bh_error_t bhSurrogateLossGrad(BhSurrogate* s, const float* devInputs, const void* devTargets, int n,
                               float* lossHost, float* gradHost, bh_stream_t stream)
{
    if (!s || n < 1 || !devInputs || !devTargets) return hipErrorInvalidValue;
    hipStream_t st = (hipStream_t)stream;
    std::lock_guard<std::mutex> lk(s->mu);
    hipError_t e = ensureWorkspace(s, n, st);
    float* dl = nullptr;
    if (e == hipSuccess) e = dalloc(&dl, 3);
    if (e == hipSuccess) e = bhk::launchZero(dl, 3, st);
    if (e == hipSuccess) e = bhk::launchZero(s->g, kL.total, st);
    if (e == hipSuccess) e = forward(s, devInputs, n, st);
    if (e == hipSuccess) e = backward(s, static_cast<const bhk::MlpTarget*>(devTargets), n, dl, st);
    if (e == hipSuccess) e = hipStreamSynchronize(st);
    if (e == hipSuccess && lossHost) e = hipMemcpy(lossHost, dl, 3 * sizeof(float), hipMemcpyDeviceToHost);
    if (e == hipSuccess && gradHost) e = hipMemcpy(gradHost, s->g, kL.total * sizeof(float), hipMemcpyDeviceToHost);
    (void)hipFree(dl);
    return e;
}

bh_error_t bhMlpGemmTest(const void* A, int lda, int transA, const void* B, int ldb, int transB, float* C, int ldc,
                         int M, int N, int K, int splitK, bh_stream_t stream)
{
    if (!A || !B || !C) return hipErrorInvalidValue;
    bhk::GemmArgs g = gemm(static_cast<const __bf16*>(A), lda, transA, static_cast<const __bf16*>(B), ldb, transB, C,
                           ldc, M, N, K, splitK > 1 ? 1 : 0);
    hipStream_t st = (hipStream_t)stream;
    if (splitK <= 1 || splitK > kMaxSplit || ldc != N) return bhk::launchGemm(g, splitK, st);
    // split K as the weight gradients run it: partial products stored per
    // split, then reduced into C (C += sum)
    float* part = nullptr;
    hipError_t e = dalloc(&part, (size_t)splitK * M * N);
    if (e != hipSuccess) return e;
    g.accumulate = 0;
    g.part = part;
    e = bhk::launchGemm(g, splitK, st);
    if (e == hipSuccess) e = bhk::launchReduceParts(part, splitK, M * N, C, nullptr, st);
    const hipError_t es = hipStreamSynchronize(st);
    (void)hipFree(part);
    return e != hipSuccess ? e : es;
}

}  // extern "C"
