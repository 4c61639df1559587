// ---------------------------------------------------------------------------
// oracle/bh_oracle.cpp -- C API of the CPU oracle (liboracle.so).
//
// TEST INFRASTRUCTURE ONLY: the checker for the HIP product path.  Loaded by
// tests/, __graft_entry__.smoke() and the CPU-baseline leg of bench.py.
//
// Frames are parallelized with OpenMP over 16x8 pixel tiles (dynamic
// schedule), as the host baseline of the reference's CPU validation runs.
// ---------------------------------------------------------------------------
#include "bh_oracle.hpp"
#include <algorithm>
#include <cstring>
#include <vector>
#include <omp.h>

namespace bho {
thread_local uint64_t tl_flops = 0;
}
using namespace bho;

// ------------------------------------------------------------ host derive / sanitize
// metric.cuh:74-137 (float, host libm)
static float hOuter(float M, float a, float Q)
{
    float d = M * M - a * a - Q * Q;
    if (d < 0.f) d = 0.f;
    return M + sqrtf(d);
}
static float hInner(float M, float a, float Q)
{
    float d = M * M - a * a - Q * Q;
    if (d < 0.f) d = 0.f;
    return M - sqrtf(d);
}
static float hPhotonSph(float M, float Q)
{
    float d = 9.f * M * M - 8.f * Q * Q;
    if (d < 0.f) d = 0.f;
    return 0.5f * (3.f * M + sqrtf(d));
}
static float hPhotonKerr(float M, float aStar)
{
    float x = aStar;
    if (x > 1.f) x = 1.f;
    if (x < -1.f) x = -1.f;
    return 2.f * M * (1.f + cosf((2.f / 3.f) * acosf(-x)));
}
static float hPhotonKN(float M, float aStar, float q)
{
    float rK = hPhotonKerr(M, aStar);
    float rS = hPhotonSph(M, q * M);
    if (fabsf(aStar) < 1e-5f) return rS;
    float w = q * q;
    if (w < 0.f) w = 0.f;
    if (w > 1.f) w = 1.f;
    return (1.f - 0.35f * w) * rK + 0.35f * w * rS;
}
static float hErgo(float M, float Q)
{
    float d = M * M - Q * Q;
    if (d < 0.f) d = 0.f;
    return M + sqrtf(d);
}

struct OBHDerived { float a, Q, rPlus, rMinus, rPhoton, rErgo, rIsco; int clamped, naked; };

// metric.cuh:224-248
static float hIscoNumeric(float M, float a, float Q)
{
    float disc = M * M - a * a - Q * Q;
    float rp = (disc >= 0.f) ? hOuter(M, a, Q) : (SING_CUT_OVER_M * M);
    float best = 6.f * M, bestE = 1e30f;
    bool found = false;
    const float r0 = (rp > 1e-6f ? rp * 1.02f : 0.5f * M);
    const float r1 = 12.f * M;
    const float step = 0.001f * M;
    for (float r = r0; r <= r1; r += step)
    {
        float Om = bhOmegaCircular<float>(r, M, a, Q);
        if (Om <= 0.f) continue;
        float gtt, gtp, gpp;
        bhEquatorialMetric<float>(r, M, a, Q, gtt, gtp, gpp);
        float den = -(gtt + 2.f * Om * gtp + Om * Om * gpp);
        if (den <= 1e-6f) continue;
        float E = -(gtt + Om * gtp) / sqrtf(den);
        if (E < bestE) { bestE = E; best = r; found = true; }
    }
    if (!found) best = (6.f * M > rp * 1.5f) ? 6.f * M : rp * 1.5f;
    float lo = (rp > 1e-6f) ? rp * 1.15f : 0.5f * M;
    return best > lo ? best : lo;
}

// metric.cuh:257-329
static OBHDerived hValidateDerive(int model, float& M, float& aStar, float& q, bool allowNaked)
{
    OBHDerived d = {0.f, 0.f, 1.f, 0.f, 1.5f, 1.f, 3.f, 0, 0};
    if (!(M > 0.f) || !std::isfinite(M)) { M = 0.5f; d.clamped = 1; }
    if (!std::isfinite(aStar)) { aStar = 0.f; d.clamped = 1; }
    if (!std::isfinite(q)) { q = 0.f; d.clamped = 1; }
    const float knobMax = allowNaked ? sqrtf(NAKED_LIMIT) : EXTREMAL_LIMIT;
    if (aStar < -knobMax) { aStar = -knobMax; d.clamped = 1; }
    if (aStar > knobMax) { aStar = knobMax; d.clamped = 1; }
    if (q < 0.f) { q = 0.f; d.clamped = 1; }
    if (q > knobMax) { q = knobMax; d.clamped = 1; }
    float aEff = (model == BH_KERR || model == BH_KERR_NEWMAN) ? aStar : 0.f;
    float qEff = (model == BH_REISSNER_NORDSTROM || model == BH_KERR_NEWMAN) ? q : 0.f;
    const float bound = allowNaked ? NAKED_LIMIT : EXTREMAL_LIMIT;
    float s2 = aEff * aEff + qEff * qEff;
    if (s2 > bound)
    {
        float scale = sqrtf(bound / s2);
        aEff *= scale;
        qEff *= scale;
        if (model == BH_KERR || model == BH_KERR_NEWMAN) aStar = aEff;
        if (model == BH_REISSNER_NORDSTROM || model == BH_KERR_NEWMAN) q = qEff;
        d.clamped = 1;
    }
    d.a = aEff * M;
    d.Q = qEff * M;
    float disc = M * M - d.a * d.a - d.Q * d.Q;
    if (disc < 0.f) { d.naked = 1; d.rPlus = 0.f; d.rMinus = 0.f; }
    else { d.rPlus = hOuter(M, d.a, d.Q); d.rMinus = hInner(M, d.a, d.Q); }
    d.rErgo = hErgo(M, d.Q);
    if (d.naked) d.rErgo = 0.f;
    if (model == BH_KERR || model == BH_KERR_NEWMAN)
    {
        d.rPhoton = (model == BH_KERR_NEWMAN) ? hPhotonKN(M, aEff, qEff) : hPhotonKerr(M, aEff);
        if (d.rPlus > 0.f)
        {
            float lo = d.rPlus * 1.02f;
            if (d.rPhoton < lo) d.rPhoton = lo;
        }
        else if (d.rPhoton < SING_CUT_OVER_M * M * 2.f) d.rPhoton = 3.f * M;
    }
    else
    {
        d.rPhoton = hPhotonSph(M, d.Q);
        if (d.naked && !(d.rPhoton > 0.f)) d.rPhoton = 3.f * M;
    }
    d.rIsco = hIscoNumeric(M, d.a, d.Q);
    return d;
}

static bool finite3(const bh_float3& v) { return std::isfinite(v.x) && std::isfinite(v.y) && std::isfinite(v.z); }

// metric.cuh:344-419
static bool hSanitize(RenderParams& P)
{
    bool fixed = false;
    auto fixf = [&](float& v, float lo, float hi, float def) {
        if (!std::isfinite(v)) { v = def; fixed = true; }
        if (v < lo) { v = lo; fixed = true; }
        if (v > hi) { v = hi; fixed = true; }
    };
    if (P.width < 1) { P.width = 1; fixed = true; }
    if (P.height < 1) { P.height = 1; fixed = true; }
    fixf(P.dPhi, 1e-4f, 0.1f, 0.012f);
    if (P.maxSteps < 16) { P.maxSteps = 16; fixed = true; }
    if (P.maxSteps > 200000) { P.maxSteps = 200000; fixed = true; }
    fixf(P.exposure, 0.01f, 100.f, 1.f);
    fixf(P.M, 1e-3f, 100.f, 0.5f);
    fixf(P.tanHalfFov, 0.05f, 5.f, 0.5773503f);
    fixf(P.aspect, 0.05f, 20.f, 16.f / 9.f);
    if (P.model < 0 || P.model > 3) { P.model = 0; fixed = true; }
    float aStar = P.aSpin / P.M, q = P.Qc / P.M, M = P.M;
    bool allowNaked = (P.allowNaked != 0);
    float bound = allowNaked ? NAKED_LIMIT : EXTREMAL_LIMIT;
    bool needFix = !std::isfinite(aStar) || !std::isfinite(q) || aStar * aStar + q * q > bound + 1e-4f ||
                   q < 0.f || !std::isfinite(P.rPlus) || (!allowNaked && P.rPlus <= 0.f) ||
                   (allowNaked && P.rPlus < 0.f);
    if (needFix)
    {
        OBHDerived d = hValidateDerive(P.model, M, aStar, q, allowNaked);
        P.M = M; P.aSpin = d.a; P.Qc = d.Q;
        P.rPlus = d.rPlus; P.rPhoton = d.rPhoton; P.rErgo = d.rErgo;
        fixed = true;
    }
    if (!finite3(P.camPos) || !finite3(P.camForward) || !finite3(P.camRight) || !finite3(P.camUp))
    {
        P.camPos = {13.f, 3.2f, 8.f};
        P.camForward = {-0.81f, -0.20f, -0.55f};
        P.camRight = {0.56f, 0.f, -0.83f};
        P.camUp = {-0.17f, 0.98f, -0.11f};
        fixed = true;
    }
    float r0 = sqrtf(P.camPos.x * P.camPos.x + P.camPos.y * P.camPos.y + P.camPos.z * P.camPos.z);
    float rMin = (P.rPlus > 1e-6f) ? P.rPlus * 2.0f : (SING_CUT_OVER_M * P.M * 8.f);
    if (r0 < rMin)
    {
        float s = rMin / (r0 > 1e-6f ? r0 : 1e-6f);
        P.camPos.x *= s; P.camPos.y *= s; P.camPos.z *= s;
        fixed = true;
    }
    float diskLo = (P.rPlus > 1e-6f) ? P.rPlus * 1.05f : (SING_CUT_OVER_M * P.M * 2.f);
    fixf(P.diskInner, diskLo, 1000.f, diskLo * 3.f);
    fixf(P.diskOuter, P.diskInner * 1.2f, 2000.f, P.diskInner * 4.f);
    if (P.sampleIndex < 0) { P.sampleIndex = 0; fixed = true; }
    if (P.sampleIndex > 65536) { P.sampleIndex = 65536; fixed = true; }
    if (P.accumMode < 0 || P.accumMode > 2) { P.accumMode = 0; fixed = true; }
    fixf(P.bloomStrength, 0.f, 4.f, 0.35f);
    fixf(P.diskEmisScale, 0.f, 50.f, 1.5f);
    fixf(P.diskAbsScale, 0.f, 20.f, 2.2f);
    fixf(P.hotSpotStrength, 0.f, 8.f, 0.65f);
    if (P.hotSpotsEnabled != 0 && P.hotSpotsEnabled != 1) { P.hotSpotsEnabled = 1; fixed = true; }
    return fixed;
}

// ------------------------------------------------------------ frame helpers
template <class F> static void forTiles(int W, int rowStart, int rowStep, int nRows, int nthreads, F&& fn)
{
    // tiles of 16 px x 8 selected rows
    const int tw = 16, th = 8;
    int ntx = (W + tw - 1) / tw, nty = (nRows + th - 1) / th;
    int nt = ntx * nty;
    if (nthreads <= 0) nthreads = omp_get_max_threads();
#pragma omp parallel for schedule(dynamic, 1) num_threads(nthreads)
    for (int t = 0; t < nt; ++t)
    {
        int tx = t % ntx, ty = t / ntx;
        for (int j = ty * th; j < std::min(nRows, ty * th + th); ++j)
        {
            int py = rowStart + j * rowStep;
            for (int px = tx * tw; px < std::min(W, tx * tw + tw); ++px) fn(px, py, j);
        }
    }
}

static int nRowsOf(int H, int rowStart, int rowStep)
{
    if (rowStep < 1) rowStep = 1;
    if (rowStart >= H) return 0;
    return (H - rowStart + rowStep - 1) / rowStep;
}

// Bloom (downsample, blur H, blur V) and composite on host buffers
// (black_hole_kernel.cu:41-103, post_process.cuh:76-148).
static void postChain(const RenderParams& P, const float* accum, unsigned char* out)
{
    const int W = P.width, H = P.height;
    const int bw = (W + 1) / 2, bh = (H + 1) / 2;
    std::vector<float> bA((size_t)bw * bh * 4, 0.f), bB((size_t)bw * bh * 4, 0.f);
    if (P.bloomEnabled)
    {
        for (int y = 0; y < bh; ++y)
            for (int x = 0; x < bw; ++x)
            {
                int x0 = x * 2, y0 = y * 2;
                int x1 = (x0 + 1 < W) ? x0 + 1 : x0;
                int y1 = (y0 + 1 < H) ? y0 + 1 : y0;
                const float* a = accum + 4 * ((size_t)y0 * W + x0);
                const float* b = accum + 4 * ((size_t)y0 * W + x1);
                const float* c = accum + 4 * ((size_t)y1 * W + x0);
                const float* d = accum + 4 * ((size_t)y1 * W + x1);
                V3<float> avg = mk3(a[0] + b[0] + c[0] + d[0], a[1] + b[1] + c[1] + d[1], a[2] + b[2] + c[2] + d[2]) * 0.25f;
                V3<float> bp = bloomBrightPass(avg * P.exposure);
                float* o = &bA[4 * ((size_t)y * bw + x)];
                o[0] = bp.x; o[1] = bp.y; o[2] = bp.z; o[3] = 1.f;
            }
        for (int y = 0; y < bh; ++y)
            for (int x = 0; x < bw; ++x)
            {
                V3<float> s = mk3(0.f, 0.f, 0.f);
                for (int k = -4; k <= 4; ++k) s += fetchHalf<float>(bA.data(), x + k, y, bw, bh) * bloomGaussW(k);
                float* o = &bB[4 * ((size_t)y * bw + x)];
                o[0] = s.x; o[1] = s.y; o[2] = s.z; o[3] = 1.f;
            }
        for (int y = 0; y < bh; ++y)
            for (int x = 0; x < bw; ++x)
            {
                V3<float> s = mk3(0.f, 0.f, 0.f);
                for (int k = -4; k <= 4; ++k) s += fetchHalf<float>(bB.data(), x, y + k, bw, bh) * bloomGaussW(k);
                float* o = &bA[4 * ((size_t)y * bw + x)];
                o[0] = s.x; o[1] = s.y; o[2] = s.z; o[3] = 1.f;
            }
    }
    for (int py = 0; py < H; ++py)
        for (int px = 0; px < W; ++px)
        {
            const float* av = accum + 4 * ((size_t)py * W + px);
            V3<float> hdr = mk3(av[0], av[1], av[2]);
            V3<float> bloom = mk3(0.f, 0.f, 0.f);
            if (P.bloomEnabled)
                bloom = sampleHalfBilinear<float>(bA.data(), ((float)px + 0.5f) * 0.5f, ((float)py + 0.5f) * 0.5f, bw, bh);
            finalizePixel<float>(hdr, bloom, P, px, py, out + 4 * ((size_t)py * W + px));
        }
}

// Conserved quantities along one Kerr ray (T = float or double): returns, in
// out[0..3], the max relative drift of E-independent Carter constant
// Qc = p_th^2 + cos^2(th) (Lz^2/sin^2(th) - a^2 E^2), the max |K|/(E^2(r^2+a^2))
// (from the integrator), the initial Carter constant and the step count.
template <class T> struct CarterObs
{
    T a; T* q0; T* maxDrift; bool* init;
    void operator()(const KState<T>& y, T E, T Lz) const
    {
        T c = mcos(y.th), s = msin(y.th);
        T q = y.pth * y.pth + c * c * (Lz * Lz / (s * s) - a * a * E * E);
        if (!*init) { *q0 = q; *init = true; return; }
        T scale = mfmax(mfabs(*q0), E * E * T(0.25f));
        T dr = mfabs(q - *q0) / scale;
        if (dr > *maxDrift) *maxDrift = dr;
    }
};
extern "C" {

int bho_abi_version(void) { return 3; }

int bho_sanitize(RenderParams* P) { return hSanitize(*P) ? 1 : 0; }

void bho_validate_derive(int model, float* M, float* aStar, float* q, int allowNaked, float* out9)
{
    OBHDerived d = hValidateDerive(model, *M, *aStar, *q, allowNaked != 0);
    float o[9] = {d.a, d.Q, d.rPlus, d.rMinus, d.rPhoton, d.rErgo, d.rIsco, (float)d.clamped, (float)d.naked};
    std::memcpy(out9, o, sizeof(o));
}

float bho_isco_numeric(float M, float a, float Q) { return hIscoNumeric(M, a, Q); }

// FP64 critical-band switch for the float paths (mode 0 of bho_trace_frame,
// bho_render_pipeline, bho_render_kernel): band rays are traced in double.
static BandCfg g_band = {false, 3e-2, 1e-3};

void bho_set_band(int on, double crit, double polar)
{
    g_band.on = on != 0;
    g_band.crit = crit;
    g_band.polar = polar;
}

static bool pixelInBand(const RenderParams& P, int px, int py, float jx, float jy)
{
    if (!g_band.on) return false;
    V3<double> cam = mk3((double)P.camPos.x, (double)P.camPos.y, (double)P.camPos.z);
    return rayInBand(P, cam, cameraDirD(P, px, py, (double)jx, (double)jy), g_band);
}

// HDR sample of the float path with the band rule applied.
static V3<float> sampleBanded(const RenderParams& P, int px, int py, float jx, float jy)
{
    if (pixelInBand(P, px, py, jx, jy)) return renderSampleHDR<double>(P, px, py, (double)jx, (double)jy);
    return renderSampleHDR<float>(P, px, py, jx, jy);
}

// Band statistics for given (M, a, Q) and photon constants xi = Lz/E,
// eta = Carter/E^2 (tests of the band restatement).
double bho_band_delta(float M, float a, float Q, double xi, double eta)
{
    RenderParams P;   // defaults; only the model fields below matter
    P.model = BH_KERR_NEWMAN;
    P.M = M; P.aSpin = a; P.Qc = Q;
    return bandCriticalDelta(P, xi, eta);
}
double bho_band_polar(double a, double xi, double eta) { return bandPolarSin2(a, xi, eta); }

// Band classification of a frame: out[W*H] = 1 for band pixels; delta and
// polar (may be NULL) receive the two band statistics (polar 2.0 when not
// applicable).
void bho_band_mask(const RenderParams* Pp, double crit, double polar, int nthreads, unsigned char* out,
                   double* delta, double* pol)
{
    const RenderParams P = *Pp;
    BandCfg cfg = {true, crit, polar};
    V2<float> j = subpixelJitter(P.accumMode, P.sampleIndex);
    V3<double> cam = mk3((double)P.camPos.x, (double)P.camPos.y, (double)P.camPos.z);
    forTiles(P.width, 0, 1, P.height, nthreads, [&](int px, int py, int jr) {
        size_t o = (size_t)jr * P.width + px;
        double d = 0.0, pl = 2.0;
        out[o] = rayInBand(P, cam, cameraDirD(P, px, py, (double)j.x, (double)j.y), cfg, &d, &pl) ? 1 : 0;
        if (delta) delta[o] = d;
        if (pol) pol[o] = pl;
    });
}

// Trace a frame (or every rowStep-th row from rowStart) with the sample
// position given by accumMode/sampleIndex.  mode 0 float (band rays in
// double when the band switch is on), 1 double.
//   hdr4  [nRows*W*4]: HDR rgb + outcome (1 escaped, 0 captured)
//   esc3  [nRows*W*3] double exit directions (may be NULL)
//   steps [nRows*W]   (may be NULL)
//   disk4 [nRows*W*4]: disk accum rgb + transmittance (may be NULL)
// Parameters are used as given (call bho_sanitize first for launch semantics).
void bho_trace_frame(const RenderParams* Pp, int mode, int rowStart, int rowStep, int nthreads, float* hdr4,
                     double* esc3, int* steps, float* disk4)
{
    const RenderParams P = *Pp;
    int nRows = nRowsOf(P.height, rowStart, rowStep);
    V2<float> j = subpixelJitter(P.accumMode, P.sampleIndex);
    forTiles(P.width, rowStart, rowStep, nRows, nthreads, [&](int px, int py, int jr) {
        size_t o = (size_t)jr * P.width + px;
        V3<float> hdr;
        bool esc;
        double e[3];
        int st;
        V3<float> acc;
        float tr;
        if (mode == 1 || (mode == 0 && pixelInBand(P, px, py, j.x, j.y)))
        {
            TraceResult<double> t;
            hdr = renderSampleHDR<double>(P, px, py, (double)j.x, (double)j.y, &t);
            esc = t.escaped;
            e[0] = t.escDir.x; e[1] = t.escDir.y; e[2] = t.escDir.z;
            st = t.steps; acc = t.accum; tr = t.trans;
        }
        else
        {
            TraceResult<float> t;
            hdr = renderSampleHDR<float>(P, px, py, j.x, j.y, &t);
            esc = t.escaped;
            e[0] = t.escDir.x; e[1] = t.escDir.y; e[2] = t.escDir.z;
            st = t.steps; acc = t.accum; tr = t.trans;
        }
        hdr4[4 * o + 0] = hdr.x; hdr4[4 * o + 1] = hdr.y; hdr4[4 * o + 2] = hdr.z;
        hdr4[4 * o + 3] = esc ? 1.f : 0.f;
        if (esc3) { esc3[3 * o] = e[0]; esc3[3 * o + 1] = e[1]; esc3[3 * o + 2] = e[2]; }
        if (steps) steps[o] = st;
        if (disk4) { disk4[4 * o] = acc.x; disk4[4 * o + 1] = acc.y; disk4[4 * o + 2] = acc.z; disk4[4 * o + 3] = tr; }
    });
}

// Exact algorithmic flop counts of one frame (rows as above):
// out[0] trace (integration + disk), out[1] sky shading, out[2] sample setup
// and blend, out[3] finalize (tonemap), out[4] total RK4 steps.
// Counted-float flops of a frame: [0] trace, [1] sky, [2] camera ray and
// blend, [3] finalize, [4] RK4 steps, and with the FP64 band rule on, the
// part of [0] and [4] that falls on band pixels: [5] trace flops, [6] steps,
// [7] band pixels.  (Band pixels are counted as if traced in float.)
void bho_count_flops(const RenderParams* Pp, int rowStart, int rowStep, int nthreads, uint64_t* out8)
{
    const RenderParams P = *Pp;
    int nRows = nRowsOf(P.height, rowStart, rowStep);
    if (nthreads <= 0) nthreads = omp_get_max_threads();
    std::vector<uint64_t> acc((size_t)nthreads * 8 * 5, 0);
    V2<float> jit = subpixelJitter(P.accumMode, P.sampleIndex);
    forTiles(P.width, rowStart, rowStep, nRows, nthreads, [&](int px, int py, int) {
        uint64_t* a = &acc[(size_t)omp_get_thread_num() * 8 * 5];
        uint64_t f0 = tl_flops;
        CF fx = (CF((float)px) + CF(jit.x)) / CF((float)P.width);
        CF fy = (CF((float)py) + CF(jit.y)) / CF((float)P.height);
        V3<CF> dir = cameraRayDir<CF>(P, fx, fy);
        uint64_t f1 = tl_flops;
        V3<CF> cam = mk3(CF(P.camPos.x), CF(P.camPos.y), CF(P.camPos.z));
        TraceResult<CF> t = traceRay<CF>(P, cam, dir, false);
        uint64_t f2 = tl_flops;
        CF gCam = cameraEnergyFactor<CF>(P);
        V3<CF> bg = t.horizon ? mk3(CF(0.f), CF(0.f), CF(0.f)) : backgroundColor<CF>(t.escDir, gCam);
        uint64_t f3 = tl_flops;
        V3<CF> hdr = t.accum + bg * t.trans;
        (void)hdr;
        uint64_t f4 = tl_flops;
        unsigned char o4[4];
        V3<CF> zero = mk3(CF(0.f), CF(0.f), CF(0.f));
        finalizePixel<CF>(mk3(CF(hdr.x.v), CF(hdr.y.v), CF(hdr.z.v)), zero, P, px, py, o4);
        uint64_t f5 = tl_flops;
        a[0] += f2 - f1;
        a[1] += f3 - f2;
        a[2] += (f1 - f0) + (f4 - f3);
        a[3] += f5 - f4;
        a[4] += (uint64_t)t.steps;
        if (pixelInBand(P, px, py, jit.x, jit.y))
        {
            a[5] += f2 - f1;
            a[6] += (uint64_t)t.steps;
            a[7] += 1;
        }
    });
    for (int k = 0; k < 8; ++k)
    {
        uint64_t s = 0;
        for (int t = 0; t < nthreads; ++t) s += acc[(size_t)t * 40 + k];
        out8[k] = s;
    }
}

// Full launchRenderPipeline semantics (black_hole_kernel.cu:108-134) on host
// buffers: accum float4[W*H] in/out, out uchar4[W*H].  Parameters are copied
// and sanitized.
void bho_render_pipeline(const RenderParams* Pp, float* accum, unsigned char* out, int nthreads)
{
    RenderParams P = *Pp;
    hSanitize(P);
    const int W = P.width, H = P.height;
    V2<float> j = subpixelJitter(P.accumMode, P.sampleIndex);
    forTiles(W, 0, 1, H, nthreads, [&](int px, int py, int) {
        V3<float> cur = sampleBanded(P, px, py, j.x, j.y);
        size_t idx = (size_t)py * W + px;
        float* pv = accum + 4 * idx;
        V3<float> prev = mk3(pv[0], pv[1], pv[2]);
        if (!std::isfinite(prev.x) || !std::isfinite(prev.y) || !std::isfinite(prev.z)) prev = cur;
        V3<float> c = accumBlend(prev, cur, P.accumMode, P.sampleIndex);
        pv[0] = c.x; pv[1] = c.y; pv[2] = c.z; pv[3] = 1.f;
    });
    postChain(P, accum, out);
}

// launchRenderKernel semantics (black_hole_kernel.cu:140-163)
void bho_render_kernel(const RenderParams* Pp, unsigned char* out, int nthreads)
{
    RenderParams P = *Pp;
    hSanitize(P);
    P.bloomEnabled = 0;
    forTiles(P.width, 0, 1, P.height, nthreads, [&](int px, int py, int) {
        V3<float> hdr = sampleBanded(P, px, py, 0.5f, 0.5f);
        finalizePixel<float>(hdr, mk3(0.f, 0.f, 0.f), P, px, py, out + 4 * ((size_t)py * P.width + px));
    });
}

// Post chain only: given an accum image, bloom + composite (parameters used
// as given).
void bho_post(const RenderParams* Pp, const float* accum, unsigned char* out) { postChain(*Pp, accum, out); }

// Probes (test_main.cu:71-98).  rec: 8 floats per ray
// [outcome, esc.x, esc.y, esc.z, maxHviol, lum, steps, trans].  mode 0 float,
// 1 double.
void bho_probe_grid(const RenderParams* Pp, int nx, int ny, int mode, double* rec, int nthreads)
{
    const RenderParams P = *Pp;
    if (nthreads <= 0) nthreads = omp_get_max_threads();
#pragma omp parallel for schedule(dynamic, 4) num_threads(nthreads)
    for (int k = 0; k < nx * ny; ++k)
    {
        int i = k % nx, jj = k / nx;
        double* o = rec + 8 * (size_t)k;
        if (mode == 1)
        {
            // pixel-centre positions as the float probe computes them
            double fx = (double)(((float)i + 0.5f) / (float)nx);
            double fy = (double)(((float)jj + 0.5f) / (float)ny);
            V3<double> dir = cameraRayDir<double>(P, fx, fy);
            V3<double> cam = mk3((double)P.camPos.x, (double)P.camPos.y, (double)P.camPos.z);
            TraceResult<double> t = traceRay<double>(P, cam, dir, true);
            o[0] = t.escaped ? 1 : 0; o[1] = t.escDir.x; o[2] = t.escDir.y; o[3] = t.escDir.z;
            o[4] = t.maxHviol; o[5] = t.accum.x + t.accum.y + t.accum.z; o[6] = t.steps; o[7] = t.trans;
        }
        else
        {
            float fx = ((float)i + 0.5f) / (float)nx;
            float fy = ((float)jj + 0.5f) / (float)ny;
            V3<float> dir = cameraRayDir<float>(P, fx, fy);
            V3<float> cam = mk3(P.camPos.x, P.camPos.y, P.camPos.z);
            TraceResult<float> t = traceRay<float>(P, cam, dir, true);
            o[0] = t.escaped ? 1 : 0; o[1] = t.escDir.x; o[2] = t.escDir.y; o[3] = t.escDir.z;
            o[4] = t.maxHviol; o[5] = t.accum.x + t.accum.y + t.accum.z; o[6] = t.steps; o[7] = t.trans;
        }
    }
}

// Explicit directions (float inputs, used as-is in float; widened for double).
void bho_probe_dirs(const RenderParams* Pp, const float* dirs, int n, int mode, double* rec, int nthreads)
{
    const RenderParams P = *Pp;
    if (nthreads <= 0) nthreads = omp_get_max_threads();
#pragma omp parallel for schedule(dynamic, 1) num_threads(nthreads)
    for (int k = 0; k < n; ++k)
    {
        double* o = rec + 8 * (size_t)k;
        if (mode == 1)
        {
            V3<double> dir = mk3((double)dirs[3 * k], (double)dirs[3 * k + 1], (double)dirs[3 * k + 2]);
            V3<double> cam = mk3((double)P.camPos.x, (double)P.camPos.y, (double)P.camPos.z);
            TraceResult<double> t = traceRay<double>(P, cam, dir, true);
            o[0] = t.escaped ? 1 : 0; o[1] = t.escDir.x; o[2] = t.escDir.y; o[3] = t.escDir.z;
            o[4] = t.maxHviol; o[5] = t.accum.x + t.accum.y + t.accum.z; o[6] = t.steps; o[7] = t.trans;
        }
        else
        {
            V3<float> dir = mk3(dirs[3 * k], dirs[3 * k + 1], dirs[3 * k + 2]);
            V3<float> cam = mk3(P.camPos.x, P.camPos.y, P.camPos.z);
            TraceResult<float> t = traceRay<float>(P, cam, dir, true);
            o[0] = t.escaped ? 1 : 0; o[1] = t.escDir.x; o[2] = t.escDir.y; o[3] = t.escDir.z;
            o[4] = t.maxHviol; o[5] = t.accum.x + t.accum.y + t.accum.z; o[6] = t.steps; o[7] = t.trans;
        }
    }
}

void bho_kerr_invariants(const RenderParams* Pp, const float* dir, int mode, double* out4)
{
    const RenderParams P = *Pp;
    if (mode == 1)
    {
        double q0 = 0, md = 0; bool init = false;
        CarterObs<double> ob{(double)P.aSpin, &q0, &md, &init};
        V3<double> d = mk3((double)dir[0], (double)dir[1], (double)dir[2]);
        V3<double> cam = mk3((double)P.camPos.x, (double)P.camPos.y, (double)P.camPos.z);
        TraceResult<double> t = traceKerr<double, CarterObs<double>>(P, cam, d, true, ob);
        out4[0] = md; out4[1] = t.maxHviol; out4[2] = q0; out4[3] = t.steps;
    }
    else
    {
        float q0 = 0, md = 0; bool init = false;
        CarterObs<float> ob{P.aSpin, &q0, &md, &init};
        V3<float> d = mk3(dir[0], dir[1], dir[2]);
        V3<float> cam = mk3(P.camPos.x, P.camPos.y, P.camPos.z);
        TraceResult<float> t = traceKerr<float, CarterObs<float>>(P, cam, d, true, ob);
        out4[0] = md; out4[1] = t.maxHviol; out4[2] = q0; out4[3] = t.steps;
    }
}

// Final integrator state of one ray (double): out[0..7] = state (see
// FinalState), out[8] outcome, out[9..11] exit direction, out[12] steps.
void bho_final_state(const RenderParams* Pp, const double* dir, double* out13)
{
    const RenderParams P = *Pp;
    V3<double> d = mk3(dir[0], dir[1], dir[2]);
    V3<double> cam = mk3((double)P.camPos.x, (double)P.camPos.y, (double)P.camPos.z);
    FinalState<double> f{};
    TraceResult<double> t;
    if (P.model == BH_KERR || P.model == BH_KERR_NEWMAN)
        t = traceKerr<double, NoObserver<double>>(P, cam, d, false, NoObserver<double>(), &f);
    else t = traceSpherical<double>(P, cam, d, &f);
    for (int k = 0; k < 8; ++k) out13[k] = f.v[k];
    out13[8] = t.escaped ? 1 : 0;
    out13[9] = t.escDir.x; out13[10] = t.escDir.y; out13[11] = t.escDir.z;
    out13[12] = t.steps;
}

// ------------------------------------------------------------ scalar exports
float bho_hash12(float x, float y) { return hash12(mk2(x, y)); }
void bho_hash33(const float* p, float* out)
{
    V3<float> h = hash33(mk3(p[0], p[1], p[2]));
    out[0] = h.x; out[1] = h.y; out[2] = h.z;
}
float bho_vnoise(float x, float y) { return vnoise(mk2(x, y)); }
float bho_fbm(float x, float y) { return fbm(mk2(x, y)); }
void bho_blackbody(float k, float* rgb)
{
    V3<float> c = blackbodyRGB(k);
    rgb[0] = c.x; rgb[1] = c.y; rgb[2] = c.z;
}
void bho_aces(const float* in, float* out)
{
    V3<float> c = acesToneMap(mk3(in[0], in[1], in[2]));
    out[0] = c.x; out[1] = c.y; out[2] = c.z;
}
void bho_background(const float* dir, float gCam, float* rgb)
{
    V3<float> c = backgroundColor(mk3(dir[0], dir[1], dir[2]), gCam);
    rgb[0] = c.x; rgb[1] = c.y; rgb[2] = c.z;
}
// One star lattice (trace.cuh:131-162) at the point d (not normalised here,
// so tests can place d * cellScale anywhere in lattice space).
void bho_star_layer(const float* d, float cellScale, float intensity, float coreSharp, float* rgb)
{
    V3<float> c = starLayer(mk3(d[0], d[1], d[2]), cellScale, intensity, coreSharp);
    rgb[0] = c.x; rgb[1] = c.y; rgb[2] = c.z;
}
// The volumetric disk on one chord [a, b] (trace.cuh:269-403) for photon
// (E, Lz) and hole spin/charge (a, Q) from ctx4 = {E, Lz, a, Q}; io4 holds
// the accumulated radiance and transmittance, updated in place.
void bho_disk_segment(const RenderParams* Pp, const float* ctx4, const float* a3, const float* b3, float tEmit,
                      float* io4)
{
    PT<float> P(*Pp);
    DiskCtx<float> ctx{ctx4[0], ctx4[1], ctx4[2], ctx4[3]};
    V3<float> acc = mk3(io4[0], io4[1], io4[2]);
    float trans = io4[3];
    sampleDiskSegment<float>(P, ctx, mk3(a3[0], a3[1], a3[2]), mk3(b3[0], b3[1], b3[2]), tEmit, acc, trans);
    io4[0] = acc.x; io4[1] = acc.y; io4[2] = acc.z; io4[3] = trans;
}
void bho_finalize(const float* hdr, const float* bloom, const RenderParams* P, int px, int py, unsigned char* out)
{
    finalizePixel<float>(mk3(hdr[0], hdr[1], hdr[2]), mk3(bloom[0], bloom[1], bloom[2]), *P, px, py, out);
}
float bho_emitter_g(float r, float M, float a, float Q, float E, float Lz) { return bhEmitterGFactor(r, M, a, Q, E, Lz); }
float bho_flux_weight(float r, float rIn) { return bhThinDiskFluxWeight(r, rIn); }
float bho_omega(float r, float M, float a, float Q) { return bhOmegaCircular(r, M, a, Q); }
float bho_camera_energy(const RenderParams* P) { return cameraEnergyFactor<float>(*P); }
void bho_jitter(int accumMode, int sampleIndex, float* out2)
{
    V2<float> j = subpixelJitter(accumMode, sampleIndex);
    out2[0] = j.x; out2[1] = j.y;
}
void bho_photon_EL(const float* origin, const float* dir, float M, float a, float Q, float* out2)
{
    float E, L;
    photonConservedEL(mk3(origin[0], origin[1], origin[2]), mk3(dir[0], dir[1], dir[2]), M, a, Q, E, L);
    out2[0] = E; out2[1] = L;
}
void bho_camera_dir(const RenderParams* P, float fx, float fy, float* out3)
{
    V3<float> d = cameraRayDir<float>(*P, fx, fy);
    out3[0] = d.x; out3[1] = d.y; out3[2] = d.z;
}

}  // extern "C"
