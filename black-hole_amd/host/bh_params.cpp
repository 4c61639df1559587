// ---------------------------------------------------------------------------
// bh_params.cpp -- host-side parameter layer of libbhtrace.so:
//   * validation / derivation of the black-hole parameters and launch-time
//     sanitation, with the reference's float semantics (metric.cuh:74-419);
//   * FP64 horizon and ISCO roots (Bardeen-Press-Teukolsky closed form for
//     Kerr, a dual-number Newton/bisection root of dE/dr = 0 for Kerr-Newman);
//   * the orbit camera (main.cpp:186-228) and the named workload presets.
// ---------------------------------------------------------------------------
#include <cmath>
#include <cstdint>
#include <cstring>
#include "../../include/bh/bhtrace.h"

namespace {

constexpr float kExtremal = 0.995f;   // metric.cuh:53
constexpr float kNaked = 2.25f;       // metric.cuh:55
constexpr float kSingCut = 0.05f;     // metric.cuh:58

inline bool isModelRotating(int m) { return m == BH_KERR || m == BH_KERR_NEWMAN; }
inline bool isModelCharged(int m) { return m == BH_REISSNER_NORDSTROM || m == BH_KERR_NEWMAN; }

// r = M +/- sqrt(max(M^2 - a^2 - Q^2, 0))  (metric.cuh:74-86)
inline float horizon(float M, float a, float Q, float sign)
{
    float d = M * M - a * a - Q * Q;
    d = d < 0.f ? 0.f : d;
    return M + sign * std::sqrt(d);
}

// Circular equatorial orbit angular velocity and metric (metric.cuh:145-162)
inline float omegaCirc(float r, float M, float a, float Q)
{
    float w2 = M * r - Q * Q;
    if (w2 <= 0.f) return 0.f;
    float sw = std::sqrt(w2);
    return sw / (r * r + a * sw);
}

// Float scan for the energy minimum, 0.001 M steps (metric.cuh:224-248)
float iscoScan(float M, float a, float Q)
{
    const float disc = M * M - a * a - Q * Q;
    const float rp = (disc >= 0.f) ? horizon(M, a, Q, 1.f) : kSingCut * M;
    float best = 6.f * M, bestE = 1e30f;
    bool found = false;
    const float rStart = rp > 1e-6f ? rp * 1.02f : 0.5f * M;
    const float rEnd = 12.f * M, dr = 0.001f * M;
    for (float r = rStart; r <= rEnd; r += dr)
    {
        float Om = omegaCirc(r, M, a, Q);
        if (Om <= 0.f) continue;
        float r2 = r * r, m2r = 2.f * M * r - Q * Q;
        float gtt = -(1.f - m2r / r2), gtp = -a * m2r / r2, gpp = r2 + a * a + m2r * a * a / r2;
        float den = -(gtt + 2.f * Om * gtp + Om * Om * gpp);
        if (den <= 1e-6f) continue;
        float E = -(gtt + Om * gtp) / std::sqrt(den);
        if (E < bestE) { bestE = E; best = r; found = true; }
    }
    if (!found) best = (6.f * M > rp * 1.5f) ? 6.f * M : rp * 1.5f;
    const float lo = rp > 1e-6f ? rp * 1.15f : 0.5f * M;
    return best > lo ? best : lo;
}

// Photon-orbit markers (metric.cuh:96-129)
inline float photonSpherical(float M, float Q)
{
    float d = 9.f * M * M - 8.f * Q * Q;
    d = d < 0.f ? 0.f : d;
    return 0.5f * (3.f * M + std::sqrt(d));
}
inline float photonKerr(float M, float aStar)
{
    float x = aStar > 1.f ? 1.f : (aStar < -1.f ? -1.f : aStar);
    return 2.f * M * (1.f + std::cos((2.f / 3.f) * std::acos(-x)));
}
inline float photonKN(float M, float aStar, float q)
{
    float rK = photonKerr(M, aStar), rS = photonSpherical(M, q * M);
    if (std::fabs(aStar) < 1e-5f) return rS;
    float w = q * q;
    w = w < 0.f ? 0.f : (w > 1.f ? 1.f : w);
    return (1.f - 0.35f * w) * rK + 0.35f * w * rS;
}

// ---- dual numbers for the FP64 ISCO root
struct Dual { double v, d; };
inline Dual operator+(Dual a, Dual b) { return {a.v + b.v, a.d + b.d}; }
inline Dual operator-(Dual a, Dual b) { return {a.v - b.v, a.d - b.d}; }
inline Dual operator*(Dual a, Dual b) { return {a.v * b.v, a.d * b.v + a.v * b.d}; }
inline Dual operator/(Dual a, Dual b) { return {a.v / b.v, (a.d * b.v - a.v * b.d) / (b.v * b.v)}; }
inline Dual operator*(double s, Dual a) { return {s * a.v, s * a.d}; }
inline Dual dsqrt(Dual a) { double s = std::sqrt(a.v); return {s, a.d / (2.0 * s)}; }
inline Dual cst(double x) { return {x, 0.0}; }

// Specific energy of the prograde (+phi) circular equatorial orbit and its
// r-derivative; returns false where the orbit does not exist or is not
// timelike.
bool orbitEnergy(double r, double M, double a, double Q, double& E, double& dE)
{
    Dual R{r, 1.0};
    Dual w2 = cst(M) * R - cst(Q * Q);
    if (w2.v <= 0.0) return false;
    Dual sw = dsqrt(w2);
    Dual Om = sw / (R * R + cst(a) * sw);
    Dual r2 = R * R, m2r = 2.0 * (cst(M) * R) - cst(Q * Q);
    Dual gtt = cst(0.0) - (cst(1.0) - m2r / r2);
    Dual gtp = cst(0.0) - cst(a) * m2r / r2;
    Dual gpp = r2 + cst(a * a) + m2r * cst(a * a) / r2;
    Dual den = cst(0.0) - (gtt + 2.0 * (Om * gtp) + Om * Om * gpp);
    if (den.v <= 0.0) return false;
    Dual e = (cst(0.0) - (gtt + Om * gtp)) / dsqrt(den);
    E = e.v;
    dE = e.d;
    return true;
}

// SplitMix64 (Steele, Lea & Flood 2014): the batch generator's PRNG.
struct SplitMix64
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

}  // namespace

extern "C" {

void bhDefaultRenderParams(RenderParams* p)
{
    RenderParams d;
    *p = d;
}

void bhValidateAndDerive(int model, float* Mp, float* aStarp, float* qp, int allowNakedI, BHDerived* out)
{
    float& M = *Mp;
    float& aStar = *aStarp;
    float& q = *qp;
    const bool allowNaked = allowNakedI != 0;
    BHDerived d = {0.f, 0.f, 1.f, 0.f, 1.5f, 1.f, 3.f, 0, 0};
    // non-finite or non-positive inputs fall back to defaults (metric.cuh:262-265)
    if (!(M > 0.f) || !std::isfinite(M)) { M = 0.5f; d.clamped = 1; }
    if (!std::isfinite(aStar)) { aStar = 0.f; d.clamped = 1; }
    if (!std::isfinite(q)) { q = 0.f; d.clamped = 1; }
    // knob range (metric.cuh:267-271)
    const float knob = allowNaked ? std::sqrt(kNaked) : kExtremal;
    if (aStar < -knob) { aStar = -knob; d.clamped = 1; }
    if (aStar > knob) { aStar = knob; d.clamped = 1; }
    if (q < 0.f) { q = 0.f; d.clamped = 1; }
    if (q > knob) { q = knob; d.clamped = 1; }
    // per-model effective spin/charge, scaled back inside the bound (metric.cuh:274-287)
    float aE = isModelRotating(model) ? aStar : 0.f;
    float qE = isModelCharged(model) ? q : 0.f;
    const float bound = allowNaked ? kNaked : kExtremal;
    const float s2 = aE * aE + qE * qE;
    if (s2 > bound)
    {
        const float k = std::sqrt(bound / s2);
        aE *= k;
        qE *= k;
        if (isModelRotating(model)) aStar = aE;
        if (isModelCharged(model)) q = qE;
        d.clamped = 1;
    }
    d.a = aE * M;
    d.Q = qE * M;
    // horizons, ergosphere, photon orbit, ISCO (metric.cuh:292-327)
    if (M * M - d.a * d.a - d.Q * d.Q < 0.f) { d.naked = 1; d.rPlus = 0.f; d.rMinus = 0.f; }
    else { d.rPlus = horizon(M, d.a, d.Q, 1.f); d.rMinus = horizon(M, d.a, d.Q, -1.f); }
    d.rErgo = d.naked ? 0.f : horizon(M, 0.f, d.Q, 1.f);
    if (isModelRotating(model))
    {
        d.rPhoton = (model == BH_KERR_NEWMAN) ? photonKN(M, aE, qE) : photonKerr(M, aE);
        if (d.rPlus > 0.f) { if (d.rPhoton < d.rPlus * 1.02f) d.rPhoton = d.rPlus * 1.02f; }
        else if (d.rPhoton < kSingCut * M * 2.f) d.rPhoton = 3.f * M;
    }
    else
    {
        d.rPhoton = photonSpherical(M, d.Q);
        if (d.naked && !(d.rPhoton > 0.f)) d.rPhoton = 3.f * M;
    }
    d.rIsco = iscoScan(M, d.a, d.Q);
    *out = d;
}

int bhSanitizeRenderParams(RenderParams* pp)
{
    RenderParams& P = *pp;
    bool fixed = false;
    auto clampField = [&fixed](float& v, float lo, float hi, float fallback) {
        if (!std::isfinite(v)) { v = fallback; fixed = true; }
        if (v < lo) { v = lo; fixed = true; }
        if (v > hi) { v = hi; fixed = true; }
    };
    auto clampInt = [&fixed](int& v, int lo, int hi) {
        if (v < lo) { v = lo; fixed = true; }
        if (v > hi) { v = hi; fixed = true; }
    };
    // frame, integration and camera scalars (metric.cuh:354-363)
    if (P.width < 1) { P.width = 1; fixed = true; }
    if (P.height < 1) { P.height = 1; fixed = true; }
    clampField(P.dPhi, 1e-4f, 0.1f, 0.012f);
    clampInt(P.maxSteps, 16, 200000);
    clampField(P.exposure, 0.01f, 100.f, 1.f);
    clampField(P.M, 1e-3f, 100.f, 0.5f);
    clampField(P.tanHalfFov, 0.05f, 5.f, 0.5773503f);
    clampField(P.aspect, 0.05f, 20.f, 16.f / 9.f);
    if (P.model < 0 || P.model > 3) { P.model = 0; fixed = true; }
    // black-hole consistency (metric.cuh:366-381)
    {
        float aStar = P.aSpin / P.M, q = P.Qc / P.M, M = P.M;
        const bool naked = P.allowNaked != 0;
        const float bound = naked ? kNaked : kExtremal;
        const bool bad = !std::isfinite(aStar) || !std::isfinite(q) || aStar * aStar + q * q > bound + 1e-4f ||
                         q < 0.f || !std::isfinite(P.rPlus) || (!naked && P.rPlus <= 0.f) ||
                         (naked && P.rPlus < 0.f);
        if (bad)
        {
            BHDerived d;
            bhValidateAndDerive(P.model, &M, &aStar, &q, naked ? 1 : 0, &d);
            P.M = M; P.aSpin = d.a; P.Qc = d.Q;
            P.rPlus = d.rPlus; P.rPhoton = d.rPhoton; P.rErgo = d.rErgo;
            fixed = true;
        }
    }
    // camera (metric.cuh:383-402)
    auto fin3 = [](const bh_float3& v) { return std::isfinite(v.x) && std::isfinite(v.y) && std::isfinite(v.z); };
    if (!fin3(P.camPos) || !fin3(P.camForward) || !fin3(P.camRight) || !fin3(P.camUp))
    {
        P.camPos = bh_float3{13.f, 3.2f, 8.f};
        P.camForward = bh_float3{-0.81f, -0.20f, -0.55f};
        P.camRight = bh_float3{0.56f, 0.f, -0.83f};
        P.camUp = bh_float3{-0.17f, 0.98f, -0.11f};
        fixed = true;
    }
    {
        const float r0 = std::sqrt(P.camPos.x * P.camPos.x + P.camPos.y * P.camPos.y + P.camPos.z * P.camPos.z);
        const float rMin = (P.rPlus > 1e-6f) ? P.rPlus * 2.0f : (kSingCut * P.M * 8.f);
        if (r0 < rMin)
        {
            const float s = rMin / (r0 > 1e-6f ? r0 : 1e-6f);
            P.camPos.x *= s; P.camPos.y *= s; P.camPos.z *= s;
            fixed = true;
        }
    }
    // disk, accumulation and shading knobs (metric.cuh:404-417)
    const float diskLo = (P.rPlus > 1e-6f) ? P.rPlus * 1.05f : (kSingCut * P.M * 2.f);
    clampField(P.diskInner, diskLo, 1000.f, diskLo * 3.f);
    clampField(P.diskOuter, P.diskInner * 1.2f, 2000.f, P.diskInner * 4.f);
    clampInt(P.sampleIndex, 0, 65536);
    if (P.accumMode < 0 || P.accumMode > 2) { P.accumMode = 0; fixed = true; }
    clampField(P.bloomStrength, 0.f, 4.f, 0.35f);
    clampField(P.diskEmisScale, 0.f, 50.f, 1.5f);
    clampField(P.diskAbsScale, 0.f, 20.f, 2.2f);
    clampField(P.hotSpotStrength, 0.f, 8.f, 0.65f);
    if (P.hotSpotsEnabled != 0 && P.hotSpotsEnabled != 1) { P.hotSpotsEnabled = 1; fixed = true; }
    return fixed ? 1 : 0;
}

double bhHorizonOuterFP64(double M, double a, double Q)
{
    double d = M * M - a * a - Q * Q;
    if (d < 0.0) return 0.0;
    return M + std::sqrt(d);
}

double bhIscoFP64(double M, double a, double Q)
{
    if (!(M > 0.0)) return NAN;
    if (Q == 0.0)
    {
        // Bardeen, Press & Teukolsky (1972), orbit in +phi, signed spin
        double x = a / M;
        if (std::fabs(x) > 1.0) x = x > 0 ? 1.0 : -1.0;
        double z1 = 1.0 + std::cbrt(1.0 - x * x) * (std::cbrt(1.0 + x) + std::cbrt(1.0 - x));
        double z2 = std::sqrt(3.0 * x * x + z1 * z1);
        double sgn = x >= 0.0 ? 1.0 : -1.0;
        return M * (3.0 + z2 - sgn * std::sqrt((3.0 - z1) * (3.0 + z1 + 2.0 * z2)));
    }
    // Kerr-Newman: minimum of the specific energy, dE/dr = 0.  Coarse scan for
    // the sign change, then bisection on the exact (dual-number) derivative.
    double rp = bhHorizonOuterFP64(M, a, Q);
    double r0 = (rp > 0.0 ? rp : kSingCut * M) * 1.0001;
    double r1 = 40.0 * M;
    const int n = 40000;
    double prevR = -1.0, prevD = 0.0;
    for (int i = 0; i <= n; ++i)
    {
        double r = r0 + (r1 - r0) * i / n, E, dE;
        if (!orbitEnergy(r, M, a, Q, E, dE)) { prevR = -1.0; continue; }
        if (prevR > 0.0 && prevD < 0.0 && dE >= 0.0)
        {
            double lo = prevR, hi = r;
            for (int it = 0; it < 200 && hi - lo > 1e-15 * hi; ++it)
            {
                double mid = 0.5 * (lo + hi), Em, dm;
                if (!orbitEnergy(mid, M, a, Q, Em, dm)) { lo = mid; continue; }
                if (dm < 0.0) lo = mid; else hi = mid;
            }
            return 0.5 * (lo + hi);
        }
        prevR = r;
        prevD = dE;
    }
    return NAN;
}

// fillCamera with identity body orientation (main.cpp:186-228) and an
// explicit vertical field of view (the reference fixes 60 degrees).
void bhFillCamera(RenderParams* p, float az, float el, float dist, float fovY)
{
    const float ca = std::cos(az), sa = std::sin(az), ce = std::cos(el), se = std::sin(el);
    float px = dist * ce * ca, py = dist * se, pz = dist * ce * sa;
    const float fl = std::sqrt(px * px + py * py + pz * pz);
    float fx = -px / fl, fy = -py / fl, fz = -pz / fl;
    // right = normalize(forward x +Y)
    float rx = -fz, ry = 0.f, rz = fx;
    float rl = std::sqrt(rx * rx + ry * ry + rz * rz);
    if (rl < 1e-6f) { rx = 1.f; ry = 0.f; rz = 0.f; rl = 1.f; }
    rx /= rl;
    rz /= rl;
    auto renorm = [](float& x, float& y, float& z) {
        float l = std::sqrt(x * x + y * y + z * z);
        if (l < 1e-8f) { x = 0.f; y = 1.f; z = 0.f; return; }
        x /= l; y /= l; z /= l;
    };
    renorm(fx, fy, fz);
    renorm(rx, ry, rz);
    // up = right x forward
    float ux = ry * fz - rz * fy, uy = rz * fx - rx * fz, uz = rx * fy - ry * fx;
    renorm(ux, uy, uz);
    p->camPos = bh_float3{px, py, pz};
    p->camForward = bh_float3{fx, fy, fz};
    p->camRight = bh_float3{rx, ry, rz};
    p->camUp = bh_float3{ux, uy, uz};
    p->tanHalfFov = std::tan(fovY * 3.14159265358979f / 180.f * 0.5f);
    p->aspect = (p->height > 0) ? (float)p->width / (float)p->height : 16.f / 9.f;
}

int bhWorkloadParams(const char* name, int frame, RenderParams* out)
{
    if (!name || !out) return -1;
    RenderParams P;
    P.M = 0.5f;
    P.swapRB = 0;
    P.diskEmisScale = 1.5f;
    P.diskAbsScale = 2.2f;
    P.bloomStrength = 0.35f;
    P.allowNaked = 0;
    P.sampleIndex = 0;
    P.accumMode = 0;
    P.exposure = 1.0f;
    const float deg = 3.14159265358979f / 180.f;
    auto bh = [&P](int model, float aStar, float q) {
        float M = 0.5f, a = aStar, qq = q;
        BHDerived d;
        bhValidateAndDerive(model, &M, &a, &qq, 0, &d);
        P.model = model; P.M = M; P.aSpin = d.a; P.Qc = d.Q;
        P.rPlus = d.rPlus; P.rPhoton = d.rPhoton; P.rErgo = d.rErgo;
    };
    if (std::strcmp(name, "schw_256_balanced") == 0)
    {
        P.width = 256; P.height = 256;
        bh(BH_SCHWARZSCHILD, 0.f, 0.f);
        P.dPhi = 0.012f; P.maxSteps = 2000;
        P.diskEnabled = 1; P.diskInner = 3.0f; P.diskOuter = 12.0f; P.diskTime = 0.f;
        P.hotSpotsEnabled = 0; P.hotSpotStrength = 0.65f;
        P.bloomEnabled = 0;
        bhFillCamera(&P, 0.6f, 10.f * deg, 20.f, 60.f);
    }
    else if (std::strcmp(name, "kerr09_1080p_high") == 0)
    {
        static const float times[3] = {0.0f, 0.5f, 1.0f};
        if (frame < 0 || frame > 2) return -1;
        P.width = 1920; P.height = 1080;
        bh(BH_KERR, 0.9f, 0.f);
        P.dPhi = 0.007f; P.maxSteps = 3500;
        P.diskEnabled = 1;
        P.diskInner = (float)bhIscoFP64(P.M, P.aSpin, P.Qc);
        P.diskOuter = 20.0f; P.diskTime = times[frame];
        P.hotSpotsEnabled = 1; P.hotSpotStrength = 0.65f;
        P.bloomEnabled = 1;
        bhFillCamera(&P, 0.6f, 5.f * deg, 15.f, 50.f);
    }
    else if (std::strcmp(name, "kn_m07_q04_2160p_high") == 0)
    {
        P.width = 3840; P.height = 2160;
        bh(BH_KERR_NEWMAN, -0.7f, 0.4f);
        P.dPhi = 0.007f; P.maxSteps = 3500;
        P.diskEnabled = 1;
        P.diskInner = (float)bhIscoFP64(P.M, P.aSpin, P.Qc);
        P.diskOuter = 8.0f; P.diskTime = 0.f;
        P.hotSpotsEnabled = 1; P.hotSpotStrength = 0.65f;
        P.bloomEnabled = 1;
        bhFillCamera(&P, 0.6f, 30.f * deg, 8.f, 70.f);
    }
    else return -1;
    *out = P;
    return 0;
}

int bhSurrogateConfigs(uint64_t seed, int n, RenderParams* out)
{
    if (!out || n < 0) return -1;
    const double deg = 3.14159265358979323846 / 180.0;
    SplitMix64 rng{seed};
    for (int c = 0; c < n; ++c)
    {
        double aStar, q;
        do
        {
            aStar = rng.uniform(-0.99, 0.99);
            q = rng.uniform(0.0, 0.6);
        } while (aStar * aStar + q * q > 0.995);
        const double dist = std::exp(rng.uniform(std::log(6.0), std::log(40.0)));
        const double el = rng.uniform(-80.0, 80.0) * deg;
        RenderParams P;
        P.width = 64; P.height = 64;
        float M = 0.5f, a = (float)aStar, qq = (float)q;
        BHDerived d;
        bhValidateAndDerive(BH_KERR_NEWMAN, &M, &a, &qq, 0, &d);
        P.model = BH_KERR_NEWMAN; P.M = M; P.aSpin = d.a; P.Qc = d.Q;
        P.rPlus = d.rPlus; P.rPhoton = d.rPhoton; P.rErgo = d.rErgo;
        P.dPhi = 0.007f; P.maxSteps = 3500;
        P.diskEnabled = 1;
        P.diskInner = (float)bhIscoFP64(P.M, P.aSpin, P.Qc);
        P.diskOuter = 20.0f;
        P.diskTime = 0.f;
        bhFillCamera(&P, 0.f, (float)el, (float)dist, 50.f);
        out[c] = P;
    }
    return 0;
}

int bhWorkloadViews(const char* name, uint64_t seed, int maxViews, RenderParams* out)
{
    if (!name || std::strcmp(name, "uneven_views") != 0 || maxViews < 0) return -1;
    static const int res[4][2] = {{320, 180}, {640, 360}, {1280, 720}, {1920, 1080}};
    const double deg = 3.14159265358979323846 / 180.0;
    SplitMix64 rng{seed};
    for (int v = 0; v < maxViews; ++v)
    {
        RenderParams P;
        const int r = (int)(rng.uniform() * 4.0) & 3;
        P.width = res[r][0];
        P.height = res[r][1];
        const double aStar = rng.uniform(-0.99, 0.99);
        const double qMax = std::sqrt(std::fmax(0.98 - aStar * aStar, 0.0));
        const double q = rng.uniform(0.0, qMax);
        const double dist = std::exp(rng.uniform(std::log(6.0), std::log(40.0)));
        const double el = rng.uniform(-80.0, 80.0) * deg;
        const double fov = rng.uniform(40.0, 90.0);
        const bool disk = rng.uniform() < 0.5;
        const double t = rng.uniform(0.0, 10.0);
        float M = 0.5f, a = (float)aStar, qq = (float)q;
        BHDerived d;
        bhValidateAndDerive(BH_KERR_NEWMAN, &M, &a, &qq, 0, &d);
        P.model = BH_KERR_NEWMAN;
        P.M = M; P.aSpin = d.a; P.Qc = d.Q;
        P.rPlus = d.rPlus; P.rPhoton = d.rPhoton; P.rErgo = d.rErgo;
        P.dPhi = 0.012f; P.maxSteps = 2000;
        P.diskEnabled = disk ? 1 : 0;
        P.diskInner = (float)bhIscoFP64(P.M, P.aSpin, P.Qc);
        P.diskOuter = 8.0f;
        P.diskTime = (float)t;
        P.hotSpotsEnabled = 1; P.hotSpotStrength = 0.65f;
        P.bloomEnabled = 1; P.bloomStrength = 0.35f;
        P.exposure = 1.0f;
        bhFillCamera(&P, 0.6f, (float)el, (float)dist, (float)fov);
        out[v] = P;
    }
    return maxViews;
}

}  // extern "C"
