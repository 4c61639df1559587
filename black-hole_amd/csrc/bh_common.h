// ---------------------------------------------------------------------------
// bh_common.h -- device-side building blocks of the gfx950 geodesic tracer.
//
// Physics and shading follow the reference's trace.cuh / metric.cuh /
// vec_math.cuh / post_process.cuh (file:line cited per function).  The
// arithmetic goes through a math policy:
//
//   MathExact -- IEEE division and square root, OCML transcendentals; with
//                -ffp-contract=off the float results follow the reference's
//                host semantics operation for operation.
//   MathFast  -- hardware reciprocal / rsq / sin / cos / exp2 / log2 (about
//                1 ulp, like the reference's own --use_fast_math GPU build);
//                compiled with FMA contraction.
//
// Everything whose output feeds a hash (lattice coordinates, the hash
// arithmetic itself, the dither) is evaluated with contraction disabled in
// both policies: a 1-ulp change there reshuffles the noise or the stars.
// ---------------------------------------------------------------------------
#pragma once
#include <hip/hip_runtime.h>
#include "../../include/bh/render_params.h"

namespace bhk {

#define BHK_INL __device__ __forceinline__
#define BHK_HD __host__ __device__ inline

// ------------------------------------------------------------------ policies
struct MathExact
{
    static constexpr bool fast = false;
    BHK_INL static float div(float a, float b) { return a / b; }
    // division by a per-frame constant b whose reciprocal invb is precomputed
    BHK_INL static float divc(float a, float b, float) { return a / b; }
    BHK_INL static float rcp(float b) { return 1.0f / b; }
    BHK_INL static float sqrt_(float x) { return sqrtf(x); }
    BHK_INL static float rsq(float x) { return 1.0f / sqrtf(x); }
    BHK_INL static float sin_(float x) { return sinf(x); }
    BHK_INL static float cos_(float x) { return cosf(x); }
    BHK_INL static float exp_(float x) { return expf(x); }
    BHK_INL static float log_(float x) { return logf(x); }
    BHK_INL static float pow_(float x, float y) { return powf(x, y); }
    BHK_INL static float atan2_(float y, float x) { return atan2f(y, x); }
    BHK_INL static float acos_(float x) { return acosf(x); }
};

// atan2 for the fast policy: atan(a) = a P(a^2) on a = min/max in [0, 1]
// (degree-7 polynomial in a^2, least-squares fit with iterative minimax
// reweighting), then the octant and quadrant reflections of atan2.  With the
// hardware reciprocal and the float rounding of results near pi, the error
// is at most 3.3e-7 rad (measured over angles x magnitudes 1e-3..1e3 against
// long double; tests/test_gpu_math_bounds.py holds it to 3.5e-7).
BHK_INL float atan2Poly(float y, float x)
{
    const float ax = fabsf(x), ay = fabsf(y);
    const float mx = fmaxf(ax, ay), mn = fminf(ax, ay);
    float a = mn * __builtin_amdgcn_rcpf(mx);
    if (mx == 0.f) a = 0.f;
    const float s = a * a;
    float p = -4.054606427e-03f;
    p = p * s + 2.186309919e-02f;
    p = p * s - 5.591253191e-02f;
    p = p * s + 9.642212093e-02f;
    p = p * s - 1.390863508e-01f;
    p = p * s + 1.994656622e-01f;
    p = p * s - 3.332986236e-01f;
    p = p * s + 9.999993443e-01f;
    float r = a * p;
    if (ay > ax) r = 1.57079632679489662f - r;
    if (__builtin_signbit(x)) r = 3.14159265358979324f - r;
    return __builtin_copysignf(r, y);
}

struct MathFast
{
    static constexpr bool fast = true;
    BHK_INL static float rcp(float b) { return __builtin_amdgcn_rcpf(b); }
    BHK_INL static float div(float a, float b) { return a * __builtin_amdgcn_rcpf(b); }
    BHK_INL static float divc(float a, float, float invb) { return a * invb; }
    BHK_INL static float sqrt_(float x) { return __builtin_amdgcn_sqrtf(x); }
    BHK_INL static float rsq(float x) { return __builtin_amdgcn_rsqf(x); }
    BHK_INL static float sin_(float x) { return __sinf(x); }
    BHK_INL static float cos_(float x) { return __cosf(x); }
    BHK_INL static float exp_(float x) { return __expf(x); }
    // hardware log2 / exp2 (x > 0 or x == 0 at every call site)
    BHK_INL static float log_(float x) { return __builtin_amdgcn_logf(x) * 0.693147180559945309f; }
    BHK_INL static float pow_(float x, float y) { return __builtin_amdgcn_exp2f(y * __builtin_amdgcn_logf(x)); }
    BHK_INL static float atan2_(float y, float x) { return atan2Poly(y, x); }
    BHK_INL static float acos_(float x) { return acosf(x); }
};

// ------------------------------------------------------------------ vectors (vec_math.cuh:12-53)
struct v3 { float x, y, z; };
struct v2 { float x, y; };
BHK_INL v3 mk3(float x, float y, float z) { return v3{x, y, z}; }
BHK_INL v2 mk2(float x, float y) { return v2{x, y}; }
BHK_INL v3 operator+(v3 a, v3 b) { return mk3(a.x + b.x, a.y + b.y, a.z + b.z); }
BHK_INL v3 operator-(v3 a, v3 b) { return mk3(a.x - b.x, a.y - b.y, a.z - b.z); }
BHK_INL v3 operator*(v3 a, v3 b) { return mk3(a.x * b.x, a.y * b.y, a.z * b.z); }
BHK_INL v3 operator*(v3 a, float s) { return mk3(a.x * s, a.y * s, a.z * s); }
BHK_INL v3& operator+=(v3& a, v3 b) { a.x += b.x; a.y += b.y; a.z += b.z; return a; }
BHK_INL v2 operator+(v2 a, v2 b) { return mk2(a.x + b.x, a.y + b.y); }
BHK_INL float dot(v3 a, v3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
template <class M> BHK_INL float len3(v3 a) { return M::sqrt_(dot(a, a)); }
template <class M> BHK_INL v3 normalize(v3 a) { return a * M::rsq(fmaxf(dot(a, a), 1e-20f)); }
template <class M> BHK_INL v3 vdiv(v3 a, float s) { return a * M::rcp(s); }
BHK_INL v3 lerp3(v3 a, v3 b, float t) { return a + (b - a) * t; }
BHK_INL float clampf(float x, float lo, float hi) { return fminf(fmaxf(x, lo), hi); }
// x - floor(x).  The hardware fract equals it wherever x - floor(x) does not
// round up to 1, i.e. for every x outside (-2^-25, 0); the hash arguments
// are lattice coordinates times 0.1031/0.1030/0.0973, lattice sums scaled
// likewise, or non-negative, so none falls there.
BHK_INL float fract(float x)
{
#if defined(__HIP_DEVICE_COMPILE__)
    return __builtin_amdgcn_fractf(x);
#else
    return x - floorf(x);
#endif
}
BHK_INL float lerpf(float a, float b, float t) { return a + (b - a) * t; }
template <class M> BHK_INL float smoothstepf(float e0, float e1, float x)
{
    float t = clampf(M::div(x - e0, e1 - e0), 0.f, 1.f);
    return t * t * (3.f - 2.f * t);
}
template <class M> BHK_INL v3 pow3(v3 v, float e) { return mk3(M::pow_(v.x, e), M::pow_(v.y, e), M::pow_(v.z, e)); }
BHK_INL bool finite3(v3 v) { return isfinite(v.x) && isfinite(v.y) && isfinite(v.z); }

// ------------------------------------------------------------------ hashes (vec_math.cuh:56-98)
// "Hash without Sine" (D. Hoskins) as the reference evaluates it.
BHK_INL float hash12(v2 p)
{
#pragma clang fp contract(off)
    float ax = fract(p.x * 0.1031f), ay = fract(p.y * 0.1031f);
    float d = ax * (ay + 33.33f) + ay * (ax + 33.33f) + ax * (ax + 33.33f);
    return fract(((ax + d) + (ay + d)) * (ax + d));
}
BHK_INL v3 hash33(v3 p)
{
#pragma clang fp contract(off)
    float qx = fract(p.x * 0.1031f), qy = fract(p.y * 0.1030f), qz = fract(p.z * 0.0973f);
    float d = qx * (qy + 33.33f) + qy * (qx + 33.33f) + qz * (qz + 33.33f);
    qx = qx + d; qy = qy + d; qz = qz + d;
    return mk3(fract((qx + qy) * qz), fract((qx + qz) * qy), fract((qy + qz) * qx));
}
// 2-D value noise on the integer lattice with a cubic fade.
BHK_INL float vnoise(v2 p)
{
    float ix = floorf(p.x), iy = floorf(p.y);
    float fx = p.x - ix, fy = p.y - iy;
    float ux = fx * fx * (3.f - 2.f * fx), uy = fy * fy * (3.f - 2.f * fy);
    float a = hash12(mk2(ix, iy));
    float b = hash12(mk2(ix + 1.f, iy + 0.f));
    float c = hash12(mk2(ix + 0.f, iy + 1.f));
    float d = hash12(mk2(ix + 1.f, iy + 1.f));
    return lerpf(lerpf(a, b, ux), lerpf(c, d, ux), uy);
}
BHK_INL float fbm(v2 p)
{
    float v = 0.f, amp = 0.5f;
#pragma unroll
    for (int i = 0; i < 3; ++i)
    {
        v += amp * vnoise(p);
        p = mk2(p.x * 2.03f + 11.7f, p.y * 2.01f - 5.2f);
        amp *= 0.5f;
    }
    return v;
}

// ------------------------------------------------------------------ constants (metric.cuh:53-58, trace.cuh:66-70)
constexpr float kPi = 3.14159265358979f;
constexpr float kTInner = 13500.f;
constexpr float kNtPeakInv = 1.f / 0.075f;
constexpr float kSingCutOverM = 0.05f;

// ------------------------------------------------------------------ metric (metric.cuh:63-199)
// Omega = sqrt(M r - Q^2) / (r^2 + a sqrt(M r - Q^2)), 0 without a circular orbit.
template <class M> BHK_INL float omegaCircular(float r, float Mm, float a, float Q2)
{
    float w2 = Mm * r - Q2;
    if (w2 <= 0.f) return 0.f;
    float sw = M::sqrt_(w2);
    return M::div(sw, r * r + a * sw);
}
BHK_HD float omegaCircularHost(float r, float Mm, float a, float Q2)
{
    float w2 = Mm * r - Q2;
    if (w2 <= 0.f) return 0.f;
    float sw = sqrtf(w2);
    return sw / (r * r + a * sw);
}
BHK_HD float captureRadius(const RenderParams& P)
{
    return (P.rPlus > 1e-6f) ? P.rPlus : kSingCutOverM * P.M;
}

// ------------------------------------------------------------------ blackbody (trace.cuh:101-120)
template <class M> BHK_INL v3 blackbody(float kelvin)
{
    if constexpr (M::fast)
    {
        // The same fit with its 2.2 power taken in the log2 domain: above
        // 6600 K red and green are c (t - 60)^k clamped to 255, so
        // (min(c (t-60)^k, 255) / 255)^2.2 = exp2(2.2 min(log2(c/255) + k log2(t-60), 0))
        // (log2(329.698727446 / 255) and log2(288.1221695283 / 255) below)
        // with one log2 shared by both channels; the constant channels
        // (255 -> 1) need no power at all.  Equal to the direct form up to
        // the rounding of the hardware exp2/log2.
        const float t = clampf(kelvin, 1200.f, 40000.f) * 0.01f;
        float r, g, b;
        if (t <= 66.f)
        {
            r = 1.f;
            const float gl = clampf(99.4708025861f * M::log_(t) - 161.1195681661f, 0.f, 255.f) * (1.f / 255.f);
            g = __builtin_amdgcn_exp2f(2.2f * __builtin_amdgcn_logf(gl));
        }
        else
        {
            const float L = __builtin_amdgcn_logf(t - 60.f);
            r = __builtin_amdgcn_exp2f(2.2f * fminf(fmaf(-0.1332047592f, L, 0.370651066f), 0.f));
            g = __builtin_amdgcn_exp2f(2.2f * fminf(fmaf(-0.0755148492f, L, 0.176183432f), 0.f));
        }
        if (t >= 66.f) b = 1.f;
        else if (t <= 19.f) b = 0.f;
        else
        {
            const float bl = clampf(138.5177312231f * M::log_(t - 10.f) - 305.0447927307f, 0.f, 255.f) *
                             (1.f / 255.f);
            b = __builtin_amdgcn_exp2f(2.2f * __builtin_amdgcn_logf(bl));
        }
        return mk3(r, g, b);
    }
    float t = M::divc(clampf(kelvin, 1200.f, 40000.f), 100.f, 0.01f);
    float r, g, b;
    if (t <= 66.f) r = 255.f;
    else r = 329.698727446f * M::pow_(t - 60.f, -0.1332047592f);
    if (t <= 66.f) g = 99.4708025861f * M::log_(t) - 161.1195681661f;
    else g = 288.1221695283f * M::pow_(t - 60.f, -0.0755148492f);
    if (t >= 66.f) b = 255.f;
    else if (t <= 19.f) b = 0.f;
    else b = 138.5177312231f * M::log_(t - 10.f) - 305.0447927307f;
    constexpr float i255 = 1.f / 255.f;
    v3 srgb = mk3(M::divc(clampf(r, 0.f, 255.f), 255.f, i255), M::divc(clampf(g, 0.f, 255.f), 255.f, i255),
                  M::divc(clampf(b, 0.f, 255.f), 255.f, i255));
    return pow3<M>(srgb, 2.2f);
}

// ------------------------------------------------------------------ tetrad (trace.cuh:213-247)
// Photon conserved E, Lz for the static observer at `o` looking along `d`.
template <class M> BHK_INL void photonEL(v3 o, v3 d, float Mm, float a, float Q, float& E, float& Lz)
{
    float r0 = len3<M>(o);
    if (r0 < 1e-4f) { E = 1.f; Lz = 0.f; return; }
    float cth0 = clampf(M::div(o.y, r0), -1.f, 1.f);
    float th0 = M::acos_(cth0);
    float ph0 = M::atan2_(-o.z, o.x);
    float s = M::sin_(th0);
    float sp = M::sin_(ph0), cp = M::cos_(ph0);
    float dpc = d.x * (-sp) + d.y * 0.f + d.z * (-cp);
    float s0 = s < 1e-4f ? 1e-4f : s;
    float s02 = s0 * s0;
    float Sig0 = r0 * r0 + a * a * cth0 * cth0;
    float m2r0 = 2.f * Mm * r0 - Q * Q;
    float gtt = -(1.f - M::div(m2r0, fmaxf(Sig0, 1e-8f)));
    if (gtt > -1e-5f) gtt = -1e-5f;
    float gtp = M::div(-a * m2r0 * s02, Sig0);
    float gpp = (r0 * r0 + a * a + M::div(m2r0 * a * a * s02, Sig0)) * s02;
    float A = M::rsq(fmaxf(gpp - M::div(gtp * gtp, gtt), 1e-8f));
    float ptUp = M::rsq(-gtt) - M::div(dpc * A * gtp, gtt);
    float pfUp = dpc * A;
    E = -(gtt * ptUp + gtp * pfUp);
    Lz = gtp * ptUp + gpp * pfUp;
    if (!(E > 1e-6f) || !isfinite(E)) { E = 1.f; Lz = 0.f; }
}

// Wrap to (-pi, pi] (trace.cuh:250-255)
BHK_INL float wrapDeltaPhi(float d)
{
    d = fmodf(d + kPi, 2.f * kPi);
    if (d < 0.f) d += 2.f * kPi;
    return d - kPi;
}

// The same wrap by d - 2 pi floor((d + pi) / (2 pi)) with a fused
// multiply-add: off by a few ulp of d from the exact fmod for the |d| of
// tens of radians the hot spots see (they enter a Gaussian of width 0.18).
BHK_INL float wrapDeltaPhiFast(float d)
{
    const float e = d + kPi;
    const float q = floorf(e * 0.159154943091895f);
    float w = fmaf(-q, 2.f * kPi, e);
    w = w < 0.f ? w + 2.f * kPi : (w >= 2.f * kPi ? w - 2.f * kPi : w);
    return w - kPi;
}

// ------------------------------------------------------------------ disk (trace.cuh:269-403)
// Per-frame disk constants, computed once on the host with the same float
// operations the reference evaluates per sample (so they are bit-identical)
// and delivered as kernel arguments, i.e. scalar registers.
struct DiskParams
{
    float rin, rout, invOut, maxH, M, a, Q2, twoM, emisK, absK, spotS;
    float rinWall, routFade, rinSpot, rinPeak, sig2, invSig2;   // flux/density profile
    float rHs0, rHs1, om0, om1, inv2sR0, inv2sR1;               // hot spots
    int spots;
};
BHK_HD DiskParams makeDisk(const RenderParams& P, float a, float Q)
{
    DiskParams D;
    D.rin = fmaxf(P.diskInner, 1e-4f);
    D.rout = P.diskOuter;
    D.invOut = 1.f / fmaxf(P.diskOuter, D.rin);
    D.maxH = 0.16f * P.diskOuter;
    D.M = P.M;
    D.a = a;
    D.Q2 = Q * Q;
    D.twoM = 2.f * P.M;
    D.emisK = P.diskEmisScale;
    D.absK = P.diskAbsScale;
    D.spots = (P.hotSpotsEnabled && P.hotSpotStrength > 1e-4f) ? 1 : 0;
    D.spotS = P.hotSpotStrength;
    D.rinWall = D.rin * 1.10f;
    D.routFade = D.rout * 0.60f;
    D.rinSpot = D.rin * 1.60f;
    D.rinPeak = D.rin * 1.15f;
    const float sig = 0.20f * D.rin;
    D.sig2 = fmaxf(sig * sig, 1e-8f);
    D.invSig2 = 1.f / D.sig2;
    D.rHs0 = D.rin * 1.06f;
    D.rHs1 = D.rin * 1.18f;
    D.om0 = omegaCircularHost(D.rHs0, D.M, a, D.Q2);
    D.om1 = omegaCircularHost(D.rHs1, D.M, a, D.Q2);
    D.inv2sR0 = 1.f / (2.f * (0.09f * D.rHs0) * (0.09f * D.rHs0));
    D.inv2sR1 = 1.f / (2.f * (0.11f * D.rHs1) * (0.11f * D.rHs1));
    return D;
}

// g = 1/[u^t (E - Omega Lz)] of a circular equatorial emitter (metric.cuh:154-185)
template <class M> BHK_INL float emitterG(const DiskParams& D, float r, float E, float Lz, float Om)
{
    const float a = D.a;
    float r2 = r * r;
    float m2r = D.twoM * r - D.Q2;
    float gtt, gtp, gpp;
    if constexpr (M::fast)
    {
        const float ir2 = M::rcp(r2);
        gtt = -(1.f - m2r * ir2);
        gtp = -a * m2r * ir2;
        gpp = r2 + a * a + m2r * a * a * ir2;
    }
    else
    {
        gtt = -(1.f - M::div(m2r, r2));
        gtp = M::div(-a * m2r, r2);
        gpp = r2 + a * a + M::div(m2r * a * a, r2);
    }
    float den = -(gtt + 2.f * Om * gtp + Om * Om * gpp);
    if (den < 1e-4f) return 0.f;
    float ut = M::rcp(M::sqrt_(den));
    float Eem = ut * (E - Om * Lz);
    if (Eem < 0.04f) Eem = 0.04f;
    float g = M::rcp(Eem);
    if (g < 0.04f) g = 0.04f;
    if (g > 6.f) g = 6.f;
    return g;
}
// Novikov-Thorne flux weight with an inner peak (metric.cuh:190-199)
template <class M> BHK_INL float fluxWeight(const DiskParams& D, float r)
{
    const float rIn = D.rin;
    if (!(r > rIn) || rIn <= 0.f) return 0.f;
    float x = M::div(rIn, r);
    float nt = x * x * x * fmaxf(1.f - M::sqrt_(x), 0.f);
    float dr = r - D.rinPeak;
    float peak = M::exp_(M::divc(-0.5f * dr * dr, D.sig2, D.invSig2));
    return nt * (0.50f + 1.00f * peak);
}

// Radiance and optical depth of one volumetric sub-sample at p.  Returns
// false where the reference skips the sub-sample (`continue`).
template <class M>
BHK_INL bool diskSubSample(const DiskParams& D, float E, float Lz, v3 p, float tEmit, float ds,
                           v3& add, float& tau, float& densOut)
{
    float rc = M::sqrt_(p.x * p.x + p.z * p.z);
    if (rc < D.rin || rc > D.rout) return false;
    float HR = 0.040f + 0.015f * (rc * D.invOut);
    float H = HR * rc;
    float dz = M::div(p.y, fmaxf(H, 1e-4f));
    if (fabsf(dz) > 2.6f) return false;
    float Fr = fluxWeight<M>(D, rc) * kNtPeakInv;
    bool wantSpots = D.spots && rc < D.rinSpot;
    if (Fr < 1e-6f && !wantSpots) return false;

    float dens = M::exp_(-dz * dz);
    dens *= smoothstepf<M>(D.rin, D.rinWall, rc);
    dens *= 1.0f - smoothstepf<M>(D.routFade, D.rout, rc);

    float omega = omegaCircular<M>(rc, D.M, D.a, D.Q2);
    if (omega <= 0.f) return false;
    float g = emitterG<M>(D, rc, E, Lz, omega);
    if (g <= 0.f) return false;

    float azim = M::atan2_(-p.z, p.x);
    float azr = azim - omega * tEmit;
    float n0 = fbm(mk2(rc * 3.6f, azr * 2.3f + rc * 0.8f));
    float st = vnoise(mk2(rc * 21.f - azr * 1.7f, azr * 3.5f + rc * 2.2f));
    dens *= (0.20f + 1.30f * n0) * (0.40f + 1.00f * st);
    dens *= dens;
    float psi1 = azim - 0.55f * omega * tEmit - 1.6f * M::log_(fmaxf(rc, 0.2f));
    dens *= (1.f + 0.38f * M::cos_(psi1));

    float spotBoost = 0.f, spotHeat = 0.f;
    if (wantSpots)
    {
        const float rHs0 = D.rHs0, rHs1 = D.rHs1;
        const float om0 = D.om0 <= 0.f ? omega : D.om0;
        const float om1 = D.om1 <= 0.f ? omega : D.om1;
        const float inv2sPhi = 1.f / (2.f * 0.18f * 0.18f);
        const float inv2sR0 = D.inv2sR0, inv2sR1 = D.inv2sR1;
        float dph0, dph1;
        if constexpr (M::fast)
        {
            dph0 = wrapDeltaPhiFast(azim - (om0 * tEmit + 0.4f));
            dph1 = wrapDeltaPhiFast(azim - (om1 * tEmit + 0.4f + kPi));
        }
        else
        {
            dph0 = wrapDeltaPhi(azim - (om0 * tEmit + 0.4f));
            dph1 = wrapDeltaPhi(azim - (om1 * tEmit + 0.4f + kPi));
        }
        float dr0 = rc - rHs0;
        float dr1 = rc - rHs1;
        float flick0 = 0.72f + 0.28f * M::sin_(1.7f * tEmit + 0.9f);
        float flick1 = 0.65f + 0.35f * M::sin_(1.1f * tEmit + 2.3f);
        float g0 = M::exp_(-(dph0 * dph0) * inv2sPhi - dr0 * dr0 * inv2sR0 - dz * dz * 0.45f);
        float g1 = M::exp_(-(dph1 * dph1) * inv2sPhi - dr1 * dr1 * inv2sR1 - dz * dz * 0.45f);
        float S = D.spotS;
        spotBoost = S * (5.0f * flick0 * g0 + 3.2f * flick1 * g1);
        spotHeat = S * (0.70f * flick0 * g0 + 0.45f * flick1 * g1);
        dens += dens * 0.25f * (g0 + g1) * S;
    }
    if (dens < 1e-5f && spotBoost < 1e-4f) return false;

    float Temit = kTInner * M::pow_(clampf(Fr, 1e-6f, 1.f), 0.25f);
    Temit = clampf(Temit * (1.f + 1.6f * spotHeat), 2000.f, 26000.f);
    float Tobs = clampf(Temit * g, 1200.f, 40000.f);
    v3 col = blackbody<M>(Tobs);
    float g3 = g * g * g;
    float emis = Fr * g3 * g * (1.f + spotBoost);
    float w = dens * ds;
    add = col * (emis * D.emisK * w);
    if (spotBoost > 1e-4f)
    {
        float wS = (0.12f + dens) * ds;
        add += col * (spotBoost * g3 * D.emisK * 0.55f * wS);
    }
    tau = D.absK * w * (1.f + 0.35f * dens);
    densOut = dens;
    return true;
}

// True when a sub-sample at p passes the geometric gates of diskSubSample
// (radial range, |dz| <= 2.6 scale heights, flux or hot-spot region): only
// such sub-samples can change a ray's radiance or transmittance.
template <class M> BHK_INL bool diskSubSampleCanContribute(const DiskParams& D, v3 p)
{
    float rc = M::sqrt_(p.x * p.x + p.z * p.z);
    if (rc < D.rin || rc > D.rout) return false;
    float HR = 0.040f + 0.015f * (rc * D.invOut);
    float H = HR * rc;
    float dz = M::div(p.y, fmaxf(H, 1e-4f));
    if (fabsf(dz) > 2.6f) return false;
    float Fr = fluxWeight<M>(D, rc) * kNtPeakInv;
    bool wantSpots = D.spots && rc < D.rinSpot;
    return !(Fr < 1e-6f && !wantSpots);
}

// A disk chord [a, b] that survived the slab and radial culls
// (trace.cuh:274-291): n sub-samples of length ds at emission time tEmit.
struct Chord
{
    v3 a, b;
    float tEmit, ds;
    int n;
};

// Chord [a, b] with the cylindrical radii of its ends given (rcA, rcB):
// they enter only the culls below, which skip chords whose sub-samples all
// fail the gates of diskSubSample.
template <class M>
BHK_INL int chordPrepareRc(const DiskParams& D, v3 a, v3 b, float rcA, float rcB, float tEmit, Chord& c)
{
    c.n = 0;
    v3 seg = b - a;
    float segLen = len3<M>(seg);
    if (segLen < 1e-7f) return 0;
    if (a.y > D.maxH && b.y > D.maxH) return 0;
    if (a.y < -D.maxH && b.y < -D.maxH) return 0;
    float rcMin = fminf(rcA, rcB), rcMax = fmaxf(rcA, rcB);
    if (rcMin - segLen > D.rout) return 0;
    if (rcMax + segLen < D.rin) return 0;
    // Scale-height cull (not in the reference; it only skips chords whose
    // sub-samples all fail the |dz| <= 2.6 gate of diskSubSample, so the
    // result is unchanged).  With a and b on one side of the midplane every
    // point of the chord has |y| >= min(|a.y|, |b.y|); its cylindrical radius
    // is at most rcMax (a norm of an affine function is convex), and inside
    // [rin, rout] the scale height is H = (0.040 + 0.015 rc / rout) rc <=
    // 0.055 rc, floored at 1e-4.  The 2% margin covers float rounding.
    if ((a.y > 0.f) == (b.y > 0.f) &&
        fminf(fabsf(a.y), fabsf(b.y)) > 1.02f * 2.6f * fmaxf(0.055f * rcMax, 1e-4f))
        return 0;
    int nSub = 1 + (int)(segLen * 6.5f);
    if (nSub > 12) nSub = 12;
    c.a = a;
    c.b = b;
    c.tEmit = tEmit;
    c.ds = M::div(segLen, (float)nSub);
    c.n = nSub;
    return nSub;
}
template <class M> BHK_INL int chordPrepare(const DiskParams& D, v3 a, v3 b, float tEmit, Chord& c)
{
    if (a.y > D.maxH && b.y > D.maxH) { c.n = 0; return 0; }
    if (a.y < -D.maxH && b.y < -D.maxH) { c.n = 0; return 0; }
    return chordPrepareRc<M>(D, a, b, M::sqrt_(a.x * a.x + a.z * a.z), M::sqrt_(b.x * b.x + b.z * b.z), tEmit, c);
}

// Sub-sample j of a chord: position at t = (j + 1/2) / n along [a, b].
template <class M>
BHK_INL bool chordSubSample(const DiskParams& D, float E, float Lz, v3 a, v3 b, int j, int n, float tEmit,
                            float ds, v3& add, float& tau)
{
    float t = M::div((float)j + 0.5f, (float)n);
    v3 p = lerp3(a, b, t);
    float dens;
    return diskSubSample<M>(D, E, Lz, p, tEmit, ds, add, tau, dens);
}

// Emission-absorption update of one sub-sample; false once the chord's
// transmittance has dropped below the early-out level (trace.cuh:396-401).
template <class M> BHK_INL bool accumulateSample(v3 add, float tau, v3& accum, float& trans)
{
    accum += add * trans;
    trans *= M::exp_(-tau);
    return !(trans < 0.012f);
}

// One thread, one chord: the reference's sampleDiskSegment (trace.cuh:269-403).
template <class M>
BHK_INL void diskChordSerial(const DiskParams& D, float E, float Lz, const Chord& c, v3& accum, float& trans)
{
    for (int j = 0; j < c.n; ++j)
    {
        v3 add;
        float tau;
        if (!chordSubSample<M>(D, E, Lz, c.a, c.b, j, c.n, c.tEmit, c.ds, add, tau)) continue;
        if (!accumulateSample<M>(add, tau, accum, trans)) return;
    }
}

// Wavefront-cooperative chord sampling.  Every lane calls this (uniform
// control flow).  The sub-samples of all lanes' chords are numbered in lane
// order and spread over the 64 lanes, 64 per round; the owner of each chord
// then folds its sub-samples into its own radiance in sub-sample order.
// Each sub-sample is evaluated exactly as in diskChordSerial and folded in
// the same order, so the result is bit-identical -- only the wavefront's
// idle time changes (a lane with 12 sub-samples no longer stalls 63 others).
template <class M>
BHK_INL void diskChordWave(const DiskParams& D, bool has, const Chord& c, float E, float Lz, v3& accum, float& trans)
{
    if (__ballot(has) == 0ull) return;
    const int lane = (int)__lane_id();
    const int n = has ? c.n : 0;
    int incl = n;
#pragma unroll
    for (int d = 1; d < 64; d <<= 1)
    {
        int t = __shfl_up(incl, d);
        if (lane >= d) incl += t;
    }
    const int total = __shfl(incl, 63);
    const int start = incl - n;
    bool live = has;
    for (int base = 0; base < total; base += 64)
    {
        const int k = base + lane;
        // owner: first lane whose inclusive count exceeds k
        int lo = 0, hi = 63;
#pragma unroll
        for (int it = 0; it < 6; ++it)
        {
            int mid = (lo + hi) >> 1;
            int v = __shfl(incl, mid);
            if (v > k) hi = mid;
            else lo = mid + 1;
        }
        const int owner = lo;
        v3 oa = mk3(__shfl(c.a.x, owner), __shfl(c.a.y, owner), __shfl(c.a.z, owner));
        v3 ob = mk3(__shfl(c.b.x, owner), __shfl(c.b.y, owner), __shfl(c.b.z, owner));
        const float oT = __shfl(c.tEmit, owner), oDs = __shfl(c.ds, owner);
        const float oE = __shfl(E, owner), oL = __shfl(Lz, owner);
        const int oStart = __shfl(start, owner), oN = __shfl(n, owner);
        v3 add = mk3(0.f, 0.f, 0.f);
        float tau = 0.f;
        int valid = 0;
        if (k < total) valid = chordSubSample<M>(D, oE, oL, oa, ob, k - oStart, oN, oT, oDs, add, tau) ? 1 : 0;
        // owners fold this round's share of their sub-samples, in order
        const int first = start > base ? start : base;
        const int lastX = (start + n) < (base + 64) ? (start + n) : (base + 64);
        const int cnt = (has && lastX > first) ? lastX - first : 0;
        for (int t = 0;; ++t)
        {
            const bool want = t < cnt;
            if (__ballot(want) == 0ull) break;
            int src = want ? first + t - base : lane;
            float ax = __shfl(add.x, src), ay = __shfl(add.y, src), az = __shfl(add.z, src);
            float tu = __shfl(tau, src);
            int vv = __shfl(valid, src);
            if (want && live && vv) live = accumulateSample<M>(mk3(ax, ay, az), tu, accum, trans);
        }
    }
}

// ------------------------------------------------------------------ camera (trace.cuh:81-96, 431-438)
template <class M> BHK_INL v3 cameraRayDir(const RenderParams& P, float fx, float fy)
{
    float nx = fx * 2.f - 1.f;
    float ny = 1.f - fy * 2.f;
    v3 F = mk3(P.camForward.x, P.camForward.y, P.camForward.z);
    v3 R = mk3(P.camRight.x, P.camRight.y, P.camRight.z);
    v3 U = mk3(P.camUp.x, P.camUp.y, P.camUp.z);
    return normalize<M>(F + R * (nx * P.tanHalfFov * P.aspect) + U * (ny * P.tanHalfFov));
}

BHK_INL float cameraEnergyFactor(const RenderParams& P)
{
    v3 c = mk3(P.camPos.x, P.camPos.y, P.camPos.z);
    float r = sqrtf(dot(c, c));
    if (r < 1e-4f) return 1.f;
    if (P.model == BH_KERR || P.model == BH_KERR_NEWMAN)
    {
        float cth = clampf(c.y / r, -1.f, 1.f);
        float Sig = r * r + P.aSpin * P.aSpin * cth * cth;
        float m2r = 2.f * P.M * r - P.Qc * P.Qc;
        float gtt = -(1.f - m2r / fmaxf(Sig, 1e-8f));
        float alpha = sqrtf(fmaxf(-gtt, 1e-6f));
        return clampf(1.f / alpha, 0.5f, 4.f);
    }
    float f = 1.f - 2.f * P.M / r + P.Qc * P.Qc / (r * r);
    return clampf(1.f / sqrtf(fmaxf(f, 1e-6f)), 0.5f, 4.f);
}

// Sub-pixel jitter, R2 sequence (post_process.cuh:47-53)
BHK_INL v2 subpixelJitter(int accumMode, int sampleIndex)
{
#pragma clang fp contract(off)
    if (accumMode == 0 || sampleIndex <= 0) return mk2(0.5f, 0.5f);
    float n = (float)sampleIndex;
    return mk2(fract(0.7548776662f * n + 0.5f), fract(0.5698402909f * n + 0.5f));
}

// Sample position of a launch: the host-computed jitter, or, when the
// sample index lives in device memory (graph-replayed accumulation), the R2
// point of that index.
BHK_INL float2 launchJitter(const RenderParams& P, float2 hostJit, const int* dynIndex)
{
    if (!dynIndex) return hostJit;
    const v2 j = subpixelJitter(P.accumMode, *dynIndex);
    return make_float2(j.x, j.y);
}

// ------------------------------------------------------------------ sky (trace.cuh:131-195), exact arithmetic
// One star lattice (trace.cuh:131-165), in two parts: the probe (the
// direction's cell, its star offset and magnitude, and whether the direction
// lies inside the star's core) and the shading of a star that it hits.
BHK_INL bool starLayerProbe(v3 d, float cellScale, float coreSharp, float& d2o, float& mago, float& tselo)
{
#pragma clang fp contract(off)
    v3 p = d * cellScale;
    v3 id = mk3(floorf(p.x), floorf(p.y), floorf(p.z));
    v3 f = p - id;
    v3 h = hash33(id);
    v3 sp = mk3(0.10f + 0.80f * h.x, 0.10f + 0.80f * h.y, 0.10f + 0.80f * h.z);
    v3 dv = f - sp;
    float d2 = dot(dv, dv);
    float mag = hash12(mk2(id.x + 91.7f * id.z, id.y - 33.1f * id.x));
    float cut = (mag > 0.9f) ? 60.f : 14.f;
    d2o = d2;
    mago = mag;
    tselo = hash12(mk2(id.y + 7.3f, id.z - 3.1f));
    return !(d2 * coreSharp > cut);
}

BHK_INL v3 starShade(float d2, float mag, float tsel, float intensity, float coreSharp)
{
#pragma clang fp contract(off)
    float bright = 0.05f * powf(mag, 9.f) + 9.0f * powf(mag, 50.f);
    float core = expf(-d2 * coreSharp);
    float glow = expf(-d2 * coreSharp * 0.10f);
    float kelvin = 2600.f + 12000.f * tsel * tsel;
    v3 col = blackbody<MathExact>(kelvin);
    float lum = 0.2126f * col.x + 0.7152f * col.y + 0.0722f * col.z;
    col = col * (1.f / fmaxf(lum, 1e-4f));
    return col * ((core + 0.08f * glow) * bright * intensity);
}

BHK_INL v3 starLayer(v3 d, float cellScale, float intensity, float coreSharp)
{
    float d2, mag, tsel;
    if (!starLayerProbe(d, cellScale, coreSharp, d2, mag, tsel)) return mk3(0.f, 0.f, 0.f);
    return starShade(d2, mag, tsel, intensity, coreSharp);
}

// Milky-Way band, dust lanes and airglow added to the star sum, scaled by
// the observer's energy factor (trace.cuh:181-195).
BHK_INL v3 skyDiffuse(v3 c, v3 d, float gCam)
{
#pragma clang fp contract(off)
    float Ts = clampf(gCam, 0.5f, 3.f);
    float Is = Ts * Ts * Ts * Ts;
    float s = 3.0f / (1.35f + fabsf(d.y));
    v2 q = mk2(d.x * s, d.z * s);
    float neb = fbm(q);
    float dust = fbm(mk2(q.x * 2.6f + 31.4f, q.y * 2.6f - 17.2f));
    float band = expf(-d.y * d.y * 16.f);
    float lanes = clampf(1.25f - 1.35f * dust, 0.f, 1.f);
    c += (mk3(0.020f, 0.024f, 0.038f) + mk3(0.030f, 0.022f, 0.014f) * neb) * (neb * band * lanes);
    c += mk3(0.0012f, 0.0015f, 0.0026f);
    return c * Is;
}

BHK_INL v3 backgroundColor(v3 dir, float gCam)
{
#pragma clang fp contract(off)
    v3 d = normalize<MathExact>(dir);
    v3 c = starLayer(d, 80.f, 1.00f, 320.f);
    c += starLayer(d, 190.f, 0.55f, 520.f);
    c += starLayer(d, 420.f, 0.30f, 820.f);
    c += starLayer(d, 900.f, 0.14f, 1500.f);
    return skyDiffuse(c, d, gCam);
}

// backgroundColor for a whole wavefront (every lane calls it; lanes with
// `active` false get an unused value).  A direction hits a star's core in a
// few percent of the pixels per lattice, so the star shading, evaluated in
// place, would run four times per wavefront with most lanes idle.  Here the
// hits of all lanes and lattices are numbered in lane order and shaded 64 at
// a time, one per lane; each lane then adds its own hits in lattice order.
// Every value is formed by the same operations as in backgroundColor, and a
// missing star adds +0 there, so the result is bit-identical.
BHK_INL v3 backgroundColorWave(v3 dir, float gCam, bool active)
{
#pragma clang fp contract(off)
    v3 d = normalize<MathExact>(dir);
    float d2_0, d2_1, d2_2, d2_3, m_0, m_1, m_2, m_3, t_0, t_1, t_2, t_3;
    unsigned hits = 0u;   // bit L: the direction is inside a star core of lattice L
    if (starLayerProbe(d, 80.f, 320.f, d2_0, m_0, t_0)) hits |= 1u;
    if (starLayerProbe(d, 190.f, 520.f, d2_1, m_1, t_1)) hits |= 2u;
    if (starLayerProbe(d, 420.f, 820.f, d2_2, m_2, t_2)) hits |= 4u;
    if (starLayerProbe(d, 900.f, 1500.f, d2_3, m_3, t_3)) hits |= 8u;
    if (!active) hits = 0u;
    const int lane = (int)__lane_id();
    const int n = __popc(hits);
    int incl = n;
#pragma unroll
    for (int sh = 1; sh < 64; sh <<= 1)
    {
        const int t = __shfl_up(incl, sh);
        if (lane >= sh) incl += t;
    }
    const int total = __shfl(incl, 63);
    const int start = incl - n;
    v3 c = mk3(0.f, 0.f, 0.f);
    for (int base = 0; base < total; base += 64)
    {
        // hit k = base + lane: its owner is the first lane whose inclusive
        // count exceeds k, its lattice the owner's (k - start)-th set bit
        const int k = base + lane;
        int lo = 0, hi = 63;
#pragma unroll
        for (int it = 0; it < 6; ++it)
        {
            const int mid = (lo + hi) >> 1;
            if (__shfl(incl, mid) > k) hi = mid;
            else lo = mid + 1;
        }
        const int owner = lo;
        unsigned m = (unsigned)__shfl((int)hits, owner);
        int j = k - __shfl(start, owner);
        j = j < 0 ? 0 : (j > 3 ? 3 : j);
        for (int q = 0; q < j; ++q) m &= m - 1u;
        const int L = m ? __ffs((int)m) - 1 : 0;
        const float o0 = __shfl(d2_0, owner), o1 = __shfl(d2_1, owner), o2 = __shfl(d2_2, owner),
                    o3 = __shfl(d2_3, owner);
        const float g0 = __shfl(m_0, owner), g1 = __shfl(m_1, owner), g2 = __shfl(m_2, owner),
                    g3 = __shfl(m_3, owner);
        const float s0 = __shfl(t_0, owner), s1 = __shfl(t_1, owner), s2 = __shfl(t_2, owner),
                    s3 = __shfl(t_3, owner);
        v3 col = mk3(0.f, 0.f, 0.f);
        if (k < total)
        {
            const float hd2 = L == 0 ? o0 : L == 1 ? o1 : L == 2 ? o2 : o3;
            const float hm = L == 0 ? g0 : L == 1 ? g1 : L == 2 ? g2 : g3;
            const float ht = L == 0 ? s0 : L == 1 ? s1 : L == 2 ? s2 : s3;
            const float inten = L == 0 ? 1.00f : L == 1 ? 0.55f : L == 2 ? 0.30f : 0.14f;
            const float sharp = L == 0 ? 320.f : L == 1 ? 520.f : L == 2 ? 820.f : 1500.f;
            col = starShade(hd2, hm, ht, inten, sharp);
        }
        // each lane adds the hits of this round that are its own, in order
#pragma unroll
        for (int jj = 0; jj < 4; ++jj)
        {
            const int kk = start + jj;
            const bool mine = jj < n && kk >= base && kk < base + 64;
            const int src = mine ? kk - base : lane;
            const float x = __shfl(col.x, src), y = __shfl(col.y, src), z = __shfl(col.z, src);
            if (mine) c += mk3(x, y, z);
        }
    }
    return skyDiffuse(c, d, gCam);
}

// ------------------------------------------------------------------ post (post_process.cuh), exact arithmetic
BHK_INL v3 acesToneMap(v3 x)
{
#pragma clang fp contract(off)
    const float A = 2.51f, B = 0.03f, C = 2.43f, D = 0.59f, E = 0.14f;
    v3 num = x * (x * A + mk3(B, B, B));
    v3 den = x * (x * C + mk3(D, D, D)) + mk3(E, E, E);
    return mk3(clampf(num.x / den.x, 0.f, 1.f), clampf(num.y / den.y, 0.f, 1.f), clampf(num.z / den.z, 0.f, 1.f));
}
BHK_INL v3 accumBlend(v3 prev, v3 cur, int accumMode, int sampleIndex)
{
#pragma clang fp contract(off)
    if (accumMode == 0 || sampleIndex <= 0) return cur;
    if (accumMode == 1)
    {
        float n = (float)sampleIndex;
        return prev + (cur - prev) * (1.0f / (n + 1.0f));
    }
    return lerp3(prev, cur, 0.42f);
}
BHK_INL v3 bloomBrightPass(v3 c)
{
#pragma clang fp contract(off)
    float l = 0.2126f * c.x + 0.7152f * c.y + 0.0722f * c.z;
    float w = smoothstepf<MathExact>(0.95f, 2.2f, l);
    return c * w;
}
// (unsigned char)(clamp(powf(c, 1/2.2) + dd, 0, 1) * 255 + 0.5), with the
// gamma first taken from the hardware log2/exp2 (relative error near 1e-7,
// so y below moves by < 3e-4): wherever y lies farther than 1e-3 from an
// integer its truncation is the same as with powf, and only the other
// samples pay for powf.
BHK_INL unsigned char quantizeGamma(float c, float dd)
{
#pragma clang fp contract(off)
    constexpr float g = 1.0f / 2.2f;
    float y = clampf(__builtin_amdgcn_exp2f(__builtin_amdgcn_logf(c) * g) + dd, 0.f, 1.f) * 255.f + 0.5f;
    const float fr = y - floorf(y);
    if (fr < 1e-3f || fr > 1.f - 1e-3f) y = clampf(powf(c, g) + dd, 0.f, 1.f) * 255.f + 0.5f;
    return (unsigned char)y;
}

BHK_INL uchar4 finalizePixel(v3 hdr, v3 bloom, const RenderParams& P, int px, int py)
{
#pragma clang fp contract(off)
    v3 c = hdr * P.exposure;
    if (P.bloomEnabled) c += bloom * P.bloomStrength;
    c = acesToneMap(c);
    float lum = 0.2126f * c.x + 0.7152f * c.y + 0.0722f * c.z;
    v3 lv = mk3(lum, lum, lum);
    v3 t = lv + (c - lv) * 1.12f;
    c = mk3(clampf(t.x, 0.f, 1.f), clampf(t.y, 0.f, 1.f), clampf(t.z, 0.f, 1.f));
    v2 fp = mk2((float)px, (float)py);
    float d = (hash12(fp) + hash12(fp + mk2(41.3f, 17.7f))) - 1.0f;
    float dd = d * (1.0f / 255.0f);
    // gamma 1/2.2, dither, quantize
    unsigned char r8 = quantizeGamma(c.x, dd);
    unsigned char g8 = quantizeGamma(c.y, dd);
    unsigned char b8 = quantizeGamma(c.z, dd);
    return P.swapRB ? make_uchar4(b8, g8, r8, 255) : make_uchar4(r8, g8, b8, 255);
}

}  // namespace bhk
