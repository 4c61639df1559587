// ---------------------------------------------------------------------------
// oracle/bh_oracle.hpp -- CPU restatement of the reference tracer.
//
// TEST INFRASTRUCTURE ONLY.  This is the checker for the HIP product path in
// black-hole_amd/: only tests/, __graft_entry__.smoke() and bench.py's CPU
// baseline may load it.  Nothing in the product links or calls it.
//
// Written from the reference's equations and expression order so that, built
// with `g++ -O3 -ffp-contract=off` (x86-64 SSE, no FMA, IEEE division and
// square root, libm transcendentals), the float instantiation reproduces the
// reference's host-compiled arithmetic operation for operation.  Each
// function cites the reference lines it follows.
//
// The scalar type T is a template parameter:
//   float  -- reference semantics (the parity oracle and the CPU baseline),
//   double -- FP64 trace of the same algorithm (precision-error oracle); the
//             hash/noise shading is evaluated in float on float-rounded
//             positions so the FP64 image stays comparable,
//   CF     -- counted float: bit-identical to float, counts every +,-,*,/,
//             sqrt, floor and transcendental (FMA would count 2; none occur
//             here), which gives exact algorithmic flop counts.
//
// Pinning status of each part is documented next to the tests that pin it
// (tests/test_oracle_pins.py).  Parts whose exact bits only an import of the
// reference could pin (hash lattices, starfield, disk shading constants,
// dither) carry "parity unpinned" there.
// ---------------------------------------------------------------------------
#pragma once
#include <cmath>
#include <cstdint>
#include <type_traits>
#include "../include/bh/render_params.h"

namespace bho {

// ---------------------------------------------------------------- counted float
extern thread_local uint64_t tl_flops;

struct CF
{
    float v;
    CF() = default;
    constexpr CF(float x) : v(x) {}
    explicit operator float() const { return v; }
};
inline CF operator+(CF a, CF b) { ++tl_flops; return CF(a.v + b.v); }
inline CF operator-(CF a, CF b) { ++tl_flops; return CF(a.v - b.v); }
inline CF operator*(CF a, CF b) { ++tl_flops; return CF(a.v * b.v); }
inline CF operator/(CF a, CF b) { ++tl_flops; return CF(a.v / b.v); }
inline CF operator-(CF a) { return CF(-a.v); }
inline CF& operator+=(CF& a, CF b) { a = a + b; return a; }
inline CF& operator-=(CF& a, CF b) { a = a - b; return a; }
inline CF& operator*=(CF& a, CF b) { a = a * b; return a; }
inline bool operator<(CF a, CF b) { return a.v < b.v; }
inline bool operator>(CF a, CF b) { return a.v > b.v; }
inline bool operator<=(CF a, CF b) { return a.v <= b.v; }
inline bool operator>=(CF a, CF b) { return a.v >= b.v; }
inline bool operator==(CF a, CF b) { return a.v == b.v; }
inline bool operator!=(CF a, CF b) { return a.v != b.v; }

// ---------------------------------------------------------------- math wrappers
inline float  msqrt(float x)  { return sqrtf(x); }
inline double msqrt(double x) { return std::sqrt(x); }
inline CF     msqrt(CF x)     { ++tl_flops; return CF(sqrtf(x.v)); }
inline float  msin(float x)  { return sinf(x); }
inline double msin(double x) { return std::sin(x); }
inline CF     msin(CF x)     { ++tl_flops; return CF(sinf(x.v)); }
inline float  mcos(float x)  { return cosf(x); }
inline double mcos(double x) { return std::cos(x); }
inline CF     mcos(CF x)     { ++tl_flops; return CF(cosf(x.v)); }
inline float  mexp(float x)  { return expf(x); }
inline double mexp(double x) { return std::exp(x); }
inline CF     mexp(CF x)     { ++tl_flops; return CF(expf(x.v)); }
inline float  mlog(float x)  { return logf(x); }
inline double mlog(double x) { return std::log(x); }
inline CF     mlog(CF x)     { ++tl_flops; return CF(logf(x.v)); }
inline float  mpow(float x, float y)   { return powf(x, y); }
inline double mpow(double x, double y) { return std::pow(x, y); }
inline CF     mpow(CF x, CF y)         { ++tl_flops; return CF(powf(x.v, y.v)); }
inline float  matan2(float y, float x)   { return atan2f(y, x); }
inline double matan2(double y, double x) { return std::atan2(y, x); }
inline CF     matan2(CF y, CF x)         { ++tl_flops; return CF(atan2f(y.v, x.v)); }
inline float  macos(float x)  { return acosf(x); }
inline double macos(double x) { return std::acos(x); }
inline CF     macos(CF x)     { ++tl_flops; return CF(acosf(x.v)); }
inline float  mfmod(float x, float y)   { return fmodf(x, y); }
inline double mfmod(double x, double y) { return std::fmod(x, y); }
inline CF     mfmod(CF x, CF y)         { ++tl_flops; return CF(fmodf(x.v, y.v)); }
inline float  mfloor(float x)  { return floorf(x); }
inline double mfloor(double x) { return std::floor(x); }
inline CF     mfloor(CF x)     { ++tl_flops; return CF(floorf(x.v)); }
// min/max/abs/compare are free in the flop convention
inline float  mfabs(float x)  { return fabsf(x); }
inline double mfabs(double x) { return std::fabs(x); }
inline CF     mfabs(CF x)     { return CF(fabsf(x.v)); }
inline float  mfmin(float a, float b)   { return fminf(a, b); }
inline double mfmin(double a, double b) { return std::fmin(a, b); }
inline CF     mfmin(CF a, CF b)         { return CF(fminf(a.v, b.v)); }
inline float  mfmax(float a, float b)   { return fmaxf(a, b); }
inline double mfmax(double a, double b) { return std::fmax(a, b); }
inline CF     mfmax(CF a, CF b)         { return CF(fmaxf(a.v, b.v)); }
inline bool   mfinite(float x)  { return std::isfinite(x); }
inline bool   mfinite(double x) { return std::isfinite(x); }
inline bool   mfinite(CF x)     { return std::isfinite(x.v); }
inline float  tof(float x)  { return x; }
inline float  tof(double x) { return (float)x; }
inline float  tof(CF x)     { return x.v; }

template <class T> struct Id { using type = T; };
template <class T> using Id_t = typename Id<T>::type;

// Shading type: hash/noise/sky/disk radiative transfer run in float (or the
// counted float), also for the FP64 trace.
template <class T> using Shade = typename std::conditional<std::is_same<T, double>::value, float, T>::type;

// ---------------------------------------------------------------- vectors (vec_math.cuh:12-53)
template <class T> struct V3 { T x, y, z; };
template <class T> struct V2 { T x, y; };
template <class T> inline V3<T> mk3(T x, T y, T z) { return V3<T>{x, y, z}; }
template <class T> inline V2<T> mk2(T x, T y) { return V2<T>{x, y}; }
template <class T> inline V3<T> operator+(V3<T> a, V3<T> b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
template <class T> inline V3<T> operator-(V3<T> a, V3<T> b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
template <class T> inline V3<T> operator-(V3<T> a) { return {-a.x, -a.y, -a.z}; }
template <class T> inline V3<T> operator*(V3<T> a, V3<T> b) { return {a.x * b.x, a.y * b.y, a.z * b.z}; }
template <class T> inline V3<T> operator*(V3<T> a, Id_t<T> s) { return {a.x * s, a.y * s, a.z * s}; }
template <class T> inline V3<T> operator*(Id_t<T> s, V3<T> a) { return a * s; }
template <class T> inline V3<T> operator/(V3<T> a, Id_t<T> s) { return a * (T(1.0f) / s); }
template <class T> inline V3<T>& operator+=(V3<T>& a, V3<T> b) { a.x += b.x; a.y += b.y; a.z += b.z; return a; }
template <class T> inline T dot(V3<T> a, V3<T> b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
template <class T> inline T len3(V3<T> a) { return msqrt(dot(a, a)); }
template <class T> inline T rsq(T x) { return T(1.0f) / msqrt(x); }   // host rsqrtf := 1/sqrtf
template <class T> inline V3<T> normalize(V3<T> a) { return a * rsq(mfmax(dot(a, a), T(1e-20f))); }
template <class T> inline V3<T> lerp3(V3<T> a, V3<T> b, Id_t<T> t) { return a + (b - a) * t; }
template <class T> inline V2<T> operator+(V2<T> a, V2<T> b) { return {a.x + b.x, a.y + b.y}; }
template <class T> inline V2<T> operator*(V2<T> a, Id_t<T> s) { return {a.x * s, a.y * s}; }
template <class T> inline T clampf(T x, Id_t<T> lo, Id_t<T> hi) { return mfmin(mfmax(x, lo), hi); }
template <class T> inline T fract(T x) { return x - mfloor(x); }
template <class T> inline T lerpf(T a, T b, Id_t<T> t) { return a + (b - a) * t; }
template <class T> inline T smoothstepf(T e0, T e1, T x)
{
    T t = clampf((x - e0) / (e1 - e0), T(0.f), T(1.f));
    return t * t * (T(3.f) - T(2.f) * t);
}
template <class T> inline V3<T> clamp3(V3<T> v, float lo, float hi)
{
    return {clampf(v.x, T(lo), T(hi)), clampf(v.y, T(lo), T(hi)), clampf(v.z, T(lo), T(hi))};
}
template <class T> inline V3<T> pow3(V3<T> v, Id_t<T> e) { return {mpow(v.x, e), mpow(v.y, e), mpow(v.z, e)}; }
template <class A, class B> inline V3<A> cvt3(V3<B> v) { return {A(tof(v.x)), A(tof(v.y)), A(tof(v.z))}; }

// ---------------------------------------------------------------- hashing / noise (vec_math.cuh:56-98)
template <class S> inline S hash12(V2<S> p)
{
    V3<S> p3 = mk3(fract(p.x * S(0.1031f)), fract(p.y * S(0.1031f)), fract(p.x * S(0.1031f)));
    S d = p3.x * (p3.y + S(33.33f)) + p3.y * (p3.z + S(33.33f)) + p3.z * (p3.x + S(33.33f));
    p3 = mk3(p3.x + d, p3.y + d, p3.z + d);
    return fract((p3.x + p3.y) * p3.z);
}
template <class S> inline V3<S> hash33(V3<S> p)
{
    p = mk3(fract(p.x * S(0.1031f)), fract(p.y * S(0.1030f)), fract(p.z * S(0.0973f)));
    S d = p.x * (p.y + S(33.33f)) + p.y * (p.x + S(33.33f)) + p.z * (p.z + S(33.33f));
    p = mk3(p.x + d, p.y + d, p.z + d);
    return mk3(fract((p.x + p.y) * p.z), fract((p.x + p.z) * p.y), fract((p.y + p.z) * p.x));
}
template <class S> inline S vnoise(V2<S> p)
{
    V2<S> i = mk2(mfloor(p.x), mfloor(p.y));
    V2<S> f = mk2(p.x - i.x, p.y - i.y);
    V2<S> u = mk2(f.x * f.x * (S(3.f) - S(2.f) * f.x), f.y * f.y * (S(3.f) - S(2.f) * f.y));
    S a = hash12(i);
    S b = hash12(i + mk2(S(1.f), S(0.f)));
    S c = hash12(i + mk2(S(0.f), S(1.f)));
    S d = hash12(i + mk2(S(1.f), S(1.f)));
    return lerpf(lerpf(a, b, u.x), lerpf(c, d, u.x), u.y);
}
template <class S> inline S fbm(V2<S> p)
{
    S v = S(0.f), amp = S(0.5f);
    for (int i = 0; i < 3; ++i)
    {
        v = v + amp * vnoise(p);
        p = mk2(p.x * S(2.03f) + S(11.7f), p.y * S(2.01f) - S(5.2f));
        amp = amp * S(0.5f);
    }
    return v;
}

// ---------------------------------------------------------------- constants (metric.cuh:53-58, trace.cuh:66-70)
constexpr float EXTREMAL_LIMIT = 0.995f;
constexpr float NAKED_LIMIT    = 2.25f;
constexpr float SING_CUT_OVER_M = 0.05f;
constexpr float T_INNER = 13500.f;
constexpr float PI_F    = 3.14159265358979f;
constexpr float NT_PEAK_INV = 1.f / 0.075f;

// ---------------------------------------------------------------- metric (metric.cuh:63-199)
template <class T> inline T bhSigma(T r, T cth, T a) { return r * r + a * a * cth * cth; }
template <class T> inline T bhDelta(T r, T M, T a, T Q) { return r * r - T(2.f) * M * r + a * a + Q * Q; }
template <class T> inline T bhSphF(T r, T M, T Q) { return T(1.f) - T(2.f) * M / r + Q * Q / (r * r); }
template <class T> inline T bhOmegaCircular(T r, T M, T a, T Q)
{
    T w2 = M * r - Q * Q;
    if (w2 <= T(0.f)) return T(0.f);
    T sw = msqrt(w2);
    return sw / (r * r + a * sw);
}
template <class T> inline void bhEquatorialMetric(T r, T M, T a, T Q, T& gtt, T& gtp, T& gpp)
{
    T r2 = r * r;
    T m2r = T(2.f) * M * r - Q * Q;
    gtt = -(T(1.f) - m2r / r2);
    gtp = -a * m2r / r2;
    gpp = r2 + a * a + m2r * a * a / r2;
}
template <class T> inline T bhEmitterGFactor(T r, T M, T a, T Q, T E, T Lz)
{
    T Om = bhOmegaCircular(r, M, a, Q);
    if (Om <= T(0.f)) return T(0.f);
    T gtt, gtp, gpp;
    bhEquatorialMetric(r, M, a, Q, gtt, gtp, gpp);
    T den = -(gtt + T(2.f) * Om * gtp + Om * Om * gpp);
    if (den < T(1e-4f)) return T(0.f);
    T ut = T(1.f) / msqrt(den);
    T Eem = ut * (E - Om * Lz);
    if (Eem < T(0.04f)) Eem = T(0.04f);
    T g = T(1.f) / Eem;
    if (g < T(0.04f)) g = T(0.04f);
    if (g > T(6.f)) g = T(6.f);
    return g;
}
template <class T> inline T bhThinDiskFluxWeight(T r, T rIn)
{
    if (!(r > rIn) || rIn <= T(0.f)) return T(0.f);
    T x = rIn / r;
    T nt = x * x * x * mfmax(T(1.f) - msqrt(x), T(0.f));
    T dr = r - rIn * T(1.15f);
    T sig = T(0.20f) * rIn;
    T peak = mexp(T(-0.5f) * dr * dr / mfmax(sig * sig, T(1e-8f)));
    return nt * (T(0.50f) + T(1.00f) * peak);
}

// ---------------------------------------------------------------- parameters in T
template <class T> struct PT
{
    int width, height, model, maxSteps, diskEnabled, hotSpotsEnabled;
    T camPosX, camPosY, camPosZ;
    T M, aSpin, Qc, rPlus, rPhoton, dPhi, diskInner, diskOuter, diskTime;
    T diskEmisScale, diskAbsScale, hotSpotStrength;
    explicit PT(const RenderParams& P)
        : width(P.width), height(P.height), model(P.model), maxSteps(P.maxSteps),
          diskEnabled(P.diskEnabled), hotSpotsEnabled(P.hotSpotsEnabled),
          camPosX(T(P.camPos.x)), camPosY(T(P.camPos.y)), camPosZ(T(P.camPos.z)),
          M(T(P.M)), aSpin(T(P.aSpin)), Qc(T(P.Qc)), rPlus(T(P.rPlus)), rPhoton(T(P.rPhoton)),
          dPhi(T(P.dPhi)), diskInner(T(P.diskInner)), diskOuter(T(P.diskOuter)),
          diskTime(T(P.diskTime)), diskEmisScale(T(P.diskEmisScale)),
          diskAbsScale(T(P.diskAbsScale)), hotSpotStrength(T(P.hotSpotStrength)) {}
};

// Capture surface (trace.cuh:73-77)
template <class T> inline T bhCaptureRadius(const PT<T>& P)
{
    if (P.rPlus > T(1e-6f)) return P.rPlus;
    return T(SING_CUT_OVER_M) * P.M;
}

// Static-observer camera energy factor (trace.cuh:81-96)
template <class S> inline S cameraEnergyFactor(const RenderParams& Pr)
{
    PT<S> P(Pr);
    V3<S> cp = mk3(P.camPosX, P.camPosY, P.camPosZ);
    S r = len3(cp);
    if (r < S(1e-4f)) return S(1.f);
    if (P.model == BH_KERR || P.model == BH_KERR_NEWMAN)
    {
        S cth = clampf(cp.y / r, S(-1.f), S(1.f));
        S Sig = bhSigma(r, cth, P.aSpin);
        S m2r = S(2.f) * P.M * r - P.Qc * P.Qc;
        S gtt = -(S(1.f) - m2r / mfmax(Sig, S(1e-8f)));
        S alpha = msqrt(mfmax(-gtt, S(1e-6f)));
        return clampf(S(1.f) / alpha, S(0.5f), S(4.f));
    }
    S f = bhSphF(r, P.M, P.Qc);
    return clampf(S(1.f) / msqrt(mfmax(f, S(1e-6f))), S(0.5f), S(4.f));
}

// Blackbody colour, Tanner Helland fit (trace.cuh:101-120)
template <class S> inline V3<S> blackbodyRGB(S kelvin)
{
    S t = clampf(kelvin, S(1200.f), S(40000.f)) / S(100.f);
    S r, g, b;
    if (t <= S(66.f)) r = S(255.f);
    else r = S(329.698727446f) * mpow(t - S(60.f), S(-0.1332047592f));
    if (t <= S(66.f)) g = S(99.4708025861f) * mlog(t) - S(161.1195681661f);
    else g = S(288.1221695283f) * mpow(t - S(60.f), S(-0.0755148492f));
    if (t >= S(66.f)) b = S(255.f);
    else if (t <= S(19.f)) b = S(0.f);
    else b = S(138.5177312231f) * mlog(t - S(10.f)) - S(305.0447927307f);
    V3<S> srgb = mk3(clampf(r, S(0.f), S(255.f)) / S(255.f), clampf(g, S(0.f), S(255.f)) / S(255.f),
                     clampf(b, S(0.f), S(255.f)) / S(255.f));
    return pow3(srgb, S(2.2f));
}

// One star lattice (trace.cuh:131-162)
template <class S> inline V3<S> starLayer(V3<S> d, float cellScale, float intensity, float coreSharp)
{
    V3<S> p = d * S(cellScale);
    V3<S> id = mk3(mfloor(p.x), mfloor(p.y), mfloor(p.z));
    V3<S> f = mk3(p.x - id.x, p.y - id.y, p.z - id.z);
    V3<S> h = hash33(id);
    V3<S> sp = mk3(S(0.10f) + S(0.80f) * h.x, S(0.10f) + S(0.80f) * h.y, S(0.10f) + S(0.80f) * h.z);
    V3<S> dv = f - sp;
    S d2 = dot(dv, dv);
    S mag = hash12(mk2(id.x + S(91.7f) * id.z, id.y - S(33.1f) * id.x));
    S cut = (mag > S(0.9f)) ? S(60.f) : S(14.f);
    if (d2 * S(coreSharp) > cut) return mk3(S(0.f), S(0.f), S(0.f));
    S bright = S(0.05f) * mpow(mag, S(9.f)) + S(9.0f) * mpow(mag, S(50.f));
    S core = mexp(-d2 * S(coreSharp));
    S glow = mexp(-d2 * S(coreSharp) * S(0.10f));
    S tsel = hash12(mk2(id.y + S(7.3f), id.z - S(3.1f)));
    S kelvin = S(2600.f) + S(12000.f) * tsel * tsel;
    V3<S> col = blackbodyRGB(kelvin);
    S lum = S(0.2126f) * col.x + S(0.7152f) * col.y + S(0.0722f) * col.z;
    col = col * (S(1.f) / mfmax(lum, S(1e-4f)));
    return col * ((core + S(0.08f) * glow) * bright * S(intensity));
}

// Sky: four star lattices + Milky-Way band + airglow (trace.cuh:166-195)
template <class S> inline V3<S> backgroundColor(V3<S> dir, S gCam)
{
    V3<S> d = normalize(dir);
    S Tscale = clampf(gCam, S(0.5f), S(3.f));
    S Iscale = Tscale * Tscale * Tscale * Tscale;
    V3<S> c = mk3(S(0.f), S(0.f), S(0.f));
    c += starLayer(d, 80.f, 1.00f, 320.f);
    c += starLayer(d, 190.f, 0.55f, 520.f);
    c += starLayer(d, 420.f, 0.30f, 820.f);
    c += starLayer(d, 900.f, 0.14f, 1500.f);
    V2<S> q = mk2(d.x, d.z) * (S(3.0f) / (S(1.35f) + mfabs(d.y)));
    S neb = fbm(q);
    S dust = fbm(mk2(q.x * S(2.6f) + S(31.4f), q.y * S(2.6f) - S(17.2f)));
    S band = mexp(-d.y * d.y * S(16.f));
    S lanes = clampf(S(1.25f) - S(1.35f) * dust, S(0.f), S(1.f));
    c += (mk3(S(0.020f), S(0.024f), S(0.038f)) + mk3(S(0.030f), S(0.022f), S(0.014f)) * neb) *
         (neb * band * lanes);
    c += mk3(S(0.0012f), S(0.0015f), S(0.0026f));
    return c * Iscale;
}

// Photon conserved (E, Lz) from the static-observer tetrad (trace.cuh:213-247)
template <class T> inline void photonConservedEL(V3<T> origin, V3<T> dir, T M, T a, T Q, T& E, T& Lz)
{
    T r0 = len3(origin);
    if (r0 < T(1e-4f)) { E = T(1.f); Lz = T(0.f); return; }
    T cth0 = clampf(origin.y / r0, T(-1.f), T(1.f));
    T th0 = macos(cth0);
    T ph0 = matan2(-origin.z, origin.x);
    T s = msin(th0);
    T sp = msin(ph0), cp = mcos(ph0);
    V3<T> eph = mk3(-sp, T(0.f), -cp);
    T dpc = dot(dir, eph);
    T s0 = s;
    if (s0 < T(1e-4f)) s0 = T(1e-4f);
    T s02 = s0 * s0;
    T Sig0 = bhSigma(r0, cth0, a);
    T m2r0 = T(2.f) * M * r0 - Q * Q;
    T gtt = -(T(1.f) - m2r0 / mfmax(Sig0, T(1e-8f)));
    if (gtt > T(-1e-5f)) gtt = T(-1e-5f);
    T gtp = -a * m2r0 * s02 / Sig0;
    T gpp = (r0 * r0 + a * a + m2r0 * a * a * s02 / Sig0) * s02;
    T A = rsq(mfmax(gpp - gtp * gtp / gtt, T(1e-8f)));
    T ptUp = rsq(-gtt) - dpc * A * gtp / gtt;
    T pfUp = dpc * A;
    E = -(gtt * ptUp + gtp * pfUp);
    Lz = gtp * ptUp + gpp * pfUp;
    if (!(E > T(1e-6f)) || !mfinite(E)) { E = T(1.f); Lz = T(0.f); }
}

// Wrap to (-pi, pi] (trace.cuh:250-255)
template <class S> inline S wrapDeltaPhi(S d)
{
    d = mfmod(d + S(PI_F), S(2.f * PI_F));
    if (d < S(0.f)) d = d + S(2.f * PI_F);
    return d - S(PI_F);
}

template <class S> struct DiskCtx { S E, Lz, a, Q; };

// Volumetric disk on one chord [a, b] (trace.cuh:269-403)
template <class S>
inline void sampleDiskSegment(const PT<S>& P, const DiskCtx<S>& ctx, V3<S> a, V3<S> b, S tEmit,
                              V3<S>& accum, S& trans)
{
    V3<S> seg = b - a;
    S segLen = len3(seg);
    if (segLen < S(1e-7f)) return;
    S maxH = S(0.16f) * P.diskOuter;
    if (a.y > maxH && b.y > maxH) return;
    if (a.y < -maxH && b.y < -maxH) return;
    S rcA = msqrt(a.x * a.x + a.z * a.z);
    S rcB = msqrt(b.x * b.x + b.z * b.z);
    S rcMin = mfmin(rcA, rcB), rcMax = mfmax(rcA, rcB);
    if (rcMin - segLen > P.diskOuter) return;
    if (rcMax + segLen < P.diskInner) return;

    int nSub = 1 + (int)tof(segLen * S(6.5f));
    if (nSub > 12) nSub = 12;
    S ds = segLen / S((float)nSub);
    const S emisK = P.diskEmisScale;
    const S absK = P.diskAbsScale;
    const S rin = mfmax(P.diskInner, S(1e-4f));
    const S invOut = S(1.f) / mfmax(P.diskOuter, rin);
    const S M = P.M, aSpin = ctx.a, Qc = ctx.Q;

    for (int j = 0; j < nSub; ++j)
    {
        S t = (S((float)j) + S(0.5f)) / S((float)nSub);
        V3<S> p = lerp3(a, b, t);
        S rc = msqrt(p.x * p.x + p.z * p.z);
        if (rc < rin || rc > P.diskOuter) continue;
        S HR = S(0.040f) + S(0.015f) * (rc * invOut);
        S H = HR * rc;
        S dz = p.y / mfmax(H, S(1e-4f));
        if (mfabs(dz) > S(2.6f)) continue;
        S Fr = bhThinDiskFluxWeight(rc, rin) * S(NT_PEAK_INV);
        bool wantSpots = P.hotSpotsEnabled && P.hotSpotStrength > S(1e-4f) && rc < rin * S(1.60f);
        if (Fr < S(1e-6f) && !wantSpots) continue;

        S dens = mexp(-dz * dz);
        dens = dens * smoothstepf(rin, rin * S(1.10f), rc);
        dens = dens * (S(1.0f) - smoothstepf(P.diskOuter * S(0.60f), P.diskOuter, rc));

        S omega = bhOmegaCircular(rc, M, aSpin, Qc);
        if (omega <= S(0.f)) continue;
        S g = bhEmitterGFactor(rc, M, aSpin, Qc, ctx.E, ctx.Lz);
        if (g <= S(0.f)) continue;

        S azim = matan2(-p.z, p.x);
        S azr = azim - omega * tEmit;
        S n0 = fbm(mk2(rc * S(3.6f), azr * S(2.3f) + rc * S(0.8f)));
        S st = vnoise(mk2(rc * S(21.f) - azr * S(1.7f), azr * S(3.5f) + rc * S(2.2f)));
        dens = dens * ((S(0.20f) + S(1.30f) * n0) * (S(0.40f) + S(1.00f) * st));
        dens = dens * dens;
        S psi1 = azim - S(0.55f) * omega * tEmit - S(1.6f) * mlog(mfmax(rc, S(0.2f)));
        dens = dens * (S(1.f) + S(0.38f) * mcos(psi1));

        S spotBoost = S(0.f), spotHeat = S(0.f);
        if (wantSpots)
        {
            const S rHs0 = rin * S(1.06f);
            const S rHs1 = rin * S(1.18f);
            S om0 = bhOmegaCircular(rHs0, M, aSpin, Qc);
            S om1 = bhOmegaCircular(rHs1, M, aSpin, Qc);
            if (om0 <= S(0.f)) om0 = omega;
            if (om1 <= S(0.f)) om1 = omega;
            const S inv2sPhi = S(1.f / (2.f * 0.18f * 0.18f));
            const S inv2sR0 = S(1.f) / (S(2.f) * (S(0.09f) * rHs0) * (S(0.09f) * rHs0));
            const S inv2sR1 = S(1.f) / (S(2.f) * (S(0.11f) * rHs1) * (S(0.11f) * rHs1));
            S dph0 = wrapDeltaPhi(azim - (om0 * tEmit + S(0.4f)));
            S dph1 = wrapDeltaPhi(azim - (om1 * tEmit + S(0.4f) + S(PI_F)));
            S dr0 = rc - rHs0;
            S dr1 = rc - rHs1;
            S flick0 = S(0.72f) + S(0.28f) * msin(S(1.7f) * tEmit + S(0.9f));
            S flick1 = S(0.65f) + S(0.35f) * msin(S(1.1f) * tEmit + S(2.3f));
            S g0 = mexp(-(dph0 * dph0) * inv2sPhi - dr0 * dr0 * inv2sR0 - dz * dz * S(0.45f));
            S g1 = mexp(-(dph1 * dph1) * inv2sPhi - dr1 * dr1 * inv2sR1 - dz * dz * S(0.45f));
            S Sx = P.hotSpotStrength;
            spotBoost = Sx * (S(5.0f) * flick0 * g0 + S(3.2f) * flick1 * g1);
            spotHeat = Sx * (S(0.70f) * flick0 * g0 + S(0.45f) * flick1 * g1);
            dens = dens + dens * S(0.25f) * (g0 + g1) * Sx;
        }
        if (dens < S(1e-5f) && spotBoost < S(1e-4f)) continue;

        S Temit = S(T_INNER) * mpow(clampf(Fr, S(1e-6f), S(1.f)), S(0.25f));
        Temit = clampf(Temit * (S(1.f) + S(1.6f) * spotHeat), S(2000.f), S(26000.f));
        S Tobs = clampf(Temit * g, S(1200.f), S(40000.f));
        V3<S> col = blackbodyRGB(Tobs);
        S g3 = g * g * g;
        S emis = Fr * g3 * g * (S(1.f) + spotBoost);
        S w = dens * ds;
        V3<S> add = col * (emis * emisK * w);
        if (spotBoost > S(1e-4f))
        {
            S wS = (S(0.12f) + dens) * ds;
            add += col * (spotBoost * g3 * emisK * S(0.55f) * wS);
        }
        accum += add * trans;
        S tau = absK * w * (S(1.f) + S(0.35f) * dens);
        trans = trans * mexp(-tau);
        if (trans < S(0.012f)) return;
    }
}

// ---------------------------------------------------------------- trace result (trace.cuh:418-428)
template <class T> struct TraceResult
{
    bool horizon, escaped;
    V3<T> escDir;
    V3<Shade<T>> accum;
    Shade<T> trans;
    T maxHviol;
    int steps;
};

// Camera ray (trace.cuh:431-438)
template <class T> inline V3<T> cameraRayDir(const RenderParams& P, T fx, T fy)
{
    T nx = fx * T(2.f) - T(1.f);
    T ny = T(1.f) - fy * T(2.f);
    V3<T> F = mk3(T(P.camForward.x), T(P.camForward.y), T(P.camForward.z));
    V3<T> R = mk3(T(P.camRight.x), T(P.camRight.y), T(P.camRight.z));
    V3<T> U = mk3(T(P.camUp.x), T(P.camUp.y), T(P.camUp.z));
    return normalize(F + R * (nx * T(P.tanHalfFov) * T(P.aspect)) + U * (ny * T(P.tanHalfFov)));
}

// Disk chord in the shading type.
template <class T>
inline void diskChord(const PT<Shade<T>>& PS, const DiskCtx<Shade<T>>& ctx, V3<T> a, V3<T> b, T tEmit,
                      V3<Shade<T>>& accum, Shade<T>& trans)
{
    using S = Shade<T>;
    sampleDiskSegment<S>(PS, ctx, cvt3<S>(a), cvt3<S>(b), S(tof(tEmit)), accum, trans);
}

// Spherical (Binet) integrator (trace.cuh:466-611)
// Final integrator state (tests): spherical u, u', phi; Kerr r, th, ph, p_r,
// p_th, E, Lz.
template <class T> struct FinalState { T v[8]; };

template <class T>
inline TraceResult<T> traceSpherical(const RenderParams& Pr, V3<T> origin, V3<T> dir, FinalState<T>* fin = nullptr)
{
    using S = Shade<T>;
    PT<T> P(Pr);
    PT<S> PS(Pr);
    TraceResult<T> R;
    R.horizon = false; R.escaped = false;
    R.escDir = dir;
    R.accum = mk3(S(0.f), S(0.f), S(0.f));
    R.trans = S(1.0f);
    R.maxHviol = T(0.f);
    R.steps = 0;

    const T M = P.M;
    const T Q2 = P.Qc * P.Qc;
    const T rCap = bhCaptureRadius(P);
    const T uHor = T(1.0f) / mfmax(rCap, T(1e-6f));
    const T escR = T(120.f) * M;
    const T uEscape = T(1.0f) / escR;

    T Ephot = T(1.f), Lzphot = T(0.f);
    photonConservedEL(origin, dir, M, T(0.f), P.Qc, Ephot, Lzphot);
    const DiskCtx<S> ctx = {S(tof(Ephot)), S(tof(Lzphot)), S(0.f), PS.Qc};

    V3<T> c = origin;
    T r0 = len3(c);
    V3<T> e1 = c / r0;
    T ddr = dot(dir, e1);
    V3<T> perp = dir - e1 * ddr;
    T pl = len3(perp);
    if (pl < T(1e-6f))
    {
        if (dot(dir, c) < T(0.f)) R.horizon = true;
        R.escDir = dir;
        R.escaped = !R.horizon;
        return R;
    }
    V3<T> e2 = perp / pl;
    T f0 = mfmax(bhSphF(r0, M, P.Qc), T(1e-6f));
    T bImpact = mfmax(r0 * pl / msqrt(f0), T(1e-6f));
    T u = T(1.0f) / r0;
    T du = -msqrt(f0) * ddr / (r0 * pl);
    T phi = T(0.f);
    T tFlight = T(0.f);
    V3<T> prevPos = c;
    V3<T> dir3 = dir;
    const T invNineM = T(1.f) / (T(9.f) * M);
    T h = P.dPhi;
    auto acc = [M, Q2](T uu) { return T(3.f) * M * uu * uu - T(2.f) * Q2 * uu * uu * uu - uu; };

    int step = 0;
    for (; step < P.maxSteps; ++step)
    {
        T hMax = P.dPhi * clampf(invNineM / u, T(1.f), T(6.f));
        h = mfmin(hMax, T(0.10f) / mfmax(mfabs(du), T(1e-6f)));
        if (h < P.dPhi * T(0.03f)) h = P.dPhi * T(0.03f);
        T k1u = du;
        T k1v = acc(u);
        T u2 = u + T(0.5f) * h * k1u;
        T v2 = du + T(0.5f) * h * k1v;
        T k2u = v2;
        T k2v = acc(u2);
        T u3 = u + T(0.5f) * h * k2u;
        T v3 = du + T(0.5f) * h * k2v;
        T k3u = v3;
        T k3v = acc(u3);
        T u4 = u + h * k3u;
        T v4 = du + h * k3v;
        T k4u = v4;
        T k4v = acc(u4);
        u = u + (h / T(6.f)) * (k1u + T(2.f) * k2u + T(2.f) * k3u + k4u);
        du = du + (h / T(6.f)) * (k1v + T(2.f) * k2v + T(2.f) * k3v + k4v);
        phi = phi + h;
        if (!mfinite(u) || !mfinite(du)) { R.horizon = true; break; }
        if (u >= uHor) { R.horizon = true; break; }
        if (u < T(1e-4f)) { R.escaped = true; R.escDir = dir3; break; }
        T r = T(1.0f) / u;
        T cp = mcos(phi), sp = msin(phi);
        V3<T> m = e1 * cp + e2 * sp;
        V3<T> mp = e1 * (-sp) + e2 * cp;
        T drdphi = -du / (u * u);
        V3<T> pos = m * r;
        dir3 = normalize(m * drdphi + mp * r);
        T fR = mfmax(bhSphF(r, M, P.Qc), T(1e-4f));
        tFlight = tFlight + h * (r * r) / (bImpact * fR);
        if (P.diskEnabled && R.trans > S(0.01f))
            diskChord<T>(PS, ctx, prevPos, pos, P.diskTime - tFlight, R.accum, R.trans);
        prevPos = pos;
        if (u < uEscape && du < T(0.f)) { R.escaped = true; R.escDir = dir3; break; }
    }
    R.steps = step;
    if (fin) { fin->v[0] = u; fin->v[1] = du; fin->v[2] = phi; fin->v[3] = bImpact; }
    if (!R.horizon && !R.escaped)
    {
        if (u > T(1.0f) / mfmax(T(1.5f) * P.rPhoton, T(3.f) * M)) R.horizon = true;
        else { R.escaped = true; R.escDir = dir3; }
    }
    return R;
}

// Kerr-Newman Hamiltonian integrator (trace.cuh:616-890)
template <class T> struct KState { T r, th, ph, pr, pth; };

template <class T> inline KState<T> kerrRHS(T M, T a, T Q, T E, T Lz, const KState<T>& y, T& Kout)
{
    T s = msin(y.th), cth = mcos(y.th);
    if (s < T(1e-4f)) s = T(1e-4f);
    T s2 = s * s;
    T r = y.r;
    T r2 = r * r, a2 = a * a;
    T Sig = r2 + a2 * cth * cth;
    T Del = r2 - T(2.f) * M * r + a2 + Q * Q;
    if (Del < T(1e-6f)) Del = T(1e-6f);
    T Pp = E * (r2 + a2) - a * Lz;
    T B = Lz / s - a * E * s;
    T W = B * B - Pp * Pp / Del;
    T K = Del * y.pr * y.pr + y.pth * y.pth + W;
    Kout = K;
    T dDel = T(2.f) * r - T(2.f) * M;
    T dP = T(2.f) * r * E;
    T Wr = -(T(2.f) * Pp * dP * Del - Pp * Pp * dDel) / (Del * Del);
    T dB = -Lz * cth / s2 - a * E * cth;
    T Wth = T(2.f) * B * dB;
    T invS = T(1.f) / Sig;
    T invS2 = invS * invS;
    KState<T> d;
    d.r = Del * y.pr * invS;
    d.th = y.pth * invS;
    d.ph = ((Lz / s2 - a * E) + a * Pp / Del) * invS;
    d.pr = T(-0.5f) * invS * (dDel * y.pr * y.pr + Wr) + r * invS2 * K;
    d.pth = T(-0.5f) * invS * Wth - a2 * s * cth * invS2 * K;
    return d;
}

template <class T> inline V3<T> kerrToCart(T r, T th, T ph)
{
    T s = msin(th), c = mcos(th);
    T sp = msin(ph), cp = mcos(ph);
    return mk3(r * s * cp, r * c, -r * s * sp);
}

template <class T> inline void kerrBasis(T th, T ph, V3<T>& er, V3<T>& eth, V3<T>& eph)
{
    T s = msin(th), c = mcos(th);
    T sp = msin(ph), cp = mcos(ph);
    er = mk3(s * cp, c, -s * sp);
    eth = mk3(c * cp, -s, -c * sp);
    eph = mk3(-sp, T(0.f), -cp);
}

// Optional per-ray observer: called after every accepted RK4 step with the
// new state (used by tests for Carter-constant and constraint checks).
template <class T> struct NoObserver { void operator()(const KState<T>&, T, T) const {} };

template <class T, class Obs = NoObserver<T>>
inline TraceResult<T> traceKerr(const RenderParams& Pr, V3<T> origin, V3<T> dir, bool wantStats,
                                const Obs& obs = Obs(), FinalState<T>* fin = nullptr)
{
    using S = Shade<T>;
    PT<T> P(Pr);
    PT<S> PS(Pr);
    TraceResult<T> R;
    R.horizon = false; R.escaped = false;
    R.escDir = dir;
    R.accum = mk3(S(0.f), S(0.f), S(0.f));
    R.trans = S(1.0f);
    R.maxHviol = T(0.f);
    R.steps = 0;

    const T M = P.M, a = P.aSpin, Q = P.Qc;
    const T rp = P.rPlus;
    const T rCap = bhCaptureRadius(P);
    const T rFloor = mfmax(rp, rCap);
    const T escR = T(120.f) * M;

    T r0 = len3(origin);
    T cth0 = clampf(origin.y / r0, T(-1.f), T(1.f));
    T th0 = macos(cth0);
    T ph0 = matan2(-origin.z, origin.x);
    V3<T> er, eth, eph;
    kerrBasis(th0, ph0, er, eth, eph);
    T drc = dot(dir, er);
    T dtc = dot(dir, eth);
    T dpc = dot(dir, eph);

    T s0 = msin(th0);
    if (s0 < T(1e-4f)) s0 = T(1e-4f);
    T s02 = s0 * s0;
    T Sig0 = bhSigma(r0, cth0, a);
    T Del0 = bhDelta(r0, M, a, Q);
    T m2r0 = T(2.f) * M * r0 - Q * Q;
    T gtt = -(T(1.f) - m2r0 / Sig0);
    if (gtt > T(-1e-5f)) gtt = T(-1e-5f);
    T gtp = -a * m2r0 * s02 / Sig0;
    T gpp = (r0 * r0 + a * a + m2r0 * a * a * s02 / Sig0) * s02;
    T A = rsq(mfmax(gpp - gtp * gtp / gtt, T(1e-8f)));
    T ptUp = rsq(-gtt) - dpc * A * gtp / gtt;
    T pfUp = dpc * A;
    T E = -(gtt * ptUp + gtp * pfUp);
    T Lz = gtp * ptUp + gpp * pfUp;

    KState<T> y;
    y.r = r0;
    y.th = th0;
    y.ph = ph0;
    y.pr = drc * msqrt(Sig0 / mfmax(Del0, T(1e-6f)));
    y.pth = dtc * msqrt(Sig0);

    const DiskCtx<S> ctx = {S(tof(E)), S(tof(Lz)), S(tof(a)), S(tof(Q))};
    const T E2 = E * E;
    const bool trackPos = (P.diskEnabled != 0);
    T tFlight = T(0.f);
    V3<T> prevPos = origin;
    V3<T> dir3 = dir;
    KState<T> dLast{};
    bool haveDeriv = false;
    int step = 0;

    for (; step < P.maxSteps; ++step)
    {
        T Kv;
        KState<T> d1 = kerrRHS(M, a, Q, E, Lz, y, Kv);
        if (!mfinite(y.r) || !mfinite(y.pr) || !mfinite(y.pth) || !mfinite(d1.r) || !mfinite(d1.pr))
        { R.horizon = true; break; }
        if (y.r <= rCap * T(1.002f) + T(1e-3f)) { R.horizon = true; break; }
        if (wantStats && y.r > rFloor * T(1.05f))
        {
            T viol = mfabs(Kv) / (E2 * (y.r * y.r + a * a) + T(1e-12f));
            if (viol > R.maxHviol) R.maxHviol = viol;
        }
        if (y.r >= escR && d1.r > T(0.f))
        {
            V3<T> e1b, e2b, e3b;
            kerrBasis(y.th, y.ph, e1b, e2b, e3b);
            T s = msin(y.th);
            R.escDir = normalize(e1b * d1.r + e2b * (y.r * d1.th) + e3b * (y.r * s * d1.ph));
            R.escaped = true;
            break;
        }
        T h = T(1.0f) / (mfabs(d1.r) / (T(8.0f) * P.dPhi * mfmax(y.r - rFloor, T(0.02f))) +
                         (mfabs(d1.th) + mfabs(d1.ph)) / (T(2.2f) * P.dPhi) + T(1e-5f));
        if (y.r < T(1.5f) * P.rPhoton) h = h * T(0.55f);
        T spd = mfabs(d1.r) + y.r * (mfabs(d1.th) + mfabs(d1.ph)) + T(1e-6f);
        T capLen;
        if (trackPos && y.r < P.diskOuter * T(1.3f))
        {
            T yAbs = y.r * mfabs(mcos(y.th));
            capLen = (yAbs < T(1.8f)) ? T(1.2f) : T(2.2f);
        }
        else capLen = mfmax(T(1.2f), T(0.22f) * y.r);
        h = mfmin(h, capLen / spd);

        T kv2, kv3, kv4;
        KState<T> y2 = {y.r + T(0.5f) * h * d1.r, y.th + T(0.5f) * h * d1.th, y.ph + T(0.5f) * h * d1.ph,
                        y.pr + T(0.5f) * h * d1.pr, y.pth + T(0.5f) * h * d1.pth};
        KState<T> d2 = kerrRHS(M, a, Q, E, Lz, y2, kv2);
        KState<T> y3 = {y.r + T(0.5f) * h * d2.r, y.th + T(0.5f) * h * d2.th, y.ph + T(0.5f) * h * d2.ph,
                        y.pr + T(0.5f) * h * d2.pr, y.pth + T(0.5f) * h * d2.pth};
        KState<T> d3 = kerrRHS(M, a, Q, E, Lz, y3, kv3);
        KState<T> y4 = {y.r + h * d3.r, y.th + h * d3.th, y.ph + h * d3.ph, y.pr + h * d3.pr,
                        y.pth + h * d3.pth};
        KState<T> d4 = kerrRHS(M, a, Q, E, Lz, y4, kv4);
        T h6 = h / T(6.f);
        y.r = y.r + h6 * (d1.r + T(2.f) * d2.r + T(2.f) * d3.r + d4.r);
        y.th = y.th + h6 * (d1.th + T(2.f) * d2.th + T(2.f) * d3.th + d4.th);
        y.ph = y.ph + h6 * (d1.ph + T(2.f) * d2.ph + T(2.f) * d3.ph + d4.ph);
        y.pr = y.pr + h6 * (d1.pr + T(2.f) * d2.pr + T(2.f) * d3.pr + d4.pr);
        y.pth = y.pth + h6 * (d1.pth + T(2.f) * d2.pth + T(2.f) * d3.pth + d4.pth);
        if (y.th < T(1e-3f)) { y.th = T(1e-3f); y.pth = mfabs(y.pth); }
        else if (y.th > T(PI_F) - T(1e-3f)) { y.th = T(PI_F) - T(1e-3f); y.pth = -mfabs(y.pth); }
        dLast = d4;
        haveDeriv = true;
        obs(y, E, Lz);
        {
            T s = msin(y.th), cth = mcos(y.th);
            if (s < T(1e-4f)) s = T(1e-4f);
            T Sig = bhSigma(y.r, cth, a);
            T Del = mfmax(bhDelta(y.r, M, a, Q), T(1e-5f));
            T Pp = E * (y.r * y.r + a * a) - a * Lz;
            T dtdl = (a * (Lz - a * E * s * s) + (y.r * y.r + a * a) * Pp / Del) / mfmax(Sig, T(1e-8f));
            tFlight = tFlight + h * mfabs(dtdl);
        }
        if (trackPos)
        {
            V3<T> pos = kerrToCart(y.r, y.th, y.ph);
            V3<T> stepv = pos - prevPos;
            T sl = len3(stepv);
            if (sl > T(1e-6f)) dir3 = stepv * (T(1.f) / sl);
            if (R.trans > S(0.01f)) diskChord<T>(PS, ctx, prevPos, pos, P.diskTime - tFlight, R.accum, R.trans);
            prevPos = pos;
        }
    }
    R.steps = step;
    if (fin)
    {
        fin->v[0] = y.r; fin->v[1] = y.th; fin->v[2] = y.ph; fin->v[3] = y.pr;
        fin->v[4] = y.pth; fin->v[5] = E; fin->v[6] = Lz; fin->v[7] = T(0.f);
    }
    if (!R.horizon && !R.escaped)
    {
        if (y.r < T(6.f) * M) R.horizon = true;
        else
        {
            R.escaped = true;
            if (haveDeriv)
            {
                V3<T> e1b, e2b, e3b;
                kerrBasis(y.th, y.ph, e1b, e2b, e3b);
                T s = msin(y.th);
                R.escDir = normalize(e1b * dLast.r + e2b * (y.r * dLast.th) + e3b * (y.r * s * dLast.ph));
            }
            else R.escDir = dir3;
        }
    }
    return R;
}

// Dispatch (trace.cuh:895-902)
template <class T> inline TraceResult<T> traceRay(const RenderParams& P, V3<T> origin, V3<T> dir, bool wantStats)
{
    if (P.model == BH_KERR || P.model == BH_KERR_NEWMAN) return traceKerr<T>(P, origin, dir, wantStats);
    return traceSpherical<T>(P, origin, dir);
}

// One HDR sample (trace.cuh:908-926).  Returns the trace result through *tr
// when non-null.
template <class T>
inline V3<Shade<T>> renderSampleHDR(const RenderParams& P, int px, int py, T jx, T jy, TraceResult<T>* tr = nullptr)
{
    using S = Shade<T>;
    T fx = (T((float)px) + jx) / T((float)P.width);
    T fy = (T((float)py) + jy) / T((float)P.height);
    V3<T> dir = cameraRayDir<T>(P, fx, fy);
    V3<T> cam = mk3(T(P.camPos.x), T(P.camPos.y), T(P.camPos.z));
    TraceResult<T> t = traceRay<T>(P, cam, dir, false);
    if (tr) *tr = t;
    S gCam = cameraEnergyFactor<S>(P);
    V3<S> bg = t.horizon ? mk3(S(0.f), S(0.f), S(0.f)) : backgroundColor<S>(cvt3<S>(t.escDir), gCam);
    V3<S> hdr = t.accum + bg * t.trans;
    if (!mfinite(hdr.x) || !mfinite(hdr.y) || !mfinite(hdr.z)) hdr = mk3(S(0.f), S(0.f), S(0.f));
    return hdr;
}

// ---------------------------------------------------------------- FP64 critical band (no reference counterpart)
// Restates the library's band test: a ray is traced in double when its FP64
// invariants put it near the critical curve, |min_r R(r)/(r^2+a^2)^2| <
// crit over r in [r+ (1+1e-3), max(5M, 1.5 r_lo)] (radial potential of
// Kerr-Newman at unit energy, Bardeen 1973), or, for the Kerr models, when
// the polar turning point has sin^2(theta) < polar.  Spherical models use
// the closed form 1 - b^2/b_c^2, b_c^2 = r_ph^2 / f(r_ph).
struct BandCfg { bool on; double crit, polar; };

// The widths cfg.crit / cfg.polar apply to steps dPhi <= 0.007; coarser
// steps widen the critical band by s = dPhi / 0.007 and the polar band by s^2.
inline double bandWidthScale(const RenderParams& Pr) { return std::max(1.0, (double)Pr.dPhi / 0.007); }

struct BandInv { double E, Lz, carter; bool ok; };

inline BandInv bandInvariants(const RenderParams& Pr, V3<double> o, V3<double> dir)
{
    const bool kerr = (Pr.model == BH_KERR || Pr.model == BH_KERR_NEWMAN);
    const double M = Pr.M, a = kerr ? (double)Pr.aSpin : 0.0, Q = Pr.Qc;
    const double r0 = len3(o);
    const double cth0 = clampf(o.y / r0, -1.0, 1.0);
    const double th0 = std::acos(cth0), ph0 = std::atan2(-o.z, o.x);
    V3<double> er, eth, eph;
    kerrBasis(th0, ph0, er, eth, eph);
    const double dtc = dot(dir, eth), dpc = dot(dir, eph);
    double s0 = std::sin(th0);
    if (s0 < (double)1e-4f) s0 = (double)1e-4f;
    const double s02 = s0 * s0;
    const double Sig0 = r0 * r0 + a * a * cth0 * cth0;
    const double m2r0 = 2.0 * M * r0 - Q * Q;
    double gtt = -(1.0 - m2r0 / Sig0);
    if (gtt > (double)-1e-5f) gtt = (double)-1e-5f;
    const double gtp = -a * m2r0 * s02 / Sig0;
    const double gpp = (r0 * r0 + a * a + m2r0 * a * a * s02 / Sig0) * s02;
    const double A = 1.0 / std::sqrt(std::max(gpp - gtp * gtp / gtt, (double)1e-8f));
    const double ptUp = 1.0 / std::sqrt(-gtt) - dpc * A * gtp / gtt;
    const double pfUp = dpc * A;
    BandInv I;
    I.E = -(gtt * ptUp + gtp * pfUp);
    I.Lz = gtp * ptUp + gpp * pfUp;
    const double pth = dtc * std::sqrt(Sig0);
    I.carter = pth * pth + cth0 * cth0 * (I.Lz * I.Lz / s02 - a * a * I.E * I.E);
    I.ok = (I.E > 1e-12) && std::isfinite(I.E) && std::isfinite(I.Lz);
    return I;
}

inline double bandRadialPot(double M, double a, double Q, double xi, double eta, double r)
{
    const double ra = r * r + a * a;
    const double P = ra - a * xi;
    const double Del = r * r - 2.0 * M * r + a * a + Q * Q;
    const double xa = xi - a;
    return (P * P - Del * (eta + xa * xa)) / (ra * ra);
}

// min of the normalised radial potential: 32 samples, then 24 golden-section
// steps on the bracket of the smallest sample.
inline double bandCriticalDelta(const RenderParams& Pr, double xi, double eta)
{
    const double M = Pr.M, a = Pr.aSpin, Q = Pr.Qc;
    const double disc = M * M - a * a - Q * Q;
    const double rp = disc >= 0.0 ? M + std::sqrt(disc) : 0.0;
    const double lo = rp > 0.0 ? rp * 1.001 : (double)SING_CUT_OVER_M * M;
    const double hi = std::max(5.0 * M, lo * 1.5);
    const int N = 32;
    const double dr = (hi - lo) / (double)(N - 1);
    int best = 0;
    double vBest = bandRadialPot(M, a, Q, xi, eta, lo);
    for (int i = 1; i < N; ++i)
    {
        double v = bandRadialPot(M, a, Q, xi, eta, lo + dr * (double)i);
        if (v < vBest) { vBest = v; best = i; }
    }
    double x0 = lo + dr * (double)(best > 0 ? best - 1 : 0);
    double x3 = lo + dr * (double)(best < N - 1 ? best + 1 : N - 1);
    const double gr = 0.6180339887498949;
    double x1 = x3 - gr * (x3 - x0), x2 = x0 + gr * (x3 - x0);
    double f1 = bandRadialPot(M, a, Q, xi, eta, x1), f2 = bandRadialPot(M, a, Q, xi, eta, x2);
    for (int it = 0; it < 24; ++it)
    {
        if (f1 < f2) { x3 = x2; x2 = x1; f2 = f1; x1 = x3 - gr * (x3 - x0); f1 = bandRadialPot(M, a, Q, xi, eta, x1); }
        else { x0 = x1; x1 = x2; f1 = f2; x2 = x0 + gr * (x3 - x0); f2 = bandRadialPot(M, a, Q, xi, eta, x2); }
    }
    return std::min(vBest, std::min(f1, f2));
}

// sin^2 of the polar turning angle: larger root u = cos^2 theta of
// a^2 u^2 - (a^2 - eta - xi^2) u - eta = 0.
inline double bandPolarSin2(double a, double xi, double eta)
{
    const double a2 = a * a, B = a2 - eta - xi * xi;
    const double D = std::sqrt(std::max(B * B + 4.0 * a2 * eta, 0.0));
    double u;
    if (B <= 0.0) u = (D - B > 0.0) ? 2.0 * eta / (D - B) : 1.0;
    else u = (a2 > 0.0) ? (B + D) / (2.0 * a2) : 1.0;
    return 1.0 - u;
}

inline bool rayInBand(const RenderParams& Pr, V3<double> o, V3<double> dir, const BandCfg& cfg, double* deltaOut = nullptr,
                      double* polarOut = nullptr)
{
    if (Pr.model == BH_KERR || Pr.model == BH_KERR_NEWMAN)
    {
        BandInv I = bandInvariants(Pr, o, dir);
        if (!I.ok) return true;
        const double xi = I.Lz / I.E, eta = I.carter / (I.E * I.E);
        const double d = bandCriticalDelta(Pr, xi, eta), pol = bandPolarSin2(Pr.aSpin, xi, eta);
        if (deltaOut) *deltaOut = d;
        if (polarOut) *polarOut = pol;
        const double sc = bandWidthScale(Pr), ad = std::fabs(d);
        // the polar band narrows away from the critical curve: the full
        // width below |delta| = 0.2 s, then 0.3, 0.1 and 0.03 of it from
        // 0.2 s, 0.5 s and s
        const double pw = cfg.polar * sc * sc;
        const double polW = ad < 0.2 * sc ? pw : ad < 0.5 * sc ? 0.3 * pw : ad < 1.0 * sc ? 0.1 * pw : 0.03 * pw;
        // rays that fall in (d > 0) have a critical band a quarter as wide;
        // escaping rays whose polar turning point is far from the axis
        // (sin^2 theta_min >= 0.1 s^2, >= 0.3 s^2) a half, a quarter
        const double cw = d > 0.0 ? 0.25 : pol >= 0.3 * sc * sc ? 0.25 : pol >= 0.1 * sc * sc ? 0.5 : 1.0;
        return ad < cw * cfg.crit * sc || pol < polW;
    }
    const double M = Pr.M, Q2 = (double)Pr.Qc * (double)Pr.Qc;
    const double disc = 9.0 * M * M - 8.0 * Q2;
    double bc2 = 0.0;
    if (disc >= 0.0)
    {
        const double rph = 0.5 * (3.0 * M + std::sqrt(disc));
        const double f = 1.0 - 2.0 * M / rph + Q2 / (rph * rph);
        if (f > 0.0) bc2 = rph * rph / f;
    }
    const double r0 = len3(o);
    V3<double> e1 = o / r0;
    V3<double> perp = dir - e1 * dot(dir, e1);
    const double pl = len3(perp);
    if (pl < (double)1e-6f || !(bc2 > 0.0)) return false;
    const double f0 = std::max(1.0 - 2.0 * M / r0 + Q2 / (r0 * r0), (double)1e-6f);
    const double b2 = (r0 * pl) * (r0 * pl) / f0;
    const double d = 1.0 - b2 / bc2;
    if (deltaOut) *deltaOut = d;
    return std::fabs(d) < (d > 0.0 ? 0.25 : 1.0) * cfg.crit * bandWidthScale(Pr);
}

// Camera ray of a pixel in double (trace.cuh:910-912).
inline V3<double> cameraDirD(const RenderParams& P, int px, int py, double jx, double jy)
{
    double fx = (double(float(px)) + jx) / double(float(P.width));
    double fy = (double(float(py)) + jy) / double(float(P.height));
    return cameraRayDir<double>(P, fx, fy);
}

// ---------------------------------------------------------------- post process (post_process.cuh)
template <class S> inline V3<S> acesToneMap(V3<S> x)
{
    const S A = S(2.51f), B = S(0.03f), C = S(2.43f), D = S(0.59f), E = S(0.14f);
    V3<S> num = x * (x * A + mk3(B, B, B));
    V3<S> den = x * (x * C + mk3(D, D, D)) + mk3(E, E, E);
    return clamp3(mk3(num.x / den.x, num.y / den.y, num.z / den.z), 0.f, 1.f);
}
inline V2<float> subpixelJitter(int accumMode, int sampleIndex)
{
    if (accumMode == 0 || sampleIndex <= 0) return mk2(0.5f, 0.5f);
    float n = (float)sampleIndex;
    return mk2(fract(0.7548776662f * n + 0.5f), fract(0.5698402909f * n + 0.5f));
}
template <class S> inline V3<S> accumBlend(V3<S> prev, V3<S> cur, int accumMode, int sampleIndex)
{
    if (accumMode == 0 || sampleIndex <= 0) return cur;
    if (accumMode == 1)
    {
        S n = S((float)sampleIndex);
        return prev + (cur - prev) * (S(1.0f) / (n + S(1.0f)));
    }
    return lerp3(prev, cur, S(0.42f));
}
template <class S> inline V3<S> bloomBrightPass(V3<S> c)
{
    S l = S(0.2126f) * c.x + S(0.7152f) * c.y + S(0.0722f) * c.z;
    S w = smoothstepf(S(0.95f), S(2.2f), l);
    return c * w;
}
inline float bloomGaussW(int i)
{
    static const float w[5] = {0.20236f, 0.17925f, 0.12458f, 0.06794f, 0.02907f};
    int k = i < 0 ? -i : i;
    return w[k > 4 ? 4 : k];
}
template <class S> inline V3<S> fetchHalf(const float* buf, int x, int y, int bw, int bh)
{
    if (x < 0) x = 0;
    if (x >= bw) x = bw - 1;
    if (y < 0) y = 0;
    if (y >= bh) y = bh - 1;
    const float* v = buf + 4 * ((size_t)y * bw + x);
    return mk3(S(v[0]), S(v[1]), S(v[2]));
}
template <class S> inline V3<S> sampleHalfBilinear(const float* buf, S x, S y, int bw, int bh)
{
    x = x - S(0.5f);
    y = y - S(0.5f);
    int x0 = (int)tof(mfloor(x)), y0 = (int)tof(mfloor(y));
    S tx = x - S((float)x0), ty = y - S((float)y0);
    V3<S> c00 = fetchHalf<S>(buf, x0, y0, bw, bh);
    V3<S> c10 = fetchHalf<S>(buf, x0 + 1, y0, bw, bh);
    V3<S> c01 = fetchHalf<S>(buf, x0, y0 + 1, bw, bh);
    V3<S> c11 = fetchHalf<S>(buf, x0 + 1, y0 + 1, bw, bh);
    return lerp3(lerp3(c00, c10, tx), lerp3(c01, c11, tx), ty);
}
// Final pixel (post_process.cuh:122-148); rgba written to out[0..3]
template <class S> inline void finalizePixel(V3<S> hdr, V3<S> bloom, const RenderParams& P, int px, int py,
                                             unsigned char* out)
{
    V3<S> c = hdr * S(P.exposure);
    if (P.bloomEnabled) c += bloom * S(P.bloomStrength);
    c = acesToneMap(c);
    S lum = S(0.2126f) * c.x + S(0.7152f) * c.y + S(0.0722f) * c.z;
    V3<S> lv = mk3(lum, lum, lum);
    c = clamp3(lv + (c - lv) * S(1.12f), 0.f, 1.f);
    c = pow3(c, S(1.0f / 2.2f));
    V2<S> fp = mk2(S((float)px), S((float)py));
    S d = (hash12(fp) + hash12(fp + mk2(S(41.3f), S(17.7f)))) - S(1.0f);
    S dd = d * S(1.0f / 255.0f);
    unsigned char r8 = (unsigned char)tof(clampf(c.x + dd, S(0.f), S(1.f)) * S(255.f) + S(0.5f));
    unsigned char g8 = (unsigned char)tof(clampf(c.y + dd, S(0.f), S(1.f)) * S(255.f) + S(0.5f));
    unsigned char b8 = (unsigned char)tof(clampf(c.z + dd, S(0.f), S(1.f)) * S(255.f) + S(0.5f));
    if (P.swapRB) { out[0] = b8; out[1] = g8; out[2] = r8; }
    else { out[0] = r8; out[1] = g8; out[2] = b8; }
    out[3] = 255;
}

}  // namespace bho
