// ---------------------------------------------------------------------------
// bh_surrogate_geom.h -- geometry shared by the surrogate's training set,
// its inference and its tests: network inputs of a camera ray and the
// exit-direction parametrisation.
//
// A ray leaving a static camera at Boyer-Lindquist (r_c, th_c) is described
// by psi, the angle between the ray and the direction to the hole, and
// alpha, its position angle on the image plane measured from the projected
// spin axis:  dir = -cos(psi) e_r + sin(psi) (cos(alpha) (-e_th) + sin(alpha) e_ph).
// Its screen-space impact parameter is b = r_c sin(psi).
//
// The exit direction is expressed in the ray's own frame: d0 = initial
// direction, e_a = unit component of the direction to the hole orthogonal
// to d0, e_b = d0 x e_a;  exit = cos(D) d0 + sin(D) (cos(beta) e_a +
// sin(beta) e_b).  The network predicts (D cos(beta) - D0, D sin(beta)),
// with D0 the deflection of the same ray in the spherical metric of the same
// M and Q up to the integrators' exit radius (sphDeflection), so it learns
// only the spin terms.
// ---------------------------------------------------------------------------
#pragma once
#include "bh_common.h"

namespace bhk {

constexpr float kSurBMax = 40.f;        // b / kSurBMax
constexpr float kSurRMin = 6.f;         // camera radius range (rs), log-scaled
constexpr float kSurRMax = 40.f;

struct RayFrame
{
    v3 d0, ea, eb;
    float psi, alpha, b;
};

// er, eth, eph: BL basis at the camera (trace.cuh:673-682 form).
BHK_HD void camBasis(v3 o, v3& er, v3& eth, v3& eph, float& r, float& th)
{
    r = sqrtf(o.x * o.x + o.y * o.y + o.z * o.z);
    const float c = fminf(fmaxf(o.y / r, -1.f), 1.f);
    th = acosf(c);
    const float ph = atan2f(-o.z, o.x);
    const float s = sinf(th), sp = sinf(ph), cp = cosf(ph);
    er = v3{s * cp, c, -s * sp};
    eth = v3{c * cp, -s, -c * sp};
    eph = v3{-sp, 0.f, -cp};
}

BHK_HD v3 cross3(v3 a, v3 b) { return v3{a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x}; }
BHK_HD float dot3(v3 a, v3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
BHK_HD v3 scale3(v3 a, float s) { return v3{a.x * s, a.y * s, a.z * s}; }
BHK_HD v3 add3(v3 a, v3 b) { return v3{a.x + b.x, a.y + b.y, a.z + b.z}; }
BHK_HD v3 unit3(v3 a) { return scale3(a, 1.f / sqrtf(fmaxf(dot3(a, a), 1e-30f))); }

// Frame and screen coordinates of the unit direction d from a camera with basis er, eth, eph.
BHK_HD RayFrame rayFrame(v3 d, v3 er, v3 eth, v3 eph, float rc)
{
    RayFrame F;
    F.d0 = d;
    const v3 toHole = scale3(er, -1.f);
    const float cpsi = fminf(fmaxf(dot3(d, toHole), -1.f), 1.f);
    F.psi = acosf(cpsi);
    F.b = rc * sinf(F.psi);
    F.alpha = atan2f(dot3(d, eph), -dot3(d, eth));
    v3 ea = add3(toHole, scale3(d, -cpsi));
    F.ea = unit3(ea);
    F.eb = cross3(d, F.ea);
    return F;
}

BHK_HD float weakDeflection(float M, float b, float psi) { return 2.f * M / fmaxf(b, 1e-3f) * (1.f + cosf(psi)); }

// Deflection of an ingoing Schwarzschild / Reissner-Nordstrom null ray from
// a static camera at radius rc (angle psi to the direction of the hole) up
// to its exit point at radius escR, measured like the integrators' exit
// direction (coordinate velocity e_r dr + e_ph r dph), minus the
// straight-line sweep pi - psi: the analytic base the network corrects for
// spin.  Impact parameter of the static observer's ray b = rc sin(psi) /
// sqrt(f(rc)).  Binet form phi = int du / sqrt(G(u)), G = 1/b^2 - u^2 (1 -
// 2Mu + Q^2 u^2), over [u_c, u_t] and [u_e, u_t] with u_t the turning point
// (Newton from 1/b); u = u_t (1 - t^2) removes the endpoint singularity and
// 16-point Gauss-Legendre integrates each leg; the exit tangent adds
// atan(u_e / sqrt(G(u_e))).  FP64.  Returns the first-order weak-field value
// when no turning point is found.
BHK_HD double sphG(double M, double Q2, double B, double u) { return B - u * u * (1.0 - 2.0 * M * u + Q2 * u * u); }

BHK_HD float sphDeflection(float Mf, float Qf, float rcf, float psif, float escRf)
{
    const double t16[16] = {0.0052995325041750307, 0.0277124884633837,   0.067184398806084122, 0.1222977958224985,
                            0.19106187779867811,   0.27099161117138632,  0.35919822461037054,  0.45249374508118129,
                            0.54750625491881877,   0.64080177538962946,  0.72900838882861363,  0.80893812220132189,
                            0.87770220417750155,   0.93281560119391593,  0.9722875115366163,   0.99470046749582497};
    const double w16[16] = {0.013576229705877019, 0.031126761969323853, 0.047579255841246296, 0.062314485627767015,
                            0.074797994408288382, 0.08457825969750131,  0.091301707522461806, 0.094725305227534293,
                            0.094725305227534293, 0.091301707522461806, 0.08457825969750131,  0.074797994408288382,
                            0.062314485627767015, 0.047579255841246296, 0.031126761969323853, 0.013576229705877019};
    const double M = Mf, Q2 = (double)Qf * (double)Qf, rc = rcf, psi = psif;
    const double f = 1.0 - 2.0 * M / rc + Q2 / (rc * rc);
    const double b = rc * sin(psi) / sqrt(fmax(f, 1e-12));
    const float weak = weakDeflection(Mf, (float)(rc * sin(psi)), psif);
    if (!(b > 1e-3) || !(f > 0.0)) return weak;
    const double B = 1.0 / (b * b), uc = 1.0 / rc, ue = 1.0 / (double)escRf;
    double u = fmax(1.0 / b, uc);
    bool ok = false;
    for (int it = 0; it < 60; ++it)
    {
        const double g = sphG(M, Q2, B, u);
        const double dg = -2.0 * u + 6.0 * M * u * u - 4.0 * Q2 * u * u * u;
        if (!(dg < 0.0)) break;
        const double un = u - g / dg;
        if (!(un > 0.0)) break;
        if (fabs(un - u) <= 1e-15 * un) { u = un; ok = true; break; }
        u = un;
    }
    if (!ok || u < uc || u < ue) return weak;
    const double ut = u;
    const double tc = sqrt(fmax(1.0 - uc / ut, 0.0)), te = sqrt(fmax(1.0 - ue / ut, 0.0));
    double I = 0.0;
    for (int k = 0; k < 16; ++k)
    {
        const double t1 = te * t16[k], t2 = tc * t16[k];
        const double g1 = sphG(M, Q2, B, ut * (1.0 - t1 * t1));
        const double g2 = sphG(M, Q2, B, ut * (1.0 - t2 * t2));
        if (!(g1 > 0.0) || !(g2 > 0.0)) return weak;
        I += w16[k] * (te * 2.0 * ut * t1 / sqrt(g1) + tc * 2.0 * ut * t2 / sqrt(g2));
    }
    const double ge = sphG(M, Q2, B, ue);
    if (!(ge > 0.0)) return weak;
    return (float)(I + atan(ue / sqrt(ge)) - (3.14159265358979323846 - psi));
}

// Normalised inputs: b / 40, alpha / pi, a*, q / 0.6, log(r_c / 6) / log(40 / 6), th_c / pi.
BHK_HD void surrogateInputs(const RayFrame& F, float aStar, float q, float rc, float thc, float* x)
{
    x[0] = F.b / kSurBMax;
    x[1] = F.alpha * 0.318309886183791f;
    x[2] = aStar;
    x[3] = q / 0.6f;
    x[4] = logf(rc / kSurRMin) / logf(kSurRMax / kSurRMin);
    x[5] = thc * 0.318309886183791f;
}

// Deflection components of an exit direction e in the ray frame.
BHK_HD void exitDeflection(const RayFrame& F, v3 e, float& dA, float& dB)
{
    e = unit3(e);
    const float c = dot3(e, F.d0);
    const float ta = dot3(e, F.ea), tb = dot3(e, F.eb);
    const float s = sqrtf(ta * ta + tb * tb);
    const float D = atan2f(s, c);
    if (s > 0.f) { dA = D * ta / s; dB = D * tb / s; }
    else { dA = 0.f; dB = 0.f; }
}

// Exit direction from deflection components.
BHK_HD v3 exitFromDeflection(const RayFrame& F, float dA, float dB)
{
    const float D = sqrtf(dA * dA + dB * dB);
    if (D <= 0.f) return F.d0;
    const float sD = sinf(D), cD = cosf(D);
    return unit3(add3(scale3(F.d0, cD), add3(scale3(F.ea, sD * dA / D), scale3(F.eb, sD * dB / D))));
}

}  // namespace bhk
