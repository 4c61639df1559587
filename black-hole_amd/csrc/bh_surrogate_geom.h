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
// with the weak-field deflection D0 = (2M/b)(1 + cos(psi)) of a ray from
// radius r_c to infinity (first order in M/b) removed.
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
BHK_HD inline void camBasis(v3 o, v3& er, v3& eth, v3& eph, float& r, float& th)
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

BHK_HD inline v3 cross3(v3 a, v3 b) { return v3{a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x}; }
BHK_HD inline float dot3(v3 a, v3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
BHK_HD inline v3 scale3(v3 a, float s) { return v3{a.x * s, a.y * s, a.z * s}; }
BHK_HD inline v3 add3(v3 a, v3 b) { return v3{a.x + b.x, a.y + b.y, a.z + b.z}; }
BHK_HD inline v3 unit3(v3 a) { return scale3(a, 1.f / sqrtf(fmaxf(dot3(a, a), 1e-30f))); }

// Frame and screen coordinates of the unit direction d from a camera with basis er, eth, eph.
BHK_HD inline RayFrame rayFrame(v3 d, v3 er, v3 eth, v3 eph, float rc)
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

BHK_HD inline float weakDeflection(float M, float b, float psi) { return 2.f * M / fmaxf(b, 1e-3f) * (1.f + cosf(psi)); }

// Normalised inputs: b / 40, alpha / pi, a*, q / 0.6, log(r_c / 6) / log(40 / 6), th_c / pi.
BHK_HD inline void surrogateInputs(const RayFrame& F, float aStar, float q, float rc, float thc, float* x)
{
    x[0] = F.b / kSurBMax;
    x[1] = F.alpha * 0.318309886183791f;
    x[2] = aStar;
    x[3] = q / 0.6f;
    x[4] = logf(rc / kSurRMin) / logf(kSurRMax / kSurRMin);
    x[5] = thc * 0.318309886183791f;
}

// Deflection components of an exit direction e in the ray frame.
BHK_HD inline void exitDeflection(const RayFrame& F, v3 e, float& dA, float& dB)
{
    e = unit3(e);
    const float c = dot3(e, F.d0);
    const float ta = dot3(e, F.ea), tb = dot3(e, F.eb);
    const float s = sqrtf(ta * ta + tb * tb);
    const float D = atan2f(s, c);
    if (s > 0.f) { dA = D * ta / s; dB = D * tb / s; }
    else { dA = 0.f; dB = 0.f; }
}

// Exit direction from predicted deflection components.
BHK_HD inline v3 exitFromDeflection(const RayFrame& F, float dA, float dB)
{
    const float D = sqrtf(dA * dA + dB * dB);
    if (D <= 0.f) return F.d0;
    const float sD = sinf(D), cD = cosf(D);
    return unit3(add3(scale3(F.d0, cD), add3(scale3(F.ea, sD * dA / D), scale3(F.eb, sD * dB / D))));
}

}  // namespace bhk
