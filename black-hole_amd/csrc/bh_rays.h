// ---------------------------------------------------------------------------
// bh_rays.h -- per-ray integrator state machines.
//
// Each ray is advanced one RK4 step per call so a persistent wavefront can
// interleave rays of very different lengths: lanes whose ray ends take a new
// ray from the work queue between step chunks.  The per-ray arithmetic is
// exactly the reference integrator's (same operations in the same order),
// so the output does not depend on the schedule.
//
//   SphRay  -- Binet equation u'' = 3 M u^2 - 2 Q^2 u^3 - u in the orbital
//              plane (trace.cuh:466-611), Schwarzschild and Reissner-Nordstrom.
//   KerrRay -- Boyer-Lindquist Hamiltonian RK4 for (r, th, ph, p_r, p_th)
//              with conserved E, Lz (trace.cuh:616-890), Kerr and Kerr-Newman.
// ---------------------------------------------------------------------------
#pragma once
#include "bh_common.h"

namespace bhk {

enum : int { RAY_RUNNING = 0, RAY_CAPTURED = 1, RAY_ESCAPED = 2 };

// Per-frame constants shared by every ray of a launch.  Built on the host
// (makeFrame) with exactly the float operations the reference repeats per
// step, so they are bit-identical, and passed by value as kernel arguments
// where the wavefront keeps them in scalar registers.
struct FrameK
{
    float M, a, Q, a2, Q2, twoM, threeM, twoQ2;
    float rCap, rFloor, uHor, escR, uEscape, invNineM, uBudget;
    float capR;                      // r <= capR: captured (trace.cuh:761)
    float dPhi, hFloor, rPhoton, rPhotonRefine, diskOuter, diskOuterCap, diskTime, sixM, rStatsFloor;
    float kRad, kAng, invKAng;       // adaptive-step budgets 8 dPhi and 2.2 dPhi (trace.cuh:793-795)
    int maxSteps, disk;
    DiskParams D;
};

BHK_HD FrameK makeFrame(const RenderParams& P)
{
    FrameK K;
    const bool kerr = (P.model == BH_KERR || P.model == BH_KERR_NEWMAN);
    K.M = P.M;
    K.a = kerr ? P.aSpin : 0.f;
    K.Q = P.Qc;
    K.a2 = K.a * K.a;
    K.Q2 = P.Qc * P.Qc;
    K.twoM = 2.f * P.M;
    K.threeM = 3.f * P.M;
    K.twoQ2 = 2.f * K.Q2;
    K.rCap = captureRadius(P);
    K.rFloor = fmaxf(P.rPlus, K.rCap);
    K.uHor = 1.0f / fmaxf(K.rCap, 1e-6f);
    K.escR = 120.f * P.M;
    K.uEscape = 1.0f / K.escR;
    K.invNineM = 1.f / (9.f * P.M);
    K.uBudget = 1.0f / fmaxf(1.5f * P.rPhoton, 3.f * P.M);
    K.capR = K.rCap * 1.002f + 1e-3f;
    K.dPhi = P.dPhi;
    K.hFloor = P.dPhi * 0.03f;
    K.rPhoton = P.rPhoton;
    K.rPhotonRefine = 1.5f * P.rPhoton;
    K.diskOuter = P.diskOuter;
    K.diskOuterCap = P.diskOuter * 1.3f;
    K.diskTime = P.diskTime;
    K.sixM = 6.f * P.M;
    K.rStatsFloor = K.rFloor * 1.05f;
    K.kRad = 8.0f * P.dPhi;
    K.kAng = 2.2f * P.dPhi;
    K.invKAng = 1.f / K.kAng;
    K.maxSteps = P.maxSteps;
    K.disk = P.diskEnabled != 0;
    K.D = makeDisk(P, K.a, P.Qc);
    return K;
}

// ========================================================================= spherical
template <class M> struct SphRay
{
    float u, du, phi, tFlight, bImpact, E, Lz, trans;
    v3 e1, e2, prevPos, dir3, acc, escDir;
    int step, status;

    BHK_INL void init(const FrameK& K, v3 origin, v3 dir)
    {
        status = RAY_RUNNING;
        step = 0;
        escDir = dir;
        acc = mk3(0.f, 0.f, 0.f);
        trans = 1.0f;
        photonEL<M>(origin, dir, K.M, 0.f, K.Q, E, Lz);
        float r0 = len3<M>(origin);
        e1 = vdiv<M>(origin, r0);
        float ddr = dot(dir, e1);
        v3 perp = dir - e1 * ddr;
        float pl = len3<M>(perp);
        if (pl < 1e-6f)
        {
            // exactly radial ray (trace.cuh:447-459, 500-505)
            status = (dot(dir, origin) < 0.f) ? RAY_CAPTURED : RAY_ESCAPED;
            return;
        }
        e2 = vdiv<M>(perp, pl);
        float f0 = fmaxf(1.f - M::div(K.twoM, r0) + M::div(K.Q2, r0 * r0), 1e-6f);
        bImpact = fmaxf(M::div(r0 * pl, M::sqrt_(f0)), 1e-6f);
        u = M::rcp(r0);
        du = M::div(-M::sqrt_(f0) * ddr, r0 * pl);
        phi = 0.f;
        tFlight = 0.f;
        prevPos = origin;
        dir3 = dir;
    }

    BHK_INL float acc_(const FrameK& K, float uu) const
    {
        return K.threeM * uu * uu - K.twoQ2 * uu * uu * uu - uu;
    }

    // One RK4 step; returns true once the ray has terminated.  The step's
    // disk chord, if it survives the culls, is returned in `ch` (ch.n > 0)
    // and must be sampled before the next step.
    BHK_INL bool advance(const FrameK& K, Chord& ch)
    {
        ch.n = 0;
        if (step >= K.maxSteps)
        {
            // budget exhausted (trace.cuh:603-609)
            if (u > K.uBudget) status = RAY_CAPTURED;
            else { status = RAY_ESCAPED; escDir = dir3; }
            return true;
        }
        float hMax = K.dPhi * clampf(M::div(K.invNineM, u), 1.f, 6.f);
        float h = fminf(hMax, M::div(0.10f, fmaxf(fabsf(du), 1e-6f)));
        if (h < K.hFloor) h = K.hFloor;
        float k1u = du;
        float k1v = acc_(K, u);
        float u2 = u + 0.5f * h * k1u;
        float v2 = du + 0.5f * h * k1v;
        float k2u = v2;
        float k2v = acc_(K, u2);
        float u3 = u + 0.5f * h * k2u;
        float v3_ = du + 0.5f * h * k2v;
        float k3u = v3_;
        float k3v = acc_(K, u3);
        float u4 = u + h * k3u;
        float v4 = du + h * k3v;
        float k4u = v4;
        float k4v = acc_(K, u4);
        float h6 = M::div(h, 6.f);
        u += h6 * (k1u + 2.f * k2u + 2.f * k3u + k4u);
        du += h6 * (k1v + 2.f * k2v + 2.f * k3v + k4v);
        phi += h;
        if (!isfinite(u) || !isfinite(du)) { status = RAY_CAPTURED; return true; }
        if (u >= K.uHor) { status = RAY_CAPTURED; return true; }
        if (u < 1e-4f) { status = RAY_ESCAPED; escDir = dir3; return true; }
        float r = M::rcp(u);
        float cp = M::cos_(phi), sp = M::sin_(phi);
        v3 m = e1 * cp + e2 * sp;
        v3 mp = e1 * (-sp) + e2 * cp;
        float drdphi = M::div(-du, u * u);
        v3 pos = m * r;
        dir3 = normalize<M>(m * drdphi + mp * r);
        float fR = fmaxf(1.f - M::div(K.twoM, r) + M::div(K.Q2, r * r), 1e-4f);
        tFlight += M::div(h * (r * r), bImpact * fR);
        if (K.disk && trans > 0.01f) chordPrepare<M>(K.D, prevPos, pos, K.diskTime - tFlight, ch);
        prevPos = pos;
        if (u < K.uEscape && du < 0.f) { status = RAY_ESCAPED; escDir = dir3; return true; }
        ++step;
        return false;
    }
};

// ========================================================================= Kerr-Newman
struct KDeriv { float r, th, ph, pr, pth; };

// Fast-policy right-hand side that also returns P, P/Delta, 1/Sigma and the
// unclamped Delta of the state (reused by the slow-light integral).
template <class M, bool kWant>
BHK_INL KDeriv kerrRHSFastX(const FrameK& K, float E, float Lz, float r, float pr, float pth, float s, float cth,
                            float& Kout, float& PpO, float& PDO, float& invSO, float& DelO)
{
    if (s < 1e-4f) s = 1e-4f;
    const float a = K.a, a2 = K.a2;
    KDeriv d;
    const float aE = a * E;
    const float r2 = r * r;
    const float is = M::rcp(s), is2 = is * is;
    const float Pp = fmaf(E, r2, E * a2 - a * Lz);
    const float DelRaw = fmaf(r, r - K.twoM, a2 + K.Q2);
    const float Del = DelRaw < 1e-6f ? 1e-6f : DelRaw;
    const float Sig = fmaf(a2 * cth, cth, r2);
    const float iDel = M::rcp(Del), invS = M::rcp(Sig);
    const float PD = Pp * iDel;
    const float B = fmaf(Lz, is, -aE * s);
    const float W = fmaf(B, B, -Pp * PD);
    const float Dpr = Del * pr;
    const float Kc = fmaf(Dpr, pr, fmaf(pth, pth, W));
    if (kWant) Kout = Kc;
    const float dDel = 2.f * r - K.twoM;
    const float Wr = fmaf(PD * PD, dDel, -4.f * E * r * PD);
    const float dB = -cth * fmaf(Lz, is2, aE);
    const float iSK = invS * Kc;
    d.r = Dpr * invS;
    d.th = pth * invS;
    d.ph = fmaf(a, PD, fmaf(Lz, is2, -aE)) * invS;
    d.pr = invS * fmaf(-0.5f, fmaf(dDel * pr, pr, Wr), r * iSK);
    d.pth = invS * fmaf(-B, dB, -(a2 * s * cth) * iSK);
    PpO = Pp; PDO = PD; invSO = invS; DelO = DelRaw;
    return d;
}

// Hamiltonian right-hand side (trace.cuh:620-659); s, c = sin/cos of th.
template <class M, bool kWant>
BHK_INL KDeriv kerrRHS(const FrameK& K, float E, float Lz, float r, float pr, float pth, float s, float cth,
                       float& Kout)
{
    if (s < 1e-4f) s = 1e-4f;
    const float a = K.a, a2 = K.a2;
    KDeriv d;
    if constexpr (M::fast)
    {
        // The same Hamiltonian right-hand side, reduced algebraically: three
        // reciprocals (1/sin th, 1/Delta, 1/Sigma), P/Delta formed once, and
        // W_r = -(2 P P' Delta - P^2 Delta') / Delta^2 = -4 r E (P/Delta) + (P/Delta)^2 Delta'.
        const float aE = a * E;
        const float r2 = r * r;
        const float is = M::rcp(s), is2 = is * is;
        const float Pp = fmaf(E, r2, E * a2 - a * Lz);
        float Del = fmaf(r, r - K.twoM, a2 + K.Q2);
        if (Del < 1e-6f) Del = 1e-6f;
        const float Sig = fmaf(a2 * cth, cth, r2);
        const float iDel = M::rcp(Del), invS = M::rcp(Sig);
        const float PD = Pp * iDel;
        const float B = fmaf(Lz, is, -aE * s);
        const float W = fmaf(B, B, -Pp * PD);
        const float Dpr = Del * pr;
        const float Kc = fmaf(Dpr, pr, fmaf(pth, pth, W));
        if (kWant) Kout = Kc;
        const float dDel = 2.f * r - K.twoM;
        const float Wr = fmaf(PD * PD, dDel, -4.f * E * r * PD);
        const float dB = -cth * fmaf(Lz, is2, aE);
        const float iSK = invS * Kc;
        d.r = Dpr * invS;
        d.th = pth * invS;
        d.ph = fmaf(a, PD, fmaf(Lz, is2, -aE)) * invS;
        d.pr = invS * fmaf(-0.5f, fmaf(dDel * pr, pr, Wr), r * iSK);
        d.pth = invS * fmaf(-B, dB, -(a2 * s * cth) * iSK);
        return d;
    }
    float s2 = s * s;
    float r2 = r * r;
    float Sig = r2 + a2 * cth * cth;
    float Del = r2 - K.twoM * r + a2 + K.Q2;
    if (Del < 1e-6f) Del = 1e-6f;
    float Pp = E * (r2 + a2) - a * Lz;
    float B = M::div(Lz, s) - a * E * s;
    float W = B * B - M::div(Pp * Pp, Del);
    float Kc = Del * pr * pr + pth * pth + W;
    if (kWant) Kout = Kc;
    float dDel = 2.f * r - K.twoM;
    float dP = 2.f * r * E;
    float Wr = M::div(-(2.f * Pp * dP * Del - Pp * Pp * dDel), Del * Del);
    float dB = M::div(-Lz * cth, s2) - a * E * cth;
    float Wth = 2.f * B * dB;
    float invS = M::rcp(Sig);
    float invS2 = invS * invS;
    d.r = Del * pr * invS;
    d.th = pth * invS;
    d.ph = ((M::div(Lz, s2) - a * E) + M::div(a * Pp, Del)) * invS;
    d.pr = -0.5f * invS * (dDel * pr * pr + Wr) + r * invS2 * Kc;
    d.pth = -0.5f * invS * Wth - a2 * s * cth * invS2 * Kc;
    return d;
}

template <class M, bool kWant> struct KerrRay
{
    float r, th, ph, pr, pth;     // Boyer-Lindquist state
    float sth, cth;               // sin/cos of th at the current state
    float E, Lz, tFlight, trans, maxHviol;
    float dLr, dLth, dLph;        // last RK4 stage derivative (budget exit)
    v3 prevPos, dir3, acc, escDir;
    int step, status, haveDeriv;
    KDeriv nx;                    // fast policy: right-hand side at the current state
    float nxK;
    float prevRc, phPrev;         // fast policy: cylindrical radius and azimuth of prevPos
    bool prevXZ;                  // fast policy: prevPos.x, .z are current

    BHK_INL void init(const FrameK& K, v3 origin, v3 dir)
    {
        status = RAY_RUNNING;
        step = 0;
        haveDeriv = 0;
        escDir = dir;
        acc = mk3(0.f, 0.f, 0.f);
        trans = 1.0f;
        maxHviol = 0.f;
        const float a = K.a, Mm = K.M, Q = K.Q;
        float r0 = len3<M>(origin);
        float cth0 = clampf(M::div(origin.y, r0), -1.f, 1.f);
        float th0 = M::acos_(cth0);
        float ph0 = M::atan2_(-origin.z, origin.x);
        float s = M::sin_(th0), c = M::cos_(th0);
        float sp = M::sin_(ph0), cp = M::cos_(ph0);
        // orthonormal BL basis at the camera (trace.cuh:673-682)
        v3 er = mk3(s * cp, c, -s * sp);
        v3 eth = mk3(c * cp, -s, -c * sp);
        v3 eph = mk3(-sp, 0.f, -cp);
        float drc = dot(dir, er), dtc = dot(dir, eth), dpc = dot(dir, eph);
        // static-observer tetrad (trace.cuh:716-739)
        float s0 = s < 1e-4f ? 1e-4f : s;
        float s02 = s0 * s0;
        float Sig0 = r0 * r0 + a * a * cth0 * cth0;
        float Del0 = r0 * r0 - 2.f * Mm * r0 + a * a + Q * Q;
        float m2r0 = 2.f * Mm * r0 - Q * Q;
        float gtt = -(1.f - M::div(m2r0, Sig0));
        if (gtt > -1e-5f) gtt = -1e-5f;
        float gtp = M::div(-a * m2r0 * s02, Sig0);
        float gpp = (r0 * r0 + a * a + M::div(m2r0 * a * a * s02, Sig0)) * s02;
        float A = M::rsq(fmaxf(gpp - M::div(gtp * gtp, gtt), 1e-8f));
        float ptUp = M::rsq(-gtt) - M::div(dpc * A * gtp, gtt);
        float pfUp = dpc * A;
        E = -(gtt * ptUp + gtp * pfUp);
        Lz = gtp * ptUp + gpp * pfUp;
        r = r0; th = th0; ph = ph0;
        pr = drc * M::sqrt_(M::div(Sig0, fmaxf(Del0, 1e-6f)));
        pth = dtc * M::sqrt_(Sig0);
        sth = s; cth = c;
        tFlight = 0.f;
        prevPos = origin;
        prevRc = M::sqrt_(origin.x * origin.x + origin.z * origin.z);
        phPrev = ph0;
        prevXZ = true;
        dir3 = dir;
        // fast policy: the first step's first stage (each step then leaves
        // the next step's first stage in nx)
        if constexpr (M::fast) nx = kerrRHS<M, kWant>(K, E, Lz, r, pr, pth, sth, cth, nxK);
    }

    BHK_INL v3 bodyDir(float dr_, float dth_, float dph_, float sp, float cp) const
    {
        // e_r dr + e_th r dth + e_ph r sin(th) dph (trace.cuh:776-779)
        v3 e1b = mk3(sth * cp, cth, -sth * sp);
        v3 e2b = mk3(cth * cp, -sth, -cth * sp);
        v3 e3b = mk3(-sp, 0.f, -cp);
        return normalize<M>(e1b * dr_ + e2b * (r * dth_) + e3b * (r * sth * dph_));
    }

    BHK_INL bool advance(const FrameK& K, Chord& ch)
    {
        ch.n = 0;
        if (step >= K.maxSteps)
        {
            // budget exhausted (trace.cuh:871-888)
            if (r < K.sixM) status = RAY_CAPTURED;
            else
            {
                status = RAY_ESCAPED;
                if (haveDeriv) escDir = bodyDir(dLr, dLth, dLph, M::sin_(ph), M::cos_(ph));
                else escDir = dir3;
            }
            return true;
        }
        float Kv = 0.f;
        KDeriv d1;
        if constexpr (M::fast) { d1 = nx; Kv = nxK; }
        else d1 = kerrRHS<M, kWant>(K, E, Lz, r, pr, pth, sth, cth, Kv);
        if (!isfinite(r) || !isfinite(pr) || !isfinite(pth) || !isfinite(d1.r) || !isfinite(d1.pr))
        { status = RAY_CAPTURED; return true; }
        if (r <= K.capR) { status = RAY_CAPTURED; return true; }
        if (kWant && r > K.rStatsFloor)
        {
            float viol = M::div(fabsf(Kv), E * E * (r * r + K.a * K.a) + 1e-12f);
            if (viol > maxHviol) maxHviol = viol;
        }
        if (r >= K.escR && d1.r > 0.f)
        {
            escDir = bodyDir(d1.r, d1.th, d1.ph, M::sin_(ph), M::cos_(ph));
            status = RAY_ESCAPED;
            return true;
        }
        // adaptive affine step (trace.cuh:793-809)
        const float spd = fabsf(d1.r) + r * (fabsf(d1.th) + fabsf(d1.ph)) + 1e-6f;
        float capLen;
        if (K.disk && r < K.diskOuterCap) capLen = (r * fabsf(cth) < 1.8f) ? 1.2f : 2.2f;
        else capLen = fmaxf(1.2f, 0.22f * r);
        float h;
        if constexpr (M::fast)
        {
            // h = min(f Bq / (|dr| + C Bq), capLen / spd) with one reciprocal:
            // 1 / (|dr| / Bq + C) = Bq / (|dr| + C Bq), and the smaller of two
            // positive ratios is picked by cross-multiplying
            const float Bq = K.kRad * fmaxf(r - K.rFloor, 0.02f);
            const float C = fmaf(fabsf(d1.th) + fabsf(d1.ph), K.invKAng, 1e-5f);
            const float D = fmaf(C, Bq, fabsf(d1.r));
            const float n1 = (r < K.rPhotonRefine) ? 0.55f * Bq : Bq;
            const bool first = n1 * spd <= capLen * D;
            h = (first ? n1 : capLen) * M::rcp(first ? D : spd);
        }
        else
        {
            h = M::rcp(M::div(fabsf(d1.r), K.kRad * fmaxf(r - K.rFloor, 0.02f)) +
                       M::divc(fabsf(d1.th) + fabsf(d1.ph), K.kAng, K.invKAng) + 1e-5f);
            if (r < K.rPhotonRefine) h *= 0.55f;
            h = fminf(h, M::div(capLen, spd));
        }

        // classic RK4; the weighted stage sum (d1 + 2 d2) + 2 d3 + d4 is
        // accumulated as each stage completes (same operations, same order)
        const float hh = 0.5f * h;
        float kd;
        KDeriv sum = d1;
        float yth = th + hh * d1.th;
        KDeriv dk = kerrRHS<M, false>(K, E, Lz, r + hh * d1.r, pr + hh * d1.pr, pth + hh * d1.pth,
                                      M::sin_(yth), M::cos_(yth), kd);
        sum.r += 2.f * dk.r; sum.th += 2.f * dk.th; sum.ph += 2.f * dk.ph; sum.pr += 2.f * dk.pr; sum.pth += 2.f * dk.pth;
        yth = th + hh * dk.th;
        dk = kerrRHS<M, false>(K, E, Lz, r + hh * dk.r, pr + hh * dk.pr, pth + hh * dk.pth, M::sin_(yth),
                               M::cos_(yth), kd);
        sum.r += 2.f * dk.r; sum.th += 2.f * dk.th; sum.ph += 2.f * dk.ph; sum.pr += 2.f * dk.pr; sum.pth += 2.f * dk.pth;
        yth = th + h * dk.th;
        dk = kerrRHS<M, false>(K, E, Lz, r + h * dk.r, pr + h * dk.pr, pth + h * dk.pth, M::sin_(yth), M::cos_(yth),
                               kd);
        sum.r += dk.r; sum.th += dk.th; sum.ph += dk.ph; sum.pr += dk.pr; sum.pth += dk.pth;
        const float h6 = M::div(h, 6.f);
        r += h6 * sum.r;
        th += h6 * sum.th;
        ph += h6 * sum.ph;
        pr += h6 * sum.pr;
        pth += h6 * sum.pth;
        if (th < 1e-3f) { th = 1e-3f; pth = fabsf(pth); }
        else if (th > kPi - 1e-3f) { th = kPi - 1e-3f; pth = -fabsf(pth); }
        dLr = dk.r; dLth = dk.th; dLph = dk.ph;
        haveDeriv = 1;
        sth = M::sin_(th);
        cth = M::cos_(th);
        if constexpr (M::fast)
        {
            // right-hand side at the new state (the next step's first stage);
            // slow light from its P/Delta and 1/Sigma (trace.cuh:840-851;
            // Delta is clamped at 1e-5 there)
            float Pp, PD, iS, DelRaw;
            nx = kerrRHSFastX<M, kWant>(K, E, Lz, r, pr, pth, sth, cth, nxK, Pp, PD, iS, DelRaw);
            const float s = sth < 1e-4f ? 1e-4f : sth;
            const float a = K.a;
            const float PDs = DelRaw < 1e-5f ? Pp * 1e5f : PD;
            const float dtdl = fmaf(a, Lz - a * E * s * s, (r * r + K.a2) * PDs) * iS;
            tFlight += h * fabsf(dtdl);
        }
        else
        {
            // slow light (trace.cuh:840-851)
            float s = sth < 1e-4f ? 1e-4f : sth;
            float a = K.a;
            float Sig = r * r + K.a2 * cth * cth;
            float Del = fmaxf(r * r - K.twoM * r + K.a2 + K.Q2, 1e-5f);
            float Pp = E * (r * r + K.a2) - a * Lz;
            float dtdl = M::div(a * (Lz - a * E * s * s) + M::div((r * r + K.a2) * Pp, Del), fmaxf(Sig, 1e-8f));
            tFlight += h * fabsf(dtdl);
        }
        if (K.disk)
        {
            // the reference also keeps the unit step direction here
            // (trace.cuh:859-862), but its only reader, the disk segment
            // sampler, ignores it (trace.cuh:270), and the exit direction
            // comes from the coordinate velocity: it is not formed
            if constexpr (M::fast)
            {
                // The new point's cylindrical radius r sin(th) and height
                // r cos(th) need no azimuth, and the chord culls need no more
                // (the previous point's are carried).  The culls are first
                // applied from them alone: the slab and scale-height tests
                // exactly as in chordPrepareRc, the radial test against an
                // upper bound of the chord length, |b - a|^2 <= (rcB - rcA)^2
                // + rcA rcB dph^2 + (yB - yA)^2 (from 2 (1 - cos x) <= x^2),
                // widened by 1% + 1e-4 rcMax^2 for the rounding of the
                // hardware sin/cos and by 0.01 in the gap.  Only chords that
                // may survive form their Cartesian ends: the same products
                // as r sin(th) cos(ph) etc., so the ends, the chord and the
                // exact culls of chordPrepareRc are unchanged.
                const float rcB = r * sth, yB = r * cth;
                const float rcA = prevRc, yA = prevPos.y;
                const DiskParams& D = K.D;
                bool may = trans > 0.01f;
                if (may)
                {
                    const float rcMax = fmaxf(rcA, rcB);
                    if ((yA > D.maxH && yB > D.maxH) || (yA < -D.maxH && yB < -D.maxH)) may = false;
                    else if ((yA > 0.f) == (yB > 0.f) &&
                             fminf(fabsf(yA), fabsf(yB)) > 1.02f * 2.6f * fmaxf(0.055f * rcMax, 1e-4f))
                        may = false;
                    else
                    {
                        const float dph = ph - phPrev, drc = rcB - rcA, dy = yB - yA;
                        const float seg2 = fmaf(1.01f * rcA * rcB, dph * dph, 1.01f * fmaf(drc, drc, dy * dy)) +
                                           1e-4f * rcMax * rcMax;
                        const float outGap = fminf(rcA, rcB) - D.rout - 0.01f;
                        const float inGap = D.rin - rcMax - 0.01f;
                        if ((outGap > 0.f && outGap * outGap > seg2) || (inGap > 0.f && inGap * inGap > seg2))
                            may = false;
                    }
                }
                if (may)
                {
                    const float sp = M::sin_(ph), cp = M::cos_(ph);
                    const v3 pos = mk3(rcB * cp, yB, -rcB * sp);
                    if (!prevXZ)
                    {
                        // rounded products, as when the end is carried from
                        // the step before (not fused into the chord's sums)
#pragma clang fp contract(off)
                        const float spa = M::sin_(phPrev), cpa = M::cos_(phPrev);
                        prevPos.x = rcA * cpa;
                        prevPos.z = -rcA * spa;
                    }
                    chordPrepareRc<M>(D, prevPos, pos, rcA, rcB, K.diskTime - tFlight, ch);
                    prevPos = pos;
                    prevXZ = true;
                }
                else
                {
                    prevPos.y = yB;
                    prevXZ = false;
                }
                prevRc = rcB;
                phPrev = ph;
            }
            else
            {
                float sp = M::sin_(ph), cp = M::cos_(ph);
                v3 pos = mk3(r * sth * cp, r * cth, -r * sth * sp);
                if (trans > 0.01f) chordPrepare<M>(K.D, prevPos, pos, K.diskTime - tFlight, ch);
                prevPos = pos;
            }
        }
        ++step;
        return false;
    }
};

template <class M> BHK_INL float rayMaxHviol(const SphRay<M>&) { return 0.f; }
template <class M, bool W> BHK_INL float rayMaxHviol(const KerrRay<M, W>& R) { return R.maxHviol; }

}  // namespace bhk
