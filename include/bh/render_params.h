/*
 * bh/render_params.h -- the per-frame parameter block of the geodesic tracer.
 *
 * Layout-identical to the reference's RenderParams (render_params.h:21-87):
 * same field order, same types, 160 bytes, 4-byte alignment.  The vector
 * members use this header's own POD structs, so a caller needs neither HIP
 * nor CUDA headers to fill one in.
 *
 * Units: G = c = 1, default M = 0.5 so the Schwarzschild radius 2M = 1 code
 * unit.  Spin a = a* M (signed), charge Q = q M (>= 0).
 *
 * Compiles as C99 and as C++.  In C++ every field carries the reference's
 * default value (render_params.h:24-86); C callers start from
 * bhDefaultRenderParams() declared in bh/bhtrace.h.
 */
#ifndef BH_RENDER_PARAMS_H
#define BH_RENDER_PARAMS_H

#ifdef __cplusplus
#define BH_DEF(v) = v
#else
#define BH_DEF(v)
#endif

typedef struct bh_float3 { float x, y, z; } bh_float3;
typedef struct bh_float4 { float x, y, z, w; } __attribute__((aligned(16))) bh_float4;
typedef struct bh_uchar4 { unsigned char x, y, z, w; } bh_uchar4;

/* Model ids (reference metric.cuh:32-38). */
enum { BH_SCHWARZSCHILD = 0, BH_REISSNER_NORDSTROM = 1, BH_KERR = 2, BH_KERR_NEWMAN = 3 };

typedef struct RenderParams
{
    /* Framebuffer */
    int       width   BH_DEF(1280);
    int       height  BH_DEF(720);
    int       swapRB  BH_DEF(0);          /* 1: write B,G,R,A instead of R,G,B,A */

    /* Camera (world frame; black hole at the origin, spin axis +Y) */
    bh_float3 camPos     BH_DEF((bh_float3{0.f, 0.f, 0.f}));
    bh_float3 camForward BH_DEF((bh_float3{0.f, 0.f, -1.f}));
    bh_float3 camRight   BH_DEF((bh_float3{1.f, 0.f, 0.f}));
    bh_float3 camUp      BH_DEF((bh_float3{0.f, 1.f, 0.f}));
    float     tanHalfFov BH_DEF(0.5773503f);
    float     aspect     BH_DEF(1280.f / 720.f);

    /* Black hole */
    int       model      BH_DEF(0);       /* BH_SCHWARZSCHILD .. BH_KERR_NEWMAN */
    float     M          BH_DEF(0.5f);
    float     aSpin      BH_DEF(0.f);     /* a = a* M */
    float     Qc         BH_DEF(0.f);     /* Q = q M */
    float     rPlus      BH_DEF(1.0f);    /* outer horizon, 0 when naked */
    float     rPhoton    BH_DEF(1.5f);
    float     rErgo      BH_DEF(1.0f);
    int       allowNaked BH_DEF(0);

    /* Integration */
    float     dPhi       BH_DEF(0.012f);
    int       maxSteps   BH_DEF(2000);

    /* Accretion disk */
    int       diskEnabled   BH_DEF(1);
    float     diskInner     BH_DEF(3.0f);
    float     diskOuter     BH_DEF(8.0f);
    float     diskTime      BH_DEF(0.f);
    float     diskEmisScale BH_DEF(1.5f);
    float     diskAbsScale  BH_DEF(2.2f);

    /* ISCO hot spots */
    int       hotSpotsEnabled BH_DEF(1);
    float     hotSpotStrength BH_DEF(0.65f);

    /* Post processing */
    float     exposure    BH_DEF(1.0f);

    /* Progressive accumulation: mode 0 overwrite, 1 running mean, 2 EMA */
    int       sampleIndex BH_DEF(0);
    int       accumMode   BH_DEF(0);

    /* Bloom */
    int       bloomEnabled  BH_DEF(1);
    float     bloomStrength BH_DEF(0.35f);
} RenderParams;

#undef BH_DEF

#ifdef __cplusplus
#include <cstddef>
static_assert(sizeof(RenderParams) == 160, "RenderParams must stay 160 bytes");
static_assert(alignof(RenderParams) == 4, "RenderParams must stay 4-byte aligned");
static_assert(offsetof(RenderParams, width) == 0 && offsetof(RenderParams, height) == 4 &&
              offsetof(RenderParams, swapRB) == 8 && offsetof(RenderParams, camPos) == 12 &&
              offsetof(RenderParams, camForward) == 24 && offsetof(RenderParams, camRight) == 36 &&
              offsetof(RenderParams, camUp) == 48 && offsetof(RenderParams, tanHalfFov) == 60 &&
              offsetof(RenderParams, aspect) == 64 && offsetof(RenderParams, model) == 68 &&
              offsetof(RenderParams, M) == 72 && offsetof(RenderParams, aSpin) == 76 &&
              offsetof(RenderParams, Qc) == 80 && offsetof(RenderParams, rPlus) == 84 &&
              offsetof(RenderParams, rPhoton) == 88 && offsetof(RenderParams, rErgo) == 92 &&
              offsetof(RenderParams, allowNaked) == 96 && offsetof(RenderParams, dPhi) == 100 &&
              offsetof(RenderParams, maxSteps) == 104 && offsetof(RenderParams, diskEnabled) == 108 &&
              offsetof(RenderParams, diskInner) == 112 && offsetof(RenderParams, diskOuter) == 116 &&
              offsetof(RenderParams, diskTime) == 120 && offsetof(RenderParams, diskEmisScale) == 124 &&
              offsetof(RenderParams, diskAbsScale) == 128 &&
              offsetof(RenderParams, hotSpotsEnabled) == 132 &&
              offsetof(RenderParams, hotSpotStrength) == 136 && offsetof(RenderParams, exposure) == 140 &&
              offsetof(RenderParams, sampleIndex) == 144 && offsetof(RenderParams, accumMode) == 148 &&
              offsetof(RenderParams, bloomEnabled) == 152 && offsetof(RenderParams, bloomStrength) == 156,
              "RenderParams offsets must match the reference layout");
static_assert(sizeof(bh_float4) == 16 && alignof(bh_float4) == 16, "float4 layout");
static_assert(sizeof(bh_uchar4) == 4, "uchar4 layout");
#endif

#endif /* BH_RENDER_PARAMS_H */
