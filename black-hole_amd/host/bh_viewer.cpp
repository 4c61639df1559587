// ---------------------------------------------------------------------------
// bh_viewer.cpp -- caller-side producer of RenderParams for an interactive or
// scripted viewer: orbit camera, black-hole body orientation, model knobs,
// quality presets, disk clock and progressive-accumulation bookkeeping.
//
// Follows the reference application loop (main.cpp):
//   AppState defaults ............... main.cpp:48-88
//   applyBlackHole (model knobs) .... main.cpp:92-134
//   clampCamera ..................... main.cpp:136-145
//   applyBodyOrientation R=RzRyRx ... main.cpp:147-167
//   fillCamera ...................... main.cpp:186-228
//   mouse / key / held-key knobs .... main.cpp:230-337
//   per-frame parameter assembly and
//   accumulation restart rules ...... main.cpp:418-478
// Window-system keys (ESC, F11) have no meaning without a window and are
// ignored.  The disk inner edge is the FP64 ISCO root wherever the hole has a
// horizon (the reference scans in float with a 0.001 M step; both agree to
// that step), and the float scan for naked singularities.
// ---------------------------------------------------------------------------
#include <cmath>
#include <new>

#include "../../include/bh/bhtrace.h"

namespace {

constexpr float kDefaultAzimuth = 0.6f;
constexpr float kDefaultElevation = 0.20f;
constexpr float kDefaultDistance = 16.0f;
constexpr float kExtremal = 0.995f;
constexpr float kNaked = 2.25f;
constexpr float kMass = 0.5f;      // rs = 2M = 1 code unit
constexpr float kPi = 3.14159265358979f;

// Snapshot of everything that changes the traced geometry; exposure and bloom
// act after accumulation and are left out.
struct GeomKey
{
    int width, height;
    bh_float3 pos, fwd;
    int model;
    float a, Q, dPhi;
    int maxSteps, disk, naked, hot;
    float oriX, oriY, oriZ;
    bool operator==(const GeomKey& o) const
    {
        return width == o.width && height == o.height && pos.x == o.pos.x && pos.y == o.pos.y &&
               pos.z == o.pos.z && fwd.x == o.fwd.x && fwd.y == o.fwd.y && fwd.z == o.fwd.z &&
               model == o.model && a == o.a && Q == o.Q && dPhi == o.dPhi && maxSteps == o.maxSteps &&
               disk == o.disk && naked == o.naked && hot == o.hot && oriX == o.oriX && oriY == o.oriY &&
               oriZ == o.oriZ;
    }
};

void normalize3(float& x, float& y, float& z)
{
    const float l = std::sqrt(x * x + y * y + z * z);
    if (l < 1e-8f) { x = 0.f; y = 1.f; z = 0.f; return; }
    x /= l;
    y /= l;
    z /= l;
}

// R = Rz(oz) Ry(oy) Rx(ox) about fixed world axes.
void rotateBody(float ox, float oy, float oz, float& x, float& y, float& z)
{
    const float cx = std::cos(ox), sx = std::sin(ox);
    const float cy = std::cos(oy), sy = std::sin(oy);
    const float cz = std::cos(oz), sz = std::sin(oz);
    const float y1 = cx * y - sx * z, z1 = sx * y + cx * z;      // Rx
    const float x2 = cy * x + sy * z1, z2 = -sy * x + cy * z1;   // Ry
    x = cz * x2 - sz * y1;                                       // Rz
    y = sz * x2 + cz * y1;
    z = z2;
}

}  // namespace

struct BhViewer
{
    BhViewState s;
    int width = 1280, height = 720, swapRB = 0;
    bool dirty = true;         // model knobs changed since the last derive
    bool havePrev = false;
    GeomKey prev{};
    RenderParams hole;         // model block of the last derive
};

namespace {

void clampView(BhViewState& s)
{
    const float elMax = 1.45f;   // keeps the up vector away from the poles
    if (s.elevation > elMax) s.elevation = elMax;
    if (s.elevation < -elMax) s.elevation = -elMax;
    const float rMin = s.allowNaked ? 1.5f : 2.2f;
    if (s.distance < rMin) s.distance = rMin;
    if (s.distance > 55.f) s.distance = 55.f;
    if (s.exposure < 0.05f) s.exposure = 0.05f;
    if (s.exposure > 20.f) s.exposure = 20.f;
}

void deriveHole(BhViewer& v)
{
    BhViewState& s = v.s;
    const float knob = s.allowNaked ? std::sqrt(kNaked) : kExtremal;
    if (s.aStar < -knob) s.aStar = -knob;
    if (s.aStar > knob) s.aStar = knob;
    if (s.q < 0.f) s.q = 0.f;
    if (s.q > knob) s.q = knob;
    float M = kMass, aS = s.aStar, qq = s.q;
    BHDerived d;
    bhValidateAndDerive(s.model, &M, &aS, &qq, s.allowNaked ? 1 : 0, &d);
    if (d.clamped) { s.aStar = aS; s.q = qq; }
    RenderParams& p = v.hole;
    p.model = s.model;
    p.M = M;
    p.aSpin = d.a;
    p.Qc = d.Q;
    p.rPlus = d.rPlus;
    p.rPhoton = d.rPhoton;
    p.rErgo = d.rErgo;
    p.allowNaked = s.allowNaked ? 1 : 0;
    float isco = d.rIsco;
    if (!d.naked)
    {
        const double r = bhIscoFP64(M, d.a, d.Q);
        if (std::isfinite(r)) isco = (float)r;
    }
    p.diskInner = isco;
    p.diskOuter = 8.0f;
    v.dirty = false;
}

}  // namespace

extern "C" {

void bhFillCameraOriented(RenderParams* p, float az, float el, float dist, float fovY, float oriX, float oriY,
                          float oriZ)
{
    const float ca = std::cos(az), sa = std::sin(az), ce = std::cos(el), se = std::sin(el);
    float px = dist * ce * ca, py = dist * se, pz = dist * ce * sa;
    const float fl = std::sqrt(px * px + py * py + pz * pz);
    float fx = -px / fl, fy = -py / fl, fz = -pz / fl;
    // right = normalize(forward x +Y) in the body frame
    float rx = -fz, ry = 0.f, rz = fx;
    float rl = std::sqrt(rx * rx + ry * ry + rz * rz);
    if (rl < 1e-6f) { rx = 1.f; ry = 0.f; rz = 0.f; rl = 1.f; }
    rx /= rl;
    rz /= rl;
    // body frame to world
    rotateBody(oriX, oriY, oriZ, px, py, pz);
    rotateBody(oriX, oriY, oriZ, fx, fy, fz);
    rotateBody(oriX, oriY, oriZ, rx, ry, rz);
    normalize3(fx, fy, fz);
    normalize3(rx, ry, rz);
    // up = right x forward, taken after the rotation so the basis stays orthogonal
    float ux = ry * fz - rz * fy, uy = rz * fx - rx * fz, uz = rx * fy - ry * fx;
    normalize3(ux, uy, uz);
    p->camPos = bh_float3{px, py, pz};
    p->camForward = bh_float3{fx, fy, fz};
    p->camRight = bh_float3{rx, ry, rz};
    p->camUp = bh_float3{ux, uy, uz};
    p->tanHalfFov = std::tan(fovY * kPi / 180.f * 0.5f);
    p->aspect = (p->height > 0) ? (float)p->width / (float)p->height : 16.f / 9.f;
}

void bhViewStateDefaults(BhViewState* s)
{
    if (!s) return;
    s->azimuth = kDefaultAzimuth;
    s->elevation = kDefaultElevation;
    s->distance = kDefaultDistance;
    s->oriX = s->oriY = s->oriZ = 0.f;
    s->model = BH_SCHWARZSCHILD;
    s->aStar = 0.70f;
    s->q = 0.40f;
    s->allowNaked = 0;
    s->exposure = 0.85f;
    s->diskOn = 1;
    s->bloomOn = 1;
    s->hotSpotsOn = 1;
    s->animPaused = 0;
    s->dPhi = 0.012f;
    s->maxSteps = 2000;
    s->fovY = 60.f;
    s->diskTime = 0.f;
    s->sampleIndex = 0;
}

int bhViewerCreate(int width, int height, BhViewer** out)
{
    if (!out || width < 1 || height < 1) return -1;
    BhViewer* v = new (std::nothrow) BhViewer;
    if (!v) return -1;
    bhViewStateDefaults(&v->s);
    v->width = width;
    v->height = height;
    bhDefaultRenderParams(&v->hole);
    clampView(v->s);
    *out = v;
    return 0;
}

void bhViewerDestroy(BhViewer* v) { delete v; }

int bhViewerGetState(const BhViewer* v, BhViewState* s)
{
    if (!v || !s) return -1;
    *s = v->s;
    return 0;
}

int bhViewerSetState(BhViewer* v, const BhViewState* s)
{
    if (!v || !s) return -1;
    const BhViewState& o = v->s;
    if (s->model != o.model || s->aStar != o.aStar || s->q != o.q || s->allowNaked != o.allowNaked) v->dirty = true;
    v->s = *s;
    if (v->s.model < BH_SCHWARZSCHILD || v->s.model > BH_KERR_NEWMAN) v->s.model = BH_SCHWARZSCHILD;
    if (!(v->s.dPhi > 0.f)) v->s.dPhi = 0.012f;
    if (v->s.maxSteps < 1) v->s.maxSteps = 1;
    if (v->s.sampleIndex < 0) v->s.sampleIndex = 0;
    clampView(v->s);
    return 0;
}

int bhViewerResize(BhViewer* v, int width, int height, int swapRB)
{
    if (!v || width < 1 || height < 1) return -1;
    v->width = width;
    v->height = height;
    v->swapRB = swapRB ? 1 : 0;
    return 0;
}

void bhViewerReset(BhViewer* v)
{
    if (!v) return;
    v->s.azimuth = kDefaultAzimuth;
    v->s.elevation = kDefaultElevation;
    v->s.distance = kDefaultDistance;
    v->s.oriX = v->s.oriY = v->s.oriZ = 0.f;
    clampView(v->s);
}

void bhViewerOrbit(BhViewer* v, float dxPixels, float dyPixels)
{
    if (!v) return;
    v->s.azimuth += dxPixels * 0.005f;
    v->s.elevation += dyPixels * 0.005f;
    clampView(v->s);
}

void bhViewerZoom(BhViewer* v, float wheelSteps)
{
    if (!v) return;
    v->s.distance *= std::pow(0.92f, wheelSteps);
    clampView(v->s);
}

int bhViewerKey(BhViewer* v, int key, int action)
{
    if (!v || (action != BH_KEY_PRESS && action != BH_KEY_REPEAT)) return 0;
    BhViewState& s = v->s;
    // repeatable knobs
    switch (key)
    {
    case '[': s.aStar -= 0.05f; v->dirty = true; return 1;
    case ']': s.aStar += 0.05f; v->dirty = true; return 1;
    case ',': s.q -= 0.05f; v->dirty = true; return 1;
    case '.': s.q += 0.05f; v->dirty = true; return 1;
    default: break;
    }
    if (action != BH_KEY_PRESS) return 0;
    switch (key)
    {
    case BH_KEY_F2: s.model = BH_SCHWARZSCHILD; v->dirty = true; return 1;
    case BH_KEY_F3: s.model = BH_REISSNER_NORDSTROM; v->dirty = true; return 1;
    case BH_KEY_F4: s.model = BH_KERR; v->dirty = true; return 1;
    case BH_KEY_F5: s.model = BH_KERR_NEWMAN; v->dirty = true; return 1;
    case BH_KEY_F1: s.diskOn = !s.diskOn; return 1;
    case 'B': case 'b': s.bloomOn = !s.bloomOn; return 1;
    case 'N': case 'n': s.allowNaked = !s.allowNaked; v->dirty = true; clampView(s); return 1;
    case 'H': case 'h': s.hotSpotsOn = !s.hotSpotsOn; return 1;
    case 'R': case 'r': bhViewerReset(v); return 1;
    case ' ': s.animPaused = !s.animPaused; return 1;
    case '1': s.dPhi = 0.020f; s.maxSteps = 1200; return 1;
    case '2': s.dPhi = 0.012f; s.maxSteps = 2000; return 1;
    case '3': s.dPhi = 0.007f; s.maxSteps = 3500; return 1;
    default: return 0;
    }
}

int bhViewerFrame(BhViewer* v, float dt, unsigned held, RenderParams* out)
{
    if (!v || !out) return -1;
    BhViewState& s = v->s;
    if (!(dt >= 0.f)) dt = 0.f;
    if (dt > 0.1f) dt = 0.1f;
    // held keys (main.cpp:311-337)
    if (held & BH_HELD_W) s.distance *= std::pow(0.5f, dt);
    if (held & BH_HELD_S) s.distance *= std::pow(2.0f, dt);
    if (held & BH_HELD_A) s.azimuth -= 1.2f * dt;
    if (held & BH_HELD_D) s.azimuth += 1.2f * dt;
    if (held & BH_HELD_Q) s.elevation -= 0.9f * dt;
    if (held & BH_HELD_E) s.elevation += 0.9f * dt;
    if (held & BH_HELD_EXPOSURE_DOWN) s.exposure *= std::pow(0.4f, dt);
    if (held & BH_HELD_EXPOSURE_UP) s.exposure *= std::pow(2.5f, dt);
    const float rate = 1.1f * dt, sgn = (held & BH_HELD_SHIFT) ? -1.f : 1.f;
    if (held & BH_HELD_X) s.oriX += sgn * rate;
    if (held & BH_HELD_Y) s.oriY += sgn * rate;
    if (held & BH_HELD_Z) s.oriZ += sgn * rate;
    clampView(s);
    if (!s.animPaused) s.diskTime += dt;
    if (v->dirty) deriveHole(*v);

    RenderParams p;
    bhDefaultRenderParams(&p);
    const RenderParams& h = v->hole;
    p.width = v->width;
    p.height = v->height;
    p.swapRB = v->swapRB;
    p.model = h.model;
    p.M = h.M;
    p.aSpin = h.aSpin;
    p.Qc = h.Qc;
    p.rPlus = h.rPlus;
    p.rPhoton = h.rPhoton;
    p.rErgo = h.rErgo;
    p.allowNaked = h.allowNaked;
    p.diskInner = h.diskInner;
    p.diskOuter = h.diskOuter;
    bhFillCameraOriented(&p, s.azimuth, s.elevation, s.distance, s.fovY, s.oriX, s.oriY, s.oriZ);
    p.exposure = s.exposure;
    p.diskEnabled = s.diskOn ? 1 : 0;
    p.diskTime = s.diskTime;
    p.dPhi = s.dPhi;
    p.maxSteps = s.maxSteps;
    p.bloomEnabled = s.bloomOn ? 1 : 0;
    p.hotSpotsEnabled = (s.hotSpotsOn && s.diskOn) ? 1 : 0;
    p.hotSpotStrength = 0.65f;

    // progressive accumulation restarts whenever the traced geometry changes
    const GeomKey key{p.width, p.height, p.camPos, p.camForward, p.model, p.aSpin, p.Qc, p.dPhi, p.maxSteps,
                      p.diskEnabled, p.allowNaked, p.hotSpotsEnabled, s.oriX, s.oriY, s.oriZ};
    const bool changed = !v->havePrev || !(key == v->prev);
    v->prev = key;
    v->havePrev = true;
    if (changed) s.sampleIndex = 0;
    p.sampleIndex = s.sampleIndex;
    p.accumMode = changed ? 0 : ((s.animPaused || !s.diskOn) ? 1 : 2);
    if (s.sampleIndex < 4096) ++s.sampleIndex;
    *out = p;
    return changed ? 1 : 0;
}

}  // extern "C"
