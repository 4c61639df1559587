// ---------------------------------------------------------------------------
// bhrender -- headless renderer: the reference application's frame loop
// (main.cpp:341-540) without a window.  Each frame takes its parameters from
// the viewer layer (or a named workload), renders through
// launchRenderPipeline into device buffers the tool owns, and prints one JSON
// line with the frame's sample index, accumulation mode and kernel times.
// The last frame is written as PNG or PPM (by extension) and, on request, its
// HDR accumulation buffer as PFM.
//
//   bhrender [--workload NAME [--wframe K]] [--size WxH] [--frames N] [--dt S]
//            [--model schw|rn|kerr|kn] [--spin A] [--charge Q] [--naked]
//            [--az RAD] [--el RAD] [--dist D] [--ori X,Y,Z] [--fov DEG]
//            [--quality 1|2|3] [--exposure E] [--no-disk] [--no-bloom]
//            [--no-hotspots] [--paused] [--fp64-band 0|1]
//            [--out FILE.png|FILE.ppm] [--hdr FILE.pfm]
//
// Only the C ABI of libbhtrace.so is used.
// ---------------------------------------------------------------------------
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/bh/bhtrace.h"

namespace {

struct Options
{
    std::string workload;
    int wframe = 0;
    int width = 1280, height = 720;
    int frames = 1;
    float dt = 1.f / 60.f;
    int model = -1;
    float spin = -99.f, charge = -99.f;
    bool naked = false;
    float az = -99.f, el = -99.f, dist = -99.f, fov = -99.f;
    float ori[3] = {0.f, 0.f, 0.f};
    int quality = 0;
    float exposure = -1.f;
    bool noDisk = false, noBloom = false, noHot = false, paused = false;
    int band = -1;
    std::string out, hdr;
};

int usage(const char* msg)
{
    std::fprintf(stderr, "bhrender: %s\n", msg);
    std::fprintf(stderr,
                 "usage: bhrender [--workload NAME [--wframe K]] [--size WxH] [--frames N] [--dt S]\n"
                 "                [--model schw|rn|kerr|kn] [--spin A] [--charge Q] [--naked]\n"
                 "                [--az RAD] [--el RAD] [--dist D] [--ori X,Y,Z] [--fov DEG]\n"
                 "                [--quality 1|2|3] [--exposure E] [--no-disk] [--no-bloom]\n"
                 "                [--no-hotspots] [--paused] [--fp64-band 0|1]\n"
                 "                [--out FILE.png|FILE.ppm] [--hdr FILE.pfm]\n");
    return 2;
}

bool parse(int argc, char** argv, Options& o, const char*& err)
{
    for (int i = 1; i < argc; ++i)
    {
        const std::string a = argv[i];
        auto val = [&](const char*& v) {
            if (i + 1 >= argc) { err = "missing value"; return false; }
            v = argv[++i];
            return true;
        };
        const char* v = nullptr;
        if (a == "--workload") { if (!val(v)) return false; o.workload = v; }
        else if (a == "--wframe") { if (!val(v)) return false; o.wframe = std::atoi(v); }
        else if (a == "--size")
        {
            if (!val(v) || std::sscanf(v, "%dx%d", &o.width, &o.height) != 2 || o.width < 1 || o.height < 1)
            { err = "bad --size"; return false; }
        }
        else if (a == "--frames") { if (!val(v)) return false; o.frames = std::atoi(v); }
        else if (a == "--dt") { if (!val(v)) return false; o.dt = (float)std::atof(v); }
        else if (a == "--model")
        {
            if (!val(v)) return false;
            const std::string m = v;
            o.model = m == "schw" ? BH_SCHWARZSCHILD : m == "rn" ? BH_REISSNER_NORDSTROM
                    : m == "kerr" ? BH_KERR : m == "kn" ? BH_KERR_NEWMAN : -2;
            if (o.model == -2) { err = "bad --model"; return false; }
        }
        else if (a == "--spin") { if (!val(v)) return false; o.spin = (float)std::atof(v); }
        else if (a == "--charge") { if (!val(v)) return false; o.charge = (float)std::atof(v); }
        else if (a == "--naked") o.naked = true;
        else if (a == "--az") { if (!val(v)) return false; o.az = (float)std::atof(v); }
        else if (a == "--el") { if (!val(v)) return false; o.el = (float)std::atof(v); }
        else if (a == "--dist") { if (!val(v)) return false; o.dist = (float)std::atof(v); }
        else if (a == "--fov") { if (!val(v)) return false; o.fov = (float)std::atof(v); }
        else if (a == "--ori")
        {
            if (!val(v) || std::sscanf(v, "%f,%f,%f", &o.ori[0], &o.ori[1], &o.ori[2]) != 3)
            { err = "bad --ori"; return false; }
        }
        else if (a == "--quality") { if (!val(v)) return false; o.quality = std::atoi(v); }
        else if (a == "--exposure") { if (!val(v)) return false; o.exposure = (float)std::atof(v); }
        else if (a == "--no-disk") o.noDisk = true;
        else if (a == "--no-bloom") o.noBloom = true;
        else if (a == "--no-hotspots") o.noHot = true;
        else if (a == "--paused") o.paused = true;
        else if (a == "--fp64-band") { if (!val(v)) return false; o.band = std::atoi(v); }
        else if (a == "--out") { if (!val(v)) return false; o.out = v; }
        else if (a == "--hdr") { if (!val(v)) return false; o.hdr = v; }
        else { err = "unknown option"; return false; }
    }
    if (o.frames < 1) { err = "--frames must be >= 1"; return false; }
    if (o.quality < 0 || o.quality > 3) { err = "bad --quality"; return false; }
    return true;
}

bool endsWith(const std::string& s, const char* suf)
{
    const size_t n = std::strlen(suf);
    return s.size() >= n && s.compare(s.size() - n, n, suf) == 0;
}

#define BH_TRY(call)                                                                        \
    do                                                                                      \
    {                                                                                       \
        const bh_error_t e_ = (call);                                                       \
        if (e_ != 0)                                                                        \
        {                                                                                   \
            std::fprintf(stderr, "bhrender: %s failed: %s\n", #call, bhErrorString(e_));    \
            return 1;                                                                       \
        }                                                                                   \
    } while (0)

}  // namespace

int main(int argc, char** argv)
{
    Options o;
    const char* err = nullptr;
    if (!parse(argc, argv, o, err)) return usage(err);
    if (bhDeviceCount() < 1) { std::fprintf(stderr, "bhrender: no GPU\n"); return 1; }

    if (o.band >= 0)
    {
        BhPrecision pr;
        bhGetPrecision(&pr);
        pr.fp64Band = o.band;
        BH_TRY(bhSetPrecision(&pr));
    }

    // parameter producer: a named workload, or the viewer
    RenderParams P;
    BhViewer* viewer = nullptr;
    if (!o.workload.empty())
    {
        if (bhWorkloadParams(o.workload.c_str(), o.wframe, &P) != 0) return usage("unknown workload or frame");
        o.width = P.width;
        o.height = P.height;
    }
    else
    {
        if (bhViewerCreate(o.width, o.height, &viewer) != 0) return usage("bad size");
        BhViewState s;
        bhViewerGetState(viewer, &s);
        if (o.model >= 0) s.model = o.model;
        if (o.spin > -99.f) s.aStar = o.spin;
        if (o.charge > -99.f) s.q = o.charge;
        if (o.naked) s.allowNaked = 1;
        if (o.az > -99.f) s.azimuth = o.az;
        if (o.el > -99.f) s.elevation = o.el;
        if (o.dist > -99.f) s.distance = o.dist;
        if (o.fov > -99.f) s.fovY = o.fov;
        s.oriX = o.ori[0];
        s.oriY = o.ori[1];
        s.oriZ = o.ori[2];
        if (o.exposure > 0.f) s.exposure = o.exposure;
        if (o.noDisk) s.diskOn = 0;
        if (o.noBloom) s.bloomOn = 0;
        if (o.noHot) s.hotSpotsOn = 0;
        if (o.paused) s.animPaused = 1;
        bhViewerSetState(viewer, &s);
        if (o.quality) bhViewerKey(viewer, '0' + o.quality, BH_KEY_PRESS);
    }

    const size_t npx = (size_t)o.width * o.height;
    const size_t nhalf = (size_t)((o.width + 1) / 2) * ((o.height + 1) / 2);
    void *out = nullptr, *accum = nullptr, *bloomA = nullptr, *bloomB = nullptr;
    bh_stream_t st = nullptr;
    BH_TRY(bhDeviceAlloc(&out, npx * sizeof(bh_uchar4)));
    BH_TRY(bhDeviceAlloc(&accum, npx * sizeof(bh_float4)));
    BH_TRY(bhDeviceAlloc(&bloomA, nhalf * sizeof(bh_float4)));
    BH_TRY(bhDeviceAlloc(&bloomB, nhalf * sizeof(bh_float4)));
    BH_TRY(bhDeviceFill(accum, 0, npx * sizeof(bh_float4)));
    BH_TRY(bhStreamCreate(&st));
    bhEnableKernelTimers(1);

    for (int f = 0; f < o.frames; ++f)
    {
        int restarted = 0;
        if (viewer) restarted = bhViewerFrame(viewer, o.dt, 0u, &P);
        else
        {
            restarted = f == 0;
            P.sampleIndex = f;
            P.accumMode = f == 0 ? 0 : 1;
        }
        BH_TRY(launchRenderPipeline((bh_uchar4*)out, (bh_float4*)accum, (bh_float4*)bloomA, (bh_float4*)bloomB, P,
                                    st));
        float tr = 0.f, sh = 0.f, po = 0.f;
        BH_TRY(bhLastKernelTimes(&tr, &sh, &po));
        std::printf("{\"frame\": %d, \"restart\": %d, \"sampleIndex\": %d, \"accumMode\": %d, \"diskTime\": %.6f, "
                    "\"trace_ms\": %.3f, \"shade_ms\": %.3f, \"post_ms\": %.3f}\n",
                    f, restarted, P.sampleIndex, P.accumMode, P.diskTime, tr, sh, po);
        std::fflush(stdout);
    }
    BH_TRY(bhStreamSynchronize(st));

    int rc = 0;
    if (!o.out.empty())
    {
        std::vector<bh_uchar4> host(npx);
        BH_TRY(bhCopyToHost(host.data(), out, npx * sizeof(bh_uchar4)));
        const int w = endsWith(o.out, ".png") ? bhWritePNG(o.out.c_str(), host.data(), o.width, o.height)
                                              : bhWritePPM(o.out.c_str(), host.data(), o.width, o.height);
        if (w != 0) { std::fprintf(stderr, "bhrender: writing %s failed\n", o.out.c_str()); rc = 1; }
    }
    if (!o.hdr.empty())
    {
        std::vector<bh_float4> host(npx);
        BH_TRY(bhCopyToHost(host.data(), accum, npx * sizeof(bh_float4)));
        if (bhWritePFM(o.hdr.c_str(), host.data(), o.width, o.height) != 0)
        {
            std::fprintf(stderr, "bhrender: writing %s failed\n", o.hdr.c_str());
            rc = 1;
        }
    }
    bhStreamDestroy(st);
    bhDeviceFree(out);
    bhDeviceFree(accum);
    bhDeviceFree(bloomA);
    bhDeviceFree(bloomB);
    bhViewerDestroy(viewer);
    return rc;
}
