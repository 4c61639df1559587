// ---------------------------------------------------------------------------
// bh_noalloc.hip -- launchRenderPipelineNoAlloc: launchRenderPipeline for
// callers that cannot allocate device memory (code under HIP graph capture,
// tight serving loops, harnesses that forbid hipMalloc).
//
// Every buffer a frame needs beyond the caller's (trace records, counters,
// FP64 band mask/lists/sort histogram, tile keys/order/histogram) is carved
// from one static device array, `scratch`, of kNoAllocBytes.  Each GPU the
// library runs on holds its own copy until the process ends.  Host code never
// uses the array's name as an address: its device address is looked up once
// per device with hipGetSymbolAddress and kept in g_base.
//
// Every call on a GPU shares that GPU's array, and the array keeps the last
// call's contents (the counters are cleared by kernels each frame), so calls
// that overlap on one GPU -- on different streams, or from several host
// threads -- race.
// ---------------------------------------------------------------------------
#include "bh_internal.h"
#include <mutex>

namespace {

// 96 MiB: a 1920x1080 frame needs 89.3 MiB (45 bytes per pixel and 8 per
// 8x8 tile, each buffer 256-byte aligned, plus 5 KiB of counters and
// histograms).
constexpr size_t kNoAllocBytes = size_t(96) << 20;
constexpr int kMaxDevices = 64;

__device__ __align__(256) unsigned char scratch[kNoAllocBytes];

std::mutex g_mu;
unsigned char* g_base[kMaxDevices] = {};

// Lays the frame's buffers out from `base` (or only sizes them when `fb` is
// null); returns the bytes used.
size_t layoutFrame(int W, int H, unsigned char* base, bhk::FrameBufs* fb)
{
    const size_t n = (size_t)W * (size_t)H;
    const size_t tiles = (size_t)((W + 7) / 8) * (size_t)((H + 7) / 8);
    size_t off = 0;
    auto take = [&](size_t bytes) {
        unsigned char* p = base ? base + off : nullptr;
        off += (bytes + 255) & ~size_t(255);
        return p;
    };
    unsigned char* counter = take(bhk::kFrameCounterWords * sizeof(unsigned));
    unsigned char* rec = take(n * sizeof(BhRayRecord));
    unsigned char* mask = take(n);
    unsigned char* list = take(n * sizeof(unsigned));
    unsigned char* key = take(n * sizeof(unsigned));
    unsigned char* sorted = take(n * sizeof(unsigned));
    unsigned char* bandHist = take(512 * sizeof(unsigned));
    unsigned char* tileKey = take(tiles * sizeof(unsigned));
    unsigned char* tileOrder = take(tiles * sizeof(unsigned));
    unsigned char* tileHist = take(512 * sizeof(unsigned));
    if (fb)
    {
        fb->counter = reinterpret_cast<unsigned*>(counter);
        fb->rec = reinterpret_cast<BhRayRecord*>(rec);
        fb->bandMask = mask;
        fb->bandList = reinterpret_cast<unsigned*>(list);
        fb->bandKey = reinterpret_cast<unsigned*>(key);
        fb->bandSorted = reinterpret_cast<unsigned*>(sorted);
        fb->bandHist = reinterpret_cast<unsigned*>(bandHist);
        fb->tileKey = reinterpret_cast<unsigned*>(tileKey);
        fb->tileOrder = reinterpret_cast<unsigned*>(tileOrder);
        fb->tileHist = reinterpret_cast<unsigned*>(tileHist);
    }
    return off;
}

// Device address of this device's `scratch`, looked up on the first call on
// the device (with the device current) and kept.
hipError_t scratchBase(int dev, unsigned char** out)
{
    if (dev < 0 || dev >= kMaxDevices) return hipErrorInvalidDevice;
    std::lock_guard<std::mutex> lk(g_mu);
    if (!g_base[dev])
    {
        void* p = nullptr;
        const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess) return e;
        if (!p) return hipErrorInvalidSymbol;
        g_base[dev] = static_cast<unsigned char*>(p);
    }
    *out = g_base[dev];
    return hipSuccess;
}

}  // namespace

extern "C" bh_error_t launchRenderPipelineNoAlloc(bh_uchar4* out, bh_float4* accum, bh_float4* bloomA,
                                                  bh_float4* bloomB, const RenderParams& p, bh_stream_t stream)
{
    RenderParams P = p;
    bhSanitizeRenderParams(&P);   // the sanitized size decides the layout
    if (layoutFrame(P.width, P.height, nullptr, nullptr) > kNoAllocBytes) return hipErrorInvalidValue;
    int dev = 0;
    hipError_t e = hipGetDevice(&dev);
    if (e != hipSuccess) return e;
    unsigned char* base = nullptr;
    e = scratchBase(dev, &base);
    if (e != hipSuccess) return e;
    bhk::FrameBufs fb;
    layoutFrame(P.width, P.height, base, &fb);
    return (bh_error_t)bhk::renderPipelineFixed(reinterpret_cast<uchar4*>(out), reinterpret_cast<float4*>(accum),
                                                reinterpret_cast<float4*>(bloomA), reinterpret_cast<float4*>(bloomB),
                                                p, (hipStream_t)stream, fb);
}
