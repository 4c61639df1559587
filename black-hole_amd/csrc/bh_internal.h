// ---------------------------------------------------------------------------
// bh_internal.h -- launch descriptors shared by the host layer and the
// kernel objects of libbhtrace.so (not part of the public ABI).
// ---------------------------------------------------------------------------
#pragma once
#include <hip/hip_runtime.h>
#include <cstddef>
#include "../../include/bh/bhtrace.h"
#include "bh_rays.h"
#include "bh_band.h"
#include "bh_mlp.h"

namespace bhk {

struct TraceArgs
{
    RenderParams P;         // sanitized
    FrameK K;               // per-frame constants (filled by the launcher)
    BhRayRecord* rec;       // [W*H] per-pixel trace records
    unsigned* counter;      // ray queue head (zero before launch)
    int tilesX;             // ceil(W/8)
    int totalIds;           // ceil(W/8)*ceil(H/8)*64
    int chunk;              // max RK4 steps between refill checks
    int refillMin;          // idle lanes that trigger a refill
    int coop;               // 1: wavefront-cooperative disk sampling
    float2 jitter;          // sub-pixel sample position
    const unsigned char* bandMask;  // [W*H] 1: traced by the FP64 band kernel (or null)
    const int* dynIndex;    // device sample index overriding P.sampleIndex (or null)
    const unsigned* tileOrder;      // [tiles] queue position -> tile, longest first (or null: row-major)
};

// FP64 critical-band pass: classification of every pixel, then the listed
// rays traced in double precision.
struct BandArgs
{
    RenderParams P;         // sanitized
    BandK B;                // FP64 frame constants
    DiskParams D;           // float disk constants (as in the FP32 trace)
    BhRayRecord* rec;       // [W*H]
    unsigned char* mask;    // [W*H]
    unsigned* list;         // [W*H] band pixel indices
    unsigned* listKey;      // [W*H] cost key of each listed ray (or null: no ordering)
    unsigned* sorted;       // [W*H] the list, costliest rays first (or null: trace in list order)
    unsigned* listHist;     // [512] bucket counts and cursors of the sort (zero before classification)
    const BhRayRecord* prevRec;     // [W*H] the previous frame's records of this same view: cost keys from
                                    // their step counts (or null: keys from the band statistics)
    unsigned* count;        // list length (zero before classification)
    unsigned* queue;        // band-trace queue head (zero before the band trace)
    unsigned* tileKey;      // [tiles] cost key of each 8x8 tile (all ones before classification, or null)
    int tilesX;             // ceil(W/8)
    float2 jitter;
    const int* dynIndex;    // device sample index overriding P.sampleIndex (or null)
    // learned surrogate: escaping non-band rays inside the training domain
    // become candidates (inputs + pixel index appended; null when off)
    float* candInputs;      // [W*H x 6]
    unsigned* candList;     // [W*H]
    unsigned* candCount;
    float psiMax;
    unsigned classifyBase;  // batch: first classification block of this view (blocks of 256 pixels)
};

// Surrogate inference and routing over the candidate list: accepted rays get
// their record (predicted exit direction, no disk) and mask value 2.
struct RouteArgs
{
    RenderParams P;
    float2 jitter;
    const int* dynIndex;
    const float* inputs;    // [cap x 6]
    const unsigned* list;   // [cap]
    const unsigned* count;  // device row count (or null: `rows`)
    int rows;
    float* outRows;         // forward only: [rows x 5] head outputs, no routing (or null)
    const __bf16* W;        // network parameters (bf16 weights)
    const float* Pf;        // network parameters (f32, biases)
    int wOff[kMlpLayers], bOff[kMlpLayers];
    float sigmaMax;         // accept when the predicted exit spread is below this (rad)
    float diskProbMax;      // ... and the predicted disk-hit probability below this
    BhRayRecord* rec;
    unsigned char* mask;
    unsigned* accepted;     // accepted-ray count (zero before launch)
    float4* accList;        // [cap] accepted rays: (pixel bits, deflection outputs x2, 0)
};

struct ProbeArgs
{
    RenderParams P;
    FrameK K;               // filled by the launcher
    BhProbe* out;
    const bh_float3* dirs;  // explicit directions, or null for an nx*ny grid
    int n, nx, ny;
};

// One view of a batched launch (device-resident descriptor table).
struct ViewDesc
{
    RenderParams P;         // sanitized
    FrameK K;
    BhRayRecord* rec;       // per-pixel trace records of this view
    uchar4* out;
    float4* accum;
    float4* bloomA;
    float4* bloomB;
    const unsigned char* bandMask;  // [W*H] FP64 band pixels of this view (skipped by the FP32 trace), or null
    int tilesX;             // ceil(W/8)
    int tileBase;           // first queue index of this view
    int tileCount;          // ceil(W/8) * ceil(H/8)
    int pad;
    float2 jitter;
};

// The host stages the table and the kernels read it: one layout in both
// compilation passes (the offsets the batch kernels' scalar loads use).
static_assert(sizeof(RenderParams) == 160, "RenderParams layout");
static_assert(offsetof(ViewDesc, rec) == 160 + sizeof(FrameK) && offsetof(ViewDesc, tilesX) == offsetof(ViewDesc, rec) + 48,
              "ViewDesc layout");
static_assert(offsetof(ViewDesc, tileBase) == offsetof(ViewDesc, tilesX) + 4 &&
                  offsetof(ViewDesc, tileCount) == offsetof(ViewDesc, tilesX) + 8 &&
                  offsetof(ViewDesc, jitter) == offsetof(ViewDesc, tilesX) + 16 && sizeof(ViewDesc) % 8 == 0,
              "ViewDesc layout");

struct BatchArgs
{
    const ViewDesc* views;  // device table, ordered by tileBase
    int nViews;
    int totalTiles;
    unsigned* counter;      // tile queue head
    unsigned* viewDone;     // [nViews] finished tiles per view
    unsigned long long* stamps;   // [nViews + 1]: completion per view (the later of its FP32 tiles and its
                                  // FP64 band rays), [nViews] = start
    const unsigned* order;  // [totalTiles] queue position -> batch tile, costliest first (or null: view order)
};

// FP64 band pass of a persistent batch: every view is classified (one
// launch over all views), then one launch traces the band rays of all views
// from a queue of 64-ray chunks of the views' lists, beside the FP32 batch
// trace, which takes the batch's tiles costliest first.
struct BatchBandArgs
{
    const BandArgs* views;      // [nViews] device table: each view's FP64 constants, records, mask, list, count
    int nViews;
    unsigned* counts;           // [nViews] band list lengths (each view's BandArgs::count points here)
    unsigned* chunkBase;        // [nViews + 1] first chunk of each view's list; [nViews]: number of chunks
    unsigned* bandDone;         // [nViews] chunks finished per view
    unsigned* queue;            // chunk queue head
    unsigned long long* stamps; // [nViews] completion per view (shared with the FP32 batch trace)
    unsigned* hist;             // [nViews x 512] each view's band sort histogram (BandArgs::listHist), or null:
                                // lists traced unsorted
    unsigned classifyBlocks;    // classification blocks of all views (BandArgs::classifyBase)
    unsigned* tileKey;          // [nTiles] cost key of every tile of the batch (each view's BandArgs::tileKey
                                // points at its first tile), or null: tiles in view order
    int nTiles;
    unsigned* tileHist;         // [512] the tile sort's bucket counts and cursors
};
constexpr int kBandChunkRays = 64;

// Surrogate training set: rays of many configurations traced by the plain
// FP32 Kerr integrator (fast policy), labelled for the network.
struct DatasetArgs
{
    const FrameK* frames;   // [nCfg] per-configuration frame constants
    const RenderParams* params;   // [nCfg] (camera position, M, a, Q)
    int nCfg, raysPerCfg;   // raysPerCfg: multiple of 64
    unsigned long long seed;
    float psiMax;           // largest angle between a ray and the direction to the hole
    const BandK* bands;     // [nCfg] band constants (critical / polar band rays are flagged)
    float* inputs;          // [n x 6] normalised network inputs
    MlpTarget* targets;     // [n]
    float* escDir;          // [n x 3] exit directions of the plain trace (or null)
};

struct LaunchShape
{
    int wgSize;
    int blocks;
};

hipError_t launchTraceExact(const TraceArgs& A, const LaunchShape& S, hipStream_t st);
hipError_t launchTraceFast(const TraceArgs& A, const LaunchShape& S, hipStream_t st);
int traceOccupancyExact(int kerr, int wgSize, int coop, int tiles);
int traceOccupancyFast(int kerr, int wgSize, int coop, int tiles);
hipError_t launchBandClassify(const BandArgs& A, hipStream_t st);
// Every view of a batch classified in one launch (A.views: the device table).
hipError_t launchBandClassifyBatch(const BatchBandArgs& A, hipStream_t st);
// Tile cost keys from the step counts of a view's previous records (the band
// pass off, so no classification gives them), for one view and for every
// view of a batch (the batch's tile numbering; maxPix: the largest view).
hipError_t launchTileKeys(const BhRayRecord* rec, int W, int H, int tilesX, unsigned* tileKey, hipStream_t st);
hipError_t launchTileKeysBatch(const ViewDesc* views, int nViews, int maxPix, unsigned* tileKeys, hipStream_t st);
hipError_t launchSurrogateDataset(const DatasetArgs& A, hipStream_t st);
hipError_t launchSurrogateRoute(const RouteArgs& A, int maxRows, int numCU, hipStream_t st);

// Device view of a surrogate's network (host/bh_surrogate.cpp).
struct SurrogateNet
{
    const __bf16* W;
    const float* P;
    int wOff[kMlpLayers], bOff[kMlpLayers];
};
SurrogateNet surrogateNet(const BhSurrogate* s);

// FP64 frame constants of the band test and the FP64 integrators (host).
BandK makeBandK(const RenderParams& P, float critBand, float polarBand);
hipError_t launchBandTraceExact(const BandArgs& A, int blocks, hipStream_t st);
hipError_t launchBandTraceFast(const BandArgs& A, int blocks, hipStream_t st);
hipError_t launchProbeExact(const ProbeArgs& A, hipStream_t st);
hipError_t launchProbeFast(const ProbeArgs& A, hipStream_t st);
hipError_t launchTraceBatchExact(const BatchArgs& A, const LaunchShape& S, hipStream_t st);
// band pass of a persistent batch: list lengths zeroed (before the views'
// classifications), chunk offsets and queue set (after them), the traces
hipError_t launchBatchBandReset(const BatchBandArgs& A, hipStream_t st);
hipError_t launchBatchBandScan(const BatchBandArgs& A, hipStream_t st);
// every view's band list sorted by its cost keys, longest rays first (with hist)
hipError_t launchBatchBandSort(const BatchBandArgs& A, hipStream_t st);
hipError_t launchBandTraceBatchExact(const BatchBandArgs& A, int blocks, hipStream_t st);
hipError_t launchBandTraceBatchFast(const BatchBandArgs& A, int blocks, hipStream_t st);
hipError_t launchTraceBatchFast(const BatchArgs& A, const LaunchShape& S, hipStream_t st);
int traceBatchOccupancyExact(int wgSize);
int traceBatchOccupancyFast(int wgSize);

// post chain (bh_post.hip)
hipError_t launchResetCounter(unsigned* counter, hipStream_t st);
// records -> HDR sample -> blend into accum (pipeline) ...
hipError_t launchShadeBlend(const RenderParams& P, const BhRayRecord* rec, float4* accum, const int* dynIndex,
                            hipStream_t st);
// dynIndex[0] += 1 (end of a graph-replayed accumulation step)
hipError_t launchBumpIndex(int* dynIndex, hipStream_t st);
// dynIndex[0] = value
hipError_t launchSetIndex(int* dynIndex, int value, hipStream_t st);
// ... or records -> HDR sample -> tonemapped pixel (single pass, no bloom)
hipError_t launchShadeFinalize(const RenderParams& P, const BhRayRecord* rec, uchar4* out, hipStream_t st);
// Bloom (when P.bloomEnabled) and composite in one launch, bit-identical to
// the reference's downsample, blur and composite launches; writes only out.
hipError_t launchPostFused(const RenderParams& P, const float4* accum, uchar4* out, hipStream_t st);
// batched post chain over a device ViewDesc table (max sizes bound the grids)
hipError_t launchResetBatch(const BatchArgs& A, hipStream_t st);
// Tile order for the FP32 tile tracer from the classification's per-tile
// keys: fill (before classification), then a counting sort into `order`
// (hist: 512 counters of scratch).
hipError_t launchFrameReset(unsigned* counter, int nWords, unsigned* tileKey, int nTiles, unsigned* hist,
                            unsigned* listHist, hipStream_t st);
hipError_t launchBandSort(const BandArgs& A, hipStream_t st);
hipError_t launchTileOrder(const unsigned* tileKey, int nTiles, unsigned* hist, unsigned* order, hipStream_t st);
hipError_t launchShadeBlendBatch(const ViewDesc* views, int nViews, int maxW, int maxH, hipStream_t st);
hipError_t launchBloomBatch(const ViewDesc* views, int nViews, int maxW, int maxH, hipStream_t st);

// Every device buffer one frame of launchRenderPipeline uses beyond the
// caller's (records, counters, band pass, tile order), handed in by a caller
// that owns them instead of the library's per-stream scratch.
struct FrameBufs
{
    BhRayRecord* rec;          // [W*H]
    unsigned* counter;         // [kFrameCounterWords]
    unsigned char* bandMask;   // [W*H]
    unsigned* bandList;        // [W*H]
    unsigned* bandKey;         // [W*H]
    unsigned* bandSorted;      // [W*H]
    unsigned* bandHist;        // [512]
    unsigned* tileKey;         // [ceil(W/8)*ceil(H/8)]
    unsigned* tileOrder;       // [ceil(W/8)*ceil(H/8)]
    unsigned* tileHist;        // [512]
};
constexpr int kFrameCounterWords = 128;   // 512 bytes of queue heads and list lengths

// launchRenderPipeline on caller-owned scratch: the same kernels and results
// (the learned surrogate is not applied), the band and tile traces one after
// the other on `st`, no kernel timers, and no allocation or stream/event
// creation (host/bh_api.cpp).
hipError_t renderPipelineFixed(uchar4* out, float4* accum, float4* bloomA, float4* bloomB, const RenderParams& p,
                               hipStream_t st, const FrameBufs& fb);

}  // namespace bhk
