// ---------------------------------------------------------------------------
// bh_band.hip -- FP64 critical-band classification of every pixel.
//
// One thread per pixel evaluates the ray's FP64 invariants and band tests
// (bh_band.h), writes the pixel's mask byte (read by the FP32 tile tracer,
// which skips band pixels) and appends band pixels to a list with one
// atomic per wavefront.  Pure FP64 arithmetic without contraction; compiled
// once.
// ---------------------------------------------------------------------------
#include "bh_internal.h"
#include "bh_surrogate_geom.h"

namespace bhk {

// Lowers tileKey[tile] to `key` (every lane of the wavefront comes in; lanes
// with tile < 0 or an all-ones key have none).  Eight consecutive lanes
// usually hold eight pixels of one tile row: such groups reduce their keys
// first and issue one atomic.
BHK_INL void tileKeyMin(unsigned* tileKey, int tile, unsigned key)
{
    const unsigned lane = __lane_id();
    const int lead = __shfl(tile, (int)(lane & ~7u));
    const unsigned long long same = __ballot(tile == lead);
    if (((same >> (lane & ~7u)) & 0xFFull) == 0xFFull)
    {
        key = min(key, (unsigned)__shfl_xor((int)key, 1));
        key = min(key, (unsigned)__shfl_xor((int)key, 2));
        key = min(key, (unsigned)__shfl_xor((int)key, 4));
        if ((lane & 7u) == 0u && tile >= 0 && key != 0xFFFFFFFFu) atomicMin(tileKey + tile, key);
    }
    else if (tile >= 0 && key != 0xFFFFFFFFu) atomicMin(tileKey + tile, key);
}

// Cost key of pixel i of a view traced again, from the RK4 step count the
// previous frame left in its record: 64 / (steps + 1) (smaller: longer).
BHK_INL unsigned stepKey(const BhRayRecord* rec, int i)
{
    int steps = reinterpret_cast<const int*>(rec + i)[3] >> 8;
    steps = steps < 0 ? 0 : (steps > (1 << 20) ? (1 << 20) : steps);
    return __float_as_uint(64.0f / (float)(steps + 1));
}

// Pixel i of A's view (every lane of the wavefront comes in; lanes past the
// image take part in the wavefront's list appends only).
BHK_INL void classifyPixel(const BandArgs& A, int i)
{
    const int W = A.P.width, H = A.P.height;
    bool band = false, cand = false;
    float x[kMlpInputs];
    int tile = -1;               // tile of this pixel (-1: none)
    unsigned key = 0xFFFFFFFFu;  // its cost key (all ones: no FP32 ray)
    unsigned bandKey = 0u;       // cost key of a band ray (0: invariants unusable)
    if (i < W * H)
    {
        const int px = i % W, py = i / W;
        const float2 jit = launchJitter(A.P, A.jitter, A.dynIndex);
        const d3 o = dk3(A.P.camPos.x, A.P.camPos.y, A.P.camPos.z);
        double delta, polar;
        const d3 dir = cameraDir64(A.P, px, py, jit);
        // Kerr-family frames with a horizon: an FP32 estimate settles the
        // rays far from the band; the rest are evaluated in FP64
        bool exact = true;
        if (A.B.kerr && A.B.rPlus > 1e-6)
        {
            float d32, p32;
            if (bandStatsF32(A.B, dir, d32, p32) && !(fabsf(d32) < (float)A.B.critBand + 0.01f) &&
                !(p32 < (float)A.B.polarBand + 1e-3f))
            {
                delta = d32;
                polar = p32;
                exact = false;
            }
        }
        const bool ok = exact ? bandStats(A.B, o, dir, delta, polar) : true;
        band = bandDecision(A.B, ok, delta, polar);
        A.mask[i] = band ? 1 : 0;
        tile = (py >> 3) * A.tilesX + (px >> 3);
        // cost key of the ray (smaller: longer), reduced to the tile's
        // smallest.  For a view traced again, the previous frame's RK4 step
        // count of the pixel, as 64 / (steps + 1); otherwise |delta|: rays
        // near the critical curve wind the longest.  (The polar statistic
        // does not predict length: among band rays its rank correlation with
        // the step count is -0.5 to +0.2, |delta|'s 0.6 to 0.99.)
        // Non-negative floats order like their bit patterns.  Keys only
        // order work: any key gives the same frame.
        unsigned cost;
        if (A.prevRec) cost = stepKey(A.prevRec, i);
        else cost = ok ? __float_as_uint((float)fabs(delta)) : 0u;
        if (!band) key = cost;
        else bandKey = cost;
        if (A.candList && !band && delta < -A.B.critBand)
        {
            // candidate of the surrogate: escaping ingoing ray inside the training domain
            const float fx = MathExact::div((float)px + jit.x, (float)W);
            const float fy = MathExact::div((float)py + jit.y, (float)H);
            const v3 d = cameraRayDir<MathExact>(A.P, fx, fy);
            v3 er, eth, eph;
            float rc, thc;
            camBasis(mk3(A.P.camPos.x, A.P.camPos.y, A.P.camPos.z), er, eth, eph, rc, thc);
            const RayFrame F = rayFrame(d, er, eth, eph, rc);
            const float aStar = A.B.kerr ? A.P.aSpin / A.P.M : 0.f, q = A.P.Qc / A.P.M;
            cand = F.psi <= A.psiMax && rc >= kSurRMin && rc <= kSurRMax && q >= 0.f && q <= 0.6f &&
                   fabsf(aStar) <= 0.99f;
            if (cand) surrogateInputs(F, aStar, q, rc, thc, x);
        }
    }
    const unsigned lane = __lane_id();
    if (A.tileKey) tileKeyMin(A.tileKey, tile, key);
    const unsigned long long m = __ballot(band);
    if (m != 0ull)
    {
        unsigned base = 0;
        if (lane == 0) base = atomicAdd(A.count, (unsigned)__popcll(m));
        base = (unsigned)__shfl((int)base, 0);
        if (band)
        {
            const unsigned slot = base + (unsigned)__popcll(m & ((1ull << lane) - 1ull));
            A.list[slot] = (unsigned)i;
            if (A.listKey) A.listKey[slot] = bandKey;
        }
    }
    const unsigned long long mc = __ballot(cand);
    if (mc != 0ull)
    {
        unsigned base = 0;
        if (lane == 0) base = atomicAdd(A.candCount, (unsigned)__popcll(mc));
        base = (unsigned)__shfl((int)base, 0);
        if (cand)
        {
            const unsigned slot = base + (unsigned)__popcll(mc & ((1ull << lane) - 1ull));
            A.candList[slot] = (unsigned)i;
            for (int k = 0; k < kMlpInputs; ++k) A.candInputs[(size_t)slot * kMlpInputs + k] = x[k];
        }
    }
}

__global__ __launch_bounds__(256) void bandClassifyKernel(BandArgs A)
{
    classifyPixel(A, blockIdx.x * blockDim.x + threadIdx.x);
}

// All views of a batch in one launch: block b belongs to the last view whose
// classifyBase <= b (views without pixels share the next view's base and
// come before it), and covers 256 of its pixels.
__global__ __launch_bounds__(256) void bandClassifyBatchKernel(const BandArgs* views, int nViews)
{
    const unsigned b = blockIdx.x;
    int lo = 0, hi = nViews - 1;
    while (lo < hi)
    {
        const int mid = (lo + hi + 1) >> 1;
        if (views[mid].classifyBase <= b) lo = mid;
        else hi = mid - 1;
    }
    const int v = __builtin_amdgcn_readfirstlane(lo);
    const BandArgs& A = views[v];
    classifyPixel(A, (int)(b - A.classifyBase) * 256 + (int)threadIdx.x);
}

// ---------------------------------------------------------------- tile order
// The FP32 tile tracer takes tiles from a queue; a tile costs as many RK4
// steps as its longest ray.  Taking the costliest tiles first keeps the last
// tiles of the frame short, so the persistent grid drains evenly.  Tiles are
// counting-sorted into 256 buckets of their key (1/16-octave steps of
// log2 key, smallest key first); order within a bucket is arbitrary, which
// changes no result (tiles are independent).
constexpr int kTileBuckets = 256;

BHK_INL int tileBucket(unsigned keyBits)
{
    const float k = __uint_as_float(keyBits);
    if (!(k < 3.0e38f)) return kTileBuckets - 1;        // no FP32 ray in the tile (or no key)
    const float b = (log2f(fmaxf(k, 1e-30f)) + 8.f) * 16.f;
    return b <= 0.f ? 0 : (b >= (float)(kTileBuckets - 2) ? kTileBuckets - 2 : (int)b);
}

// counts per bucket: a workgroup histogram in LDS, then one global atomic
// per bucket and workgroup (many tiles share a bucket)
__global__ __launch_bounds__(kTileBuckets) void tileHistKernel(const unsigned* tileKey, int n, unsigned* hist)
{
    __shared__ unsigned h[kTileBuckets];
    const int t = threadIdx.x, i = blockIdx.x * kTileBuckets + t;
    h[t] = 0u;
    __syncthreads();
    if (i < n) atomicAdd(h + tileBucket(tileKey[i]), 1u);
    __syncthreads();
    if (h[t]) atomicAdd(hist + t, h[t]);
}

// one workgroup of 256: exclusive scan of the bucket counts into the
// scatter cursors hist[256..511]
__global__ __launch_bounds__(kTileBuckets) void tileScanKernel(unsigned* hist)
{
    __shared__ unsigned s[kTileBuckets];
    const int t = threadIdx.x;
    const unsigned v = hist[t];
    s[t] = v;
    __syncthreads();
    for (int d = 1; d < kTileBuckets; d <<= 1)
    {
        const unsigned add = t >= d ? s[t - d] : 0u;
        __syncthreads();
        s[t] += add;
        __syncthreads();
    }
    hist[kTileBuckets + t] = s[t] - v;
}

// each workgroup reserves one range per bucket from the bucket cursors and
// places its tiles by their rank inside the workgroup
__global__ __launch_bounds__(kTileBuckets) void tileScatterKernel(const unsigned* tileKey, int n, unsigned* hist,
                                                                  unsigned* order)
{
    __shared__ unsigned h[kTileBuckets], base[kTileBuckets];
    const int t = threadIdx.x, i = blockIdx.x * kTileBuckets + t;
    h[t] = 0u;
    __syncthreads();
    int b = 0;
    unsigned rank = 0u;
    if (i < n)
    {
        b = tileBucket(tileKey[i]);
        rank = atomicAdd(h + b, 1u);
    }
    __syncthreads();
    if (h[t]) base[t] = atomicAdd(hist + kTileBuckets + t, h[t]);
    __syncthreads();
    if (i < n) order[base[b] + rank] = (unsigned)i;
}

// ---------------------------------------------------------------- band ray order
// The FP64 band kernel hands its rays out from a queue, 64 to a wavefront.
// Sorted longest first (smallest key: closest to the critical curve, or
// most steps in the previous frame), each wavefront holds rays of similar
// length: only the first few run as long as the longest ray, the rest end
// early and give their SIMD slots (each resident band wavefront displaces
// two wavefronts of the FP32 tile grid) back to the tile kernel.  The list is counting-sorted
// into 256 buckets of 1/8 octave of the key (2^-24 .. 2^8); order within a
// bucket is arbitrary, which changes no result (rays are independent).
constexpr int kBandBuckets = 256;

BHK_INL int bandBucket(unsigned keyBits)
{
    const float k = __uint_as_float(keyBits);
    if (!(k < 3.0e38f)) return kBandBuckets - 1;
    const float b = (log2f(fmaxf(k, 1e-30f)) + 24.f) * 8.f;
    return b <= 0.f ? 0 : (b >= (float)(kBandBuckets - 1) ? kBandBuckets - 1 : (int)b);
}

// Three launches, as for the tiles: per-workgroup LDS histograms added into
// hist[0..255], one workgroup's exclusive scan into the cursors
// hist[256..511] (tileScanKernel), and a scatter in which each workgroup
// ranks a 256-entry chunk in LDS and reserves its range of every bucket with
// one atomic.  The list length is read on the device; the grid strides over
// it.
constexpr unsigned kBandSortBlocks = 64;

__device__ void bandHist(const unsigned* count, const unsigned* key, unsigned* hist)
{
    __shared__ unsigned h[kBandBuckets];
    const int t = threadIdx.x;
    const unsigned n = *count;
    h[t] = 0u;
    __syncthreads();
    for (unsigned i = blockIdx.x * (unsigned)kBandBuckets + t; i < n; i += gridDim.x * (unsigned)kBandBuckets)
        atomicAdd(h + bandBucket(key[i]), 1u);
    __syncthreads();
    if (h[t]) atomicAdd(hist + t, h[t]);
}

__global__ __launch_bounds__(kBandBuckets) void bandHistKernel(const unsigned* count, const unsigned* key,
                                                              unsigned* hist)
{
    bandHist(count, key, hist);
}

__device__ void bandScatter(const unsigned* count, const unsigned* list, const unsigned* key, unsigned* hist,
                            unsigned* sorted)
{
    __shared__ unsigned h[kBandBuckets], base[kBandBuckets];
    const int t = threadIdx.x;
    const unsigned n = *count;
    // chunk start is the same for the whole workgroup, so every thread
    // reaches every barrier
    for (unsigned c0 = blockIdx.x * (unsigned)kBandBuckets; c0 < n; c0 += gridDim.x * (unsigned)kBandBuckets)
    {
        h[t] = 0u;
        __syncthreads();
        const unsigned i = c0 + t;
        int b = 0;
        unsigned rank = 0u;
        if (i < n)
        {
            b = bandBucket(key[i]);
            rank = atomicAdd(h + b, 1u);
        }
        __syncthreads();
        if (h[t]) base[t] = atomicAdd(hist + kBandBuckets + t, h[t]);
        __syncthreads();
        if (i < n) sorted[base[b] + rank] = list[i];
        __syncthreads();
    }
}

__global__ __launch_bounds__(kBandBuckets) void bandScatterKernel(const unsigned* count, const unsigned* list,
                                                                 const unsigned* key, unsigned* hist,
                                                                 unsigned* sorted)
{
    bandScatter(count, list, key, hist, sorted);
}

// The same sort for every view of a batch at once: blockIdx.y is the view,
// whose list, keys, histogram and sorted list are its descriptor's.
__global__ __launch_bounds__(kBandBuckets) void bandHistBatchKernel(const BandArgs* views)
{
    const BandArgs& A = views[blockIdx.y];
    bandHist(A.count, A.listKey, A.listHist);
}
__global__ __launch_bounds__(kBandBuckets) void bandScanBatchKernel(const BandArgs* views)
{
    __shared__ unsigned sh[kBandBuckets];
    unsigned* hist = views[blockIdx.y].listHist;
    const int t = threadIdx.x;
    const unsigned v = hist[t];
    sh[t] = v;
    __syncthreads();
    for (int d = 1; d < kBandBuckets; d <<= 1)
    {
        const unsigned add = t >= d ? sh[t - d] : 0u;
        __syncthreads();
        sh[t] += add;
        __syncthreads();
    }
    hist[kBandBuckets + t] = sh[t] - v;
}
__global__ __launch_bounds__(kBandBuckets) void bandScatterBatchKernel(const BandArgs* views)
{
    const BandArgs& A = views[blockIdx.y];
    bandScatter(A.count, A.list, A.listKey, A.listHist, A.sorted);
}

hipError_t launchBandSort(const BandArgs& A, hipStream_t st)
{
    hipLaunchKernelGGL(bandHistKernel, dim3(kBandSortBlocks), dim3(kBandBuckets), 0, st, A.count, A.listKey,
                       A.listHist);
    hipLaunchKernelGGL(tileScanKernel, dim3(1), dim3(kTileBuckets), 0, st, A.listHist);
    hipLaunchKernelGGL(bandScatterKernel, dim3(kBandSortBlocks), dim3(kBandBuckets), 0, st, A.count, A.list,
                       A.listKey, A.listHist, A.sorted);
    return hipGetLastError();
}

// Start of a band-pass frame in one launch: zero the frame's counter words
// (tile queue head, band list length and queue head, surrogate counters)
// and, with tile ordering, set every tile key to "no FP32 ray" and clear
// the tile histogram.
__global__ void frameResetKernel(unsigned* counter, int nWords, unsigned* tileKey, int nTiles, unsigned* hist,
                                 unsigned* listHist)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < nWords) counter[i] = 0u;
    if (listHist && i < 2 * kBandBuckets) listHist[i] = 0u;
    if (tileKey)
    {
        if (i < nTiles) tileKey[i] = 0xFFFFFFFFu;
        if (i < 2 * kTileBuckets) hist[i] = 0u;
    }
}

hipError_t launchFrameReset(unsigned* counter, int nWords, unsigned* tileKey, int nTiles, unsigned* hist,
                            unsigned* listHist, hipStream_t st)
{
    int n = nWords > 2 * kBandBuckets ? nWords : 2 * kBandBuckets;
    if (tileKey)
    {
        n = nTiles > n ? nTiles : n;
        n = 2 * kTileBuckets > n ? 2 * kTileBuckets : n;
    }
    hipLaunchKernelGGL(frameResetKernel, dim3((n + 255) / 256), dim3(256), 0, st, counter, nWords, tileKey, nTiles,
                       hist, listHist);
    return hipGetLastError();
}

hipError_t launchTileOrder(const unsigned* tileKey, int nTiles, unsigned* hist, unsigned* order, hipStream_t st)
{
    if (nTiles <= 0) return hipSuccess;
    const dim3 g((nTiles + kTileBuckets - 1) / kTileBuckets), b(kTileBuckets);
    hipLaunchKernelGGL(tileHistKernel, g, b, 0, st, tileKey, nTiles, hist);
    hipLaunchKernelGGL(tileScanKernel, dim3(1), dim3(kTileBuckets), 0, st, hist);
    hipLaunchKernelGGL(tileScatterKernel, g, b, 0, st, tileKey, nTiles, hist, order);
    return hipGetLastError();
}

// Band pass of a persistent batch.  Before the views' classifications:
// every list length to zero.  After them: each view's first 64-ray chunk
// (an exclusive scan of the chunk counts, one thread: batches have tens to
// hundreds of views), the number of chunks, per-view chunk completions and
// the queue head to zero.
__global__ void batchBandResetKernel(BatchBandArgs A)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < A.nViews) A.counts[i] = 0u;
    if (A.hist && i < A.nViews * 2 * kBandBuckets) A.hist[i] = 0u;
    if (A.tileKey && i < A.nTiles) A.tileKey[i] = 0xFFFFFFFFu;
    if (A.tileKey && i < 2 * kTileBuckets) A.tileHist[i] = 0u;
}

__global__ void batchBandScanKernel(BatchBandArgs A)
{
    if (blockIdx.x != 0 || threadIdx.x != 0) return;
    unsigned base = 0u;
    for (int v = 0; v < A.nViews; ++v)
    {
        A.chunkBase[v] = base;
        A.bandDone[v] = 0u;
        base += (A.counts[v] + (unsigned)(kBandChunkRays - 1)) / (unsigned)kBandChunkRays;
    }
    A.chunkBase[A.nViews] = base;
    *A.queue = 0u;
}

hipError_t launchBatchBandReset(const BatchBandArgs& A, hipStream_t st)
{
    if (A.nViews <= 0) return hipSuccess;
    int n = A.hist ? A.nViews * 2 * kBandBuckets : A.nViews;
    if (A.tileKey)
    {
        n = A.nTiles > n ? A.nTiles : n;
        n = 2 * kTileBuckets > n ? 2 * kTileBuckets : n;
    }
    hipLaunchKernelGGL(batchBandResetKernel, dim3((n + 255) / 256), dim3(256), 0, st, A);
    return hipGetLastError();
}

hipError_t launchBatchBandSort(const BatchBandArgs& A, hipStream_t st)
{
    if (A.nViews <= 0 || !A.hist) return hipSuccess;
    const dim3 g(kBandSortBlocks, (unsigned)A.nViews), g1(1, (unsigned)A.nViews), b(kBandBuckets);
    hipLaunchKernelGGL(bandHistBatchKernel, g, b, 0, st, A.views);
    hipLaunchKernelGGL(bandScanBatchKernel, g1, b, 0, st, A.views);
    hipLaunchKernelGGL(bandScatterBatchKernel, g, b, 0, st, A.views);
    return hipGetLastError();
}

hipError_t launchBatchBandScan(const BatchBandArgs& A, hipStream_t st)
{
    hipLaunchKernelGGL(batchBandScanKernel, dim3(1), dim3(64), 0, st, A);
    return hipGetLastError();
}

// Tile keys of a view traced again with the band pass off (no
// classification): the smallest step key over each tile's pixels.
__global__ __launch_bounds__(256) void tileKeysKernel(const BhRayRecord* rec, int W, int H, int tilesX,
                                                      unsigned* tileKey)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    int tile = -1;
    unsigned key = 0xFFFFFFFFu;
    if (i < W * H)
    {
        tile = ((i / W) >> 3) * tilesX + ((i % W) >> 3);
        key = stepKey(rec, i);
    }
    tileKeyMin(tileKey, tile, key);
}

// The same for every view of a batch (blockIdx.z = view; blocks past a
// view's pixels leave whole), into the batch's tile numbering.
__global__ __launch_bounds__(256) void tileKeysBatchKernel(const ViewDesc* views, unsigned* tileKeys)
{
    const ViewDesc& V = views[blockIdx.z];
    const int W = V.P.width, n = W * V.P.height;
    if ((int)(blockIdx.x * blockDim.x) >= n) return;
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    int tile = -1;
    unsigned key = 0xFFFFFFFFu;
    if (i < n)
    {
        tile = ((i / W) >> 3) * V.tilesX + ((i % W) >> 3);
        key = stepKey(V.rec, i);
    }
    tileKeyMin(tileKeys + V.tileBase, tile, key);
}

hipError_t launchTileKeys(const BhRayRecord* rec, int W, int H, int tilesX, unsigned* tileKey, hipStream_t st)
{
    const int n = W * H;
    if (n <= 0) return hipSuccess;
    hipLaunchKernelGGL(tileKeysKernel, dim3((n + 255) / 256), dim3(256), 0, st, rec, W, H, tilesX, tileKey);
    return hipGetLastError();
}

hipError_t launchTileKeysBatch(const ViewDesc* views, int nViews, int maxPix, unsigned* tileKeys, hipStream_t st)
{
    if (nViews <= 0 || maxPix <= 0) return hipSuccess;
    hipLaunchKernelGGL(tileKeysBatchKernel, dim3((maxPix + 255) / 256, 1, nViews), dim3(256), 0, st, views, tileKeys);
    return hipGetLastError();
}

hipError_t launchBandClassifyBatch(const BatchBandArgs& A, hipStream_t st)
{
    if (A.nViews <= 0 || A.classifyBlocks == 0u) return hipSuccess;
    hipLaunchKernelGGL(bandClassifyBatchKernel, dim3(A.classifyBlocks), dim3(256), 0, st, A.views, A.nViews);
    return hipGetLastError();
}

hipError_t launchBandClassify(const BandArgs& A, hipStream_t st)
{
    const int n = A.P.width * A.P.height;
    if (n <= 0) return hipSuccess;
    hipLaunchKernelGGL(bandClassifyKernel, dim3((n + 255) / 256), dim3(256), 0, st, A);
    return hipGetLastError();
}

}  // namespace bhk
