// ---------------------------------------------------------------------------
// bh_post.hip -- per-pixel shading and post chain (exact IEEE arithmetic),
// single-view kernels and their batched forms (one grid z-slice per view).
//
//   shade      trace record -> sky (4 star lattices + Milky-Way band, once per
//              escaped ray) -> HDR sample -> accumulation blend   (pixel order,
//              full wavefronts; trace.cuh:908-926, black_hole_kernel.cu:22-39)
//   post       bloom and composite in one launch: each workgroup rebuilds the
//              half-resolution bloom image under its output tile in LDS
//              (bright pass, horizontal and vertical 9-tap blurs) and
//              tonemaps its pixels (black_hole_kernel.cu:41-103,
//              post_process.cuh:76-148).  The batched views use the same
//              steps as two passes through the half-resolution images.
//
// HBM traffic per full-resolution pixel (float4 accum, uchar4 out):
//   shade 32 (record) + 16 (accum read) + 16 (accum write)
//   post  16 (accum read) + 4 (out).
// ---------------------------------------------------------------------------
#include "bh_common.h"
#include "bh_internal.h"

namespace bhk {

__global__ void resetCounterKernel(unsigned* c) { c[0] = 0u; }

// HDR sample of a record for a whole wavefront (every lane calls it; lanes
// past the image get an unused value): the sky through
// backgroundColorWave (bit-identical to backgroundColor per lane), the
// record's disk light plus the sky weighted by its transmittance.
BHK_INL v3 sampleFromRecordWave(const BhRayRecord* rec, int idx, bool valid, float gCam)
{
    float4 a = make_float4(0.f, 0.f, 1.f, 0.f), b = make_float4(0.f, 0.f, 0.f, 0.f);
    if (valid)
    {
        const float4* r = reinterpret_cast<const float4*>(rec + idx);
        a = r[0];
        b = r[1];
    }
    const bool escaped = valid && (__float_as_int(a.w) & 1);
    const v3 sky = backgroundColorWave(mk3(a.x, a.y, a.z), gCam, escaped);
    const v3 bg = escaped ? sky : mk3(0.f, 0.f, 0.f);
    v3 hdr;
    {
#pragma clang fp contract(off)
        hdr = mk3(b.x, b.y, b.z) + bg * b.w;
    }
    if (!finite3(hdr)) hdr = mk3(0.f, 0.f, 0.f);
    return hdr;
}

__global__ __launch_bounds__(256) void shadeBlendKernel(RenderParams P, const BhRayRecord* rec, float4* accum,
                                                        const int* dynIndex)
{
    const int idx = blockIdx.x * blockDim.x + threadIdx.x;
    const bool valid = idx < P.width * P.height;
    const v3 cur = sampleFromRecordWave(rec, idx, valid, cameraEnergyFactor(P));
    if (!valid) return;
    const int sampleIndex = dynIndex ? *dynIndex : P.sampleIndex;
    float4 pv = accum[idx];
    v3 prev = mk3(pv.x, pv.y, pv.z);
    if (!finite3(prev)) prev = cur;
    v3 c = accumBlend(prev, cur, P.accumMode, sampleIndex);
    accum[idx] = make_float4(c.x, c.y, c.z, 1.f);
}

__global__ void bumpIndexKernel(int* i) { i[0] += 1; }
__global__ void setIndexKernel(int* i, int v) { i[0] = v; }

// The batched views' shading, as shadeBlendKernel shades one view (the
// wavefront's star hits shaded together): blocks past a view's pixels leave
// whole, so every wavefront that enters has all its lanes.
__global__ __launch_bounds__(256) void shadeBlendBatchKernel(const ViewDesc* views)
{
    const ViewDesc& V = views[blockIdx.z];
    const int n = V.P.width * V.P.height;
    if ((int)(blockIdx.x * blockDim.x) >= n) return;
    const int idx = blockIdx.x * blockDim.x + threadIdx.x;
    const bool valid = idx < n;
    const v3 cur = sampleFromRecordWave(V.rec, idx, valid, cameraEnergyFactor(V.P));
    if (!valid) return;
    float4 pv = V.accum[idx];
    v3 prev = mk3(pv.x, pv.y, pv.z);
    if (!finite3(prev)) prev = cur;
    v3 c = accumBlend(prev, cur, V.P.accumMode, V.P.sampleIndex);
    V.accum[idx] = make_float4(c.x, c.y, c.z, 1.f);
}

__global__ __launch_bounds__(256) void shadeFinalizeKernel(RenderParams P, const BhRayRecord* rec, uchar4* out)
{
    const int idx = blockIdx.x * blockDim.x + threadIdx.x;
    const bool valid = idx < P.width * P.height;
    const v3 hdr = sampleFromRecordWave(rec, idx, valid, cameraEnergyFactor(P));
    if (!valid) return;
    int px = idx % P.width, py = idx / P.width;
    out[idx] = finalizePixel(hdr, mk3(0.f, 0.f, 0.f), P, px, py);
}

BHK_INL float gaussW(int k)
{
    k = k < 0 ? -k : k;
    return k == 0 ? 0.20236f : k == 1 ? 0.17925f : k == 2 ? 0.12458f : k == 3 ? 0.06794f : 0.02907f;
}

// One workgroup = one half-res row segment of kSegX outputs: 2x2 box
// average of the exposed accumulation, bright pass, then the horizontal
// 9-tap blur from an LDS row with a 4-texel clamped halo on each side.
constexpr int kSegX = 256;
BHK_INL void bloomRowSegment(const RenderParams& P, const float4* accum, float4* dstH, int bw, int bh, int seg, int y,
                             float3* row)
{
#pragma clang fp contract(off)
    const int x0 = seg * kSegX;
    const int W = P.width, H = P.height;
    const int yA = y * 2, yB = (yA + 1 < H) ? yA + 1 : yA;
    for (int t = threadIdx.x; t < kSegX + 8; t += kSegX)
    {
        int x = x0 - 4 + t;
        x = x < 0 ? 0 : (x >= bw ? bw - 1 : x);
        int xa = x * 2, xb = (xa + 1 < W) ? xa + 1 : xa;
        float4 a = accum[yA * W + xa], b = accum[yA * W + xb];
        float4 c = accum[yB * W + xa], d = accum[yB * W + xb];
        v3 avg = mk3(a.x + b.x + c.x + d.x, a.y + b.y + c.y + d.y, a.z + b.z + c.z + d.z) * 0.25f;
        v3 bp = bloomBrightPass(avg * P.exposure);
        row[t] = make_float3(bp.x, bp.y, bp.z);
    }
    __syncthreads();
    const int x = x0 + threadIdx.x;
    if (x >= bw) return;
    v3 s = mk3(0.f, 0.f, 0.f);
#pragma unroll
    for (int k = -4; k <= 4; ++k)
    {
        int xs = x + k;
        xs = xs < 0 ? 0 : (xs >= bw ? bw - 1 : xs);
        float3 v = row[xs - (x0 - 4)];
        s += mk3(v.x, v.y, v.z) * gaussW(k);
    }
    dstH[y * bw + x] = make_float4(s.x, s.y, s.z, 1.f);
}

__global__ __launch_bounds__(kSegX) void bloomBatchKernel(const ViewDesc* views)
{
    __shared__ float3 row[kSegX + 8];
    const ViewDesc& V = views[blockIdx.z];
    if (!V.P.bloomEnabled) return;
    const int bw = (V.P.width + 1) / 2, bh = (V.P.height + 1) / 2;
    if ((int)blockIdx.x * kSegX >= bw || (int)blockIdx.y >= bh) return;
    bloomRowSegment(V.P, V.accum, V.bloomB, bw, bh, blockIdx.x, blockIdx.y, row);
}

// Output tile of kTW x kTH full-resolution pixels: vertical 9-tap blur of
// the horizontally blurred half-res image into LDS, then bilinear
// upsampling and the tonemap of each pixel.  Without bloom the tile is a
// plain per-pixel tonemap.
constexpr int kTW = 64, kTH = 16;
constexpr int kHX = kTW / 2 + 2;      // half-res columns touched by the bilinear taps
constexpr int kHY = kTH / 2 + 2;      // half-res rows touched
constexpr int kBY = kHY + 8;          // rows of the horizontally blurred image needed

struct CompositeLds
{
    float3 tileH[kBY][kHX];
    float3 tileV[kHY][kHX];
};

BHK_INL void compositeTile(const RenderParams& P, const float4* accum, const float4* srcH, float4* dstV, uchar4* out,
                           int tx, int ty, CompositeLds& L)
{
#pragma clang fp contract(off)
    const int X0 = tx * kTW, Y0 = ty * kTH;
    const int bw = (P.width + 1) / 2, bh = (P.height + 1) / 2;
    const int hx0 = X0 / 2 - 1, hy0 = Y0 / 2 - 1;   // first half-res tap (unclamped)
    const int by0 = hy0 - 4;
    const bool bloom = P.bloomEnabled != 0;
    if (bloom)
    {
        for (int t = threadIdx.x; t < kBY * kHX; t += blockDim.x)
        {
            int ry = t / kHX, cx = t % kHX;
            int yy = by0 + ry, xx = hx0 + cx;
            yy = yy < 0 ? 0 : (yy >= bh ? bh - 1 : yy);
            xx = xx < 0 ? 0 : (xx >= bw ? bw - 1 : xx);
            float4 v = srcH[yy * bw + xx];
            L.tileH[ry][cx] = make_float3(v.x, v.y, v.z);
        }
        __syncthreads();
        for (int t = threadIdx.x; t < kHY * kHX; t += blockDim.x)
        {
            int ry = t / kHX, cx = t % kHX;
            int hy = hy0 + ry;
            int hyc = hy < 0 ? 0 : (hy >= bh ? bh - 1 : hy);
            v3 s = mk3(0.f, 0.f, 0.f);
#pragma unroll
            for (int k = -4; k <= 4; ++k)
            {
                int yy = hyc + k;
                yy = yy < 0 ? 0 : (yy >= bh ? bh - 1 : yy);
                float3 v = L.tileH[yy - by0][cx];
                s += mk3(v.x, v.y, v.z) * gaussW(k);
            }
            L.tileV[ry][cx] = make_float3(s.x, s.y, s.z);
            int hx = hx0 + cx;
            if (dstV && hy == hyc && hx >= 0 && hx < bw && ry >= 1 && ry <= kHY - 2 && cx >= 1 && cx <= kHX - 2)
                dstV[hy * bw + hx] = make_float4(s.x, s.y, s.z, 1.f);
        }
        __syncthreads();
    }
    for (int t = threadIdx.x; t < kTW * kTH; t += blockDim.x)
    {
        int px = X0 + (t % kTW), py = Y0 + (t / kTW);
        if (px >= P.width || py >= P.height) continue;
        v3 bl = mk3(0.f, 0.f, 0.f);
        if (bloom)
        {
            float x = ((float)px + 0.5f) * 0.5f - 0.5f;
            float y = ((float)py + 0.5f) * 0.5f - 0.5f;
            int ix = (int)floorf(x), iy = (int)floorf(y);
            float fx = x - (float)ix, fy = y - (float)iy;
            auto tap = [&](int xx, int yy) {
                xx = xx < 0 ? 0 : (xx >= bw ? bw - 1 : xx);
                yy = yy < 0 ? 0 : (yy >= bh ? bh - 1 : yy);
                float3 v = L.tileV[yy - hy0][xx - hx0];
                return mk3(v.x, v.y, v.z);
            };
            v3 c00 = tap(ix, iy), c10 = tap(ix + 1, iy), c01 = tap(ix, iy + 1), c11 = tap(ix + 1, iy + 1);
            bl = lerp3(lerp3(c00, c10, fx), lerp3(c01, c11, fx), fy);
        }
        float4 av = accum[py * P.width + px];
        out[py * P.width + px] = finalizePixel(mk3(av.x, av.y, av.z), bl, P, px, py);
    }
}

// ---------------------------------------------------------------- fused post chain
// Bloom and composite in one launch.  Each workgroup owns a kPW x kPH tile of
// output pixels and rebuilds the part of the half-resolution bloom image its
// bilinear taps need, with its blur halos, in LDS:
//   bright  2x2 box average of the exposed accumulation + bright pass at the
//           (kPHX + 8) x (kPHY + 8) half-res texels under the two blurs,
//   blurH   horizontal 9-tap of bright on the kPHX columns,
//   blurV   vertical 9-tap of blurH on the kPHX x kPHY texels,
// then upsamples and tonemaps its pixels.  Every value is formed with the
// same operations, in the same order and with the same edge clamps as the
// two-pass form above (bloomRowSegment, compositeTile), so the frame is
// bit-identical to it.  The halo texels are recomputed by neighbouring tiles
// instead of going through memory as half-resolution images: the kernel
// writes only the uchar4 frame, and the accumulation rows that neighbouring
// tiles share are re-read through the caches (a tile's halo region is about
// 2.1 times its own pixels).
constexpr int kPW = 64, kPH = 32;
constexpr int kPHX = kPW / 2 + 2;     // half-res columns under the bilinear taps
constexpr int kPHY = kPH / 2 + 2;     // half-res rows under the bilinear taps
constexpr int kPBX = kPHX + 8;        // bright-pass columns (horizontal halo)
constexpr int kPBY = kPHY + 8;        // bright-pass / blurH rows (vertical halo)

struct PostLds
{
    union
    {
        float3 bright[kPBY][kPBX];
        float3 blurV[kPHY][kPHX];
    } u;
    float3 blurH[kPBY][kPHX];
};

__global__ __launch_bounds__(256) void postFusedKernel(RenderParams P, const float4* __restrict__ accum,
                                                       uchar4* __restrict__ out)
{
#pragma clang fp contract(off)
    __shared__ PostLds L;
    const int X0 = blockIdx.x * kPW, Y0 = blockIdx.y * kPH;
    const int W = P.width, H = P.height;
    const int bw = (W + 1) / 2, bh = (H + 1) / 2;
    const int hx0 = X0 / 2 - 1, hy0 = Y0 / 2 - 1;   // first half-res tap (unclamped)
    const bool bloom = P.bloomEnabled != 0;
    if (bloom)
    {
        // bright pass at half-res (clamp(hx0 - 4 + cx), clamp(hy0 - 4 + ry))
        for (int t = threadIdx.x; t < kPBY * kPBX; t += blockDim.x)
        {
            const int ry = t / kPBX, cx = t % kPBX;
            int y = hy0 - 4 + ry, x = hx0 - 4 + cx;
            y = y < 0 ? 0 : (y >= bh ? bh - 1 : y);
            x = x < 0 ? 0 : (x >= bw ? bw - 1 : x);
            const int yA = y * 2, yB = (yA + 1 < H) ? yA + 1 : yA;
            const int xa = x * 2, xb = (xa + 1 < W) ? xa + 1 : xa;
            const float4 a = accum[yA * W + xa], b = accum[yA * W + xb];
            const float4 c = accum[yB * W + xa], d = accum[yB * W + xb];
            const v3 avg = mk3(a.x + b.x + c.x + d.x, a.y + b.y + c.y + d.y, a.z + b.z + c.z + d.z) * 0.25f;
            const v3 bp = bloomBrightPass(avg * P.exposure);
            L.u.bright[ry][cx] = make_float3(bp.x, bp.y, bp.z);
        }
        __syncthreads();
        // horizontal 9-tap at column hx0 + cx of every bright row.  Entry c
        // of a bright row holds the texel at clamp(hx0 - 4 + c), so for an
        // in-image column the taps clamp(hx0 + cx + k) are entries cx + k + 4
        // without further clamping.  Columns outside the image are computed
        // the same way but never read: the bilinear taps below read only
        // in-image texels.
        for (int t = threadIdx.x; t < kPBY * kPHX; t += blockDim.x)
        {
            const int ry = t / kPHX, cx = t % kPHX;
            v3 s = mk3(0.f, 0.f, 0.f);
#pragma unroll
            for (int k = -4; k <= 4; ++k)
            {
                const float3 v = L.u.bright[ry][cx + k + 4];
                s += mk3(v.x, v.y, v.z) * gaussW(k);
            }
            L.blurH[ry][cx] = make_float3(s.x, s.y, s.z);
        }
        __syncthreads();
        // vertical 9-tap at row hy0 + ry (overwrites the bright texels); row
        // r of blurH holds clamp(hy0 - 4 + r), so the same argument applies
        for (int t = threadIdx.x; t < kPHY * kPHX; t += blockDim.x)
        {
            const int ry = t / kPHX, cx = t % kPHX;
            v3 s = mk3(0.f, 0.f, 0.f);
#pragma unroll
            for (int k = -4; k <= 4; ++k)
            {
                const float3 v = L.blurH[ry + k + 4][cx];
                s += mk3(v.x, v.y, v.z) * gaussW(k);
            }
            L.u.blurV[ry][cx] = make_float3(s.x, s.y, s.z);
        }
        __syncthreads();
    }
    for (int t = threadIdx.x; t < kPW * kPH; t += blockDim.x)
    {
        const int px = X0 + (t % kPW), py = Y0 + (t / kPW);
        if (px >= W || py >= H) continue;
        v3 bl = mk3(0.f, 0.f, 0.f);
        if (bloom)
        {
            const float x = ((float)px + 0.5f) * 0.5f - 0.5f;
            const float y = ((float)py + 0.5f) * 0.5f - 0.5f;
            const int ix = (int)floorf(x), iy = (int)floorf(y);
            const float fx = x - (float)ix, fy = y - (float)iy;
            auto tap = [&](int xx, int yy) {
                xx = xx < 0 ? 0 : (xx >= bw ? bw - 1 : xx);
                yy = yy < 0 ? 0 : (yy >= bh ? bh - 1 : yy);
                const float3 v = L.u.blurV[yy - hy0][xx - hx0];
                return mk3(v.x, v.y, v.z);
            };
            const v3 c00 = tap(ix, iy), c10 = tap(ix + 1, iy), c01 = tap(ix, iy + 1), c11 = tap(ix + 1, iy + 1);
            bl = lerp3(lerp3(c00, c10, fx), lerp3(c01, c11, fx), fy);
        }
        const float4 av = accum[py * W + px];
        out[py * W + px] = finalizePixel(mk3(av.x, av.y, av.z), bl, P, px, py);
    }
}

__global__ __launch_bounds__(256) void compositeBatchKernel(const ViewDesc* views)
{
    __shared__ CompositeLds L;
    const ViewDesc& V = views[blockIdx.z];
    if ((int)blockIdx.x * kTW >= V.P.width || (int)blockIdx.y * kTH >= V.P.height) return;
    compositeTile(V.P, V.accum, V.bloomB, nullptr, V.out, blockIdx.x, blockIdx.y, L);
}

__global__ void resetBatchKernel(BatchArgs A)
{
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i == 0)
    {
        A.counter[0] = 0u;
        A.stamps[A.nViews] = (unsigned long long)wall_clock64();
    }
    if (i < A.nViews)
    {
        A.viewDone[i] = 0u;
        A.stamps[i] = 0ull;
    }
}

// ------------------------------------------------------------------ launchers
hipError_t launchResetCounter(unsigned* counter, hipStream_t st)
{
    hipLaunchKernelGGL(resetCounterKernel, dim3(1), dim3(1), 0, st, counter);
    return hipGetLastError();
}

static inline unsigned blocksFor(long long n, int b) { return (unsigned)((n + b - 1) / b); }

hipError_t launchShadeBlend(const RenderParams& P, const BhRayRecord* rec, float4* accum, const int* dynIndex,
                            hipStream_t st)
{
    long long n = (long long)P.width * P.height;
    hipLaunchKernelGGL(shadeBlendKernel, dim3(blocksFor(n, 256)), dim3(256), 0, st, P, rec, accum, dynIndex);
    return hipGetLastError();
}

hipError_t launchBumpIndex(int* dynIndex, hipStream_t st)
{
    hipLaunchKernelGGL(bumpIndexKernel, dim3(1), dim3(1), 0, st, dynIndex);
    return hipGetLastError();
}

hipError_t launchSetIndex(int* dynIndex, int value, hipStream_t st)
{
    hipLaunchKernelGGL(setIndexKernel, dim3(1), dim3(1), 0, st, dynIndex, value);
    return hipGetLastError();
}

hipError_t launchShadeFinalize(const RenderParams& P, const BhRayRecord* rec, uchar4* out, hipStream_t st)
{
    long long n = (long long)P.width * P.height;
    hipLaunchKernelGGL(shadeFinalizeKernel, dim3(blocksFor(n, 256)), dim3(256), 0, st, P, rec, out);
    return hipGetLastError();
}

hipError_t launchPostFused(const RenderParams& P, const float4* accum, uchar4* out, hipStream_t st)
{
    dim3 grid(blocksFor(P.width, kPW), blocksFor(P.height, kPH));
    hipLaunchKernelGGL(postFusedKernel, grid, dim3(256), 0, st, P, accum, out);
    return hipGetLastError();
}

hipError_t launchResetBatch(const BatchArgs& A, hipStream_t st)
{
    hipLaunchKernelGGL(resetBatchKernel, dim3(blocksFor(A.nViews + 1, 64)), dim3(64), 0, st, A);
    return hipGetLastError();
}

hipError_t launchShadeBlendBatch(const ViewDesc* views, int nViews, int maxW, int maxH, hipStream_t st)
{
    dim3 grid(blocksFor((long long)maxW * maxH, 256), 1, nViews);
    hipLaunchKernelGGL(shadeBlendBatchKernel, grid, dim3(256), 0, st, views);
    return hipGetLastError();
}

hipError_t launchBloomBatch(const ViewDesc* views, int nViews, int maxW, int maxH, hipStream_t st)
{
    const int bw = (maxW + 1) / 2, bh = (maxH + 1) / 2;
    hipLaunchKernelGGL(bloomBatchKernel, dim3(blocksFor(bw, kSegX), bh, nViews), dim3(kSegX), 0, st, views);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL(compositeBatchKernel, dim3(blocksFor(maxW, kTW), blocksFor(maxH, kTH), nViews), dim3(256), 0,
                       st, views);
    return hipGetLastError();
}

}  // namespace bhk
