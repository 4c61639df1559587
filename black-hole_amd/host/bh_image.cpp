// ---------------------------------------------------------------------------
// bh_image.cpp -- headless output formats and image comparison.
//
// Replaces the reference's presentation path (the Vulkan interop buffer and
// the per-swapchain-image buffer-to-image copy, vulkan_context.cpp:446-484
// and 529-588; cuda_interop.cpp:61-74) with files a check can read:
//   PPM  (P6, 8-bit RGB), PNG (8-bit RGBA, zlib stream of stored deflate
//   blocks, CRC-32 and Adler-32 written here), PFM (32-bit float RGB HDR,
//   little-endian, rows bottom to top as the format defines).
// Plus per-channel LSB statistics of two uchar4 images and exit-direction
// angle statistics of two ray-record buffers.  Host memory only.
// ---------------------------------------------------------------------------
#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <vector>

#include "../../include/bh/bhtrace.h"

namespace {

uint32_t crc32Update(uint32_t crc, const unsigned char* p, size_t n)
{
    static uint32_t table[256];
    static bool init = false;
    if (!init)
    {
        for (uint32_t i = 0; i < 256; ++i)
        {
            uint32_t c = i;
            for (int k = 0; k < 8; ++k) c = (c & 1) ? 0xEDB88320u ^ (c >> 1) : c >> 1;
            table[i] = c;
        }
        init = true;
    }
    crc = ~crc;
    for (size_t i = 0; i < n; ++i) crc = table[(crc ^ p[i]) & 0xFFu] ^ (crc >> 8);
    return ~crc;
}

void put32(std::vector<unsigned char>& v, uint32_t x)
{
    v.push_back((unsigned char)(x >> 24));
    v.push_back((unsigned char)(x >> 16));
    v.push_back((unsigned char)(x >> 8));
    v.push_back((unsigned char)x);
}

bool writeChunk(FILE* f, const char* type, const std::vector<unsigned char>& data)
{
    std::vector<unsigned char> buf;
    put32(buf, (uint32_t)data.size());
    buf.insert(buf.end(), type, type + 4);
    buf.insert(buf.end(), data.begin(), data.end());
    put32(buf, crc32Update(0, buf.data() + 4, buf.size() - 4));
    return std::fwrite(buf.data(), 1, buf.size(), f) == buf.size();
}

}  // namespace

extern "C" {

int bhWritePPM(const char* path, const bh_uchar4* rgba, int w, int h)
{
    if (!path || !rgba || w < 1 || h < 1) return -1;
    FILE* f = std::fopen(path, "wb");
    if (!f) return -1;
    std::fprintf(f, "P6\n%d %d\n255\n", w, h);
    std::vector<unsigned char> row((size_t)w * 3);
    bool ok = true;
    for (int y = 0; y < h && ok; ++y)
    {
        for (int x = 0; x < w; ++x)
        {
            const bh_uchar4& p = rgba[(size_t)y * w + x];
            row[(size_t)x * 3 + 0] = p.x;
            row[(size_t)x * 3 + 1] = p.y;
            row[(size_t)x * 3 + 2] = p.z;
        }
        ok = std::fwrite(row.data(), 1, row.size(), f) == row.size();
    }
    return (std::fclose(f) == 0 && ok) ? 0 : -1;
}

int bhWritePNG(const char* path, const bh_uchar4* rgba, int w, int h)
{
    if (!path || !rgba || w < 1 || h < 1) return -1;
    // raw scanlines: filter byte 0 + RGBA
    const size_t rowBytes = (size_t)w * 4 + 1;
    std::vector<unsigned char> raw(rowBytes * (size_t)h);
    for (int y = 0; y < h; ++y)
    {
        unsigned char* r = raw.data() + rowBytes * (size_t)y;
        r[0] = 0;
        std::memcpy(r + 1, rgba + (size_t)y * w, (size_t)w * 4);
    }
    // zlib stream: header, stored deflate blocks of <= 65535 bytes, Adler-32
    std::vector<unsigned char> z;
    z.push_back(0x78);
    z.push_back(0x01);
    size_t off = 0;
    do
    {
        const size_t n = std::min<size_t>(65535, raw.size() - off);
        const bool last = off + n == raw.size();
        z.push_back(last ? 1 : 0);
        z.push_back((unsigned char)(n & 0xFF));
        z.push_back((unsigned char)(n >> 8));
        z.push_back((unsigned char)(~n & 0xFF));
        z.push_back((unsigned char)((~n >> 8) & 0xFF));
        z.insert(z.end(), raw.begin() + (long)off, raw.begin() + (long)(off + n));
        off += n;
    } while (off < raw.size());
    uint32_t a = 1, b = 0;
    for (unsigned char c : raw)
    {
        a = (a + c) % 65521u;
        b = (b + a) % 65521u;
    }
    put32(z, (b << 16) | a);
    FILE* f = std::fopen(path, "wb");
    if (!f) return -1;
    static const unsigned char sig[8] = {0x89, 'P', 'N', 'G', '\r', '\n', 0x1A, '\n'};
    bool ok = std::fwrite(sig, 1, 8, f) == 8;
    std::vector<unsigned char> ihdr;
    put32(ihdr, (uint32_t)w);
    put32(ihdr, (uint32_t)h);
    ihdr.push_back(8);   // bit depth
    ihdr.push_back(6);   // RGBA
    ihdr.push_back(0);
    ihdr.push_back(0);
    ihdr.push_back(0);
    ok = ok && writeChunk(f, "IHDR", ihdr) && writeChunk(f, "IDAT", z) && writeChunk(f, "IEND", {});
    return (std::fclose(f) == 0 && ok) ? 0 : -1;
}

int bhWritePFM(const char* path, const bh_float4* hdr, int w, int h)
{
    if (!path || !hdr || w < 1 || h < 1) return -1;
    FILE* f = std::fopen(path, "wb");
    if (!f) return -1;
    std::fprintf(f, "PF\n%d %d\n-1.0\n", w, h);   // negative scale: little-endian
    std::vector<float> row((size_t)w * 3);
    bool ok = true;
    for (int y = h - 1; y >= 0 && ok; --y)
    {
        for (int x = 0; x < w; ++x)
        {
            const bh_float4& p = hdr[(size_t)y * w + x];
            row[(size_t)x * 3 + 0] = p.x;
            row[(size_t)x * 3 + 1] = p.y;
            row[(size_t)x * 3 + 2] = p.z;
        }
        ok = std::fwrite(row.data(), sizeof(float), row.size(), f) == row.size();
    }
    return (std::fclose(f) == 0 && ok) ? 0 : -1;
}

int bhImageDiff(const bh_uchar4* a, const bh_uchar4* b, int w, int h, BhDiffStats* out)
{
    if (!a || !b || !out || w < 1 || h < 1) return -1;
    std::memset(out, 0, sizeof(*out));
    const size_t n = (size_t)w * h;
    for (size_t i = 0; i < n; ++i)
    {
        const unsigned char pa[3] = {a[i].x, a[i].y, a[i].z}, pb[3] = {b[i].x, b[i].y, b[i].z};
        int m = 0;
        for (int c = 0; c < 3; ++c)
        {
            const int d = std::abs((int)pa[c] - (int)pb[c]);
            if (d > out->maxDiff[c]) out->maxDiff[c] = d;
            if (d > m) m = d;
        }
        // histogram of the per-pixel max channel difference: 0, 1, 2, 3, 4-7, 8-15, 16-63, 64+
        const int bin = m <= 3 ? m : (m <= 7 ? 4 : (m <= 15 ? 5 : (m <= 63 ? 6 : 7)));
        out->hist[bin] += 1;
    }
    out->pixels = (long long)n;
    out->fracWithin1 = (double)(out->hist[0] + out->hist[1]) / (double)n;
    return 0;
}

int bhExitAngleDiff(const BhRayRecord* a, const BhRayRecord* b, int n, float* anglesOut, BhAngleStats* out)
{
    if (!a || !b || !out || n < 0) return -1;
    std::memset(out, 0, sizeof(*out));
    std::vector<double> ang;
    ang.reserve((size_t)n);
    for (int i = 0; i < n; ++i)
    {
        const bool ea = (a[i].flags & 1) != 0, eb = (b[i].flags & 1) != 0;
        float angle = -1.f;   // -1: not both escaped
        if (ea != eb) out->outcomeMismatch += 1;
        else if (ea)
        {
            double u[3], v[3], nu = 0, nv = 0;
            for (int k = 0; k < 3; ++k)
            {
                u[k] = a[i].escDir[k];
                v[k] = b[i].escDir[k];
                nu += u[k] * u[k];
                nv += v[k] * v[k];
            }
            nu = std::sqrt(nu);
            nv = std::sqrt(nv);
            double d2 = 0;
            for (int k = 0; k < 3; ++k)
            {
                const double d = u[k] / nu - v[k] / nv;
                d2 += d * d;
            }
            const double t = std::min(1.0, std::sqrt(d2) * 0.5);
            const double g = 2.0 * std::asin(t);
            angle = (float)g;
            ang.push_back(g);
        }
        if (anglesOut) anglesOut[i] = angle;
    }
    out->bothEscaped = (long long)ang.size();
    if (!ang.empty())
    {
        std::sort(ang.begin(), ang.end());
        auto q = [&](double p) { return ang[std::min(ang.size() - 1, (size_t)(p * (double)(ang.size() - 1) + 0.5))]; };
        out->median = q(0.5);
        out->p999 = q(0.999);
        out->max = ang.back();
        size_t over = 0;
        for (double g : ang) over += g > 1e-3 ? 1 : 0;
        out->over1mrad = (long long)over;
    }
    return 0;
}

}  // extern "C"
