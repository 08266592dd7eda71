// lmx_image.hip -- saving a device film as the reference's Film_HDR::Save
// does (src/liblightmetrica/asset/film/film_hdr.cpp:47-164 SaveImage,
// :260-280 the suffix rule) without FreeImage, and reading images back for
// golden-image exchange.
//
// The pixels are encoded on the GPU in the file's row order (one thread per
// pixel: 12 bytes in, 3-12 bytes out; HBM-bound), copied to the host once
// and written there:
//   radiancehdr -- Radiance RGBE: Ward's shared-exponent encoding (v = the
//                  largest channel, v = m 2^e with m in [0.5, 1), bytes
//                  c * (m 256 / v) truncated and e + 128; zero below 1e-32),
//                  scanlines run-length encoded in the format's "new" form
//                  (2, 2, width; each channel as runs of >= 4 equal bytes
//                  and literal stretches of <= 128), flat outside widths
//                  8..32767; header "#?RADIANCE", FORMAT=32-bit_rle_rgbe,
//                  "-Y H +X W";
//   openexr     -- single-part scanline OpenEXR 2.0: channels B, G, R as
//                  32-bit float, no compression, increasing y;
//   png         -- 8-bit RGB of the reference's tone map,
//                  (int)(pow((double)v, 1 / 2.2) * 255) clamped to
//                  [0, 255], in a zlib stream of stored blocks;
//   pfm         -- no reference counterpart: Portable Float Map "PF",
//                  little-endian, rows bottom to top, the values as rendered.
// As the reference does, radiancehdr and openexr clamp negative values to
// zero, and every format but pfm stores the film's last row (the top of the
// image, film row 0 being the bottom) first.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "lmx.h"
#include "lmx_internal.h"

using namespace lmx;

namespace {

// one pixel's channels in file order: film row h-1-r for file row r
__device__ __forceinline__ const float* film_px(const float* film, int w, int h, size_t i, bool flip)
{
    const size_t r = i / (size_t)w, x = i % (size_t)w;
    const size_t fr = flip ? (size_t)h - 1 - r : r;
    return film + 3 * (fr * (size_t)w + x);
}

__global__ void k_enc_rgbe(const float* __restrict__ film, int w, int h, uint8_t* __restrict__ out)
{
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= (size_t)w * (size_t)h) return;
    const float* p = film_px(film, w, h, i, true);
    const float c0 = fmaxf(p[0], 0.0f), c1 = fmaxf(p[1], 0.0f), c2 = fmaxf(p[2], 0.0f);
    float v = c0;
    if (c1 > v) v = c1;
    if (c2 > v) v = c2;
    uchar4 o = make_uchar4(0, 0, 0, 0);
    if (v >= 1e-32f) {
        // Ward's scale frexp(v) 256 / v is exactly 2^(8 - e) for v = m 2^e,
        // a normal float for every v in [1e-32, FLT_MAX]: computed in FP32
        // as that power of two, it equals the double-precision formula bit
        // for bit
        int e;
        (void)frexpf(v, &e);
        const float sc = ldexpf(1.0f, 8 - e);
        o.x = (uint8_t)fminf(c0 * sc, 255.0f);
        o.y = (uint8_t)fminf(c1 * sc, 255.0f);
        o.z = (uint8_t)fminf(c2 * sc, 255.0f);
        o.w = (uint8_t)(e + 128);
    }
    reinterpret_cast<uchar4*>(out)[i] = o;
}

// The PNG tone map byte(v) = clamp((int)(pow((double)v, 1/2.2) 255), 0, 255)
// (film_hdr.cpp:135-137) is a non-decreasing step function of v >= 0; the
// host finds its 255 steps (t[b] = the smallest float whose byte is >= b) in
// double precision once, and the kernel, in FP32 only, counts the steps at
// or below v by binary search.  NaN and negative v (pow gives NaN, taken as
// 0) map to 0, +inf to 255.
struct ToneSteps { float t[256]; };

__device__ __forceinline__ uint8_t tonemap(float v, const ToneSteps& ts)
{
    if (!(v >= 0.0f)) return 0;
    uint32_t lo = 0, hi = 255;             // largest b in [0, 255] with t[b] <= v (t[0] = 0)
    while (lo < hi) {
        const uint32_t mid = (lo + hi + 1) >> 1;
        if (ts.t[mid] <= v) lo = mid; else hi = mid - 1;
    }
    return (uint8_t)lo;
}

__global__ void k_enc_png(const float* __restrict__ film, int w, int h, uint8_t* __restrict__ out, ToneSteps ts)
{
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= (size_t)w * (size_t)h) return;
    const float* p = film_px(film, w, h, i, true);
    out[3 * i] = tonemap(p[0], ts);
    out[3 * i + 1] = tonemap(p[1], ts);
    out[3 * i + 2] = tonemap(p[2], ts);
}

// host: the byte of the reference's formula, and its steps
int tone_byte_host(float v)
{
    const double t = std::pow((double)v, 1.0 / 2.2) * 255.0;
    const int k = t != t ? 0 : (t >= 2147483647.0 ? 2147483647 : (int)t);
    return k < 0 ? 0 : (k > 255 ? 255 : k);
}

ToneSteps tone_steps()
{
    ToneSteps ts;
    ts.t[0] = 0.0f;
    for (int b = 1; b < 256; b++) {
        // non-negative floats order as their bit patterns: bisect over
        // [0, +inf] for the first pattern whose byte reaches b
        uint32_t lo = 0, hi = 0x7f800000u;
        while (lo < hi) {
            const uint32_t mid = lo + (hi - lo) / 2;
            float v;
            std::memcpy(&v, &mid, 4);
            if (tone_byte_host(v) >= b) hi = mid; else lo = mid + 1;
        }
        std::memcpy(&ts.t[b], &lo, 4);
    }
    return ts;
}

// openexr: per file row, the B, G, R planes (clamped at zero); pfm: rows
// bottom to top, RGB interleaved, unclamped
__global__ void k_enc_f32(const float* __restrict__ film, int w, int h, int exr, float* __restrict__ out)
{
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= (size_t)w * (size_t)h) return;
    if (exr) {
        const float* p = film_px(film, w, h, i, true);
        const size_t r = i / (size_t)w, x = i % (size_t)w;
        float* o = out + 3 * r * (size_t)w;
        o[x] = fmaxf(p[2], 0.0f);
        o[(size_t)w + x] = fmaxf(p[1], 0.0f);
        o[2 * (size_t)w + x] = fmaxf(p[0], 0.0f);
    } else {
        const float* p = film + 3 * i;
        out[3 * i] = p[0]; out[3 * i + 1] = p[1]; out[3 * i + 2] = p[2];
    }
}

// ---- host: file bodies ----------------------------------------------------

void put_u32le(std::vector<uint8_t>& b, uint32_t v) { for (int k = 0; k < 4; k++) b.push_back((uint8_t)(v >> (8 * k))); }
void put_u64le(std::vector<uint8_t>& b, uint64_t v) { for (int k = 0; k < 8; k++) b.push_back((uint8_t)(v >> (8 * k))); }
void put_str(std::vector<uint8_t>& b, const char* s) { b.insert(b.end(), s, s + std::strlen(s) + 1); }   // with the NUL

// one channel of a scanline in the new run-length form: runs of >= 4 equal
// bytes as (128 + n, byte), n <= 127; the bytes between as (n, bytes...),
// n <= 128
void rle_channel(std::vector<uint8_t>& o, const uint8_t* d, int n)
{
    int cur = 0;
    while (cur < n) {
        // the next run of >= 4 equal bytes at or after cur
        int beg = cur, run = 0;
        while (beg < n) {
            run = 1;
            while (beg + run < n && run < 127 && d[beg + run] == d[beg]) run++;
            if (run >= 4) break;
            beg += run;
        }
        if (beg >= n) { beg = n; run = 0; }
        // literals before it
        while (cur < beg) {
            const int k = std::min(128, beg - cur);
            o.push_back((uint8_t)k);
            o.insert(o.end(), d + cur, d + cur + k);
            cur += k;
        }
        if (run >= 4) {
            o.push_back((uint8_t)(128 + run));
            o.push_back(d[beg]);
            cur = beg + run;
        }
    }
}

std::vector<uint8_t> hdr_file(const uint8_t* rgbe, int w, int h)
{
    std::vector<uint8_t> b;
    char head[128];
    const int n = snprintf(head, sizeof head, "#?RADIANCE\nFORMAT=32-bit_rle_rgbe\n\n-Y %d +X %d\n", h, w);
    b.insert(b.end(), head, head + n);
    const bool rle = w >= 8 && w <= 32767;
    std::vector<uint8_t> ch(w);
    for (int r = 0; r < h; r++) {
        const uint8_t* row = rgbe + 4 * (size_t)r * (size_t)w;
        if (!rle) { b.insert(b.end(), row, row + 4 * (size_t)w); continue; }
        b.push_back(2); b.push_back(2); b.push_back((uint8_t)(w >> 8)); b.push_back((uint8_t)(w & 255));
        for (int c = 0; c < 4; c++) {
            for (int x = 0; x < w; x++) ch[x] = row[4 * x + c];
            rle_channel(b, ch.data(), w);
        }
    }
    return b;
}

std::vector<uint8_t> exr_file(const float* planes, int w, int h)
{
    std::vector<uint8_t> b;
    put_u32le(b, 20000630u);                 // magic
    put_u32le(b, 2u);                        // version 2, single-part scanline
    auto attr = [&](const char* name, const char* type, const std::vector<uint8_t>& v) {
        put_str(b, name); put_str(b, type); put_u32le(b, (uint32_t)v.size());
        b.insert(b.end(), v.begin(), v.end());
    };
    std::vector<uint8_t> v;
    for (const char* c : {"B", "G", "R"}) {  // chlist: name, pixel type 2 (FLOAT), pLinear, 3 reserved, x/y sampling 1
        put_str(v, c); put_u32le(v, 2u); v.push_back(0); v.push_back(0); v.push_back(0); v.push_back(0);
        put_u32le(v, 1u); put_u32le(v, 1u);
    }
    v.push_back(0);
    attr("channels", "chlist", v);
    attr("compression", "compression", std::vector<uint8_t>{0});
    v.clear(); put_u32le(v, 0); put_u32le(v, 0); put_u32le(v, (uint32_t)(w - 1)); put_u32le(v, (uint32_t)(h - 1));
    attr("dataWindow", "box2i", v);
    attr("displayWindow", "box2i", v);
    attr("lineOrder", "lineOrder", std::vector<uint8_t>{0});
    float one = 1.0f, zero = 0.0f;
    v.assign(4, 0); std::memcpy(v.data(), &one, 4);
    attr("pixelAspectRatio", "float", v);
    v.assign(8, 0); std::memcpy(v.data(), &zero, 4); std::memcpy(v.data() + 4, &zero, 4);
    attr("screenWindowCenter", "v2f", v);
    v.assign(4, 0); std::memcpy(v.data(), &one, 4);
    attr("screenWindowWidth", "float", v);
    b.push_back(0);                          // end of header
    const size_t line = 8 + 12 * (size_t)w;
    const uint64_t first = (uint64_t)b.size() + 8 * (uint64_t)h;
    for (int y = 0; y < h; y++) put_u64le(b, first + (uint64_t)y * line);
    for (int y = 0; y < h; y++) {
        put_u32le(b, (uint32_t)y);
        put_u32le(b, (uint32_t)(12 * w));
        const uint8_t* p = reinterpret_cast<const uint8_t*>(planes + 3 * (size_t)y * (size_t)w);
        b.insert(b.end(), p, p + 12 * (size_t)w);
    }
    return b;
}

uint32_t crc32(const uint8_t* d, size_t n, uint32_t c = 0)
{
    static uint32_t tab[256];
    static bool init = false;
    if (!init) {
        for (uint32_t i = 0; i < 256; i++) {
            uint32_t k = i;
            for (int j = 0; j < 8; j++) k = (k & 1) ? 0xEDB88320u ^ (k >> 1) : k >> 1;
            tab[i] = k;
        }
        init = true;
    }
    c = ~c;
    for (size_t i = 0; i < n; i++) c = tab[(c ^ d[i]) & 255] ^ (c >> 8);
    return ~c;
}

void put_u32be(std::vector<uint8_t>& b, uint32_t v) { for (int k = 3; k >= 0; k--) b.push_back((uint8_t)(v >> (8 * k))); }

void png_chunk(std::vector<uint8_t>& b, const char* type, const std::vector<uint8_t>& data)
{
    put_u32be(b, (uint32_t)data.size());
    const size_t at = b.size();
    b.insert(b.end(), type, type + 4);
    b.insert(b.end(), data.begin(), data.end());
    put_u32be(b, crc32(b.data() + at, 4 + data.size()));
}

std::vector<uint8_t> png_file(const uint8_t* rgb, int w, int h)
{
    std::vector<uint8_t> b = {0x89, 'P', 'N', 'G', '\r', '\n', 0x1a, '\n'};
    std::vector<uint8_t> ihdr;
    put_u32be(ihdr, (uint32_t)w); put_u32be(ihdr, (uint32_t)h);
    ihdr.push_back(8); ihdr.push_back(2); ihdr.push_back(0); ihdr.push_back(0); ihdr.push_back(0);
    png_chunk(b, "IHDR", ihdr);
    // raw scanlines (filter 0) in a zlib stream of stored deflate blocks
    const size_t stride = 3 * (size_t)w + 1;
    std::vector<uint8_t> raw(stride * (size_t)h);
    for (int y = 0; y < h; y++) {
        raw[y * stride] = 0;
        std::memcpy(&raw[y * stride + 1], rgb + 3 * (size_t)y * (size_t)w, 3 * (size_t)w);
    }
    std::vector<uint8_t> z = {0x78, 0x01};
    size_t pos = 0;
    do {
        const size_t k = std::min<size_t>(65535, raw.size() - pos);
        z.push_back(pos + k == raw.size() ? 1 : 0);
        z.push_back((uint8_t)(k & 255)); z.push_back((uint8_t)(k >> 8));
        z.push_back((uint8_t)(~k & 255)); z.push_back((uint8_t)((~k >> 8) & 255));
        z.insert(z.end(), raw.begin() + pos, raw.begin() + pos + k);
        pos += k;
    } while (pos < raw.size());
    uint32_t a = 1, bsum = 0;                // Adler-32
    for (uint8_t c : raw) { a = (a + c) % 65521u; bsum = (bsum + a) % 65521u; }
    put_u32be(z, (bsum << 16) | a);
    png_chunk(b, "IDAT", z);
    png_chunk(b, "IEND", {});
    return b;
}

std::vector<uint8_t> pfm_file(const float* rgb, int w, int h)
{
    std::vector<uint8_t> b;
    char head[64];
    const int n = snprintf(head, sizeof head, "PF\n%d %d\n-1.0\n", w, h);
    b.insert(b.end(), head, head + n);
    const uint8_t* p = reinterpret_cast<const uint8_t*>(rgb);
    b.insert(b.end(), p, p + 12 * (size_t)w * (size_t)h);
    return b;
}

int write_file(const std::string& path, const std::vector<uint8_t>& b)
{
    FILE* f = std::fopen(path.c_str(), "wb");
    if (!f) return set_error(LMX_EIO, "lmx_film_save: cannot open " + path);
    const size_t k = std::fwrite(b.data(), 1, b.size(), f);
    const int c = std::fclose(f);
    if (k != b.size() || c != 0) return set_error(LMX_EIO, "lmx_film_save: write failed: " + path);
    return LMX_OK;
}

// ---- host: readers ----------------------------------------------------------

bool read_all(const char* path, std::vector<uint8_t>& b)
{
    FILE* f = std::fopen(path, "rb");
    if (!f) return false;
    uint8_t buf[1 << 16];
    size_t k;
    while ((k = std::fread(buf, 1, sizeof buf, f)) > 0) b.insert(b.end(), buf, buf + k);
    std::fclose(f);
    return true;
}

bool ends_with(const std::string& s, const char* suf)
{
    const size_t n = std::strlen(suf);
    return s.size() >= n && s.compare(s.size() - n, n, suf) == 0;
}

int load_hdr(const std::vector<uint8_t>& b, float* rgb, int32_t* W, int32_t* H)
{
    size_t p = 0;
    auto line = [&](std::string& s) {
        s.clear();
        while (p < b.size() && b[p] != '\n') s.push_back((char)b[p++]);
        if (p >= b.size()) return false;
        p++;
        return true;
    };
    std::string s;
    if (!line(s) || (s != "#?RADIANCE" && s != "#?RGBE")) return set_error(LMX_EINVAL, "lmx_image_load: not a Radiance file");
    while (line(s) && !s.empty())
        if (s.rfind("FORMAT=", 0) == 0 && s != "FORMAT=32-bit_rle_rgbe")
            return set_error(LMX_EINVAL, "lmx_image_load: unsupported Radiance format " + s);
    int w = 0, h = 0;
    if (!line(s) || std::sscanf(s.c_str(), "-Y %d +X %d", &h, &w) != 2 || w <= 0 || h <= 0)
        return set_error(LMX_EINVAL, "lmx_image_load: unsupported Radiance resolution line");
    *W = w; *H = h;
    if (!rgb) return LMX_OK;
    std::vector<uint8_t> row(4 * (size_t)w);
    for (int r = 0; r < h; r++) {
        if (p + 4 <= b.size() && b[p] == 2 && b[p + 1] == 2 && (b[p + 2] & 0x80) == 0 && w >= 8 && w <= 32767) {
            if (((b[p + 2] << 8) | b[p + 3]) != w) return set_error(LMX_EINVAL, "lmx_image_load: bad scanline width");
            p += 4;
            for (int c = 0; c < 4; c++) {
                int x = 0;
                while (x < w) {
                    if (p >= b.size()) return set_error(LMX_EINVAL, "lmx_image_load: truncated scanline");
                    int k = b[p++];
                    if (k > 128) {
                        k -= 128;
                        if (x + k > w || p >= b.size()) return set_error(LMX_EINVAL, "lmx_image_load: bad run");
                        for (int j = 0; j < k; j++) row[4 * (x + j) + c] = b[p];
                        p++;
                    } else {
                        if (k == 0 || x + k > w || p + k > b.size()) return set_error(LMX_EINVAL, "lmx_image_load: bad literal");
                        for (int j = 0; j < k; j++) row[4 * (x + j) + c] = b[p + j];
                        p += k;
                    }
                    x += k;
                }
            }
        } else {
            if (p + 4 * (size_t)w > b.size()) return set_error(LMX_EINVAL, "lmx_image_load: truncated file");
            std::memcpy(row.data(), &b[p], 4 * (size_t)w);
            p += 4 * (size_t)w;
        }
        float* o = rgb + 3 * (size_t)(h - 1 - r) * (size_t)w;
        for (int x = 0; x < w; x++) {
            const uint8_t* e = &row[4 * x];
            const float f = e[3] ? std::ldexp(1.0f, (int)e[3] - (128 + 8)) : 0.0f;
            o[3 * x] = e[0] * f; o[3 * x + 1] = e[1] * f; o[3 * x + 2] = e[2] * f;
        }
    }
    return LMX_OK;
}

int load_pfm(const std::vector<uint8_t>& b, float* rgb, int32_t* W, int32_t* H)
{
    int w = 0, h = 0, n = 0;
    double scale = 0.0;
    char tag[3] = {0, 0, 0};
    std::string head(b.begin(), b.begin() + std::min<size_t>(b.size(), 128));
    if (std::sscanf(head.c_str(), "%2s %d %d %lf%n", tag, &w, &h, &scale, &n) != 4 || w <= 0 || h <= 0 ||
        (std::strcmp(tag, "PF") != 0 && std::strcmp(tag, "Pf") != 0))
        return set_error(LMX_EINVAL, "lmx_image_load: not a PFM file");
    size_t p = (size_t)n + 1;                // one whitespace byte ends the header
    const int nc = tag[1] == 'F' ? 3 : 1;
    if (scale > 0) return set_error(LMX_EINVAL, "lmx_image_load: big-endian PFM is not supported");
    if (p + 4 * (size_t)nc * w * h > b.size()) return set_error(LMX_EINVAL, "lmx_image_load: truncated PFM");
    *W = w; *H = h;
    if (!rgb) return LMX_OK;
    for (size_t i = 0; i < (size_t)w * h; i++)
        for (int c = 0; c < 3; c++) std::memcpy(&rgb[3 * i + c], &b[p + 4 * (nc * i + (nc == 3 ? c : 0))], 4);
    return LMX_OK;
}

int load_exr(const std::vector<uint8_t>& b, float* rgb, int32_t* W, int32_t* H)
{
    auto u32 = [&](size_t at) { uint32_t v; std::memcpy(&v, &b[at], 4); return v; };
    if (b.size() < 8 || u32(0) != 20000630u || (u32(4) & 0xff) != 2 || (u32(4) & ~0xffu) != 0)
        return set_error(LMX_EINVAL, "lmx_image_load: not a single-part scanline OpenEXR file");
    size_t p = 8;
    int32_t box[4] = {0, 0, -1, -1};
    int comp = -1, order = 0;
    std::vector<std::string> chans;
    std::vector<uint32_t> types;
    while (p < b.size() && b[p] != 0) {
        std::string name, type;
        while (p < b.size() && b[p]) name.push_back((char)b[p++]);
        p++;
        while (p < b.size() && b[p]) type.push_back((char)b[p++]);
        p++;
        if (p + 4 > b.size()) return set_error(LMX_EINVAL, "lmx_image_load: truncated OpenEXR header");
        const uint32_t sz = u32(p);
        p += 4;
        if (p + sz > b.size()) return set_error(LMX_EINVAL, "lmx_image_load: truncated OpenEXR header");
        if (name == "channels") {
            size_t q = p;
            while (q < p + sz && b[q]) {
                std::string c;
                while (b[q]) c.push_back((char)b[q++]);
                q++;
                chans.push_back(c);
                types.push_back(u32(q));
                q += 16;
            }
        } else if (name == "compression") comp = b[p];
        else if (name == "dataWindow") std::memcpy(box, &b[p], 16);
        else if (name == "lineOrder") order = b[p];
        p += sz;
    }
    p++;
    const int w = box[2] - box[0] + 1, h = box[3] - box[1] + 1;
    if (comp != 0 || order != 0 || w <= 0 || h <= 0 || chans != std::vector<std::string>{"B", "G", "R"} ||
        types != std::vector<uint32_t>{2u, 2u, 2u})
        return set_error(LMX_EINVAL, "lmx_image_load: only uncompressed FLOAT B, G, R OpenEXR scanlines are read");
    *W = w; *H = h;
    if (!rgb) return LMX_OK;
    for (int y = 0; y < h; y++) {
        uint64_t off;
        std::memcpy(&off, &b[p + 8 * (size_t)y], 8);
        if (off + 8 + 12 * (uint64_t)w > b.size() || (int32_t)u32((size_t)off) != y + box[1])
            return set_error(LMX_EINVAL, "lmx_image_load: bad OpenEXR scanline");
        const uint8_t* d = &b[(size_t)off + 8];
        float* o = rgb + 3 * (size_t)(h - 1 - y) * (size_t)w;
        for (int x = 0; x < w; x++)
            for (int c = 0; c < 3; c++) std::memcpy(&o[3 * x + (2 - c)], d + 4 * ((size_t)c * w + x), 4);
    }
    return LMX_OK;
}

}  // namespace

extern "C" {

int lmx_film_save(const float* film_rgb, int32_t width, int32_t height, int32_t type, const char* path, void* stream)
{
    if (!film_rgb || !path) return set_error(LMX_EINVAL, "lmx_film_save: NULL argument");
    if (width <= 0 || height <= 0 || (int64_t)width * height > (1ll << 30))
        return set_error(LMX_EINVAL, "lmx_film_save: width and height must be positive, W H <= 2^30");
    static const char* suffix[] = {".hdr", ".exr", ".png", ".pfm"};
    if (type < LMX_IMAGE_RADIANCEHDR || type > LMX_IMAGE_PFM) return set_error(LMX_EINVAL, "lmx_film_save: unknown type");
    const std::string file = std::string(path) + suffix[type];
    const size_t n = (size_t)width * (size_t)height;
    const size_t per = type == LMX_IMAGE_RADIANCEHDR ? 4 : (type == LMX_IMAGE_PNG ? 3 : 12);
    hipStream_t st = (hipStream_t)stream;
    void* dev = nullptr;
    hipError_t e = hipMalloc(&dev, n * per);
    if (e != hipSuccess) return hip_error(e, "lmx_film_save: device buffer");
    const int grid = (int)((n + 255) / 256);
    if (type == LMX_IMAGE_RADIANCEHDR) k_enc_rgbe<<<grid, 256, 0, st>>>(film_rgb, width, height, (uint8_t*)dev);
    else if (type == LMX_IMAGE_PNG) {
        static const ToneSteps ts = tone_steps();
        k_enc_png<<<grid, 256, 0, st>>>(film_rgb, width, height, (uint8_t*)dev, ts);
    }
    else k_enc_f32<<<grid, 256, 0, st>>>(film_rgb, width, height, type == LMX_IMAGE_OPENEXR, (float*)dev);
    std::vector<uint8_t> host(n * per);
    if ((e = hipGetLastError()) == hipSuccess &&
        (e = hipMemcpyAsync(host.data(), dev, n * per, hipMemcpyDeviceToHost, st)) == hipSuccess)
        e = hipStreamSynchronize(st);
    (void)hipFree(dev);
    if (e != hipSuccess) return hip_error(e, "lmx_film_save: encode");
    std::vector<uint8_t> body;
    if (type == LMX_IMAGE_RADIANCEHDR) body = hdr_file(host.data(), width, height);
    else if (type == LMX_IMAGE_OPENEXR) body = exr_file(reinterpret_cast<const float*>(host.data()), width, height);
    else if (type == LMX_IMAGE_PNG) body = png_file(host.data(), width, height);
    else body = pfm_file(reinterpret_cast<const float*>(host.data()), width, height);
    return write_file(file, body);
}

int lmx_image_load(const char* path, float* rgb, int32_t* width, int32_t* height)
{
    if (!path || !width || !height) return set_error(LMX_EINVAL, "lmx_image_load: NULL argument");
    std::vector<uint8_t> b;
    if (!read_all(path, b)) return set_error(LMX_EIO, std::string("lmx_image_load: cannot read ") + path);
    const std::string s(path);
    if (ends_with(s, ".hdr")) return load_hdr(b, rgb, width, height);
    if (ends_with(s, ".pfm")) return load_pfm(b, rgb, width, height);
    if (ends_with(s, ".exr")) return load_exr(b, rgb, width, height);
    return set_error(LMX_EINVAL, "lmx_image_load: the file name must end in .hdr, .pfm or .exr");
}

}  // extern "C"
