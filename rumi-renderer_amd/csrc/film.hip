// C ABI: the film output path -- Film::Save's float-to-byte conversion
// (src/Camera/film.cpp:152-203) and FirefliesFiltering
// (src/Scene/scene.cpp:261-431) as GPU kernels over device buffers, and
// SaveBmp's header / LoadBmp's reading rule (src/Kernel/bmp.cpp:3-97) on the
// host.  Pixels are bytes B, G, R (and a fourth byte for 32-bit files) in the
// reference's buffer order.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdio>
#include <cstring>
#include <vector>

#include "internal.h"

using namespace rumi;

namespace {

constexpr int kFilmBlock = 256;

// powf(x, y) of Film::Save, correctly rounded: the double-precision power
// rounded once to float (x >= 0 here except for negative film values, which
// give NaN as powf does).
__device__ __forceinline__ float pow_cr(float x, float y) { return (float)pow((double)x, (double)y); }

// The byte of one channel: (unsigned int) of the clamped value; a NaN (a
// negative film value) becomes 0.
__device__ __forceinline__ uint32_t to_byte(float v) {
    v = v > 255.0f ? 255.0f : (v < 0.0f ? 0.0f : v);
    return v != v ? 0u : (uint32_t)v;
}

// One thread per pixel.  BGR_SUM: film holds B, G, R sums of the samples and
// is scaled by inv_n; RGB_MEAN: R, G, B means taken as they are.
template <bool BGR_SUM, int BPP>
__global__ __launch_bounds__(kFilmBlock) void k_film_encode(const float* __restrict__ film, uint32_t n_pix,
                                                            float inv_n, float inv_gamma,
                                                            uint8_t* __restrict__ out) {
    const uint32_t i = blockIdx.x * kFilmBlock + threadIdx.x;
    if (i >= n_pix) return;
    const float a = film[3 * (size_t)i + 0], b = film[3 * (size_t)i + 1], c = film[3 * (size_t)i + 2];
    float bgr[3];
    if (BGR_SUM) {
        bgr[0] = a * inv_n;
        bgr[1] = b * inv_n;
        bgr[2] = c * inv_n;
    } else {
        bgr[0] = c;
        bgr[1] = b;
        bgr[2] = a;
    }
    uint32_t px = 0;
#pragma unroll
    for (int k = 0; k < 3; ++k) px |= to_byte(pow_cr(bgr[k], inv_gamma) * 255.0f) << (8 * k);
    if (BPP == 4) {
        reinterpret_cast<uint32_t*>(out)[i] = px;
    } else {
        out[3 * (size_t)i + 0] = (uint8_t)px;
        out[3 * (size_t)i + 1] = (uint8_t)(px >> 8);
        out[3 * (size_t)i + 2] = (uint8_t)(px >> 16);
    }
}

// FirefliesFiltering over a loaded image (width w, height h, bpp bytes per
// pixel; one thread per pixel).  Its rules, kept: image row y lives in buffer
// row h - 1 - y; a pixel with no channel at 255 is copied; otherwise its
// neighbours are gathered left, right, then the three of buffer row - 1, then
// the three of buffer row + 1, each under the reference's condition (which
// tests the image row against the image's edges, not the buffer row), the
// first n_copy of them (8, 5 on an edge, 3 in a corner) are checked for a
// white (255, 255, 255) one -- then the pixel is kept -- and otherwise
// averaged in FP32 and truncated.  A gathered position outside the buffer
// (the reference reads outside its allocation there: image rows 0 and h - 1)
// reads as 0.  The fourth byte of a 32-bit pixel is written as 0.
template <int BPP>
__global__ __launch_bounds__(kFilmBlock) void k_fireflies(const uint8_t* __restrict__ src, uint32_t w, uint32_t h,
                                                          uint8_t* __restrict__ dst) {
    const uint32_t i = blockIdx.x * kFilmBlock + threadIdx.x;
    if (i >= w * h) return;
    const int x = (int)(i % w), y = (int)(i / w);
    const int W = (int)w, H = (int)h;
    const long long row = (long long)W * BPP, total = row * H;
    const long long base = (long long)(H - (y + 1)) * row + (long long)x * BPP;
    auto load = [&](long long o, float c[3]) {
        const bool in = o >= 0 && o + 2 < total;
        c[0] = in ? (float)src[o] : 0.0f;
        c[1] = in ? (float)src[o + 1] : 0.0f;
        c[2] = in ? (float)src[o + 2] : 0.0f;
    };
    auto store = [&](uint32_t b, uint32_t g, uint32_t r) {
        dst[base] = (uint8_t)b;
        dst[base + 1] = (uint8_t)g;
        dst[base + 2] = (uint8_t)r;
        if (BPP == 4) dst[base + 3] = 0;
    };
    float t[3];
    load(base, t);
    if (t[0] < 255.0f && t[1] < 255.0f && t[2] < 255.0f) {
        store((uint32_t)t[0], (uint32_t)t[1], (uint32_t)t[2]);
        return;
    }
    int n_copy = 8;
    if (x == 0 || y == 0 || x == W - 1 || y == H - 1) {
        n_copy -= 3;
        if (x == 0 && y == 0) n_copy -= 2;
        if (x == 0 && y == H - 1) n_copy -= 2;
        if (x == W - 1 && y == 0) n_copy -= 2;
        if (x == W - 1 && y == H - 1) n_copy -= 2;
    }
    const bool cond[8] = {x != 0,          x != W - 1,          y != 0 || x != 0, y != 0,
                          y != 0 || x != W - 1, y != H - 1 || x != 0, y != H - 1, y != H - 1 || x != W - 1};
    const long long off[8] = {base - BPP,       base + BPP,       base - row - BPP, base - row,
                              base - row + BPP, base + row - BPP, base + row,       base + row + BPP};
    float f[8][3];
    int n = 0;
#pragma unroll
    for (int k = 0; k < 8; ++k)
        if (cond[k]) load(off[k], f[n++]);
    bool firefly = true;
    for (int k = 0; k < n_copy && k < n; ++k)
        if (f[k][0] == 255.0f && f[k][1] == 255.0f && f[k][2] == 255.0f) firefly = false;
    if (!firefly) {
        store((uint32_t)t[0], (uint32_t)t[1], (uint32_t)t[2]);
        return;
    }
    float s[3] = {0.0f, 0.0f, 0.0f};
    for (int k = 0; k < n_copy && k < n; ++k) {
        s[0] += f[k][0];
        s[1] += f[k][1];
        s[2] += f[k][2];
    }
    const float inv = 1.0f / (float)n_copy;
    uint32_t b[3];
#pragma unroll
    for (int k = 0; k < 3; ++k) b[k] = to_byte(s[k] * inv);
    store(b[0], b[1], b[2]);
}

uint32_t blocks_for(uint64_t n) { return (uint32_t)((n + kFilmBlock - 1) / kFilmBlock); }

void put32(uint8_t* p, uint32_t v) {
    for (int k = 0; k < 4; ++k) p[k] = (uint8_t)(v >> (8 * k));
}

void put16(uint8_t* p, uint32_t v) {
    p[0] = (uint8_t)v;
    p[1] = (uint8_t)(v >> 8);
}

uint32_t get32(const uint8_t* p) { return p[0] | (p[1] << 8) | (p[2] << 16) | ((uint32_t)p[3] << 24); }

}  // namespace

extern "C" int rumi_film_encode(rumi_ctx* ctx, const float* film, uint32_t width, uint32_t height,
                                int32_t layout, uint32_t n_samples, float gamma, uint32_t bytes_per_pixel,
                                uint8_t* out, void* stream) {
    clear_error();
    if (!ctx) return fail(RUMI_ERR_NO_DEVICE, "rumi_film_encode: null context");
    if (layout != RUMI_FILM_BGR_SUM && layout != RUMI_FILM_RGB_MEAN)
        return fail(RUMI_ERR_INVALID, "rumi_film_encode: unknown layout %d", layout);
    if (bytes_per_pixel != 3 && bytes_per_pixel != 4)
        return fail(RUMI_ERR_INVALID, "rumi_film_encode: bytes_per_pixel must be 3 or 4");
    if (layout == RUMI_FILM_BGR_SUM && n_samples == 0)
        return fail(RUMI_ERR_INVALID, "rumi_film_encode: n_samples must be > 0");
    if (!(gamma > 0.0f)) return fail(RUMI_ERR_INVALID, "rumi_film_encode: gamma must be > 0");
    const uint64_t n = (uint64_t)width * height;
    if (n > 0xffffffffull / 4) return fail(RUMI_ERR_INVALID, "rumi_film_encode: image too large");
    if (n == 0) return RUMI_OK;
    if (!film || !out) return fail(RUMI_ERR_INVALID, "rumi_film_encode: null buffer");
    if (bytes_per_pixel == 4 && ((uintptr_t)out & 3u))
        return fail(RUMI_ERR_INVALID, "rumi_film_encode: 32-bit output must be 4-byte aligned");
    if (hipSetDevice(ctx->device) != hipSuccess) return fail(RUMI_ERR_HIP, "hipSetDevice failed");
    // Film::Film: _sampling_depth = (int)sqrt((float)n); Save divides by its square
    const int depth = (int)sqrtf((float)n_samples);
    const float inv_n = 1.0f / (float)(depth * depth);
    const float inv_gamma = 1.0f / gamma;
    hipStream_t s = (hipStream_t)stream;
    const bool sum = layout == RUMI_FILM_BGR_SUM;
    if (sum && bytes_per_pixel == 4)
        hipLaunchKernelGGL((k_film_encode<true, 4>), dim3(blocks_for(n)), dim3(kFilmBlock), 0, s, film,
                           (uint32_t)n, inv_n, inv_gamma, out);
    else if (sum)
        hipLaunchKernelGGL((k_film_encode<true, 3>), dim3(blocks_for(n)), dim3(kFilmBlock), 0, s, film,
                           (uint32_t)n, inv_n, inv_gamma, out);
    else if (bytes_per_pixel == 4)
        hipLaunchKernelGGL((k_film_encode<false, 4>), dim3(blocks_for(n)), dim3(kFilmBlock), 0, s, film,
                           (uint32_t)n, inv_n, inv_gamma, out);
    else
        hipLaunchKernelGGL((k_film_encode<false, 3>), dim3(blocks_for(n)), dim3(kFilmBlock), 0, s, film,
                           (uint32_t)n, inv_n, inv_gamma, out);
    const hipError_t e = hipGetLastError();
    if (e != hipSuccess) return fail(RUMI_ERR_HIP, "rumi_film_encode launch failed: %s", hipGetErrorString(e));
    return RUMI_OK;
}

extern "C" int rumi_film_fireflies(rumi_ctx* ctx, const uint8_t* src, uint32_t width, uint32_t height,
                                   uint32_t bytes_per_pixel, uint8_t* dst, void* stream) {
    clear_error();
    if (!ctx) return fail(RUMI_ERR_NO_DEVICE, "rumi_film_fireflies: null context");
    if (bytes_per_pixel != 3 && bytes_per_pixel != 4)
        return fail(RUMI_ERR_INVALID, "rumi_film_fireflies: bytes_per_pixel must be 3 or 4");
    const uint64_t n = (uint64_t)width * height;
    if (n * 4 > 0x7fffffffull) return fail(RUMI_ERR_INVALID, "rumi_film_fireflies: image too large");
    if (n == 0) return RUMI_OK;
    if (!src || !dst) return fail(RUMI_ERR_INVALID, "rumi_film_fireflies: null buffer");
    if (src == dst) return fail(RUMI_ERR_INVALID, "rumi_film_fireflies: src and dst must differ");
    if (hipSetDevice(ctx->device) != hipSuccess) return fail(RUMI_ERR_HIP, "hipSetDevice failed");
    hipStream_t s = (hipStream_t)stream;
    if (bytes_per_pixel == 4)
        hipLaunchKernelGGL(k_fireflies<4>, dim3(blocks_for(n)), dim3(kFilmBlock), 0, s, src, width, height, dst);
    else
        hipLaunchKernelGGL(k_fireflies<3>, dim3(blocks_for(n)), dim3(kFilmBlock), 0, s, src, width, height, dst);
    const hipError_t e = hipGetLastError();
    if (e != hipSuccess)
        return fail(RUMI_ERR_HIP, "rumi_film_fireflies launch failed: %s", hipGetErrorString(e));
    return RUMI_OK;
}

extern "C" int rumi_bmp_header(uint32_t width, uint32_t height, uint32_t bytes_per_pixel, uint8_t* out) {
    clear_error();
    if (!out) return fail(RUMI_ERR_INVALID, "rumi_bmp_header: null output");
    const uint64_t data = (uint64_t)width * height * bytes_per_pixel;
    if (data + 54 > 0xffffffffull) return fail(RUMI_ERR_INVALID, "rumi_bmp_header: image too large");
    std::memset(out, 0, 54);
    out[0] = 'B';
    out[1] = 'M';
    put32(out + 2, (uint32_t)(54 + data));  // file size
    put32(out + 10, 54);                     // pixel data offset
    put32(out + 14, 40);                     // info header size
    put32(out + 18, width);
    put32(out + 22, height);
    put16(out + 26, 1);                      // planes
    put16(out + 28, bytes_per_pixel * 8);    // bits per pixel
    put32(out + 34, (uint32_t)data);         // pixel data size
    return RUMI_OK;
}

extern "C" int rumi_bmp_texels(const uint8_t* pixels, uint32_t width, uint32_t height,
                               uint32_t bytes_per_pixel, float gamma, float* rgb) {
    clear_error();
    if (!pixels || !rgb) return fail(RUMI_ERR_INVALID, "rumi_bmp_texels: null buffer");
    if (bytes_per_pixel != 3 && bytes_per_pixel != 4)
        return fail(RUMI_ERR_INVALID, "rumi_bmp_texels: bytes_per_pixel must be 3 or 4");
    // ImageTexture::Pixel (imagetexture.cpp:32-61): byte * (1 / 255), raised
    // to gamma unless gamma is 1 (powf taken correctly rounded); B, G, R
    // bytes to R, G, B floats, rows as they are (row 0 = v = 0)
    const float inv = 1.0f / 255.0f;
    const uint64_t n = (uint64_t)width * height;
    for (uint64_t i = 0; i < n; ++i)
        for (int c = 0; c < 3; ++c) {
            const float x = (float)pixels[i * bytes_per_pixel + (2 - c)] * inv;
            rgb[3 * i + c] = gamma == 1.0f ? x : (float)std::pow((double)x, (double)gamma);
        }
    return RUMI_OK;
}

namespace {

// A line as fgets(buf, 256) reads it from a memory image: up to 255 bytes,
// through the first '\n'.  Returns false at the end of the data.
bool mem_line(const uint8_t* data, uint64_t n, uint64_t& at, char out[256]) {
    if (at >= n) return false;
    int k = 0;
    while (k < 255 && at < n) {
        const char c = (char)data[at++];
        out[k++] = c;
        if (c == '\n') break;
    }
    out[k] = 0;
    return true;
}

// ScanLine (hdr.cpp:114-177): a new-style run-length scanline of `w` RGBE
// pixels into four planes.  The width word is compared with char
// arithmetic, as the reference does with a signed char (a low byte of 128 or
// more never matches).
bool scan_line(const uint8_t* data, uint64_t n, uint64_t& at, int w, uint8_t* planes) {
    if (w < 8 || w > 0x7fff) return false;
    if (at + 4 > n) return false;
    const signed char h0 = (signed char)data[at], h1 = (signed char)data[at + 1];
    const signed char h2 = (signed char)data[at + 2], h3 = (signed char)data[at + 3];
    at += 4;
    if (h0 != 2 || h1 != 2 || (h2 & 0x80)) return false;
    if ((((int)h2 << 8) | (int)h3) != w) return false;
    uint8_t* p = planes;
    for (int c = 0; c < 4; ++c) {
        uint8_t* end = planes + (size_t)(c + 1) * w;
        while (p < end) {
            if (at + 2 > n) return false;
            const uint8_t a = data[at], b = data[at + 1];
            at += 2;
            if (a > 128) {
                const int count = a - 128;
                if (count > end - p) return false;
                for (int k = 0; k < count; ++k) *p++ = b;
            } else {
                const int count = a;
                if (count == 0 || count > end - p) return false;
                *p++ = b;
                if (count > 1) {
                    if (at + (uint64_t)(count - 1) > n) return false;
                    std::memcpy(p, data + at, (size_t)(count - 1));
                    at += (uint64_t)(count - 1);
                    p += count - 1;
                }
            }
        }
    }
    return true;
}

}  // namespace

extern "C" int rumi_hdr_read(const uint8_t* file, uint64_t file_bytes, uint32_t* width, uint32_t* height,
                             float* rgb, uint64_t rgb_capacity) {
    clear_error();
    if (!file) return fail(RUMI_ERR_INVALID, "rumi_hdr_read: null data");
    // LoadHDR (hdr.cpp:4-112): the header lines, then the resolution line
    char line[256];
    uint64_t at = 0;
    if (!mem_line(file, file_bytes, at, line) || std::strcmp(line, "#?RADIANCE\n") != 0)
        return fail(RUMI_ERR_INVALID, "rumi_hdr_read: no #?RADIANCE line");
    do {
        if (!mem_line(file, file_bytes, at, line)) return fail(RUMI_ERR_INVALID, "rumi_hdr_read: header ends");
    } while (line[0] == '#');
    if (std::strcmp(line, "FORMAT=32-bit_rle_rgbe\n") != 0)
        return fail(RUMI_ERR_INVALID, "rumi_hdr_read: not FORMAT=32-bit_rle_rgbe");
    for (;;) {  // EXPOSURE, GAMMA and any other line until an empty one
        if (!mem_line(file, file_bytes, at, line)) return fail(RUMI_ERR_INVALID, "rumi_hdr_read: header ends");
        if (line[0] == '\n') break;
    }
    int h = 0, w = 0;
    if (!mem_line(file, file_bytes, at, line) || std::sscanf(line, "-Y %d + X  %d", &h, &w) != 2 || h <= 0 ||
        w <= 0)
        return fail(RUMI_ERR_INVALID, "rumi_hdr_read: no -Y h +X w line");
    if (width) *width = (uint32_t)w;
    if (height) *height = (uint32_t)h;
    if (!rgb) return RUMI_OK;
    if (rgb_capacity < 3ull * (uint64_t)w * (uint64_t)h)
        return fail(RUMI_ERR_INVALID, "rumi_hdr_read: rgb buffer too small");
    std::vector<uint8_t> planes(4 * (size_t)w);
    for (int y = 0; y < h; ++y) {
        if (!scan_line(file, file_bytes, at, w, planes.data()))
            return fail(RUMI_ERR_INVALID, "rumi_hdr_read: scanline %d is not one the reference reads", y);
        for (int x = 0; x < w; ++x) {
            const int e = planes[3 * (size_t)w + x];
            float* o = rgb + 3 * ((size_t)y * w + x);
            if (e) {
                const float f = (float)std::ldexp(1.0, e - (128 + 8));
                for (int c = 0; c < 3; ++c) o[c] = (float)planes[(size_t)c * w + x] * f;
            } else {
                o[0] = o[1] = o[2] = 0.0f;
            }
        }
    }
    return RUMI_OK;
}

extern "C" int rumi_bmp_read(const uint8_t* file, uint64_t file_bytes, uint32_t* width, uint32_t* height,
                             uint32_t* bytes_per_pixel, uint8_t* pixels, uint64_t pixel_capacity) {
    clear_error();
    if (!file || file_bytes < 54) return fail(RUMI_ERR_INVALID, "rumi_bmp_read: not a BMP (too short)");
    const uint32_t w = get32(file + 18), h = get32(file + 22);
    const uint32_t bpp = (file[28] | (file[29] << 8)) / 8;
    const uint32_t offset = get32(file + 10);
    if (width) *width = w;
    if (height) *height = h;
    if (bytes_per_pixel) *bytes_per_pixel = bpp;
    if (!pixels) return RUMI_OK;
    const uint64_t row = (uint64_t)w * bpp, need = row * h;
    if (pixel_capacity < need) return fail(RUMI_ERR_INVALID, "rumi_bmp_read: pixel buffer too small");
    // LoadBmp: rows back to back when the width is a multiple of 4, else
    // each row followed by (width % 4) skipped bytes
    const uint64_t skip = (w % 4) ? (w % 4) : 0;
    for (uint32_t y = 0; y < h; ++y) {
        const uint64_t at = offset + (uint64_t)y * (row + skip);
        if (at + row > file_bytes) return fail(RUMI_ERR_INVALID, "rumi_bmp_read: file ends inside row %u", y);
        std::memcpy(pixels + (uint64_t)y * row, file + at, row);
    }
    return RUMI_OK;
}
