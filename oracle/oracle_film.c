/* Oracle: the film output path, restated from Film::Save
 * (src/Camera/film.cpp:152-203), SaveBmp (src/Kernel/bmp.cpp:59-97) and
 * FirefliesFiltering (src/Scene/scene.cpp:261-431).  Test infrastructure
 * only: the tests compare the library's GPU kernels with this and pin this
 * against closed forms and hand-made images.
 *
 * powf is taken correctly rounded (the double power rounded once to float);
 * NaN channel values become 0; neighbours the reference reads outside its
 * buffer read as 0; the fourth byte of a 32-bit pixel is 0.
 */
#include "oracle.h"

#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

static unsigned char channel_byte(float v)
{
    if (v > 255.0f) v = 255.0f;
    else if (v < 0.0f) v = 0.0f;
    if (isnan(v)) return 0;
    return (unsigned char)(unsigned int)v;
}

void or_film_encode(const float* film, int width, int height, int bgr_sum, int n_samples,
                    float gamma, int bpp, unsigned char* out)
{
    int depth = (int)sqrtf((float)n_samples);
    float inv_n = 1.0f / (float)(depth * depth);
    float g = 1.0f / gamma;
    long i, n = (long)width * height;
    memset(out, 0, (size_t)n * (size_t)bpp);
    for (i = 0; i < n; ++i) {
        const float* p = film + 3 * i;
        float c[3];
        int k;
        if (bgr_sum) {
            c[0] = p[0] * inv_n;
            c[1] = p[1] * inv_n;
            c[2] = p[2] * inv_n;
        } else {
            c[0] = p[2];
            c[1] = p[1];
            c[2] = p[0];
        }
        for (k = 0; k < 3; ++k) {
            float v = (float)pow((double)c[k], (double)g) * 255.0f;
            out[i * bpp + k] = channel_byte(v);
        }
    }
}

void or_bmp_header(int width, int height, int bpp, unsigned char* h)
{
    unsigned int data = (unsigned int)width * (unsigned int)height * (unsigned int)bpp;
    unsigned int f[13];
    int k;
    memset(h, 0, 54);
    h[0] = 'B';
    h[1] = 'M';
    /* file header (12 bytes) and info header (40), little-endian */
    f[0] = 54 + data; /* file size */
    f[1] = 0;         /* two 16-bit reserved words */
    f[2] = 54;        /* data offset */
    f[3] = 40;
    f[4] = (unsigned int)width;
    f[5] = (unsigned int)height;
    f[6] = 1u | ((unsigned int)(bpp * 8) << 16); /* planes, bits per pixel */
    f[7] = 0;
    f[8] = data;
    f[9] = f[10] = f[11] = f[12] = 0;
    for (k = 0; k < 13; ++k) {
        h[2 + 4 * k + 0] = (unsigned char)(f[k]);
        h[2 + 4 * k + 1] = (unsigned char)(f[k] >> 8);
        h[2 + 4 * k + 2] = (unsigned char)(f[k] >> 16);
        h[2 + 4 * k + 3] = (unsigned char)(f[k] >> 24);
    }
}

void or_fireflies(const unsigned char* src, int w, int h, int bpp, unsigned char* dst)
{
    long row = (long)w * bpp, total = row * h;
    int y, x;
    for (y = 0; y < h; ++y)
        for (x = 0; x < w; ++x) {
            long at = (long)(h - (y + 1)) * row + (long)x * bpp;
            long pos[8];
            int use[8], k, m = 0, nc = 8, white = 0;
            float t[3], nb[8][3], s[3] = {0, 0, 0}, scale;
            for (k = 0; k < 3; ++k) t[k] = (float)src[at + k];
            if (bpp == 4) dst[at + 3] = 0;
            if (t[0] < 255.0f && t[1] < 255.0f && t[2] < 255.0f) {
                for (k = 0; k < 3; ++k) dst[at + k] = (unsigned char)(int)t[k];
                continue;
            }
            if (x == 0 || y == 0 || x == w - 1 || y == h - 1) {
                nc -= 3;
                if (x == 0 && y == 0) nc -= 2;
                if (x == 0 && y == h - 1) nc -= 2;
                if (x == w - 1 && y == 0) nc -= 2;
                if (x == w - 1 && y == h - 1) nc -= 2;
            }
            /* left, right; buffer row above: left, centre, right; below */
            use[0] = x != 0;                 pos[0] = at - bpp;
            use[1] = x != w - 1;             pos[1] = at + bpp;
            use[2] = y != 0 || x != 0;       pos[2] = at - row - bpp;
            use[3] = y != 0;                 pos[3] = at - row;
            use[4] = y != 0 || x != w - 1;   pos[4] = at - row + bpp;
            use[5] = y != h - 1 || x != 0;   pos[5] = at + row - bpp;
            use[6] = y != h - 1;             pos[6] = at + row;
            use[7] = y != h - 1 || x != w - 1; pos[7] = at + row + bpp;
            for (k = 0; k < 8; ++k) {
                int c;
                if (!use[k]) continue;
                for (c = 0; c < 3; ++c)
                    nb[m][c] = (pos[k] >= 0 && pos[k] + 2 < total) ? (float)src[pos[k] + c] : 0.0f;
                m++;
            }
            for (k = 0; k < nc && k < m; ++k)
                if (nb[k][0] == 255.0f && nb[k][1] == 255.0f && nb[k][2] == 255.0f) white = 1;
            if (white) {
                for (k = 0; k < 3; ++k) dst[at + k] = (unsigned char)(int)t[k];
                continue;
            }
            for (k = 0; k < nc && k < m; ++k) {
                s[0] += nb[k][0];
                s[1] += nb[k][1];
                s[2] += nb[k][2];
            }
            scale = 1.0f / (float)nc;
            for (k = 0; k < 3; ++k) dst[at + k] = channel_byte(s[k] * scale);
        }
}

/* ImageTexture::Pixel's texel values (src/Texture/imagetexture.cpp:32-61):
 * byte * (1.0f / 255.0f), raised to gamma unless it is 1; R, G, B from the
 * B, G, R bytes. */
void or_bmp_texels(const unsigned char* px, int w, int h, int bpp, float gamma, float* rgb)
{
    const float inv = 1.0f / 255.0f;
    long i, n = (long)w * h;
    int c;
    for (i = 0; i < n; ++i)
        for (c = 0; c < 3; ++c) {
            float x = (float)px[i * bpp + 2 - c] * inv;
            rgb[3 * i + c] = gamma == 1.0f ? x : (float)pow((double)x, (double)gamma);
        }
}

/* LoadHDR + ScanLine (src/Kernel/hdr.cpp:4-177) over a file image in memory:
 * rows top-down, (r, g, b) * 2^(e - 136).  Returns -1 where the reference
 * would fail or read garbage (see the library's rumi_hdr_read). */
static int next_line(const unsigned char* d, size_t n, size_t* at, char* out)
{
    int k = 0;
    if (*at >= n) return 0;
    while (k < 255 && *at < n) {
        char c = (char)d[(*at)++];
        out[k++] = c;
        if (c == '\n') break;
    }
    out[k] = 0;
    return 1;
}

int or_hdr_read(const unsigned char* d, size_t n, int* width, int* height, float* rgb)
{
    char line[256];
    size_t at = 0;
    int w = 0, h = 0, y, x, c;
    if (!next_line(d, n, &at, line) || strcmp(line, "#?RADIANCE\n") != 0) return -1;
    do {
        if (!next_line(d, n, &at, line)) return -1;
    } while (line[0] == '#');
    if (strcmp(line, "FORMAT=32-bit_rle_rgbe\n") != 0) return -1;
    do {
        if (!next_line(d, n, &at, line)) return -1;
    } while (line[0] != '\n');
    if (!next_line(d, n, &at, line) || sscanf(line, "-Y %d + X  %d", &h, &w) != 2 || h <= 0 || w <= 0)
        return -1;
    *width = w;
    *height = h;
    if (!rgb) return 0;
    {
        unsigned char* pl = (unsigned char*)malloc(4 * (size_t)w);
        if (!pl) return -1;
        for (y = 0; y < h; ++y) {
            signed char hd[4];
            unsigned char* p = pl;
            if (w < 8 || w > 0x7fff || at + 4 > n) { free(pl); return -1; }
            for (c = 0; c < 4; ++c) hd[c] = (signed char)d[at + c];
            at += 4;
            if (hd[0] != 2 || hd[1] != 2 || (hd[2] & 0x80) || ((((int)hd[2]) << 8) | (int)hd[3]) != w) {
                free(pl);
                return -1;
            }
            for (c = 0; c < 4; ++c) {
                unsigned char* end = pl + (size_t)(c + 1) * w;
                while (p < end) {
                    int cnt;
                    unsigned char a, b;
                    if (at + 2 > n) { free(pl); return -1; }
                    a = d[at];
                    b = d[at + 1];
                    at += 2;
                    if (a > 128) {
                        cnt = a - 128;
                        if (cnt > end - p) { free(pl); return -1; }
                        while (cnt-- > 0) *p++ = b;
                    } else {
                        cnt = a;
                        if (cnt == 0 || cnt > end - p) { free(pl); return -1; }
                        *p++ = b;
                        if (--cnt > 0) {
                            if (at + (size_t)cnt > n) { free(pl); return -1; }
                            memcpy(p, d + at, (size_t)cnt);
                            at += (size_t)cnt;
                            p += cnt;
                        }
                    }
                }
            }
            for (x = 0; x < w; ++x) {
                float* o = rgb + 3 * ((size_t)y * w + x);
                int e = pl[3 * (size_t)w + x];
                if (e) {
                    float f = (float)ldexp(1.0, e - 136);
                    for (c = 0; c < 3; ++c) o[c] = (float)pl[(size_t)c * w + x] * f;
                } else {
                    o[0] = o[1] = o[2] = 0.0f;
                }
            }
        }
        free(pl);
    }
    return 0;
}
