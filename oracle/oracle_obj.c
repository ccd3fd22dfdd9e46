/* Oracle: OBJ model loading, restated from Model::LoadObjectFile
 * (src/Shape/model.cpp:72-338), Shape::Shape's transposed inverse
 * (src/Shape/shape.cpp:29-37), Inverse / Transpose / Mat4x4 * Point3 /
 * Mat4x4 * Vector3 (src/Kernel/matrix.cpp:42-147) and Normalize / Cross
 * (src/Kernel/vector.h:44-68).  Test infrastructure only: the tests compare
 * the library's loader with this, and pin this against closed forms and the
 * reference's own model files.
 *
 * Output: per triangle 9 vertex floats and 9 normal floats (corners 0, 1, 2).
 * Faces whose reference behaviour is undefined (fewer than three corners, a
 * fourth '/' field, an index outside the records) make the call return -1.
 */
#include "oracle.h"

#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

static void mat_inverse(const float* a, float* inv)
{
    float s[16];
    int c, r, k;
    memcpy(s, a, sizeof(s));
    for (k = 0; k < 16; ++k) inv[k] = (k % 5 == 0) ? 1.0f : 0.0f;
    /* below the diagonal */
    for (c = 0; c < 4; ++c)
        for (r = c + 1; r < 4; ++r) {
            float x = -s[r * 4 + c] / s[c * 4 + c];
            for (k = 0; k < 4; ++k) {
                s[r * 4 + k] += s[c * 4 + k] * x;
                inv[r * 4 + k] += inv[c * 4 + k] * x;
            }
        }
    /* above it, nearest row first */
    for (c = 0; c < 4; ++c)
        for (r = c - 1; r >= 0; --r) {
            float x = -s[r * 4 + c] / s[c * 4 + c];
            for (k = 0; k < 4; ++k) {
                s[r * 4 + k] += s[c * 4 + k] * x;
                inv[r * 4 + k] += inv[c * 4 + k] * x;
            }
        }
    for (r = 0; r < 4; ++r) {
        float ip = 1.0f / s[r * 5];
        for (k = 0; k < 4; ++k) inv[r * 4 + k] *= ip;
    }
}

void or_mat4_inverse(const float* a, float* out) { mat_inverse(a, out); }

static void mat_apply(const float* m, const float* p, float* o)
{
    int r;
    for (r = 0; r < 3; ++r) {
        float acc = m[r * 4] * p[0];
        acc = acc + m[r * 4 + 1] * p[1];
        acc = acc + m[r * 4 + 2] * p[2];
        o[r] = acc + m[r * 4 + 3];
    }
}

static void unit(float* v)
{
    float inv = 1.0f / sqrtf(v[0] * v[0] + v[1] * v[1] + v[2] * v[2]);
    v[0] = v[0] * inv;
    v[1] = v[1] * inv;
    v[2] = v[2] * inv;
}

typedef struct {
    float* a;
    size_t n, cap;
} fvec;

static int push(fvec* v, float x)
{
    if (v->n == v->cap) {
        size_t nc = v->cap ? 2 * v->cap : 1024;
        float* p = (float*)realloc(v->a, nc * sizeof(float));
        if (!p) return -1;
        v->a = p;
        v->cap = nc;
    }
    v->a[v->n++] = x;
    return 0;
}

/* The face text after "f ": three corners of (v, vt, vn) indices. */
static int face_corners(const char* b, const int cnt[3], int out[9])
{
    float d[3] = {0, 0, 0};
    int fld = 0, n = 0, i, neg = 0;
    for (i = 0; i < 256 && n < 3; ++i) {
        char c = b[i];
        if (c >= '0' && c <= '9') {
            if (fld > 2) return -1;
            d[fld] = d[fld] * 10.0f + (float)(c - 48);
        } else if (c == '-') {
            neg = 1;
        } else if (c == '/') {
            fld++;
            neg = 0;
        } else if (c == ' ' || c == '\n' || c == 0) {
            int j;
            if (c == 0 && i > 0 && b[i - 1] == 0) return -1;
            for (j = 0; j < 3; ++j)
                out[3 * n + j] = neg ? (int)(-d[j] + (float)cnt[j]) : (int)(d[j] - 1);
            n++;
            d[0] = d[1] = d[2] = 0;
            fld = 0;
            neg = 0;
        }
    }
    return n == 3 ? 0 : -1;
}

int or_obj_load(const char* path, const float* xf, float** out_xyz, float** out_n, int* n_tri,
                int counts[4])
{
    FILE* f = fopen(path, "rb");
    char line[256];
    int cnt[4] = {0, 0, 0, 0};
    fvec v = {0, 0, 0}, vn = {0, 0, 0}, idx = {0, 0, 0};
    float m[16], mi[16], nt_m[16];
    int rc = 0, t, k, r, c;
    if (!f) return -1;
    if (xf) memcpy(m, xf, sizeof(m));
    else
        for (k = 0; k < 16; ++k) m[k] = (k % 5 == 0) ? 1.0f : 0.0f;
    while (fgets(line, 256, f)) {
        if (line[0] == 'v' && line[1] == ' ') cnt[0]++;
        if (line[0] == 'v' && line[1] == 't') cnt[1]++;
        if (line[0] == 'v' && line[1] == 'n') cnt[2]++;
        if (line[0] == 'f' && line[1] == ' ') cnt[3]++;
    }
    rewind(f);
    for (;;) {
        signed char tag[4] = {0, 0, 0, 0};
        int i;
        for (i = 0; i < 4; ++i) {
            tag[i] = (signed char)fgetc(f);
            if (tag[i] == ' ' || tag[i] == '\n') break;
        }
        if (tag[0] == -1) break;
        if (tag[0] == 'v' && (tag[1] == ' ' || ((tag[1] == 'n' || tag[1] == 't') && tag[2] == ' '))) {
            float x[3] = {0, 0, 0};
            int want = tag[1] == 't' ? 2 : 3;
            if (want == 3) {
                if (fscanf(f, "%f %f %f", &x[0], &x[1], &x[2]) < 0) { /* fields stay 0 */ }
            } else {
                if (fscanf(f, "%f %f", &x[0], &x[1]) < 0) { /* texture coordinates: dropped */ }
            }
            if (tag[1] == ' ')
                for (i = 0; i < 3; ++i) rc |= push(&v, x[i]);
            else if (tag[1] == 'n')
                for (i = 0; i < 3; ++i) rc |= push(&vn, x[i]);
        } else if (tag[0] == 'f' && tag[1] == ' ') {
            char b[256];
            int cor[9];
            memset(b, 0, sizeof(b));
            if (fgets(b, 256, f)) {
                if (face_corners(b, cnt, cor) != 0) { rc = -1; break; }
                for (i = 0; i < 9; ++i) rc |= push(&idx, (float)cor[i]);
            }
        }
        if (rc) break;
    }
    fclose(f);
    if (rc) goto done;
    mat_inverse(m, mi);
    for (r = 0; r < 4; ++r)
        for (c = 0; c < 4; ++c) nt_m[c * 4 + r] = mi[r * 4 + c];
    {
        int ntri = (int)(idx.n / 9), nv = (int)(v.n / 3), nn = (int)(vn.n / 3);
        float* xyz = (float*)malloc(sizeof(float) * 9 * (size_t)(ntri > 0 ? ntri : 1));
        float* nrm = (float*)malloc(sizeof(float) * 9 * (size_t)(ntri > 0 ? ntri : 1));
        if (!xyz || !nrm) { free(xyz); free(nrm); rc = -1; goto done; }
        for (t = 0; t < ntri && !rc; ++t) {
            const float* cor = idx.a + 9 * (size_t)t;
            float* p = xyz + 9 * (size_t)t;
            float* n = nrm + 9 * (size_t)t;
            float e1[3], e2[3], fn[3];
            for (k = 0; k < 3; ++k) {
                int vi = (int)cor[3 * k];
                if (vi < 0 || vi >= nv) { rc = -1; break; }
                mat_apply(m, v.a + 3 * (size_t)vi, p + 3 * k);
            }
            if (rc) break;
            for (k = 0; k < 3; ++k) {
                e1[k] = p[3 + k] - p[k];
                e2[k] = p[6 + k] - p[k];
            }
            fn[0] = e1[1] * e2[2] - e2[1] * e1[2];
            fn[1] = e1[2] * e2[0] - e2[2] * e1[0];
            fn[2] = e1[0] * e2[1] - e2[0] * e1[1];
            unit(fn);
            for (k = 0; k < 3; ++k) {
                float src[3];
                int ni = (int)cor[3 * k + 2];
                if (nn == 0) {
                    memcpy(n + 3 * k, fn, sizeof(fn));
                    continue;
                }
                if (ni < 0 || ni >= nn) { rc = -1; break; }
                memcpy(src, vn.a + 3 * (size_t)ni, sizeof(src));
                if (src[0] == 0.0f && src[1] == 0.0f && src[2] == 0.0f) memcpy(src, fn, sizeof(src));
                mat_apply(nt_m, src, n + 3 * k);
                unit(n + 3 * k);
            }
        }
        if (rc) { free(xyz); free(nrm); goto done; }
        *out_xyz = xyz;
        *out_n = nrm;
        *n_tri = ntri;
        memcpy(counts, cnt, sizeof(cnt));
    }
done:
    free(v.a);
    free(vn.a);
    free(idx.a);
    return rc ? -1 : 0;
}

void or_obj_free(float* p) { free(p); }
