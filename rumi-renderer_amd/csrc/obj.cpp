// C ABI: OBJ models -- Model::LoadObjectFile (src/Shape/model.cpp:72-338)
// with the reference's lexical rules, index arithmetic and transforms, so a
// loaded model holds the same FP32 triangles and vertex normals that the
// reference's Model builds.  The renderer takes the result as a
// RUMI_SHAPE_MESH (one SAH tree over the scene, no second-level BVH: the
// closest hit is the same; see the two-level note in the header).
#include <cmath>
#include <cstdio>
#include <cstring>
#include <new>
#include <vector>

#include "internal.h"

struct rumi_obj {
    uint32_t n_triangles = 0;
    uint32_t counts[4] = {0, 0, 0, 0};  // v, vt, vn lines; f lines
    std::vector<float> xyz;             // 9 per triangle (corners 0, 1, 2)
    std::vector<float> normals;         // 9 per triangle
};

using namespace rumi;

namespace {

// Row-major 4x4 (Mat4x4::_m, src/Kernel/matrix.h).
struct M4 {
    float m[16];
};

M4 identity() {
    M4 r{};
    r.m[0] = r.m[5] = r.m[10] = r.m[15] = 1.0f;
    return r;
}

// Inverse (src/Kernel/matrix.cpp:67-133): Gaussian elimination without
// pivoting -- below the diagonal column by column, then above it, then each
// row scaled by the reciprocal of its pivot; the same operations applied to
// an identity matrix give the inverse.
M4 inverse(const M4& a) {
    M4 s = a, inv = identity();
    auto eliminate = [&](int col, int row) {
        const float f = -s.m[row * 4 + col] / s.m[col * 4 + col];
        for (int k = 0; k < 4; ++k) {
            s.m[row * 4 + k] += s.m[col * 4 + k] * f;
            inv.m[row * 4 + k] += inv.m[col * 4 + k] * f;
        }
    };
    for (int col = 0; col < 4; ++col)
        for (int row = col + 1; row < 4; ++row) eliminate(col, row);
    for (int col = 0; col < 4; ++col)
        for (int row = col - 1; row >= 0; --row) eliminate(col, row);
    for (int r = 0; r < 4; ++r) {
        const float ip = 1.0f / s.m[r * 4 + r];
        for (int k = 0; k < 4; ++k) inv.m[r * 4 + k] *= ip;
    }
    return inv;
}

M4 transpose(const M4& a) {
    M4 t;
    for (int r = 0; r < 4; ++r)
        for (int c = 0; c < 4; ++c) t.m[c * 4 + r] = a.m[r * 4 + c];
    return t;
}

// Mat4x4 * Point3 and Mat4x4 * Vector3 (matrix.cpp:42-64): both add the
// fourth column, summed left to right.
void apply(const M4& a, const float in[3], float out[3]) {
    for (int r = 0; r < 3; ++r)
        out[r] = a.m[r * 4 + 0] * in[0] + a.m[r * 4 + 1] * in[1] + a.m[r * 4 + 2] * in[2] + a.m[r * 4 + 3];
}

// Normalize / Cross (src/Kernel/vector.h:44-68).
void normalize(float v[3]) {
    const float il = 1.0f / sqrtf(v[0] * v[0] + v[1] * v[1] + v[2] * v[2]);
    v[0] *= il;
    v[1] *= il;
    v[2] *= il;
}

void face_normal(const float* p0, const float* p1, const float* p2, float n[3]) {
    const float u[3] = {p1[0] - p0[0], p1[1] - p0[1], p1[2] - p0[2]};
    const float w[3] = {p2[0] - p0[0], p2[1] - p0[1], p2[2] - p0[2]};
    n[0] = u[1] * w[2] - w[1] * u[2];
    n[1] = u[2] * w[0] - w[2] * u[0];
    n[2] = u[0] * w[1] - w[0] * u[1];
    normalize(n);
}

// One corner of a face record: "v", "v/vt", "v//vn", "v/vt/vn", digits
// accumulated in FP32, a '-' anywhere after the last '/' making the whole
// group relative to the file's total counts (the reference's rule, not the
// OBJ standard's "relative to the current position").
struct Corner {
    int v = 0, vt = 0, vn = 0;
};

// Parses the text after "f " (a 256-byte line buffer, NUL-padded) into three
// corners; returns false when the buffer ends before the third.
bool parse_face(const char* buf, const uint32_t counts[4], Corner out[3]) {
    float data[3] = {0.0f, 0.0f, 0.0f};
    int field = 0, got = 0;
    bool negative = false;
    for (int i = 0; i < 256 && got < 3; ++i) {
        const char c = buf[i];
        if (c == '-') {
            negative = true;
        } else if (c >= '0' && c <= '9') {
            if (field > 2) return false;  // a fourth field: the reference writes past its array
            data[field] = data[field] * 10.0f + (float)(c - '0');
        } else if (c == '/') {
            ++field;
            negative = false;
        } else if (c == ' ' || c == '\n' || c == '\0') {
            if (c == '\0' && i > 0 && buf[i - 1] == '\0') return false;  // past the line's end
            Corner& k = out[got++];
            if (!negative) {
                k.v = (int)(data[0] - 1);
                k.vt = (int)(data[1] - 1);
                k.vn = (int)(data[2] - 1);
            } else {
                k.v = (int)(-data[0] + (float)counts[0]);
                k.vt = (int)(-data[1] + (float)counts[1]);
                k.vn = (int)(-data[2] + (float)counts[2]);
            }
            field = 0;
            data[0] = data[1] = data[2] = 0.0f;
            negative = false;
        }
    }
    return got == 3;
}

int load(FILE* f, const M4& xf, rumi_obj* m) {
    // pass 1: records by their first two bytes, as fgets(256) chunks them
    char line[256];
    uint32_t nv = 0, nvt = 0, nvn = 0, nf = 0;
    while (fgets(line, sizeof(line), f)) {
        if (line[0] == 'v' && line[1] == ' ') ++nv;
        if (line[0] == 'v' && line[1] == 't') ++nvt;
        if (line[0] == 'v' && line[1] == 'n') ++nvn;
        if (line[0] == 'f' && line[1] == ' ') ++nf;
    }
    m->counts[0] = nv;
    m->counts[1] = nvt;
    m->counts[2] = nvn;
    m->counts[3] = nf;
    if (fseek(f, 0, SEEK_SET) != 0) return fail(RUMI_ERR_INVALID, "rumi_obj_load: file not seekable");

    // pass 2: a record tag is up to four bytes ending at a space or newline
    std::vector<float> v, vn;
    std::vector<Corner> corners;
    v.reserve(3 * (size_t)nv);
    vn.reserve(3 * (size_t)nvn);
    corners.reserve(3 * (size_t)nf);
    char face[256];
    while (true) {
        signed char tag[4] = {0, 0, 0, 0};
        for (int i = 0; i < 4; ++i) {
            tag[i] = (signed char)fgetc(f);
            if (tag[i] == '\n' || tag[i] == ' ') break;
        }
        if (tag[0] == (signed char)EOF) break;
        const bool is_v = tag[0] == 'v' && tag[1] == ' ';
        const bool is_vt = tag[0] == 'v' && tag[1] == 't' && tag[2] == ' ';
        const bool is_vn = tag[0] == 'v' && tag[1] == 'n' && tag[2] == ' ';
        const bool is_f = tag[0] == 'f' && tag[1] == ' ';
        if (is_v || is_vn) {
            // fields fscanf cannot read stay 0, and the record still counts
            float x = 0.0f, y = 0.0f, z = 0.0f;
            (void)!fscanf(f, "%f %f %f", &x, &y, &z);
            std::vector<float>& dst = is_v ? v : vn;
            dst.push_back(x);
            dst.push_back(y);
            dst.push_back(z);
        } else if (is_vt) {
            float s = 0.0f, t = 0.0f;
            (void)!fscanf(f, "%f %f", &s, &t);  // read and dropped: triangles carry no uv
        } else if (is_f) {
            std::memset(face, 0, sizeof(face));
            if (fgets(face, sizeof(face), f)) {
                Corner c[3];
                if (!parse_face(face, m->counts, c))
                    return fail(RUMI_ERR_INVALID, "rumi_obj_load: face %zu has fewer than 3 corners",
                                corners.size() / 3);
                corners.insert(corners.end(), c, c + 3);
            }
        }
    }

    // triangles: vertices through the transform, normals through the
    // transposed inverse (Shape::Shape, src/Shape/shape.cpp:29-37), zero or
    // missing normals replaced by the transformed face's normal
    const M4 nxf = transpose(inverse(xf));
    const size_t nt = corners.size() / 3;
    const uint32_t nvv = (uint32_t)(v.size() / 3), nnn = (uint32_t)(vn.size() / 3);
    m->xyz.resize(9 * nt);
    m->normals.resize(9 * nt);
    for (size_t t = 0; t < nt; ++t) {
        const Corner* c = &corners[3 * t];
        float* p = &m->xyz[9 * t];
        float* n = &m->normals[9 * t];
        for (int k = 0; k < 3; ++k) {
            if (c[k].v < 0 || (uint32_t)c[k].v >= nvv)
                return fail(RUMI_ERR_INVALID, "rumi_obj_load: face %zu: vertex index %d out of range", t,
                            c[k].v + 1);
            apply(xf, &v[3 * (size_t)c[k].v], p + 3 * k);
        }
        float fn[3];
        face_normal(p, p + 3, p + 6, fn);
        for (int k = 0; k < 3; ++k) {
            if (nvn == 0) {
                std::memcpy(n + 3 * k, fn, sizeof(fn));
                continue;
            }
            if (c[k].vn < 0 || (uint32_t)c[k].vn >= nnn)
                return fail(RUMI_ERR_INVALID, "rumi_obj_load: face %zu: normal index %d out of range", t,
                            c[k].vn + 1);
            float src[3];
            std::memcpy(src, &vn[3 * (size_t)c[k].vn], sizeof(src));
            if (src[0] == 0.0f && src[1] == 0.0f && src[2] == 0.0f) std::memcpy(src, fn, sizeof(src));
            apply(nxf, src, n + 3 * k);
            normalize(n + 3 * k);
        }
    }
    m->n_triangles = (uint32_t)nt;
    return RUMI_OK;
}

}  // namespace

extern "C" int rumi_obj_load(const char* path, const float* transform, rumi_obj** out) {
    clear_error();
    if (!out) return fail(RUMI_ERR_INVALID, "rumi_obj_load: null output");
    *out = nullptr;
    if (!path) return fail(RUMI_ERR_INVALID, "rumi_obj_load: null path");
    M4 xf = identity();
    if (transform) std::memcpy(xf.m, transform, sizeof(xf.m));
    FILE* f = fopen(path, "rb");
    if (!f) return fail(RUMI_ERR_INVALID, "rumi_obj_load: cannot open %s", path);
    rumi_obj* m = new (std::nothrow) rumi_obj;
    if (!m) {
        fclose(f);
        return fail(RUMI_ERR_NO_MEMORY, "rumi_obj_load: out of host memory");
    }
    int rc;
    try {
        rc = load(f, xf, m);
    } catch (const std::bad_alloc&) {
        rc = fail(RUMI_ERR_NO_MEMORY, "rumi_obj_load: out of host memory");
    }
    fclose(f);
    if (rc != RUMI_OK) {
        delete m;
        return rc;
    }
    *out = m;
    return RUMI_OK;
}

extern "C" int rumi_obj_get(const rumi_obj* m, uint32_t* n_triangles, const float** xyz,
                            const float** normals, uint32_t* counts) {
    clear_error();
    if (!m) return fail(RUMI_ERR_INVALID, "rumi_obj_get: null model");
    if (n_triangles) *n_triangles = m->n_triangles;
    if (xyz) *xyz = m->xyz.data();
    if (normals) *normals = m->normals.data();
    if (counts) std::memcpy(counts, m->counts, sizeof(m->counts));
    return RUMI_OK;
}

extern "C" int rumi_obj_destroy(rumi_obj* m) {
    clear_error();
    delete m;
    return RUMI_OK;
}
