// C ABI: scene creation -- expands shapes into primitives, builds the SAH BVH
// over all of them and uploads the leaf records and shading tables.
#include <hip/hip_runtime_api.h>

#include <cfloat>
#include <cmath>
#include <cstring>
#include <new>

#include "scene.h"

using namespace rumi;

namespace {

constexpr double kPi = 3.14159265358979323846;

struct Prim {
    int32_t kind;      // RUMI_SHAPE_TRIANGLE / SPHERE / RECT
    int32_t material;
    int32_t texture;
    int32_t shape;     // index of the rumi_shape it came from
    float v[9];
    float n[9];
    float uv[6];
    float radius;
};

void tri_normal_default(Prim& p) {
    bool zero = true;
    for (int k = 0; k < 9; ++k) zero = zero && p.n[k] == 0.0f;
    if (!zero) return;
    const float e1[3] = {p.v[3] - p.v[0], p.v[4] - p.v[1], p.v[5] - p.v[2]};
    const float e2[3] = {p.v[6] - p.v[0], p.v[7] - p.v[1], p.v[8] - p.v[2]};
    float c[3] = {e1[1] * e2[2] - e2[1] * e1[2], e1[2] * e2[0] - e2[2] * e1[0],
                  e1[0] * e2[1] - e2[0] * e1[1]};
    const float inv = 1.0f / sqrtf(c[0] * c[0] + c[1] * c[1] + c[2] * c[2]);
    for (int j = 0; j < 3; ++j)
        for (int k = 0; k < 3; ++k) p.n[3 * j + k] = c[k] * inv;
}

template <typename T>
int upload(void** dst, const std::vector<T>& v) {
    *dst = nullptr;
    if (v.empty()) return RUMI_OK;
    const size_t bytes = sizeof(T) * v.size();
    if (hipMalloc(dst, bytes) != hipSuccess)
        return fail(RUMI_ERR_NO_MEMORY, "hipMalloc of %zu scene bytes failed", bytes);
    if (hipMemcpy(*dst, v.data(), bytes, hipMemcpyHostToDevice) != hipSuccess)
        return fail(RUMI_ERR_HIP, "hipMemcpy of scene data failed");
    return RUMI_OK;
}

void release(rumi_scene* s) {
    if (s->device >= 0) {
        int prev = 0;
        (void)hipGetDevice(&prev);
        (void)hipSetDevice(s->device);
        void* ptrs[] = {s->nodes, s->prims, s->rects, s->shapes, s->materials,
                        s->lights, s->textures, s->texels, s->work, s->adj};
        for (void* p : ptrs)
            if (p) (void)hipFree(p);
        (void)hipSetDevice(prev);
    }
}

}  // namespace

extern "C" int rumi_scene_create(rumi_ctx* ctx, const rumi_scene_desc* d, rumi_scene** out) {
    clear_error();
    if (!out) return fail(RUMI_ERR_INVALID, "rumi_scene_create: null output");
    *out = nullptr;
    if (!ctx) return fail(RUMI_ERR_NO_DEVICE, "rumi_scene_create: null context");
    if (!d) return fail(RUMI_ERR_INVALID, "rumi_scene_create: null description");
    if (d->n_shapes && !d->shapes) return fail(RUMI_ERR_INVALID, "scene: null shapes");
    if (d->n_materials == 0 || !d->materials) return fail(RUMI_ERR_INVALID, "scene: no materials");
    if (d->n_textures && !d->textures) return fail(RUMI_ERR_INVALID, "scene: null textures");

    // ------------------------------------------------------------ materials
    std::vector<DevMaterial> mats(d->n_materials);
    uint32_t kinds = 0;
    for (uint32_t i = 0; i < d->n_materials; ++i) {
        const rumi_material& m = d->materials[i];
        if (m.kind < RUMI_MAT_MATTE || m.kind > RUMI_MAT_GGX)
            return fail(RUMI_ERR_INVALID, "material %u: unknown kind %d", i, m.kind);
        DevMaterial& o = mats[i];
        std::memset(&o, 0, sizeof(o));
        o.kind = m.kind;
        kinds |= 1u << m.kind;
        o.opaque = !(m.kind == RUMI_MAT_DIELECTRIC || m.kind == RUMI_MAT_SUBSURFACE);
        o.kd = m.kd;
        o.ks = m.ks;
        o.e = m.exponent;
        o.kd_pdf = (m.kd + m.ks) > 0.0f ? m.kd / (m.kd + m.ks) : 0.0f;
        o.phong_norm = (float)(((double)m.exponent + 1.0) / (2.0 * kPi));
        o.eta_in = m.eta_in;
        o.eta_out = m.eta_out;
        o.density = m.density > 0.0f ? m.density : 1.0f;
        o.alpha = m.roughness;
        if (m.kind == RUMI_MAT_GGX && !(m.roughness > 0.0f && m.roughness <= 1.0f))
            return fail(RUMI_ERR_INVALID, "material %u: GGX roughness must be in (0, 1]", i);
        for (int k = 0; k < 3; ++k) {
            o.color[k] = m.color[k];
            o.absorption[k] = m.absorption[k];
            o.sigma_s[k] = m.sigma_s[k];
            o.sigma_a[k] = m.sigma_a[k];
            o.sigma_t[k] = m.sigma_s[k] + m.sigma_a[k];
        }
        if (m.kind == RUMI_MAT_SUBSURFACE) {
            for (int k = 0; k < 3; ++k)
                if (!(o.sigma_t[k] > 0.0f))
                    return fail(RUMI_ERR_INVALID, "material %u: subsurface sigma_t must be > 0", i);
            // Subsurface ctor (src/Material/subsurface.cpp:34-47)
            o.avg_sigmat = (o.sigma_t[2] + o.sigma_t[1] + o.sigma_t[0]) / 3.0f * o.density;
            o.absorb = ((o.sigma_a[2] / o.sigma_t[2]) + (o.sigma_a[1] / o.sigma_t[1]) +
                        (o.sigma_a[0] / o.sigma_t[0])) / 3.0f;
        }
    }

    // ------------------------------------------------------------ primitives
    std::vector<Prim> prims;
    for (uint32_t si = 0; si < d->n_shapes; ++si) {
        const rumi_shape& s = d->shapes[si];
        if (s.material < 0 || (uint32_t)s.material >= d->n_materials)
            return fail(RUMI_ERR_INVALID, "shape %u: material %d out of range", si, s.material);
        if (s.texture >= 0 && (uint32_t)s.texture >= d->n_textures)
            return fail(RUMI_ERR_INVALID, "shape %u: texture %d out of range", si, s.texture);
        Prim p;
        std::memset(&p, 0, sizeof(p));
        p.material = s.material;
        p.texture = s.texture < 0 ? -1 : s.texture;
        p.shape = (int32_t)si;
        switch (s.kind) {
        case RUMI_SHAPE_TRIANGLE:
            p.kind = RUMI_SHAPE_TRIANGLE;
            std::memcpy(p.v, s.v, sizeof(p.v));
            std::memcpy(p.n, s.n, sizeof(p.n));
            tri_normal_default(p);
            prims.push_back(p);
            break;
        case RUMI_SHAPE_SPHERE:
            if (!(s.radius > 0.0f)) return fail(RUMI_ERR_INVALID, "shape %u: radius must be > 0", si);
            p.kind = RUMI_SHAPE_SPHERE;
            std::memcpy(p.v, s.v, sizeof(float) * 3);
            p.radius = s.radius;
            // texture basis: the rows of Shape::_transform's 3x3 (zero: identity)
            std::memcpy(p.n, s.n, sizeof(p.n));
            {
                bool zero = true;
                for (int k = 0; k < 9; ++k) zero = zero && p.n[k] == 0.0f;
                if (zero) p.n[0] = p.n[4] = p.n[8] = 1.0f;
            }
            prims.push_back(p);
            break;
        case RUMI_SHAPE_RECT:
            p.kind = RUMI_SHAPE_RECT;
            std::memcpy(p.v, s.v, sizeof(p.v));
            prims.push_back(p);
            break;
        case RUMI_SHAPE_MESH:
            if (s.nt && (!s.xyz || !s.idx)) return fail(RUMI_ERR_INVALID, "shape %u: null mesh", si);
            for (uint32_t t = 0; t < s.nt; ++t) {
                Prim q = p;
                q.kind = RUMI_SHAPE_TRIANGLE;
                for (int j = 0; j < 3; ++j) {
                    const uint32_t vi = s.idx[3 * (size_t)t + j];
                    if (vi >= s.nv)
                        return fail(RUMI_ERR_INVALID, "shape %u: index %u >= nv %u", si, vi, s.nv);
                    for (int k = 0; k < 3; ++k) {
                        q.v[3 * j + k] = s.xyz[3 * (size_t)vi + k];
                        q.n[3 * j + k] = s.normals ? s.normals[3 * (size_t)vi + k] : 0.0f;
                    }
                    if (s.uv) {
                        q.uv[2 * j + 0] = s.uv[2 * (size_t)vi + 0];
                        q.uv[2 * j + 1] = s.uv[2 * (size_t)vi + 1];
                    }
                }
                tri_normal_default(q);
                prims.push_back(q);
            }
            break;
        default:
            return fail(RUMI_ERR_INVALID, "shape %u: unknown kind %d", si, s.kind);
        }
    }
    if (prims.size() > (size_t)kMaxLeafFirst) return fail(RUMI_ERR_INVALID, "scene too large");
    const uint32_t n = (uint32_t)prims.size();

    // bounds: Triangle::Bound / Sphere::Bound; rects use the full parallelogram
    std::vector<float> lo(3 * (size_t)n), hi(3 * (size_t)n);
    for (uint32_t i = 0; i < n; ++i) {
        const Prim& p = prims[i];
        for (int k = 0; k < 3; ++k) {
            float a, b;
            if (p.kind == RUMI_SHAPE_TRIANGLE) {
                a = p.v[k] < p.v[3 + k] ? p.v[k] : p.v[3 + k];
                if (a > p.v[6 + k]) a = p.v[6 + k];
                b = p.v[k] > p.v[3 + k] ? p.v[k] : p.v[3 + k];
                if (b < p.v[6 + k]) b = p.v[6 + k];
            } else if (p.kind == RUMI_SHAPE_SPHERE) {
                a = p.v[k] - p.radius;
                b = p.v[k] + p.radius;
            } else {
                const float l = p.v[3 + k], w = p.v[6 + k];
                a = p.v[k] + (l < 0 ? l : 0.0f) + (w < 0 ? w : 0.0f);
                b = p.v[k] + (l > 0 ? l : 0.0f) + (w > 0 ? w : 0.0f);
            }
            lo[3 * (size_t)i + k] = a;
            hi[3 * (size_t)i + k] = b;
        }
    }
    BuildResult br;
    build_sah(lo.data(), hi.data(), (int)n, 4, 32, &br);
    FlatBvh flat;
    flatten(br, &flat);
    if (flat.max_depth > kMaxDepth)
        return fail(RUMI_ERR_UNSUPPORTED, "scene BVH depth %d exceeds %d", flat.max_depth, kMaxDepth);
    // the renderer traverses the tree's 4-wide form (near-first, up to three
    // stack entries per level)
    FlatBvh4 wide;
    widen(flat, &wide);
    if (3 * wide.max_depth > kMaxDepth)
        return fail(RUMI_ERR_UNSUPPORTED, "scene BVH: 4-wide depth %d needs more than %d stack entries",
                    wide.max_depth, kMaxDepth);

    // rect side table (Rect ctor, src/Shape/rect.cpp:6-15) and leaf records
    std::vector<float> rects;
    std::vector<int32_t> rect_of(n, -1);
    for (uint32_t i = 0; i < n; ++i) {
        const Prim& p = prims[i];
        if (p.kind != RUMI_SHAPE_RECT) continue;
        const float* P = p.v;
        const float* L = p.v + 3;
        const float* W = p.v + 6;
        const float cx = L[1] * W[2] - W[1] * L[2];
        const float cy = L[2] * W[0] - W[2] * L[0];
        const float cz = L[0] * W[1] - W[0] * L[1];
        const float inv = 1.0f / sqrtf(cx * cx + cy * cy + cz * cz);
        const float nx = cx * inv, ny = cy * inv, nz = cz * inv;
        const float length = sqrtf(L[0] * L[0] + L[1] * L[1] + L[2] * L[2]);
        const float width = sqrtf(W[0] * W[0] + W[1] * W[1] + W[2] * W[2]);
        const float inv_area = 1.0f / (length * width);
        rect_of[i] = (int32_t)(rects.size() / 16);
        const float rec[16] = {P[0], P[1], P[2], nx, ny, nz, L[0], L[1], L[2], W[0], W[1], W[2],
                               inv_area * width, inv_area * length, length, width};
        rects.insert(rects.end(), rec, rec + 16);
    }
    std::vector<DevTri> leaf(n);
    std::vector<DevShape> shapes(n);
    for (uint32_t i = 0; i < n; ++i) {
        const Prim& p = prims[i];
        DevShape& s = shapes[i];
        std::memset(&s, 0, sizeof(s));
        float g[12] = {0};
        if (p.kind == RUMI_SHAPE_TRIANGLE) {
            const float* v0 = p.v;
            const float* v1 = p.v + 3;
            const float* v2 = p.v + 6;
            const float rec[9] = {v0[0], v0[1], v0[2], v1[0] - v0[0], v1[1] - v0[1], v1[2] - v0[2],
                                  v2[0] - v0[0], v2[1] - v0[1], v2[2] - v0[2]};
            std::memcpy(g, rec, sizeof(rec));
        } else if (p.kind == RUMI_SHAPE_SPHERE) {
            g[0] = p.v[0];
            g[1] = p.v[1];
            g[2] = p.v[2];
            g[3] = p.radius;
        } else {
            std::memcpy(g, rects.data() + 16 * (size_t)rect_of[i], sizeof(float) * 12);
        }
        std::memcpy(s.g0, g, sizeof(float) * 4);
        std::memcpy(s.g1, g + 4, sizeof(float) * 4);
        std::memcpy(s.g2, g + 8, sizeof(float) * 4);
        for (int k = 0; k < 3; ++k) {
            s.n0[k] = p.n[k];
            s.n1[k] = p.n[3 + k];
            s.n2[k] = p.n[6 + k];
        }
        s.n0[3] = p.uv[0];
        s.n1[3] = p.uv[1];
        s.n2[3] = p.uv[2];
        s.uv2[0] = p.uv[3];
        s.uv2[1] = p.uv[4];
        s.uv2[2] = p.uv[5];
        s.info[0] = p.kind;
        s.info[1] = p.material;
        s.info[2] = p.texture;
        s.info[3] = p.kind == RUMI_SHAPE_RECT ? rect_of[i] : -1;
    }
    for (uint32_t sl = 0; sl < n; ++sl) {
        const int32_t i = flat.leaf_prims[sl];
        const DevShape& s = shapes[i];
        float rec[12];
        std::memcpy(rec, s.g0, 16);
        std::memcpy(rec + 4, s.g1, 16);
        std::memcpy(rec + 8, s.g2, 16);
        const int32_t tag = s.info[0] == RUMI_SHAPE_TRIANGLE ? 0 : (s.info[0] == RUMI_SHAPE_SPHERE ? 1 : 2);
        // rect: emissive (no back-face cull); sphere: emissive and textured (the
        // image-based-light root swap of Sphere::Intersect, sphere.cpp:85-92)
        const int32_t flags = ((tag == 2 && mats[s.info[1]].kind == RUMI_MAT_EMISSIVE) ||
                               (tag == 1 && mats[s.info[1]].kind == RUMI_MAT_EMISSIVE && s.info[2] >= 0))
                                  ? 1
                                  : 0;
        if (tag == 2) std::memcpy(&rec[8], &s.info[3], 4);
        std::memcpy(&rec[9], &i, 4);
        std::memcpy(&rec[10], &tag, 4);
        std::memcpy(&rec[11], &flags, 4);
        std::memcpy(&leaf[sl], rec, sizeof(rec));
    }

    // ------------------------------------------------------------ lights
    std::vector<DevLight> lights;
    bool sc_has_ibl = false;
    for (uint32_t i = 0; i < n; ++i) {
        const Prim& p = prims[i];
        if (!d->shapes[p.shape].is_light) continue;
        const DevMaterial& m = mats[p.material];
        if (m.kind != RUMI_MAT_EMISSIVE)
            return fail(RUMI_ERR_INVALID, "shape %d: an area light needs an emissive material", p.shape);
        DevLight L;
        std::memset(&L, 0, sizeof(L));
        L.prim = (int32_t)i;
        for (int k = 0; k < 3; ++k) L.le[k] = m.color[k] * m.kd;
        if (p.kind == RUMI_SHAPE_RECT) {
            const float* r = rects.data() + 16 * (size_t)rect_of[i];
            L.kind = RUMI_SHAPE_RECT;
            for (int k = 0; k < 3; ++k) {
                L.p[k] = r[k];
                L.l[k] = r[6 + k];
                L.w[k] = r[9 + k];
            }
            L.n[0] = r[3];
            L.n[1] = r[4];
            L.n[2] = r[5];
            L.p[3] = 1.0f / (r[14] * r[15]);  // 1 / area (Rect::Pdf)
        } else if (p.kind == RUMI_SHAPE_SPHERE && d->shapes[p.shape].is_light == 2) {
            // ImageBasedLight(sphere, emissive, texture) (imagebasedlight.cpp:24-33)
            if (p.texture < 0)
                return fail(RUMI_ERR_INVALID, "shape %d: an image-based light needs a texture", p.shape);
            L.kind = kLightIbl;
            for (int k = 0; k < 3; ++k) L.p[k] = p.v[k];
            L.l[0] = p.radius;
            sc_has_ibl = true;
        } else if (p.kind == RUMI_SHAPE_SPHERE) {
            L.kind = RUMI_SHAPE_SPHERE;
            for (int k = 0; k < 3; ++k) L.p[k] = p.v[k];
            L.l[0] = p.radius;
            L.p[3] = (float)(1.0 / (4.0 * kPi * p.radius * p.radius));
        } else {
            return fail(RUMI_ERR_UNSUPPORTED, "shape %d: only rect and sphere area lights", p.shape);
        }
        lights.push_back(L);
    }
    if (d->n_point_lights && !d->point_lights)
        return fail(RUMI_ERR_INVALID, "rumi_scene_create: null point_lights with n_point_lights > 0");
    for (uint32_t i = 0; i < d->n_point_lights; ++i) {
        // PointLight(position, ls, rgb) (pointlight.cpp:16-21): L = cl * ls
        const rumi_point_light& q = d->point_lights[i];
        DevLight L;
        std::memset(&L, 0, sizeof(L));
        L.kind = kLightPoint;
        L.prim = -1;
        for (int k = 0; k < 3; ++k) {
            L.p[k] = q.position[k];
            L.le[k] = q.color[k] * q.ls;
        }
        L.p[3] = 1.0f;  // PointLight::Pdf
        lights.push_back(L);
    }

    // light index of each emitter primitive (word uv2[3]; -1 = not a light)
    for (uint32_t i = 0; i < n; ++i) {
        const int32_t none = -1;
        std::memcpy(&shapes[i].uv2[3], &none, 4);
    }
    for (size_t li = 0; li < lights.size(); ++li) {
        if (lights[li].prim < 0) continue;
        const int32_t idx = (int32_t)li;
        std::memcpy(&shapes[lights[li].prim].uv2[3], &idx, 4);
    }

    // ------------------------------------------------------------ textures
    std::vector<DevTexture> texs(d->n_textures);
    std::vector<float> texels;
    for (uint32_t i = 0; i < d->n_textures; ++i) {
        const rumi_texture& t = d->textures[i];
        texs[i] = DevTexture{};
        texs[i].kind = t.kind;
        if (t.kind == RUMI_TEXTURE_CHECKER) {
            texs[i].cw = t.checker_w;
            texs[i].ch = t.checker_h;
            std::memcpy(texs[i].c1, t.color_1, sizeof(texs[i].c1));
            std::memcpy(texs[i].c2, t.color_2, sizeof(texs[i].c2));
            texs[i].offset = (uint32_t)texels.size();
            continue;
        }
        if (t.kind != RUMI_TEXTURE_IMAGE && t.kind != RUMI_TEXTURE_HDR)
            return fail(RUMI_ERR_INVALID, "texture %u: unknown kind %d", i, t.kind);
        if (t.width <= 0 || t.height <= 0 || !t.rgb)
            return fail(RUMI_ERR_INVALID, "texture %u: bad size or data", i);
        texs[i].width = t.width;
        texs[i].height = t.height;
        texs[i].offset = (uint32_t)texels.size();
        texels.insert(texels.end(), t.rgb, t.rgb + 3 * (size_t)t.width * t.height);
    }

    rumi_scene* sc = new (std::nothrow) rumi_scene;
    if (!sc) return fail(RUMI_ERR_NO_MEMORY, "rumi_scene_create: out of host memory");
    sc->n_prims = n;
    sc->n_lights = (uint32_t)lights.size();
    sc->max_depth = flat.max_depth;
    sc->stack_need = 3 * wide.max_depth;
    sc->root_ref = wide.root_ref;
    std::memcpy(sc->root_lo, flat.root_lo, sizeof(sc->root_lo));
    std::memcpy(sc->root_hi, flat.root_hi, sizeof(sc->root_hi));
    sc->material_kinds = kinds;
    sc->has_ibl = sc_has_ibl;
    sc->has_point = d->n_point_lights > 0;
    sc->leaf_prims = flat.leaf_prims;

    // ------------------------------------------------------------ medium + camera
    const rumi_medium& md = d->medium;
    DevMedium& M = sc->medium;
    std::memset(&M, 0, sizeof(M));
    M.bound = md.bound;
    if (md.bound != RUMI_MEDIUM_NONE) {
        if (md.bound != RUMI_MEDIUM_BOX && md.bound != RUMI_MEDIUM_SPHERE) {
            delete sc;
            return fail(RUMI_ERR_INVALID, "medium: unknown bound %d", md.bound);
        }
        for (int k = 0; k < 3; ++k) {
            M.lo[k] = md.lo[k];
            M.hi[k] = md.hi[k];
            M.center[k] = md.center[k];
            M.sigma_s[k] = md.sigma_s[k];
            M.sigma_a[k] = md.sigma_a[k];
            M.sigma_t[k] = md.sigma_a[k] + md.sigma_s[k];
            if (!(M.sigma_t[k] > 0.0f)) {
                delete sc;
                return fail(RUMI_ERR_INVALID, "medium: sigma_t must be > 0");
            }
        }
        M.radius = md.radius;
        M.g = md.g;
        if (!(M.g > -1.0f && M.g < 1.0f)) {
            delete sc;
            return fail(RUMI_ERR_INVALID, "medium: g must be in (-1, 1)");
        }
        // HomogeneousVolume ctor (homogeneousvolume.cpp:27-38), Average = sum * 0.33333333333f
        M.avg_sigma_t = (M.sigma_t[2] + M.sigma_t[1] + M.sigma_t[0]) * 0.33333333333f;
        M.absorb = ((M.sigma_a[2] / M.sigma_t[2]) + (M.sigma_a[1] / M.sigma_t[1]) +
                    (M.sigma_a[0] / M.sigma_t[0])) * 0.33333333333f;
    }
    DevCamera& C = sc->camera;
    std::memcpy(C.p, d->camera.position, sizeof(C.p));
    std::memcpy(C.R, d->camera.basis, sizeof(C.R));
    // Pinhole ctor: _tanfov = tan(PI / 180.f * _fov * 0.5f) (pinhole.cpp:45)
    C.tanfov = (float)std::tan(kPi / 180.0f * (double)d->camera.fov_deg * 0.5f);
    C.y_ratio = d->camera.y_ratio;
    C.focal = d->camera.focal_distance;
    C.lens = d->camera.lens_radius;

    int rc = RUMI_OK;
    do {
        if (hipSetDevice(ctx->device) != hipSuccess) {
            rc = fail(RUMI_ERR_HIP, "hipSetDevice(%d) failed", ctx->device);
            break;
        }
        sc->device = ctx->device;
        if ((rc = upload(&sc->nodes, wide.nodes))) break;
        if ((rc = upload(&sc->prims, leaf))) break;
        if ((rc = upload(&sc->rects, rects))) break;
        if ((rc = upload(&sc->shapes, shapes))) break;
        if ((rc = upload(&sc->materials, mats))) break;
        if ((rc = upload(&sc->lights, lights))) break;
        if ((rc = upload(&sc->textures, texs))) break;
        if ((rc = upload(&sc->texels, texels))) break;
        sc->texel_count = (uint32_t)(texels.size() / 3);
        sc->tree_f4[0] = (uint32_t)(sizeof(DevNode4) / 16 * wide.nodes.size());
        sc->tree_f4[1] = (uint32_t)(sizeof(DevTri) / 16 * leaf.size());
        sc->tree_f4[2] = (uint32_t)(rects.size() * sizeof(rects[0]) / 16);
    } while (0);
    if (rc != RUMI_OK) {
        release(sc);
        delete sc;
        return rc;
    }
    *out = sc;
    return RUMI_OK;
}

extern "C" int rumi_scene_destroy(rumi_scene* s) {
    if (!s) return RUMI_OK;
    release(s);
    delete s;
    return RUMI_OK;
}
