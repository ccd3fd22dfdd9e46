// rumi_renderer.hpp -- C++ classes with the reference renderer's names and
// constructor arguments, over the C ABI of librumi (include/rumi.h).
//
// A reference scene is written the way Scene::Build writes it
// (src/Scene/scene.cpp:135-204): shapes with materials (and textures), area
// lights, an optional HomogeneousVolume, a Pinhole, a Film and an
// Integrator; Scene::Rendering (src/Scene/scene.cpp:206-258) then renders
// the whole frame on the GPU with one rumi_render call and fills the Film.
// BVHAccel keeps its constructor and Travel; TravelBatch is the form a GPU
// caller uses.  Header-only: link with -lrumi.
//
// Differences a user of the reference sees:
//  - nothing runs on the CPU: with no HIP device Scene::Rendering records the
//    error (Scene::Status) and leaves the film untouched, as the reference
//    leaves its film on a failed file load;
//  - Color keeps the reference's member order (_b, _g, _r) and constructor
//    order (r, g, b);
//  - objects handed to a Scene are owned by it, as in the reference
//    (scene.cpp:42-74); a Shape owns its material and texture
//    (shape.cpp:8-15), except that an AreaLight's material is the shape's.
#pragma once

#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <memory>
#include <stdexcept>
#include <string>
#include <vector>

#include <hip/hip_runtime_api.h>

#include "rumi.h"

namespace rumi {

struct Error : std::runtime_error {
    int code;
    Error(int c, const std::string& what) : std::runtime_error(what), code(c) {}
};

inline void check(int rc, const char* what) {
    if (rc != RUMI_OK) throw Error(rc, std::string(what) + ": " + rumi_last_error());
}

// ------------------------------------------------------------------ kernel
struct Vector3 {  // src/Kernel/vector.h
    float _x = 0, _y = 0, _z = 0;
    Vector3() = default;
    Vector3(float x, float y, float z) : _x(x), _y(y), _z(z) {}
};
using Point3 = Vector3;

struct Color {  // src/Kernel/color.h:7-15: stored b, g, r; built r, g, b
    float _b = 0, _g = 0, _r = 0;
    Color() = default;
    explicit Color(float v) : _b(v), _g(v), _r(v) {}
    Color(float r, float g, float b) : _b(b), _g(g), _r(r) {}
};

struct Mat4x4 {  // src/Kernel/matrix.h: row-major 4 x 4
    float _m[16] = {1, 0, 0, 0, 0, 1, 0, 0, 0, 0, 1, 0, 0, 0, 0, 1};
};

// RotationY(angle) (src/Kernel/matrix.cpp): the Cornell camera uses RotationY(PI).
inline Mat4x4 RotationY(float a) {
    Mat4x4 m;
    const float c = std::cos(a), s = std::sin(a);
    m._m[0] = c;
    m._m[2] = s;
    m._m[8] = -s;
    m._m[10] = c;
    return m;
}

inline Mat4x4 RotationX(float a) {
    Mat4x4 m;
    const float c = std::cos(a), s = std::sin(a);
    m._m[5] = c;
    m._m[6] = -s;
    m._m[9] = s;
    m._m[10] = c;
    return m;
}

inline Mat4x4 RotationZ(float a) {
    Mat4x4 m;
    const float c = std::cos(a), s = std::sin(a);
    m._m[0] = c;
    m._m[1] = -s;
    m._m[4] = s;
    m._m[5] = c;
    return m;
}

inline Mat4x4 Translation(float x, float y, float z) {
    Mat4x4 m;
    m._m[3] = x;
    m._m[7] = y;
    m._m[11] = z;
    return m;
}

inline Mat4x4 Scaling(float x, float y, float z) {
    Mat4x4 m;
    m._m[0] = x;
    m._m[5] = y;
    m._m[10] = z;
    return m;
}

// Mat4x4::operator * (src/Kernel/matrix.cpp:18-40): each entry summed left
// to right in FP32.
inline Mat4x4 operator*(const Mat4x4& a, const Mat4x4& b) {
    Mat4x4 o;
    for (int i = 0; i < 4; ++i)
        for (int j = 0; j < 4; ++j) {
            float acc = a._m[i * 4] * b._m[j];
            for (int k = 1; k < 4; ++k) acc = acc + a._m[i * 4 + k] * b._m[k * 4 + j];
            o._m[i * 4 + j] = acc;
        }
    return o;
}

struct BBox {  // src/Kernel/bbox.h
    Point3 _min, _max;
    BBox() = default;
    BBox(const Point3& lo, const Point3& hi) : _min(lo), _max(hi) {}
};

// ---------------------------------------------------------------- materials
class Material {
  public:
    virtual ~Material() = default;
    rumi_material desc{};

  protected:
    explicit Material(int kind) { desc.kind = kind; desc.density = 1.0f; }
    void color(const Color& c) {
        desc.color[0] = c._r;
        desc.color[1] = c._g;
        desc.color[2] = c._b;
    }
    static void rgb(float* out, const Color& c) {
        out[0] = c._r;
        out[1] = c._g;
        out[2] = c._b;
    }
};

class Matte : public Material {  // src/Material/matte.cpp
  public:
    Matte(float kd, const Color& cd) : Material(RUMI_MAT_MATTE) { desc.kd = kd; color(cd); }
};

class Glossy : public Material {  // src/Material/glossy.cpp
  public:
    Glossy(float kd, float ks, float e, const Color& c) : Material(RUMI_MAT_GLOSSY) {
        desc.kd = kd;
        desc.ks = ks;
        desc.exponent = e;
        color(c);
    }
};

class Specular : public Material {  // src/Material/specular.cpp
  public:
    Specular(float kr, const Color& cr) : Material(RUMI_MAT_SPECULAR) { desc.kd = kr; color(cr); }
};

class Emissive : public Material {  // src/Material/emissive.cpp
  public:
    Emissive(float ls, const Color& ce) : Material(RUMI_MAT_EMISSIVE) { desc.kd = ls; color(ce); }
};

class Dielectric : public Material {  // src/Material/dielectric.cpp
  public:
    Dielectric(float eta_in, float eta_out, const Color& absorption) : Material(RUMI_MAT_DIELECTRIC) {
        desc.eta_in = eta_in;
        desc.eta_out = eta_out;
        rgb(desc.absorption, absorption);
    }
};

class Subsurface : public Material {  // src/Material/subsurface.cpp
  public:
    Subsurface(float eta_in, float eta_out, const Color& sigma_s, const Color& sigma_a,
               float density = 1.0f)
        : Material(RUMI_MAT_SUBSURFACE) {
        desc.eta_in = eta_in;
        desc.eta_out = eta_out;
        rgb(desc.sigma_s, sigma_s);
        rgb(desc.sigma_a, sigma_a);
        desc.density = density;
    }
};

// Extension (no reference counterpart): GGX microfacet reflection.
class GGX : public Material {
  public:
    GGX(float roughness, const Color& f0) : Material(RUMI_MAT_GGX) {
        desc.roughness = roughness;
        color(f0);
    }
};

// ----------------------------------------------------------------- textures
class Texture {
  public:
    virtual ~Texture() = default;
    virtual rumi_texture Desc() const = 0;
};

// ImageTexture (src/Texture/imagetexture.cpp): from RGB floats (row 0 = v =
// 0), or as the reference builds it, from a BMP file and a gamma.
class ImageTexture : public Texture {
  public:
    ImageTexture(int width, int height, const float* rgb)
        : _width(width), _height(height), _rgb(rgb, rgb + 3 * (size_t)width * height) {}
    explicit ImageTexture(const char* file_name, float gamma = 1.0f) {
        FILE* f = std::fopen(file_name, "rb");
        if (!f) {
            std::fprintf(stderr, "ImageTexture: cannot open %s\n", file_name);
            return;
        }
        std::vector<uint8_t> bytes;
        uint8_t buf[65536];
        size_t n;
        while ((n = std::fread(buf, 1, sizeof(buf), f)) > 0) bytes.insert(bytes.end(), buf, buf + n);
        std::fclose(f);
        uint32_t w = 0, h = 0, b = 0;
        if (rumi_bmp_read(bytes.data(), bytes.size(), &w, &h, &b, nullptr, 0) != RUMI_OK) return;
        std::vector<uint8_t> pix((size_t)w * h * b);
        if (rumi_bmp_read(bytes.data(), bytes.size(), &w, &h, &b, pix.data(), pix.size()) != RUMI_OK ||
            (b != 3 && b != 4))
            return;
        _rgb.resize(3 * (size_t)w * h);
        if (rumi_bmp_texels(pix.data(), w, h, b, gamma, _rgb.data()) != RUMI_OK) {
            _rgb.clear();
            return;
        }
        _width = (int)w;
        _height = (int)h;
    }
    rumi_texture Desc() const override {
        rumi_texture t{};
        t.width = _width;
        t.height = _height;
        t.rgb = _rgb.data();
        t.kind = RUMI_TEXTURE_IMAGE;
        return t;
    }
    int _width = 0, _height = 0;
    std::vector<float> _rgb;
};

// Checker(width, height, color_1, color_2) (src/Texture/checker.cpp:16-32).
class Checker : public Texture {
  public:
    Checker(float w, float h, const Color& c1, const Color& c2) : _w(w), _h(h), _c1(c1), _c2(c2) {}
    rumi_texture Desc() const override {
        rumi_texture t{};
        t.kind = RUMI_TEXTURE_CHECKER;
        t.checker_w = _w;
        t.checker_h = _h;
        t.color_1[0] = _c1._r, t.color_1[1] = _c1._g, t.color_1[2] = _c1._b;
        t.color_2[0] = _c2._r, t.color_2[1] = _c2._g, t.color_2[2] = _c2._b;
        return t;
    }
    float _w, _h;
    Color _c1, _c2;
};

// ------------------------------------------------------------------- shapes
class Shape {
  public:
    virtual ~Shape() {
        if (_owns_material) delete _material_ptr;
        delete _texture_ptr;
    }
    void AddingMaterial(Material* m) {
        if (_owns_material) delete _material_ptr;
        _material_ptr = m;
        _owns_material = true;
    }
    void AddingTexture(Texture* t) {
        delete _texture_ptr;
        _texture_ptr = t;
    }
    Material* MaterialPtr() const { return _material_ptr; }
    Texture* TexturePtr() const { return _texture_ptr; }
    virtual rumi_shape Desc() const = 0;
    virtual bool IsTriangle() const { return false; }

    bool _is_light = false;

  protected:
    Material* _material_ptr = nullptr;
    Texture* _texture_ptr = nullptr;
    bool _owns_material = false;
    friend class AreaLight;
    friend class Scene;
};

class Triangle : public Shape {  // src/Shape/triangle.cpp
  public:
    Triangle(const Point3& p0, const Point3& p1, const Point3& p2) : _v{p0, p1, p2} {}
    void Normal(const Vector3& n0, const Vector3& n1, const Vector3& n2) { _n[0] = n0, _n[1] = n1, _n[2] = n2; }
    rumi_shape Desc() const override {
        rumi_shape s{};
        s.kind = RUMI_SHAPE_TRIANGLE;
        for (int k = 0; k < 3; ++k) {
            s.v[3 * k + 0] = _v[k]._x, s.v[3 * k + 1] = _v[k]._y, s.v[3 * k + 2] = _v[k]._z;
            s.n[3 * k + 0] = _n[k]._x, s.n[3 * k + 1] = _n[k]._y, s.n[3 * k + 2] = _n[k]._z;
        }
        return s;
    }
    bool IsTriangle() const override { return true; }
    Point3 _v[3];
    Vector3 _n[3];  // zero until Normal() (triangle.cpp:22-30)
};

class Sphere : public Shape {  // src/Shape/sphere.cpp
  public:
    Sphere(const Point3& c, float r) : _c(c), _r(r) {}
    rumi_shape Desc() const override {
        rumi_shape s{};
        s.kind = RUMI_SHAPE_SPHERE;
        s.v[0] = _c._x, s.v[1] = _c._y, s.v[2] = _c._z;
        s.radius = _r;
        return s;
    }
    Point3 _c;
    float _r;
};

class Rect : public Shape {  // src/Shape/rect.cpp
  public:
    Rect(const Point3& p, const Vector3& l, const Vector3& w) : _p(p), _l(l), _w(w) {}
    rumi_shape Desc() const override {
        rumi_shape s{};
        s.kind = RUMI_SHAPE_RECT;
        const Vector3 a[3] = {_p, _l, _w};
        for (int k = 0; k < 3; ++k) s.v[3 * k] = a[k]._x, s.v[3 * k + 1] = a[k]._y, s.v[3 * k + 2] = a[k]._z;
        return s;
    }
    Point3 _p;
    Vector3 _l, _w;
};

// An indexed triangle mesh with one material: the triangle list a Model
// (src/Shape/model.cpp) holds, without the OBJ loader.
class Mesh : public Shape {
  public:
    Mesh(std::vector<float> xyz, std::vector<uint32_t> idx, std::vector<float> uv = {})
        : _xyz(std::move(xyz)), _idx(std::move(idx)), _uv(std::move(uv)) {}
    rumi_shape Desc() const override {
        rumi_shape s{};
        s.kind = RUMI_SHAPE_MESH;
        s.xyz = _xyz.data();
        s.nv = (uint32_t)(_xyz.size() / 3);
        s.idx = _idx.data();
        s.nt = (uint32_t)(_idx.size() / 3);
        s.uv = _uv.empty() ? nullptr : _uv.data();
        return s;
    }
    std::vector<float> _xyz;
    std::vector<uint32_t> _idx;
    std::vector<float> _uv;
};

// Model(file, transform) (src/Shape/model.cpp:35-43): an OBJ file loaded by
// rumi_obj_load with the reference's parsing and transforms; its triangles
// join the scene as one mesh with per-corner normals.
class Model : public Shape {
  public:
    Model(const char* file, const Mat4x4& transform) {
        rumi_obj* m = nullptr;
        if (rumi_obj_load(file, transform._m, &m) != RUMI_OK) {
            std::fprintf(stderr, "Model: %s\n", rumi_last_error());
            return;
        }
        uint32_t n = 0;
        const float *xyz = nullptr, *nrm = nullptr;
        rumi_obj_get(m, &n, &xyz, &nrm, nullptr);
        _xyz.assign(xyz, xyz + 9 * (size_t)n);
        _normals.assign(nrm, nrm + 9 * (size_t)n);
        _idx.resize(3 * (size_t)n);
        for (size_t i = 0; i < _idx.size(); ++i) _idx[i] = (uint32_t)i;
        rumi_obj_destroy(m);
    }
    rumi_shape Desc() const override {
        rumi_shape s{};
        s.kind = RUMI_SHAPE_MESH;
        s.xyz = _xyz.data();
        s.nv = (uint32_t)(_xyz.size() / 3);
        s.idx = _idx.data();
        s.nt = (uint32_t)(_idx.size() / 3);
        s.normals = _normals.empty() ? nullptr : _normals.data();
        return s;
    }
    uint32_t Triangles() const { return (uint32_t)(_idx.size() / 3); }
    std::vector<float> _xyz, _normals;
    std::vector<uint32_t> _idx;
};

// ------------------------------------------------------------------- lights
class Light {  // src/Light/light.h
  public:
    virtual ~Light() = default;
};

class AreaLight : public Light {  // src/Light/arealight.cpp: borrows its shape and material
  public:
    AreaLight(Shape* shape, Material* material) : _shape(shape) {
        shape->_is_light = true;
        if (material && material != shape->_material_ptr) shape->AddingMaterial(material);
    }
    Shape* _shape;
};

class PointLight : public Light {  // src/Light/pointlight.cpp:16-21; no shape
  public:
    PointLight(const Point3& position, float ls, const Color& rgb) {
        desc.position[0] = position._x, desc.position[1] = position._y, desc.position[2] = position._z;
        desc.ls = ls;
        desc.color[0] = rgb._r, desc.color[1] = rgb._g, desc.color[2] = rgb._b;
    }
    rumi_point_light desc{};
};

// -------------------------------------------------------------------- media
class HomogeneousVolume {  // src/Volume/homogeneousvolume.cpp:27-38
  public:
    HomogeneousVolume(const BBox& bound, const Color& sigma_s, const Color& sigma_a, float g = 0.0f) {
        desc.bound = RUMI_MEDIUM_BOX;
        desc.lo[0] = bound._min._x, desc.lo[1] = bound._min._y, desc.lo[2] = bound._min._z;
        desc.hi[0] = bound._max._x, desc.hi[1] = bound._max._y, desc.hi[2] = bound._max._z;
        set(sigma_s, sigma_a, g);
    }
    // Extension: a sphere-bounded medium.
    HomogeneousVolume(const Point3& c, float r, const Color& sigma_s, const Color& sigma_a, float g) {
        desc.bound = RUMI_MEDIUM_SPHERE;
        desc.center[0] = c._x, desc.center[1] = c._y, desc.center[2] = c._z;
        desc.radius = r;
        set(sigma_s, sigma_a, g);
    }
    rumi_medium desc{};

  private:
    void set(const Color& s, const Color& a, float g) {
        desc.sigma_s[0] = s._r, desc.sigma_s[1] = s._g, desc.sigma_s[2] = s._b;
        desc.sigma_a[0] = a._r, desc.sigma_a[1] = a._g, desc.sigma_a[2] = a._b;
        desc.g = g;
    }
};

// ------------------------------------------------------------- camera, film
class Pinhole {  // src/Camera/pinhole.cpp:30-51
  public:
    Pinhole(const Point3& pos, const Mat4x4& basis, float y_ratio) : _pos(pos), _basis(basis), _y_ratio(y_ratio) {}
    rumi_camera Desc() const {
        rumi_camera c{};
        c.position[0] = _pos._x, c.position[1] = _pos._y, c.position[2] = _pos._z;
        for (int r = 0; r < 3; ++r)
            for (int k = 0; k < 3; ++k) c.basis[3 * r + k] = _basis._m[4 * r + k];
        c.fov_deg = _fov;
        c.y_ratio = _y_ratio;
        c.focal_distance = _focal_distance;
        c.lens_radius = _lens_radius;
        return c;
    }
    // the thin lens (pinhole.cpp:89-95); focal_distance 0 = a pinhole
    void DepthOfField(float focal_distance, float lens_radius) {
        _focal_distance = focal_distance;
        _lens_radius = lens_radius;
    }
    Point3 _pos;
    Mat4x4 _basis;
    float _y_ratio;
    float _focal_distance = 0.0f, _lens_radius = 0.0f;
    float _fov = 20.0f;  // the reference's fixed field of view (pinhole.cpp:36); public, as there
};

class Film {  // src/Camera/film.cpp: RGB float accumulation, row 0 at the top
  public:
    Film(unsigned width, unsigned height) : _width(width), _height(height), _rgb(3 * (size_t)width * height, 0.f) {}
    Color Pixel(unsigned x, unsigned y) const {
        const float* p = &_rgb[3 * ((size_t)y * _width + x)];
        return Color(p[0], p[1], p[2]);
    }
    unsigned _width, _height;
    std::vector<float> _rgb;  // r, g, b per pixel (the per-pixel mean radiance)
};

// -------------------------------------------------------------- integrators
class Scene;
class Integrator {
  public:
    virtual ~Integrator() = default;
    virtual void Fill(rumi_render_desc& d) const = 0;

  protected:
    explicit Integrator(Scene* scene) : _scene(scene) {}
    Scene* _scene;
};

class PathTracer : public Integrator {  // src/Integrator/pathtracer.cpp
  public:
    PathTracer(Scene* scene, int max_depth) : Integrator(scene), _max_depth(max_depth) {}
    void Fill(rumi_render_desc& d) const override {
        d.integrator = RUMI_INTEGRATOR_PT;
        d.max_depth = _max_depth;
    }
    int _max_depth;
};

class BidirectionalPathTracer : public Integrator {  // src/Integrator/bidirectionalpathtracer.cpp
  public:
    BidirectionalPathTracer(Scene* scene, int n_light, int n_eye, float mis_beta = 1.0f)
        : Integrator(scene), _n_light(n_light), _n_eye(n_eye), _mis_beta(mis_beta) {}
    void Fill(rumi_render_desc& d) const override {
        d.integrator = RUMI_INTEGRATOR_BDPT;
        d.n_light = _n_light;
        d.n_eye = _n_eye;
        d.max_depth = _n_light + _n_eye - 1;  // longest path the two subpaths make
        d.mis_beta = _mis_beta;
    }
    int _n_light, _n_eye;
    float _mis_beta;  // 1 = the reference's balance heuristic
};

// ------------------------------------------------------------------ device
// The GPU a Scene or BVHAccel runs on (rumi_ctx).
class Device {
  public:
    explicit Device(int ordinal = 0) { check(rumi_ctx_create(ordinal, &_ctx), "rumi_ctx_create"); }
    ~Device() { rumi_ctx_destroy(_ctx); }
    Device(const Device&) = delete;
    Device& operator=(const Device&) = delete;
    rumi_ctx* get() const { return _ctx; }

  private:
    rumi_ctx* _ctx = nullptr;
};

// ---------------------------------------------------------------------- BVH
// BVHAccel(objs, n) over triangle shapes (src/Acceleration/bvh.cpp:40-148).
// Travel answers one ray (a GPU round trip; slow, kept for the reference's
// call sites); TravelBatch answers a device batch.  Hit indices are the
// caller's shape indices (the reference returns its permuted index).
class BVHAccel {
  public:
    BVHAccel(Device& dev, Shape** objs, const int& n, int mode = RUMI_BVH_SAH_BINNED) : _dev(dev) {
        std::vector<float> xyz;
        std::vector<uint32_t> idx;
        for (int i = 0; i < n; ++i) {
            if (!objs[i]->IsTriangle()) throw Error(RUMI_ERR_UNSUPPORTED, "BVHAccel: triangle shapes only");
            const rumi_shape s = objs[i]->Desc();
            for (int k = 0; k < 9; ++k) xyz.push_back(s.v[k]);
            for (int k = 0; k < 3; ++k) idx.push_back((uint32_t)(3 * i + k));
        }
        rumi_bvh_params p{mode, mode == RUMI_BVH_SAH_BINNED ? 4 : 0, 32, 0};
        check(rumi_bvh_build_mesh(dev.get(), xyz.data(), (uint32_t)(xyz.size() / 3), idx.data(), (uint32_t)n, &p,
                                  &_bvh),
              "BVHAccel");
    }
    ~BVHAccel() { rumi_bvh_destroy(_bvh); }
    BVHAccel(const BVHAccel&) = delete;
    BVHAccel& operator=(const BVHAccel&) = delete;

    // Closest hit (isect != nullptr semantics) or any hit: returns the shape
    // index or -1; *t_hit receives the hit distance (the reference's _tmin).
    int Travel(const rumi_ray& ray, float* t_hit, bool any = false) const {
        rumi_ray* d_ray = nullptr;
        rumi_hit* d_hit = nullptr;
        int32_t* d_any = nullptr;
        if (hipMalloc((void**)&d_ray, sizeof(rumi_ray)) != hipSuccess ||
            hipMalloc((void**)&d_hit, sizeof(rumi_hit)) != hipSuccess ||
            hipMalloc((void**)&d_any, sizeof(int32_t)) != hipSuccess) {
            (void)hipFree(d_ray), (void)hipFree(d_hit), (void)hipFree(d_any);
            throw Error(RUMI_ERR_NO_MEMORY, "BVHAccel::Travel: hipMalloc failed");
        }
        rumi_hit h{-1, 0.f};
        int32_t a = -1;
        int rc = hipMemcpy(d_ray, &ray, sizeof(ray), hipMemcpyHostToDevice) == hipSuccess ? RUMI_OK : RUMI_ERR_HIP;
        if (rc == RUMI_OK) rc = any ? rumi_trace_any(_dev.get(), _bvh, d_ray, 1, d_any, 0, nullptr)
                                    : rumi_trace_closest(_dev.get(), _bvh, d_ray, 1, d_hit, nullptr, 0, nullptr);
        if (rc == RUMI_OK && hipDeviceSynchronize() == hipSuccess) {
            (void)hipMemcpy(&h, d_hit, sizeof(h), hipMemcpyDeviceToHost);
            (void)hipMemcpy(&a, d_any, sizeof(a), hipMemcpyDeviceToHost);
        }
        (void)hipFree(d_ray), (void)hipFree(d_hit), (void)hipFree(d_any);
        check(rc, "BVHAccel::Travel");
        if (t_hit) *t_hit = h.t;
        return any ? a : h.prim;
    }
    // Batch form: device rays and hits, asynchronous on stream.
    void TravelBatch(const rumi_ray* dev_rays, uint32_t n, rumi_hit* dev_hits, void* stream = nullptr) const {
        check(rumi_trace_closest(_dev.get(), _bvh, dev_rays, n, dev_hits, nullptr, 0, stream), "TravelBatch");
    }
    void TravelBatchAny(const rumi_ray* dev_rays, uint32_t n, int32_t* dev_prim, void* stream = nullptr) const {
        check(rumi_trace_any(_dev.get(), _bvh, dev_rays, n, dev_prim, 0, stream), "TravelBatchAny");
    }

  private:
    Device& _dev;
    rumi_bvh* _bvh = nullptr;
};

// -------------------------------------------------------------------- scene
// Scene (src/Scene/scene.h:16-53): public object pointers as in the
// reference; Build() is the user's (the reference hard-codes it), Rendering
// renders the frame on the GPU into _image.
class Scene {
  public:
    // dev may be null for a scene that is only described (Describe()).
    Scene(Device* dev, int spp = 16, uint64_t seed = 1) : _dev(dev), _spp(spp), _seed(seed) {}
    virtual ~Scene() {  // scene.cpp:42-74
        delete _tracer_ptr;
        for (Light* l : _lights) delete l;
        for (Shape* s : _objects) delete s;
        delete _volume_ptr;
        delete _camera;
        delete _image;
        for (Material* m : _shared) delete m;
    }
    Scene(const Scene&) = delete;
    Scene& operator=(const Scene&) = delete;

    virtual void Build() {}

    void AddObject(Shape* s) { _objects.push_back(s); }
    void AddLight(Light* l) { _lights.push_back(l); }
    // A material several shapes use: owned by the scene, not by the shapes.
    Material* Share(Material* m) {
        _shared.push_back(m);
        return m;
    }
    void Bind(Shape* s, Material* shared) {
        if (s->_owns_material) delete s->_material_ptr;
        s->_material_ptr = shared;
        s->_owns_material = false;
    }

    // The rumi scene description of the object graph (host memory valid
    // while this Scene lives).
    const rumi_scene_desc& Describe() {
        _mats.clear();
        _shapes.clear();
        _texs.clear();
        std::vector<const Material*> seen;
        std::vector<const Texture*> tseen;
        for (Shape* o : _objects) {
            rumi_shape s = o->Desc();
            const Material* m = o->MaterialPtr();
            if (!m) throw Error(RUMI_ERR_INVALID, "Scene: a shape has no material");
            int mi = -1;
            for (size_t k = 0; k < seen.size(); ++k)
                if (seen[k] == m) mi = (int)k;
            if (mi < 0) {
                mi = (int)seen.size();
                seen.push_back(m);
                _mats.push_back(m->desc);
            }
            s.material = mi;
            s.is_light = o->_is_light ? 1 : 0;
            s.texture = -1;
            if (const Texture* t = o->TexturePtr()) {
                int ti = -1;
                for (size_t k = 0; k < tseen.size(); ++k)
                    if (tseen[k] == t) ti = (int)k;
                if (ti < 0) {
                    ti = (int)tseen.size();
                    tseen.push_back(t);
                    _texs.push_back(t->Desc());
                }
                s.texture = ti;
            }
            _shapes.push_back(s);
        }
        _points.clear();
        for (Light* l : _lights)
            if (const PointLight* p = dynamic_cast<const PointLight*>(l)) _points.push_back(p->desc);
        _desc = rumi_scene_desc{};
        _desc.point_lights = _points.empty() ? nullptr : _points.data();
        _desc.n_point_lights = (uint32_t)_points.size();
        _desc.shapes = _shapes.data();
        _desc.n_shapes = (uint32_t)_shapes.size();
        _desc.materials = _mats.data();
        _desc.n_materials = (uint32_t)_mats.size();
        _desc.textures = _texs.empty() ? nullptr : _texs.data();
        _desc.n_textures = (uint32_t)_texs.size();
        if (!_camera) throw Error(RUMI_ERR_INVALID, "Scene: no camera");
        _desc.camera = _camera->Desc();
        if (_volume_ptr) _desc.medium = _volume_ptr->desc;
        return _desc;
    }

    // Scene::Rendering (scene.cpp:206-258): the whole frame in one call.
    // *seconds receives the device time.  On failure the film is left as it
    // was and Status() / Message() tell why.
    void Rendering(float* seconds) {
        _status = RUMI_OK;
        _message.clear();
        try {
            if (!_dev) throw Error(RUMI_ERR_NO_DEVICE, "Scene: no device");
            if (!_image || !_tracer_ptr) throw Error(RUMI_ERR_INVALID, "Scene: no film or integrator");
            const rumi_scene_desc& d = Describe();
            rumi_scene* sc = nullptr;
            check(rumi_scene_create(_dev->get(), &d, &sc), "rumi_scene_create");
            rumi_render_desc rd{};
            _tracer_ptr->Fill(rd);
            rd.width = (int32_t)_image->_width;
            rd.height = (int32_t)_image->_height;
            rd.spp = _spp;
            rd.seed = _seed;
            if (rd.mis_beta == 0.0f) rd.mis_beta = 1.0f;
            float* d_rgb = nullptr;
            const size_t bytes = sizeof(float) * _image->_rgb.size();
            if (hipMalloc((void**)&d_rgb, bytes) != hipSuccess) {
                rumi_scene_destroy(sc);
                throw Error(RUMI_ERR_NO_MEMORY, "Scene::Rendering: hipMalloc failed");
            }
            rumi_render_stats st{};
            const int rc = rumi_render(_dev->get(), sc, &rd, d_rgb, &st, nullptr);
            const std::string msg = rc == RUMI_OK ? "" : rumi_last_error();
            if (rc == RUMI_OK && hipMemcpy(_image->_rgb.data(), d_rgb, bytes, hipMemcpyDeviceToHost) != hipSuccess) {
                (void)hipFree(d_rgb);
                rumi_scene_destroy(sc);
                throw Error(RUMI_ERR_HIP, "Scene::Rendering: copy of the film failed");
            }
            (void)hipFree(d_rgb);
            rumi_scene_destroy(sc);
            if (rc != RUMI_OK) throw Error(rc, "rumi_render: " + msg);
            if (seconds) *seconds = (float)(st.milliseconds * 1e-3);
            _stats = st;
        } catch (const Error& e) {
            _status = e.code;
            _message = e.what();
            std::fprintf(stderr, "Scene::Rendering: %s\n", e.what());
        }
    }

    int Status() const { return _status; }
    const std::string& Message() const { return _message; }
    const rumi_render_stats& Stats() const { return _stats; }

    Film* _image = nullptr;
    Pinhole* _camera = nullptr;
    std::vector<Shape*> _objects;
    std::vector<Light*> _lights;
    std::vector<rumi_point_light> _points;
    HomogeneousVolume* _volume_ptr = nullptr;
    Integrator* _tracer_ptr = nullptr;

  protected:
    Device* _dev;
    int _spp;
    uint64_t _seed;

  private:
    std::vector<Material*> _shared;
    std::vector<rumi_material> _mats;
    std::vector<rumi_shape> _shapes;
    std::vector<rumi_texture> _texs;
    rumi_scene_desc _desc{};
    rumi_render_stats _stats{};
    int _status = RUMI_OK;
    std::string _message;
};

}  // namespace rumi
