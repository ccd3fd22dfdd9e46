// Device-resident scene layout shared by the scene builder (scene.cpp) and
// the renderer kernels (render.hip).  Not part of the C ABI.
#pragma once

#include <cstdint>
#include <vector>

#include "internal.h"

namespace rumi {

// Material record, 144 B.
struct alignas(16) DevMaterial {
    int32_t kind, opaque, pad0, pad1;
    float kd, ks, e, kd_pdf;                 // kd_pdf = kd / (kd + ks)
    float phong_norm, eta_in, eta_out, avg_sigmat;  // phong_norm = (e + 1) / 2pi
    float absorb, density, alpha, pad3;      // absorb = avg(sigma_a / sigma_t); GGX alpha
    float color[4];
    float absorption[4];
    float sigma_s[4];
    float sigma_a[4];
    float sigma_t[4];
};

// Per-primitive shading record in input order, 128 B.
//   g0..g2: the leaf record (triangle v0/e1/e2, sphere centre+radius)
//   n0..n2: triangle vertex normals .xyz and uv in .w (u0, v0, u1)
//   uv2:    (v1, u2, v2, 0)
//   info:   kind, material, texture, rect/light index
struct alignas(16) DevShape {
    float g0[4], g1[4], g2[4];
    float n0[4], n1[4], n2[4];
    float uv2[4];
    int32_t info[4];
};

// Area light, 96 B: rect (p, l, w, n) or sphere (centre, radius).
constexpr int32_t kLightIbl = 16;    // DevLight::kind of the image-based light (a sphere)
constexpr int32_t kLightPoint = 32;  // DevLight::kind of a PointLight: p, le; prim -1

struct alignas(16) DevLight {
    int32_t kind, prim, pad0, pad1;
    float p[4];   // rect p / sphere centre + radius in w
    float l[4];
    float w[4];
    float n[4];
    float le[4];  // emitted radiance ce * ls
};

struct DevMedium {
    int32_t bound;  // 0 none, 1 box, 2 sphere
    float g;
    float avg_sigma_t, absorb;
    float lo[3], hi[3];
    float center[3], radius;
    float sigma_s[3], sigma_a[3], sigma_t[3];
};

struct DevCamera {
    float p[3];
    float R[9];
    float tanfov, y_ratio;
    float focal, lens;  // Pinhole::DepthOfField; focal 0 = pinhole
};

struct DevTexture {
    int32_t width, height;
    uint32_t offset;  // float offset of the RGB data in the texel pool
    int32_t kind;     // RUMI_TEXTURE_*
    float cw, ch;     // checker: Checker(w, h, c1, c2)
    float c1[3], c2[3];
};

}  // namespace rumi

struct rumi_scene {
    int device = -1;
    uint32_t n_prims = 0;
    uint32_t n_lights = 0;
    int32_t max_depth = 0;   // binary tree depth
    int32_t stack_need = 0;  // traversal stack entries of the 4-wide form (3 per level)
    int32_t root_ref = 0;    // root of the 4-wide form
    uint32_t tree_f4[3] = {0, 0, 0};  // float4s of the device nodes, primitive records, rect records
    float root_lo[3] = {0, 0, 0}, root_hi[3] = {0, 0, 0};
    rumi::DevMedium medium{};
    rumi::DevCamera camera{};
    uint32_t material_kinds = 0;  // bit mask of RUMI_MAT_* present
    bool has_point = false;       // a PointLight (not for the bidirectional tracer)
    bool has_ibl = false;         // an image-based light (path tracer and RayTracer only, no medium)
    // device arrays
    void* nodes = nullptr;      // DevNode4[]: the scene tree in its 4-wide form
    void* prims = nullptr;      // leaf-ordered 48-B records
    void* rects = nullptr;      // 64-B rect records
    void* shapes = nullptr;     // DevShape[n_prims] in input order
    void* materials = nullptr;  // DevMaterial[]
    void* lights = nullptr;     // DevLight[]
    void* textures = nullptr;   // DevTexture[]
    void* texels = nullptr;     // float RGB pool
    uint32_t texel_count = 0;   // texels in the pool (3 floats each)
    // wavefront work buffers, grown on demand
    uint32_t capacity = 0;
    void* work = nullptr;
    size_t work_bytes = 0;
    // adjoint buffers (per-sample radiance, gradient stripes), grown on demand
    void* adj = nullptr;
    size_t adj_bytes = 0;
    // host-side copies kept for inspection
    std::vector<int32_t> leaf_prims;
};
