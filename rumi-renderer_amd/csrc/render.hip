// Wavefront path tracer for gfx950: PathTracer::Trace(w, h, camera, film)
// (src/Integrator/pathtracer.cpp:117-354) re-organised as queue stages.
//
// The reference evaluates a path in two passes: a forward loop that records
// each vertex, then a backward loop L = direct_k + vol_k fr_k L cos_k / pdf_k.
// Expanding that recursion gives the same estimator as a forward throughput
// sum, L = sum_k beta_k direct_k with beta_{k+1} = beta_k vol_k fr_k cos_k /
// pdf_k, which needs no per-vertex storage.  Each wavefront iteration runs
//   regen   -- new camera paths into free queue slots (per-pixel PCG32),
//   extend  -- closest hit of every queued ray (traverse.h),
//   bucket  -- per-material index lists (wave ballot + one atomic per wave),
//   shade   -- media, emission, next-event estimation, BSDF sampling; the
//              surviving paths and the shadow rays are appended compactly
//              to the next queue and the shadow queue,
//   shadow  -- any-hit of the shadow queue; unoccluded contributions are
//              added to the film with float atomics.
// Path state lives in HBM as structure-of-arrays buffers sized by the
// queue capacity.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cfloat>
#include <cstring>
#include <vector>

#include "scene.h"
#include "traverse.h"

using namespace rumi;
using namespace rumi::dev;

namespace {

constexpr int kBlock = 128;
constexpr int kStack = 16;
constexpr int kKeys = 8;     // material kinds 0..5, 6 = miss
constexpr int kKeyMiss = 6;
constexpr float kPiF = 3.14159265358979323846f;
constexpr float kInvPiF = 0.3183098865475128f;

// ------------------------------------------------------------------ RNG
struct Rng {
    uint64_t state, inc;
    __device__ uint32_t next() {
        const uint64_t old = state;
        state = old * 6364136223846793005ULL + inc;
        const uint32_t xs = (uint32_t)(((old >> 18u) ^ old) >> 27u);
        const uint32_t rot = (uint32_t)(old >> 59u);
        return (xs >> rot) | (xs << ((32u - rot) & 31u));
    }
    __device__ float uniform() { return (float)(next() >> 8) * (1.0f / 16777216.0f); }
};

__device__ inline Rng rng_seed(uint64_t initstate, uint64_t initseq) {
    Rng g;
    g.state = 0;
    g.inc = (initseq << 1u) | 1u;
    g.next();
    g.state += initstate;
    g.next();
    return g;
}

// ------------------------------------------------------------------ vectors
struct V3 {
    float x, y, z;
};
__device__ inline V3 v3(float x, float y, float z) { return V3{x, y, z}; }
__device__ inline V3 operator+(V3 a, V3 b) { return v3(a.x + b.x, a.y + b.y, a.z + b.z); }
__device__ inline V3 operator-(V3 a, V3 b) { return v3(a.x - b.x, a.y - b.y, a.z - b.z); }
__device__ inline V3 operator*(V3 a, float s) { return v3(a.x * s, a.y * s, a.z * s); }
__device__ inline V3 operator-(V3 a) { return v3(-a.x, -a.y, -a.z); }
__device__ inline float dot(V3 a, V3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
__device__ inline V3 cross(V3 u, V3 v) {
    return v3(u.y * v.z - v.y * u.z, u.z * v.x - v.z * u.x, u.x * v.y - v.x * u.y);
}
__device__ inline float length(V3 a) { return sqrtf(a.x * a.x + a.y * a.y + a.z * a.z); }
__device__ inline V3 normalize(V3 a) {
    const float inv = 1.0f / sqrtf(a.x * a.x + a.y * a.y + a.z * a.z);
    return v3(a.x * inv, a.y * inv, a.z * inv);
}
__device__ inline bool is_nan(V3 a) { return a.x != a.x || a.y != a.y || a.z != a.z; }
__device__ inline V3 cmul(V3 a, V3 b) { return v3(a.x * b.x, a.y * b.y, a.z * b.z); }
__device__ inline V3 cdiv(V3 a, V3 b) { return v3(a.x / b.x, a.y / b.y, a.z / b.z); }
__device__ inline V3 ld3(const float* p) { return v3(p[0], p[1], p[2]); }

// OrthonormalBasis (src/Kernel/vector.h:70-95).
__device__ inline void basis(V3 v0, V3& v1, V3& v2) {
    if (v0.x == 0.0f && v0.z == 0.0f && (v0.y == 1.0f || v0.y == -1.0f)) {
        v1 = v3(1, 0, 0);
        v2 = v0.y == 0.0f ? v3(0, 0, 1) : v3(0, 0, -1);
    } else {
        v1 = normalize(cross(v0, v3(0, 1, 0)));
        if (is_nan(v1)) v1 = v3(1, 0, 0);
        v2 = cross(v0, v1);
    }
}

__device__ inline V3 frame(V3 u, V3 v, V3 w, float su, float sv, float sw) {
    return (u * su + v * sv) + w * sw;
}

// ------------------------------------------------------------------ scene view
struct SceneView {
    const float4* nodes;
    const float4* prims;
    const float4* rects;
    const DevShape* shapes;
    const DevMaterial* mats;
    const DevLight* lights;
    const DevTexture* textures;
    const float* texels;
    RootInfo root;
    uint32_t n_lights;
    DevMedium med;
    DevCamera cam;
};

struct PathBuf {
    float4* ray;       // 2 per path: o.xyz tmin | d.xyz tmax
    float4* beta;      // throughput rgb
    uint64_t* rng;     // PCG32 state (increment derived from the pixel)
    uint32_t* pixel;
    uint32_t* flags;   // bit 0 le, bit 1 in_object, bits 8.. depth
};

struct ShadowBuf {
    float4* ray;       // 2 per entry
    float4* contrib;   // rgb, pixel (int bits) in w
};

struct Params {
    int32_t width, height, max_depth;
    uint32_t pixels;
    uint64_t seed;
};

__device__ inline uint32_t lane_id() { return __lane_id(); }

// Wave-aggregated append: one atomic per wave, lanes get consecutive slots.
__device__ inline uint32_t wave_append(uint32_t* counter, bool pred) {
    const uint64_t mask = __ballot(pred);
    if (mask == 0) return 0;
    const uint32_t cnt = (uint32_t)__popcll(mask);
    const int leader = __ffsll((unsigned long long)mask) - 1;
    uint32_t base = 0;
    if ((int)lane_id() == leader) base = atomicAdd(counter, cnt);
    base = __shfl(base, leader);
    const uint64_t below = lane_id() == 0 ? 0ull : (mask & (~0ull >> (64 - lane_id())));
    return base + (uint32_t)__popcll(below);
}

__device__ inline void film_add(float* film, uint32_t pixel, V3 c) {
    if (is_nan(c)) return;  // Scene::Rendering drops NaN radiance (scene.cpp:229-238)
    float* f = film + 3 * (size_t)pixel;
    atomicAdd(f + 0, c.x);
    atomicAdd(f + 1, c.y);
    atomicAdd(f + 2, c.z);
}

// ------------------------------------------------------------------ kernels
__global__ void k_zero(float* p, size_t n) {
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = 0.0f;
}

__global__ void k_zero_u32(uint32_t* p, uint32_t n) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = 0u;
}

__global__ void k_scale(float* p, size_t n, float s) {
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] *= s;
}

// Pinhole::CastViewRay (src/Camera/pinhole.cpp:97-136) for sample k:
// pixel = k mod P, per-pixel PCG32 stream.
__global__ __launch_bounds__(kBlock) void k_regen(PathBuf buf, uint32_t start, uint32_t count,
                                                  uint64_t sample_base, Params prm, DevCamera cam) {
    const uint32_t j = blockIdx.x * kBlock + threadIdx.x;
    if (j >= count) return;
    const uint32_t slot = start + j;
    const uint64_t k = sample_base + j;
    const uint32_t pixel = (uint32_t)(k % prm.pixels);
    const uint64_t s = k / prm.pixels;
    Rng g = rng_seed(prm.seed + s * 0x9E3779B97F4A7C15ull, pixel);
    const int x = (int)(pixel % (uint32_t)prm.width);
    const int y = (int)(pixel / (uint32_t)prm.width);
    const float sx = g.uniform(), sy = g.uniform();
    const float hw = prm.width * 0.5f, hh = prm.height * 0.5f;
    const float px = -(((float)x - hw) + sx) / hw * cam.tanfov;
    const float py = -(((float)y - hh) + sy) / hh * cam.tanfov * cam.y_ratio;
    const V3 l = normalize(v3(px, py, 1.0f));
    const V3 d = v3(cam.R[0] * l.x + cam.R[1] * l.y + cam.R[2] * l.z,
                    cam.R[3] * l.x + cam.R[4] * l.y + cam.R[5] * l.z,
                    cam.R[6] * l.x + cam.R[7] * l.y + cam.R[8] * l.z);
    buf.ray[2 * (size_t)slot + 0] = make_float4(cam.p[0], cam.p[1], cam.p[2], 0.0f);
    buf.ray[2 * (size_t)slot + 1] = make_float4(d.x, d.y, d.z, FLT_MAX);
    buf.beta[slot] = make_float4(1.0f, 1.0f, 1.0f, 0.0f);
    buf.rng[slot] = g.state;
    buf.pixel[slot] = pixel;
    buf.flags[slot] = 1u;  // le = true, outside objects, depth 0
}

template <bool SPILL>
__global__ __launch_bounds__(kBlock) void k_extend(PathBuf buf, uint32_t n, SceneView sv,
                                                   int2* hits, float2* uvs) {
    __shared__ int32_t s_ref[kStack * kBlock];
    __shared__ float s_tn[kStack * kBlock];
    const uint32_t tid = threadIdx.x;
    const uint32_t i = blockIdx.x * kBlock + tid;
    if (i >= n) return;
    const float4 r0 = buf.ray[2 * (size_t)i + 0];
    const float4 r1 = buf.ray[2 * (size_t)i + 1];
    Ray r{r0.x, r0.y, r0.z, r1.x, r1.y, r1.z, r1.w, r0.w};
    const bool in_object = (buf.flags[i] & 2u) != 0;
    float hu = 0.0f, hv = 0.0f;
    int32_t nn = 0, np = 0;
    const int32_t hit = traverse<kBlock, kStack, false, true, true, false, SPILL>(
        sv.nodes, sv.prims, sv.rects, sv.root, r, in_object, hu, hv, nn, np, s_ref, s_tn, tid);
    hits[i] = make_int2(hit, __float_as_int(r.tmin));
    uvs[i] = make_float2(hu, hv);
}

__global__ __launch_bounds__(kBlock) void k_bucket(uint32_t n, const int2* hits, SceneView sv,
                                                   uint32_t* lists, uint32_t* counts,
                                                   uint32_t cap) {
    const uint32_t i = blockIdx.x * kBlock + threadIdx.x;
    const bool active = i < n;
    int key = kKeyMiss;
    if (active) {
        const int32_t prim = hits[i].x;
        if (prim >= 0) key = sv.mats[sv.shapes[prim].info[1]].kind;
    }
    for (int k = 0; k < kKeys; ++k) {
        const bool mine = active && key == k;
        const uint32_t slot = wave_append(&counts[k], mine);
        if (mine) lists[(size_t)k * cap + slot] = i;
    }
}

// ------------------------------------------------------------ media helpers
__device__ inline bool slab_f(const float* lo, const float* hi, V3 o, V3 d, float& tn, float& tf) {
    return slab(lo[0], lo[1], lo[2], hi[0], hi[1], hi[2], o.x, o.y, o.z, 1.0f / d.x, 1.0f / d.y,
                1.0f / d.z, tn, tf);
}

// HomogeneousVolume::Intersect (homogeneousvolume.cpp:113-133); the sphere
// bound is an extension.  thit = the ray's closest surface hit (0 = none).
__device__ inline bool med_intersect(const DevMedium& m, V3 o, V3 d, float thit, float& tn,
                                     float& tf) {
    tn = 0.0f;
    tf = 0.0f;
    if (m.bound == 1) {
        if (!slab_f(m.lo, m.hi, o, d, tn, tf)) return false;
    } else {
        const V3 oc = o - ld3(m.center);
        const float B = dot(d, oc), C = dot(oc, oc) - m.radius * m.radius, A = dot(d, d);
        const float disc = B * B - A * C;
        if (disc < 0.0f) return false;
        const float sq = sqrtf(disc);
        tn = (-B - sq) / A;
        tf = (-B + sq) / A;
    }
    if (tn != tn || tf != tf) return false;
    if (0.0f < thit && thit < tn) return false;
    if (tf < 0.0f) return false;
    return true;
}

// HomogeneousVolume::MaxDistance (homogeneousvolume.cpp:135-166).
__device__ inline float med_maxdist(float thit, float tn, float tf) {
    if (tn < 0.0f) {
        if (thit == 0.0f) return tf;
        return thit < tf ? thit : tf;
    }
    if (thit == 0.0f) return tf - tn;
    return thit < tf ? thit - tn : tf - tn;
}

__device__ inline V3 exp_tau(const float* st, float d) {
    return v3(expf(st[0] * -d), expf(st[1] * -d), expf(st[2] * -d));
}

enum { kPass = 0, kScatter = 1, kAbsorb = 2 };

// HomogeneousVolume::DistanceSampling (homogeneousvolume.cpp:168-215) and
// Subsurface::DistanceSampling (subsurface.cpp:49-99).
__device__ inline int distance_sample(const float* st, float avg, float absorb, float smax,
                                      Rng& g, float& s, V3& ratio) {
    const float xi = g.uniform();
    V3 tau, pdf;
    if (xi >= 1.0f - expf(avg * -smax)) {
        tau = exp_tau(st, smax);
        if (tau.x == 0.0f && tau.y == 0.0f && tau.z == 0.0f) tau = v3(1e-4f, 1e-4f, 1e-4f);
        ratio = cdiv(tau, tau);
        return kPass;
    }
    const float ds = -logf(1.0f - xi) / avg;
    if (smax - ds < 1e-4f) {
        tau = exp_tau(st, smax);
        if (tau.x == 0.0f && tau.y == 0.0f && tau.z == 0.0f) tau = v3(1e-4f, 1e-4f, 1e-4f);
        ratio = cdiv(tau, tau);
        return kPass;
    }
    if (g.uniform() < absorb || ds < 1e-4f) return kAbsorb;
    s = ds;
    tau = exp_tau(st, ds);
    pdf = cmul(tau, ld3(st));
    if (pdf.x == 0.0f && pdf.y == 0.0f && pdf.z == 0.0f) pdf = v3(1e-4f, 1e-4f, 1e-4f);
    ratio = cdiv(tau, pdf);
    return kScatter;
}

__device__ inline float hg_eval(float g, float c) {
    const float den = 1.0f + g * g - 2.0f * g * c;
    return (1.0f - g * g) / (4.0f * kPiF * den * sqrtf(den));
}

// Isotropic::IsotropicSampling (phase.cpp:40-56) for g = 0; Henyey-Greenstein
// about the propagation direction d otherwise.  Returns the new direction and
// the phase value (= its pdf).
__device__ inline V3 phase_sample(float g, V3 d, Rng& rng, float& pdf) {
    V3 u, w;
    basis(d, u, w);
    const float phi = rng.uniform() * 2.0f * kPiF;
    float ct;
    if (g == 0.0f) {
        ct = cosf(acosf(rng.uniform() * 2.0f - 1.0f));
        pdf = 1.0f / (4.0f * kPiF);
    } else {
        const float x = rng.uniform();
        const float sq = (1.0f - g * g) / (1.0f - g + 2.0f * g * x);
        ct = fminf(1.0f, fmaxf(-1.0f, (1.0f + g * g - sq * sq) / (2.0f * g)));
        pdf = hg_eval(g, ct);
    }
    const float st = sqrtf(fmaxf(0.0f, 1.0f - ct * ct));
    return frame(u, d, w, st * sinf(phi), ct, st * cosf(phi));
}

// HomogeneousVolume::ShadowRayTr (homogeneousvolume.cpp:56-93): grey
// Bernoulli transmittance of a shadow segment, 0 or 1.
__device__ inline float shadow_tr(const DevMedium& m, V3 o, V3 d, float tmax, Rng& g) {
    if (m.bound == 0) return 1.0f;
    float tn, tf;
    if (!med_intersect(m, o, d, 0.0f, tn, tf)) return 1.0f;
    const float smax = med_maxdist(tmax, tn, tf);
    const float xi = g.uniform();
    if (xi < 1.0f - expf(m.avg_sigma_t * -smax)) return 0.0f;
    return 1.0f;
}

// ------------------------------------------------------------ light sampling
struct LightSample {
    V3 p, n, le;
    float pdf;
};

// Scene::VisibilityTest (scene.cpp:99-123) up to the shadow ray: uniform
// light choice, a uniform point on the emitter (the reference draws from a
// fixed stratified set), the one-sided test.  Returns false if unlit.
__device__ inline bool sample_light(const SceneView& sv, V3 x, Rng& g, LightSample& ls, V3& wi,
                                    float& dist) {
    if (sv.n_lights == 0) return false;
    uint32_t li = (uint32_t)(g.uniform() * (float)sv.n_lights);
    if (li >= sv.n_lights) li = sv.n_lights - 1;
    const DevLight& L = sv.lights[li];
    if (L.kind == RUMI_SHAPE_RECT) {
        const float a = g.uniform(), b = g.uniform();
        ls.p = (ld3(L.p) + ld3(L.l) * a) + ld3(L.w) * b;
        ls.n = ld3(L.n);
    } else {
        const float z = 1.0f - 2.0f * g.uniform(), phi = 2.0f * kPiF * g.uniform();
        const float r = sqrtf(fmaxf(0.0f, 1.0f - z * z));
        ls.p = ld3(L.p) + v3(r * cosf(phi), r * sinf(phi), z) * L.l[0];
        ls.n = normalize(ls.p - ld3(L.p));
    }
    ls.pdf = 1.0f / (float)sv.n_lights * L.p[3];
    ls.le = ld3(L.le);
    wi = normalize(ls.p - x);
    dist = length(ls.p - x);
    return !(dot(wi, ls.n) >= 0.0f);
}

// AreaLight::G (arealight.cpp:106-112).
__device__ inline float light_g(V3 x, const LightSample& ls, V3 wi) {
    const float d = length(x - ls.p);
    const float c = fmaxf(-dot(wi, ls.n), 0.0f);
    return c / (d * d);
}

// ------------------------------------------------------------ textures
__device__ inline V3 texel(const SceneView& sv, int tex, float u, float v) {
    const DevTexture t = sv.textures[tex];
    // ImageTexture::Pixel (imagetexture.cpp:32-61): nearest texel
    const int row = (int)((1.0f - v) * (float)(t.height - 1));
    const int col = (int)(u * (float)(t.width - 1));
    const int y = min(max(t.height - 1 - row, 0), t.height - 1);
    const int x = min(max(col, 0), t.width - 1);
    const float* p = sv.texels + t.offset + 3 * ((size_t)y * t.width + x);
    return v3(p[0], p[1], p[2]);
}

// Shape::TextureMapping overrides: Rect (rect.cpp:183-197), Sphere
// (sphere.cpp:172-195, identity transform); triangles interpolate vertex uv.
__device__ inline V3 pattern(const SceneView& sv, const DevShape& s, V3 p, float bu, float bv) {
    const int tex = s.info[2];
    if (tex < 0) return v3(1, 1, 1);
    float u = 0.0f, v = 0.0f;
    if (s.info[0] == RUMI_SHAPE_TRIANGLE) {
        const float w0 = 1.0f - bu - bv;
        u = w0 * s.n0[3] + bu * s.n2[3] + bv * s.uv2[1];
        v = w0 * s.n1[3] + bu * s.uv2[0] + bv * s.uv2[2];
    } else if (s.info[0] == RUMI_SHAPE_SPHERE) {
        const V3 pc = normalize(p - v3(s.g0[0], s.g0[1], s.g0[2]));
        const float theta = acosf(pc.y);
        const V3 pxz = pc * sinf(theta);
        const float phi = atan2f(pxz.x, pxz.z);
        u = phi >= 0.0f ? phi / (2.0f * kPiF) : (phi + 2.0f * kPiF) / (2.0f * kPiF);
        v = 1.0f - theta / kPiF;
    } else {
        const float* r = &s.g0[0];  // rect: p in g0.xyz, l in g1.zw g2.x, w in g2.yzw
        const V3 onp = p - v3(r[0], r[1], r[2]);
        const V3 L = v3(s.g1[2], s.g1[3], s.g2[0]);
        const V3 W = v3(s.g2[1], s.g2[2], s.g2[3]);
        const float len_l = length(L), len_w = length(W);
        u = fminf(fmaxf(dot(onp, L) / (len_l * len_l), 0.0f), 1.0f);
        v = fminf(fmaxf(dot(onp, W) / (len_w * len_w), 0.0f), 1.0f);
    }
    return texel(sv, tex, u, v);
}

// ------------------------------------------------------------ BSDF pieces
// Lambertian::BSDFSampling (bsdf.cpp:43-69).
__device__ inline V3 lambert_sample(V3 n, Rng& g, float& cost, float& pdf) {
    V3 u, w;
    basis(n, u, w);
    const float phi = g.uniform() * 2.0f * kPiF;
    const float theta = acosf(1.0f - g.uniform() * 2.0f) * 0.5f;
    const float st = sinf(theta);
    const V3 wi = frame(u, n, w, st * sinf(phi), cosf(theta), st * cosf(phi));
    cost = dot(wi, n);
    if (cost <= 0.0f) cost = 1e-4f;
    pdf = kInvPiF * cost;
    if (pdf <= 0.0f) pdf = 1e-4f;
    return wi;
}

// Phong::F (bsdf.cpp:91-104): white lobe.
__device__ inline float phong_f(const DevMaterial& m, V3 n, V3 wi, V3 wo) {
    float c = dot(wi, n);
    if (c <= 0.0f) c = 1e-4f;
    const V3 r = -wi + n * (c * 2.0f);
    const float ca = powf(dot(wo, r), m.e);
    return ca >= 1e-4f ? m.phong_norm * ca : m.phong_norm * 1e-4f;
}

// Phong::BSDFSampling (bsdf.cpp:127-170): the returned cosine is 1.
__device__ inline V3 phong_sample(const DevMaterial& m, V3 wo, V3 n, Rng& g, float& cost,
                                  float& pdf) {
    float won = dot(wo, n);
    if (won <= 0.0f) won = 1e-4f;
    const V3 v = -wo + n * won * 2.0f;
    V3 u, w;
    basis(v, u, w);
    const float phi = g.uniform() * 2.0f * kPiF;
    const float theta = acosf(powf(g.uniform(), 1.0f / (1.0f + m.e)));
    const float su = sinf(theta) * sinf(phi), sv = cosf(theta), sw = sinf(theta) * cosf(phi);
    V3 wi = frame(u, v, w, su, sv, sw);
    cost = dot(wi, n);
    if (cost < 0.0f) {
        cost = -cost;
        wi = frame(u, v, w, -su, sv, -sw);
    }
    if (cost == 0.0f) cost = 1e-4f;
    const float ca = powf(dot(wi, v), m.e);
    pdf = ca >= 1e-4f ? m.phong_norm * ca : m.phong_norm * 1e-4f;
    if (pdf <= 0.0f) pdf = 1e-4f;
    cost = 1.0f;
    return wi;
}

// FresnelReflection::Fresnel (bsdf.cpp:255-283).
__device__ inline float fresnel(V3 dir, V3 n, float eta_in, float eta_out) {
    float eta_i = eta_out, eta_t = eta_in, ci = -dot(dir, n);
    if (ci < 0.0f) {
        eta_i = eta_in;
        eta_t = eta_out;
        ci = -ci;
    }
    const float inv_eta = 1.0f / (eta_t / eta_i);
    float ct = 1.0f - (1.0f - ci * ci) * (inv_eta * inv_eta);
    if (ct <= 0.0f) return 1.0f;
    ct = sqrtf(ct);
    const float rs = (eta_i * ci - eta_t * ct) / (eta_i * ci + eta_t * ct);
    const float rp = (eta_i * ct - eta_t * ci) / (eta_i * ct + eta_t * ci);
    return (rs * rs + rp * rp) * 0.5f;
}

// ------------------------------------------------------------ shading
struct ShadeOut {
    uint32_t pixel;
    bool alive;
    float4 r0, r1, beta;
    uint64_t rng;
    uint32_t flags;
    bool shadow;
    float4 s0, s1, contrib;
};

__device__ inline void push_shadow(ShadeOut& o, V3 x, V3 wi, float dist, V3 c, uint32_t pixel) {
    o.shadow = true;
    o.s0 = make_float4(x.x, x.y, x.z, 0.0f);
    o.s1 = make_float4(wi.x, wi.y, wi.z, dist);
    o.contrib = make_float4(c.x, c.y, c.z, __int_as_float((int)pixel));
}

__device__ inline void push_next(ShadeOut& o, V3 p, V3 w, V3 beta, uint32_t flags) {
    o.alive = true;
    o.r0 = make_float4(p.x, p.y, p.z, 0.0f);
    o.r1 = make_float4(w.x, w.y, w.z, FLT_MAX);
    o.beta = make_float4(beta.x, beta.y, beta.z, 0.0f);
    o.flags = flags;
}

__device__ inline uint32_t mk_flags(bool le, bool in_obj, uint32_t depth) {
    return (le ? 1u : 0u) | (in_obj ? 2u : 0u) | (depth << 8);
}

// Matte::DirectLighting (matte.cpp:54-82) without the occlusion test, which
// the shadow stage performs.
__device__ inline bool matte_direct(const SceneView& sv, const DevMaterial& m, V3 p, V3 n, Rng& g,
                                    V3& lo, V3& wi, float& dist) {
    LightSample ls;
    if (!sample_light(sv, p, g, ls, wi, dist)) return false;
    const V3 f = ld3(m.color) * (kInvPiF * m.kd);
    const float c = fmaxf(dot(wi, n), 0.0f);
    lo = cmul(f, ls.le) * (c * light_g(p, ls, wi) / ls.pdf);
    if (sv.med.bound) lo = lo * shadow_tr(sv.med, p, wi, dist, g);
    return true;
}

__device__ void shade(const SceneView& sv, const Params& prm, V3 o, V3 d, uint32_t flags, V3 beta,
                      Rng& g, uint32_t pixel, int32_t prim, float t, float bu, float bv, float* film,
                      ShadeOut& out) {
    const bool le = (flags & 1u) != 0;
    const bool in_obj = (flags & 2u) != 0;
    const uint32_t depth = flags >> 8;
    const bool last = (int)depth + 1 > prm.max_depth - 1;
    const DevMedium& md = sv.med;

    // Homogeneous medium along the segment (pathtracer.cpp:156-178).
    if (md.bound && !in_obj) {
        float tn, tf;
        if (med_intersect(md, o, d, prim >= 0 ? t : 0.0f, tn, tf)) {
            const float smax = med_maxdist(prim >= 0 ? t : 0.0f, tn, tf);
            float s = 0.0f;
            V3 ratio;
            const int ev = distance_sample(md.sigma_t, md.avg_sigma_t, md.absorb, smax, g, s, ratio);
            if (ev == kAbsorb) return;
            beta = cmul(beta, ratio);
            if (ev == kScatter) {
                const V3 p = o + d * (tn < 0.0f ? s : tn + s);
                // HomogeneousVolume::DirectLighting (homogeneousvolume.cpp:231-250)
                LightSample ls;
                V3 wi;
                float dist;
                if (sample_light(sv, p, g, ls, wi, dist)) {
                    const float tr = shadow_tr(md, p, wi, dist, g);
                    const float fp = md.g == 0.0f ? 1.0f / (4.0f * kPiF) : hg_eval(md.g, dot(d, wi));
                    const V3 c = cmul(cmul(beta, ls.le), ld3(md.sigma_s)) *
                                 (tr * fp * light_g(p, ls, wi) / ls.pdf);
                    if (tr > 0.0f) push_shadow(out, p, wi, dist, c, pixel);
                }
                if (last) return;
                float pdf;
                const V3 w = phase_sample(md.g, d, g, pdf);
                // Fp / pdf = sigma_s (phase value equals its pdf)
                beta = cmul(beta, ld3(md.sigma_s));
                push_next(out, p, w, beta, mk_flags(false, false, depth + 1));
                return;
            }
        }
    }
    if (prim < 0) return;

    const DevShape& s = sv.shapes[prim];
    const DevMaterial& m = sv.mats[s.info[1]];
    // Surface point and normal: Triangle/Sphere/Rect::Intersect vertex writes
    // (triangle.cpp:104-114, sphere.cpp:110-129, rect.cpp:112-118).
    V3 p = o + d * t;
    V3 n;
    if (s.info[0] == RUMI_SHAPE_TRIANGLE) {
        n = (ld3(s.n0) * (1.0f - bu - bv) + ld3(s.n1) * bu) + ld3(s.n2) * bv;
        const V3 e1 = v3(s.g0[3], s.g1[0], s.g1[1]);
        const V3 e2 = v3(s.g1[2], s.g1[3], s.g2[0]);
        const V3 pv = v3(d.y * e2.z - e2.y * d.z, d.z * e2.x - e2.z * d.x, d.x * e2.y - e2.x * d.y);
        const float det = pv.x * e1.x + pv.y * e1.y + pv.z * e1.z;
        if (det < 0.0f && m.opaque) n = -n;
    } else if (s.info[0] == RUMI_SHAPE_SPHERE) {
        const V3 c = v3(s.g0[0], s.g0[1], s.g0[2]);
        if (m.kind == RUMI_MAT_EMISSIVE && !(length(o - c) > s.g0[3]))
            n = normalize(c - p);
        else
            n = normalize(p - c);
    } else {
        n = v3(s.g0[3], s.g1[0], s.g1[1]);
    }

    if (m.kind == RUMI_MAT_EMISSIVE) {
        // Emissive::DirectLighting (emissive.cpp:60-74)
        if (!in_obj && le && -dot(d, n) > 0.0f) film_add(film, pixel, cmul(beta, ld3(m.color) * m.kd));
        return;
    }

    V3 vol = v3(1, 1, 1);  // Subsurface pass-through tau / pdf at the exit boundary
    if (m.kind == RUMI_MAT_SUBSURFACE && in_obj) {
        // Subsurface::PhaseSampling (subsurface.cpp:135-155)
        float sd = 0.0f;
        V3 ratio;
        const int ev = distance_sample(m.sigma_t, m.avg_sigmat, m.absorb, t, g, sd, ratio);
        if (ev == kAbsorb) return;
        if (ev == kScatter) {
            // vol * Fp / pdf_w with Fp = sigma_s / 4pi and pdf_w = 1 / 4pi
            beta = cmul(cmul(beta, ratio), ld3(m.sigma_s));
            if (last) return;
            const V3 ps = o + d * sd;
            float pdf;
            const V3 w = phase_sample(0.0f, d, g, pdf);
            push_next(out, ps, w, beta, mk_flags(true, true, depth + 1));
            return;
        }
        vol = ratio;
    }

    const V3 tex = pattern(sv, s, p, bu, bv);
    const V3 wo = -d;
    float xi = 1.0f, cost = 1.0f, pdf_w = 1.0f, pdf_path = 1.0f;
    bool le2 = false, in2 = false;
    V3 w;
    switch (m.kind) {
    case RUMI_MAT_MATTE:
        w = lambert_sample(n, g, cost, pdf_w);
        break;
    case RUMI_MAT_GLOSSY:
        xi = g.uniform();
        if (m.kd != 0.0f && xi <= m.kd_pdf) {
            w = lambert_sample(n, g, cost, pdf_w);
            pdf_path = m.kd_pdf;
        } else {
            w = phong_sample(m, wo, n, g, cost, pdf_w);
            le2 = true;
            pdf_path = 1.0f - m.kd_pdf;
        }
        break;
    case RUMI_MAT_SPECULAR: {
        // PerfectReflection::BSDFSampling (bsdf.cpp:197-207)
        float won = dot(wo, n);
        if (won <= 0.0f) won = 1e-4f;
        w = -wo + n * (2.0f * won);
        le2 = true;
        break;
    }
    default: {  // dielectric, subsurface boundary: Dielectric::BSDFSampling (dielectric.cpp:43-71)
        const float kr = fresnel(-wo, n, m.eta_in, m.eta_out);
        const float kt = 1.0f - kr;
        if (kr != 0.0f && g.uniform() <= kr) {
            // FresnelReflection::BSDFSampling (bsdf.cpp:232-253)
            const float c = dot(wo, n);
            float won = c;
            V3 nn = n;
            if (won < 1e-4f) {
                if (won > -1e-4f)
                    won = 1e-4f;
                else {
                    nn = -n;
                    won = -won;
                }
            }
            w = -wo + nn * (2.0f * won);
            in2 = c < 0.0f;
            xi = 1.0f;
            pdf_path = kr;
        } else {
            // FresnelTransparency::BSDFSampling (bsdf.cpp:322-342)
            float eta = m.eta_in / m.eta_out;
            V3 nn = n;
            const float c = dot(wo, n);
            float ci = c;
            if (ci < 0.0f) {
                nn = -n;
                eta = m.eta_out / m.eta_in;
                ci = -ci;
            }
            const float inv = 1.0f / eta;
            const float ct = sqrtf(1.0f - (1.0f - ci * ci) * (inv * inv));
            w = -wo * inv + nn * (ci * inv - ct);
            in2 = c > 0.0f;
            xi = 0.0f;
            pdf_path = kt;
        }
        le2 = true;
        break;
    }
    }

    // Next-event estimation at this vertex (Material::DirectLighting).
    if (m.kind == RUMI_MAT_MATTE || (m.kind == RUMI_MAT_GLOSSY && m.kd != 0.0f && xi <= m.kd_pdf)) {
        V3 lo, wi;
        float dist;
        if (matte_direct(sv, m, p, n, g, lo, wi, dist)) {
            if (m.kind == RUMI_MAT_GLOSSY) lo = lo * (1.0f / m.kd_pdf);
            const V3 c = cmul(cmul(beta, tex), lo);
            if (c.x != 0.0f || c.y != 0.0f || c.z != 0.0f) push_shadow(out, p, wi, dist, c, pixel);
        }
    }
    if (last) return;

    // fr of this vertex (Material::Fr) with the sampled lobe.
    V3 fr;
    switch (m.kind) {
    case RUMI_MAT_MATTE:
        fr = ld3(m.color) * (kInvPiF * m.kd);
        break;
    case RUMI_MAT_GLOSSY:
        if (m.kd != 0.0f && xi <= m.kd_pdf)
            fr = ld3(m.color) * (kInvPiF * m.kd);
        else {
            const float f = phong_f(m, n, w, wo) * m.ks;
            fr = v3(f, f, f);
        }
        break;
    case RUMI_MAT_SPECULAR:
        fr = ld3(m.color) * m.kd;
        break;
    default: {
        float k;
        if (xi == 1.0f) {
            k = pdf_path;
        } else {
            const float eta = in_obj ? m.eta_in / m.eta_out : m.eta_out / m.eta_in;
            k = 1.0f / (eta * eta) * pdf_path;
        }
        fr = v3(k, k, k);
        if (in_obj) {
            if (m.kind == RUMI_MAT_DIELECTRIC)
                fr = cmul(fr, v3(expf(-m.absorption[0] * t), expf(-m.absorption[1] * t),
                                 expf(-m.absorption[2] * t)));
            else
                fr = cmul(fr, vol);
        }
        break;
    }
    }
    const float pdf = pdf_w * pdf_path;
    beta = cmul(cmul(beta, tex), fr * (cost / pdf));
    push_next(out, p, w, beta, mk_flags(le2, in2, depth + 1));
}

__global__ __launch_bounds__(kBlock) void k_shade(PathBuf in, PathBuf out, const uint32_t* lists,
                                                  uint32_t cap, int4 off_lo, int4 off_hi,
                                                  uint32_t total, const int2* hits,
                                                  const float2* uvs, SceneView sv, Params prm,
                                                  float* film, ShadowBuf sh, uint32_t* counters) {
    const uint32_t j = blockIdx.x * kBlock + threadIdx.x;
    ShadeOut so;
    so.alive = false;
    so.shadow = false;
    if (j < total) {
        const int offs[kKeys + 1] = {0, off_lo.x, off_lo.y, off_lo.z, off_lo.w,
                                     off_hi.x, off_hi.y, off_hi.z, off_hi.w};
        int b = 0;
        while (b + 1 < kKeys && (int)j >= offs[b + 1]) ++b;
        const uint32_t i = lists[(size_t)b * cap + (j - (uint32_t)offs[b])];
        const float4 r0 = in.ray[2 * (size_t)i + 0];
        const float4 r1 = in.ray[2 * (size_t)i + 1];
        const float4 bt = in.beta[i];
        const uint32_t pixel = in.pixel[i];
        Rng g;
        g.state = in.rng[i];
        g.inc = ((uint64_t)pixel << 1u) | 1u;
        const int2 h = hits[i];
        const float2 uv = uvs[i];
        shade(sv, prm, v3(r0.x, r0.y, r0.z), v3(r1.x, r1.y, r1.z), in.flags[i],
              v3(bt.x, bt.y, bt.z), g, pixel, h.x, __int_as_float(h.y), uv.x, uv.y, film, so);
        so.rng = g.state;
        so.pixel = pixel;
    }
    const uint32_t ns = wave_append(&counters[0], so.alive);
    if (so.alive) {
        out.ray[2 * (size_t)ns + 0] = so.r0;
        out.ray[2 * (size_t)ns + 1] = so.r1;
        out.beta[ns] = make_float4(so.beta.x, so.beta.y, so.beta.z, 0.0f);
        out.rng[ns] = so.rng;
        out.pixel[ns] = so.pixel;
        out.flags[ns] = so.flags;
    }
    const uint32_t ss = wave_append(&counters[1], so.shadow);
    if (so.shadow) {
        sh.ray[2 * (size_t)ss + 0] = so.s0;
        sh.ray[2 * (size_t)ss + 1] = so.s1;
        sh.contrib[ss] = so.contrib;
    }
}

template <bool SPILL>
__global__ __launch_bounds__(kBlock) void k_shadow(ShadowBuf sh, uint32_t n, SceneView sv,
                                                   float* film) {
    __shared__ int32_t s_ref[kStack * kBlock];
    __shared__ float s_tn[kStack * kBlock];
    const uint32_t tid = threadIdx.x;
    const uint32_t i = blockIdx.x * kBlock + tid;
    if (i >= n) return;
    const float4 r0 = sh.ray[2 * (size_t)i + 0];
    const float4 r1 = sh.ray[2 * (size_t)i + 1];
    Ray r{r0.x, r0.y, r0.z, r1.x, r1.y, r1.z, r1.w, 0.0f};
    float hu, hv;
    int32_t nn = 0, np = 0;
    traverse<kBlock, kStack, true, true, true, false, SPILL>(sv.nodes, sv.prims, sv.rects, sv.root,
                                                             r, false, hu, hv, nn, np, s_ref, s_tn,
                                                             tid);
    // Scene::VisibilityTest: in shadow iff the any-hit query found a hit.
    if (r.tmin > 0.0f) return;
    const float4 c = sh.contrib[i];
    film_add(film, (uint32_t)__float_as_int(c.w), v3(c.x, c.y, c.z));
}

inline uint32_t grid_of(uint64_t n) { return (uint32_t)((n + kBlock - 1) / kBlock); }

}  // namespace

extern "C" int rumi_render(rumi_ctx* ctx, const rumi_scene* scene_c, const rumi_render_desc* desc,
                           float* film, rumi_render_stats* stats, void* stream_v) {
    clear_error();
    rumi_scene* sc = const_cast<rumi_scene*>(scene_c);
    if (!ctx) return fail(RUMI_ERR_NO_DEVICE, "rumi_render: null context");
    if (!sc || !desc || !film) return fail(RUMI_ERR_INVALID, "rumi_render: null argument");
    if (sc->device != ctx->device) return fail(RUMI_ERR_INVALID, "rumi_render: scene is on another device");
    if (desc->integrator != RUMI_INTEGRATOR_PT)
        return fail(RUMI_ERR_UNSUPPORTED, "rumi_render: integrator %d not available", desc->integrator);
    if (desc->width <= 0 || desc->height <= 0 || desc->spp <= 0 || desc->max_depth <= 0)
        return fail(RUMI_ERR_INVALID, "rumi_render: width, height, spp and max_depth must be > 0");
    if (desc->max_depth > 1000) return fail(RUMI_ERR_INVALID, "rumi_render: max_depth > 1000");
    if ((uint64_t)desc->width * desc->height > 0xffffffffull / 4)
        return fail(RUMI_ERR_INVALID, "rumi_render: image too large");
    if (hipSetDevice(ctx->device) != hipSuccess) return fail(RUMI_ERR_HIP, "hipSetDevice failed");
    hipStream_t stream = (hipStream_t)stream_v;

    const uint32_t pixels = (uint32_t)desc->width * (uint32_t)desc->height;
    const uint64_t total = (uint64_t)pixels * (uint64_t)desc->spp;
    uint32_t cap = desc->queue_capacity ? desc->queue_capacity : (1u << 21);
    if ((uint64_t)cap > total) cap = (uint32_t)total;
    if (cap < 64) cap = 64;
    uint32_t regen = desc->regen_threshold ? desc->regen_threshold : cap;

    // Work buffers: 2 path sets, hits, bucket lists, shadow queue, counters.
    const size_t per_path = 2 * (32 + 16 + 8 + 4 + 4) + 8 + 8 + 4 * kKeys + 32 + 16;
    const size_t need = per_path * (size_t)cap + 256 * sizeof(uint32_t) + 1024;
    if (sc->work_bytes < need) {
        if (sc->work) (void)hipFree(sc->work);
        sc->work = nullptr;
        sc->work_bytes = 0;
        if (hipMalloc(&sc->work, need) != hipSuccess)
            return fail(RUMI_ERR_NO_MEMORY, "rumi_render: hipMalloc of %zu work bytes failed", need);
        sc->work_bytes = need;
    }
    char* w = (char*)sc->work;
    auto take = [&](size_t bytes) {
        char* p = w;
        w += (bytes + 255) & ~(size_t)255;
        return (void*)p;
    };
    PathBuf pb[2];
    for (int k = 0; k < 2; ++k) {
        pb[k].ray = (float4*)take(32 * (size_t)cap);
        pb[k].beta = (float4*)take(16 * (size_t)cap);
        pb[k].rng = (uint64_t*)take(8 * (size_t)cap);
        pb[k].pixel = (uint32_t*)take(4 * (size_t)cap);
        pb[k].flags = (uint32_t*)take(4 * (size_t)cap);
    }
    int2* hits = (int2*)take(8 * (size_t)cap);
    float2* uvs = (float2*)take(8 * (size_t)cap);
    uint32_t* lists = (uint32_t*)take(4 * (size_t)kKeys * cap);
    ShadowBuf sh;
    sh.ray = (float4*)take(32 * (size_t)cap);
    sh.contrib = (float4*)take(16 * (size_t)cap);
    uint32_t* counters = (uint32_t*)take(256 * sizeof(uint32_t));  // [0] next [1] shadow [8..16) buckets
    if ((size_t)(w - (char*)sc->work) > sc->work_bytes)
        return fail(RUMI_ERR_INVALID, "rumi_render: internal work-buffer layout overflow");

    SceneView sv;
    sv.nodes = (const float4*)sc->nodes;
    sv.prims = (const float4*)sc->prims;
    sv.rects = (const float4*)sc->rects;
    sv.shapes = (const DevShape*)sc->shapes;
    sv.mats = (const DevMaterial*)sc->materials;
    sv.lights = (const DevLight*)sc->lights;
    sv.textures = (const DevTexture*)sc->textures;
    sv.texels = (const float*)sc->texels;
    for (int k = 0; k < 3; ++k) {
        sv.root.lo[k] = sc->root_lo[k];
        sv.root.hi[k] = sc->root_hi[k];
    }
    sv.root.ref = sc->root_ref;
    sv.n_lights = sc->n_lights;
    sv.med = sc->medium;
    sv.cam = sc->camera;
    Params prm;
    prm.width = desc->width;
    prm.height = desc->height;
    prm.max_depth = desc->max_depth;
    prm.pixels = pixels;
    prm.seed = desc->seed;
    const bool spill = sc->max_depth > kStack;

    hipEvent_t ev0, ev1;
    (void)hipEventCreate(&ev0);
    (void)hipEventCreate(&ev1);
    (void)hipEventRecord(ev0, stream);
    const size_t film_n = 3 * (size_t)pixels;
    hipLaunchKernelGGL(k_zero, dim3((uint32_t)((film_n + 255) / 256)), dim3(256), 0, stream, film, film_n);

    uint64_t issued = 0, ext_rays = 0, shadow_rays = 0, iters = 0;
    uint32_t n_cur = 0;
    int cur = 0;
    uint32_t host_cnt[16];
    int rc = RUMI_OK;
    while (true) {
        if (issued < total && n_cur < cap && n_cur <= regen) {
            const uint32_t cnt = (uint32_t)std::min<uint64_t>(cap - n_cur, total - issued);
            hipLaunchKernelGGL(k_regen, dim3(grid_of(cnt)), dim3(kBlock), 0, stream, pb[cur], n_cur,
                               cnt, issued, prm, sc->camera);
            n_cur += cnt;
            issued += cnt;
        }
        if (n_cur == 0) break;
        ++iters;
        ext_rays += n_cur;
        hipLaunchKernelGGL(k_zero_u32, dim3(1), dim3(64), 0, stream, counters, 16u);
        if (spill)
            hipLaunchKernelGGL(k_extend<true>, dim3(grid_of(n_cur)), dim3(kBlock), 0, stream, pb[cur],
                               n_cur, sv, hits, uvs);
        else
            hipLaunchKernelGGL(k_extend<false>, dim3(grid_of(n_cur)), dim3(kBlock), 0, stream,
                               pb[cur], n_cur, sv, hits, uvs);
        hipLaunchKernelGGL(k_bucket, dim3(grid_of(n_cur)), dim3(kBlock), 0, stream, n_cur, hits, sv,
                           lists, counters + 8, cap);
        if (hipMemcpyAsync(host_cnt + 8, counters + 8, kKeys * sizeof(uint32_t), hipMemcpyDeviceToHost,
                           stream) != hipSuccess ||
            hipStreamSynchronize(stream) != hipSuccess) {
            rc = fail(RUMI_ERR_HIP, "rumi_render: bucket count read failed: %s",
                      hipGetErrorString(hipGetLastError()));
            break;
        }
        int offs[kKeys + 1];
        offs[0] = 0;
        for (int k = 0; k < kKeys; ++k) offs[k + 1] = offs[k] + (int)host_cnt[8 + k];
        const int4 off_lo = make_int4(offs[1], offs[2], offs[3], offs[4]);
        const int4 off_hi = make_int4(offs[5], offs[6], offs[7], offs[8]);
        hipLaunchKernelGGL(k_shade, dim3(grid_of(n_cur)), dim3(kBlock), 0, stream, pb[cur], pb[cur ^ 1],
                           lists, cap, off_lo, off_hi, n_cur, hits, uvs, sv, prm, film, sh, counters);
        if (hipMemcpyAsync(host_cnt, counters, 2 * sizeof(uint32_t), hipMemcpyDeviceToHost, stream) !=
                hipSuccess ||
            hipStreamSynchronize(stream) != hipSuccess) {
            rc = fail(RUMI_ERR_HIP, "rumi_render: counter read failed: %s",
                      hipGetErrorString(hipGetLastError()));
            break;
        }
        const uint32_t n_next = host_cnt[0], n_shadow = host_cnt[1];
        if (n_shadow) {
            shadow_rays += n_shadow;
            if (spill)
                hipLaunchKernelGGL(k_shadow<true>, dim3(grid_of(n_shadow)), dim3(kBlock), 0, stream, sh,
                                   n_shadow, sv, film);
            else
                hipLaunchKernelGGL(k_shadow<false>, dim3(grid_of(n_shadow)), dim3(kBlock), 0, stream,
                                   sh, n_shadow, sv, film);
        }
        cur ^= 1;
        n_cur = n_next;
    }
    if (rc == RUMI_OK) {
        hipLaunchKernelGGL(k_scale, dim3((uint32_t)((film_n + 255) / 256)), dim3(256), 0, stream, film,
                           film_n, 1.0f / (float)desc->spp);
        (void)hipEventRecord(ev1, stream);
        if (hipStreamSynchronize(stream) != hipSuccess)
            rc = fail(RUMI_ERR_HIP, "rumi_render: %s", hipGetErrorString(hipGetLastError()));
        else if (hipGetLastError() != hipSuccess)
            rc = fail(RUMI_ERR_HIP, "rumi_render: a kernel launch failed");
    }
    float ms = 0.0f;
    if (rc == RUMI_OK) (void)hipEventElapsedTime(&ms, ev0, ev1);
    (void)hipEventDestroy(ev0);
    (void)hipEventDestroy(ev1);
    if (stats) {
        stats->samples = total;
        stats->extension_rays = ext_rays;
        stats->shadow_rays = shadow_rays;
        stats->iterations = iters;
        stats->milliseconds = ms;
    }
    return rc;
}
