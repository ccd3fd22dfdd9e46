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
//   shade   -- per tile of 512 paths, a counting sort by material key in LDS
//              (so a wave mostly runs one BSDF), then media, emission,
//              next-event estimation, BSDF sampling; the surviving paths and
//              the shadow rays are appended compactly to the next queue and
//              the shadow queue (one atomic per tile and queue),
//   shadow  -- any-hit of the shadow queue; unoccluded contributions are
//              added to the film with float atomics.
// Path state lives in HBM as structure-of-arrays buffers sized by the
// queue capacity.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cfloat>
#include <chrono>
#include <cstring>
#include <mutex>
#include <type_traits>
#include <vector>

#include "scene.h"
#include "traverse.h"

using namespace rumi;
using namespace rumi::dev;

namespace {

constexpr int kBlock = 128;

// Traversal stack of the renderer's persistent stages (k_extend, k_shadow):
// 16 LDS entries per lane and private-memory entries beyond them, chosen
// per scene from its 4-wide tree's bound (3 entries per level, stack_need):
// variant 0: 16 + 0, 1: 16 + 16, 2: 16 + 32, 3: 16 + 80.  The private array is
// allocated for every lane whether it is reached or not, so it is kept to
// what the bound needs.  (32 LDS entries cost more in occupancy than the
// spills they save: the batch graph went from 734 to 864 ms.)
constexpr int sv_lds(int) { return 16; }
constexpr int sv_priv(int v) { return v == 1 ? 16 : (v == 2 ? 32 : (v == 3 ? kMaxDepth - 16 : 0)); }
inline int stack_variant(int need) { return need <= 16 ? 0 : need <= 32 ? 1 : need <= 48 ? 2 : 3; }

// Variant 4: variant 0 with the whole scene tree (nodes, primitive and rect
// records, at most kTreeF4 float4s) copied into LDS by every block of the
// single-scene extension and shadow stages, so their node and primitive
// fetches are LDS reads (the Cornell box and media scenes: ~1.5 KB).
constexpr uint32_t kTreeF4 = 256;  // 4 KB
constexpr bool sv_tree(int v) { return v == 4; }

template <class F>
void with_stack(int v, F&& f) {
    switch (v) {
    case 0: f(std::integral_constant<int, 0>{}); break;
    case 1: f(std::integral_constant<int, 1>{}); break;
    case 2: f(std::integral_constant<int, 2>{}); break;
    case 4: f(std::integral_constant<int, 4>{}); break;
    default: f(std::integral_constant<int, 3>{}); break;
    }
}
constexpr int kKeys = 8;     // material kinds 0..6, 7 = miss
constexpr int kKeyMiss = 7;
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

// The stream of a vertex's light sample: a second PCG32 stream seeded from
// the path stream's state on arrival at the vertex, so the numbers the path
// itself draws (distances, lobes, directions) do not depend on how many its
// light samples take.  Both renderers of this path number their draws this
// way, whatever order they evaluate light samples in.
constexpr uint64_t kNeeMix = 0xDA3E39CB94B95BDBull;
__device__ inline Rng nee_stream(const Rng& path, uint64_t key) {
    Rng n;
    n.state = 0;
    n.inc = path.inc ^ 2u;
    n.next();
    n.state += key ^ kNeeMix;
    n.next();
    return n;
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

// OrthonormalBasis (src/Kernel/vector.h:70-95).  Its special case for an axis of
// exactly +-Y (tangents X and -Z) is taken within 1e-6 here: an interpolated
// normal of an axis-aligned surface lands on or beside +-Y by rounding, and
// beside it the general case turns rounding noise into the frame's
// orientation.  The orientation about the axis changes no distribution
// sampled in the frame (every sampler here is symmetric about its axis), only
// which direction a given pair of random numbers maps to.
__device__ inline void basis(V3 v0, V3& v1, V3& v2) {
    if (fabsf(v0.x) <= 1e-6f && fabsf(v0.z) <= 1e-6f && fabsf(fabsf(v0.y) - 1.0f) <= 1e-6f) {
        v1 = v3(1, 0, 0);
        v2 = v3(0, 0, -1);
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
    const float4* nodes;  // the scene tree's 4-wide form (DevNode4)
    const float4* prims;
    const float4* rects;
    const DevShape* shapes;
    const DevMaterial* mats;
    const DevLight* lights;
    const DevTexture* textures;
    const float* texels;
    RootInfo root;
    uint32_t n_lights;
    uint32_t tree_f4[3];  // float4s of nodes, primitive records, rect records
    DevMedium med;
    DevCamera cam;
};

struct PathBuf {
    float4* ray;       // 2 per path: o.xyz tmin | d.xyz tmax
    float4* beta;      // throughput rgb
    uint64_t* rng;     // PCG32 state (increment derived from the pixel)
    uint32_t* pixel;   // pixel within its scene (selects the PCG stream)
    uint32_t* sample;  // sample index within the frame (the batch, for batches)
    uint32_t* flags;   // bit 0 le, bit 1 in_object, bits 8.. depth
    float* pend;       // adjoint replay: score of the sampling that produced this ray
    uint32_t* scene;   // batches: the path's scene
};

struct ShadowBuf {
    float4* ray;       // 2 per entry
    float4* contrib;   // rgb, film pixel (int bits) in w
    uint32_t* sample;  // adjoint passes: the sample the contribution belongs to
    float2* score;     // adjoint replay: the connection's own score terms
    int32_t* tex;      // adjoint replay: the image-based light's texel (scene pool index) or -1
    uint32_t* scene;   // batches: the entry's scene
};

struct Params {
    int32_t width, height, max_depth;
    uint32_t pixels;
    uint64_t seed;
    uint32_t n_scenes;  // batches
    int32_t whitted;    // RayTracer: only Specular and Dielectric continue a path
};

// One scene of a batch: its tables, image and offsets into the batch's
// concatenated sample, film and texel-gradient arrays.
struct SceneTab {
    SceneView sv;
    int32_t width, height, spp;
    uint32_t pixels;
    uint32_t sample_base;
    uint32_t pixel_base;
    uint32_t texel_base;
    uint32_t pad;
};

// Device-side control block of the wavefront loop: queue sizes, bucket
// counts and offsets live here so that no stage needs the host to read them.
struct Ctl {
    uint32_t queue[2];      // paths in path buffer 0 / 1
    uint32_t shadow;        // entries in the shadow queue
    uint32_t fetch[2];      // ray counters of the persistent extension / shadow stages
    uint32_t pad0;
    // a frame's totals, added up at each iteration's start (k_clear_ctl):
    // extension rays, shadow entries of the finished iterations, iterations
    unsigned long long total[3];
    uint32_t pad1[16];
};
static_assert(sizeof(Ctl) == 112, "Ctl layout");

__device__ inline uint32_t lane_id() { return __lane_id(); }

// Adds a lane's traversal counts to two 64-bit totals, one atomic per wave.
__device__ inline void add_counts(unsigned long long* counts, uint32_t nodes, uint32_t prims) {
    uint64_t a = nodes, b = prims;
    for (int off = 32; off > 0; off >>= 1) {
        a += __shfl_xor((unsigned long long)a, off);
        b += __shfl_xor((unsigned long long)b, off);
    }
    if (__lane_id() == 0) {
        if (a) atomicAdd(counts + 0, (unsigned long long)a);
        if (b) atomicAdd(counts + 1, (unsigned long long)b);
    }
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
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
         i += (size_t)gridDim.x * blockDim.x)
        p[i] = 0.0f;
}

__global__ void k_zero_u32(uint32_t* p, uint32_t n) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = 0u;
}

__global__ void k_scale(float* p, size_t n, float s) {
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] *= s;
}

// The scene a path belongs to: the only one, or its entry in the batch table.
template <bool MULTI>
__device__ __forceinline__ const SceneView& view_of(const SceneView& sv, const SceneTab* tab,
                                                    uint32_t scene) {
    if (MULTI) return tab[scene].sv;
    return sv;
}

// Grid-stride loop over n items in wave-sized steps: every lane of a wave
// runs the same iterations (so wave-wide ballots are safe); i = base + lane.
#define RUMI_WAVE_LOOP(n)                                                                      \
    for (uint32_t base_ = blockIdx.x * kBlock + (threadIdx.x & ~63u); base_ < (n);            \
         base_ += gridDim.x * kBlock)

// Pinhole::CastViewRay (src/Camera/pinhole.cpp:97-136) for sample k of the
// frame: pixel = k mod P, s = k / P, PCG32 stream (seed + s * golden, pixel).
// Batches number samples scene after scene; a scene's samples and streams are
// the ones it has when rendered alone.  Writes count paths from slot start
// and sets the queue size to start + count.
template <bool MULTI>
__global__ __launch_bounds__(kBlock) void k_regen(PathBuf buf, uint32_t start, uint32_t count,
                                                  uint64_t sample_base, Params prm, DevCamera cam,
                                                  const SceneTab* tab, uint32_t* queue_size) {
    if (blockIdx.x == 0 && threadIdx.x == 0) *queue_size = start + count;
    for (uint32_t j = blockIdx.x * kBlock + threadIdx.x; j < count; j += gridDim.x * kBlock) {
        const uint32_t slot = start + j;
        const uint64_t k = sample_base + j;
        uint32_t scene = 0, pixels = prm.pixels, width = (uint32_t)prm.width,
                 height = (uint32_t)prm.height;
        uint64_t local = k;
        const DevCamera* cp = &cam;
        if (MULTI) {
            uint32_t lo = 0, hi = prm.n_scenes;  // last scene with sample_base <= k
            while (hi - lo > 1) {
                const uint32_t mid = (lo + hi) >> 1;
                if ((uint64_t)tab[mid].sample_base <= k) lo = mid; else hi = mid;
            }
            scene = lo;
            const SceneTab& t = tab[scene];
            local = k - t.sample_base;
            pixels = t.pixels;
            width = (uint32_t)t.width;
            height = (uint32_t)t.height;
            cp = &t.sv.cam;
        }
        const DevCamera& c = *cp;
        const uint32_t pixel = (uint32_t)(local % pixels);
        const uint64_t s = local / pixels;
        Rng g = rng_seed(prm.seed + s * 0x9E3779B97F4A7C15ull, pixel);
        const int x = (int)(pixel % width);
        const int y = (int)(pixel / width);
        const float sx = g.uniform(), sy = g.uniform();
        const float hw = (float)(int)width * 0.5f, hh = (float)(int)height * 0.5f;
        const float px = -(((float)x - hw) + sx) / hw * c.tanfov;
        const float py = -(((float)y - hh) + sy) / hh * c.tanfov * c.y_ratio;
        const V3 l = normalize(v3(px, py, 1.0f));
        V3 d = v3(c.R[0] * l.x + c.R[1] * l.y + c.R[2] * l.z,
                  c.R[3] * l.x + c.R[4] * l.y + c.R[5] * l.z,
                  c.R[6] * l.x + c.R[7] * l.y + c.R[8] * l.z);
        V3 o = v3(c.p[0], c.p[1], c.p[2]);
        if (c.focal > 0.0f) {
            // the thin lens of Pinhole::CastViewRay (pinhole.cpp:111-122): a
            // point on the lens disk aiming at the focal-plane point
            const V3 fp = o + d * (c.focal / l.z);
            const float rr = sqrtf(g.uniform()) * c.lens;
            const float th = g.uniform() * 2.0f * kPiF;
            const V3 up = v3(c.R[1], c.R[4], c.R[7]), right = v3(c.R[0], c.R[3], c.R[6]);
            o = (o + (up * rr) * sinf(th)) + (right * rr) * cosf(th);
            d = normalize(fp - o);
        }
        buf.ray[2 * (size_t)slot + 0] = make_float4(o.x, o.y, o.z, 0.0f);
        buf.ray[2 * (size_t)slot + 1] = make_float4(d.x, d.y, d.z, FLT_MAX);
        buf.beta[slot] = make_float4(1.0f, 1.0f, 1.0f, 0.0f);
        buf.rng[slot] = g.state;
        buf.pixel[slot] = pixel;
        buf.sample[slot] = (uint32_t)k;
        buf.pend[slot] = 0.0f;
        buf.flags[slot] = 1u;  // le = true, outside objects, depth 0
        if (MULTI) buf.scene[slot] = scene;
    }
}

// Closest hit of every queued path (BVHAccel::Travel with the path's
// in_object flag, bvh.cpp:150-213), persistent lanes (run_lanes).
template <int SV, bool MULTI, bool COUNT = false>
struct ExtendItem {
    static constexpr bool kWaveStep = false;
    PathBuf buf;
    const SceneTab* tab;
    const float4 *nodes, *prims, *rects;
    RootInfo root;
    int2* hits;
    float2* uvs;
    int2* s_stk;
    int2* p_stk;
    uint32_t tid, id, nid, nflags, nscene;
    bool in_obj;
    float4 n0, n1;
    __device__ void prefetch(uint32_t k) {
        n0 = buf.ray[2 * (size_t)k + 0];
        n1 = buf.ray[2 * (size_t)k + 1];
        nflags = buf.flags[k];
        if (MULTI) nscene = buf.scene[k];
        nid = k;
    }
    __device__ bool start(TravState& s) {
        s.r = Ray{n0.x, n0.y, n0.z, n1.x, n1.y, n1.z, n1.w, n0.w};
        id = nid;
        in_obj = (nflags & 2u) != 0;
        if (MULTI) {
            const SceneView& v = tab[nscene].sv;
            nodes = v.nodes;
            prims = v.prims;
            rects = v.rects;
            return trav_begin(v.root, s);
        }
        return trav_begin(root, s);
    }
    __device__ bool step(TravState& s) {
        int32_t nn = 0, np = 0;
        const bool go = trav_step<kBlock, sv_lds(SV), false, true, true, COUNT, (sv_priv(SV) > 0), true>(
            nodes, prims, rects, s, in_obj, nn, np, s_stk, tid, p_stk);
        if (COUNT) cn += (uint32_t)nn, cp += (uint32_t)np;
        return go;
    }
    __device__ void finish(const TravState& s) {
        hits[id] = make_int2(s.hit, __float_as_int(s.r.tmin));
        uvs[id] = make_float2(s.hu, s.hv);
    }
    uint32_t cn = 0, cp = 0;  // COUNT: interior nodes entered and primitives tested
};

// Copies a single scene's tree into LDS (stack variant 4) and points
// nodes/prims/rects there; every thread of the block calls it.
__device__ inline void tree_to_lds(const SceneView& sv, float4* s_tree, const float4*& nodes,
                                   const float4*& prims, const float4*& rects) {
    const uint32_t nf = sv.tree_f4[0], pf = sv.tree_f4[1], rf = sv.tree_f4[2];
    for (uint32_t k = threadIdx.x; k < nf; k += blockDim.x) s_tree[k] = sv.nodes[k];
    for (uint32_t k = threadIdx.x; k < pf; k += blockDim.x) s_tree[nf + k] = sv.prims[k];
    for (uint32_t k = threadIdx.x; k < rf; k += blockDim.x) s_tree[nf + pf + k] = sv.rects[k];
    __syncthreads();
    nodes = s_tree;
    prims = s_tree + nf;
    rects = s_tree + nf + pf;
}

// COUNT: also adds every lane's node visits and primitive tests to counts[0..1].
template <int SV, bool MULTI, bool COUNT = false>
__global__ __launch_bounds__(kBlock) void k_extend(PathBuf buf, const uint32_t* n_ptr,
                                                   uint32_t* counter, SceneView sv,
                                                   const SceneTab* tab, int2* hits, float2* uvs, bool lanes,
                                                   unsigned long long* counts = nullptr) {
    __shared__ int2 s_stk[sv_lds(SV) * kBlock];
    int2 p_stk[sv_priv(SV) > 0 ? sv_priv(SV) : 1];
    ExtendItem<SV, MULTI, COUNT> it;
    it.buf = buf;
    it.tab = tab;
    it.nodes = sv.nodes;
    it.prims = sv.prims;
    it.rects = sv.rects;
    if constexpr (sv_tree(SV) && !MULTI) {
        __shared__ float4 s_tree[kTreeF4];
        tree_to_lds(sv, s_tree, it.nodes, it.prims, it.rects);
    }
    it.root = sv.root;
    it.hits = hits;
    it.uvs = uvs;
    it.s_stk = s_stk;
    it.p_stk = p_stk;
    it.tid = threadIdx.x;
    it.id = it.nid = it.nflags = it.nscene = 0;
    it.in_obj = false;
    if (!MULTI && (sv_tree(SV) || !lanes)) {
        // one ray per lane at a time, without the persistent lanes'
        // bookkeeping: a tree in LDS always (its rays are short and cost
        // about the same), other trees when the caller asks (!lanes)
        const uint32_t n = *n_ptr;
        for (uint32_t k = blockIdx.x * kBlock + threadIdx.x; k < n; k += gridDim.x * kBlock) {
            it.prefetch(k);
            TravState ts;
            bool act = it.start(ts);
            while (act) act = it.step(ts);
            it.finish(ts);
        }
    } else {
        run_lanes(*n_ptr, counter, it);
    }
    if (COUNT) {  // counts: [0] nodes, [1] primitives, [4] rays
        add_counts(counts, it.cn, it.cp);
        if (blockIdx.x == 0 && threadIdx.x == 0) atomicAdd(counts + 4, (unsigned long long)*n_ptr);
    }
}

// Clears the counters the next wavefront iteration appends to.
__global__ void k_clear_ctl(Ctl* ctl, int next_queue) {
    const uint32_t t = threadIdx.x;
    if (t == 0) {
        const uint32_t n = ctl->queue[next_queue ^ 1];
        ctl->total[0] += n;
        ctl->total[2] += n ? 1u : 0u;
        ctl->queue[next_queue] = 0u;
    }
    if (t == 1) {
        ctl->total[1] += ctl->shadow;
        ctl->shadow = 0u;
    }
    if (t == 2) ctl->fetch[0] = 0u;
    if (t == 3) ctl->fetch[1] = 0u;
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
__device__ inline float shadow_tr(const DevMedium& m, V3 o, V3 d, float tmax, Rng& g, float& smax_out) {
    smax_out = 0.0f;
    if (m.bound == 0) return 1.0f;
    float tn, tf;
    if (!med_intersect(m, o, d, 0.0f, tn, tf)) return 1.0f;
    const float smax = med_maxdist(tmax, tn, tf);
    const float xi = g.uniform();
    if (xi < 1.0f - expf(m.avg_sigma_t * -smax)) return 0.0f;
    smax_out = smax;
    return 1.0f;
}

// ------------------------------------------------------------ light sampling
struct LightSample {
    V3 p, n, le;
    float pdf;
    float g;    // the light's G toward x (AreaLight::G; 1 for the image-based light)
    int tex;    // image-based light: the texel le came from (scene pool index), else -1
};

__device__ inline V3 pattern(const SceneView& sv, const DevShape& s, V3 p, float bu, float bv, int& index);

// AreaLight::G (arealight.cpp:106-112).
__device__ inline float light_g(V3 x, const LightSample& ls, V3 wi) {
    const float d = length(x - ls.p);
    const float c = fmaxf(-dot(wi, ls.n), 0.0f);
    return c / (d * d);
}

// ImageBasedLight::UniformSampling (imagebasedlight.cpp:35-64): a
// cosine-weighted direction about the normal, the sphere's far side along
// it, its texture there; pdf cos / pi, G = 1.  Compiled only into the
// shading kernels of scenes that hold one (the IBL template flag): its
// sphere test and texture lookup would raise every scene's shading kernel
// from about 90 VGPRs to 150 (5 waves per SIMD to 3).
__device__ inline bool sample_ibl(const SceneView& sv, const DevLight& L, V3 x, V3 nrm, Rng& g,
                                  LightSample& ls, V3& wi, float& dist) {
    V3 u, w;
    basis(nrm, u, w);
    const float phi = g.uniform() * 2.0f * kPiF;
    const float theta = acosf(sqrtf(1.0f - g.uniform()));
    const float st = sinf(theta);
    wi = frame(u, nrm, w, st * sinf(phi), cosf(theta), st * cosf(phi));
    Ray r{x.x, x.y, x.z, wi.x, wi.y, wi.z, FLT_MAX, 0.0f};
    hit_sphere(make_float4(L.p[0], L.p[1], L.p[2], L.l[0]), r, false, true);
    ls.p = x + wi * r.tmin;
    ls.n = -wi;
    int index = -1;
    ls.le = pattern(sv, sv.shapes[L.prim], ls.p, 0.0f, 0.0f, index);
    ls.tex = index;
    float c = dot(wi, nrm);
    if (c <= 0.0f) c = 1e-4f;
    float pdf = kInvPiF * c;
    if (pdf <= 0.0f) pdf = 1e-4f;
    ls.pdf = 1.0f / (float)sv.n_lights * pdf;
    ls.g = 1.0f;
    dist = length(ls.p - x);
    return !(dot(wi, ls.n) >= 0.0f);
}

// Scene::VisibilityTest (scene.cpp:99-123) up to the shadow ray: uniform
// light choice, a uniform point on the emitter (the reference draws from a
// fixed stratified set), the one-sided test.  Returns false if unlit.  n:
// the shading normal, which the image-based light samples around.  IBL:
// the scene may hold an image-based light.
template <bool IBL>
__device__ inline bool sample_light(const SceneView& sv, V3 x, V3 nrm, Rng& g, LightSample& ls, V3& wi,
                                    float& dist) {
    ls.tex = -1;
    if (sv.n_lights == 0) return false;
    uint32_t li = (uint32_t)(g.uniform() * (float)sv.n_lights);
    if (li >= sv.n_lights) li = sv.n_lights - 1;
    const DevLight& L = sv.lights[li];
    if (L.kind == kLightPoint) {
        // PointLight::P / Wi / N / G / Pdf (pointlight.cpp:44-72): the point
        // itself, N from the light toward x, G = 1 / d^2, Pdf 1
        ls.p = ld3(L.p);
        wi = normalize(ls.p - x);
        ls.n = normalize(x - ls.p);
        ls.le = ld3(L.le);
        ls.pdf = 1.0f / (float)sv.n_lights;
        dist = length(ls.p - x);
        ls.g = 1.0f / (dist * dist);
        return !(dot(wi, ls.n) >= 0.0f);
    }
    if (IBL && L.kind == kLightIbl) return sample_ibl(sv, L, x, nrm, g, ls, wi, dist);
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
    ls.g = light_g(x, ls, wi);
    return !(dot(wi, ls.n) >= 0.0f);
}

// ------------------------------------------------------------ textures
__device__ inline V3 texel(const SceneView& sv, int tex, float u, float v, int& index) {
    const DevTexture t = sv.textures[tex];
    if (t.kind == RUMI_TEXTURE_CHECKER) {
        // Checker::Pixel (checker.cpp:24-32)
        index = -1;
        const int a = (int)(u * t.cw) % 2, b = (int)(v * t.ch) % 2;
        const bool second = (a == 0 && b == 0) || (a != 0 && b != 0);
        return second ? v3(t.c2[0], t.c2[1], t.c2[2]) : v3(t.c1[0], t.c1[1], t.c1[2]);
    }
    if (t.kind == RUMI_TEXTURE_HDR) {
        // HDRTexture::Pixel (hdrtexture.cpp:29-41): rows top-down as loaded, u mirrored
        const int row = (int)((1.0f - v) * (float)(t.height - 1));
        const int col = (int)((1.0f - u) * (float)(t.width - 1));
        const int y = min(max(row, 0), t.height - 1), x = min(max(col, 0), t.width - 1);
        index = (int)(t.offset / 3u) + y * t.width + x;
        const float* p = sv.texels + t.offset + 3 * ((size_t)y * t.width + x);
        return v3(p[0], p[1], p[2]);
    }
    // ImageTexture::Pixel (imagetexture.cpp:32-61): nearest texel
    const int row = (int)((1.0f - v) * (float)(t.height - 1));
    const int col = (int)(u * (float)(t.width - 1));
    const int y = min(max(t.height - 1 - row, 0), t.height - 1);
    const int x = min(max(col, 0), t.width - 1);
    index = (int)(t.offset / 3u) + y * t.width + x;
    const float* p = sv.texels + t.offset + 3 * ((size_t)y * t.width + x);
    return v3(p[0], p[1], p[2]);
}

// Shape::TextureMapping overrides: Rect (rect.cpp:183-197), Sphere
// (sphere.cpp:172-195: angles against the rows uu, vv, ww of the shape's
// transform, kept in n0..n2); triangles interpolate vertex uv.
__device__ inline V3 pattern(const SceneView& sv, const DevShape& s, V3 p, float bu, float bv,
                             int& index) {
    const int tex = s.info[2];
    index = -1;
    if (tex < 0) return v3(1, 1, 1);
    float u = 0.0f, v = 0.0f;
    if (s.info[0] == RUMI_SHAPE_TRIANGLE) {
        const float w0 = 1.0f - bu - bv;
        u = w0 * s.n0[3] + bu * s.n2[3] + bv * s.uv2[1];
        v = w0 * s.n1[3] + bu * s.uv2[0] + bv * s.uv2[2];
    } else if (s.info[0] == RUMI_SHAPE_SPHERE) {
        const V3 pc = normalize(p - v3(s.g0[0], s.g0[1], s.g0[2]));
        const V3 uu = ld3(s.n0), vv = ld3(s.n1), ww = ld3(s.n2);
        const float theta = acosf(dot(pc, vv));
        const V3 pxz = pc * sinf(theta);
        const float phi = atan2f(dot(pxz, uu), dot(pxz, ww));
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
    return texel(sv, tex, u, v, index);
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

// GGX microfacet reflection (Walter et al. 2007): D with roughness alpha,
// height-correlated Smith G2, Schlick Fresnel with F0 = colour.  Extension,
// no reference counterpart.
__device__ inline float ggx_d(float alpha, float cos_h) {
    const float a2 = alpha * alpha;
    const float d = cos_h * cos_h * (a2 - 1.0f) + 1.0f;
    return a2 / (kPiF * d * d);
}

__device__ inline float ggx_lambda(float alpha, float c) {
    const float c2 = c * c;
    const float t2 = (1.0f - c2) / c2;
    return 0.5f * (-1.0f + sqrtf(1.0f + alpha * alpha * t2));
}

__device__ inline V3 ggx_f(const DevMaterial& m, V3 n, V3 wi, V3 wo) {
    const float ci = dot(n, wi), co = dot(n, wo);
    if (ci <= 0.0f || co <= 0.0f) return v3(0, 0, 0);
    const V3 h = normalize(wi + wo);
    const float D = ggx_d(m.alpha, dot(n, h));
    const float G = 1.0f / (1.0f + ggx_lambda(m.alpha, ci) + ggx_lambda(m.alpha, co));
    const float x = 1.0f - fmaxf(dot(wi, h), 0.0f);
    const float x5 = x * x * x * x * x;
    const V3 F = v3(m.color[0] + (1.0f - m.color[0]) * x5, m.color[1] + (1.0f - m.color[1]) * x5,
                    m.color[2] + (1.0f - m.color[2]) * x5);
    return F * (D * G / (4.0f * ci * co));
}

// Solid-angle pdf of sampling wi from wo with D(h) cos(theta_h).
__device__ inline float ggx_pdf(const DevMaterial& m, V3 n, V3 wi, V3 wo) {
    const float ci = dot(n, wi), co = dot(n, wo);
    if (ci <= 0.0f || co <= 0.0f) return 0.0f;
    const V3 h = normalize(wi + wo);
    const float ch = dot(n, h);
    return ggx_d(m.alpha, ch) * ch / (4.0f * dot(wo, h));
}

__device__ inline V3 ggx_sample(const DevMaterial& m, V3 n, V3 wo, Rng& g) {
    const float u1 = g.uniform(), u2 = g.uniform();
    const float t2 = m.alpha * m.alpha * u1 / (1.0f - u1);
    const float ch = 1.0f / sqrtf(1.0f + t2);
    const float sh = sqrtf(fmaxf(0.0f, 1.0f - ch * ch));
    const float phi = 2.0f * kPiF * u2;
    V3 u, w;
    basis(n, u, w);
    const V3 h = frame(u, n, w, sh * sinf(phi), ch, sh * cosf(phi));
    return h * (2.0f * dot(wo, h)) - wo;
}

// ------------------------------------------------------------ shading
// Modes of the shading and shadow stages: plain rendering; the primal pass
// of the adjoint, which also sums every sample's radiance; the adjoint
// replay, which re-runs every path with the same random numbers and
// accumulates parameter gradients instead of the film.
enum { kRender = 0, kPrimal = 1, kReplay = 2 };
constexpr int kStripes = 64;  // gradient partial sums spread over 64 addresses

struct AdjCtx {
    float* L_total;     // 3 per sample: the primal pass's radiance
    float* L_sofar;     // 3 per sample: radiance of the vertices replayed so far
    const float* dimg;  // dLoss/dImage, rgb per pixel
    float inv_spp;      // the image is the mean over spp samples
    float* gstripes;    // kStripes x 8 partial sums of the 7 medium gradients
    float* tex_grad;    // texel-pool gradient (rgb) or null
};

// d(log) of the path weight with respect to the medium parameters
// (sigma_a rgb, sigma_s rgb, g), times dLoss/dL x L, summed in registers.
struct Grad7 {
    float v[7];
};

__device__ inline void grad_grey(Grad7& G, float dlog_dsigma_bar, float S) {
    // sigma_bar = (sigma_t.r + sigma_t.g + sigma_t.b) / 3, sigma_t = sigma_a + sigma_s
    const float t = dlog_dsigma_bar * (1.0f / 3.0f) * S;
#pragma unroll
    for (int k = 0; k < 6; ++k) G.v[k] += t;
}

// d log HG(c; g) / d g, c the cosine between propagation and scattered
// direction (3c at g = 0, where the isotropic phase is HG's limit).
__device__ inline float dlog_hg(float g, float c) {
    return -2.0f * g / (1.0f - g * g) - 3.0f * (g - c) / (1.0f + g * g - 2.0f * g * c);
}

__device__ inline void flush_grad(Grad7& G, float* stripes) {
#pragma unroll
    for (int k = 0; k < 7; ++k) {
        float v = G.v[k];
        for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off);
        G.v[k] = v;
    }
    if (lane_id() == 0) {
        const uint32_t wave = blockIdx.x * (kBlock / 64) + threadIdx.x / 64;
        float* dst = stripes + 8 * (wave % kStripes);
#pragma unroll
        for (int k = 0; k < 7; ++k)
            if (G.v[k] != 0.0f) atomicAdd(dst + k, G.v[k]);
    }
}

__device__ inline void add3(float* p, V3 c) {
    if (is_nan(c)) return;
    atomicAdd(p + 0, c.x);
    atomicAdd(p + 1, c.y);
    atomicAdd(p + 2, c.z);
}

struct ShadeOut {
    uint32_t pixel, sample;
    bool alive;
    float4 r0, r1, beta;
    uint64_t rng;
    uint32_t flags;
    float pend;     // score of the sampling decision made at this vertex for the next one
    bool shadow;
    float4 s0, s1, contrib;
    float2 score;   // shadow entry: medium length of the segment (Bernoulli), d log fp / d g
    int32_t ltex;   // shadow entry: the image-based light's texel, or -1
};

__device__ inline void push_shadow(ShadeOut& o, V3 x, V3 wi, float dist, V3 c, uint32_t pixel,
                                   float smax_s, float dlog_fp, int ltex) {
    o.shadow = true;
    o.s0 = make_float4(x.x, x.y, x.z, 0.0f);
    o.s1 = make_float4(wi.x, wi.y, wi.z, dist);
    o.contrib = make_float4(c.x, c.y, c.z, __int_as_float((int)pixel));
    o.score = make_float2(smax_s, dlog_fp);
    o.ltex = ltex;
}

__device__ inline void push_next(ShadeOut& o, V3 p, V3 w, V3 beta, uint32_t flags, float pend) {
    o.alive = true;
    o.r0 = make_float4(p.x, p.y, p.z, 0.0f);
    o.r1 = make_float4(w.x, w.y, w.z, FLT_MAX);
    o.beta = make_float4(beta.x, beta.y, beta.z, 0.0f);
    o.flags = flags;
    o.pend = pend;
}

__device__ inline uint32_t mk_flags(bool le, bool in_obj, uint32_t depth) {
    return (le ? 1u : 0u) | (in_obj ? 2u : 0u) | (depth << 8);
}

// Matte::DirectLighting (matte.cpp:54-82) without the occlusion test, which
// the shadow stage performs.  smax_s: the shadow segment's length inside
// the medium when its Bernoulli transmittance survived (0 if none).
template <bool IBL>
__device__ inline bool matte_direct(const SceneView& sv, const DevMaterial& m, V3 p, V3 n, Rng& g,
                                    V3& lo, V3& wi, float& dist, float& smax_s, int& ltex) {
    LightSample ls;
    smax_s = 0.0f;
    const bool lit = sample_light<IBL>(sv, p, n, g, ls, wi, dist);
    ltex = ls.tex;
    if (!lit) return false;
    const V3 f = ld3(m.color) * (kInvPiF * m.kd);
    const float c = fmaxf(dot(wi, n), 0.0f);
    lo = cmul(f, ls.le) * (c * ls.g / ls.pdf);
    if (sv.med.bound) lo = lo * shadow_tr(sv.med, p, wi, dist, g, smax_s);
    return true;
}

template <int MODE, bool IBL>
// pixel: the film pixel (the batch's, for batches); tex_base: the scene's
// first texel in the batch's texel-gradient pool.
__device__ void shade(const SceneView& sv, const Params& prm, V3 o, V3 d, uint32_t flags, V3 beta,
                      Rng& g, uint32_t pixel, uint32_t sample, uint32_t tex_base, float pend,
                      int32_t prim, float t, float bu, float bv, float* film, const AdjCtx& adj,
                      Grad7& G, ShadeOut& out) {
    const bool le = (flags & 1u) != 0;
    const bool in_obj = (flags & 2u) != 0;
    const uint32_t depth = flags >> 8;
    const bool last = (int)depth + 1 > prm.max_depth - 1;
    const DevMedium& md = sv.med;
    const uint64_t nee_key = g.state;  // this vertex's light-sample stream

    // Replay: the radiance this sample still collects from here on.
    V3 w3 = v3(0, 0, 0);
    float S = 0.0f;
    if (MODE == kReplay) {
        const float* Lt = adj.L_total + 3 * (size_t)sample;
        const float* Ls = adj.L_sofar + 3 * (size_t)sample;
        const V3 dl = ld3(adj.dimg + 3 * (size_t)pixel) * adj.inv_spp;
        w3 = cmul(dl, v3(Lt[0] - Ls[0], Lt[1] - Ls[1], Lt[2] - Ls[2]));
        S = w3.x + w3.y + w3.z;
        if (pend != 0.0f) G.v[6] += pend * S;  // the previous vertex's phase direction
    }

    // Homogeneous medium along the segment (pathtracer.cpp:156-178); the
    // RayTracer samples none (its only medium term is the shadow rays').
    if (md.bound && !in_obj && !prm.whitted) {
        float tn, tf;
        if (med_intersect(md, o, d, prim >= 0 ? t : 0.0f, tn, tf)) {
            const float smax = med_maxdist(prim >= 0 ? t : 0.0f, tn, tf);
            float s = 0.0f;
            V3 ratio;
            const int ev = distance_sample(md.sigma_t, md.avg_sigma_t, md.absorb, smax, g, s, ratio);
            if (ev == kAbsorb) return;
            beta = cmul(beta, ratio);
            if (MODE == kReplay) {
                if (ev == kPass) {
                    // P(pass) = exp(-sigma_bar (smax - 1e-4))
                    grad_grey(G, -fmaxf(smax - 1e-4f, 0.0f), S);
                } else {
                    // distance density sigma_bar exp(-sigma_bar s), survival of the
                    // absorption roulette (1 - A), albedo weight sigma_s / sigma_t
                    grad_grey(G, 1.0f / md.avg_sigma_t - s, S);
                    const float inv1a = 1.0f / (1.0f - md.absorb);
                    const float wc[3] = {w3.x, w3.y, w3.z};
#pragma unroll
                    for (int c = 0; c < 3; ++c) {
                        const float st = md.sigma_t[c];
                        const float dA_da = md.sigma_s[c] / (st * st) * (1.0f / 3.0f);
                        const float dA_ds = -md.sigma_a[c] / (st * st) * (1.0f / 3.0f);
                        G.v[c] += -dA_da * inv1a * S - wc[c] / st;
                        G.v[3 + c] += -dA_ds * inv1a * S;
                        if (md.sigma_s[c] > 0.0f) G.v[3 + c] += wc[c] * (1.0f / md.sigma_s[c] - 1.0f / st);
                    }
                }
            }
            if (ev == kScatter) {
                const V3 p = o + d * (tn < 0.0f ? s : tn + s);
                // HomogeneousVolume::DirectLighting (homogeneousvolume.cpp:231-250)
                LightSample ls;
                V3 wi;
                float dist;
                Rng gn = nee_stream(g, nee_key);
                if (sample_light<false>(sv, p, v3(0.0f, 1.0f, 0.0f), gn, ls, wi, dist)) {  // no image-based light with media
                    float smax_s;
                    const float tr = shadow_tr(md, p, wi, dist, gn, smax_s);
                    const float cl = dot(d, wi);
                    const float fp = md.g == 0.0f ? 1.0f / (4.0f * kPiF) : hg_eval(md.g, cl);
                    const V3 c = cmul(cmul(beta, ls.le), ld3(md.sigma_s)) *
                                 (tr * fp * ls.g / ls.pdf);
                    if (tr > 0.0f) push_shadow(out, p, wi, dist, c, pixel, smax_s, dlog_hg(md.g, cl), -1);
                }
                if (last) return;
                float pdf;
                const V3 w = phase_sample(md.g, d, g, pdf);
                // Fp / pdf = sigma_s (phase value equals its pdf)
                beta = cmul(beta, ld3(md.sigma_s));
                push_next(out, p, w, beta, mk_flags(false, false, depth + 1),
                          MODE == kReplay ? dlog_hg(md.g, dot(d, w)) : 0.0f);
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

    // The shape's texture multiplies the radiance leaving every surface
    // vertex, emitters and subsurface interior events included
    // (PathTracer::Trace, pathtracer.cpp:343-344).  In the replay, the texel
    // gains dLoss/dL x L_rem / texel.
    auto texture_at = [&](V3 at) -> V3 {
        int index = -1;
        const V3 tx = pattern(sv, s, at, bu, bv, index);
        if (MODE == kReplay && adj.tex_grad && index >= 0) {
            float* tg = adj.tex_grad + 3 * ((size_t)tex_base + (size_t)index);
            if (tx.x > 0.0f && w3.x != 0.0f) atomicAdd(tg + 0, w3.x / tx.x);
            if (tx.y > 0.0f && w3.y != 0.0f) atomicAdd(tg + 1, w3.y / tx.y);
            if (tx.z > 0.0f && w3.z != 0.0f) atomicAdd(tg + 2, w3.z / tx.z);
        }
        return tx;
    };

    if (m.kind == RUMI_MAT_EMISSIVE) {
        // Emissive::DirectLighting (emissive.cpp:60-74)
        if (!in_obj && le && -dot(d, n) > 0.0f) {
            V3 c = cmul(beta, ld3(m.color) * m.kd);
            if (s.info[2] >= 0) c = cmul(c, texture_at(p));
            if (MODE != kReplay) film_add(film, pixel, c);
            if (MODE == kPrimal) add3(adj.L_total + 3 * (size_t)sample, c);
            if (MODE == kReplay) add3(adj.L_sofar + 3 * (size_t)sample, c);
        }
        return;
    }

    // RayTracer::Trace (raytracer.cpp:37-78): Specular and Dielectric trace
    // on (GlobalIllumination); every other material returns its
    // DirectLighting and ends the path -- Matte's light sample, Glossy's
    // only when the lobe variable, left at 1 by the RayTracer
    // (intersection.cpp:18), is within kd / (kd + ks); Subsurface and GGX
    // none.
    if (prm.whitted && m.kind != RUMI_MAT_SPECULAR && m.kind != RUMI_MAT_DIELECTRIC) {
        if (m.kind == RUMI_MAT_MATTE || (m.kind == RUMI_MAT_GLOSSY && m.kd != 0.0f && 1.0f <= m.kd_pdf)) {
            V3 lo, wi;
            float dist, smax_s;
            int ltex;
            Rng gn = nee_stream(g, nee_key);
            if (matte_direct<IBL>(sv, m, p, n, gn, lo, wi, dist, smax_s, ltex)) {
                if (m.kind == RUMI_MAT_GLOSSY) lo = lo * (1.0f / m.kd_pdf);
                const V3 tex = s.info[2] >= 0 ? texture_at(p) : v3(1, 1, 1);
                const V3 c = cmul(cmul(beta, tex), lo);
                if (c.x != 0.0f || c.y != 0.0f || c.z != 0.0f)
                    push_shadow(out, p, wi, dist, c, pixel, smax_s, 0.0f, ltex);
            }
        }
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
            const V3 ps = o + d * sd;
            beta = cmul(cmul(beta, ratio), ld3(m.sigma_s));
            if (s.info[2] >= 0) beta = cmul(beta, texture_at(ps));
            if (last) return;
            float pdf;
            const V3 w = phase_sample(0.0f, d, g, pdf);
            push_next(out, ps, w, beta, mk_flags(true, true, depth + 1), 0.0f);
            return;
        }
        vol = ratio;
    }

    const V3 tex = s.info[2] >= 0 ? texture_at(p) : v3(1, 1, 1);
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
    case RUMI_MAT_GGX:
        // emission is counted on BSDF-sampled hits (le), as after the
        // reference's Phong lobe; no next-event estimation
        w = ggx_sample(m, n, wo, g);
        cost = dot(n, w);
        pdf_w = ggx_pdf(m, n, w, wo);
        le2 = true;
        if (!(cost > 0.0f && pdf_w > 0.0f)) return;
        break;
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
        float dist, smax_s;
        int ltex;
        Rng gn = nee_stream(g, nee_key);
        if (matte_direct<IBL>(sv, m, p, n, gn, lo, wi, dist, smax_s, ltex)) {
            if (m.kind == RUMI_MAT_GLOSSY) lo = lo * (1.0f / m.kd_pdf);
            const V3 c = cmul(cmul(beta, tex), lo);
            if (c.x != 0.0f || c.y != 0.0f || c.z != 0.0f)
                push_shadow(out, p, wi, dist, c, pixel, smax_s, 0.0f, ltex);
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
    case RUMI_MAT_GGX:
        fr = ggx_f(m, n, w, wo);
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
    push_next(out, p, w, beta, mk_flags(le2, in2, depth + 1), 0.0f);
}

// Per-scene completion of a batch: samples still running per scene, and the
// device clock when a scene's last path ended.
struct Track {
    uint32_t* left;
    uint64_t* done;
};

// Counts this wave's path terminations against their scenes: one atomic per
// distinct scene in the wave (a wave holds consecutive samples, so usually
// one); the lane that retires a scene's last sample stamps the clock.
__device__ inline void track_end(const Track& tr, bool ended, uint32_t scene) {
    uint64_t todo = __ballot(ended);
    while (todo) {
        const int leader = __ffsll((unsigned long long)todo) - 1;
        const uint32_t sc = __shfl(scene, leader);
        const uint64_t same = __ballot(ended && scene == sc) & todo;
        if ((int)lane_id() == leader) {
            const uint32_t cnt = (uint32_t)__popcll(same);
            if (atomicSub(&tr.left[sc], cnt) == cnt) tr.done[sc] = wall_clock64();
        }
        todo &= ~same;
    }
}

// Paths per shading tile (one block): one atomic per tile and output queue,
// and the queue counters' same-address atomics serialise, so tiles of 512
// paths beat 256 (Cornell path tracer +11%) and 1024 (fewer resident waves).
constexpr int kShadeBlock = 512;

__device__ __forceinline__ uint32_t lane_rank(uint64_t mask) {
    // set bits of mask in the lanes below this one
    return __builtin_amdgcn_mbcnt_hi((uint32_t)(mask >> 32),
                                     __builtin_amdgcn_mbcnt_lo((uint32_t)mask, 0u));
}

// Block-aggregated append: wave counts from ballots, one global atomic per
// block, consecutive slots in lane order.  Every thread of the block calls it.
template <int BLOCK>
__device__ inline uint32_t block_append(uint32_t* counter, bool pred, uint32_t* s_tmp) {
    constexpr int W = BLOCK / 64;
    const uint32_t wave = threadIdx.x >> 6;
    const uint64_t mask = __ballot(pred);
    const uint32_t rank = lane_rank(mask);
    if (lane_id() == 0) s_tmp[wave] = (uint32_t)__popcll(mask);
    __syncthreads();
    if (threadIdx.x == 0) {
        uint32_t total = 0;
        for (int w = 0; w < W; ++w) {
            const uint32_t c = s_tmp[w];
            s_tmp[w] = total;
            total += c;
        }
        s_tmp[W] = total ? atomicAdd(counter, total) : 0u;
    }
    __syncthreads();
    const uint32_t slot = s_tmp[W] + s_tmp[wave] + rank;
    __syncthreads();
    return slot;
}

// Shading of every queued path.  Each block takes a tile of 512 paths and
// orders them by material key (kind 0..6, 7 = miss) with a counting sort in
// LDS, so the waves of a tile mostly run one BSDF; continuing paths go to the
// other path buffer and shadow rays to the shadow queue.
template <int MODE, bool MULTI, bool IBL>
__global__ __launch_bounds__(kShadeBlock) void k_shade(PathBuf in, PathBuf out, const uint32_t* n_ptr,
                                                       const int2* hits, const float2* uvs,
                                                       SceneView sv, const SceneTab* tab,
                                                       Params prm, float* film, ShadowBuf sh,
                                                       uint32_t* next_count,
                                                       uint32_t* shadow_count, AdjCtx adj,
                                                       Track tr) {
    constexpr int W = kShadeBlock / 64;
    constexpr int kSortKeys = kKeys + 1;  // + 1: slots past the end of the queue
    __shared__ uint32_t s_perm[kShadeBlock];
    __shared__ uint32_t s_hist[W][kSortKeys];
    __shared__ uint32_t s_tmp[W + 1];
    const uint32_t n = *n_ptr;
    const uint32_t tid = threadIdx.x, wave = tid >> 6;
    const uint32_t base = blockIdx.x * kShadeBlock;  // one tile per block
    if (base >= n) return;
    {
        // counting sort of the tile by material key
        const uint32_t i0 = base + tid;
        int key = kKeys;
        if (i0 < n) {
            key = kKeyMiss;
            const int32_t prim = hits[i0].x;
            if (prim >= 0) {
                const SceneView& v = view_of<MULTI>(sv, tab, MULTI ? in.scene[i0] : 0u);
                key = v.mats[v.shapes[prim].info[1]].kind;
            }
        }
        uint32_t rank = 0;
        for (int k = 0; k < kSortKeys; ++k) {
            const uint64_t m = __ballot(key == k);
            if (key == k) rank = lane_rank(m);
            if (lane_id() == 0) s_hist[wave][k] = (uint32_t)__popcll(m);
        }
        __syncthreads();
        if (tid == 0) {
            uint32_t run = 0;
            for (int k = 0; k < kSortKeys; ++k)
                for (int w = 0; w < W; ++w) {
                    const uint32_t c = s_hist[w][k];
                    s_hist[w][k] = run;
                    run += c;
                }
        }
        __syncthreads();
        s_perm[s_hist[wave][key] + rank] = i0;
        __syncthreads();
        const uint32_t i = s_perm[tid];

        ShadeOut so;
        so.alive = false;
        so.shadow = false;
        Grad7 G;
#pragma unroll
        for (int k = 0; k < 7; ++k) G.v[k] = 0.0f;
        uint32_t scene = 0;
        const bool valid = i < n;
        if (valid) {
            if (MULTI) scene = in.scene[i];
            const float4 r0 = in.ray[2 * (size_t)i + 0];
            const float4 r1 = in.ray[2 * (size_t)i + 1];
            const float4 bt = in.beta[i];
            const uint32_t pixel = in.pixel[i];
            const uint32_t sample = in.sample[i];
            Rng g;
            g.state = in.rng[i];
            g.inc = ((uint64_t)pixel << 1u) | 1u;
            const int2 h = hits[i];
            const float2 uv = uvs[i];
            const uint32_t fpix = MULTI ? tab[scene].pixel_base + pixel : pixel;
            const uint32_t tbase = MULTI ? tab[scene].texel_base : 0u;
            shade<MODE, IBL>(view_of<MULTI>(sv, tab, scene), prm, v3(r0.x, r0.y, r0.z),
                        v3(r1.x, r1.y, r1.z), in.flags[i], v3(bt.x, bt.y, bt.z), g, fpix, sample,
                        tbase, MODE == kReplay ? in.pend[i] : 0.0f, h.x, __int_as_float(h.y),
                        uv.x, uv.y, film, adj, G, so);
            so.rng = g.state;
            so.pixel = pixel;
            so.sample = sample;
        }
        if (MODE == kReplay) flush_grad(G, adj.gstripes);
        if (MULTI && tr.left) track_end(tr, valid && !so.alive, scene);
        const uint32_t ns = block_append<kShadeBlock>(next_count, so.alive, s_tmp);
        if (so.alive) {
            out.ray[2 * (size_t)ns + 0] = so.r0;
            out.ray[2 * (size_t)ns + 1] = so.r1;
            out.beta[ns] = make_float4(so.beta.x, so.beta.y, so.beta.z, 0.0f);
            out.rng[ns] = so.rng;
            out.pixel[ns] = so.pixel;
            out.sample[ns] = so.sample;
            out.flags[ns] = so.flags;
            if (MODE == kReplay) out.pend[ns] = so.pend;
            if (MULTI) out.scene[ns] = scene;
        }
        const uint32_t ss = block_append<kShadeBlock>(shadow_count, so.shadow, s_tmp);
        if (so.shadow) {
            sh.ray[2 * (size_t)ss + 0] = so.s0;
            sh.ray[2 * (size_t)ss + 1] = so.s1;
            sh.contrib[ss] = so.contrib;
            if (MODE != kRender) sh.sample[ss] = so.sample;
            if (MODE == kReplay) {
                sh.score[ss] = so.score;
                sh.tex[ss] = so.ltex;
            }
            if (MULTI) sh.scene[ss] = scene;
        }
    }
}

// Any-hit test of every shadow ray (Scene::VisibilityTest, scene.cpp:91-123),
// persistent lanes; unoccluded contributions go to the film (the adjoint
// passes: to the per-sample sums and the connection's own gradient terms).
template <int SV, int MODE, bool MULTI, bool COUNT = false>
struct ShadowItem {
    static constexpr bool kWaveStep = false;
    ShadowBuf sh;
    const SceneTab* tab;
    const float4 *nodes, *prims, *rects;
    const float* texels;  // the scene's texel pool (single scene)
    RootInfo root;
    float* film;
    AdjCtx adj;
    Grad7 G;
    int2* s_stk;
    int2* p_stk;
    uint32_t tid, id, nid, nscene, scene;
    float4 n0, n1;
    __device__ void prefetch(uint32_t k) {
        n0 = sh.ray[2 * (size_t)k + 0];
        n1 = sh.ray[2 * (size_t)k + 1];
        if (MULTI) nscene = sh.scene[k];
        nid = k;
    }
    __device__ bool start(TravState& s) {
        s.r = Ray{n0.x, n0.y, n0.z, n1.x, n1.y, n1.z, n1.w, 0.0f};
        id = nid;
        if (MULTI) {
            scene = nscene;
            const SceneView& v = tab[nscene].sv;
            nodes = v.nodes;
            prims = v.prims;
            rects = v.rects;
            return trav_begin(v.root, s);
        }
        return trav_begin(root, s);
    }
    __device__ bool step(TravState& s) {
        int32_t nn = 0, np = 0;
        const bool go = trav_step<kBlock, sv_lds(SV), true, true, true, COUNT, (sv_priv(SV) > 0), true>(
            nodes, prims, rects, s, false, nn, np, s_stk, tid, p_stk);
        if (COUNT) cn += (uint32_t)nn, cp += (uint32_t)np;
        return go;
    }
    uint32_t cn = 0, cp = 0;  // COUNT: interior nodes entered and primitives tested
    __device__ void finish(const TravState& s) {
        // Scene::VisibilityTest: in shadow iff the any-hit query found a hit.
        if (s.r.tmin > 0.0f) return;
        const float4 c4 = sh.contrib[id];
        const V3 c = v3(c4.x, c4.y, c4.z);
        const uint32_t pixel = (uint32_t)__float_as_int(c4.w);
        if (MODE != kReplay) film_add(film, pixel, c);
        if (MODE == kPrimal) add3(adj.L_total + 3 * (size_t)sh.sample[id], c);
        if (MODE == kReplay && !is_nan(c)) {
            add3(adj.L_sofar + 3 * (size_t)sh.sample[id], c);
            // factors of this connection alone: the shadow segment's
            // Bernoulli transmittance and the phase value toward the light
            const V3 w3 = cmul(ld3(adj.dimg + 3 * (size_t)pixel) * adj.inv_spp, c);
            const float Ss = w3.x + w3.y + w3.z;
            const float2 sc = sh.score[id];
            if (sc.x > 0.0f) grad_grey(G, -sc.x, Ss);
            G.v[6] += sc.y * Ss;
            // an image-based light's texel scales this contribution linearly
            const int lt = sh.tex[id];
            if (lt >= 0 && adj.tex_grad) {
                const float* pool = MULTI ? tab[scene].sv.texels : texels;
                const uint32_t base = MULTI ? tab[scene].texel_base : 0u;
                const V3 tx = ld3(pool + 3 * (size_t)lt);
                float* tg = adj.tex_grad + 3 * ((size_t)base + (size_t)lt);
                if (tx.x > 0.0f && w3.x != 0.0f) atomicAdd(tg + 0, w3.x / tx.x);
                if (tx.y > 0.0f && w3.y != 0.0f) atomicAdd(tg + 1, w3.y / tx.y);
                if (tx.z > 0.0f && w3.z != 0.0f) atomicAdd(tg + 2, w3.z / tx.z);
            }
        }
    }
};

template <int SV, int MODE, bool MULTI, bool COUNT = false>
__global__ __launch_bounds__(kBlock) void k_shadow(ShadowBuf sh, const uint32_t* n_ptr,
                                                   uint32_t* counter, SceneView sv,
                                                   const SceneTab* tab, float* film, AdjCtx adj, bool lanes,
                                                   unsigned long long* counts = nullptr) {
    __shared__ int2 s_stk[sv_lds(SV) * kBlock];
    int2 p_stk[sv_priv(SV) > 0 ? sv_priv(SV) : 1];
    ShadowItem<SV, MODE, MULTI, COUNT> it;
    it.sh = sh;
    it.tab = tab;
    it.nodes = sv.nodes;
    it.prims = sv.prims;
    it.rects = sv.rects;
    if constexpr (sv_tree(SV) && !MULTI) {
        __shared__ float4 s_tree[kTreeF4];
        tree_to_lds(sv, s_tree, it.nodes, it.prims, it.rects);
    }
    it.texels = sv.texels;
    it.root = sv.root;
    it.film = film;
    it.adj = adj;
#pragma unroll
    for (int k = 0; k < 7; ++k) it.G.v[k] = 0.0f;
    it.s_stk = s_stk;
    it.p_stk = p_stk;
    it.tid = threadIdx.x;
    it.id = it.nid = it.nscene = it.scene = 0;
    if (!MULTI && (sv_tree(SV) || !lanes)) {
        // one ray per lane at a time, without the persistent lanes'
        // bookkeeping: a tree in LDS always (its rays are short and cost
        // about the same), other trees when the caller asks (!lanes)
        const uint32_t n = *n_ptr;
        for (uint32_t k = blockIdx.x * kBlock + threadIdx.x; k < n; k += gridDim.x * kBlock) {
            it.prefetch(k);
            TravState ts;
            bool act = it.start(ts);
            while (act) act = it.step(ts);
            it.finish(ts);
        }
    } else {
        run_lanes(*n_ptr, counter, it);
    }
    if (MODE == kReplay) flush_grad(it.G, adj.gstripes);
    if (COUNT) {  // counts: [2] nodes, [3] primitives, [5] rays
        add_counts(counts + 2, it.cn, it.cp);
        if (blockIdx.x == 0 && threadIdx.x == 0) atomicAdd(counts + 5, (unsigned long long)*n_ptr);
    }
}

// ============================================================ BDPT
// Bidirectional path tracing after BidirectionalPathTracer::Trace
// (src/Integrator/bidirectionalpathtracer.cpp:121-394): an eye subpath
// (ConstructViewPath, :396-711) and a light subpath (ConstructLightPath,
// :713-1016) per sample, stored in HBM with area-measure forward and reverse
// pdfs; then every (s, t) connection (:179-388) -- s light vertices, t eye
// vertices, t = 1 splatting onto the film through the pinhole
// (Pinhole::FilmPlaneIntersection, pinhole.cpp:138-167) -- weighted by MIS
// over all strategies with exponent beta (beta = 1 is the reference's
// balance heuristic, MISWeight :1096-1586).  Dirac vertices (mirror,
// dielectric) are never connected and enter the MIS sums with remapped
// pdfs.  Stages: k_bdpt_init and rounds of k_extend + k_bdpt_walk build both
// subpaths of every sample of a batch as one wavefront; k_bdpt_hist and
// k_bdpt_scatter order the samples by their subpath lengths; k_bdpt_connect
// (one block per 64 samples of equal subpath lengths, one wave per strategy
// those lengths allow) evaluates the connections and queues their
// visibility tests for k_shadow.
//
// A glossy vertex keeps the one lobe Glossy::BSDFSampling picks
// (src/Material/glossy.cpp:56-78).  Lambertian lobe (probability kd_pdf): a
// matte vertex whose f and throughput carry 1/kd_pdf (Glossy::Fr :80-90 over
// _pdf_path).  Phong lobe: Dirac per Glossy::IsDirac (:35-41), sampled with
// Phong::BSDFSampling's rules (bsdf.cpp:127-170: mirrored below-horizon
// samples, a returned cosine of 1) -- the path tracer's integrand, which only
// the eye side reproduces: a light walk ends at a Phong vertex (kVPhong) and
// the MIS sums stop there, every strategy beyond needing a light walk
// through it.

constexpr uint32_t kVDelta = 1u, kVEmitter = 2u, kVCamera = 4u, kVLight = 8u, kVPhong = 16u,
                   kVMedium = 32u;
// MIS records only: the vertex before this one is a Dirac vertex
constexpr uint32_t kVPrevDelta = 64u;
constexpr uint64_t kConnMix = 0x9FB21C651E98DF25ull;  // connection streams

struct BVert {
    float4 a;  // p.xyz, pdf_fwd (area)
    float4 b;  // shading normal.xyz, w: an emitter vertex's primitive id
    float4 c;  // beta.rgb, flags
    float4 d;  // texture multiplier rgb, material (int bits)
};

struct BdptBufs {
    BVert* eye;       // [sample][vertex], vslot()
    BVert* light;
    float4* eye_mis;  // [sample][vertex]: pdf_fwd, the MIS prefix of the vertex before, flags (+
                      // kVPrevDelta), pdf_rev of the vertex before -- what the MIS weights read
    float4* light_mis;
    uint8_t* n_eye;
    uint8_t* n_light;
    uint32_t nb;      // samples in the batch (row stride)
};

struct BdptParams {
    Params base;
    int32_t n_eye, n_light, max_len;
    float beta;       // MIS exponent
    float film_area;  // (2 tan)(2 tan y_ratio) at plane distance 1
    uint64_t light_seed;
};

// A subpath vertex's slot: the vertices of one sample's subpath are
// contiguous (cap = the subpath's vertex cap), so the connections of a
// sample read its few vertices and MIS records from a few cache lines.
__device__ inline size_t vslot(uint32_t j, int v, uint32_t cap) { return (size_t)j * cap + (uint32_t)v; }

__device__ inline V3 bv_p(const BVert& v) { return v3(v.a.x, v.a.y, v.a.z); }
__device__ inline V3 bv_n(const BVert& v) { return v3(v.b.x, v.b.y, v.b.z); }
__device__ inline uint32_t bv_flags(const BVert& v) { return (uint32_t)__float_as_int(v.c.w); }
__device__ inline int bv_mat(const BVert& v) { return __float_as_int(v.d.w); }

// Medium length of the segment a -> b (0 without a medium).
__device__ __forceinline__ float seg_medium(const DevMedium& md, V3 a, V3 b) {
    if (!md.bound) return 0.0f;
    const V3 dv = b - a;
    const float dist = length(dv);
    if (!(dist > 0.0f)) return 0.0f;
    float tn, tf;
    if (!med_intersect(md, a, dv * (1.0f / dist), dist, tn, tf)) return 0.0f;
    return fmaxf(med_maxdist(dist, tn, tf), 0.0f);
}

// Density of reaching a vertex across a segment of medium length L, as the
// path tracer samples segments (distance_sample): a surface vertex passes the
// medium (probability exp(-sbar L)), a medium vertex is a sampled distance
// that scattered (sbar (1 - A) exp(-sbar L)).
__device__ inline float seg_density(const DevMedium& md, float L, bool medium_vertex) {
    if (!md.bound) return 1.0f;
    const float tr = expf(-md.avg_sigma_t * L);
    return medium_vertex ? md.avg_sigma_t * (1.0f - md.absorb) * tr : tr;
}

// Area (surface vertex) or volume (medium vertex) density at p of a direction
// sampled at `from` with solid-angle pdf pdf_w, the segment holding L of medium.
__device__ inline float to_meas(const DevMedium& md, float pdf_w, V3 from, V3 p, V3 n, bool medium,
                                float L) {
    const V3 d = p - from;
    const float d2 = dot(d, d);
    const float c = medium ? 1.0f : fabsf(dot(n, d)) / sqrtf(d2);
    return pdf_w * c / d2 * seg_density(md, L, medium);
}

__device__ inline float to_meas(const DevMedium& md, float pdf_w, V3 from, const BVert& v, float L) {
    return to_meas(md, pdf_w, from, bv_p(v), bv_n(v), (bv_flags(v) & kVMedium) != 0, L);
}

// Phase value (HomogeneousVolume::Fp without sigma_s) between the directions
// leaving a medium vertex toward its neighbours: light arrives along -wi and
// leaves along wo.
__device__ inline float phase_v(const DevMedium& md, V3 wi, V3 wo) {
    return md.g == 0.0f ? 1.0f / (4.0f * kPiF) : hg_eval(md.g, -dot(wi, wo));
}

// Non-delta BSDF value f(wi, wo) at a stored vertex (matte, the Lambertian
// lobe of glossy, GGX; sigma_s times the phase value in a medium).
__device__ inline V3 bsdf_f(const SceneView& sv, const BVert& v, V3 wi, V3 wo) {
    if (bv_flags(v) & kVMedium) return ld3(sv.med.sigma_s) * phase_v(sv.med, wi, wo);
    const DevMaterial& m = sv.mats[bv_mat(v)];
    const V3 tex = v3(v.d.x, v.d.y, v.d.z);
    if (m.kind == RUMI_MAT_MATTE) return cmul(ld3(m.color), tex) * (kInvPiF * m.kd);
    if (m.kind == RUMI_MAT_GLOSSY) return cmul(ld3(m.color), tex) * (kInvPiF * m.kd / m.kd_pdf);
    if (m.kind == RUMI_MAT_GGX) return cmul(ggx_f(m, bv_n(v), wi, wo), tex);
    return v3(0, 0, 0);
}

// Solid-angle pdf of sampling wi given wo at a stored vertex.
__device__ inline float bsdf_pdf(const SceneView& sv, const BVert& v, V3 wi, V3 wo) {
    if (bv_flags(v) & kVMedium) return phase_v(sv.med, wi, wo);
    const DevMaterial& m = sv.mats[bv_mat(v)];
    if (m.kind == RUMI_MAT_MATTE || m.kind == RUMI_MAT_GLOSSY) {
        // Lambertian::Pdfw (bsdf.cpp:34-41)
        const float p = kInvPiF * dot(bv_n(v), wi);
        return p <= 0.0f ? 1e-4f : p;
    }
    if (m.kind == RUMI_MAT_GGX) return ggx_pdf(m, bv_n(v), wi, wo);
    return 0.0f;
}

// Camera direction pdf (Pinhole::Pdfw, pinhole.cpp:58-62) and importance
// (Pinhole::We, :84-87) for world direction dc; raster position out.
__device__ inline bool camera_project(const DevCamera& cam, int W, int H, V3 dc, uint32_t& pixel,
                                      float& cos_c) {
    const float lx = cam.R[0] * dc.x + cam.R[3] * dc.y + cam.R[6] * dc.z;
    const float ly = cam.R[1] * dc.x + cam.R[4] * dc.y + cam.R[7] * dc.z;
    const float lz = cam.R[2] * dc.x + cam.R[5] * dc.y + cam.R[8] * dc.z;
    if (!(lz > 0.0f)) return false;
    const float px = lx / lz, py = ly / lz;
    // invert px = -((x - W/2) + sx) / (W/2) * tan: raster x = W/2 - px (W/2) / tan
    const float hw = W * 0.5f, hh = H * 0.5f;
    const float rx = hw - px * hw / cam.tanfov;
    const float ry = hh - py * hh / (cam.tanfov * cam.y_ratio);
    if (!(rx >= 0.0f && rx < (float)W && ry >= 0.0f && ry < (float)H)) return false;
    const int ix = min((int)rx, W - 1), iy = min((int)ry, H - 1);
    pixel = (uint32_t)iy * (uint32_t)W + (uint32_t)ix;
    cos_c = lz;
    return true;
}

// Surface point, shading normal and texture multiplier of a closest hit
// (the PT shading stage's rules).
__device__ inline void surface_at(const SceneView& sv, int32_t prim, V3 o, V3 d, float t, float bu,
                                  float bv, V3& p, V3& n, V3& tex) {
    const DevShape& s = sv.shapes[prim];
    const DevMaterial& m = sv.mats[s.info[1]];
    p = o + d * t;
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
    int tex_index;
    tex = pattern(sv, s, p, bu, bv, tex_index);
}

__device__ inline float remap0(float f) { return f != 0.0f ? f : 1.0f; }

__device__ inline float mis_pow(float r, float beta) {
    return beta == 1.0f ? r : (beta == 2.0f ? r * r : powf(r, beta));
}

__device__ inline void store_vert(BVert* arr, float4* mis, size_t idx, V3 p, float pdf_fwd, V3 n,
                                  V3 beta, uint32_t flags, V3 tex, int mat, float bw = 0.0f,
                                  float prev_rev = 0.0f, float prefix = 0.0f, uint32_t mis_bits = 0u) {
    BVert v;
    v.a = make_float4(p.x, p.y, p.z, pdf_fwd);
    v.b = make_float4(n.x, n.y, n.z, bw);
    v.c = make_float4(beta.x, beta.y, beta.z, __int_as_float((int)flags));
    v.d = make_float4(tex.x, tex.y, tex.z, __int_as_float(mat));
    arr[idx] = v;
    mis[idx] = make_float4(pdf_fwd, prefix, __int_as_float((int)(flags | mis_bits)), prev_rev);
}

// One segment of a walk (eye: radiance transport; light: importance
// transport, no 1/eta^2 on refraction) given its closest hit (prim < 0: a
// miss; t the hit distance): the medium along it, the vertex it reaches
// (stored at st.nv) and the next direction.  Returns true when the walk
// goes on with st's next segment.  The body of the reference's vertex loop
// (bidirectionalpathtracer.cpp:396-711, :713-1016), one traversal apart.
struct WalkState {
    V3 o, d, beta, prev_p;
    V3 prev_n;      // the previous vertex's stored normal (medium vertices: the direction)
    bool prev_med;  // the previous vertex is a medium vertex
    // the MIS prefix recursion (mis_prefix): the previous vertex's pdf_fwd,
    // Dirac and Phong flags, the Dirac flag of the vertex before it, and the
    // prefix of the vertex before it
    float prev_fwd, prev_sum;
    bool prev_delta, prev_phong, prevprev_delta;
    float pdf_w;
    int nv;
    bool in_obj;
};

// The MIS weight of a connection sums, over the other strategies of its
// path, products of ratios R_k = pdf_rev(k) / pdf_fwd(k) (0 read as 1) over
// the subpath vertices k, raised to beta (connect_one).  The part that does
// not depend on the connection is a prefix of each subpath, made when the
// vertex after k is stored (its reverse pdf is known then) and kept in that
// vertex's MIS record:
//   eye:   E_k = R_k^b ([c_k] + [k not Phong] E_(k-1)),  E_0 = 0 (the camera)
//   light: L_k = R_k^b ([c_k] + L_(k-1)),               L_(-1) = 0
// c_k: neither vertex k nor k - 1 is a Dirac vertex.  An eye sum stops at a
// Phong vertex, as the MIS walk of the reference stops there.
template <bool EYE>
__device__ inline float mis_prefix(const WalkState& st, int k, float rev, float beta) {
    if (EYE && k < 1) return 0.0f;
    const float rb = mis_pow(remap0(rev) / remap0(st.prev_fwd), beta);
    const float c = !st.prev_delta && !st.prevprev_delta ? 1.0f : 0.0f;
    return rb * (c + (EYE && st.prev_phong ? 0.0f : st.prev_sum));
}

// The state after vertex flags/pdf_fwd were stored with MIS prefix `prefix`.
__device__ inline void mis_advance(WalkState& st, float pdf_fwd, uint32_t flags, float prefix) {
    st.prevprev_delta = st.prev_delta;
    st.prev_delta = (flags & kVDelta) != 0u;
    st.prev_phong = (flags & kVPhong) != 0u;
    st.prev_fwd = pdf_fwd;
    st.prev_sum = prefix;
}

template <bool EYE>
__device__ bool walk_segment(const SceneView& sv, BVert* arr, float4* mis, uint32_t cap, uint32_t j,
                             WalkState& st, int32_t prim, float t, float bu, float bv, Rng& g, float mis_beta) {
    int& nv = st.nv;
    bool& in_obj = st.in_obj;
    V3& prev_p = st.prev_p;
    V3& prev_n = st.prev_n;
    bool& prev_med = st.prev_med;
    V3& o = st.o;
    V3& d = st.d;
    V3& beta = st.beta;
    float& pdf_w = st.pdf_w;
    // the medium along the segment, sampled as the path tracer does
    // (pathtracer.cpp:156-178; bidirectionalpathtracer.cpp:452-473)
    const DevMedium& md = sv.med;
    float seg_L = 0.0f;
    if (md.bound && !in_obj) {
        const float thit = prim >= 0 ? t : 0.0f;
        float tn, tf;
        if (med_intersect(md, o, d, thit, tn, tf)) {
            const float smax = med_maxdist(thit, tn, tf);
            float sd = 0.0f;
            V3 ratio;
            const int ev = distance_sample(md.sigma_t, md.avg_sigma_t, md.absorb, smax, g, sd, ratio);
            if (ev == kAbsorb) return false;
            beta = cmul(beta, ratio);
            if (ev == kScatter) {
                // a medium vertex: HomogeneousVolume::PhaseSampling, Fp / pdf = sigma_s
                const V3 mp = o + d * (tn < 0.0f ? sd : tn + sd);
                const size_t mslot = vslot(j, nv, cap);
                float pdf_ph;
                const V3 w = phase_sample(md.g, d, g, pdf_ph);
                // the previous vertex's reverse pdf (area measure there),
                // kept in this vertex's MIS record
                const float prev_rev = nv >= 1 ? to_meas(md, pdf_ph, mp, prev_p, prev_n, prev_med, sd) : 0.0f;
                const float pre = mis_prefix<EYE>(st, nv - 1, prev_rev, mis_beta);
                const float fwd = to_meas(md, pdf_w, prev_p, mp, d, true, sd);
                store_vert(arr, mis, mslot, mp, fwd, d, beta, kVMedium, v3(1, 1, 1), -1, 0.0f, prev_rev, pre,
                           st.prev_delta ? kVPrevDelta : 0u);
                mis_advance(st, fwd, kVMedium, pre);
                ++nv;
                beta = cmul(beta, ld3(md.sigma_s));
                pdf_w = pdf_ph;
                prev_p = mp;
                prev_n = d;
                prev_med = true;
                o = mp;
                d = w;
                in_obj = false;
                return true;
            }
            seg_L = smax;
        }
    }
    if (prim < 0) return false;
    V3 p, n, tex;
    surface_at(sv, prim, o, d, t, bu, bv, p, n, tex);
    const int mi = sv.shapes[prim].info[1];
    const DevMaterial& m = sv.mats[mi];
    V3 vol = v3(1, 1, 1);  // subsurface pass-through tau / pdf at the exit boundary
    if (m.kind == RUMI_MAT_SUBSURFACE && in_obj) {
        // Subsurface::PhaseSampling (subsurface.cpp:135-155): the interior
        // walk over the segment to the boundary, as the path tracer takes
        // it.  Interior vertices are never connected (a connection from
        // inside ends on the boundary), so they enter the MIS walk like
        // Dirac vertices.
        float sd = 0.0f;
        V3 ratio;
        const int ev = distance_sample(m.sigma_t, m.avg_sigmat, m.absorb, t, g, sd, ratio);
        if (ev == kAbsorb) return false;
        if (ev == kScatter) {
            const V3 ps = o + d * sd;
            int tix;
            const V3 txs = sv.shapes[prim].info[2] >= 0 ? pattern(sv, sv.shapes[prim], ps, bu, bv, tix)
                                                       : v3(1, 1, 1);
            // (a Dirac vertex: the previous vertex's reverse pdf is 0)
            const float pre = mis_prefix<EYE>(st, nv - 1, 0.0f, mis_beta);
            store_vert(arr, mis, vslot(j, nv, cap), ps, 0.0f, d, beta, kVDelta, txs, mi, 0.0f, 0.0f, pre,
                       st.prev_delta ? kVPrevDelta : 0u);
            mis_advance(st, 0.0f, kVDelta, pre);
            ++nv;
            beta = cmul(cmul(cmul(beta, ratio), ld3(m.sigma_s)), txs);
            float pdf_ph;
            const V3 w = phase_sample(0.0f, d, g, pdf_ph);
            pdf_w = 0.0f;
            prev_p = ps;
            prev_n = d;
            prev_med = false;
            o = ps;
            d = w;
            return true;
        }
        vol = ratio;
    }
    const float pdf_fwd = to_meas(md, pdf_w, prev_p, p, n, false, seg_L);
    if (m.kind == RUMI_MAT_EMISSIVE) {
        // The walk ends on an emitter (bidirectionalpathtracer.cpp:693-699).
        // an emitter vertex is never scattered from, so its reverse-pdf
        // word carries the primitive id instead
        store_vert(arr, mis, vslot(j, nv, cap), p, pdf_fwd, n, beta, kVEmitter, tex, mi, __int_as_float(prim),
                   0.0f, mis_prefix<EYE>(st, nv - 1, 0.0f, mis_beta), st.prev_delta ? kVPrevDelta : 0u);
        ++nv;
        return false;
    }
    const size_t slot = vslot(j, nv, cap);
    ++nv;
    const bool was_in = in_obj;
    const V3 wo = -d;
    V3 w;
    float pdf_next = 0.0f, pdf_rev = 0.0f;
    bool delta = false;
    uint32_t vflags = 0u;
    V3 weight;  // f |cos| / pdf of this bounce
    if (m.kind == RUMI_MAT_MATTE) {
        float cost, pdf;
        w = lambert_sample(n, g, cost, pdf);
        weight = cmul(ld3(m.color), tex) * (m.kd * kInvPiF * cost / pdf);
        pdf_next = pdf;
        const float pr = kInvPiF * dot(n, wo);
        pdf_rev = pr <= 0.0f ? 1e-4f : pr;
    } else if (m.kind == RUMI_MAT_GLOSSY) {
        if (m.kd != 0.0f && g.uniform() <= m.kd_pdf) {
            float cost, pdf;
            w = lambert_sample(n, g, cost, pdf);
            weight = cmul(ld3(m.color), tex) * (m.kd * kInvPiF * cost / (pdf * m.kd_pdf));
            pdf_next = pdf;
            const float pr = kInvPiF * dot(n, wo);
            pdf_rev = pr <= 0.0f ? 1e-4f : pr;
        } else {
            if (!EYE) {  // a light walk ends at a Phong-lobe vertex
                store_vert(arr, mis, slot, p, pdf_fwd, n, beta, kVDelta | kVPhong, tex, mi, 0.0f, 0.0f,
                           mis_prefix<EYE>(st, nv - 2, 0.0f, mis_beta), st.prev_delta ? kVPrevDelta : 0u);
                return false;
            }
            float cost, pdf;
            w = phong_sample(m, wo, n, g, cost, pdf);
            weight = tex * (phong_f(m, n, w, wo) * m.ks * cost / (pdf * (1.0f - m.kd_pdf)));
            delta = true;
            vflags = kVPhong;
        }
    } else if (m.kind == RUMI_MAT_GGX) {
        w = ggx_sample(m, n, wo, g);
        const float cost = dot(n, w);
        pdf_next = ggx_pdf(m, n, w, wo);
        if (!(cost > 0.0f && pdf_next > 0.0f)) {
            store_vert(arr, mis, slot, p, pdf_fwd, n, beta, 0u, tex, mi, 0.0f, 0.0f,
                       mis_prefix<EYE>(st, nv - 2, 0.0f, mis_beta), st.prev_delta ? kVPrevDelta : 0u);
            return false;
        }
        weight = cmul(ggx_f(m, n, w, wo), tex) * (cost / pdf_next);
        pdf_rev = ggx_pdf(m, n, wo, w);
    } else if (m.kind == RUMI_MAT_SPECULAR) {
        float won = dot(wo, n);
        if (won <= 0.0f) won = 1e-4f;
        w = -wo + n * (2.0f * won);
        weight = cmul(ld3(m.color), tex) * m.kd;
        delta = true;
    } else {  // dielectric, subsurface boundary
        const float kr = fresnel(-wo, n, m.eta_in, m.eta_out);
        const float c = dot(wo, n);
        delta = true;
        if (kr != 0.0f && g.uniform() <= kr) {
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
            in_obj = c < 0.0f;
            weight = v3(1, 1, 1);
        } else {
            float eta = m.eta_in / m.eta_out;
            V3 nn = n;
            float ci = c;
            if (ci < 0.0f) {
                nn = -n;
                eta = m.eta_out / m.eta_in;
                ci = -ci;
            }
            const float inv = 1.0f / eta;
            const float ct = sqrtf(1.0f - (1.0f - ci * ci) * (inv * inv));
            w = -wo * inv + nn * (ci * inv - ct);
            in_obj = c > 0.0f;
            // Dielectric::Fr (dielectric.cpp:73-98): radiance is scaled by
            // 1/eta^2, eta = eta_in/eta_out from inside, eta_out/eta_in from
            // outside; importance carried by light paths is not scaled.
            const float er = was_in ? m.eta_in / m.eta_out : m.eta_out / m.eta_in;
            const float scale = EYE ? 1.0f / (er * er) : 1.0f;
            weight = v3(scale, scale, scale);
        }
        weight = cmul(weight, tex);
        // Beer-Lambert along a segment travelled inside (dielectric.cpp:94-95);
        // a subsurface interior passed through without scattering
        if (was_in && m.kind == RUMI_MAT_DIELECTRIC)
            weight = cmul(weight, v3(expf(-m.absorption[0] * t), expf(-m.absorption[1] * t),
                                     expf(-m.absorption[2] * t)));
        if (was_in && m.kind == RUMI_MAT_SUBSURFACE) weight = cmul(weight, vol);
    }
    // the previous vertex's reverse pdf (area measure there), kept in this
    // vertex's MIS record
    const float prev_rev = nv >= 2 && !delta ? to_meas(md, pdf_rev, p, prev_p, prev_n, prev_med, seg_L) : 0.0f;
    const uint32_t fl = vflags | (delta ? kVDelta : 0u);
    const float pre = mis_prefix<EYE>(st, nv - 2, prev_rev, mis_beta);
    store_vert(arr, mis, slot, p, pdf_fwd, n, beta, fl, tex, mi, 0.0f, prev_rev, pre,
               st.prev_delta ? kVPrevDelta : 0u);
    mis_advance(st, pdf_fwd, fl, pre);
    beta = cmul(beta, weight);
    pdf_w = delta ? 0.0f : pdf_next;
    prev_p = p;
    prev_n = n;
    prev_med = false;
    o = p;
    d = w;
    if (m.kind != RUMI_MAT_DIELECTRIC && m.kind != RUMI_MAT_SUBSURFACE) in_obj = false;
    return true;
}

// Walk state of the wavefront BDPT walks: eye and light walks of a batch in
// one queue, double-buffered.  ray and flags are laid out as PathBuf's, so
// the path tracer's persistent-lane extension stage (k_extend) traces them.
struct WalkBuf {
    float4* ray;       // 2 per walk: o.xyz 0 | d.xyz FLT_MAX
    uint32_t* flags;   // bit 1 in_object, bit 2 light walk, bits 8.. vertices stored
    float4* beta;      // throughput rgb, pdf_w of the segment's direction
    float4* prev;      // the previous vertex's position
    float4* prevn;     // the previous vertex's stored normal, w: 1 for a medium vertex
    uint64_t* rng;     // PCG32 state (increment from the pixel)
    uint32_t* slot;    // sample index within the batch
};

constexpr uint32_t kWalkLight = 4u;

__device__ inline void write_walk(const WalkBuf& wb, uint32_t q, const WalkState& st, uint64_t rng, uint32_t light,
                                  uint32_t j) {
    wb.ray[2 * (size_t)q + 0] = make_float4(st.o.x, st.o.y, st.o.z, 0.0f);
    wb.ray[2 * (size_t)q + 1] = make_float4(st.d.x, st.d.y, st.d.z, FLT_MAX);
    wb.flags[q] = (st.in_obj ? 2u : 0u) | light | ((uint32_t)st.nv << 8) | (st.prev_med ? 1u << 16 : 0u) |
                  (st.prev_delta ? 1u << 17 : 0u) | (st.prev_phong ? 1u << 18 : 0u) |
                  (st.prevprev_delta ? 1u << 19 : 0u);
    wb.beta[q] = make_float4(st.beta.x, st.beta.y, st.beta.z, st.pdf_w);
    wb.prev[q] = make_float4(st.prev_p.x, st.prev_p.y, st.prev_p.z, st.prev_fwd);
    wb.prevn[q] = make_float4(st.prev_n.x, st.prev_n.y, st.prev_n.z, st.prev_sum);
    wb.rng[q] = rng;
    wb.slot[q] = j;
}

// The first vertex of every walk of the batch (thread w < nb: the eye walk
// of sample w from the camera, ConstructViewPath :396-420; w >= nb: the
// light walk of sample w - nb from a light point with a cosine-weighted
// direction, Scene::LightsUniformSampling scene.cpp:91-97 +
// AreaLight::SolidAngleSampling arealight.cpp:31-60), queued for the walk
// stages.
__device__ inline bool init_walk(const BdptBufs& bb, uint64_t base, const BdptParams& bp, const SceneView& sv,
                                 uint32_t w, WalkState& st, Rng& g, uint32_t& j_out, bool& eye_out) {
    const Params& prm = bp.base;
    const bool eye = w < bb.nb;
    const uint32_t j = eye ? w : w - bb.nb;
    const uint64_t k = base + j;
    const uint32_t pixel = (uint32_t)(k % prm.pixels);
    const uint64_t s = k / prm.pixels;
    st.nv = 1;
    st.in_obj = false;
    st.prev_sum = 0.0f;
    st.prev_fwd = 1.0f;
    st.prev_delta = st.prev_phong = st.prevprev_delta = st.prev_med = false;
    bool go = false;
    if (eye) {
        g = rng_seed(prm.seed + s * 0x9E3779B97F4A7C15ull, pixel);
        const DevCamera& cam = sv.cam;
        const int x = (int)(pixel % (uint32_t)prm.width);
        const int y = (int)(pixel / (uint32_t)prm.width);
        const float sx = g.uniform(), sy = g.uniform();
        const float hw = prm.width * 0.5f, hh = prm.height * 0.5f;
        const float px = -(((float)x - hw) + sx) / hw * cam.tanfov;
        const float py = -(((float)y - hh) + sy) / hh * cam.tanfov * cam.y_ratio;
        const V3 l = normalize(v3(px, py, 1.0f));
        st.d = v3(cam.R[0] * l.x + cam.R[1] * l.y + cam.R[2] * l.z, cam.R[3] * l.x + cam.R[4] * l.y + cam.R[5] * l.z,
                  cam.R[6] * l.x + cam.R[7] * l.y + cam.R[8] * l.z);
        const V3 cp = v3(cam.p[0], cam.p[1], cam.p[2]);
        const V3 cdir = v3(cam.R[2], cam.R[5], cam.R[8]);
        store_vert(bb.eye, bb.eye_mis, vslot(j, 0, bp.n_eye), cp, 1.0f, cdir, v3(1, 1, 1), kVCamera, v3(1, 1, 1),
                   -1);
        // Pinhole::Pdfw: 1 / (cos^3 film_area) at plane distance 1
        st.pdf_w = 1.0f / (l.z * l.z * l.z * bp.film_area);
        st.o = st.prev_p = cp;
        st.prev_n = cdir;
        st.prev_med = false;
        st.prev_fwd = 1.0f;
        st.beta = v3(1, 1, 1);
        bb.n_eye[j] = 1;
        go = bp.n_eye > 1;
    } else {
        g = rng_seed(bp.light_seed + s * 0x9E3779B97F4A7C15ull, pixel);
        if (sv.n_lights == 0) {
            bb.n_light[j] = 0;
        } else {
            uint32_t li = (uint32_t)(g.uniform() * (float)sv.n_lights);
            if (li >= sv.n_lights) li = sv.n_lights - 1;
            const DevLight& L = sv.lights[li];
            V3 p0, n0;
            if (L.kind == RUMI_SHAPE_RECT) {
                const float a = g.uniform(), b = g.uniform();
                p0 = (ld3(L.p) + ld3(L.l) * a) + ld3(L.w) * b;
                n0 = ld3(L.n);
            } else {
                const float z = 1.0f - 2.0f * g.uniform(), phi = 2.0f * kPiF * g.uniform();
                const float rr = sqrtf(fmaxf(0.0f, 1.0f - z * z));
                p0 = ld3(L.p) + v3(rr * cosf(phi), rr * sinf(phi), z) * L.l[0];
                n0 = normalize(p0 - ld3(L.p));
            }
            const float pdf_pos = 1.0f / (float)sv.n_lights * L.p[3];
            const V3 le = ld3(L.le);
            store_vert(bb.light, bb.light_mis, vslot(j, 0, bp.n_light), p0, pdf_pos, n0, le * (1.0f / pdf_pos),
                       kVLight, v3(1, 1, 1), -1);
            float cost, pdf_dir;
            st.d = lambert_sample(n0, g, cost, pdf_dir);
            st.beta = le * (cost / (pdf_pos * pdf_dir));
            st.pdf_w = pdf_dir;
            st.o = st.prev_p = p0;
            st.prev_n = n0;
            st.prev_med = false;
            st.prev_fwd = pdf_pos;
            bb.n_light[j] = 1;
            go = bp.n_light > 1;
        }
    }
    j_out = j;
    eye_out = eye;
    return go;
}

// Every eye walk goes on after its first vertex (n_eye >= 2), and every
// light walk does when the scene has a light and n_light >= 2, so the
// queue's slots follow from the walk's index: eye walks first, then light
// walks, and no counter is needed.
__global__ __launch_bounds__(kBlock) void k_bdpt_init(BdptBufs bb, uint64_t base, BdptParams bp,
                                                      SceneView sv, WalkBuf wb, uint32_t* count) {
    const uint32_t w = blockIdx.x * kBlock + threadIdx.x;
    const bool eye_go = bp.n_eye > 1, light_go = sv.n_lights > 0 && bp.n_light > 1;
    if (w == 0) *count = (eye_go ? bb.nb : 0u) + (light_go ? bb.nb : 0u);
    if (w >= 2 * bb.nb) return;
    WalkState st;
    Rng g;
    uint32_t j = 0;
    bool eye = true;
    const bool go = init_walk(bb, base, bp, sv, w, st, g, j, eye);
    if (go) write_walk(wb, eye ? j : (eye_go ? bb.nb : 0u) + j, st, g.state, eye ? 0u : kWalkLight, j);
}

// The segment of queued walk i, its closest hit already found by k_extend
// (k_bdpt_walk); true when the walk goes on, else its vertex count is
// recorded.
__device__ inline bool walk_one(const BdptBufs& bb, uint64_t base, const BdptParams& bp, const SceneView& sv,
                                const WalkBuf& in, uint32_t i, const int2* hits, const float2* uvs, WalkState& st,
                                Rng& g, uint32_t& j_out, bool& light_out) {
    const uint32_t flags = in.flags[i];
    const uint32_t j = in.slot[i];
    const bool light = (flags & kWalkLight) != 0;
    const uint32_t pixel = (uint32_t)((base + j) % bp.base.pixels);
    g.state = in.rng[i];
    g.inc = ((uint64_t)pixel << 1u) | 1u;
    const float4 r0 = in.ray[2 * (size_t)i + 0], r1 = in.ray[2 * (size_t)i + 1];
    const float4 b4 = in.beta[i], p4 = in.prev[i];
    st.o = v3(r0.x, r0.y, r0.z);
    st.d = v3(r1.x, r1.y, r1.z);
    st.beta = v3(b4.x, b4.y, b4.z);
    st.pdf_w = b4.w;
    st.prev_p = v3(p4.x, p4.y, p4.z);
    st.prev_fwd = p4.w;
    const float4 n4 = in.prevn[i];
    st.prev_n = v3(n4.x, n4.y, n4.z);
    st.prev_sum = n4.w;
    st.prev_med = (flags >> 16) & 1u;
    st.prev_delta = (flags >> 17) & 1u;
    st.prev_phong = (flags >> 18) & 1u;
    st.prevprev_delta = (flags >> 19) & 1u;
    st.nv = (int)((flags >> 8) & 0xffu);
    st.in_obj = (flags & 2u) != 0;
    const int2 h = hits[i];
    const float2 uv = uvs[i];
    const float t = __int_as_float(h.y);
    bool go = light ? walk_segment<false>(sv, bb.light, bb.light_mis, bp.n_light, j, st, h.x, t, uv.x, uv.y, g,
                                          bp.beta)
                    : walk_segment<true>(sv, bb.eye, bb.eye_mis, bp.n_eye, j, st, h.x, t, uv.x, uv.y, g, bp.beta);
    go = go && st.nv < (light ? bp.n_light : bp.n_eye);
    if (!go) {
        if (light)
            bb.n_light[j] = (uint8_t)st.nv;
        else
            bb.n_eye[j] = (uint8_t)st.nv;
    }
    j_out = j;
    light_out = light;
    return go;
}

// One segment of every queued walk; walks that go on are appended to the
// next queue (one atomic per block of 256: the queue counter's same-address
// atomics serialise).
constexpr int kWalkBlock = 256;

// Three waves per SIMD (168 VGPRs, a few spills): left to itself the
// compiler takes 180 and two waves (whole-frame BDPT 345 -> 341 ms; four
// waves spill more, 345 ms).
__global__ __launch_bounds__(kWalkBlock) __attribute__((amdgpu_waves_per_eu(3))) void k_bdpt_walk(BdptBufs bb, uint64_t base, BdptParams bp, SceneView sv,
                                                          WalkBuf in, const uint32_t* n_ptr, const int2* hits,
                                                          const float2* uvs, WalkBuf out, uint32_t* next_count) {
    __shared__ uint32_t s_tmp[kWalkBlock / 64 + 1];
    const uint32_t n = *n_ptr;
    // block-uniform trip count: every thread reaches the block-aggregated append
    for (uint32_t i0 = blockIdx.x * kWalkBlock; i0 < n; i0 += gridDim.x * kWalkBlock) {
        const uint32_t i = i0 + threadIdx.x;
        bool go = false;
        WalkState st;
        Rng g;
        uint32_t j = 0;
        bool light = false;
        if (i < n) go = walk_one(bb, base, bp, sv, in, i, hits, uvs, st, g, j, light);
        const uint32_t q = block_append<kWalkBlock>(next_count, go, s_tmp);
        if (go) write_walk(out, q, st, g.state, light ? kWalkLight : 0u, j);
    }
}


// Connection work list.  After the walks, the samples of a batch are put in
// the order of their subpath lengths, (n_eye, n_light) both descending
// (k_bdpt_hist counts each pair, the host turns the counts into the first
// position of each pair, k_bdpt_scatter writes every sample's index at its
// position).  k_bdpt_connect then takes the samples of one pair 64 at a time
// ("tiles"): every sample of a tile can make the same strategies, so each
// wave runs one strategy over a whole tile (no idle lanes, no divergence
// between strategies), and the block's waves go through all of the tile's
// strategies while its subpaths sit in the caches.  Both kernels give every
// block the same contiguous range of samples.
constexpr uint32_t kMaxBins = 65u * 65u;  // (n_eye, n_light) pairs, caps of 64

__device__ inline void bin_range(uint32_t nb, uint32_t& j0, uint32_t& j1) {
    const uint32_t per = (nb + gridDim.x - 1) / gridDim.x;
    j0 = blockIdx.x * per;
    j1 = min(nb, j0 + per);
}

__global__ __launch_bounds__(kBlock) void k_bdpt_hist(const uint8_t* n_eye, const uint8_t* n_light, uint32_t nb,
                                                      uint32_t l1, uint32_t nbins, uint32_t* hist) {
    __shared__ uint32_t h[kMaxBins];
    for (uint32_t i = threadIdx.x; i < nbins; i += kBlock) h[i] = 0u;
    __syncthreads();
    uint32_t j0, j1;
    bin_range(nb, j0, j1);
    for (uint32_t j = j0 + threadIdx.x; j < j1; j += kBlock) atomicAdd(&h[n_eye[j] * l1 + n_light[j]], 1u);
    __syncthreads();
    for (uint32_t i = threadIdx.x; i < nbins; i += kBlock)
        if (h[i]) atomicAdd(hist + i, h[i]);
}

// cursor[bin] holds the bin's first position on entry; each block reserves
// a run per bin for its samples, then places them.
__global__ __launch_bounds__(kBlock) void k_bdpt_scatter(const uint8_t* n_eye, const uint8_t* n_light,
                                                         uint32_t nb, uint32_t l1, uint32_t nbins,
                                                         uint32_t* cursor, uint32_t* perm) {
    __shared__ uint32_t h[kMaxBins];
    for (uint32_t i = threadIdx.x; i < nbins; i += kBlock) h[i] = 0u;
    __syncthreads();
    uint32_t j0, j1;
    bin_range(nb, j0, j1);
    for (uint32_t j = j0 + threadIdx.x; j < j1; j += kBlock) atomicAdd(&h[n_eye[j] * l1 + n_light[j]], 1u);
    __syncthreads();
    for (uint32_t i = threadIdx.x; i < nbins; i += kBlock)
        if (h[i]) h[i] = atomicAdd(cursor + i, h[i]);
    __syncthreads();
    for (uint32_t j = j0 + threadIdx.x; j < j1; j += kBlock)
        perm[atomicAdd(&h[n_eye[j] * l1 + n_light[j]], 1u)] = j;
}


// A queued visibility test of a connection: the segment (origin, unit
// direction, length) and the weighted contribution with its pixel.
struct ShadowEntry {
    float4 o, d, c;
};

// Connection (s, t) of sample j, whose subpaths have at least t eye and s
// light vertices: a contribution that needs no visibility test goes to the
// film; for the others the visibility test is returned in e (true), to be
// queued for k_shadow, which resolves them with persistent lanes as the
// path tracer's shadow rays are.
__device__ __forceinline__ bool connect_one(const BdptBufs& bb, uint64_t base, const BdptParams& bp, const SceneView& sv, int s,
                            int t, uint32_t j, float* film, ShadowEntry& e) {
    const uint32_t nEc = (uint32_t)bp.n_eye, nLc = (uint32_t)bp.n_light;
    const Params& prm = bp.base;
    const BVert pt = bb.eye[vslot(j, t - 1, nEc)];
    const uint32_t ptf = bv_flags(pt);
    V3 c;
    // the sample's own pixel; a t = 1 connection splats elsewhere
    uint32_t pixel = (uint32_t)((base + j) % prm.pixels);
    bool mis_trivial = false;
    float pt_rev = 0.0f, ptm_rev = 0.0f, qs_rev = 0.0f, qsm_rev = 0.0f;
    bool visible_test = false;
    V3 va, vb;
    const DevMedium& md = sv.med;
    float medL = 0.0f;  // medium length of the connecting segment
    if (s == 0) {
        if (!(ptf & kVEmitter) || t < 2) return false;
        const BVert pm = bb.eye[vslot(j, t - 2, nEc)];
        const V3 to_prev = normalize(bv_p(pm) - bv_p(pt));
        const DevMaterial& m = sv.mats[bv_mat(pt)];
        // Emissive::DirectLighting front-face rule (emissive.cpp:60-74)
        if (!(dot(bv_n(pt), to_prev) > 0.0f)) return false;
        c = cmul(v3(pt.c.x, pt.c.y, pt.c.z), ld3(m.color) * m.kd);
        const int prim = __float_as_int(pt.b.w);
        const int li = __float_as_int(sv.shapes[prim].uv2[3]);
        if (li < 0) {
            // an emitter that is not a sampled light: no other strategy
            mis_trivial = true;
        } else {
            const DevLight& L = sv.lights[li];
            pt_rev = 1.0f / (float)sv.n_lights * L.p[3];
            const float cl = fmaxf(dot(bv_n(pt), to_prev), 0.0f);
            if (!(bv_flags(pm) & kVCamera))
                ptm_rev = to_meas(md, kInvPiF * cl, bv_p(pt), pm, seg_medium(md, bv_p(pt), bv_p(pm)));
        }
    } else {
        const BVert qs = bb.light[vslot(j, s - 1, nLc)];
        const uint32_t qsf = bv_flags(qs);
        if ((qsf & kVDelta) || (qsf & kVEmitter)) return false;
        if (t == 1) {
            // light tracing: connect qs to the pinhole, splat (t_idx == 0 case, :367-380)
            const V3 cp = v3(sv.cam.p[0], sv.cam.p[1], sv.cam.p[2]);
            const V3 dc = normalize(bv_p(qs) - cp);
            float cos_c;
            if (!camera_project(sv.cam, prm.width, prm.height, dc, pixel, cos_c)) return false;
            const V3 to_cam = -dc;
            const float d2 = dot(bv_p(qs) - cp, bv_p(qs) - cp);
            V3 fL;
            const float cos_q = (qsf & kVMedium) ? 1.0f : dot(bv_n(qs), to_cam);
            if (!(cos_q > 0.0f)) return false;
            if (s == 1) {
                fL = v3(1, 1, 1);
            } else {
                const BVert qm = bb.light[vslot(j, s - 2, nLc)];
                fL = bsdf_f(sv, qs, to_cam, normalize(bv_p(qm) - bv_p(qs)));
                qsm_rev = to_meas(md, bsdf_pdf(sv, qs, normalize(bv_p(qm) - bv_p(qs)), to_cam), bv_p(qs), qm,
                                  seg_medium(md, bv_p(qs), bv_p(qm)));
            }
            // Pinhole::We = 1 / (A cos^4) (pinhole.cpp:84-87); G = cos_q cos_c / d^2
            const float We = 1.0f / (bp.film_area * cos_c * cos_c * cos_c * cos_c);
            const float G = cos_q * cos_c / d2;
            c = cmul(v3(qs.c.x, qs.c.y, qs.c.z), fL) * (We * G);
            medL = seg_medium(md, bv_p(qs), cp);
            qs_rev = to_meas(md, 1.0f / (cos_c * cos_c * cos_c * bp.film_area), cp, qs, medL);
            visible_test = true;
            va = bv_p(qs);
            vb = cp;
        } else {
            if ((ptf & kVDelta) || (ptf & kVEmitter)) return false;
            const BVert pm = bb.eye[vslot(j, t - 2, nEc)];
            const V3 d_pq = bv_p(qs) - bv_p(pt);
            const float d2 = dot(d_pq, d_pq);
            const V3 w_pq = d_pq * (1.0f / sqrtf(d2));
            const float cos_p = (ptf & kVMedium) ? 1.0f : dot(bv_n(pt), w_pq);
            const float cos_q = (qsf & kVMedium) ? 1.0f : -dot(bv_n(qs), w_pq);
            if (!(cos_p > 0.0f && cos_q > 0.0f)) return false;
            const V3 to_pm = normalize(bv_p(pm) - bv_p(pt));
            const V3 fE = bsdf_f(sv, pt, w_pq, to_pm);
            V3 fL;
            medL = seg_medium(md, bv_p(pt), bv_p(qs));
            if (s == 1) {
                fL = v3(1, 1, 1);
                pt_rev = to_meas(md, kInvPiF * cos_q, bv_p(qs), pt, medL);
            } else {
                const BVert qm = bb.light[vslot(j, s - 2, nLc)];
                const V3 to_qm = normalize(bv_p(qm) - bv_p(qs));
                fL = bsdf_f(sv, qs, -w_pq, to_qm);
                pt_rev = to_meas(md, bsdf_pdf(sv, qs, -w_pq, to_qm), bv_p(qs), pt, medL);
                qsm_rev = to_meas(md, bsdf_pdf(sv, qs, to_qm, -w_pq), bv_p(qs), qm,
                                  seg_medium(md, bv_p(qs), bv_p(qm)));
            }
            const float G = cos_p * cos_q / d2;
            c = cmul(cmul(cmul(v3(pt.c.x, pt.c.y, pt.c.z), fE), fL), v3(qs.c.x, qs.c.y, qs.c.z)) * G;
            qs_rev = to_meas(md, bsdf_pdf(sv, pt, w_pq, to_pm), bv_p(pt), qs, medL);
            if (!(bv_flags(pm) & kVCamera))
                ptm_rev = to_meas(md, bsdf_pdf(sv, pt, to_pm, w_pq), bv_p(pt), pm,
                                  seg_medium(md, bv_p(pt), bv_p(pm)));
            visible_test = true;
            va = bv_p(qs);
            vb = bv_p(pt);
        }
    }
    if (c.x == 0.0f && c.y == 0.0f && c.z == 0.0f) return false;
    // MIS weight over all strategies of this path (power heuristic with exponent beta)
    float w = 1.0f;
    if (s + t > 2 && !mis_trivial) {
        // the other strategies of this path through the subpaths' MIS
        // prefixes (mis_prefix): the two vertices at each end of the
        // connection take the connection's reverse pdfs, the rest is the
        // prefix kept in the record of the second vertex from the end
        float sum = 0.0f;
        if (t >= 2) {
            const float4 e1 = bb.eye_mis[vslot(j, t - 1, nEc)], e2 = bb.eye_mis[vslot(j, t - 2, nEc)];
            const uint32_t f1 = (uint32_t)__float_as_int(e1.z), f2 = (uint32_t)__float_as_int(e2.z);
            const float a = remap0(pt_rev) / remap0(e1.x);
            if (!(f2 & kVDelta)) sum += mis_pow(a, bp.beta);
            if (!(f1 & kVPhong) && t >= 3) {
                const float ab = a * (remap0(ptm_rev) / remap0(e2.x));
                const float abb = mis_pow(ab, bp.beta);
                if (!(f2 & (kVDelta | kVPrevDelta))) sum += abb;
                if (!(f2 & kVPhong) && t >= 4) sum += abb * e2.y;
            }
        }
        if (s >= 1) {
            const float4 l1 = bb.light_mis[vslot(j, s - 1, nLc)];
            const float4 l2 = s >= 2 ? bb.light_mis[vslot(j, s - 2, nLc)] : make_float4(0, 0, 0, 0);
            const uint32_t f2 = (uint32_t)__float_as_int(l2.z);
            const float a = remap0(qs_rev) / remap0(l1.x);
            if (!(f2 & kVDelta)) sum += mis_pow(a, bp.beta);
            if (s >= 2) {
                const float ab = a * (remap0(qsm_rev) / remap0(l2.x));
                const float abb = mis_pow(ab, bp.beta);
                if (!(f2 & (kVDelta | kVPrevDelta))) sum += abb;
                if (s >= 3) sum += abb * l2.y;
            }
        }
        w = 1.0f / (1.0f + sum);
    }
    // the connection's medium: a grey Bernoulli survival, as VolumetricTest
    // (bidirectionalpathtracer.cpp:1050-1062) and the path tracer's shadow
    // rays; one number from the connection's own stream
    if (medL > 0.0f) {
        Rng gc = rng_seed((prm.seed ^ kConnMix) + (base + j) * 0x9E3779B97F4A7C15ull,
                          ((uint64_t)s << 8) | (uint64_t)t);
        if (gc.uniform() < 1.0f - expf(-md.avg_sigma_t * medL)) return false;
    }
    const V3 cw = c * w;
    if (!visible_test) {
        film_add(film, pixel, cw);
        return false;
    }
    // VisibilityTest (bidirectionalpathtracer.cpp:1018-1048): the segment
    // va -> vb as an any-hit query
    const V3 dv = vb - va;
    const float dist = length(dv);
    const V3 dir = dv * (1.0f / dist);
    e.o = make_float4(va.x, va.y, va.z, 0.0f);
    e.d = make_float4(dir.x, dir.y, dir.z, dist);
    e.c = make_float4(cw.x, cw.y, cw.z, __int_as_float((int32_t)pixel));
    return true;
}

constexpr int kConnBlock = 256;  // 4 waves per tile of 64 samples
constexpr uint32_t kConnStage = 640;  // staged visibility tests per block (30 KB of LDS)

// The connections of tiles tile0 .. tile0 + gridDim.x - 1, one tile per
// block.  bins[k] = (n_eye, n_light, first position, sample count) of the
// k-th non-empty length pair in position order, bin_tile[k] its first tile.
// The strategies of a tile are all (s, t) with t <= n_eye, s <= n_light,
// s + t >= 2, s + t - 1 <= max_len, except (1, 1) (the camera ray reaching
// the light covers that path); the block takes them four at a time, one per
// wave.
__global__ __launch_bounds__(kConnBlock) void k_bdpt_connect(BdptBufs bb, uint64_t base, BdptParams bp,
                                                             SceneView sv, const int4* bins,
                                                             const uint32_t* bin_tile, uint32_t n_bins,
                                                             uint32_t tile0, const uint32_t* perm, float* film,
                                                             ShadowBuf sh, uint32_t* sh_count) {
    // visibility tests are staged in LDS and written out kConnStage at a
    // time, so the block takes its queue slots with few atomics
    __shared__ float4 s_stage[3 * kConnStage];  // ray halves [0, 2 kConnStage), contributions after
    __shared__ uint32_t s_tmp[kConnBlock / 64 + 1];
    __shared__ uint32_t s_fill, s_base;
    const uint32_t tile = tile0 + blockIdx.x;
    uint32_t lo = 0, hi = n_bins - 1;
    while (lo < hi) {
        const uint32_t mid = (lo + hi + 1) >> 1;
        if (bin_tile[mid] <= tile)
            lo = mid;
        else
            hi = mid - 1;
    }
    const int4 bn = bins[lo];
    const uint32_t first = (tile - bin_tile[lo]) * 64u;
    const uint32_t lane = threadIdx.x & 63u, wave = threadIdx.x >> 6;
    const bool live = first + lane < (uint32_t)bn.w;
    const uint32_t j = live ? perm[(uint32_t)bn.z + first + lane] : 0u;
    const int e = bn.x, l = bn.y;
    if (threadIdx.x == 0) s_fill = 0u;
    __syncthreads();
    // every thread calls these two at the same points
    auto drain = [&]() {
        const uint32_t n = s_fill;
        if (threadIdx.x == 0) s_base = atomicAdd(sh_count, n);
        __syncthreads();
        const size_t b = s_base;
        for (uint32_t k = threadIdx.x; k < 2 * n; k += kConnBlock) sh.ray[2 * b + k] = s_stage[k];
        for (uint32_t k = threadIdx.x; k < n; k += kConnBlock) sh.contrib[b + k] = s_stage[2 * kConnStage + k];
        __syncthreads();
        if (threadIdx.x == 0) s_fill = 0u;
        __syncthreads();
    };
    auto stage = [&](bool q, const ShadowEntry& se) {
        const uint64_t mask = __ballot(q);
        if (lane == 0) s_tmp[wave] = (uint32_t)__popcll(mask);
        __syncthreads();
        if (threadIdx.x == 0) {
            uint32_t tot = 0;
            for (int w = 0; w < kConnBlock / 64; ++w) {
                const uint32_t c = s_tmp[w];
                s_tmp[w] = tot;
                tot += c;
            }
            s_tmp[kConnBlock / 64] = tot;
        }
        __syncthreads();
        const uint32_t tot = s_tmp[kConnBlock / 64];
        if (s_fill + tot > kConnStage) drain();
        if (q) {
            const uint32_t k = s_fill + s_tmp[wave] + lane_rank(mask);
            s_stage[2 * k + 0] = se.o;
            s_stage[2 * k + 1] = se.d;
            s_stage[2 * kConnStage + k] = se.c;
        }
        __syncthreads();
        if (threadIdx.x == 0) s_fill += tot;
        __syncthreads();
    };
    // the block-uniform enumeration; (ms, mt) = this wave's strategy of the
    // current group of four
    uint32_t k = 0;
    int ms = -1, mt = 0;
    auto flush = [&]() {
        ShadowEntry se;
        const bool q = live && ms >= 0 && connect_one(bb, base, bp, sv, ms, mt, j, film, se);
        stage(q, se);
        ms = -1;
    };
    for (int t = 1; t <= e; ++t)
        for (int s = 0; s <= l && s + t - 1 <= bp.max_len; ++s) {
            if (s + t < 2 || (s == 1 && t == 1)) continue;
            if ((k & 3u) == wave) {
                ms = s;
                mt = t;
            }
            if ((k & 3u) == 3u) flush();
            ++k;
        }
    if (k & 3u) flush();
    if (s_fill > 0) drain();
}

inline uint32_t grid_of(uint64_t n) { return (uint32_t)((n + kBlock - 1) / kBlock); }

// A single scene's stack variant: 4 (tree in LDS) when the tree is shallow
// enough for variant 0 and fits kTreeF4.
inline int scene_stack_variant(const rumi_scene* sc) {
    const int v = stack_variant(sc->stack_need);
    const uint64_t f4 = (uint64_t)sc->tree_f4[0] + sc->tree_f4[1] + sc->tree_f4[2];
    return v == 0 && f4 <= kTreeF4 ? 4 : v;
}

uint32_t grid_for(uint64_t n, uint32_t cap_blocks) {
    uint64_t g = (n + kBlock - 1) / kBlock;
    if (g < 1) g = 1;
    return (uint32_t)(g < cap_blocks ? g : cap_blocks);
}

// Blocks of kernel K (kBlock threads) that stay resident on one compute
// unit.  The single-scene path tracer's persistent-lane stages get that
// many per CU and no more: a block beyond them would only start once others
// end, and its first draw on the work counter (same-address atomics
// serialise) would come at the tail (Cornell +5%, media +8%).  Batches and
// the bidirectional tracer keep 16 blocks per CU, which measured faster
// there (batch graph 536 against 561 ms, BDPT 146 against 140 Msamples/s).
template <auto K>
uint32_t resident_per_cu() {
    static const uint32_t per_cu = [] {
        int v = 0;
        if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&v, K, kBlock, 0) != hipSuccess || v < 1) v = 1;
        return (uint32_t)v;
    }();
    return per_cu;
}

// Clears the walk queue the next iteration appends to and k_extend's counter.
__global__ void k_walk_ctl(uint32_t* wctl, int next) {
    if (threadIdx.x == 0) wctl[next] = 0u;
    if (threadIdx.x == 1) wctl[2] = 0u;
}
SceneView scene_view(const rumi_scene* sc) {
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
    for (int k = 0; k < 3; ++k) sv.tree_f4[k] = sc->tree_f4[k];
    sv.med = sc->medium;
    sv.cam = sc->camera;
    return sv;
}


// One launch of k_bdpt_connect: tiles tile0 .. tile1 - 1 of the batch.
struct ConnChunk {
    uint32_t tile0, tile1;
    size_t conns;  // connections of the chunk's tiles (64 per tile and strategy)
};

// Plans a batch's connections from its (n_eye, n_light) counts (hist, index
// n_eye * (nL + 1) + n_light; strategies as k_bdpt_connect enumerates them).
// cursor[0 .. nbins) = the first position of each pair in the (n_eye,
// n_light) descending order (k_bdpt_scatter's cursors); bins = (n_eye,
// n_light, first position, count) of each non-empty pair in that order with
// its first tile in bin_tile; chunks of tiles whose connections (64 per
// tile and strategy) fit max_conn.
void conn_plan(const std::vector<uint32_t>& hist, int nE, int nL, int max_len, size_t max_conn,
               std::vector<uint32_t>& cursor, std::vector<int4>& bins, std::vector<uint32_t>& bin_tile,
               std::vector<ConnChunk>& chunks) {
    const uint32_t l1 = (uint32_t)nL + 1u, nbins = (uint32_t)(nE + 1) * l1;
    cursor.assign(nbins, 0u);
    bins.clear();
    bin_tile.clear();
    chunks.clear();
    uint32_t pos = 0, tiles = 0;
    size_t in_chunk = 0;
    ConnChunk cur{0, 0, 0};
    for (int e = nE; e >= 0; --e)
        for (int l = nL; l >= 0; --l) {
            const uint32_t cnt = hist[e * l1 + l];
            cursor[e * l1 + l] = pos;
            if (cnt > 0) {
                // the strategies k_bdpt_connect enumerates for this pair
                size_t n_s = 0;
                for (int t = 1; t <= e; ++t)
                    for (int s = 0; s <= l && s + t - 1 <= max_len; ++s)
                        if (s + t >= 2 && !(s == 1 && t == 1)) ++n_s;
                bins.push_back(make_int4(e, l, (int)pos, (int)cnt));
                bin_tile.push_back(tiles);
                const uint32_t nt = (cnt + 63u) / 64u;
                for (uint32_t k = 0; k < nt; ++k) {
                    if (in_chunk > 0 && in_chunk + 64 * n_s > max_conn) {
                        cur.tile1 = tiles + k;
                        cur.conns = in_chunk;
                        chunks.push_back(cur);
                        cur.tile0 = tiles + k;
                        in_chunk = 0;
                    }
                    in_chunk += 64 * n_s;
                }
                tiles += nt;
            }
            pos += cnt;
        }
    cur.tile1 = tiles;
    cur.conns = in_chunk;
    if (cur.tile1 > cur.tile0) chunks.push_back(cur);
}

int render_bdpt(rumi_ctx* ctx, rumi_scene* sc, const rumi_render_desc* desc, float* film,
                rumi_render_stats* stats, hipStream_t stream) {
    const int nE = desc->n_eye > 0 ? desc->n_eye : desc->max_depth + 1;
    const int nL = desc->n_light > 0 ? desc->n_light : desc->max_depth + 1;
    if (nE < 2 || nE > 64 || nL < 1 || nL > 64)
        return fail(RUMI_ERR_INVALID, "rumi_render: BDPT vertex caps must be in [2,64] (eye), [1,64] (light)");
    if (!(desc->mis_beta > 0.0f)) return fail(RUMI_ERR_INVALID, "rumi_render: mis_beta must be > 0");
    // strategies: all (s, t) with t >= 1, s + t >= 2, path edges
    // s + t - 1 <= max_len, except (1, 1), that the sample's subpaths reach
    // (k_bdpt_connect)
    const int max_len = desc->max_depth;
    const uint32_t pixels = (uint32_t)desc->width * (uint32_t)desc->height;
    const uint64_t total = (uint64_t)pixels * (uint64_t)desc->spp;
    // the connection work list: (n_eye, n_light) counts and cursors, the
    // samples in that order, the non-empty pairs and their first tiles
    const uint32_t l1 = (uint32_t)nL + 1u, nbins = (uint32_t)(nE + 1) * l1;
    // device bytes of a batch of b samples (the layout taken below)
    auto batch_bytes = [&](uint32_t b) {
        const size_t conn = std::max<size_t>(4 * (size_t)b, (size_t)1 << 25);
        return (sizeof(BVert) + sizeof(float4)) * (size_t)b * (size_t)(nE + nL) + 6 * (size_t)b +
               (4 + 4 + 16 + 4) * (size_t)nbins + 3 * sizeof(float4) * conn +
               2 * (size_t)b * (2 * 96 + 16) + 40 * 256;
    };
    uint32_t nb = desc->queue_capacity;
    if (!nb) {
        // 2^24 samples (about 35 GB at 9 + 9 vertices; the 1024^2 x 64 frame
        // renders at 170 Msamples/s with it, 150 with 2^22), or the largest
        // power of two down to 2^20 whose buffers fit in 90% of the free
        // device memory and what this scene already holds
        nb = total < (1u << 24) ? (uint32_t)total : (1u << 24);
        size_t free_b = 0, total_b = 0;
        if (hipMemGetInfo(&free_b, &total_b) == hipSuccess) {
            const size_t avail = (free_b + sc->work_bytes) / 10 * 9;
            while (nb > (1u << 20) && batch_bytes(nb) > avail) nb >>= 1;
        } else {
            (void)hipGetLastError();
            nb = 1u << 22;
        }
    }
    if ((uint64_t)nb > total) nb = (uint32_t)total;
    const size_t vbytes = sizeof(BVert) * (size_t)nb;
    const size_t mbytes = sizeof(float4) * (size_t)nb;
    // connections are made in chunks of at most max_conn (sample, strategy)
    // pairs; each chunk's visibility tests are queued (48 B each) and
    // resolved before the next.  Four per sample: a 2^26-sample batch with
    // 11.7 connections per sample takes three chunks (one per sample: 12,
    // 1% slower)
    const size_t max_conn = std::max<size_t>(4 * (size_t)nb, (size_t)1 << 25);
    // walks: eye and light walks of a batch in one double-buffered queue
    // (80 B per walk and buffer) with their closest hits (16 B)
    const size_t walks = 2 * (size_t)nb;
    const size_t need = batch_bytes(nb);
    if (sc->work_bytes < need) {
        if (sc->work) (void)hipFree(sc->work);
        sc->work = nullptr;
        sc->work_bytes = 0;
        if (hipMalloc(&sc->work, need) != hipSuccess)
            return fail(RUMI_ERR_NO_MEMORY, "rumi_render: hipMalloc of %zu BDPT bytes failed", need);
        sc->work_bytes = need;
    }
    char* w = (char*)sc->work;
    auto take = [&](size_t bytes) {
        char* p = w;
        w += (bytes + 255) & ~(size_t)255;
        return (void*)p;
    };
    BdptBufs bb;
    bb.eye = (BVert*)take(vbytes * (size_t)nE);
    bb.light = (BVert*)take(vbytes * (size_t)nL);
    bb.eye_mis = (float4*)take(mbytes * (size_t)nE);
    bb.light_mis = (float4*)take(mbytes * (size_t)nL);
    bb.n_eye = (uint8_t*)take(nb);
    bb.n_light = (uint8_t*)take(nb);
    uint32_t* perm = (uint32_t*)take(4 * (size_t)nb);
    uint32_t* d_hist = (uint32_t*)take(4 * (size_t)nbins);
    uint32_t* d_cursor = (uint32_t*)take(4 * (size_t)nbins);
    int4* d_bins = (int4*)take(16 * (size_t)nbins);
    uint32_t* d_bin_tile = (uint32_t*)take(4 * (size_t)nbins);
    ShadowBuf sh{};
    sh.ray = (float4*)take(2 * sizeof(float4) * max_conn);
    sh.contrib = (float4*)take(sizeof(float4) * max_conn);
    uint32_t* qctl = (uint32_t*)take(256);  // [0] queued connections, [1] k_shadow's fetch counter
    WalkBuf wb[2];
    for (int q = 0; q < 2; ++q) {
        wb[q].ray = (float4*)take(2 * sizeof(float4) * walks);
        wb[q].flags = (uint32_t*)take(4 * walks);
        wb[q].beta = (float4*)take(sizeof(float4) * walks);
        wb[q].prev = (float4*)take(sizeof(float4) * walks);
        wb[q].prevn = (float4*)take(sizeof(float4) * walks);
        wb[q].rng = (uint64_t*)take(8 * walks);
        wb[q].slot = (uint32_t*)take(4 * walks);
    }
    int2* whits = (int2*)take(8 * walks);
    float2* wuvs = (float2*)take(8 * walks);
    uint32_t* wctl = (uint32_t*)take(256);  // [0], [1] walk queue sizes, [2] k_extend's fetch counter
    unsigned long long* counts = (unsigned long long*)take(64);  // counting render: traversal totals
    const bool counting = (desc->flags & RUMI_RENDER_COUNT_TRAVERSAL) != 0;
    if ((size_t)(w - (char*)sc->work) > sc->work_bytes)
        return fail(RUMI_ERR_INVALID, "rumi_render: internal BDPT work-buffer layout overflow");

    SceneView sv = scene_view(sc);
    BdptParams bp{};
    bp.base.width = desc->width;
    bp.base.height = desc->height;
    bp.base.max_depth = desc->max_depth;
    bp.base.pixels = pixels;
    bp.base.seed = desc->seed;
    bp.n_eye = nE;
    bp.n_light = nL;
    bp.max_len = max_len;
    bp.beta = desc->mis_beta;
    bp.film_area = (2.0f * sc->camera.tanfov) * (2.0f * sc->camera.tanfov * sc->camera.y_ratio);
    bp.light_seed = desc->seed ^ 0x2545F4914F6CDD1Dull;
    const int stack = scene_stack_variant(sc);

    hipEvent_t ev0, ev1;
    (void)hipEventCreate(&ev0);
    (void)hipEventCreate(&ev1);
    (void)hipEventRecord(ev0, stream);
    const size_t film_n = 3 * (size_t)pixels;
    hipLaunchKernelGGL(k_zero, dim3((uint32_t)((film_n + 255) / 256)), dim3(256), 0, stream, film, film_n);
    if (counting) hipLaunchKernelGGL(k_zero_u32, dim3(1), dim3(64), 0, stream, (uint32_t*)counts, 12u);
    uint64_t batches = 0;
    int rc = RUMI_OK;
    std::vector<uint32_t> h_hist(nbins), h_cursor, h_bin_tile;
    std::vector<int4> h_bins;
    std::vector<ConnChunk> chunks;
    for (uint64_t base = 0; base < total; base += nb) {
        bb.nb = (uint32_t)std::min<uint64_t>(nb, total - base);
        // the walks as a wavefront: each iteration finds the closest hit of
        // every live walk with persistent lanes (k_extend), then takes its
        // segment (k_bdpt_walk); every iteration stores one vertex per live
        // walk, so max(nE, nL) - 1 iterations end them all
        const uint32_t nw = 2 * bb.nb;
        const uint32_t gw = std::min<uint32_t>(grid_of(nw), (uint32_t)ctx->cu_count * 16u);
        hipLaunchKernelGGL(k_zero_u32, dim3(1), dim3(64), 0, stream, wctl, 3u);
        hipLaunchKernelGGL(k_bdpt_init, dim3(grid_of(nw)), dim3(kBlock), 0, stream, bb, base, bp, sv, wb[0],
                           wctl);
        for (int it = 0, cur = 0; it < std::max(nE, nL) - 1; ++it, cur ^= 1) {
            hipLaunchKernelGGL(k_walk_ctl, dim3(1), dim3(64), 0, stream, wctl, cur ^ 1);
            PathBuf pb{};
            pb.ray = wb[cur].ray;
            pb.flags = wb[cur].flags;
            with_stack(stack, [&](auto v) {
                if (counting)
                    hipLaunchKernelGGL((k_extend<decltype(v)::value, false, true>), dim3(gw), dim3(kBlock), 0,
                                       stream, pb, wctl + cur, wctl + 2, sv, (const SceneTab*)nullptr, whits, wuvs,
                                       true, counts);
                else
                    hipLaunchKernelGGL((k_extend<decltype(v)::value, false>), dim3(gw), dim3(kBlock), 0, stream,
                                       pb, wctl + cur, wctl + 2, sv, (const SceneTab*)nullptr, whits, wuvs,
                                       true, (unsigned long long*)nullptr);
            });
            hipLaunchKernelGGL(k_bdpt_walk, dim3((gw + 1) / 2), dim3(kWalkBlock), 0, stream, bb, base, bp, sv, wb[cur],
                               (const uint32_t*)(wctl + cur), (const int2*)whits, (const float2*)wuvs,
                               wb[cur ^ 1], wctl + (cur ^ 1));
        }
        // the samples in (n_eye, n_light) order, and the chunks of tiles
        const uint32_t gb = std::min<uint32_t>(grid_of(bb.nb), 1024u);
        hipLaunchKernelGGL(k_zero_u32, dim3((nbins + 255) / 256), dim3(256), 0, stream, d_hist, nbins);
        hipLaunchKernelGGL(k_bdpt_hist, dim3(gb), dim3(kBlock), 0, stream, (const uint8_t*)bb.n_eye,
                           (const uint8_t*)bb.n_light, bb.nb, l1, nbins, d_hist);
        if (hipMemcpyAsync(h_hist.data(), d_hist, 4 * (size_t)nbins, hipMemcpyDeviceToHost, stream) !=
                hipSuccess ||
            hipStreamSynchronize(stream) != hipSuccess) {
            rc = fail(RUMI_ERR_HIP, "rumi_render: subpath length counts: %s", hipGetErrorString(hipGetLastError()));
            break;
        }
        conn_plan(h_hist, nE, nL, max_len, max_conn, h_cursor, h_bins, h_bin_tile, chunks);
        if (hipMemcpyAsync(d_cursor, h_cursor.data(), 4 * h_cursor.size(), hipMemcpyHostToDevice, stream) !=
                hipSuccess ||
            (!h_bins.empty() &&
             (hipMemcpyAsync(d_bins, h_bins.data(), sizeof(int4) * h_bins.size(), hipMemcpyHostToDevice,
                             stream) != hipSuccess ||
              hipMemcpyAsync(d_bin_tile, h_bin_tile.data(), 4 * h_bin_tile.size(), hipMemcpyHostToDevice,
                             stream) != hipSuccess))) {
            rc = fail(RUMI_ERR_HIP, "rumi_render: work-list upload failed");
            break;
        }
        hipLaunchKernelGGL(k_bdpt_scatter, dim3(gb), dim3(kBlock), 0, stream, (const uint8_t*)bb.n_eye,
                           (const uint8_t*)bb.n_light, bb.nb, l1, nbins, d_cursor, perm);
        for (const ConnChunk& ch : chunks) {
            const uint32_t tiles = ch.tile1 - ch.tile0;
            hipLaunchKernelGGL(k_zero_u32, dim3(1), dim3(64), 0, stream, qctl, 2u);
            hipLaunchKernelGGL(k_bdpt_connect, dim3(tiles), dim3(kConnBlock), 0, stream, bb, base, bp, sv,
                               (const int4*)d_bins, (const uint32_t*)d_bin_tile, (uint32_t)h_bins.size(), ch.tile0,
                               (const uint32_t*)perm, film, sh, qctl);
            const uint32_t g = (uint32_t)std::min<uint64_t>((ch.conns + kBlock - 1) / kBlock,
                                                            (uint64_t)ctx->cu_count * 16u);
            with_stack(stack, [&](auto v) {
                if (counting)
                    hipLaunchKernelGGL((k_shadow<decltype(v)::value, kRender, false, true>), dim3(g), dim3(kBlock),
                                       0, stream, sh, qctl, qctl + 1, sv, (const SceneTab*)nullptr, film, AdjCtx{},
                                       true, counts);
                else
                    hipLaunchKernelGGL((k_shadow<decltype(v)::value, kRender, false>), dim3(g), dim3(kBlock), 0,
                                       stream, sh, qctl, qctl + 1, sv, (const SceneTab*)nullptr, film, AdjCtx{},
                                       true, (unsigned long long*)nullptr);
            });
        }
        ++batches;
    }
    if (rc == RUMI_OK)
        hipLaunchKernelGGL(k_scale, dim3((uint32_t)((film_n + 255) / 256)), dim3(256), 0, stream, film, film_n,
                           1.0f / (float)desc->spp);
    (void)hipEventRecord(ev1, stream);
    if (hipStreamSynchronize(stream) != hipSuccess && rc == RUMI_OK)
        rc = fail(RUMI_ERR_HIP, "rumi_render: %s", hipGetErrorString(hipGetLastError()));
    else if (hipGetLastError() != hipSuccess && rc == RUMI_OK)
        rc = fail(RUMI_ERR_HIP, "rumi_render: a BDPT kernel launch failed");
    float ms = 0.0f;
    if (rc == RUMI_OK) (void)hipEventElapsedTime(&ms, ev0, ev1);
    (void)hipEventDestroy(ev0);
    (void)hipEventDestroy(ev1);
    if (stats) {
        stats->samples = total;
        stats->iterations = batches;
        stats->milliseconds = ms;
        // walk segments and connection rays are counted on the device only
        // by counting renders
        unsigned long long c[6] = {0, 0, 0, 0, 0, 0};
        if (rc == RUMI_OK && counting && hipMemcpy(c, counts, sizeof(c), hipMemcpyDeviceToHost) != hipSuccess)
            rc = fail(RUMI_ERR_HIP, "rumi_render: traversal count read failed");
        stats->extension_rays = c[4];
        stats->shadow_rays = c[5];
        stats->ext_nodes = c[0];
        stats->ext_prims = c[1];
        stats->shadow_nodes = c[2];
        stats->shadow_prims = c[3];
    }
    return rc;
}


// Shared checks of rumi_render and rumi_adjoint.
int check_render_args(rumi_ctx* ctx, const rumi_scene* sc, const rumi_render_desc* desc, const char* fn) {
    if (!ctx) return fail(RUMI_ERR_NO_DEVICE, "%s: null context", fn);
    if (!sc || !desc) return fail(RUMI_ERR_INVALID, "%s: null argument", fn);
    if (sc->device != ctx->device) return fail(RUMI_ERR_INVALID, "%s: scene is on another device", fn);
    if (desc->integrator != RUMI_INTEGRATOR_PT && desc->integrator != RUMI_INTEGRATOR_BDPT &&
        desc->integrator != RUMI_INTEGRATOR_WHITTED)
        return fail(RUMI_ERR_UNSUPPORTED, "%s: unknown integrator %d", fn, desc->integrator);
    if (desc->width <= 0 || desc->height <= 0 || desc->spp <= 0 || desc->max_depth <= 0)
        return fail(RUMI_ERR_INVALID, "%s: width, height, spp and max_depth must be > 0", fn);
    if (sc->has_point && desc->integrator == RUMI_INTEGRATOR_BDPT)
        return fail(RUMI_ERR_UNSUPPORTED, "%s: the bidirectional tracer takes area lights, not point lights", fn);
    if (sc->has_ibl && (desc->integrator == RUMI_INTEGRATOR_BDPT || sc->medium.bound))
        return fail(RUMI_ERR_UNSUPPORTED, "%s: the image-based light works with the path tracer and the "
                    "RayTracer, without a medium", fn);
    if (desc->max_depth > 1000) return fail(RUMI_ERR_INVALID, "%s: max_depth > 1000", fn);
    if ((uint64_t)desc->width * desc->height > 0xffffffffull / 4)
        return fail(RUMI_ERR_INVALID, "%s: image too large", fn);
    if (hipSetDevice(ctx->device) != hipSuccess) return fail(RUMI_ERR_HIP, "%s: hipSetDevice failed", fn);
    return RUMI_OK;
}

// ------------------------------------------------------------ wavefront host
// Device buffers of the wavefront path tracer, carved from one allocation.
struct Wave {
    PathBuf pb[2];
    int2* hits;
    float2* uvs;
    ShadowBuf sh;
    Ctl* ctl;
    unsigned long long* counts;  // counting renders: extension nodes, prims; shadow nodes, prims; rays
    uint32_t cap;
};

// Bytes per queued path: 2 path sets (ray 32, beta 16, rng 8, pixel, sample,
// flags, pend, scene 4 each), hit 8, barycentrics 8, shadow entry (ray 32,
// contrib 16, sample 4, score 8, texel 4, scene 4).
constexpr size_t kBytesPerPath = 2 * (32 + 16 + 8 + 4 * 5) + 8 + 8 + (32 + 16 + 4 + 8 + 4 + 4);
constexpr int kCarved = 26;  // arrays carved below, each aligned to 256 B

size_t wave_bytes(uint32_t cap) { return kBytesPerPath * (size_t)cap + sizeof(Ctl) + 64 + 256 * kCarved; }

bool wave_carve(Wave& w, void* mem, size_t bytes, uint32_t cap) {
    char* p = (char*)mem;
    auto take = [&](size_t n) {
        char* q = p;
        p += (n + 255) & ~(size_t)255;
        return (void*)q;
    };
    for (int k = 0; k < 2; ++k) {
        w.pb[k].ray = (float4*)take(32 * (size_t)cap);
        w.pb[k].beta = (float4*)take(16 * (size_t)cap);
        w.pb[k].rng = (uint64_t*)take(8 * (size_t)cap);
        w.pb[k].pixel = (uint32_t*)take(4 * (size_t)cap);
        w.pb[k].sample = (uint32_t*)take(4 * (size_t)cap);
        w.pb[k].flags = (uint32_t*)take(4 * (size_t)cap);
        w.pb[k].pend = (float*)take(4 * (size_t)cap);
        w.pb[k].scene = (uint32_t*)take(4 * (size_t)cap);
    }
    w.hits = (int2*)take(8 * (size_t)cap);
    w.uvs = (float2*)take(8 * (size_t)cap);
    w.sh.ray = (float4*)take(32 * (size_t)cap);
    w.sh.contrib = (float4*)take(16 * (size_t)cap);
    w.sh.sample = (uint32_t*)take(4 * (size_t)cap);
    w.sh.score = (float2*)take(8 * (size_t)cap);
    w.sh.tex = (int32_t*)take(4 * (size_t)cap);
    w.sh.scene = (uint32_t*)take(4 * (size_t)cap);
    w.ctl = (Ctl*)take(sizeof(Ctl));
    w.counts = (unsigned long long*)take(64);
    w.cap = cap;
    return (size_t)(p - (char*)mem) <= bytes;
}

// What the stages of one pass share.
struct Pass {
    SceneView sv;         // single scene
    const SceneTab* tab;  // batches
    Params prm;
    float* film;
    AdjCtx adj;
    Track tr;
    int stack;            // traversal stack variant (stack_variant)
    unsigned long long* counts;  // counting render: traversal totals, else null
    bool ibl;             // some scene holds an image-based light
    int mode;
    uint32_t grid_cap;    // most blocks a grid-stride stage is given
    uint32_t cu;          // compute units (persistent-lane grids: resident blocks per CU x cu)
};


template <int MODE, bool MULTI>
// The single-scene path tracer traces one ray per lane (lanes = false): its
// queues are compacted every iteration, so lanes rarely idle and the
// persistent lanes' bookkeeping costs more than it saves (the bidirectional
// tracer's walks and connections, of very uneven length, keep them).
void stages(const Wave& w, int cur, uint32_t bound, const Pass& ps, hipStream_t s) {
    Ctl* ctl = w.ctl;
    hipLaunchKernelGGL(k_clear_ctl, dim3(1), dim3(64), 0, s, ctl, cur ^ 1);
    with_stack(ps.stack, [&](auto v) {
        constexpr int SV = decltype(v)::value;
        if constexpr (MODE == kRender && !MULTI) {
            if (ps.counts) {
                const uint32_t g = grid_for(bound, resident_per_cu<k_extend<SV, false, true>>() * ps.cu);
                hipLaunchKernelGGL((k_extend<SV, false, true>), dim3(g), dim3(kBlock), 0, s, w.pb[cur],
                                   &ctl->queue[cur], &ctl->fetch[0], ps.sv, ps.tab, w.hits, w.uvs, false, ps.counts);
                return;
            }
        }
        const uint32_t g = grid_for(bound, MULTI ? ps.grid_cap : resident_per_cu<k_extend<SV, MULTI>>() * ps.cu);
        hipLaunchKernelGGL((k_extend<SV, MULTI>), dim3(g), dim3(kBlock), 0, s, w.pb[cur], &ctl->queue[cur],
                           &ctl->fetch[0], ps.sv, ps.tab, w.hits, w.uvs, false, (unsigned long long*)nullptr);
    });
    // one tile per block; blocks past the queue's size return at once
    const uint32_t gs = (uint32_t)((bound + kShadeBlock - 1) / kShadeBlock);
    if (ps.ibl)
        hipLaunchKernelGGL((k_shade<MODE, MULTI, true>), dim3(gs < 1 ? 1 : gs), dim3(kShadeBlock), 0, s,
                           w.pb[cur], w.pb[cur ^ 1], &ctl->queue[cur], w.hits, w.uvs, ps.sv, ps.tab, ps.prm,
                           ps.film, w.sh, &ctl->queue[cur ^ 1], &ctl->shadow, ps.adj, ps.tr);
    else
        hipLaunchKernelGGL((k_shade<MODE, MULTI, false>), dim3(gs < 1 ? 1 : gs), dim3(kShadeBlock), 0, s,
                           w.pb[cur], w.pb[cur ^ 1], &ctl->queue[cur], w.hits, w.uvs, ps.sv, ps.tab, ps.prm,
                           ps.film, w.sh, &ctl->queue[cur ^ 1], &ctl->shadow, ps.adj, ps.tr);
    with_stack(ps.stack, [&](auto v) {
        constexpr int SV = decltype(v)::value;
        if constexpr (MODE == kRender && !MULTI) {
            if (ps.counts) {
                const uint32_t g = grid_for(bound, resident_per_cu<k_shadow<SV, kRender, false, true>>() * ps.cu);
                hipLaunchKernelGGL((k_shadow<SV, kRender, false, true>), dim3(g), dim3(kBlock), 0, s, w.sh,
                                   &ctl->shadow, &ctl->fetch[1], ps.sv, ps.tab, ps.film, ps.adj, false, ps.counts);
                return;
            }
        }
        const uint32_t g = grid_for(bound, MULTI ? ps.grid_cap : resident_per_cu<k_shadow<SV, MODE, MULTI>>() * ps.cu);
        hipLaunchKernelGGL((k_shadow<SV, MODE, MULTI>), dim3(g), dim3(kBlock), 0, s, w.sh, &ctl->shadow,
                           &ctl->fetch[1], ps.sv, ps.tab, ps.film, ps.adj, false, (unsigned long long*)nullptr);
    });
}

// One wavefront iteration over queue `cur` (at most `bound` paths).
template <bool MULTI>
void iterate(const Wave& w, int cur, uint32_t bound, const Pass& ps, hipStream_t s) {
    if (ps.mode == kRender)
        stages<kRender, MULTI>(w, cur, bound, ps, s);
    else if (ps.mode == kPrimal)
        stages<kPrimal, MULTI>(w, cur, bound, ps, s);
    else
        stages<kReplay, MULTI>(w, cur, bound, ps, s);
}

template <bool MULTI>
void regen(const Wave& w, int cur, uint32_t start, uint32_t count, uint64_t sample_base,
           const Pass& ps, hipStream_t s) {
    hipLaunchKernelGGL((k_regen<MULTI>), dim3(grid_for(count, ps.grid_cap)), dim3(kBlock), 0, s,
                       w.pb[cur], start, count, sample_base, ps.prm, ps.sv.cam, ps.tab,
                       &w.ctl->queue[cur]);
}

// The wavefront path tracer: regeneration, extension, material buckets,
// shading, shadow rays, until every sample of the frame has terminated.
// MODE selects plain rendering, the adjoint's primal pass or its replay.
// The stages read their sizes from the device control block; the host reads
// the queue size after each iteration while samples remain to be issued (to
// decide on regeneration), then after batches of up to kDrainIters
// iterations that stop at the paths' depth limit (queues only shrink then,
// so the last size read bounds the launches, and stages whose queue has
// emptied end at once).
// `fixed` (or nullptr): a caller-provided work buffer of fixed_bytes that the
// queue must fit (the no-allocation entry point); the pass then allocates
// nothing -- no work buffer and no HIP events (its time is taken on the host,
// the call being synchronous).
// Iterations between host reads once every sample of the frame is issued
// (at most; fewer when the paths' depth limit says fewer remain: the
// RayTracer's 12 iterations go out in one batch, 4774 -> 5050 Msamples/s
// against batches of 8).
constexpr int kDrainIters = 24;

int render_pt(rumi_ctx* ctx, rumi_scene* sc, const rumi_render_desc* desc, float* film,
              rumi_render_stats* stats, hipStream_t stream, int mode, const AdjCtx& adj,
              void* fixed = nullptr, size_t fixed_bytes = 0) {
    const uint32_t pixels = (uint32_t)desc->width * (uint32_t)desc->height;
    const uint64_t total = (uint64_t)pixels * (uint64_t)desc->spp;
    // 2^23 paths (236 B each, about 2 GB) unless the caller says otherwise:
    // the 512^2 x 16 Cornell frame renders at 1722 Msamples/s with it and
    // 1493 with 2^21, the 512^2 x 32 media frame at 964 and 847.
    uint32_t cap = desc->queue_capacity ? desc->queue_capacity : (1u << 23);
    if ((uint64_t)cap > total) cap = (uint32_t)total;
    if (cap < 64) cap = 64;
    uint32_t regen_at = desc->regen_threshold ? desc->regen_threshold : cap;

    const size_t need = wave_bytes(cap);
    void* work = fixed;
    size_t work_bytes = fixed_bytes;
    if (!fixed) {
        if (sc->work_bytes < need) {
            if (sc->work) (void)hipFree(sc->work);
            sc->work = nullptr;
            sc->work_bytes = 0;
            if (hipMalloc(&sc->work, need) != hipSuccess)
                return fail(RUMI_ERR_NO_MEMORY, "rumi_render: hipMalloc of %zu work bytes failed", need);
            sc->work_bytes = need;
        }
        work = sc->work;
        work_bytes = sc->work_bytes;
    } else if (need > fixed_bytes) {
        return fail(RUMI_ERR_INVALID, "rumi_render: queue of %u paths needs %zu work bytes, %zu available",
                    cap, need, fixed_bytes);
    }
    Wave w;
    if (!wave_carve(w, work, work_bytes, cap))
        return fail(RUMI_ERR_INVALID, "rumi_render: internal work-buffer layout overflow");

    Pass ps{};
    ps.sv = scene_view(sc);
    ps.tab = nullptr;
    ps.prm.width = desc->width;
    ps.prm.height = desc->height;
    // RayTracer(scene, D) traces while trace_depth <= D: D + 1 surfaces
    const bool whitted = desc->integrator == RUMI_INTEGRATOR_WHITTED;
    ps.prm.max_depth = whitted ? desc->max_depth + 1 : desc->max_depth;
    ps.prm.pixels = pixels;
    ps.prm.seed = desc->seed;
    ps.prm.n_scenes = 1;
    ps.prm.whitted = whitted ? 1 : 0;
    ps.film = film;
    ps.adj = adj;
    ps.tr = Track{nullptr, nullptr};
    ps.stack = scene_stack_variant(sc);
    const bool counting = (desc->flags & RUMI_RENDER_COUNT_TRAVERSAL) != 0 && mode == kRender;
    ps.counts = counting ? w.counts : nullptr;
    ps.ibl = sc->has_ibl;
    ps.mode = mode;
    ps.grid_cap = (uint32_t)ctx->cu_count * 16u;
    ps.cu = (uint32_t)ctx->cu_count;

    const bool events = fixed == nullptr;
    hipEvent_t ev0 = nullptr, ev1 = nullptr;
    const auto host_t0 = std::chrono::steady_clock::now();
    if (events) {
        (void)hipEventCreate(&ev0);
        (void)hipEventCreate(&ev1);
        (void)hipEventRecord(ev0, stream);
    }
    const size_t film_n = 3 * (size_t)pixels;
    if (mode != kReplay)
        hipLaunchKernelGGL(k_zero, dim3((uint32_t)((film_n + 255) / 256)), dim3(256), 0, stream, film, film_n);
    if (counting) hipLaunchKernelGGL(k_zero_u32, dim3(1), dim3(64), 0, stream, (uint32_t*)w.counts, 12u);

    uint64_t issued = 0, ext_rays = 0, shadow_rays = 0, iters = 0;
    uint32_t n_cur = 0;
    int cur = 0, since_regen = 0;
    Ctl host_ctl{};
    int rc = RUMI_OK;
    hipLaunchKernelGGL(k_zero_u32, dim3(1), dim3(64), 0, stream, (uint32_t*)w.ctl->total, 6u);
    while (true) {
        if (issued < total && n_cur < cap && n_cur <= regen_at) {
            const uint32_t cnt = (uint32_t)std::min<uint64_t>(cap - n_cur, total - issued);
            regen<false>(w, cur, n_cur, cnt, issued, ps, stream);
            n_cur += cnt;
            issued += cnt;
            since_regen = 0;
        }
        if (n_cur == 0) break;
        // a path ends within max_depth + 1 iterations of its start (past
        // that bound, one iteration per read again)
        const int left = ps.prm.max_depth + 1 - since_regen;
        const int batch = issued >= total ? std::min(kDrainIters, std::max(1, left)) : 1;
        since_regen += batch;
        for (int k = 0; k < batch; ++k) {
            iterate<false>(w, cur, n_cur, ps, stream);
            cur ^= 1;
        }
        if (hipMemcpyAsync(&host_ctl, w.ctl, sizeof(Ctl), hipMemcpyDeviceToHost, stream) != hipSuccess ||
            hipStreamSynchronize(stream) != hipSuccess) {
            rc = fail(RUMI_ERR_HIP, "rumi_render: counter read failed: %s",
                      hipGetErrorString(hipGetLastError()));
            break;
        }
        n_cur = host_ctl.queue[cur];
    }
    if (rc == RUMI_OK) {
        // the last iteration's shadow entries were not added up yet
        ext_rays = host_ctl.total[0];
        shadow_rays = host_ctl.total[1] + host_ctl.shadow;
        iters = host_ctl.total[2];
    }
    if (rc == RUMI_OK) {
        if (mode != kReplay)
            hipLaunchKernelGGL(k_scale, dim3((uint32_t)((film_n + 255) / 256)), dim3(256), 0, stream, film,
                               film_n, 1.0f / (float)desc->spp);
        if (events) (void)hipEventRecord(ev1, stream);
        if (hipStreamSynchronize(stream) != hipSuccess)
            rc = fail(RUMI_ERR_HIP, "rumi_render: %s", hipGetErrorString(hipGetLastError()));
        else if (hipGetLastError() != hipSuccess)
            rc = fail(RUMI_ERR_HIP, "rumi_render: a kernel launch failed");
    }
    float ms = 0.0f;
    if (events) {
        if (rc == RUMI_OK) (void)hipEventElapsedTime(&ms, ev0, ev1);
        (void)hipEventDestroy(ev0);
        (void)hipEventDestroy(ev1);
    } else {
        ms = std::chrono::duration<float, std::milli>(std::chrono::steady_clock::now() - host_t0).count();
    }
    if (stats) {
        stats->samples = total;
        stats->extension_rays = ext_rays;
        stats->shadow_rays = shadow_rays;
        stats->iterations = iters;
        stats->milliseconds = ms;
        unsigned long long c[6] = {0, 0, 0, 0, 0, 0};
        if (rc == RUMI_OK && counting && hipMemcpy(c, w.counts, sizeof(c), hipMemcpyDeviceToHost) != hipSuccess)
            rc = fail(RUMI_ERR_HIP, "rumi_render: traversal count read failed");
        stats->ext_nodes = c[0];
        stats->ext_prims = c[1];
        stats->shadow_nodes = c[2];
        stats->shadow_prims = c[3];
    }
    return rc;
}

__global__ void k_reduce_stripes(const float* stripes, float* out) {
    // one wave: lane k < 7 sums parameter k over the stripes in a fixed order
    const int k = threadIdx.x;
    if (k < 7) {
        float a = 0.0f;
        for (int j = 0; j < kStripes; ++j) a += stripes[8 * j + k];
        out[k] = a;
    }
}

// ============================================================ batches
// Many scenes in one wavefront: samples are numbered scene after scene, every
// path carries its scene, and the stages index a device table of the scenes'
// BVHs, materials, lights, textures and cameras.  The schedule is fixed when
// the batch is built -- generations of at most `cap` samples, each drained by
// exactly max_depth iterations (every path ends by then) -- so one run is a
// straight sequence of launches with no host round trip, which is captured
// once as a hipGraph and replayed.

__global__ void k_batch_begin(uint32_t* left, const uint32_t* samples, uint64_t* done, uint32_t n) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) {
        left[i] = samples[i];
        done[i] = 0ull;
    }
    if (i == 0) done[n] = wall_clock64();
}

// dLoss/dImage of the L2 loss sum (I - T)^2, and the loss per item: grid
// (blocks, items), each block striding over its item's pixels, one atomic per
// wave.
__global__ void k_l2_grad(const float* img, const float* target, float* dimg,
                          const uint32_t* pixel_base, uint32_t n_items, uint64_t total_pixels,
                          float* loss) {
    const uint32_t item = blockIdx.y;
    const size_t lo = 3 * (size_t)pixel_base[item];
    const size_t hi = 3 * (item + 1 < n_items ? (size_t)pixel_base[item + 1] : (size_t)total_pixels);
    float acc = 0.0f;
    for (size_t i = lo + (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < hi;
         i += (size_t)gridDim.x * blockDim.x) {
        const float d = img[i] - target[i];
        dimg[i] = 2.0f * d;
        acc += d * d;
    }
    if (loss) {
        for (int off = 32; off > 0; off >>= 1) acc += __shfl_xor(acc, off);
        if ((threadIdx.x & 63u) == 0 && acc != 0.0f) atomicAdd(loss + item, acc);
    }
}

}  // namespace

struct rumi_batch {
    int device = -1;
    int cu_count = 0;
    uint32_t n = 0;
    rumi_batch_desc desc{};
    uint32_t cap = 0, generations = 0;
    uint64_t total_samples = 0, total_pixels = 0, total_texels = 0;
    int stack = 0;        // traversal stack variant of the deepest scene
    bool ibl = false;  // some scene holds an image-based light
    std::vector<uint32_t> item_samples, item_pixels;
    // device
    void* mem = nullptr;
    size_t mem_bytes = 0;
    SceneTab* tab = nullptr;
    uint32_t* samples = nullptr;     // per item
    uint32_t* pixel_base = nullptr;  // per item
    uint32_t* left = nullptr;        // per item
    uint64_t* done = nullptr;        // per item + start stamp
    float* scratch_film = nullptr;   // the adjoint's primal image
    float* dimg = nullptr;
    float* L_total = nullptr;
    float* L_sofar = nullptr;
    float* stripes = nullptr;
    Wave w{};
    // captured graph and the io it was captured with
    hipGraphExec_t exec = nullptr;
    hipGraph_t graph = nullptr;
    rumi_batch_io graph_io{};
    double clock_khz = 0.0;
};

namespace {

// Every launch of one batch run, in stream order (eagerly or under capture).
void batch_enqueue(rumi_batch* b, const rumi_batch_io& io, hipStream_t s) {
    const uint32_t n = b->n;
    const size_t film_n = 3 * (size_t)b->total_pixels;
    auto zero = [&](float* p, size_t cnt) {
        if (!p || cnt == 0) return;
        const size_t blocks = (cnt + 255) / 256;
        const size_t cap_blocks = (size_t)b->cu_count * 64;
        hipLaunchKernelGGL(k_zero, dim3((uint32_t)(blocks < cap_blocks ? blocks : cap_blocks)), dim3(256), 0,
                           s, p, cnt);
    };
    hipLaunchKernelGGL(k_batch_begin, dim3((n + 255) / 256), dim3(256), 0, s, b->left, b->samples, b->done, n);
    zero(io.dev_images, film_n);

    Pass ps{};
    ps.sv = SceneView{};
    ps.tab = b->tab;
    ps.prm.width = 0;
    ps.prm.height = 0;
    ps.prm.max_depth = b->desc.max_depth;
    ps.prm.pixels = 0;
    ps.prm.n_scenes = n;
    ps.stack = b->stack;
    ps.ibl = b->ibl;
    ps.grid_cap = (uint32_t)b->cu_count * 16u;
    ps.cu = (uint32_t)b->cu_count;
    ps.adj = AdjCtx{};
    ps.adj.gstripes = b->stripes;
    ps.adj.inv_spp = 1.0f / (float)b->desc.spp;

    auto pass = [&](int mode, uint64_t seed, float* film, Track tr) {
        ps.mode = mode;
        ps.prm.seed = seed;
        ps.film = film;
        ps.tr = tr;
        for (uint32_t g = 0; g < b->generations; ++g) {
            const uint64_t first = (uint64_t)g * b->cap;
            const uint32_t cnt = (uint32_t)std::min<uint64_t>(b->cap, b->total_samples - first);
            regen<true>(b->w, 0, 0, cnt, first, ps, s);
            int cur = 0;
            for (int d = 0; d < b->desc.max_depth; ++d) {
                iterate<true>(b->w, cur, cnt, ps, s);
                cur ^= 1;
            }
        }
    };
    pass(kRender, b->desc.seed, io.dev_images, Track{b->left, b->done});
    hipLaunchKernelGGL(k_scale, dim3((uint32_t)((film_n + 255) / 256)), dim3(256), 0, s, io.dev_images,
                       film_n, 1.0f / (float)b->desc.spp);
    if (b->desc.adjoint) {
        if (io.dev_loss) zero(io.dev_loss, n);
        hipLaunchKernelGGL(k_l2_grad, dim3(64, n), dim3(256), 0, s, io.dev_images, io.dev_targets,
                           b->dimg, b->pixel_base, n, b->total_pixels, io.dev_loss);
        zero(b->L_total, 3 * (size_t)b->total_samples);
        zero(b->L_sofar, 3 * (size_t)b->total_samples);
        zero(b->stripes, 8 * kStripes);
        zero(io.dev_texel_grads, 3 * (size_t)b->total_texels);
        ps.adj.L_total = b->L_total;
        ps.adj.L_sofar = b->L_sofar;
        ps.adj.dimg = b->dimg;
        ps.adj.tex_grad = io.dev_texel_grads;
        zero(b->scratch_film, film_n);
        pass(kPrimal, b->desc.seed + 1, b->scratch_film, Track{nullptr, nullptr});
        pass(kReplay, b->desc.seed + 1, nullptr, Track{nullptr, nullptr});
    }
}

void batch_free(rumi_batch* b) {
    if (b->device >= 0) {
        int prev = 0;
        (void)hipGetDevice(&prev);
        (void)hipSetDevice(b->device);
        if (b->exec) (void)hipGraphExecDestroy(b->exec);
        if (b->graph) (void)hipGraphDestroy(b->graph);
        if (b->mem) (void)hipFree(b->mem);
        (void)hipSetDevice(prev);
    }
    b->exec = nullptr;
    b->graph = nullptr;
    b->mem = nullptr;
}

}  // namespace

extern "C" int rumi_render(rumi_ctx* ctx, const rumi_scene* scene_c, const rumi_render_desc* desc,
                           float* film, rumi_render_stats* stats, void* stream_v) {
    clear_error();
    rumi_scene* sc = const_cast<rumi_scene*>(scene_c);
    int rc = check_render_args(ctx, sc, desc, "rumi_render");
    if (rc != RUMI_OK) return rc;
    if (!film) return fail(RUMI_ERR_INVALID, "rumi_render: null film");
    hipStream_t stream = (hipStream_t)stream_v;
    if (desc->integrator == RUMI_INTEGRATOR_BDPT) return render_bdpt(ctx, sc, desc, film, stats, stream);
    AdjCtx none{};
    return render_pt(ctx, sc, desc, film, stats, stream, kRender, none);
}

// ------------------------------------------------------------ no allocation
// The work buffer of rumi_render_noalloc: one statically sized array per
// GPU, resident from the library's load until the process ends.  Its
// device address comes from hipGetSymbolAddress once per device (the
// array's name is only a host-side stand-in); every call on a device shares
// the array, so calls that overlap on one device race.
namespace {
constexpr uint32_t kNoAllocPaths = RUMI_NOALLOC_MAX_PATHS;
constexpr size_t kScratchBytes = kBytesPerPath * (size_t)kNoAllocPaths + sizeof(Ctl) + 64 + 256 * kCarved;
static __device__ __align__(256) unsigned char scratch[kScratchBytes];

constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
void* g_scratch_ptr[kMaxDevices] = {};

// The calling thread's current device must be `device`.
int scratch_for(int device, void** out) {
    if (device < 0 || device >= kMaxDevices)
        return fail(RUMI_ERR_INVALID, "rumi_render_noalloc: device %d out of range", device);
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (!g_scratch_ptr[device]) {
        void* ptr = nullptr;
        const hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !ptr)
            return fail(RUMI_ERR_HIP, "rumi_render_noalloc: hipGetSymbolAddress(scratch) failed: %s",
                        hipGetErrorString(e));
        g_scratch_ptr[device] = ptr;
    }
    *out = g_scratch_ptr[device];
    return RUMI_OK;
}
}  // namespace

extern "C" int rumi_render_noalloc(rumi_ctx* ctx, const rumi_scene* scene_c, const rumi_render_desc* desc,
                                   float* film, rumi_render_stats* stats, void* stream_v) {
    clear_error();
    rumi_scene* sc = const_cast<rumi_scene*>(scene_c);
    int rc = check_render_args(ctx, sc, desc, "rumi_render_noalloc");
    if (rc != RUMI_OK) return rc;
    if (!film) return fail(RUMI_ERR_INVALID, "rumi_render_noalloc: null film");
    if (desc->integrator == RUMI_INTEGRATOR_BDPT)
        return fail(RUMI_ERR_UNSUPPORTED,
                    "rumi_render_noalloc: the path tracer and the RayTracer only, not the bidirectional tracer");
    if (desc->queue_capacity > kNoAllocPaths)
        return fail(RUMI_ERR_INVALID, "rumi_render_noalloc: queue_capacity %u exceeds the maximum of %u paths",
                    desc->queue_capacity, kNoAllocPaths);
    void* work = nullptr;
    if ((rc = scratch_for(ctx->device, &work)) != RUMI_OK) return rc;
    rumi_render_desc d = *desc;
    if (d.queue_capacity == 0) d.queue_capacity = kNoAllocPaths;
    AdjCtx none{};
    return render_pt(ctx, sc, &d, film, stats, (hipStream_t)stream_v, kRender, none, work, kScratchBytes);
}

extern "C" int rumi_adjoint(rumi_ctx* ctx, const rumi_scene* scene_c, const rumi_render_desc* desc,
                            const float* dloss_dimage, const rumi_grad_buffers* grads,
                            rumi_render_stats* stats, void* stream_v) {
    clear_error();
    rumi_scene* sc = const_cast<rumi_scene*>(scene_c);
    int rc = check_render_args(ctx, sc, desc, "rumi_adjoint");
    if (rc != RUMI_OK) return rc;
    if (!dloss_dimage || !grads) return fail(RUMI_ERR_INVALID, "rumi_adjoint: null argument");
    if ((uint64_t)desc->width * desc->height * (uint64_t)desc->spp > 0xffffffffull)
        return fail(RUMI_ERR_INVALID, "rumi_adjoint: more than 2^32 - 1 samples in one frame");
    if (desc->integrator != RUMI_INTEGRATOR_PT)
        return fail(RUMI_ERR_UNSUPPORTED, "rumi_adjoint: only the path tracer is differentiable");
    if (grads->d_textures && sc->texel_count == 0)
        return fail(RUMI_ERR_INVALID, "rumi_adjoint: texture gradients asked of a scene without textures");
    hipStream_t stream = (hipStream_t)stream_v;
    const uint32_t pixels = (uint32_t)desc->width * (uint32_t)desc->height;
    const uint64_t total = (uint64_t)pixels * (uint64_t)desc->spp;

    // Per-sample radiance of the primal pass and of the replay so far,
    // the gradient stripes, and the primal image when none is asked for.
    const size_t l_bytes = 12 * (size_t)total;
    const size_t need = 2 * ((l_bytes + 255) & ~(size_t)255) + 8 * kStripes * sizeof(float) + 256 +
                        (grads->d_image ? 0 : ((12 * (size_t)pixels + 255) & ~(size_t)255));
    if (sc->adj_bytes < need) {
        if (sc->adj) (void)hipFree(sc->adj);
        sc->adj = nullptr;
        sc->adj_bytes = 0;
        if (hipMalloc(&sc->adj, need) != hipSuccess)
            return fail(RUMI_ERR_NO_MEMORY, "rumi_adjoint: hipMalloc of %zu bytes failed", need);
        sc->adj_bytes = need;
    }
    char* w = (char*)sc->adj;
    auto take = [&](size_t bytes) {
        char* p = w;
        w += (bytes + 255) & ~(size_t)255;
        return (void*)p;
    };
    AdjCtx adj;
    adj.L_total = (float*)take(l_bytes);
    adj.L_sofar = (float*)take(l_bytes);
    adj.gstripes = (float*)take(8 * kStripes * sizeof(float));
    float* image = grads->d_image ? grads->d_image : (float*)take(12 * (size_t)pixels);
    adj.dimg = dloss_dimage;
    adj.inv_spp = 1.0f / (float)desc->spp;
    adj.tex_grad = grads->d_textures;

    hipLaunchKernelGGL(k_zero, dim3((uint32_t)((3 * total + 255) / 256)), dim3(256), 0, stream, adj.L_total,
                       3 * (size_t)total);
    hipLaunchKernelGGL(k_zero, dim3((uint32_t)((3 * total + 255) / 256)), dim3(256), 0, stream, adj.L_sofar,
                       3 * (size_t)total);
    hipLaunchKernelGGL(k_zero, dim3(2), dim3(256), 0, stream, adj.gstripes, (size_t)(8 * kStripes));
    if (grads->d_textures)
        hipLaunchKernelGGL(k_zero, dim3((uint32_t)((3 * sc->texel_count + 255) / 256)), dim3(256), 0, stream,
                           grads->d_textures, 3 * (size_t)sc->texel_count);
    rumi_render_stats st0{}, st1{};
    rc = render_pt(ctx, sc, desc, image, &st0, stream, kPrimal, adj);
    if (rc == RUMI_OK) rc = render_pt(ctx, sc, desc, nullptr, &st1, stream, kReplay, adj);
    if (rc == RUMI_OK && grads->d_medium) {
        hipLaunchKernelGGL(k_reduce_stripes, dim3(1), dim3(64), 0, stream, adj.gstripes, grads->d_medium);
        if (hipStreamSynchronize(stream) != hipSuccess)
            rc = fail(RUMI_ERR_HIP, "rumi_adjoint: %s", hipGetErrorString(hipGetLastError()));
    }
    if (stats) {
        stats->samples = total;
        stats->extension_rays = st0.extension_rays + st1.extension_rays;
        stats->shadow_rays = st0.shadow_rays + st1.shadow_rays;
        stats->iterations = st0.iterations + st1.iterations;
        stats->milliseconds = st0.milliseconds + st1.milliseconds;
    }
    return rc;
}

extern "C" int rumi_batch_create(rumi_ctx* ctx, const rumi_batch_item* items, uint32_t n,
                                 const rumi_batch_desc* desc, rumi_batch** out) {
    clear_error();
    if (!out) return fail(RUMI_ERR_INVALID, "rumi_batch_create: null output");
    *out = nullptr;
    if (!ctx) return fail(RUMI_ERR_NO_DEVICE, "rumi_batch_create: null context");
    if (!items || n == 0 || !desc) return fail(RUMI_ERR_INVALID, "rumi_batch_create: no items");
    if (desc->spp <= 0 || desc->max_depth <= 0 || desc->max_depth > 1000)
        return fail(RUMI_ERR_INVALID, "rumi_batch_create: spp > 0 and max_depth in [1, 1000] required");
    if (hipSetDevice(ctx->device) != hipSuccess) return fail(RUMI_ERR_HIP, "rumi_batch_create: hipSetDevice failed");
    rumi_batch* b = new (std::nothrow) rumi_batch;
    if (!b) return fail(RUMI_ERR_NO_MEMORY, "rumi_batch_create: out of host memory");
    b->n = n;
    b->desc = *desc;
    b->cu_count = ctx->cu_count;
    std::vector<SceneTab> tab(n);
    std::vector<uint32_t> pbase(n);
    uint64_t samples = 0, pixels = 0, texels = 0;
    for (uint32_t i = 0; i < n; ++i) {
        const rumi_batch_item& it = items[i];
        const rumi_scene* sc = it.scene;
        if (!sc || sc->device != ctx->device) {
            delete b;
            return fail(RUMI_ERR_INVALID, "rumi_batch_create: item %u has no scene on the context's device", i);
        }
        if (it.width <= 0 || it.height <= 0) {
            delete b;
            return fail(RUMI_ERR_INVALID, "rumi_batch_create: item %u has an empty image", i);
        }
        if (sc->has_ibl && sc->medium.bound) {
            delete b;
            return fail(RUMI_ERR_UNSUPPORTED, "rumi_batch_create: item %u: image-based light with a medium", i);
        }
        b->ibl = b->ibl || sc->has_ibl;
        SceneTab& t = tab[i];
        t.sv = scene_view(sc);
        t.width = it.width;
        t.height = it.height;
        t.spp = desc->spp;
        t.pixels = (uint32_t)it.width * (uint32_t)it.height;
        t.sample_base = (uint32_t)samples;
        t.pixel_base = (uint32_t)pixels;
        t.texel_base = (uint32_t)texels;
        t.pad = 0;
        pbase[i] = (uint32_t)pixels;
        b->item_samples.push_back(t.pixels * (uint32_t)desc->spp);
        b->item_pixels.push_back(t.pixels);
        samples += (uint64_t)t.pixels * (uint64_t)desc->spp;
        pixels += t.pixels;
        texels += sc->texel_count;
        b->stack = std::max(b->stack, stack_variant(sc->stack_need));
        if (samples > 0xffffffffull || pixels > 0xffffffffull / 4) {
            delete b;
            return fail(RUMI_ERR_INVALID, "rumi_batch_create: more than 2^32 - 1 samples in the batch");
        }
    }
    b->total_samples = samples;
    b->total_pixels = pixels;
    b->total_texels = texels;
    // one allocation: table, per-item arrays, wave buffers, adjoint buffers
    auto al = [](size_t x) { return (x + 255) & ~(size_t)255; };
    size_t fixed = al(sizeof(SceneTab) * n) + 3 * al(4 * (size_t)n) + al(8 * ((size_t)n + 1)) +
                   al(4 * 8 * kStripes);
    if (desc->adjoint)
        fixed += 2 * al(12 * (size_t)pixels) + 2 * al(12 * (size_t)samples);
    uint32_t cap = desc->queue_capacity;
    if (!cap) {
        // 2^26 paths per generation (about 16 GB; the 256-scene forward +
        // adjoint graph takes 470 ms with it, 867 with 2^22), or the largest
        // power of two down to 2^22 that fits in 90% of the free device memory
        cap = samples < (1u << 26) ? (uint32_t)samples : (1u << 26);
        size_t free_b = 0, total_b = 0;
        if (hipMemGetInfo(&free_b, &total_b) == hipSuccess) {
            const size_t avail = free_b / 10 * 9;
            while (cap > (1u << 22) && fixed + al(wave_bytes(cap)) > avail) cap >>= 1;
        } else {
            (void)hipGetLastError();
            cap = 1u << 22;
        }
    }
    if ((uint64_t)cap > samples) cap = (uint32_t)samples;
    if (cap < 64) cap = 64;
    b->cap = cap;
    b->generations = (uint32_t)((samples + cap - 1) / cap);
    const size_t wbytes = wave_bytes(cap);
    const size_t bytes = fixed + al(wbytes);
    if (hipMalloc(&b->mem, bytes) != hipSuccess) {
        delete b;
        return fail(RUMI_ERR_NO_MEMORY, "rumi_batch_create: hipMalloc of %zu bytes failed", bytes);
    }
    b->device = ctx->device;
    b->mem_bytes = bytes;
    char* p = (char*)b->mem;
    auto take = [&](size_t x) {
        char* q = p;
        p += al(x);
        return (void*)q;
    };
    b->tab = (SceneTab*)take(sizeof(SceneTab) * n);
    b->samples = (uint32_t*)take(4 * (size_t)n);
    b->pixel_base = (uint32_t*)take(4 * (size_t)n);
    b->left = (uint32_t*)take(4 * (size_t)n);
    b->done = (uint64_t*)take(8 * ((size_t)n + 1));
    void* wmem = take(wbytes);
    b->stripes = (float*)take(4 * 8 * kStripes);
    if (desc->adjoint) {
        b->scratch_film = (float*)take(12 * (size_t)pixels);
        b->dimg = (float*)take(12 * (size_t)pixels);
        b->L_total = (float*)take(12 * (size_t)samples);
        b->L_sofar = (float*)take(12 * (size_t)samples);
    }
    if ((size_t)(p - (char*)b->mem) > bytes || !wave_carve(b->w, wmem, wbytes, cap)) {
        batch_free(b);
        delete b;
        return fail(RUMI_ERR_INVALID, "rumi_batch_create: internal layout overflow");
    }
    if (hipMemcpy(b->tab, tab.data(), sizeof(SceneTab) * n, hipMemcpyHostToDevice) != hipSuccess ||
        hipMemcpy(b->samples, b->item_samples.data(), 4 * (size_t)n, hipMemcpyHostToDevice) != hipSuccess ||
        hipMemcpy(b->pixel_base, pbase.data(), 4 * (size_t)n, hipMemcpyHostToDevice) != hipSuccess) {
        batch_free(b);
        delete b;
        return fail(RUMI_ERR_HIP, "rumi_batch_create: table upload failed");
    }
    int khz = 0;
    if (hipDeviceGetAttribute(&khz, hipDeviceAttributeWallClockRate, ctx->device) != hipSuccess || khz <= 0)
        khz = 100000;
    b->clock_khz = (double)khz;
    *out = b;
    return RUMI_OK;
}

extern "C" int rumi_batch_run(rumi_batch* b, const rumi_batch_io* io, int32_t use_graph,
                              rumi_batch_stats* stats, void* stream_v) {
    clear_error();
    if (!b || !io || !io->dev_images) return fail(RUMI_ERR_INVALID, "rumi_batch_run: null argument");
    if (b->desc.adjoint && (!io->dev_targets || (b->total_texels && !io->dev_texel_grads)))
        return fail(RUMI_ERR_INVALID, "rumi_batch_run: the adjoint needs targets and a texel-gradient buffer");
    if (hipSetDevice(b->device) != hipSuccess) return fail(RUMI_ERR_HIP, "rumi_batch_run: hipSetDevice failed");
    hipStream_t s = (hipStream_t)stream_v;
    hipEvent_t ev0, ev1;
    (void)hipEventCreate(&ev0);
    (void)hipEventCreate(&ev1);
    (void)hipEventRecord(ev0, s);
    int rc = RUMI_OK;
    if (use_graph) {
        const bool same = b->exec && b->graph_io.dev_images == io->dev_images &&
                          b->graph_io.dev_targets == io->dev_targets &&
                          b->graph_io.dev_texel_grads == io->dev_texel_grads &&
                          b->graph_io.dev_loss == io->dev_loss;
        if (!same) {
            if (b->exec) (void)hipGraphExecDestroy(b->exec);
            if (b->graph) (void)hipGraphDestroy(b->graph);
            b->exec = nullptr;
            b->graph = nullptr;
            hipStream_t cs;
            if (hipStreamCreateWithFlags(&cs, hipStreamNonBlocking) != hipSuccess)
                rc = fail(RUMI_ERR_HIP, "rumi_batch_run: capture stream creation failed");
            if (rc == RUMI_OK && hipStreamBeginCapture(cs, hipStreamCaptureModeThreadLocal) != hipSuccess)
                rc = fail(RUMI_ERR_HIP, "rumi_batch_run: hipStreamBeginCapture failed");
            if (rc == RUMI_OK) {
                batch_enqueue(b, *io, cs);
                if (hipStreamEndCapture(cs, &b->graph) != hipSuccess || !b->graph)
                    rc = fail(RUMI_ERR_HIP, "rumi_batch_run: capture failed: %s", hipGetErrorString(hipGetLastError()));
                else if (hipGraphInstantiate(&b->exec, b->graph, nullptr, nullptr, 0) != hipSuccess)
                    rc = fail(RUMI_ERR_HIP, "rumi_batch_run: hipGraphInstantiate failed");
            }
            (void)hipStreamDestroy(cs);
            if (rc == RUMI_OK) b->graph_io = *io;
        }
        if (rc == RUMI_OK && hipGraphLaunch(b->exec, s) != hipSuccess)
            rc = fail(RUMI_ERR_HIP, "rumi_batch_run: hipGraphLaunch failed");
    } else {
        batch_enqueue(b, *io, s);
    }
    (void)hipEventRecord(ev1, s);
    if (hipStreamSynchronize(s) != hipSuccess)
        rc = fail(RUMI_ERR_HIP, "rumi_batch_run: %s", hipGetErrorString(hipGetLastError()));
    else if (rc == RUMI_OK && hipGetLastError() != hipSuccess)
        rc = fail(RUMI_ERR_HIP, "rumi_batch_run: a kernel launch failed");
    float ms = 0.0f;
    if (rc == RUMI_OK) (void)hipEventElapsedTime(&ms, ev0, ev1);
    (void)hipEventDestroy(ev0);
    (void)hipEventDestroy(ev1);
    if (stats) {
        stats->total_samples = b->total_samples;
        stats->total_pixels = b->total_pixels;
        stats->total_texels = b->total_texels;
        stats->n_items = b->n;
        stats->generations = b->generations;
        stats->milliseconds = ms;
        if (rc == RUMI_OK && stats->host_item_ms) {
            std::vector<uint64_t> done(b->n + 1);
            if (hipMemcpy(done.data(), b->done, 8 * ((size_t)b->n + 1), hipMemcpyDeviceToHost) != hipSuccess)
                return fail(RUMI_ERR_HIP, "rumi_batch_run: completion stamps unreadable");
            for (uint32_t i = 0; i < b->n; ++i)
                stats->host_item_ms[i] =
                    done[i] >= done[b->n] ? (double)(done[i] - done[b->n]) / b->clock_khz : -1.0;
        }
    }
    return rc;
}

extern "C" int rumi_batch_destroy(rumi_batch* b) {
    if (!b) return RUMI_OK;
    batch_free(b);
    delete b;
    return RUMI_OK;
}
