/*
 * rumi.h -- C ABI of the MI355X (gfx950) rendering core.
 *
 * The reference (Rumi-Renderer) has no FFI: its boundary is a C++ class
 * hierarchy called per ray and per vertex.  Each entry point below replaces
 * the reference member function named in its comment (paths relative to the
 * reference source tree); the C++ drop-in classes in
 * rumi-renderer_amd/host/ call these, and so does any FFI binding (ctypes in
 * the test harness).
 *
 * Conventions
 *   - Every function returns RUMI_OK (0) or a positive error code; the
 *     message of the last failure on the calling thread is rumi_last_error().
 *   - Pointers named dev_* are device pointers (HBM) owned by the caller;
 *     pointers named host_* are host memory.  Sizes are element counts.
 *   - stream is a hipStream_t passed as void* (NULL = the null stream).
 *     Calls are asynchronous on that stream unless stated otherwise.
 *   - Handles are not shared between threads without external locking.
 */
#ifndef RUMI_H
#define RUMI_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---------------------------------------------------------------- status */
enum {
    RUMI_OK = 0,
    RUMI_ERR_INVALID = 1,    /* bad argument, size or handle */
    RUMI_ERR_HIP = 2,        /* a HIP runtime call failed */
    RUMI_ERR_NO_MEMORY = 3,  /* host or device allocation failed */
    RUMI_ERR_NO_DEVICE = 4,  /* no GPU / no device handle */
    RUMI_ERR_UNSUPPORTED = 5 /* combination not implemented */
};

/* Thread-local text of the last error ("" when none). */
const char* rumi_last_error(void);

/* Library version string. */
const char* rumi_version(void);

/* --------------------------------------------------------------- context */
typedef struct rumi_ctx rumi_ctx;

/* Binds a GPU (HIP device ordinal).  No reference counterpart: the
 * reference is single-threaded CPU code (src/main.cpp:5-20). */
int rumi_ctx_create(int device, rumi_ctx** out);
int rumi_ctx_destroy(rumi_ctx* ctx);

/* ------------------------------------------------------------------ rays */
/* Ray record, 32 B: the reference's Ray (src/Kernel/ray.h:9-31).
 * tmin is the reference's "closest hit so far" (0 = none yet), tmax the
 * segment limit (FLT_MAX or +inf for an unbounded ray). */
typedef struct {
    float ox, oy, oz, tmin;
    float dx, dy, dz, tmax;
} rumi_ray;

/* Closest-hit record, 8 B: primitive index in the caller's input order
 * (-1 = miss) and the hit distance (the ray's tmin on a miss). */
typedef struct {
    int32_t prim;
    float t;
} rumi_hit;

/* ------------------------------------------------------------------- BVH */
typedef struct rumi_bvh rumi_bvh;

enum {
    /* Binned SAH with FP64 cost sums; leaves of at most max_leaf. */
    RUMI_BVH_SAH_BINNED = 0,
    /* The reference's builder: longest-axis median split after a shell sort
     * of box centres, leaves of at most 6 (src/Acceleration/bvh.cpp:48-148). */
    RUMI_BVH_REF_MEDIAN6 = 1
};

typedef struct {
    int32_t mode;     /* RUMI_BVH_* */
    int32_t max_leaf; /* SAH: 1..15 (default 4); ignored for REF_MEDIAN6 */
    int32_t n_bins;   /* SAH bins per axis, 4..256 (default 32) */
    int32_t reserved;
} rumi_bvh_params;

typedef struct {
    int32_t mode;
    int32_t n_prims;
    int32_t n_inner;   /* interior nodes (64 B each on the device) */
    int32_t n_leaves;
    int32_t max_depth; /* interior nodes on the deepest root-to-leaf path */
    int32_t on_device; /* 1 when uploaded to a context's GPU */
    double sah_cost;   /* SAH cost (C_node = C_prim = 1) relative to the root area */
    double build_seconds;
} rumi_bvh_info;

/* Builds a BVH over an indexed triangle mesh: xyz is nv x 3 floats, idx is
 * nt x 3 vertex indices (host memory).  Replaces BVHAccel::BVHAccel +
 * BuildBVH (src/Acceleration/bvh.cpp:40-45, 129-148) over Triangle shapes
 * (src/Shape/triangle.cpp:120-148 for the per-primitive bounds).  Unlike the
 * reference it does not permute the caller's arrays: hits report input
 * triangle indices.  ctx may be NULL: the BVH then stays on the host (for
 * inspection and export only); otherwise it is uploaded to ctx's GPU. */
int rumi_bvh_build_mesh(rumi_ctx* ctx, const float* host_xyz, uint32_t nv,
                        const uint32_t* host_idx, uint32_t nt,
                        const rumi_bvh_params* params, rumi_bvh** out);
int rumi_bvh_destroy(rumi_bvh* bvh);
int rumi_bvh_get_info(const rumi_bvh* bvh, rumi_bvh_info* info);

/* Copies the BVH's layout to host arrays (any may be NULL):
 *   child_boxes  n_inner x 12 floats (child 0 lo.xyz hi.xyz, child 1 lo.xyz hi.xyz)
 *   child_refs   n_inner x 2: >= 0 interior node, < 0 leaf ~((first << 4) | count)
 *   root_box     6 floats, root_ref 1 int
 *   leaf_prims   n_prims: input primitive index of each leaf slot. */
int rumi_bvh_export(const rumi_bvh* bvh, float* host_child_boxes, int32_t* host_child_refs,
                    float* host_root_box, int32_t* host_root_ref, int32_t* host_leaf_prims);

/* ---------------------------------------------------------------- traversal */
enum {
    /* Visiting order: DEFAULT = ORDERED for SAH trees, REF_ORDER for
     * REF_MEDIAN6 trees. */
    RUMI_TRACE_DEFAULT = 0,
    /* The reference's order: push right, descend left (bvh.cpp:188-200). */
    RUMI_TRACE_REF_ORDER = 1,
    /* Near child first by box entry distance. */
    RUMI_TRACE_ORDERED = 2
};

/* Closest hit of n rays: BVHAccel::Travel(ray, isect != NULL)
 * (src/Acceleration/bvh.cpp:150-212) with Triangle::Intersect's acceptance
 * rules (src/Shape/triangle.cpp:41-118).  dev_rays n x rumi_ray,
 * dev_hits n x rumi_hit; dev_uv (n x 2 floats, barycentric u, v of the hit)
 * may be NULL. */
int rumi_trace_closest(rumi_ctx* ctx, const rumi_bvh* bvh, const rumi_ray* dev_rays,
                       uint32_t n, rumi_hit* dev_hits, float* dev_uv, uint32_t flags,
                       void* stream);

/* Shadow-ray (any-hit) query of n rays: BVHAccel::Travel(ray, NULL), which
 * returns after the first leaf that produced a hit (bvh.cpp:182-184) and
 * serves Scene::VisibilityTest (src/Scene/scene.cpp:99-123).  dev_prim
 * receives the index of a hit primitive or -1. */
int rumi_trace_any(rumi_ctx* ctx, const rumi_bvh* bvh, const rumi_ray* dev_rays,
                   uint32_t n, int32_t* dev_prim, uint32_t flags, void* stream);

/* Instrumented traversal: the closest-hit (any = 0) or any-hit (any = 1)
 * kernel that also writes, per ray, the interior nodes entered and the
 * primitives tested (the per-ray work the roofline's algorithmic bytes
 * count).  Outputs as rumi_trace_closest / rumi_trace_any. */
int rumi_trace_count(rumi_ctx* ctx, const rumi_bvh* bvh, const rumi_ray* dev_rays,
                     uint32_t n, int32_t any, rumi_hit* dev_hits, int32_t* dev_node_count,
                     int32_t* dev_prim_count, uint32_t flags, void* stream);

/* Launch shape of rumi_trace_closest / rumi_trace_any on this context (no
 * counterpart in the reference, whose traversal is one call per ray).  By
 * default every lane steps its own ray and swaps in a ray it has already
 * prefetched when its ray ends (persistent 2); persistent 1 instead refills
 * a wave from a shared counter once `refill` of its lanes are idle; -1 is one
 * ray per lane.  width 4 traverses SAH trees in their 4-wide form (near-first
 * order; persistent 1 then runs as 2).  Zero fields keep the defaults: block
 * 256 threads, stack 16 LDS entries per lane (deeper entries spill to private
 * memory), refill 16, persistent 2, width 4.  Within one width every setting
 * returns the same bits.  Between widths 2 and 4 the hit distances agree and
 * only which of two primitives at exactly equal distance is reported can
 * differ, except for rays with a zero direction component that start on the
 * plane of a box face: the reference's box test rejects such a box when the
 * plane is an x face, so the hit can depend on which boxes the tree holds
 * (rumi_trace_count follows the width). */
typedef struct {
    int32_t block;       /* 64, 128 or 256 */
    int32_t stack;       /* 8 or 16 */
    int32_t refill;      /* 1..64 (persistent 1) */
    int32_t persistent;  /* 2 prefetching lanes, 1 threshold refill, -1 one ray per lane */
    int32_t width;       /* 2 binary nodes, 4 four-wide nodes (SAH trees) */
} rumi_trace_tuning;

int rumi_trace_set_tuning(rumi_ctx* ctx, const rumi_trace_tuning* tuning);
int rumi_trace_get_tuning(const rumi_ctx* ctx, rumi_trace_tuning* tuning);

/* ----------------------------------------------------------------- scenes */
/* A scene replaces the reference's Scene object graph (src/Scene/scene.h:
 * 16-53): shapes with materials, area lights, an optional homogeneous medium
 * and a pinhole camera, uploaded once and rendered many times. */
typedef struct rumi_scene rumi_scene;

enum {
    RUMI_SHAPE_TRIANGLE = 0, /* Triangle(v0, v1, v2) + Normal(n0, n1, n2) (src/Shape/triangle.cpp) */
    RUMI_SHAPE_SPHERE = 1,   /* Sphere(centre, radius) (src/Shape/sphere.cpp) */
    RUMI_SHAPE_RECT = 2,     /* Rect(p, l, w) (src/Shape/rect.cpp) */
    RUMI_SHAPE_MESH = 3      /* indexed triangle mesh, one material */
};

enum {
    RUMI_MAT_MATTE = 0,      /* Matte(kd, color) (src/Material/matte.cpp) */
    RUMI_MAT_GLOSSY = 1,     /* Glossy(kd, ks, exponent, color) (src/Material/glossy.cpp) */
    RUMI_MAT_SPECULAR = 2,   /* Specular(kr = kd, color) (src/Material/specular.cpp) */
    RUMI_MAT_EMISSIVE = 3,   /* Emissive(ls = kd, color) (src/Material/emissive.cpp) */
    RUMI_MAT_DIELECTRIC = 4, /* Dielectric(eta_in, eta_out, absorption) (src/Material/dielectric.cpp) */
    RUMI_MAT_SUBSURFACE = 5, /* Subsurface(eta_in, eta_out, sigma_s, sigma_a, density) (src/Material/subsurface.cpp) */
    RUMI_MAT_GGX = 6         /* extension: GGX microfacet reflection, Schlick Fresnel with F0 = color,
                                roughness = alpha (Walter et al. 2007); no reference counterpart */
};

typedef struct {
    int32_t kind;        /* RUMI_MAT_* */
    float kd;            /* Matte/Glossy kd, Specular kr, Emissive ls */
    float ks;            /* Glossy ks */
    float exponent;      /* Glossy Phong exponent */
    float color[3];      /* RGB albedo / specular colour / emitted colour */
    float eta_in, eta_out;
    float absorption[3]; /* Dielectric Beer-Lambert coefficients */
    float sigma_s[3], sigma_a[3];
    float density;       /* Subsurface sampling density (0 = 1) */
    float roughness;     /* GGX alpha */
    int32_t reserved[3];
} rumi_material;

typedef struct {
    int32_t kind;       /* RUMI_SHAPE_* */
    int32_t material;   /* index into rumi_scene_desc.materials */
    int32_t is_light;   /* 1: an AreaLight on this shape (rect or sphere, emissive material);
                           2: an ImageBasedLight (src/Light/imagebasedlight.cpp:24-64) on
                           a sphere with an emissive material and a texture: light
                           samples cosine-distributed about the shading normal, the
                           texture's value where they leave the sphere, pdf cos/pi;
                           path tracer and RayTracer only, no medium */
    int32_t texture;    /* index into rumi_scene_desc.textures or -1 (multiplies radiance,
                           PathTracer::Trace pathtracer.cpp:343-344) */
    float v[9];         /* triangle v0 v1 v2 | sphere centre in v[0..2] | rect p, l, w */
    float n[9];         /* triangle vertex normals (all zero: geometric normal); sphere:
                           the rows of Shape::_transform's 3x3 its TextureMapping measures
                           angles against (sphere.cpp:172-195; all zero: identity) */
    float radius;       /* sphere */
    /* RUMI_SHAPE_MESH (host memory, copied at scene creation): */
    const float* xyz;   uint32_t nv;
    const uint32_t* idx; uint32_t nt;
    const float* normals; /* nv x 3 or NULL (geometric normals) */
    const float* uv;      /* nv x 2 texture coordinates or NULL */
} rumi_shape;

typedef struct {
    float position[3];
    float basis[9];     /* row-major rotation: world direction = basis * camera direction;
                           the camera looks along its +z (Pinhole, src/Camera/pinhole.cpp:30-51) */
    float fov_deg;      /* full field of view; the reference fixes 20 (pinhole.cpp:36) */
    float y_ratio;      /* image height / width */
    /* Pinhole::DepthOfField(focal_distance, lens_radius) (pinhole.cpp:89-95,
     * 111-122): with focal_distance > 0 each camera ray starts on a lens disk
     * (r = sqrt(u) lens_radius, angle 2 pi v, along the camera's up and
     * right axes) and aims at its pinhole ray's point on the focal plane.
     * 0 = a pinhole.  The bidirectional tracer ignores the lens, as the
     * reference's BidirectionalPathTracer resets it to a pinhole
     * (bidirectionalpathtracer.cpp:56). */
    float focal_distance, lens_radius;
} rumi_camera;

enum { RUMI_MEDIUM_NONE = 0, RUMI_MEDIUM_BOX = 1, RUMI_MEDIUM_SPHERE = 2 };

typedef struct {
    int32_t bound;      /* RUMI_MEDIUM_*: BOX is HomogeneousVolume(BBox, ...) (src/Volume/
                           homogeneousvolume.cpp:27-38); SPHERE is an extension */
    float lo[3], hi[3];
    float center[3], radius;
    float sigma_s[3], sigma_a[3];
    float g;            /* Henyey-Greenstein anisotropy; 0 = the reference's Isotropic phase */
} rumi_medium;

enum { RUMI_TEXTURE_IMAGE = 0, RUMI_TEXTURE_CHECKER = 1, RUMI_TEXTURE_HDR = 2 };

typedef struct {
    int32_t width, height;
    const float* rgb;   /* IMAGE: host, width * height * 3, row 0 = v = 0 (ImageTexture::Pixel,
                           src/Texture/imagetexture.cpp:32-61: nearest texel) */
    int32_t kind;       /* RUMI_TEXTURE_* (0 = IMAGE); HDR: HDRTexture (src/Texture/
                           hdrtexture.cpp:21-41): rgb rows top-down as rumi_hdr_read gives
                           them, texel (int)((1 - v)(h - 1)), (int)((1 - u)(w - 1)) */
    /* CHECKER: Checker(w, h, color_1, color_2) (src/Texture/checker.cpp:16-32):
     * color_2 where (int)(u w) and (int)(v h) are both even or both odd, else
     * color_1; no texels (no texture gradient) */
    float checker_w, checker_h;
    float color_1[3], color_2[3];
} rumi_texture;

/* PointLight(position, ls, rgb) (src/Light/pointlight.cpp:16-72): a light
 * with no shape (never hit, not in the scene tree).  Its light sample is the
 * point itself: Le = rgb * ls, G = 1 / d^2, Pdf = 1, so next-event estimation
 * from it is deterministic.  The path tracer and the RayTracer take it, with
 * or without a medium; the bidirectional tracer rejects it. */
typedef struct {
    float position[3];
    float ls;
    float color[3];
} rumi_point_light;

typedef struct {
    const rumi_shape* shapes;       uint32_t n_shapes;
    const rumi_material* materials; uint32_t n_materials;
    const rumi_texture* textures;   uint32_t n_textures;
    rumi_camera camera;
    rumi_medium medium;
    const rumi_point_light* point_lights; uint32_t n_point_lights;  /* after the shapes' lights */
} rumi_scene_desc;

/* Builds the scene's BVH (binned SAH over all primitives) and uploads it
 * with the shading tables.  Replaces Scene::Build's object graph + BVHAccel
 * (src/Scene/scene.cpp:135-204). */
int rumi_scene_create(rumi_ctx* ctx, const rumi_scene_desc* desc, rumi_scene** out);
int rumi_scene_destroy(rumi_scene* scene);

/* ---------------------------------------------------------------- render */
/* RUMI_INTEGRATOR_WHITTED: RayTracer(scene, max_depth)
 * (src/Integrator/raytracer.cpp:30-86): Specular and Dielectric surfaces
 * trace on; every other surface ends the path with its DirectLighting (a
 * light sample for Matte, and for Glossy only when ks = 0, the RayTracer
 * leaving the lobe variable at 1); emitters count when hit directly or
 * through specular/dielectric chains; no medium sampling along rays
 * (shadow rays keep the medium's transmittance); up to max_depth + 1
 * surfaces. */
enum { RUMI_INTEGRATOR_PT = 0, RUMI_INTEGRATOR_BDPT = 1, RUMI_INTEGRATOR_WHITTED = 2 };

typedef struct {
    int32_t integrator;     /* RUMI_INTEGRATOR_* */
    int32_t width, height;
    int32_t spp;
    int32_t max_depth;      /* PathTracer(scene, max_depth): surface/medium vertices per path */
    int32_t n_light, n_eye; /* BidirectionalPathTracer(scene, nL, nE) vertex caps */
    float mis_beta;         /* BDPT MIS exponent: 1 balance (the reference's), 2 power */
    uint64_t seed;
    uint32_t queue_capacity;    /* wavefront paths in flight (BDPT: samples per batch), 0 = automatic
                                   (2^23 paths, at most the frame's samples;
                                   BDPT 2^24 samples or what 90% of the free
                                   device memory holds, down to 2^20) */
    uint32_t regen_threshold;   /* refill the queue when fewer paths are alive, 0 = always */
    uint32_t flags;             /* RUMI_RENDER_* */
} rumi_render_desc;

/* rumi_render_desc.flags: count the traversal work of every extension and
 * shadow ray into rumi_render_stats (interior nodes entered, primitives
 * tested; for the bidirectional tracer also its walk segments and
 * connection rays) -- the per-ray work behind the renderer's algorithmic
 * bytes.  Same image, slower kernels; path tracer and bidirectional tracer. */
#define RUMI_RENDER_COUNT_TRAVERSAL 1u

typedef struct {
    uint64_t samples;
    uint64_t extension_rays;  /* BDPT: walk segments, counting renders only */
    uint64_t shadow_rays;     /* BDPT: connection rays, counting renders only */
    uint64_t iterations;
    double milliseconds;   /* device time between the first and last stage (HIP events) */
    /* RUMI_RENDER_COUNT_TRAVERSAL only (else 0): interior nodes entered and
     * primitives tested by all extension (BDPT: walk) rays and all shadow
     * (BDPT: connection) rays */
    uint64_t ext_nodes, ext_prims, shadow_nodes, shadow_prims;
} rumi_render_stats;

/* Renders one frame: Scene::Rendering's pixel x sample loop
 * (src/Scene/scene.cpp:206-258) over Integrator::Trace(w, h, camera, film)
 * (PathTracer src/Integrator/pathtracer.cpp:117-354; BidirectionalPathTracer
 * src/Integrator/bidirectionalpathtracer.cpp:121-394, all materials, the
 * homogeneous medium and textures; for BDPT max_depth is the path length in
 * edges and n_light / n_eye default to max_depth + 1).  dev_rgb receives the
 * per-pixel mean radiance, height x width x 3 floats RGB, row 0 at the top of
 * the image as Film::Pixel indexes it (src/Camera/film.cpp:83-97).  Samples
 * whose radiance is NaN are dropped, as Scene::Rendering drops them.
 * stats may be NULL.  Synchronises the stream before returning. */
int rumi_render(rumi_ctx* ctx, const rumi_scene* scene, const rumi_render_desc* desc,
                float* dev_rgb, rumi_render_stats* stats, void* stream);

/* rumi_render without any device allocation.  No counterpart in the
 * reference: it is for callers that cannot allocate device memory (code
 * under HIP graph capture, tight serving loops, harnesses that forbid
 * hipMalloc).  The path tracer and the RayTracer only (BDPT:
 * RUMI_ERR_UNSUPPORTED).  The wavefront queues live in one statically sized
 * device array per GPU (RUMI_NOALLOC_MAX_PATHS paths, about 61 MB, held
 * until the process ends); desc->queue_capacity 0 means that maximum, and a
 * larger one is rejected with RUMI_ERR_INVALID.  The image equals
 * rumi_render's with the same seed and queue capacity (up to the order of
 * float atomics).  Calls on one GPU share the array: calls that overlap on
 * one GPU (other threads, other streams) race.  stats->milliseconds is host
 * time around the call (no HIP events are created).  Synchronises the
 * stream before returning. */
#define RUMI_NOALLOC_MAX_PATHS (1u << 18)
int rumi_render_noalloc(rumi_ctx* ctx, const rumi_scene* scene, const rumi_render_desc* desc,
                        float* dev_rgb, rumi_render_stats* stats, void* stream);

/* -------------------------------------------------------------- autotuner */
/* The launch-shape autotuner (no counterpart in the reference, which has one
 * code path per call): times candidate settings on the caller's own
 * workload with HIP events and keeps the fastest.  Every candidate gives the
 * same results (traversal: the same hit distances, the same bits within one
 * node width; rendering: the same image up to the order of float atomics).
 * Synchronous on stream. */
typedef struct {
    uint32_t n_candidates;        /* out: candidates timed */
    float best_us;                /* out: the winner's median time */
    float* host_candidate_us;     /* caller array or NULL: each candidate's median time in us
                                     (-2: skipped as equal to the previous candidate,
                                     -3: skipped, its buffers did not fit in device memory) */
    uint32_t candidates_capacity; /* entries host_candidate_us holds */
    int32_t repeats;              /* timed runs per candidate after one untimed run (0: 3 for
                                     traversal, 1 for rendering) */
} rumi_autotune_report;

/* The traversal candidates in the order rumi_autotune_trace times them:
 * one ray per lane; prefetching lanes for width 2 and 4, blocks 64/128/256
 * and LDS stacks 8/16; threshold refill for every block, stack and refill
 * 16/32/48/64.  out may be NULL (count only). */
int rumi_autotune_trace_candidates(rumi_trace_tuning* out, uint32_t capacity, uint32_t* count);

/* Times rumi_trace_closest (any = 0; dev_out: n rumi_hit) or rumi_trace_any
 * (any = 1; dev_out: n int32) on the caller's ray batch under every
 * candidate and leaves the fastest set on the context (rumi_trace_set_tuning);
 * best and report may be NULL.  The batch the reference would trace with
 * BVHAccel::Travel (src/Acceleration/bvh.cpp:150-212) ray by ray. */
int rumi_autotune_trace(rumi_ctx* ctx, const rumi_bvh* bvh, const rumi_ray* dev_rays, uint32_t n,
                        int32_t any, uint32_t flags, void* dev_out, rumi_trace_tuning* best,
                        rumi_autotune_report* report, void* stream);

/* Renders desc's frame with candidate queue sizes and writes the fastest
 * into desc (path tracer and RayTracer: queue_capacity 2^19..2^23 with the
 * refill threshold at the capacity or half of it; bidirectional tracer:
 * samples per batch 2^20..2^26, about 2 KB of device memory per sample
 * at 9 + 9 vertices; a candidate that does not fit in device memory is
 * skipped); dev_rgb receives the last render.  The
 * frame of Scene::Rendering (src/Scene/scene.cpp:206-258). */
int rumi_autotune_render(rumi_ctx* ctx, const rumi_scene* scene, rumi_render_desc* desc, float* dev_rgb,
                         rumi_autotune_report* report, void* stream);

/* ---------------------------------------------------------------- adjoint */
/* Device buffers that rumi_adjoint overwrites; any may be NULL. */
typedef struct {
    float* d_medium;    /* 7 floats: dLoss/d sigma_a (rgb), sigma_s (rgb), g of the medium */
    float* d_textures;  /* 3 floats per texel of every texture, in scene order and texel order */
    float* d_image;     /* the primal pass's image (height x width x 3), as rumi_render writes it */
} rumi_grad_buffers;

/* Path-replay adjoint of rumi_render with the path tracer (the reference has
 * no differentiable renderer; this extends PathTracer::Trace,
 * src/Integrator/pathtracer.cpp:117-354).  Given dLoss/dImage (device,
 * height x width x 3, the derivative with respect to the per-pixel mean
 * that rumi_render writes), runs a primal pass that records every sample's
 * radiance, then replays each path with the same random numbers and
 * accumulates the derivative of the expected image with respect to the
 * medium's sigma_a, sigma_s, g and to every texel the paths read.  Gradients
 * follow the estimator as rendered, including its sampling decisions, so
 * they match finite differences of rumi_render.  Holds 24 bytes per sample
 * of device memory (width x height x spp < 2^32).  Synchronises the stream
 * before returning. */
int rumi_adjoint(rumi_ctx* ctx, const rumi_scene* scene, const rumi_render_desc* desc,
                 const float* dev_dloss_dimage, const rumi_grad_buffers* grads,
                 rumi_render_stats* stats, void* stream);

/* ---------------------------------------------------------------- batches */
/* Many scenes rendered (and differentiated) as one wavefront.  No reference
 * counterpart: the reference renders one hard-coded scene per process
 * (Scene::Build + Scene::Rendering, src/Scene/scene.cpp:135-258).  Samples
 * are numbered item after item; each item's image equals the image
 * rumi_render gives its scene alone with the same seed (up to the order of
 * float atomics).  The launch schedule is fixed at creation (generations of
 * queue_capacity samples, max_depth iterations each), so a run needs no host
 * round trip; rumi_batch_run with use_graph captures it once as a hipGraph
 * and replays it. */
typedef struct rumi_batch rumi_batch;

typedef struct {
    const rumi_scene* scene;  /* the item's scene and camera */
    int32_t width, height;
} rumi_batch_item;

typedef struct {
    int32_t spp;              /* samples per pixel of every item */
    int32_t max_depth;        /* PathTracer max depth of every item */
    uint64_t seed;            /* forward pass; the adjoint passes use seed + 1 */
    uint32_t queue_capacity;  /* samples per generation (0 = 2^26, or what 90% of the free
                                 device memory holds, down to 2^22) */
    int32_t adjoint;          /* 1: the run also forms dLoss/dImage = 2 (I - target)
                                 and runs the path-replay adjoint */
} rumi_batch_desc;

typedef struct {
    float* dev_images;         /* out: every item's image, concatenated in item order */
    const float* dev_targets;  /* adjoint: the target images, same layout */
    float* dev_texel_grads;    /* adjoint, out: dLoss/dtexel, items' texel pools concatenated */
    float* dev_loss;           /* adjoint, out (may be NULL): sum (I - T)^2 per item */
} rumi_batch_io;

typedef struct {
    uint64_t total_samples, total_pixels, total_texels;
    uint32_t n_items, generations;
    double milliseconds;   /* device time of the run */
    double* host_item_ms;  /* caller array (n_items) or NULL: when each item's last
                              forward path ended, ms after the run began */
} rumi_batch_stats;

int rumi_batch_create(rumi_ctx* ctx, const rumi_batch_item* items, uint32_t n,
                      const rumi_batch_desc* desc, rumi_batch** out);
/* Runs the batch on stream (use_graph: replay the captured hipGraph,
 * capturing it first if the io pointers changed).  Synchronises the stream. */
int rumi_batch_run(rumi_batch* batch, const rumi_batch_io* io, int32_t use_graph,
                   rumi_batch_stats* stats, void* stream);
int rumi_batch_destroy(rumi_batch* batch);

/* ------------------------------------------------------------ workloads */
/* Seeded geodesic icosphere of frequency f (20 f^2 triangles, 10 f^2 + 2
 * vertices) with each vertex at radius 1 + amplitude * (2u - 1), u from
 * PCG32(seed, stream = vertex index).  Host memory. */
int rumi_make_icosphere(uint32_t freq, uint64_t seed, float amplitude, float* host_xyz,
                        uint32_t* host_idx);

/* n rays, origins uniform in [-half_extent, half_extent]^3, directions
 * uniform on the sphere (z = 1 - 2u, phi = 2 pi v), tmin = 0,
 * tmax = +inf; PCG32(seed, stream = ray index).  Host memory. */
int rumi_make_rays_uniform(uint64_t n, uint64_t seed, float half_extent, rumi_ray* host_rays);

/* -------------------------------------------------------------------- film */
/* Film::Save's conversion of the float film to pixel bytes
 * (src/Camera/film.cpp:152-203), on the GPU: per channel
 *   v = film value (RUMI_FILM_BGR_SUM: the reference's _pixel_buffer, B, G, R
 *       sums, times 1 / d^2 with d = (int)sqrt((float)n_samples), as Save
 *       divides; RUMI_FILM_RGB_MEAN: rumi_render's image, R, G, B means,
 *       taken as they are),
 *   v = powf(v, 1 / gamma) * 255, clamped to [0, 255], truncated;
 * powf is taken correctly rounded (the double power rounded once), and a
 * NaN (a negative film value) gives 0.  dev_out: width * height *
 * bytes_per_pixel bytes B, G, R (+ 0 for 4 bytes per pixel), in the film's
 * row order -- the pixel data SaveBmp writes (src/Kernel/bmp.cpp:59-97). */
enum { RUMI_FILM_BGR_SUM = 0, RUMI_FILM_RGB_MEAN = 1 };

int rumi_film_encode(rumi_ctx* ctx, const float* dev_film, uint32_t width, uint32_t height,
                     int32_t layout, uint32_t n_samples, float gamma, uint32_t bytes_per_pixel,
                     uint8_t* dev_out, void* stream);

/* FirefliesFiltering (src/Scene/scene.cpp:261-431) over the pixel bytes of a
 * loaded image, on the GPU: a pixel with a channel at 255 none of whose
 * first n_copy gathered neighbours is white is replaced by their FP32 mean
 * (truncated); the reference's gathering rules are kept, including its
 * flipped row order and its edge conditions.  Where the reference reads
 * outside its buffer (saturated pixels of the first and last image rows),
 * those neighbours read as 0; the fourth byte of a 32-bit pixel is written
 * as 0 (the reference leaves it unset).  dev_src and dev_dst must differ. */
int rumi_film_fireflies(rumi_ctx* ctx, const uint8_t* dev_src, uint32_t width, uint32_t height,
                        uint32_t bytes_per_pixel, uint8_t* dev_dst, void* stream);

/* SaveBmp's 54-byte header (src/Kernel/bmp.cpp:59-97): "BM", sizes, offset
 * 54, a 40-byte info header with positive height and no row padding counted.
 * Host memory. */
int rumi_bmp_header(uint32_t width, uint32_t height, uint32_t bytes_per_pixel, uint8_t* host_out);

/* LoadBmp's reading rule (src/Kernel/bmp.cpp:3-57) over a file image in host
 * memory: size fields from the header; rows back to back when the width is a
 * multiple of 4, else each row followed by (width % 4) skipped bytes.
 * pixels NULL: only the size fields. */
int rumi_bmp_read(const uint8_t* host_file, uint64_t file_bytes, uint32_t* width, uint32_t* height,
                  uint32_t* bytes_per_pixel, uint8_t* host_pixels, uint64_t pixel_capacity);

/* LoadHDR (src/Kernel/hdr.cpp:4-177) over a Radiance RGBE file image in host
 * memory: "#?RADIANCE", '#' lines, "FORMAT=32-bit_rle_rgbe", lines up to an
 * empty one, "-Y h +X w", then h new-style run-length scanlines; each pixel
 * (r, g, b, e) -> (r, g, b) * 2^(e - 136), 0 when e = 0.  rgb: w * h * 3
 * floats, rows top-down as in the file (a RUMI_TEXTURE_HDR's data); NULL:
 * only the size.  A file the reference cannot read (another format, old
 * run-length scanlines, a width below 8 or above 32767 or whose low byte is
 * 128 or more -- the reference compares it as a signed char) is rejected. */
int rumi_hdr_read(const uint8_t* host_file, uint64_t file_bytes, uint32_t* width, uint32_t* height,
                  float* host_rgb, uint64_t rgb_capacity);

/* ImageTexture(file, gamma)'s texel values (src/Texture/imagetexture.cpp:
 * 21-61) from loaded pixel bytes: byte * (1 / 255) in FP32, raised to gamma
 * (powf taken correctly rounded) unless gamma is 1; rgb: width * height * 3
 * floats R, G, B in the buffer's row order (row 0 = v = 0), a
 * RUMI_TEXTURE_IMAGE's data.  Host memory. */
int rumi_bmp_texels(const uint8_t* host_pixels, uint32_t width, uint32_t height,
                    uint32_t bytes_per_pixel, float gamma, float* host_rgb);

/* ------------------------------------------------------------------ models */
/* An OBJ model, loaded as Model::Model(file, transform) loads it
 * (Model::LoadObjectFile, src/Shape/model.cpp:72-338): "v", "vt", "vn" and
 * "f" records found by the reference's lexical rules (a record tag is up to
 * four bytes ending at a space or newline; a face keeps its first three
 * corners; face indices are parsed digit by digit in FP32, a '-' makes a
 * corner relative to the file's total counts); vertices go through the
 * row-major 4x4 `transform` (Mat4x4 * Point3, fourth column added), normals
 * through the transposed inverse computed by the reference's unpivoted
 * elimination (Shape::Shape, src/Shape/shape.cpp:29-37; Inverse,
 * src/Kernel/matrix.cpp:67-133), then normalised; a file without "vn" lines,
 * or a zero normal, gets the transformed face's normal.  Texture
 * coordinates are read and dropped, as the reference's triangles carry none.
 *
 * The reference places a Model in its scene tree as one shape with its own
 * BVHAccel inside (Model::Intersect, model.cpp:71-74).  Here the triangles
 * join the scene as one RUMI_SHAPE_MESH (xyz = 3 vertices per triangle,
 * idx = 0, 1, 2, ..., normals = the vertex normals) under the scene's single
 * tree: the same closest hit, the same traversal kernels.
 *
 * Malformed faces whose reference behaviour is undefined (fewer than three
 * corners, a fourth '/' field, an index outside the file's records) are
 * rejected with RUMI_ERR_INVALID.  transform NULL = identity. */
typedef struct rumi_obj rumi_obj;

int rumi_obj_load(const char* path, const float* transform, rumi_obj** out);
/* n_triangles; xyz and normals: 9 floats per triangle (corners 0, 1, 2),
 * host memory owned by the model; counts: the file's v, vt, vn and f record
 * counts (4 values).  Any output may be NULL. */
int rumi_obj_get(const rumi_obj* model, uint32_t* n_triangles, const float** xyz,
                 const float** normals, uint32_t* counts);
int rumi_obj_destroy(rumi_obj* model);

/* PCG32 (O'Neill, pcg32_srandom_r / pcg32_random_r): writes n outputs of
 * the generator seeded with (initstate, initseq).  Host memory. */
int rumi_pcg32_sequence(uint64_t initstate, uint64_t initseq, uint32_t n, uint32_t* host_out);

#ifdef __cplusplus
}
#endif
#endif
