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

/* PCG32 (O'Neill, pcg32_srandom_r / pcg32_random_r): writes n outputs of
 * the generator seeded with (initstate, initseq).  Host memory. */
int rumi_pcg32_sequence(uint64_t initstate, uint64_t initseq, uint32_t n, uint32_t* host_out);

#ifdef __cplusplus
}
#endif
#endif
