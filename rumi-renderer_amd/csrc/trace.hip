// BVH traversal kernels for gfx950 and their C-ABI launchers.
//
// One ray per lane (wave64).  The traversal stack lives in LDS, one column
// per lane (entry k of lane t at k * BLOCK + t, so the 64 lanes of a wave
// touch 64 consecutive dwords whatever their depths); trees deeper than the
// LDS stack spill the excess to a private per-lane array.  A stack entry is
// (child ref, entry distance of its box): the box test runs when the parent
// is entered, the "closer than the current hit" test when the entry is
// popped, which is the moment BVHAccel::Travel tests a popped node
// (src/Acceleration/bvh.cpp:164-169).
//
// Every FP32 expression keeps the reference's operation order and the file
// is compiled with -ffp-contract=off and correctly rounded division, so the
// box and triangle tests give the bits BBox::Intersect (src/Kernel/bbox.cpp:
// 30-92) and Triangle::Intersect (src/Shape/triangle.cpp:41-118) give.
#include <hip/hip_runtime.h>

#include "internal.h"

using namespace rumi;

namespace {

constexpr float kEps = 1e-4f;  // Ray::_epsilon (src/Kernel/ray.cpp:4)

struct RootInfo {
    float lo[3];
    float hi[3];
    int32_t ref;
};

// BBox::Intersect with a precomputed reciprocal direction (1.0f / d is the
// same value whether computed per node or once per ray).
__device__ __forceinline__ bool slab(float lx, float ly, float lz, float hx, float hy, float hz,
                                     float ox, float oy, float oz, float ix, float iy, float iz,
                                     float& tn, float& tf) {
    float t0 = (lx - ox) * ix;
    float t1 = (hx - ox) * ix;
    float tmp;
    if (t0 > t1) { tmp = t0; t0 = t1; t1 = tmp; }
    tn = t0;
    tf = t1;
    t0 = (ly - oy) * iy;
    t1 = (hy - oy) * iy;
    if (t0 > t1) { tmp = t0; t0 = t1; t1 = tmp; }
    if (tn > t1 || t0 > tf) return false;
    if (tn < t0) tn = t0;
    if (tf > t1) tf = t1;
    t0 = (lz - oz) * iz;
    t1 = (hz - oz) * iz;
    if (t0 > t1) { tmp = t0; t0 = t1; t1 = tmp; }
    if (tn > t1 || t0 > tf) return false;
    if (tn < t0) tn = t0;
    if (tf > t1) tf = t1;
    return true;
}

// The tmin-independent part of the node accept test (bvh.cpp:167-168).
__device__ __forceinline__ bool box_static(float lx, float ly, float lz, float hx, float hy,
                                           float hz, float ox, float oy, float oz, float ix,
                                           float iy, float iz, float tmax, float& tn) {
    float tf;
    return slab(lx, ly, lz, hx, hy, hz, ox, oy, oz, ix, iy, iz, tn, tf) && tf >= 0.0f &&
           tmax >= tn;
}

// The tmin part (bvh.cpp:169): no hit yet, or the box starts before it.
__device__ __forceinline__ bool closer(float tmin, float tn) {
    return tmin == 0.0f || tmin > tn;
}

template <int BLOCK, int STACK, bool ANY, bool ORDERED, bool COUNT, bool SPILL>
__global__ __launch_bounds__(BLOCK) void k_trace(
    const float4* __restrict__ rays, uint32_t n, const float4* __restrict__ nodes,
    const float4* __restrict__ tris, RootInfo root, int2* __restrict__ hits,
    float2* __restrict__ uv, int32_t* __restrict__ any_out, int32_t* __restrict__ cnt_nodes,
    int32_t* __restrict__ cnt_prims) {
    __shared__ int32_t s_ref[STACK * BLOCK];
    __shared__ float s_tn[STACK * BLOCK];
    constexpr int kSpill = SPILL ? kMaxDepth - STACK : 1;
    int32_t p_ref[kSpill];
    float p_tn[kSpill];

    const uint32_t tid = threadIdx.x;
    const uint32_t i = blockIdx.x * BLOCK + tid;
    if (i >= n) return;

    const float4 r0 = rays[2 * (size_t)i + 0];
    const float4 r1 = rays[2 * (size_t)i + 1];
    const float ox = r0.x, oy = r0.y, oz = r0.z;
    const float dx = r1.x, dy = r1.y, dz = r1.z, tmax = r1.w;
    float tmin = r0.w;
    int32_t hit = -1;
    float hu = 0.0f, hv = 0.0f;
    int32_t nn = 0, np = 0;

    if (!(dx == 0.0f && dy == 0.0f && dz == 0.0f)) {
        const float ix = 1.0f / dx;
        const float iy = 1.0f / dy;
        const float iz = 1.0f / dz;
        float tn;
        bool go = box_static(root.lo[0], root.lo[1], root.lo[2], root.hi[0], root.hi[1],
                             root.hi[2], ox, oy, oz, ix, iy, iz, tmax, tn) &&
                  closer(tmin, tn);
        int32_t cur = root.ref;
        int sp = 0;
        while (go) {
            if (cur >= 0) {
                const float4* np4 = nodes + 4 * (size_t)cur;
                const float4 a = np4[0];
                const float4 b = np4[1];
                const float4 c = np4[2];
                const float4 d = np4[3];
                if (COUNT) ++nn;
                const int32_t ref0 = __float_as_int(d.x);
                const int32_t ref1 = __float_as_int(d.y);
                float tn0, tn1;
                const bool h0 = box_static(a.x, a.y, a.z, a.w, b.x, b.y, ox, oy, oz, ix, iy, iz,
                                           tmax, tn0);
                const bool h1 = box_static(b.z, b.w, c.x, c.y, c.z, c.w, ox, oy, oz, ix, iy, iz,
                                           tmax, tn1);
                const bool a0 = h0 && closer(tmin, tn0);
                const bool a1 = h1 && closer(tmin, tn1);
                int32_t push_ref = 0;
                float push_tn = 0.0f;
                bool do_push = false;
                bool descend = true;
                if (ORDERED) {
                    if (a0 && a1) {
                        const bool swap = tn1 < tn0;
                        cur = swap ? ref1 : ref0;
                        push_ref = swap ? ref0 : ref1;
                        push_tn = swap ? tn0 : tn1;
                        do_push = true;
                    } else if (a0) {
                        cur = ref0;
                    } else if (a1) {
                        cur = ref1;
                    } else {
                        descend = false;
                    }
                } else {
                    // Reference order: the right child waits on the stack,
                    // the left child is visited next.
                    do_push = a1;
                    push_ref = ref1;
                    push_tn = tn1;
                    if (a0)
                        cur = ref0;
                    else
                        descend = false;
                }
                if (do_push) {
                    if (!SPILL || sp < STACK) {
                        s_ref[sp * BLOCK + tid] = push_ref;
                        s_tn[sp * BLOCK + tid] = push_tn;
                    } else {
                        p_ref[sp - STACK] = push_ref;
                        p_tn[sp - STACK] = push_tn;
                    }
                    ++sp;
                }
                if (descend) continue;
            } else {
                const int32_t enc = ~cur;
                const int32_t first = enc >> 4;
                const int32_t count = enc & 15;
                for (int32_t k = 0; k < count; ++k) {
                    const float4* t4 = tris + 3 * (size_t)(first + k);
                    const float4 ta = t4[0];
                    const float4 tb = t4[1];
                    const float4 tc = t4[2];
                    if (COUNT) ++np;
                    const float v0x = ta.x, v0y = ta.y, v0z = ta.z;
                    const float e1x = ta.w, e1y = tb.x, e1z = tb.y;
                    const float e2x = tb.z, e2y = tb.w, e2z = tc.x;
                    // p = Cross(d, e2)
                    const float px = dy * e2z - e2y * dz;
                    const float py = dz * e2x - e2z * dx;
                    const float pz = dx * e2y - e2x * dy;
                    const float det = px * e1x + py * e1y + pz * e1z;
                    if (-1e-8f < det && det < 1e-8f) continue;
                    const float inv_det = 1.0f / det;
                    const float tx = ox - v0x, ty = oy - v0y, tz = oz - v0z;
                    const float u = (px * tx + py * ty + pz * tz) * inv_det;
                    if (u < 0.0f || u > 1.0f) continue;
                    // q = Cross(T, e1)
                    const float qx = ty * e1z - e1y * tz;
                    const float qy = tz * e1x - e1z * tx;
                    const float qz = tx * e1y - e1x * ty;
                    const float v = (qx * dx + qy * dy + qz * dz) * inv_det;
                    if (v < 0.0f || u + v > 1.0f) continue;
                    const float t = (qx * e2x + qy * e2y + qz * e2z) * inv_det;
                    if (t < kEps) continue;
                    if (tmin > kEps && t > tmin) continue;
                    if (tmax - t < kEps) continue;
                    tmin = t;
                    hit = __float_as_int(tc.y);
                    hu = u;
                    hv = v;
                }
                if (ANY && tmin > 0.0f) break;
            }
            // Pop until an entry still starts before the current hit.
            go = false;
            while (sp > 0) {
                --sp;
                int32_t r;
                float t;
                if (!SPILL || sp < STACK) {
                    r = s_ref[sp * BLOCK + tid];
                    t = s_tn[sp * BLOCK + tid];
                } else {
                    r = p_ref[sp - STACK];
                    t = p_tn[sp - STACK];
                }
                if (closer(tmin, t)) {
                    cur = r;
                    go = true;
                    break;
                }
            }
        }
    }
    if (ANY && any_out) any_out[i] = hit;
    if (hits) hits[i] = make_int2(hit, __float_as_int(tmin));
    if (!ANY && uv) uv[i] = make_float2(hu, hv);
    if (COUNT) {
        cnt_nodes[i] = nn;
        cnt_prims[i] = np;
    }
}

struct TraceArgs {
    const float4* rays;
    uint32_t n;
    const float4* nodes;
    const float4* tris;
    RootInfo root;
    int2* hits;
    float2* uv;
    int32_t* any_out;
    int32_t* cnt_nodes;
    int32_t* cnt_prims;
};

constexpr int kBlock = 128;

template <int STACK, bool ANY, bool ORDERED, bool COUNT, bool SPILL>
hipError_t launch(const TraceArgs& a, hipStream_t s) {
    const uint32_t grid = (a.n + kBlock - 1) / kBlock;
    hipLaunchKernelGGL((k_trace<kBlock, STACK, ANY, ORDERED, COUNT, SPILL>), dim3(grid),
                       dim3(kBlock), 0, s, a.rays, a.n, a.nodes, a.tris, a.root, a.hits, a.uv,
                       a.any_out, a.cnt_nodes, a.cnt_prims);
    return hipGetLastError();
}

template <bool ANY, bool ORDERED, bool COUNT>
hipError_t launch_depth(const TraceArgs& a, int depth, hipStream_t s) {
    if (depth <= 16) return launch<16, ANY, ORDERED, COUNT, false>(a, s);
    if (depth <= 32) return launch<32, ANY, ORDERED, COUNT, false>(a, s);
    return launch<32, ANY, ORDERED, COUNT, true>(a, s);
}

template <bool ANY, bool COUNT>
hipError_t launch_order(const TraceArgs& a, bool ordered, int depth, hipStream_t s) {
    return ordered ? launch_depth<ANY, true, COUNT>(a, depth, s)
                   : launch_depth<ANY, false, COUNT>(a, depth, s);
}

int check_common(rumi_ctx* ctx, const rumi_bvh* bvh, const void* rays, uint32_t n,
                 uint32_t flags, bool* ordered) {
    if (!ctx) return fail(RUMI_ERR_NO_DEVICE, "trace: null context");
    if (!bvh) return fail(RUMI_ERR_INVALID, "trace: null BVH");
    if (bvh->device != ctx->device)
        return fail(RUMI_ERR_INVALID, "trace: BVH is not resident on the context's device %d",
                    ctx->device);
    if (n && !rays) return fail(RUMI_ERR_INVALID, "trace: null ray buffer");
    if (((uintptr_t)rays & 15u) != 0) return fail(RUMI_ERR_INVALID, "trace: rays must be 16-byte aligned");
    if (flags & ~3u) return fail(RUMI_ERR_INVALID, "trace: unknown flags 0x%x", flags);
    if ((flags & RUMI_TRACE_REF_ORDER) && (flags & RUMI_TRACE_ORDERED))
        return fail(RUMI_ERR_INVALID, "trace: REF_ORDER and ORDERED are exclusive");
    if (flags & RUMI_TRACE_ORDERED)
        *ordered = true;
    else if (flags & RUMI_TRACE_REF_ORDER)
        *ordered = false;
    else
        *ordered = bvh->mode == RUMI_BVH_SAH_BINNED;
    if (hipSetDevice(ctx->device) != hipSuccess)
        return fail(RUMI_ERR_HIP, "hipSetDevice(%d) failed", ctx->device);
    return RUMI_OK;
}

TraceArgs base_args(const rumi_bvh* bvh, const rumi_ray* rays, uint32_t n) {
    TraceArgs a{};
    a.rays = reinterpret_cast<const float4*>(rays);
    a.n = n;
    a.nodes = reinterpret_cast<const float4*>(bvh->dev_nodes);
    a.tris = reinterpret_cast<const float4*>(bvh->dev_tris);
    for (int k = 0; k < 3; ++k) {
        a.root.lo[k] = bvh->flat.root_lo[k];
        a.root.hi[k] = bvh->flat.root_hi[k];
    }
    a.root.ref = bvh->flat.root_ref;
    return a;
}

int finish(hipError_t e, const char* what) {
    if (e != hipSuccess) return fail(RUMI_ERR_HIP, "%s launch failed: %s", what, hipGetErrorString(e));
    return RUMI_OK;
}

}  // namespace

extern "C" int rumi_trace_closest(rumi_ctx* ctx, const rumi_bvh* bvh, const rumi_ray* rays,
                                  uint32_t n, rumi_hit* hits, float* uv, uint32_t flags,
                                  void* stream) {
    clear_error();
    bool ordered = false;
    if (int rc = check_common(ctx, bvh, rays, n, flags, &ordered)) return rc;
    if (n && !hits) return fail(RUMI_ERR_INVALID, "rumi_trace_closest: null hit buffer");
    if (((uintptr_t)hits & 7u) || ((uintptr_t)uv & 7u))
        return fail(RUMI_ERR_INVALID, "rumi_trace_closest: hits and uv must be 8-byte aligned");
    if (n == 0) return RUMI_OK;
    TraceArgs a = base_args(bvh, rays, n);
    a.hits = reinterpret_cast<int2*>(hits);
    a.uv = reinterpret_cast<float2*>(uv);
    return finish(launch_order<false, false>(a, ordered, bvh->flat.max_depth, (hipStream_t)stream),
                  "rumi_trace_closest");
}

extern "C" int rumi_trace_any(rumi_ctx* ctx, const rumi_bvh* bvh, const rumi_ray* rays,
                              uint32_t n, int32_t* prim, uint32_t flags, void* stream) {
    clear_error();
    bool ordered = false;
    if (int rc = check_common(ctx, bvh, rays, n, flags, &ordered)) return rc;
    if (n && !prim) return fail(RUMI_ERR_INVALID, "rumi_trace_any: null output");
    if (n == 0) return RUMI_OK;
    TraceArgs a = base_args(bvh, rays, n);
    a.any_out = prim;
    return finish(launch_order<true, false>(a, ordered, bvh->flat.max_depth, (hipStream_t)stream),
                  "rumi_trace_any");
}

extern "C" int rumi_trace_count(rumi_ctx* ctx, const rumi_bvh* bvh, const rumi_ray* rays,
                                uint32_t n, int32_t any, rumi_hit* hits, int32_t* node_count,
                                int32_t* prim_count, uint32_t flags, void* stream) {
    clear_error();
    bool ordered = false;
    if (int rc = check_common(ctx, bvh, rays, n, flags, &ordered)) return rc;
    if (n && (!hits || !node_count || !prim_count))
        return fail(RUMI_ERR_INVALID, "rumi_trace_count: null output");
    if ((uintptr_t)hits & 7u) return fail(RUMI_ERR_INVALID, "rumi_trace_count: hits must be 8-byte aligned");
    if (n == 0) return RUMI_OK;
    TraceArgs a = base_args(bvh, rays, n);
    a.cnt_nodes = node_count;
    a.cnt_prims = prim_count;
    a.hits = reinterpret_cast<int2*>(hits);
    const hipError_t e =
        any ? launch_order<true, true>(a, ordered, bvh->flat.max_depth, (hipStream_t)stream)
            : launch_order<false, true>(a, ordered, bvh->flat.max_depth, (hipStream_t)stream);
    return finish(e, "rumi_trace_count");
}
