// Ray-query kernels for gfx950 (closest hit, any hit, instrumented) and their
// C-ABI launchers.  One ray per lane (wave64); the traversal itself is the
// shared device routine in traverse.h.
#include <hip/hip_runtime.h>

#include "internal.h"
#include "traverse.h"

using namespace rumi;
using namespace rumi::dev;

namespace {

template <int BLOCK, int STACK, bool ANY, bool ORDERED, bool COUNT, bool SPILL>
__global__ __launch_bounds__(BLOCK) void k_trace(
    const float4* __restrict__ rays, uint32_t n, const float4* __restrict__ nodes,
    const float4* __restrict__ tris, RootInfo root, int2* __restrict__ hits,
    float2* __restrict__ uv, int32_t* __restrict__ any_out, int32_t* __restrict__ cnt_nodes,
    int32_t* __restrict__ cnt_prims) {
    __shared__ int32_t s_ref[STACK * BLOCK];
    __shared__ float s_tn[STACK * BLOCK];
    const uint32_t tid = threadIdx.x;
    const uint32_t i = blockIdx.x * BLOCK + tid;
    if (i >= n) return;
    const float4 r0 = rays[2 * (size_t)i + 0];
    const float4 r1 = rays[2 * (size_t)i + 1];
    Ray r{r0.x, r0.y, r0.z, r1.x, r1.y, r1.z, r1.w, r0.w};
    float hu = 0.0f, hv = 0.0f;
    int32_t nn = 0, np = 0;
    const int32_t hit = traverse<BLOCK, STACK, ANY, ORDERED, false, COUNT, SPILL>(
        nodes, tris, nullptr, root, r, false, hu, hv, nn, np, s_ref, s_tn, tid);
    if (ANY && any_out) any_out[i] = hit;
    if (hits) hits[i] = make_int2(hit, __float_as_int(r.tmin));
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
