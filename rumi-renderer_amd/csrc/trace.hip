// Ray-query kernels for gfx950 (closest hit, any hit, instrumented) and their
// C-ABI launchers.  One ray per lane (wave64); the traversal itself is the
// shared device routine in traverse.h.
#include <hip/hip_runtime.h>

#include "internal.h"
#include "traverse.h"

using namespace rumi;
using namespace rumi::dev;

namespace {

template <int BLOCK, int STACK, bool ANY, bool ORDERED, bool COUNT, bool SPILL, bool WIDE = false>
__global__ __launch_bounds__(BLOCK) void k_trace(
    const float4* __restrict__ rays, uint32_t n, const float4* __restrict__ nodes,
    const float4* __restrict__ tris, RootInfo root, int2* __restrict__ hits,
    float2* __restrict__ uv, int32_t* __restrict__ any_out, int32_t* __restrict__ cnt_nodes,
    int32_t* __restrict__ cnt_prims) {
    __shared__ int2 s_stk[STACK * BLOCK];
    const uint32_t tid = threadIdx.x;
    const uint32_t i = blockIdx.x * BLOCK + tid;
    if (i >= n) return;
    const float4 r0 = rays[2 * (size_t)i + 0];
    const float4 r1 = rays[2 * (size_t)i + 1];
    Ray r{r0.x, r0.y, r0.z, r1.x, r1.y, r1.z, r1.w, r0.w};
    float hu = 0.0f, hv = 0.0f;
    int32_t nn = 0, np = 0;
    const int32_t hit = traverse<BLOCK, STACK, ANY, ORDERED, false, COUNT, SPILL, WIDE>(
        nodes, tris, nullptr, root, r, false, hu, hv, nn, np, s_stk, tid);
    if (ANY && any_out) any_out[i] = hit;
    if (hits) hits[i] = make_int2(hit, __float_as_int(r.tmin));
    if (!ANY && uv) uv[i] = make_float2(hu, hv);
    if (COUNT) {
        cnt_nodes[i] = nn;
        cnt_prims[i] = np;
    }
}

__device__ __forceinline__ uint32_t lane_below(uint64_t mask) {
    // set bits of mask in the lanes below this one
    return __builtin_amdgcn_mbcnt_hi((uint32_t)(mask >> 32),
                                     __builtin_amdgcn_mbcnt_lo((uint32_t)mask, 0u));
}

// Persistent-lane traversal.  Rays differ in cost by two orders of magnitude
// (half of a uniform batch misses the root box; the 99th percentile visits
// hundreds of nodes), so one ray per lane leaves a wave waiting on its
// slowest lane.  Here every lane steps its own ray one node or leaf at a
// time and, once enough lanes of the wave are idle, the wave takes that many
// new ray indices from a global counter with one atomic; a lane whose new
// ray misses the root box writes the miss at once and stays idle.  The wave
// leaves when the batch is drained and its last ray is done.  Each ray's
// arithmetic is the same as in k_trace, so the results are bit-identical.
template <int BLOCK, int STACK, bool ANY, bool ORDERED, bool SPILL>
__global__ __launch_bounds__(BLOCK) void k_trace_lanes(
    const float4* __restrict__ rays, uint32_t n, const float4* __restrict__ nodes,
    const float4* __restrict__ tris, RootInfo root, int2* __restrict__ hits,
    float2* __restrict__ uv, int32_t* __restrict__ any_out, uint32_t* __restrict__ counter,
    int refill) {
    __shared__ int2 s_stk[STACK * BLOCK];
    constexpr int kSpill = SPILL ? kMaxDepth - STACK : 1;
    int2 p_stk[kSpill];
    const uint32_t tid = threadIdx.x;
    const uint32_t lane = tid & 63u;
    TravState s;
    uint32_t id = 0;
    bool active = false, drained = false;
    int32_t nn = 0, np = 0;
    auto finish_ray = [&](uint32_t k) {
        if (ANY && any_out) any_out[k] = s.hit;
        if (hits) hits[k] = make_int2(s.hit, __float_as_int(s.r.tmin));
        if (!ANY && uv) uv[k] = make_float2(s.hu, s.hv);
    };
    while (true) {
        const uint64_t want = __ballot(!active && !drained);
        const uint64_t busy = __ballot(active);
        if (want == 0 && busy == 0) break;
        if (want != 0 && (busy == 0 || __popcll(want) >= (uint32_t)refill)) {
            const uint32_t leader = (uint32_t)__ffsll((long long)want) - 1u;
            uint32_t base = 0;
            if (lane == leader) base = atomicAdd(counter, (uint32_t)__popcll(want));
            base = __shfl(base, (int)leader);
            if (!active && !drained) {
                const uint32_t k = base + lane_below(want);
                if (k >= n) {
                    drained = true;
                } else {
                    const float4 r0 = rays[2 * (size_t)k + 0];
                    const float4 r1 = rays[2 * (size_t)k + 1];
                    s.r = Ray{r0.x, r0.y, r0.z, r1.x, r1.y, r1.z, r1.w, r0.w};
                    id = k;
                    active = trav_begin(root, s);
                    if (!active) finish_ray(k);
                }
            }
        }
        if (active) {
            active = trav_step<BLOCK, STACK, ANY, ORDERED, false, false, SPILL>(
                nodes, tris, nullptr, s, false, nn, np, s_stk, tid, p_stk);
            if (!active) finish_ray(id);
        }
    }
}

// Persistent lanes with a prefetched next ray per lane (run_lanes).
template <int BLOCK, int STACK, bool ANY, bool ORDERED, bool SPILL, bool WIDE>
struct QueryItem {
    const float4* rays;
    const float4* nodes;
    const float4* tris;
    RootInfo root;
    int2* hits;
    float2* uv;
    int32_t* any_out;
    int2* s_stk;
    int2* p_stk;
    uint8_t* slots;  // the wave's 256 int2 for spread triangle tests
    uint32_t tid, id, nid;
    float4 n0, n1;
    // the 4-wide mesh step runs on the whole wave, its triangle tests spread
    // over the lanes (1.92 -> 1.87 ms closest-hit, 1.37 -> 1.32 ms any-hit on
    // the 1.3M-triangle mesh, with four triangles per step)
    static constexpr bool kWaveStep = WIDE && ORDERED;
    __device__ void prefetch(uint32_t k) {
        n0 = rays[2 * (size_t)k + 0];
        n1 = rays[2 * (size_t)k + 1];
        nid = k;
    }
    __device__ bool start(TravState& s) {
        s.r = Ray{n0.x, n0.y, n0.z, n1.x, n1.y, n1.z, n1.w, n0.w};
        id = nid;
        return trav_begin(root, s);
    }
    __device__ bool step(TravState& s) {
        int32_t nn = 0, np = 0;
        return trav_step<BLOCK, STACK, ANY, ORDERED, false, false, SPILL, WIDE>(
            nodes, tris, nullptr, s, false, nn, np, s_stk, tid, p_stk);
    }
    __device__ bool step(TravState& s, bool live) {
        int32_t nn = 0, np = 0;
        return trav_step_near4<BLOCK, STACK, ANY, false, false, SPILL, true>(
            nodes, tris, nullptr, s, false, nn, np, s_stk, tid, p_stk, live, slots);
    }
    __device__ void finish(const TravState& s) {
        if (ANY && any_out) any_out[id] = s.hit;
        if (hits) hits[id] = make_int2(s.hit, __float_as_int(s.r.tmin));
        if (!ANY && uv) uv[id] = make_float2(s.hu, s.hv);
    }
};

template <int BLOCK, int STACK, bool ANY, bool ORDERED, bool SPILL, bool WIDE = false>
__global__ __launch_bounds__(BLOCK) void k_trace_prefetch(
    const float4* __restrict__ rays, uint32_t n, const float4* __restrict__ nodes,
    const float4* __restrict__ tris, RootInfo root, int2* __restrict__ hits,
    float2* __restrict__ uv, int32_t* __restrict__ any_out, uint32_t* __restrict__ counter) {
    __shared__ int2 s_stk[STACK * BLOCK];
    __shared__ int2 s_slots[QueryItem<BLOCK, STACK, ANY, ORDERED, SPILL, WIDE>::kWaveStep ? BLOCK * 4 : 1];
    constexpr int kSpill = SPILL ? kMaxDepth - STACK : 1;
    int2 p_stk[kSpill];
    QueryItem<BLOCK, STACK, ANY, ORDERED, SPILL, WIDE> item;
    item.slots = reinterpret_cast<uint8_t*>(
        s_slots + (QueryItem<BLOCK, STACK, ANY, ORDERED, SPILL, WIDE>::kWaveStep ? (threadIdx.x >> 6) * 256u : 0u));
    item.rays = rays;
    item.nodes = nodes;
    item.tris = tris;
    item.root = root;
    item.hits = hits;
    item.uv = uv;
    item.any_out = any_out;
    item.s_stk = s_stk;
    item.p_stk = p_stk;
    item.tid = threadIdx.x;
    item.id = item.nid = 0;
    run_lanes(n, counter, item);
}

__global__ void k_zero_counter(uint32_t* c) { *c = 0u; }

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
    uint32_t* counter;  // persistent launches: the work counter slot
    int grid_cap;       // persistent launches: compute units
    int refill;         // persistent launches: idle lanes that trigger a refill
    int block;          // persistent launches: threads per block (64, 128, 256)
    int stack;          // persistent launches: LDS stack entries per lane (8, 16)
    bool persistent;
    int persistent_kind;  // 1 refill on a threshold, 2 prefetched next ray
    bool wide;            // nodes is the 4-wide tree (near-first order only)
    int depth;            // stack entries a traversal can need
};

constexpr int kBlock = 128;

// One ray per lane (the instrumented launches, and tuning.persistent < 0).
template <int STACK, bool ANY, bool ORDERED, bool COUNT, bool SPILL, bool WIDE = false>
hipError_t launch_rays(const TraceArgs& a, hipStream_t s) {
    const uint32_t grid = (a.n + kBlock - 1) / kBlock;
    hipLaunchKernelGGL((k_trace<kBlock, STACK, ANY, ORDERED, COUNT, SPILL, WIDE>), dim3(grid),
                       dim3(kBlock), 0, s, a.rays, a.n, a.nodes, a.tris, a.root, a.hits, a.uv,
                       a.any_out, a.cnt_nodes, a.cnt_prims);
    return hipGetLastError();
}

// Persistent lanes: as many blocks as stay resident, refilled from a counter.
// The 4-wide tree runs with prefetching lanes only.
template <int BLOCK, int STACK, bool ANY, bool ORDERED, bool SPILL, bool WIDE = false>
hipError_t launch_lanes(const TraceArgs& a, hipStream_t s) {
    const uint32_t grid = (a.n + BLOCK - 1) / BLOCK;
    int per_cu = 0;
    hipLaunchKernelGGL(k_zero_counter, dim3(1), dim3(1), 0, s, a.counter);
    if (WIDE || a.persistent_kind == 2) {
        auto kern = k_trace_prefetch<BLOCK, STACK, ANY, ORDERED, SPILL, WIDE>;
        if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, kern, BLOCK, 0) != hipSuccess ||
            per_cu < 1)
            per_cu = 1;
        const uint32_t resident = (uint32_t)per_cu * (uint32_t)a.grid_cap;
        hipLaunchKernelGGL(kern, dim3(grid < resident ? grid : resident), dim3(BLOCK), 0, s, a.rays,
                           a.n, a.nodes, a.tris, a.root, a.hits, a.uv, a.any_out, a.counter);
    } else if constexpr (!WIDE) {
        auto kern = k_trace_lanes<BLOCK, STACK, ANY, ORDERED, SPILL>;
        if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, kern, BLOCK, 0) != hipSuccess ||
            per_cu < 1)
            per_cu = 1;
        const uint32_t resident = (uint32_t)per_cu * (uint32_t)a.grid_cap;
        hipLaunchKernelGGL(kern, dim3(grid < resident ? grid : resident), dim3(BLOCK), 0, s, a.rays,
                           a.n, a.nodes, a.tris, a.root, a.hits, a.uv, a.any_out, a.counter,
                           a.refill);
    }
    return hipGetLastError();
}

template <int BLOCK, int STACK, bool ANY, bool ORDERED, bool WIDE = false>
hipError_t lanes_depth(const TraceArgs& a, int depth, hipStream_t s) {
    return depth <= STACK ? launch_lanes<BLOCK, STACK, ANY, ORDERED, false, WIDE>(a, s)
                          : launch_lanes<BLOCK, STACK, ANY, ORDERED, true, WIDE>(a, s);
}

template <int BLOCK, bool ANY, bool ORDERED, bool WIDE = false>
hipError_t lanes_stack(const TraceArgs& a, int depth, hipStream_t s) {
    return a.stack == 8 ? lanes_depth<BLOCK, 8, ANY, ORDERED, WIDE>(a, depth, s)
                        : lanes_depth<BLOCK, 16, ANY, ORDERED, WIDE>(a, depth, s);
}

template <bool ANY, bool ORDERED, bool COUNT>
hipError_t launch_depth(const TraceArgs& a, int depth, hipStream_t s) {
    if constexpr (ORDERED) {
        if (a.wide) {
            if (!COUNT && a.persistent) {
                if (a.block == 64) return lanes_stack<64, ANY, true, true>(a, depth, s);
                if (a.block == 256) return lanes_stack<256, ANY, true, true>(a, depth, s);
                return lanes_stack<128, ANY, true, true>(a, depth, s);
            }
            if (depth <= 16) return launch_rays<16, ANY, true, COUNT, false, true>(a, s);
            if (depth <= 32) return launch_rays<32, ANY, true, COUNT, false, true>(a, s);
            return launch_rays<32, ANY, true, COUNT, true, true>(a, s);
        }
    }
    if (!COUNT && a.persistent) {
        if (a.block == 64) return lanes_stack<64, ANY, ORDERED>(a, depth, s);
        if (a.block == 256) return lanes_stack<256, ANY, ORDERED>(a, depth, s);
        return lanes_stack<128, ANY, ORDERED>(a, depth, s);
    }
    if (depth <= 16) return launch_rays<16, ANY, ORDERED, COUNT, false>(a, s);
    if (depth <= 32) return launch_rays<32, ANY, ORDERED, COUNT, false>(a, s);
    return launch_rays<32, ANY, ORDERED, COUNT, true>(a, s);
}

template <bool ANY, bool COUNT>
hipError_t launch_order(const TraceArgs& a, bool ordered, hipStream_t s) {
    return ordered ? launch_depth<ANY, true, COUNT>(a, a.depth, s)
                   : launch_depth<ANY, false, COUNT>(a, a.depth, s);
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
    if (n > 0x7fffffffu) return fail(RUMI_ERR_INVALID, "trace: more than 2^31 - 1 rays");
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

// ordered: near-first order (the 4-wide tree is used when the context asks
// for it and the tree has one).
TraceArgs base_args(rumi_ctx* ctx, const rumi_bvh* bvh, const rumi_ray* rays, uint32_t n,
                    bool ordered) {
    TraceArgs a{};
    a.counter = ctx->take_counter();
    a.grid_cap = ctx->cu_count;
    a.refill = ctx->tuning.refill;
    a.block = ctx->tuning.block;
    a.stack = ctx->tuning.stack;
    a.persistent = ctx->tuning.persistent >= 0;
    a.persistent_kind = ctx->tuning.persistent;
    a.rays = reinterpret_cast<const float4*>(rays);
    a.n = n;
    a.nodes = reinterpret_cast<const float4*>(bvh->dev_nodes);
    a.tris = reinterpret_cast<const float4*>(bvh->dev_tris);
    for (int k = 0; k < 3; ++k) {
        a.root.lo[k] = bvh->flat.root_lo[k];
        a.root.hi[k] = bvh->flat.root_hi[k];
    }
    a.root.ref = bvh->flat.root_ref;
    a.depth = bvh->flat.max_depth;
    a.wide = ordered && ctx->tuning.width == 4 && bvh->wide_ok &&
             (bvh->dev_nodes4 != nullptr || bvh->wide.nodes.empty());
    if (a.wide) {
        a.nodes = reinterpret_cast<const float4*>(bvh->dev_nodes4);
        a.root.ref = bvh->wide.root_ref;
        a.depth = 3 * bvh->wide.max_depth;
    }
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
    TraceArgs a = base_args(ctx, bvh, rays, n, ordered);
    a.hits = reinterpret_cast<int2*>(hits);
    a.uv = reinterpret_cast<float2*>(uv);
    return finish(launch_order<false, false>(a, ordered, (hipStream_t)stream),
                  "rumi_trace_closest");
}

extern "C" int rumi_trace_any(rumi_ctx* ctx, const rumi_bvh* bvh, const rumi_ray* rays,
                              uint32_t n, int32_t* prim, uint32_t flags, void* stream) {
    clear_error();
    bool ordered = false;
    if (int rc = check_common(ctx, bvh, rays, n, flags, &ordered)) return rc;
    if (n && !prim) return fail(RUMI_ERR_INVALID, "rumi_trace_any: null output");
    if (n == 0) return RUMI_OK;
    TraceArgs a = base_args(ctx, bvh, rays, n, ordered);
    a.any_out = prim;
    return finish(launch_order<true, false>(a, ordered, (hipStream_t)stream),
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
    TraceArgs a = base_args(ctx, bvh, rays, n, ordered);
    a.cnt_nodes = node_count;
    a.cnt_prims = prim_count;
    a.hits = reinterpret_cast<int2*>(hits);
    const hipError_t e =
        any ? launch_order<true, true>(a, ordered, (hipStream_t)stream)
            : launch_order<false, true>(a, ordered, (hipStream_t)stream);
    return finish(e, "rumi_trace_count");
}

extern "C" int rumi_trace_set_tuning(rumi_ctx* ctx, const rumi_trace_tuning* t) {
    clear_error();
    if (!ctx) return fail(RUMI_ERR_NO_DEVICE, "rumi_trace_set_tuning: null context");
    rumi_trace_tuning v = {256, 16, 16, 2, 4};
    if (t) {
        if (t->block) v.block = t->block;
        if (t->stack) v.stack = t->stack;
        if (t->refill) v.refill = t->refill;
        if (t->persistent) v.persistent = t->persistent;
        if (t->width) v.width = t->width;
    }
    if (v.width != 2 && v.width != 4)
        return fail(RUMI_ERR_INVALID, "rumi_trace_set_tuning: width %d not 2 or 4", v.width);
    if (v.block != 64 && v.block != 128 && v.block != 256)
        return fail(RUMI_ERR_INVALID, "rumi_trace_set_tuning: block %d not 64, 128 or 256", v.block);
    if (v.stack != 8 && v.stack != 16)
        return fail(RUMI_ERR_INVALID, "rumi_trace_set_tuning: stack %d not 8 or 16", v.stack);
    if (v.refill < 1 || v.refill > 64)
        return fail(RUMI_ERR_INVALID, "rumi_trace_set_tuning: refill %d not in [1, 64]", v.refill);
    if (v.persistent != 1 && v.persistent != 2 && v.persistent != -1)
        return fail(RUMI_ERR_INVALID, "rumi_trace_set_tuning: persistent %d not 1, 2 or -1",
                    v.persistent);
    ctx->tuning = v;
    return RUMI_OK;
}

extern "C" int rumi_trace_get_tuning(const rumi_ctx* ctx, rumi_trace_tuning* t) {
    clear_error();
    if (!ctx || !t) return fail(RUMI_ERR_INVALID, "rumi_trace_get_tuning: null argument");
    *t = ctx->tuning;
    return RUMI_OK;
}
