// C ABI: contexts and BVH handles (host side).
#include <hip/hip_runtime_api.h>

#include <chrono>
#include <cstring>
#include <new>

#include "internal.h"

using namespace rumi;

#define HIP_TRY(expr)                                                                     \
    do {                                                                                  \
        hipError_t e_ = (expr);                                                           \
        if (e_ != hipSuccess)                                                             \
            return fail(RUMI_ERR_HIP, "%s failed: %s", #expr, hipGetErrorString(e_));      \
    } while (0)

extern "C" int rumi_ctx_create(int device, rumi_ctx** out) {
    clear_error();
    if (!out) return fail(RUMI_ERR_INVALID, "rumi_ctx_create: null output");
    *out = nullptr;
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess || n <= 0)
        return fail(RUMI_ERR_NO_DEVICE, "rumi_ctx_create: no HIP device available");
    if (device < 0 || device >= n)
        return fail(RUMI_ERR_INVALID, "rumi_ctx_create: device %d not in [0,%d)", device, n);
    HIP_TRY(hipSetDevice(device));
    hipDeviceProp_t prop;
    HIP_TRY(hipGetDeviceProperties(&prop, device));
    rumi_ctx* c = new (std::nothrow) rumi_ctx;
    if (!c) return fail(RUMI_ERR_NO_MEMORY, "rumi_ctx_create: out of host memory");
    c->device = device;
    c->cu_count = prop.multiProcessorCount;
    if (hipMalloc(&c->dev_counters, sizeof(uint32_t) * rumi_ctx::kCounterSlots) != hipSuccess) {
        delete c;
        return fail(RUMI_ERR_NO_MEMORY, "rumi_ctx_create: hipMalloc of the counter slots failed");
    }
    *out = c;
    return RUMI_OK;
}

extern "C" int rumi_ctx_destroy(rumi_ctx* ctx) {
    if (ctx && ctx->dev_counters) {
        int prev = 0;
        (void)hipGetDevice(&prev);
        (void)hipSetDevice(ctx->device);
        (void)hipFree(ctx->dev_counters);
        (void)hipSetDevice(prev);
    }
    delete ctx;
    return RUMI_OK;
}

static void release_device(rumi_bvh* b) {
    if (b->device >= 0) {
        int prev = 0;
        (void)hipGetDevice(&prev);
        (void)hipSetDevice(b->device);
        if (b->dev_nodes) (void)hipFree(b->dev_nodes);
        if (b->dev_tris) (void)hipFree(b->dev_tris);
        if (b->dev_nodes4) (void)hipFree(b->dev_nodes4);
        (void)hipSetDevice(prev);
    }
    b->dev_nodes = b->dev_tris = b->dev_nodes4 = nullptr;
    b->device = -1;
}

extern "C" int rumi_bvh_build_mesh(rumi_ctx* ctx, const float* xyz, uint32_t nv,
                                   const uint32_t* idx, uint32_t nt,
                                   const rumi_bvh_params* params, rumi_bvh** out) {
    clear_error();
    if (!out) return fail(RUMI_ERR_INVALID, "rumi_bvh_build_mesh: null output");
    *out = nullptr;
    if (nt && (!xyz || !idx)) return fail(RUMI_ERR_INVALID, "rumi_bvh_build_mesh: null mesh arrays");
    if (nt > (uint32_t)kMaxLeafFirst)
        return fail(RUMI_ERR_INVALID, "rumi_bvh_build_mesh: %u triangles exceed %d", nt, kMaxLeafFirst);
    rumi_bvh_params p = {RUMI_BVH_SAH_BINNED, 4, 32, 0};
    if (params) p = *params;
    if (p.mode != RUMI_BVH_SAH_BINNED && p.mode != RUMI_BVH_REF_MEDIAN6)
        return fail(RUMI_ERR_INVALID, "rumi_bvh_build_mesh: unknown mode %d", p.mode);
    if (p.mode == RUMI_BVH_SAH_BINNED) {
        if (p.max_leaf == 0) p.max_leaf = 4;
        if (p.n_bins == 0) p.n_bins = 32;
        if (p.max_leaf < 1 || p.max_leaf > kMaxLeafCount)
            return fail(RUMI_ERR_INVALID, "max_leaf %d not in [1,%d]", p.max_leaf, kMaxLeafCount);
        if (p.n_bins < 4 || p.n_bins > 256)
            return fail(RUMI_ERR_INVALID, "n_bins %d not in [4,256]", p.n_bins);
    }
    for (uint64_t i = 0; i < 3ull * nt; ++i)
        if (idx[i] >= nv)
            return fail(RUMI_ERR_INVALID, "rumi_bvh_build_mesh: index %u of triangle %llu >= nv %u",
                        idx[i], (unsigned long long)(i / 3), nv);

    auto t0 = std::chrono::steady_clock::now();
    // Per-primitive bounds with Triangle::Bound's selections
    // (src/Shape/triangle.cpp:120-148).
    std::vector<float> lo(3 * (size_t)nt), hi(3 * (size_t)nt);
    for (uint32_t t = 0; t < nt; ++t) {
        const float* v0 = xyz + 3 * (size_t)idx[3 * (size_t)t + 0];
        const float* v1 = xyz + 3 * (size_t)idx[3 * (size_t)t + 1];
        const float* v2 = xyz + 3 * (size_t)idx[3 * (size_t)t + 2];
        for (int k = 0; k < 3; ++k) {
            float mn = v0[k] < v1[k] ? v0[k] : v1[k];
            if (mn > v2[k]) mn = v2[k];
            float mx = v0[k] > v1[k] ? v0[k] : v1[k];
            if (mx < v2[k]) mx = v2[k];
            lo[3 * (size_t)t + k] = mn;
            hi[3 * (size_t)t + k] = mx;
        }
    }
    BuildResult br;
    if (p.mode == RUMI_BVH_SAH_BINNED)
        build_sah(lo.data(), hi.data(), (int)nt, p.max_leaf, p.n_bins, &br);
    else
        build_median6(lo.data(), hi.data(), (int)nt, &br);

    rumi_bvh* b = new (std::nothrow) rumi_bvh;
    if (!b) return fail(RUMI_ERR_NO_MEMORY, "rumi_bvh_build_mesh: out of host memory");
    flatten(br, &b->flat);
    if (b->flat.max_depth > kMaxDepth) {
        const int depth = b->flat.max_depth;
        delete b;
        return fail(RUMI_ERR_UNSUPPORTED, "BVH depth %d exceeds the traversal limit %d", depth,
                    kMaxDepth);
    }
    b->mode = p.mode;
    b->n_prims = (int32_t)nt;
    b->sah_cost = br.sah_cost;
    if (p.mode == RUMI_BVH_SAH_BINNED) {
        widen(b->flat, &b->wide);
        b->wide_ok = 3 * b->wide.max_depth <= kMaxDepth;
    }
    b->host_tris.resize(nt);
    for (uint32_t s = 0; s < nt; ++s) {
        const int32_t prim = b->flat.leaf_prims[s];
        const float* v0 = xyz + 3 * (size_t)idx[3 * (size_t)prim + 0];
        const float* v1 = xyz + 3 * (size_t)idx[3 * (size_t)prim + 1];
        const float* v2 = xyz + 3 * (size_t)idx[3 * (size_t)prim + 2];
        DevTri& d = b->host_tris[s];
        const float e1x = v1[0] - v0[0], e1y = v1[1] - v0[1], e1z = v1[2] - v0[2];
        const float e2x = v2[0] - v0[0], e2y = v2[1] - v0[1], e2z = v2[2] - v0[2];
        d.a[0] = v0[0]; d.a[1] = v0[1]; d.a[2] = v0[2]; d.a[3] = e1x;
        d.b[0] = e1y;   d.b[1] = e1z;   d.b[2] = e2x;   d.b[3] = e2y;
        d.c[0] = e2z;
        std::memcpy(&d.c[1], &prim, sizeof(int32_t));
        d.c[2] = 0.0f;
        d.c[3] = 0.0f;
    }
    b->build_seconds =
        std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();

    if (ctx) {
        int rc = RUMI_OK;
        do {
            if (hipSetDevice(ctx->device) != hipSuccess) {
                rc = fail(RUMI_ERR_HIP, "hipSetDevice(%d) failed", ctx->device);
                break;
            }
            b->device = ctx->device;
            const size_t nb = sizeof(DevNode) * b->flat.nodes.size();
            const size_t tb = sizeof(DevTri) * b->host_tris.size();
            if (nb) {
                if (hipMalloc(&b->dev_nodes, nb) != hipSuccess) {
                    rc = fail(RUMI_ERR_NO_MEMORY, "hipMalloc of %zu node bytes failed", nb);
                    break;
                }
                if (hipMemcpy(b->dev_nodes, b->flat.nodes.data(), nb, hipMemcpyHostToDevice) != hipSuccess) {
                    rc = fail(RUMI_ERR_HIP, "hipMemcpy of BVH nodes failed");
                    break;
                }
            }
            if (tb) {
                if (hipMalloc(&b->dev_tris, tb) != hipSuccess) {
                    rc = fail(RUMI_ERR_NO_MEMORY, "hipMalloc of %zu triangle bytes failed", tb);
                    break;
                }
                if (hipMemcpy(b->dev_tris, b->host_tris.data(), tb, hipMemcpyHostToDevice) != hipSuccess) {
                    rc = fail(RUMI_ERR_HIP, "hipMemcpy of triangles failed");
                    break;
                }
            }
            const size_t wb = sizeof(DevNode4) * b->wide.nodes.size();
            if (b->wide_ok && wb) {
                if (hipMalloc(&b->dev_nodes4, wb) != hipSuccess) {
                    rc = fail(RUMI_ERR_NO_MEMORY, "hipMalloc of %zu wide-node bytes failed", wb);
                    break;
                }
                if (hipMemcpy(b->dev_nodes4, b->wide.nodes.data(), wb, hipMemcpyHostToDevice) != hipSuccess) {
                    rc = fail(RUMI_ERR_HIP, "hipMemcpy of wide BVH nodes failed");
                    break;
                }
            }
        } while (0);
        if (rc != RUMI_OK) {
            release_device(b);
            delete b;
            return rc;
        }
    }
    *out = b;
    return RUMI_OK;
}

extern "C" int rumi_bvh_destroy(rumi_bvh* bvh) {
    if (!bvh) return RUMI_OK;
    release_device(bvh);
    delete bvh;
    return RUMI_OK;
}

extern "C" int rumi_bvh_get_info(const rumi_bvh* b, rumi_bvh_info* info) {
    clear_error();
    if (!b || !info) return fail(RUMI_ERR_INVALID, "rumi_bvh_get_info: null argument");
    info->mode = b->mode;
    info->n_prims = b->n_prims;
    info->n_inner = (int32_t)b->flat.nodes.size();
    info->n_leaves = b->flat.n_leaves;
    info->max_depth = b->flat.max_depth;
    info->on_device = b->device >= 0 ? 1 : 0;
    info->sah_cost = b->sah_cost;
    info->build_seconds = b->build_seconds;
    return RUMI_OK;
}

extern "C" int rumi_bvh_export(const rumi_bvh* b, float* child_boxes, int32_t* child_refs,
                               float* root_box, int32_t* root_ref, int32_t* leaf_prims) {
    clear_error();
    if (!b) return fail(RUMI_ERR_INVALID, "rumi_bvh_export: null BVH");
    const size_t n = b->flat.nodes.size();
    for (size_t i = 0; i < n; ++i) {
        const DevNode& d = b->flat.nodes[i];
        if (child_boxes) {
            float* o = child_boxes + 12 * i;
            o[0] = d.a[0]; o[1] = d.a[1]; o[2] = d.a[2];
            o[3] = d.a[3]; o[4] = d.b[0]; o[5] = d.b[1];
            o[6] = d.b[2]; o[7] = d.b[3]; o[8] = d.c[0];
            o[9] = d.c[1]; o[10] = d.c[2]; o[11] = d.c[3];
        }
        if (child_refs) {
            child_refs[2 * i + 0] = d.d[0];
            child_refs[2 * i + 1] = d.d[1];
        }
    }
    if (root_box) {
        std::memcpy(root_box, b->flat.root_lo, sizeof(float) * 3);
        std::memcpy(root_box + 3, b->flat.root_hi, sizeof(float) * 3);
    }
    if (root_ref) *root_ref = b->flat.root_ref;
    if (leaf_prims && !b->flat.leaf_prims.empty())
        std::memcpy(leaf_prims, b->flat.leaf_prims.data(), sizeof(int32_t) * b->flat.leaf_prims.size());
    return RUMI_OK;
}
