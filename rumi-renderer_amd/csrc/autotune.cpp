// Launch-shape autotuner of the C ABI (rumi_autotune_*): times candidate
// settings of the traversal kernels and of the renderer's queue sizes on
// the caller's own workload with HIP events and keeps the fastest.  Every
// traversal candidate returns the same hit distances (the same bits within
// one node width) and every render candidate the same image up to the order
// of float atomics, so choosing among them changes only the time.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstring>
#include <vector>

#include "internal.h"

using namespace rumi;

namespace {

// The traversal candidates, in the order rumi_autotune_trace reports them:
// one ray per lane; prefetching lanes for every width, block and LDS stack;
// threshold refill (binary nodes) for every block, stack and threshold.
std::vector<rumi_trace_tuning> trace_candidates() {
    std::vector<rumi_trace_tuning> out;
    out.push_back({128, 16, 16, -1, 2});
    const int blocks[] = {64, 128, 256}, stacks[] = {8, 16}, refills[] = {16, 32, 48, 64}, widths[] = {2, 4};
    for (int w : widths)
        for (int b : blocks)
            for (int s : stacks) out.push_back({b, s, 16, 2, w});
    for (int b : blocks)
        for (int s : stacks)
            for (int r : refills) out.push_back({b, s, r, 1, 2});
    return out;
}

struct Events {
    hipEvent_t a = nullptr, b = nullptr;
    bool ok = false;
    Events() { ok = hipEventCreate(&a) == hipSuccess && hipEventCreate(&b) == hipSuccess; }
    ~Events() {
        if (a) (void)hipEventDestroy(a);
        if (b) (void)hipEventDestroy(b);
    }
};

// Median over `repeats` timed runs of fn (after one untimed run), in us;
// < 0 when fn or a HIP call failed (the failure's message is kept).
template <class F>
float time_median(F&& fn, int repeats, hipStream_t s) {
    Events ev;
    if (!ev.ok) {
        fail(RUMI_ERR_HIP, "rumi_autotune: hipEventCreate failed");
        return -1.0f;
    }
    if (fn() != RUMI_OK) return -1.0f;
    std::vector<float> t;
    for (int k = 0; k < repeats; ++k) {
        if (hipEventRecord(ev.a, s) != hipSuccess) return -1.0f;
        if (fn() != RUMI_OK) return -1.0f;
        if (hipEventRecord(ev.b, s) != hipSuccess || hipEventSynchronize(ev.b) != hipSuccess) {
            fail(RUMI_ERR_HIP, "rumi_autotune: event synchronisation failed");
            return -1.0f;
        }
        float ms = 0.0f;
        (void)hipEventElapsedTime(&ms, ev.a, ev.b);
        t.push_back(ms * 1e3f);
    }
    std::sort(t.begin(), t.end());
    return t[t.size() / 2];
}

void report_times(rumi_autotune_report* rep, const std::vector<float>& us, float best) {
    if (!rep) return;
    rep->n_candidates = (uint32_t)us.size();
    rep->best_us = best;
    if (rep->host_candidate_us)
        for (size_t k = 0; k < us.size() && k < rep->candidates_capacity; ++k) rep->host_candidate_us[k] = us[k];
}

}  // namespace

extern "C" int rumi_autotune_trace_candidates(rumi_trace_tuning* out, uint32_t capacity, uint32_t* count) {
    clear_error();
    const std::vector<rumi_trace_tuning> c = trace_candidates();
    if (count) *count = (uint32_t)c.size();
    if (out)
        for (size_t k = 0; k < c.size() && k < capacity; ++k) out[k] = c[k];
    return RUMI_OK;
}

extern "C" int rumi_autotune_trace(rumi_ctx* ctx, const rumi_bvh* bvh, const rumi_ray* dev_rays, uint32_t n,
                                   int32_t any, uint32_t flags, void* dev_out, rumi_trace_tuning* best,
                                   rumi_autotune_report* rep, void* stream) {
    clear_error();
    if (!ctx) return fail(RUMI_ERR_NO_DEVICE, "rumi_autotune_trace: null context");
    if (!bvh || !dev_rays || !dev_out || n == 0)
        return fail(RUMI_ERR_INVALID, "rumi_autotune_trace: null BVH, rays or output, or no rays");
    const int repeats = rep && rep->repeats > 0 ? rep->repeats : 3;
    if (repeats > 100) return fail(RUMI_ERR_INVALID, "rumi_autotune_trace: more than 100 repeats");
    const hipStream_t s = (hipStream_t)stream;
    const rumi_trace_tuning keep = ctx->tuning;
    const std::vector<rumi_trace_tuning> cands = trace_candidates();
    std::vector<float> us;
    int bi = -1;
    float bt = 0.0f;
    for (size_t k = 0; k < cands.size(); ++k) {
        if (int rc = rumi_trace_set_tuning(ctx, &cands[k])) {
            ctx->tuning = keep;
            return rc;
        }
        const float t = time_median(
            [&]() {
                return any ? rumi_trace_any(ctx, bvh, dev_rays, n, (int32_t*)dev_out, flags, stream)
                           : rumi_trace_closest(ctx, bvh, dev_rays, n, (rumi_hit*)dev_out, nullptr, flags, stream);
            },
            repeats, s);
        if (t < 0.0f) {
            ctx->tuning = keep;
            return RUMI_ERR_HIP;
        }
        us.push_back(t);
        if (bi < 0 || t < bt) bi = (int)k, bt = t;
    }
    ctx->tuning = cands[(size_t)bi];
    if (best) *best = ctx->tuning;
    report_times(rep, us, bt);
    return RUMI_OK;
}

extern "C" int rumi_autotune_render(rumi_ctx* ctx, const rumi_scene* scene, rumi_render_desc* desc,
                                    float* dev_rgb, rumi_autotune_report* rep, void* stream) {
    clear_error();
    if (!ctx) return fail(RUMI_ERR_NO_DEVICE, "rumi_autotune_render: null context");
    if (!scene || !desc || !dev_rgb) return fail(RUMI_ERR_INVALID, "rumi_autotune_render: null argument");
    const int repeats = rep && rep->repeats > 0 ? rep->repeats : 1;
    if (repeats > 100) return fail(RUMI_ERR_INVALID, "rumi_autotune_render: more than 100 repeats");
    const hipStream_t s = (hipStream_t)stream;
    // path tracer / RayTracer: paths in flight and the refill threshold;
    // bidirectional tracer: samples per batch (2^26 samples: about 130 GB of
    // subpaths, walks and visibility tests at 9 + 9 vertices)
    std::vector<std::pair<uint32_t, uint32_t>> cands;
    if (desc->integrator == RUMI_INTEGRATOR_BDPT) {
        for (uint32_t c : {1u << 20, 1u << 21, 1u << 22, 1u << 23, 1u << 24, 1u << 25, 1u << 26})
            cands.push_back({c, 0u});
    } else {
        for (uint32_t c : {1u << 19, 1u << 20, 1u << 21, 1u << 22, 1u << 23}) {
            cands.push_back({c, c});
            cands.push_back({c, c / 2});
        }
    }
    const uint64_t total = (uint64_t)desc->width * (uint64_t)desc->height * (uint64_t)desc->spp;
    std::vector<float> us;
    int bi = -1;
    float bt = 0.0f;
    for (size_t k = 0; k < cands.size(); ++k) {
        rumi_render_desc d = *desc;
        d.queue_capacity = cands[k].first;
        d.regen_threshold = cands[k].second;
        float t;
        if ((uint64_t)cands[k].first > total && k > 0 && (uint64_t)cands[k - 1].first >= total) {
            t = -2.0f;  // the same queue as a smaller candidate: the frame fits either
        } else {
            int rc = RUMI_OK;
            t = time_median(
                [&]() {
                    rc = rumi_render(ctx, scene, &d, dev_rgb, nullptr, stream);
                    return rc;
                },
                repeats, s);
            if (t < 0.0f) {
                // a candidate whose buffers do not fit in device memory is
                // skipped; any other failure ends the tuning
                if (rc != RUMI_ERR_NO_MEMORY) return rc != RUMI_OK ? rc : RUMI_ERR_HIP;
                (void)hipGetLastError();
                clear_error();
                t = -3.0f;
            }
        }
        us.push_back(t);
        if (t >= 0.0f && (bi < 0 || t < bt)) bi = (int)k, bt = t;
    }
    if (bi < 0) return fail(RUMI_ERR_NO_MEMORY, "rumi_autotune_render: no candidate fits in device memory");
    desc->queue_capacity = cands[(size_t)bi].first;
    desc->regen_threshold = cands[(size_t)bi].second;
    report_times(rep, us, bt);
    return RUMI_OK;
}
