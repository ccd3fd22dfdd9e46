"""Benchmark: incoherent closest-hit ray casting on one MI355X.

Workload (one step = one pass of the hot path over one batch):
  2^23 rays, origins uniform in [-2.1, 2.1]^3 and directions uniform on the
  sphere (PCG32 seed 2, stream = ray index), cast against a noisy geodesic
  icosphere of 1,310,720 triangles (frequency 256, radius 1 +- 0.05, PCG32
  seed 1) in a binned-SAH BVH with leaves of at most 4 triangles.
  Inputs are resident in HBM before the timed region; the BVH is built once
  on the host (setup, like the reference's BVHAccel constructor).

Prints ONE JSON line on stdout.  Progress goes to stderr.
"""
import argparse
import importlib
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "Mrays/s (closest-hit, incoherent) + samples/s at matched rel-MSE vs CPU; grad/s"
HBM_PEAK_GBS = 8000.0  # MI355X HBM3E, published peak
RAY_IN_BYTES = 32      # rumi_ray
HIT_OUT_BYTES = 8      # rumi_hit (closest); any-hit writes 4
NODE_BYTES = 64        # binary interior node: two child boxes + refs
NODE_BYTES4 = 128      # 4-wide interior node: four child boxes + refs
TRI_BYTES = 48         # leaf-ordered triangle: v0, e1, e2, id


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--freq", type=int, default=256)
    ap.add_argument("--rays", type=int, default=1 << 23)
    ap.add_argument("--bins", type=int, default=32, help="binned-SAH bins of the mesh BVH")
    ap.add_argument("--cpu-rays", type=int, default=1 << 20,
                    help="rays in the bounded CPU-baseline sample (0 = skip)")
    ap.add_argument("--no-render", action="store_true", help="skip the path-tracing legs")
    ap.add_argument("--only-leg", default="", metavar="NAME",
                    help="run only this rendering leg (cornell_pt, media_pt, cornell_whitted, model_pt, "
                         "media_gradient, cornell_bdpt, batch) and print its JSON: for profiling")
    ap.add_argument("--no-batch", action="store_true", help="skip the batched 256-scene leg")
    ap.add_argument("--grad-check", type=float, default=0.0, metavar="SECONDS",
                    help="only compare the media adjoint with central differences for about SECONDS")
    ap.add_argument("--grad-steps", default="0.2,0.1",
                    help="relative finite-difference steps of --grad-check (Richardson over the two smallest)")
    ap.add_argument("--batch-scenes", type=int, default=256)
    ap.add_argument("--no-tune", action="store_true",
                    help="keep the default traversal launch shape instead of autotuning it")
    ap.add_argument("--render-runs", type=int, default=3)
    ap.add_argument("--gpu-only", action="store_true",
                    help="render legs: GPU renders only, with the default queue sizes (no autotuner, "
                         "counting render, rel-MSE renders or CPU baseline): for profiling")
    ap.add_argument("--bdpt-batch", type=int, default=0, metavar="SAMPLES",
                    help="with --gpu-only: the BDPT leg's samples per batch (0: the library default)")
    ap.add_argument("--queue", type=int, default=0, metavar="PATHS",
                    help="with --gpu-only: the path-tracing legs' queue capacity (0: the library default)")
    ap.add_argument("--calibrate-copy", type=int, default=0, metavar="MIB",
                    help="only run a device-to-device copy of MIB MiB five times (the "
                         "known-size kernel that scales rocprofv3 FETCH_SIZE/WRITE_SIZE)")
    args = ap.parse_args()

    if args.calibrate_copy:
        import torch
        n = args.calibrate_copy * (1 << 20) // 4
        a = torch.ones(n, dtype=torch.float32, device="cuda")
        b = torch.empty_like(a)
        for _ in range(5):
            b.copy_(a)
        torch.cuda.synchronize()
        log(f"copied {args.calibrate_copy} MiB x 5")
        return

    if args.gpus != 1:
        print(f"bench.py measures one GPU only; --gpus {args.gpus} is not supported "
              "(the work shards by ray and sample ranges; this build measures one GPU)",
              file=sys.stderr)
        sys.exit(2)

    import numpy as np
    import torch

    rumi = importlib.import_module("rumi-renderer_amd")
    assert torch.cuda.is_available(), "bench.py needs a HIP device"

    if args.grad_check > 0:
        grad_check(rumi, rumi.Context(0), args, log)
        return
    if args.only_leg:
        sys.path.insert(0, os.path.join(ROOT, "rumi-renderer_amd"))
        import scenes
        legs = {"cornell_whitted": whitted_leg, "model_pt": model_leg, "media_gradient": gradient_leg,
                "cornell_bdpt": bdpt_leg, "batch": batch_leg}
        ctx = rumi.Context(0)
        if args.only_leg in ("cornell_pt", "media_pt"):
            res = render_legs(rumi, ctx, args, log, only=args.only_leg)
        else:
            res = legs[args.only_leg](rumi, ctx, args, log, scenes)
        print(json.dumps({args.only_leg: res}), flush=True)
        return

    t0 = time.time()
    xyz, idx = rumi.make_icosphere(args.freq, 1, 0.05)
    rays_h = rumi.make_rays_uniform(args.rays, 2, 2.1)
    ctx = rumi.Context(0)
    bvh = rumi.Bvh(ctx, xyz, idx, mode=rumi.RUMI_BVH_SAH_BINNED, max_leaf=4, n_bins=args.bins)
    info = bvh.info()
    n = rays_h.shape[0]
    rays = torch.from_numpy(rays_h).cuda()
    hits = torch.empty((n, 2), dtype=torch.int32, device="cuda")
    anyp = torch.empty(n, dtype=torch.int32, device="cuda")
    torch.cuda.synchronize()
    log(f"setup {time.time() - t0:.1f}s: {idx.shape[0]} tris, {info.n_inner} nodes, "
        f"depth {info.max_depth}, SAH build {info.build_seconds:.2f}s")

    # Per-ray work counted by the instrumented kernel over the whole batch
    # (the same traversal as the timed launches, with the context's current
    # tree width; outside the timed region).
    def work(any_hit):
        nc = torch.empty(n, dtype=torch.int32, device="cuda")
        pc = torch.empty(n, dtype=torch.int32, device="cuda")
        rumi.trace_count(ctx, bvh, rays, any_hit, hits, nc, pc)
        torch.cuda.synchronize()
        node_b = NODE_BYTES4 if ctx.trace_tuning()["width"] == 4 else NODE_BYTES
        return float(nc.double().mean()), float(pc.double().mean()), node_b

    stream = torch.cuda.current_stream()

    def timed(fn, steps, warmup):
        for _ in range(warmup):
            fn()
        torch.cuda.synchronize()
        start = torch.cuda.Event(enable_timing=True)
        end = torch.cuda.Event(enable_timing=True)
        start.record(stream)
        for _ in range(steps):
            fn()
        end.record(stream)
        torch.cuda.synchronize()
        return start.elapsed_time(end) / steps  # ms per launch

    tuning = None
    if not args.no_tune:
        # launch shape chosen by the autotuner on this workload (untimed setup;
        # every candidate returns the same bits)
        best, table = rumi.autotune_trace(ctx, bvh, rays, hits, stream=stream)
        tuning = {"closest": best,
                  "candidates_us": [[c["block"], c["stack"], c["refill"], c["persistent"],
                                     c["width"], round(t * 1e6, 1)] for c, t in table]}
        log(f"autotuned closest-hit: {best}")

    nodes_c, tris_c, node_bc = work(False)
    hit_frac = float((hits[:, 0] >= 0).double().mean())
    ms_c = timed(lambda: rumi.trace_closest(ctx, bvh, rays, hits, None, 0, stream),
                 args.steps, args.warmup)
    log(f"closest-hit: {ms_c:.3f} ms/step, {n / ms_c / 1e3:.1f} Mrays/s")
    if tuning is not None:
        best_a, _ = rumi.autotune_trace(ctx, bvh, rays, anyp, any_hit=True, stream=stream)
        tuning["any"] = best_a
        log(f"autotuned any-hit: {best_a}")
    nodes_a, tris_a, node_ba = work(True)
    ms_a = timed(lambda: rumi.trace_any(ctx, bvh, rays, anyp, 0, stream), args.steps, args.warmup)
    log(f"any-hit: {ms_a:.3f} ms/step, {n / ms_a / 1e3:.1f} Mrays/s")
    b_ray_c = RAY_IN_BYTES + HIT_OUT_BYTES + node_bc * nodes_c + TRI_BYTES * tris_c
    b_ray_a = RAY_IN_BYTES + 4 + node_ba * nodes_a + TRI_BYTES * tris_a

    mrays_c = n / ms_c / 1e3
    mrays_a = n / ms_a / 1e3
    achieved_c = b_ray_c * n / (ms_c * 1e-3) / 1e9
    achieved_a = b_ray_a * n / (ms_a * 1e-3) / 1e9

    pt = None if args.no_render else render_legs(rumi, ctx, args, log)

    cpu = None
    ref_check = None
    if args.cpu_rays > 0:
        import oracle
        k = min(args.cpu_rays, n)
        log(f"CPU baseline: reference median-6 build + Travel on {k} rays, 1 thread ...")
        ref_p, ref_t, secs, build_s = oracle.time_median6_mesh(xyz, idx, rays_h[:k])
        cpu = {"value": k / secs / 1e6, "unit": "Mrays/s", "cores": 1, "kind": "port",
               "sample": f"first {k} rays of the same batch, same mesh; the oracle's C "
                         f"restatement of BVHAccel::BuildBVH (median-6) + Travel; "
                         f"traversal timed only ({secs:.2f}s; build {build_s:.2f}s untimed)",
               "host_cpu": _cpu_model(), "nproc": os.cpu_count()}
        log(f"CPU baseline: {cpu['value']:.3f} Mrays/s")
        # for context: the same sample on the host's share of cores (at most
        # 16, the CPU share of one GPU on the box), one shared tree
        nth = max(1, min(os.cpu_count() or 1, 16))
        _, _, secs_mt, _ = oracle.time_median6_mesh(xyz, idx, rays_h[:k], nthreads=nth)
        cpu["all_cores"] = {"value": k / secs_mt / 1e6, "unit": "Mrays/s", "cores": nth}
        log(f"CPU baseline on {nth} threads: {cpu['all_cores']['value']:.3f} Mrays/s")
        # the benchmarked kernel's hits on those rays against the reference's
        # own tree and order: ids equal except at exact ties (1e-6 relative t)
        rumi.trace_closest(ctx, bvh, rays[:k], hits[:k], None, 0, stream)
        torch.cuda.synchronize()
        gp, gt = rumi.split_hits(hits[:k].cpu().numpy())
        t_same = gt.view(np.int32) == ref_t.view(np.int32)
        diff = gp != ref_p
        tie = np.abs(gt.astype(np.float64) - ref_t) <= 1e-6 * np.abs(ref_t.astype(np.float64))
        ref_check = {"rays": k, "ids_equal": int((~diff).sum()), "ties": int((diff & tie).sum()),
                     "mismatches": int((diff & ~tie).sum()), "t_bit_equal": int(t_same.sum()),
                     "tree": "reference median-6 (BVHAccel::BuildBVH), reference order"}
        log(f"reference-tree check: {ref_check}")

    traffic = measured_traffic("closest-hit", n)
    copy_gbs = hbm_copy_bandwidth(log)

    out = {
        "metric": METRIC,
        "value": round(mrays_c, 2),
        "unit": "Mrays/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_c, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {
            "workload": "incoherent closest-hit ray cast: 2^23 uniform rays vs noisy "
                        "icosphere (20*256^2 = 1,310,720 tris), binned-SAH BVH, <=4 tris/leaf",
            "rays": n, "triangles": int(idx.shape[0]), "bvh_nodes": info.n_inner,
            "bvh_depth": info.max_depth, "hit_fraction": round(hit_frac, 5),
        },
        "roofline": {
            "bound": "hbm", "achieved": round(achieved_c, 1), "peak": HBM_PEAK_GBS,
            "unit": "GB/s", "frac": round(achieved_c / HBM_PEAK_GBS, 4), "traffic": traffic,
            "bytes_per_ray": round(b_ray_c, 1), "nodes_per_ray": round(nodes_c, 3),
            "tris_per_ray": round(tris_c, 3), "node_bytes": node_bc,
            # the I/O-only floor: ray in + hit out, no tree bytes
            "io_floor_bytes_per_ray": RAY_IN_BYTES + HIT_OUT_BYTES,
            "io_floor_mrays": round(HBM_PEAK_GBS * 1e9 / (RAY_IN_BYTES + HIT_OUT_BYTES) / 1e6, 1),
            "achievable_copy_GBs": copy_gbs,
            "frac_of_copy": round(achieved_c / copy_gbs, 4) if copy_gbs else None,
        },
        "tuning": tuning,
        "reference_tree_check": ref_check,
        "cpu_baseline": cpu,
        "any_hit": {
            "value": round(mrays_a, 2), "unit": "Mrays/s", "ms_per_step": round(ms_a, 4),
            "roofline_achieved_GBs": round(achieved_a, 1),
            "roofline_frac": round(achieved_a / HBM_PEAK_GBS, 4),
            "bytes_per_ray": round(b_ray_a, 1), "nodes_per_ray": round(nodes_a, 3),
            "tris_per_ray": round(tris_a, 3), "node_bytes": node_ba,
        },
        "path_tracing": pt,
    }
    print(json.dumps(out), flush=True)


def measured_traffic(label, rays):
    """Bytes per launch that rocprofv3's FETCH_SIZE + WRITE_SIZE counters read
    for this kernel on this workload, each scaled by what the counter reads
    for a device copy of known size (profiles/k_trace_traffic.json, written
    from the two --pmc runs listed in profiles/k_trace_icosphere1p3m_8Mrays.md).
    None when no measurement of this workload is at hand."""
    path = os.path.join(ROOT, "profiles", "k_trace_traffic.json")
    try:
        with open(path) as f:
            k = json.load(f)["kernels"][label]
    except (OSError, KeyError, ValueError):
        return None
    if int(k["rays"]) != rays:
        return None
    return int(round(k["bytes_per_launch"]))


def whitted_leg(rumi, ctx, args, log, scenes):
    """The RayTracer (RUMI_INTEGRATOR_WHITTED, RayTracer(scene, 10)) on the
    512^2 x 16 spp Cornell box: GPU samples/s; the same seed at 64^2 on the
    GPU and in the oracle (relative L1 between them); CPU baseline = that
    oracle render on one thread."""
    import numpy as np
    import torch

    w = h = 512
    spp, depth = 16, 10
    scene = rumi.Scene(ctx, scenes.cornell_box(w, h))
    kw = dict(integrator=rumi.RUMI_INTEGRATOR_WHITTED)
    # queue size from the native autotuner on this frame (untimed)
    if args.gpu_only:
        qtune = {"queue_capacity": args.queue} if args.queue else {}
    else:
        qtune = scene.autotune(w, h, spp, depth, seed=1, **kw)[0]
    kw.update(qtune)
    scene.render(w, h, spp, depth, seed=1, **kw)
    torch.cuda.synchronize()
    times = []
    for k in range(args.render_runs):
        t0 = time.perf_counter()
        img, st = scene.render(w, h, spp, depth, seed=10 + k, **kw)
        torch.cuda.synchronize()
        times.append(time.perf_counter() - t0)
    scene.close()
    kw = dict(integrator=rumi.RUMI_INTEGRATOR_WHITTED)
    sec = min(times)
    sps = st.samples / sec
    import oracle
    small = scenes.cornell_box(64, 64)
    s64 = rumi.Scene(ctx, small)
    gi, _ = s64.render(64, 64, spp, depth, seed=7, **kw)
    torch.cuda.synchronize()
    gi = gi.cpu().numpy()
    s64.close()
    t0 = time.perf_counter()
    ci, _ = oracle.render_whitted(small, 64, 64, spp, depth, seed=7, nthreads=1)
    cpu_sec = time.perf_counter() - t0
    cpu_sps = 64 * 64 * spp / cpu_sec
    l1 = float(np.abs(gi - ci).sum() / max(np.abs(ci).sum(), 1e-30))
    log(f"cornell_whitted: {sps / 1e6:.2f} Msamples/s ({sec * 1e3:.2f} ms), same-seed L1 {l1:.2e}, "
        f"CPU {cpu_sps:.0f} samples/s")
    return {"samples_per_s": round(sps, 1), "unit": "samples/s", "seconds": round(sec, 5),
            "device_ms": round(st.milliseconds, 3), "width": w, "height": h, "spp": spp,
            "max_depth": depth, "queue_tuning": qtune, "same_seed_rel_l1_64px": l1,
            "cpu_baseline": {"value": round(cpu_sps, 1), "unit": "samples/s", "cores": 1, "kind": "port",
                             "sample": f"64x64 at {spp} spp, the oracle's FP64 restatement of "
                                       f"RayTracer::Trace ({cpu_sec:.2f}s)"},
            "speedup_vs_cpu": round(sps / cpu_sps, 1)}


def film_leg(rumi, ctx, img, w, h, log):
    """The film output path on the rendered image: Film::Save's byte
    conversion (rumi_film_encode, 32-bit pixels, gamma 2.2) and
    FirefliesFiltering (rumi_film_fireflies), each timed with HIP events on
    the launch stream over 50 launches; bytes moved per launch (film read +
    pixels written; pixels read + written) over that time; the oracle's C
    restatements of the same two steps timed on one host thread."""
    import numpy as np
    import torch

    stream = torch.cuda.current_stream()
    px = rumi.film_encode(ctx, img, w, h, rumi.RUMI_FILM_RGB_MEAN, 1, 2.2, 4, stream=stream)
    out = rumi.film_fireflies(ctx, px, w, h, 4, stream=stream)

    def timed(fn, reps=50):
        fn()
        e0 = torch.cuda.Event(enable_timing=True)
        e1 = torch.cuda.Event(enable_timing=True)
        e0.record(stream)
        for _ in range(reps):
            fn()
        e1.record(stream)
        torch.cuda.synchronize()
        return e0.elapsed_time(e1) / reps * 1e-3

    t_enc = timed(lambda: rumi.film_encode(ctx, img, w, h, rumi.RUMI_FILM_RGB_MEAN, 1, 2.2, 4, out=px,
                                           stream=stream))
    t_ff = timed(lambda: rumi.film_fireflies(ctx, px, w, h, 4, out=out, stream=stream))
    n = w * h
    import oracle
    host_img = img.cpu().numpy()
    t0 = time.perf_counter()
    host_px = oracle.film_encode(host_img, False, 1, 2.2, 4)
    c_enc = time.perf_counter() - t0
    t0 = time.perf_counter()
    oracle.fireflies(host_px)
    c_ff = time.perf_counter() - t0
    same = bool(np.array_equal(px.cpu().numpy(), host_px))
    res = {"pixels": n, "encode_us": round(t_enc * 1e6, 2), "encode_GBs": round(16 * n / t_enc / 1e9, 1),
           "fireflies_us": round(t_ff * 1e6, 2), "fireflies_GBs": round(8 * n / t_ff / 1e9, 1),
           "bytes_equal_oracle": same,
           "cpu_baseline": {"encode_us": round(c_enc * 1e6, 1), "fireflies_us": round(c_ff * 1e6, 1),
                            "cores": 1, "kind": "port",
                            "sample": f"the same {w}x{h} image through the oracle's C restatements"}}
    log(f"film: encode {t_enc * 1e6:.1f} us, fireflies {t_ff * 1e6:.1f} us "
        f"(CPU {c_enc * 1e6:.0f} / {c_ff * 1e6:.0f} us), bytes equal oracle: {same}")
    return res


def model_leg(rumi, ctx, args, log, scenes):
    """The reference's model scene (src/Scene/Build Model.txt): its
    100,000-triangle dragon.obj loaded by rumi_obj_load with the scene's
    transform, in glass with absorption, 1200 x 1000, depth 20; 4 spp per
    render (the reference's film takes 1).  GPU samples/s; the same seed at
    60 x 50 on the GPU and in the oracle (relative L1 between the two: the
    same samples up to rounding); CPU baseline = that oracle render on one
    thread."""
    import gzip
    import tempfile

    import numpy as np
    import torch

    w, h, spp, depth = 1200, 1000, 4, 20
    with tempfile.TemporaryDirectory() as d:
        path = os.path.join(d, "dragon.obj")
        with open(path, "wb") as f:
            f.write(gzip.open(os.path.join(ROOT, "tests", "golden", "models", "dragon.obj.gz")).read())
        t0 = time.perf_counter()
        model = rumi.load_obj(path, scenes.model_transform(0.3))
        load_s = time.perf_counter() - t0
    sc = scenes.model_scene(model, w, h)
    scene = rumi.Scene(ctx, sc)
    # queue size from the native autotuner on this frame (untimed)
    if args.gpu_only:
        qtune = {"queue_capacity": args.queue} if args.queue else {}
    else:
        qtune = scene.autotune(w, h, spp, depth, seed=1)[0]
    img, st = scene.render(w, h, spp, depth, seed=1, **qtune)
    torch.cuda.synchronize()
    times = []
    for k in range(args.render_runs):
        t0 = time.perf_counter()
        img, st = scene.render(w, h, spp, depth, seed=10 + k, **qtune)
        torch.cuda.synchronize()
        times.append(time.perf_counter() - t0)
    scene.close()
    sec = min(times)
    sps = st.samples / sec
    film = film_leg(rumi, ctx, img.contiguous(), w, h, log)
    import oracle
    cw, ch = 60, 50
    small = scenes.model_scene(model, cw, ch)
    s_small = rumi.Scene(ctx, small)
    gi, _ = s_small.render(cw, ch, spp, depth, seed=7)
    torch.cuda.synchronize()
    gi = gi.cpu().numpy()
    s_small.close()
    t0 = time.perf_counter()
    ci, _ = oracle.render_pt(small, cw, ch, spp, depth, seed=7, nthreads=1)
    cpu_sec = time.perf_counter() - t0
    cpu_sps = cw * ch * spp / cpu_sec
    l1 = float(np.abs(gi - ci).sum() / max(np.abs(ci).sum(), 1e-30))
    out = {
        "samples_per_s": round(sps, 1), "unit": "samples/s", "seconds": round(sec, 4),
        "device_ms": round(st.milliseconds, 3), "width": w, "height": h, "spp": spp,
        "max_depth": depth, "triangles": int(model["idx"].shape[0]),
        "obj_load_s": round(load_s, 3), "queue_tuning": qtune,
        "extension_rays": int(st.extension_rays), "shadow_rays": int(st.shadow_rays),
        "rays_per_s": round((st.extension_rays + st.shadow_rays) / sec, 1),
        "same_seed_rel_l1_60x50": l1,
        "cpu_baseline": {"value": round(cpu_sps, 1), "unit": "samples/s", "cores": 1, "kind": "port",
                         "sample": f"60x50 at {spp} spp, depth {depth}, the oracle's FP64 "
                                   f"restatement of PathTracer::Trace ({cpu_sec:.2f}s)"},
        "speedup_vs_cpu": round(sps / cpu_sps, 1),
        "film_output": film,
    }
    log(f"model_pt (dragon): {sps / 1e6:.2f} Msamples/s ({sec * 1e3:.1f} ms), same-seed L1 {l1:.2e}, "
        f"CPU {cpu_sps:.0f} samples/s")
    return out


def render_legs(rumi, ctx, args, log, only=""):
    """Path tracing: the 512^2 x 16 spp Cornell box (depth 5) and the 512^2 x
    32 spp media scene (depth 16), samples/s on the GPU; rel-MSE against the
    committed 4096-spp oracle reference at 64^2, GPU vs the oracle at equal
    spp; CPU baseline = the oracle on one thread over a bounded sample."""
    import numpy as np
    import torch

    sys.path.insert(0, os.path.join(ROOT, "rumi-renderer_amd"))
    import scenes

    out = {}
    legs = [("cornell_pt", scenes.cornell_box, 512, 512, 16, 5, "cornell64_pt_d5"),
            ("media_pt", scenes.cornell_media, 512, 512, 32, 16, "media64_pt_d16")]
    for name, factory, w, h, spp, depth, golden in legs:
        if only and name != only:
            continue
        sc = factory(w, h)
        scene = rumi.Scene(ctx, sc)
        # queue sizes chosen by the native autotuner on this frame (untimed)
        if args.gpu_only:
            qtune, qtable = ({"queue_capacity": args.queue} if args.queue else {}), []
        else:
            qtune, qtable = scene.autotune(w, h, spp, depth, seed=1)
        img, st = scene.render(w, h, spp, depth, seed=1, **qtune)  # warm-up
        torch.cuda.synchronize()
        times = []
        for k in range(args.render_runs):
            t0 = time.perf_counter()
            img, st = scene.render(w, h, spp, depth, seed=10 + k, **qtune)
            torch.cuda.synchronize()
            times.append(time.perf_counter() - t0)
        sec = min(times)
        sps = st.samples / sec
        if args.gpu_only:
            out[name] = {"samples_per_s": round(sps, 1), "unit": "samples/s", "seconds": round(sec, 4)}
            log(f"{name}: {sps / 1e6:.2f} Msamples/s ({sec * 1e3:.1f} ms)")
            scene.close()
            continue
        # one counting render (untimed): the per-ray work behind the roofline
        _, cst = scene.render(w, h, spp, depth, seed=10, count_traversal=True, **qtune)
        torch.cuda.synchronize()
        roof = render_roofline(cst, sps)
        # rel-MSE at equal spp against the committed 64^2 4096-spp reference,
        # GPU and CPU on disjoint seeds (independent renders): n_ab GPU renders,
        # the 4 timed one-thread CPU renders below and n_ab - 4 more on all
        # threads (the squared error is heavy-tailed: caustic paths)
        n_ab = 64 if name == "cornell_pt" else 40
        ref = np.load(os.path.join(ROOT, "tests", "golden", golden + ".npz"))
        sc64 = factory(64, 64)
        s64 = rumi.Scene(ctx, sc64)
        g_vals = []
        for k in range(n_ab):
            gi, _ = s64.render(64, 64, spp, depth, seed=200 + k)
            torch.cuda.synchronize()
            g_vals.append(rel_mse(gi.cpu().numpy(), ref["mean"]))
        s64.close()
        # CPU baseline leg: the oracle (FP64 restatement of PathTracer::Trace)
        # on one thread, 4 renders of 64^2 at the same spp
        import oracle
        c_vals = []
        t0 = time.perf_counter()
        for k in range(4):
            ci, _ = oracle.render_pt(sc64, 64, 64, spp, depth, seed=600 + k, nthreads=1)
            c_vals.append(rel_mse(ci, ref["mean"]))
        cpu_sec = time.perf_counter() - t0
        cpu_sps = 4 * 64 * 64 * spp / cpu_sec
        threads = min(16, os.cpu_count() or 1)
        for k in range(4, n_ab):
            ci, _ = oracle.render_pt(sc64, 64, 64, spp, depth, seed=600 + k, nthreads=threads)
            c_vals.append(rel_mse(ci, ref["mean"]))
        ab = rel_mse_ab(g_vals, c_vals)
        if name == "cornell_pt":
            # the reference's own estimator (fixed pixel and light point sets,
            # Q9), restated in the oracle
            r_vals = [rel_mse(oracle.render_pt_refsampling(sc64, 64, 64, spp, depth, seed=900 + k,
                                                           nthreads=threads)[0], ref["mean"])
                      for k in range(n_ab)]
            ab["vs_reference_sampling"] = rel_mse_ab(g_vals, r_vals)
        out[name] = {
            "samples_per_s": round(sps, 1), "unit": "samples/s", "seconds": round(sec, 4),
            "device_ms": round(st.milliseconds, 3), "width": w, "height": h, "spp": spp,
            "max_depth": depth, "extension_rays": int(st.extension_rays),
            "shadow_rays": int(st.shadow_rays), "iterations": int(st.iterations),
            "rays_per_s": round((st.extension_rays + st.shadow_rays) / sec, 1),
            "rel_mse_64px": dict(ab, spp=spp),
            "roofline": roof,
            "queue_tuning": dict(qtune, candidates_ms=[None if t is None else round(t * 1e3, 2) for t in qtable]),
            "cpu_baseline": {"value": round(cpu_sps, 1), "unit": "samples/s", "cores": 1,
                             "kind": "port",
                             "sample": f"4 renders of 64x64 at {spp} spp, the oracle's FP64 "
                                       f"restatement of PathTracer::Trace ({cpu_sec:.2f}s)"},
            "speedup_vs_cpu": round(sps / cpu_sps, 1),
        }
        log(f"{name}: {sps / 1e6:.2f} Msamples/s ({sec * 1e3:.1f} ms), "
            f"rel-MSE ratio {ab['ratio']:.3f} +- {ab['ratio_se']:.3f}, CPU {cpu_sps:.0f} samples/s")
        scene.close()
    if only:
        return out[only]
    out["cornell_whitted"] = whitted_leg(rumi, ctx, args, log, scenes)
    out["model_pt"] = model_leg(rumi, ctx, args, log, scenes)
    out["media_gradient"] = gradient_leg(rumi, ctx, args, log, scenes)
    out["cornell_bdpt"] = bdpt_leg(rumi, ctx, args, log, scenes)
    if not args.no_batch:
        out["batch"] = batch_leg(rumi, ctx, args, log, scenes)
    return out


def batch_leg(rumi, ctx, args, log, scenes):
    """256 scenes (noisy icospheres of 1k-200k triangles, log-uniform, each
    with a 128^2 albedo texture, a floor and an area light; image sides
    log-uniform in 128..512) at 8 spp, depth 6, as one wavefront: the
    forward render, dLoss/dImage of the L2 loss against targets (the same
    scenes with textures x 0.8) and the adjoint with respect to every texel,
    captured as one hipGraph.  The generation size is autotuned.  Reports
    samples/s of the forward pass alone, grad/s of the whole graph and the
    per-scene completion spread of the forward pass inside it."""
    import numpy as np
    import torch
    from concurrent.futures import ThreadPoolExecutor
    sys.path.insert(0, os.path.join(ROOT, "rumi-renderer_amd"))
    import autotune
    spp, depth = 8, 6
    params = scenes.batch_params(args.batch_scenes)
    t0 = time.perf_counter()

    def make(i, scale):
        _, freq, side = params[i]
        return rumi.Scene(ctx, scenes.batch_scene(i, freq, side, tex_scale=scale))

    with ThreadPoolExecutor(16) as ex:
        objs = list(ex.map(lambda i: make(i, 1.0), range(len(params))))
        tobjs = list(ex.map(lambda i: make(i, 0.8), range(len(params))))
    items = [(o, side, side) for o, (_, _, side) in zip(objs, params)]
    tris = int(sum(20 * f * f + 2 for _, f, _ in params))
    log(f"batch: {len(items)} scenes, {tris} triangles, setup {time.perf_counter() - t0:.1f}s")
    tb = rumi.Batch(ctx, [(o, s_, s_) for o, (_, _, s_) in zip(tobjs, params)], spp, depth, seed=77)
    targets = tb.run(use_graph=False, item_ms=False)[0].clone()
    torch.cuda.synchronize()
    tb.close()
    for o in tobjs:
        o.close()

    def timed_graph(b, **kw):
        b.run(use_graph=True, item_ms=False, **kw)  # capture + warm-up
        torch.cuda.synchronize()
        best = None
        for _ in range(max(1, args.render_runs - 1)):
            r = b.run(use_graph=True, **kw)
            torch.cuda.synchronize()
            best = r if best is None or r[3].milliseconds < best[3].milliseconds else best
        return best

    # generation size: autotuned on the full graph (one replay after capture)
    results = {}

    def build(cap):
        b = rumi.Batch(ctx, items, spp, depth, seed=5, queue_capacity=cap, adjoint=True)
        b.capacity = cap
        return b

    def measure(b):
        results[b.capacity] = timed_graph(b, targets=targets)
        return results[b.capacity][3].milliseconds * 1e-3

    best_cap, table = autotune.tune_batch(build, measure, release=lambda b: b.close())
    best_run = results[best_cap]
    best_ms = best_run[3].milliseconds
    table = [[c, round(t * 1e3, 2)] for c, t in table]
    log(f"batch: generation sizes {table} -> {best_cap}")
    _, tg, loss, st, item_ms = best_run
    grad_s = 1000.0 / best_ms
    fb = rumi.Batch(ctx, items, spp, depth, seed=5, queue_capacity=best_cap)
    fr = timed_graph(fb)
    fwd_ms = fr[3].milliseconds
    fb.close()
    lat = np.array(item_ms)
    res = {
        "scenes": len(items), "triangles": tris, "pixels": int(st.total_pixels),
        "texels": int(st.total_texels), "spp": spp, "max_depth": depth,
        "samples": int(st.total_samples), "generation_size": best_cap,
        "generations": int(st.generations), "generation_sizes_ms": table,
        "forward_samples_per_s": round(st.total_samples / (fwd_ms * 1e-3), 1),
        "forward_ms": round(fwd_ms, 2),
        "graph_ms": round(best_ms, 2), "grads_per_s": round(grad_s, 4),
        "unit": "grad/s (forward + dLoss/dImage + adjoint for all texels of all scenes)",
        "graph_samples_per_s": round(3 * st.total_samples / (best_ms * 1e-3), 1),
        "scene_latency_ms": {"p50": round(float(np.percentile(lat, 50)), 2),
                             "p95": round(float(np.percentile(lat, 95)), 2),
                             "max": round(float(lat.max()), 2),
                             "min": round(float(lat.min()), 2)},
        "loss_total": float(loss.double().sum()),
        "texel_grad_l1": float(tg.double().abs().sum()),
    }
    # CPU baseline: the oracle's FP64 render + adjoint of scene 0 at 64^2, one thread
    import oracle
    _, f0, _ = params[0]
    sc0 = scenes.batch_scene(0, f0, 64)
    t64 = oracle.render_pt(scenes.batch_scene(0, f0, 64, tex_scale=0.8), 64, 64, spp, depth,
                           seed=3, nthreads=8)[0]
    t0 = time.perf_counter()
    i64 = oracle.render_pt(sc0, 64, 64, spp, depth, seed=4, nthreads=1)[0]
    oracle.adjoint_pt(sc0, 2.0 * (i64 - t64), spp, depth, seed=5, nthreads=1, textures=True)
    cpu_sec = time.perf_counter() - t0
    cpu_sps = 3 * 64 * 64 * spp / cpu_sec
    res["cpu_baseline"] = {"value": round(cpu_sps, 1), "unit": "samples/s (render + adjoint)",
                           "cores": 1, "kind": "port",
                           "sample": f"scene 0 at 64x64, {spp} spp: the oracle's FP64 render and "
                                     f"forward-mode texel adjoint ({cpu_sec:.2f}s)"}
    res["speedup_vs_cpu"] = round(res["graph_samples_per_s"] / cpu_sps, 1)
    log(f"batch: forward {res['forward_samples_per_s'] / 1e6:.1f} Msamples/s, graph {best_ms:.0f} ms "
        f"= {grad_s:.3f} grad/s, latency p50/p95/max {res['scene_latency_ms']}")
    for o in objs:
        o.close()
    return res


def grad_check(rumi, ctx, args, log):
    """Adjoint vs central finite differences on the differentiable-media
    workload (512^2, 32 spp, depth 16), common random numbers (one seed per
    replicate, shared by every step and sign).  The loss is linearised at a
    fixed weight image w = 2 (I0 - T) (I0 at theta, T at theta x 1.10, both
    256 spp), so the adjoint and the differences estimate the same quantity,
    sum w dE[I]/dtheta.

    A difference with common random numbers is dominated by sampling
    decisions that flip between theta - h and theta + h (a distance, lobe or
    roulette choice): its variance grows as 1/h, so a step of 1e-3 cannot
    resolve 1% in minutes.  The steps here are relative (h_k = step theta_k),
    and the estimate is the Richardson combination (4 D(h/2) - D(h)) / 3 of
    the two smallest, whose remaining bias is O(h^4); the O(h^2) bias of D(h)
    itself, (D(h) - D(h/2)) 4/3, is reported beside it.  Runs for about the
    given seconds; prints one JSON line."""
    import numpy as np
    import torch
    sys.path.insert(0, os.path.join(ROOT, "rumi-renderer_amd"))
    import scenes
    w = h = 512
    spp, depth = 32, 16
    steps = sorted({float(x) for x in args.grad_steps.split(",")}, reverse=True)
    theta = np.array([0.05, 0.2, 0.4, 1.0, 1.0, 1.0, 0.7])
    names = ["sigma_a.r", "sigma_a.g", "sigma_a.b", "sigma_s.r", "sigma_s.g", "sigma_s.b", "g"]

    def scene(p):
        return rumi.Scene(ctx, scenes.cornell_media(w, h, sigma_a=tuple(p[:3]), sigma_s=tuple(p[3:6]),
                                                    g=float(p[6])))

    s0, st = scene(theta), scene(theta * 1.10)
    i0 = s0.render(w, h, 256, depth, seed=11)[0].clone()
    tg = st.render(w, h, 256, depth, seed=12)[0].clone()
    st.close()
    wgt = (2.0 * (i0 - tg)).contiguous()
    wgt64 = wgt.double()
    torch.cuda.synchronize()
    adj = []
    t0 = time.perf_counter()
    while time.perf_counter() - t0 < args.grad_check * 0.15 or len(adj) < 8:
        adj.append(s0.adjoint(wgt, spp, depth, seed=1000 + len(adj))[1].cpu().double().numpy())
    log(f"grad check: {len(adj)} adjoint runs")
    adj = np.array(adj)
    # scenes at every perturbed point, built once
    pts = {}
    for k in range(7):
        for hs in steps:
            for sg in (1, -1):
                p = theta.copy()
                p[k] += sg * hs * theta[k]
                pts[(k, hs, sg)] = scene(p)
    img = torch.empty((h, w, 3), dtype=torch.float32, device="cuda")
    fd = []   # replicate x step x param
    t0 = time.perf_counter()
    budget = args.grad_check * 0.85
    while time.perf_counter() - t0 < budget or len(fd) < 4:
        r = len(fd)
        rep = np.zeros((len(steps), 7))
        for k in range(7):
            for j, hs in enumerate(steps):
                lp = float((pts[(k, hs, 1)].render(w, h, spp, depth, seed=5000 + r, out=img)[0]
                            .double() * wgt64).sum())
                lm = float((pts[(k, hs, -1)].render(w, h, spp, depth, seed=5000 + r, out=img)[0]
                            .double() * wgt64).sum())
                rep[j, k] = (lp - lm) / (2 * hs * theta[k])
        fd.append(rep)
        if r % 10 == 9:
            log(f"grad check: {r + 1} difference rounds, {time.perf_counter() - t0:.0f}s")
    for o in pts.values():
        o.close()
    s0.close()
    fd = np.array(fd)
    n = fd.shape[0]
    a_m, a_se = adj.mean(0), adj.std(0, ddof=1) / np.sqrt(len(adj))
    per_step = {str(hs): {"fd": fd[:, j].mean(0).round(6).tolist(),
                          "fd_se": (fd[:, j].std(0, ddof=1) / np.sqrt(n)).round(6).tolist()}
                for j, hs in enumerate(steps)}
    rich = (4.0 * fd[:, -1] - fd[:, -2]) / 3.0 if len(steps) > 1 else fd[:, -1]
    r_m, r_se = rich.mean(0), rich.std(0, ddof=1) / np.sqrt(n)
    bias_h = ((fd[:, -2] - fd[:, -1]).mean(0) * 4.0 / 3.0) if len(steps) > 1 else np.zeros(7)
    rel = np.abs(a_m - r_m) / np.abs(r_m)
    rel_se = np.sqrt(a_se ** 2 + r_se ** 2) / np.abs(r_m)
    # the whole gradient: relative L2 distance between the adjoint's and the
    # differences' 7-vectors, with a bootstrap over adjoint runs and
    # difference replicates for its spread
    vec = float(np.linalg.norm(a_m - r_m) / np.linalg.norm(r_m))
    rng = np.random.default_rng(7)
    boot = []
    for _ in range(2000):
        am = adj[rng.integers(0, len(adj), len(adj))].mean(0)
        rm = rich[rng.integers(0, n, n)].mean(0)
        boot.append(np.linalg.norm(am - rm) / np.linalg.norm(rm))
    boot = np.array(boot)
    # what pure noise gives: the same distance with the adjoint replaced by
    # the differences' own mean (a perfect adjoint), bootstrapped
    null = []
    for _ in range(2000):
        rm = rich[rng.integers(0, n, n)].mean(0)
        null.append(np.linalg.norm(r_m - rm) / np.linalg.norm(rm))
    null = np.array(null)
    out = {"metric": "adjoint vs central FD (Richardson over relative steps), media 512^2 32 spp depth 16",
           "params": names, "theta": theta.tolist(), "steps": steps,
           "adjoint": a_m.round(6).tolist(), "adjoint_se": a_se.round(6).tolist(),
           "fd_richardson": r_m.round(6).tolist(), "fd_richardson_se": r_se.round(6).tolist(),
           "fd_rel_se": (r_se / np.abs(r_m)).round(5).tolist(),
           "bias_of_largest_step_estimate": bias_h.round(6).tolist(),
           "per_step": per_step,
           "rel_diff": rel.round(5).tolist(), "rel_diff_se": rel_se.round(5).tolist(),
           "vector_rel_l2": round(vec, 5),
           "vector_rel_l2_bootstrap_p50_p95": [round(float(np.percentile(boot, 50)), 5),
                                               round(float(np.percentile(boot, 95)), 5)],
           "vector_rel_l2_noise_only_p50_p95": [round(float(np.percentile(null, 50)), 5),
                                                round(float(np.percentile(null, 95)), 5)],
           "adjoint_runs": len(adj), "fd_replicates": n, "spp": spp}
    print(json.dumps(out), flush=True)


def gradient_leg(rumi, ctx, args, log, scenes):
    """Differentiable media: 512^2 x 32 spp, depth 16, L2 loss against the
    image rendered at every medium parameter x 1.10.  One gradient = the
    forward render, dLoss/dImage = 2 (I - T) on the device, and rumi_adjoint
    (primal pass + path replay) for all 7 parameters (sigma_a rgb, sigma_s
    rgb, g).  CPU baseline: the oracle's FP64 render + adjoint on one thread
    over a 64^2 x 32 spp sample, scaled by the sample ratio."""
    import numpy as np
    import torch
    w = h = 512
    spp, depth = 32, 16
    theta = dict(sigma_a=(0.05, 0.2, 0.4), sigma_s=(1.0, 1.0, 1.0), g=0.7)
    target = {k: tuple(1.1 * x for x in v) if isinstance(v, tuple) else 1.1 * v
              for k, v in theta.items()}
    st_t = rumi.Scene(ctx, scenes.cornell_media(w, h, **target))
    tgt, _ = st_t.render(w, h, spp, depth, seed=7)
    torch.cuda.synchronize()
    st_t.close()
    scene = rumi.Scene(ctx, scenes.cornell_media(w, h, **theta))
    dimg = torch.empty_like(tgt)
    # the frame's queue size from the native autotuner (untimed); the
    # adjoint's primal and replay passes run on the same queue
    qtune, _ = scene.autotune(w, h, spp, depth, seed=1)

    def one_grad(seed):
        img, _ = scene.render(w, h, spp, depth, seed=seed, **qtune)
        torch.sub(img, tgt, out=dimg)
        dimg.mul_(2.0)
        return scene.adjoint(dimg, spp, depth, seed=seed + 1000, **qtune)

    one_grad(1)
    torch.cuda.synchronize()
    times, grads = [], []
    for k in range(max(1, args.render_runs)):
        t0 = time.perf_counter()
        _, dmed, _, st = one_grad(20 + k)
        torch.cuda.synchronize()
        times.append(time.perf_counter() - t0)
        grads.append(dmed.cpu().tolist())
    sec = min(times)
    loss = float(((scene.render(w, h, spp, depth, seed=99)[0] - tgt) ** 2).sum())
    scene.close()
    import oracle
    sc64 = scenes.cornell_media(64, 64, **theta)
    t64 = oracle.render_pt(scenes.cornell_media(64, 64, **target), 64, 64, spp, depth, seed=5,
                           nthreads=8)[0]
    t0 = time.perf_counter()
    i64 = oracle.render_pt(sc64, 64, 64, spp, depth, seed=6, nthreads=1)[0]
    oracle.adjoint_pt(sc64, 2.0 * (i64 - t64), spp, depth, seed=7, nthreads=1)
    cpu_sec = time.perf_counter() - t0
    scale = (w * h) / (64 * 64)
    res = {
        "grads_per_s": round(1.0 / sec, 3), "unit": "grad/s", "seconds": round(sec, 4),
        "width": w, "height": h, "spp": spp, "max_depth": depth,
        "params": ["sigma_a.r", "sigma_a.g", "sigma_a.b", "sigma_s.r", "sigma_s.g", "sigma_s.b", "g"],
        "gradient": [round(float(x), 6) for x in np.mean(np.array(grads), 0)],
        "loss": loss, "samples_per_grad": 3 * w * h * spp, "queue_tuning": qtune,
        "cpu_baseline": {"value": round(1.0 / (cpu_sec * scale), 6), "unit": "grad/s", "cores": 1,
                         "kind": "port",
                         "sample": f"64x64 at {spp} spp (1/{scale:.0f} of the image): the oracle's "
                                   f"FP64 render + forward-mode adjoint, {cpu_sec:.2f}s, scaled"},
    }
    res["speedup_vs_cpu"] = round(res["grads_per_s"] / res["cpu_baseline"]["value"], 1)
    log(f"media_gradient: {res['grads_per_s']:.2f} grad/s ({sec * 1e3:.1f} ms), "
        f"CPU {res['cpu_baseline']['value']:.5f} grad/s")
    return res


def bdpt_leg(rumi, ctx, args, log, scenes):
    """Bidirectional path tracing on the 1024^2 Cornell box with the
    200k-triangle GGX icosphere and the glass sphere, path length <= 8,
    power heuristic, 64 spp; samples/s and a CPU baseline of the oracle's
    FP64 BDPT on one thread over a bounded sample."""
    import torch
    w = h = 1024
    spp, max_len = 64, 8
    sc = scenes.cornell_bdpt(w, h)
    scene = rumi.Scene(ctx, sc)
    kw = dict(integrator=rumi.RUMI_INTEGRATOR_BDPT, mis_beta=2.0, n_eye=9, n_light=9)
    # samples per batch chosen by the native autotuner on this frame (untimed)
    if args.gpu_only:
        qtune, qtable = {"queue_capacity": args.bdpt_batch}, []
    else:
        qtune, qtable = scene.autotune(w, h, spp, max_len, seed=1, **kw)
    log(f"cornell_bdpt: batch sizes {qtable} -> {qtune}")
    times = []
    for k in range(max(1, args.render_runs - 1)):
        t0 = time.perf_counter()
        img, st = scene.render(w, h, spp, max_len, seed=10 + k, queue_capacity=qtune["queue_capacity"], **kw)
        torch.cuda.synchronize()
        times.append(time.perf_counter() - t0)
    sec = min(times)
    sps = st.samples / sec
    if args.gpu_only:
        scene.close()
        log(f"cornell_bdpt: {sps / 1e6:.3f} Msamples/s ({sec * 1e3:.1f} ms)")
        return {"samples_per_s": round(sps, 1), "unit": "samples/s", "seconds": round(sec, 4)}
    _, cst = scene.render(w, h, spp, max_len, seed=10, queue_capacity=qtune["queue_capacity"],
                          count_traversal=True, **kw)
    torch.cuda.synchronize()
    roof = render_roofline(cst, sps)
    scene.close()
    # rel-MSE at equal spp against the committed 64^2 4096-spp oracle
    # reference, GPU and CPU on disjoint seeds (independent renders)
    import numpy as np
    ref = np.load(os.path.join(ROOT, "tests", "golden", "bdpt64_l8.npz"))
    sc64 = scenes.cornell_bdpt(64, 64)
    s64 = rumi.Scene(ctx, sc64)
    g_vals = []
    for k in range(24):
        gi, _ = s64.render(64, 64, spp, max_len, seed=300 + k, **kw)
        torch.cuda.synchronize()
        g_vals.append(rel_mse(gi.cpu().numpy(), ref["mean"]))
    s64.close()
    # CPU baseline leg: the oracle's FP64 bidirectional tracer timed on one
    # thread over a 32^2 render at the same spp; its rel-MSE at equal spp from
    # the six 64^2 renders on all host threads
    import oracle
    t0 = time.perf_counter()
    oracle.render_bdpt(scenes.cornell_bdpt(32, 32), 32, 32, spp, max_len, n_eye=9, n_light=9,
                       mis_beta=2.0, seed=5, nthreads=1)
    cpu_sec = time.perf_counter() - t0
    cpu_sps = 32 * 32 * spp / cpu_sec
    c_vals = [rel_mse(oracle.render_bdpt(sc64, 64, 64, spp, max_len, n_eye=9, n_light=9,
                                         mis_beta=2.0, seed=700 + k, nthreads=min(16, os.cpu_count() or 1)),
                      ref["mean"]) for k in range(8)]
    ab = rel_mse_ab(g_vals, c_vals)
    log(f"cornell_bdpt: {sps / 1e6:.3f} Msamples/s ({sec * 1e3:.1f} ms), CPU {cpu_sps:.0f} samples/s, "
        f"rel-MSE ratio {ab['ratio']:.3f} +- {ab['ratio_se']:.3f}")
    return {"samples_per_s": round(sps, 1), "unit": "samples/s", "seconds": round(sec, 4),
            "device_ms": round(st.milliseconds, 3), "width": w, "height": h, "spp": spp,
            "max_path_length": max_len, "mis_beta": 2.0, "triangles": 200000 + 10,
            "rel_mse_64px": dict(ab, spp=spp),
            "roofline": roof,
            "walk_segments_per_sample": round(cst.extension_rays / cst.samples, 3),
            "connection_rays_per_sample": round(cst.shadow_rays / cst.samples, 3),
            "batch_tuning": dict(samples_per_batch=qtune["queue_capacity"],
                                 candidates_ms=[None if t is None else round(t * 1e3, 1) for t in qtable]),
            "cpu_baseline": {"value": round(cpu_sps, 1), "unit": "samples/s", "cores": 1,
                             "kind": "port",
                             "sample": f"32x32 at {spp} spp, the oracle's FP64 bidirectional "
                                       f"tracer ({cpu_sec:.2f}s)"},
            "speedup_vs_cpu": round(sps / cpu_sps, 1)}


def render_roofline(st, samples_per_s, node_bytes=NODE_BYTES4, prim_bytes=TRI_BYTES):
    """Roofline of a render leg from one counting render's stats (st:
    RenderStats with the traversal totals): algorithmic bytes per sample =
    sum over the sample's rays of B_ray (ray 32 B in, hit 8 B out -- 4 for a
    shadow ray -- 128 B per 4-wide node entered, 48 B per primitive tested)
    + 128 B of path state read and written per extension ray (per walk
    segment for BDPT) + 12 B of film; achieved = that x samples/s against
    the 8 TB/s peak.  These scenes' trees are cache-resident, so the bytes
    are touched, not DRAM traffic."""
    R = float(st.samples)
    e, sh = float(st.extension_rays), float(st.shadow_rays)
    ne = st.ext_nodes / e if e else 0.0
    pe = st.ext_prims / e if e else 0.0
    ns = st.shadow_nodes / sh if sh else 0.0
    ps = st.shadow_prims / sh if sh else 0.0
    b_ext = RAY_IN_BYTES + HIT_OUT_BYTES + node_bytes * ne + prim_bytes * pe
    b_sh = RAY_IN_BYTES + 4 + node_bytes * ns + prim_bytes * ps
    per_sample = (e * b_ext + sh * b_sh) / R + 128.0 * (e / R) + 12.0
    achieved = per_sample * samples_per_s / 1e9
    return {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
            "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": None,
            "bytes_per_sample": round(per_sample, 1),
            "extension_rays_per_sample": round(e / R, 3), "shadow_rays_per_sample": round(sh / R, 3),
            "nodes_per_extension_ray": round(ne, 3), "prims_per_extension_ray": round(pe, 3),
            "nodes_per_shadow_ray": round(ns, 3), "prims_per_shadow_ray": round(ps, 3)}


def hbm_copy_bandwidth(log, mib=4096, runs=5):
    """Achievable HBM bandwidth: a device-to-device copy of `mib` MiB (16x
    the 256 MB Infinity Cache, so it streams from HBM), read + write bytes
    over the best of `runs` timed copies (HIP events)."""
    import torch
    n = mib * (1 << 20) // 4
    a = torch.ones(n, dtype=torch.float32, device="cuda")
    b = torch.empty_like(a)
    b.copy_(a)
    torch.cuda.synchronize()
    best = float("inf")
    for _ in range(runs):
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record()
        b.copy_(a)
        e1.record()
        e1.synchronize()
        best = min(best, e0.elapsed_time(e1) * 1e-3)
    gbs = 2.0 * n * 4 / best / 1e9
    del a, b
    torch.cuda.empty_cache()
    log(f"HBM copy: {mib} MiB in {best * 1e3:.3f} ms = {gbs:.0f} GB/s (read + write)")
    return round(gbs, 1)


def rel_mse_ab(g_vals, c_vals):
    """Mean rel-MSE of independent GPU and CPU renders and their ratio with
    its standard error (delta method)."""
    import numpy as np
    g, c = np.asarray(g_vals, float), np.asarray(c_vals, float)
    gm, cm = g.mean(), c.mean()
    gse, cse = g.std(ddof=1) / np.sqrt(len(g)), c.std(ddof=1) / np.sqrt(len(c))
    r = gm / cm
    return {"gpu": float(gm), "gpu_se": float(gse), "cpu_oracle": float(cm), "cpu_se": float(cse),
            "ratio": float(r), "ratio_se": float(r * np.sqrt((gse / gm) ** 2 + (cse / cm) ** 2)),
            "ratio_of_medians": float(np.median(g) / np.median(c)),
            "renders": [len(g), len(c)], "independent_seeds": True}


def rel_mse(img, ref):
    """mean over pixels and channels of (I - R)^2 / (R^2 + 0.01)."""
    import numpy as np
    img = np.asarray(img, np.float64)
    ref = np.asarray(ref, np.float64)
    return float(np.mean((img - ref) ** 2 / (ref ** 2 + 0.01)))


def _cpu_model():
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


if __name__ == "__main__":
    main()
