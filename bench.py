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
NODE_BYTES = 64        # interior node: two child boxes + refs
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
    ap.add_argument("--cpu-rays", type=int, default=1 << 20,
                    help="rays in the bounded CPU-baseline sample (0 = skip)")
    args = ap.parse_args()

    if args.gpus != 1:
        print(f"bench.py measures one GPU only; --gpus {args.gpus} is not supported "
              "(the path shards by rays; see DESIGN.md)", file=sys.stderr)
        sys.exit(2)

    import numpy as np
    import torch

    rumi = importlib.import_module("rumi-renderer_amd")
    assert torch.cuda.is_available(), "bench.py needs a HIP device"

    t0 = time.time()
    xyz, idx = rumi.make_icosphere(args.freq, 1, 0.05)
    rays_h = rumi.make_rays_uniform(args.rays, 2, 2.1)
    ctx = rumi.Context(0)
    bvh = rumi.Bvh(ctx, xyz, idx, mode=rumi.RUMI_BVH_SAH_BINNED, max_leaf=4, n_bins=32)
    info = bvh.info()
    n = rays_h.shape[0]
    rays = torch.from_numpy(rays_h).cuda()
    hits = torch.empty((n, 2), dtype=torch.int32, device="cuda")
    anyp = torch.empty(n, dtype=torch.int32, device="cuda")
    torch.cuda.synchronize()
    log(f"setup {time.time() - t0:.1f}s: {idx.shape[0]} tris, {info.n_inner} nodes, "
        f"depth {info.max_depth}, SAH build {info.build_seconds:.2f}s")

    # Per-ray work counted by the instrumented kernel over the whole batch
    # (the same traversal; outside the timed region).
    nc = torch.empty(n, dtype=torch.int32, device="cuda")
    pc = torch.empty(n, dtype=torch.int32, device="cuda")
    rumi.trace_count(ctx, bvh, rays, False, hits, nc, pc)
    torch.cuda.synchronize()
    nodes_c = float(nc.double().mean())
    tris_c = float(pc.double().mean())
    hit_frac = float((hits[:, 0] >= 0).double().mean())
    rumi.trace_count(ctx, bvh, rays, True, hits, nc, pc)
    torch.cuda.synchronize()
    nodes_a = float(nc.double().mean())
    tris_a = float(pc.double().mean())
    del nc, pc
    b_ray_c = RAY_IN_BYTES + HIT_OUT_BYTES + NODE_BYTES * nodes_c + TRI_BYTES * tris_c
    b_ray_a = RAY_IN_BYTES + 4 + NODE_BYTES * nodes_a + TRI_BYTES * tris_a

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

    ms_c = timed(lambda: rumi.trace_closest(ctx, bvh, rays, hits, None, 0, stream),
                 args.steps, args.warmup)
    log(f"closest-hit: {ms_c:.3f} ms/step, {n / ms_c / 1e3:.1f} Mrays/s")
    ms_a = timed(lambda: rumi.trace_any(ctx, bvh, rays, anyp, 0, stream), args.steps, args.warmup)
    log(f"any-hit: {ms_a:.3f} ms/step, {n / ms_a / 1e3:.1f} Mrays/s")

    mrays_c = n / ms_c / 1e3
    mrays_a = n / ms_a / 1e3
    achieved_c = b_ray_c * n / (ms_c * 1e-3) / 1e9
    achieved_a = b_ray_a * n / (ms_a * 1e-3) / 1e9

    cpu = None
    if args.cpu_rays > 0:
        import oracle
        k = min(args.cpu_rays, n)
        log(f"CPU baseline: reference median-6 build + Travel on {k} rays, 1 thread ...")
        _, _, secs, build_s = oracle.time_median6_mesh(xyz, idx, rays_h[:k])
        cpu = {"value": k / secs / 1e6, "unit": "Mrays/s", "cores": 1, "kind": "port",
               "sample": f"first {k} rays of the same batch, same mesh; the oracle's C "
                         f"restatement of BVHAccel::BuildBVH (median-6) + Travel; "
                         f"traversal timed only ({secs:.2f}s; build {build_s:.2f}s untimed)",
               "host_cpu": _cpu_model(), "nproc": os.cpu_count()}
        log(f"CPU baseline: {cpu['value']:.3f} Mrays/s")

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
            "unit": "GB/s", "frac": round(achieved_c / HBM_PEAK_GBS, 4), "traffic": None,
            "bytes_per_ray": round(b_ray_c, 1), "nodes_per_ray": round(nodes_c, 3),
            "tris_per_ray": round(tris_c, 3),
        },
        "cpu_baseline": cpu,
        "any_hit": {
            "value": round(mrays_a, 2), "unit": "Mrays/s", "ms_per_step": round(ms_a, 4),
            "roofline_achieved_GBs": round(achieved_a, 1),
            "roofline_frac": round(achieved_a / HBM_PEAK_GBS, 4),
            "bytes_per_ray": round(b_ray_a, 1), "nodes_per_ray": round(nodes_a, 3),
            "tris_per_ray": round(tris_a, 3),
        },
    }
    print(json.dumps(out), flush=True)


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
