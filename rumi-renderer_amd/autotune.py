"""Autotuner for the launch shapes of the traversal and the wavefront renderer.

The product's tuner is native (rumi_autotune_trace / rumi_autotune_render in
the C ABI, bound as rumi.autotune_trace and Scene.autotune): it times the
same traversal candidates as trace_candidates() below, in the same order.
This module keeps the tuning policy with a pluggable clock (so tests can
drive it with a fake one) and tunes what only the Python layer builds: the
generation size of a batch (tune_batch).

Traversal: threads per block, LDS stack entries per lane, the refill
threshold of the persistent-lane kernels and the node width (binary or
4-wide) of SAH trees (rumi_trace_set_tuning).  Renderer: the wavefront queue
capacity and regeneration threshold (rumi_render_desc).  Every traversal
candidate gives the same hit distances (and, within one node width, the same
bits); renderer candidates give the same image up to the order of float
atomics; so the tuner only times.

The timer is pluggable: any callable that runs a zero-argument function and
returns its duration in seconds.  The default times on the device with HIP
events (median of `repeats` runs after one warm-up run); tests pass a fake
clock to make the choice deterministic.
"""
from __future__ import annotations

import itertools
from typing import Callable, Iterable, Sequence

Timer = Callable[[Callable[[], object]], float]

TRACE_BLOCKS = (64, 128, 256)
TRACE_STACKS = (8, 16)
TRACE_REFILLS = (16, 32, 48, 64)
TRACE_WIDTHS = (2, 4)


class HipEventTimer:
    """Median device time of fn over `repeats` runs (after `warmup` runs),
    measured with events on torch's current stream."""

    def __init__(self, repeats: int = 3, warmup: int = 1):
        self.repeats = repeats
        self.warmup = warmup

    def __call__(self, fn: Callable[[], object]) -> float:
        import torch
        for _ in range(self.warmup):
            fn()
        stream = torch.cuda.current_stream()
        times = []
        for _ in range(self.repeats):
            a = torch.cuda.Event(enable_timing=True)
            b = torch.cuda.Event(enable_timing=True)
            a.record(stream)
            fn()
            b.record(stream)
            b.synchronize()
            times.append(a.elapsed_time(b) * 1e-3)
        times.sort()
        return times[len(times) // 2]


def trace_candidates(blocks: Sequence[int] = TRACE_BLOCKS, stacks: Sequence[int] = TRACE_STACKS,
                     refills: Sequence[int] = TRACE_REFILLS, widths: Sequence[int] = TRACE_WIDTHS,
                     include_one_ray_per_lane=True):
    """Tuning dicts to try: the one-ray-per-lane launch first, then the
    prefetching lanes (persistent 2) for every width, block and stack size,
    then the threshold refill (persistent 1, binary nodes) for every block,
    stack and threshold."""
    out = []
    if include_one_ray_per_lane:
        out.append(dict(block=128, stack=16, refill=16, persistent=-1, width=2))
    for w, b, s in itertools.product(widths, blocks, stacks):
        out.append(dict(block=b, stack=s, refill=16, persistent=2, width=w))
    for b, s, r in itertools.product(blocks, stacks, refills):
        out.append(dict(block=b, stack=s, refill=r, persistent=1, width=2))
    return out


def tune_trace(ctx, run: Callable[[], object], candidates: Iterable[dict] | None = None,
               timer: Timer | None = None):
    """Times run() under every traversal tuning, leaves the fastest set on ctx.
    Returns (best, [(tuning, seconds), ...]) in candidate order; ties keep the
    earlier candidate."""
    timer = timer or HipEventTimer()
    results = []
    best, best_t = None, float("inf")
    for cand in (candidates if candidates is not None else trace_candidates()):
        ctx.set_trace_tuning(**cand)
        t = float(timer(run))
        results.append((dict(cand), t))
        if t < best_t:
            best, best_t = dict(cand), t
    if best is None:
        raise ValueError("tune_trace: no candidates")
    ctx.set_trace_tuning(**best)
    return best, results


def render_candidates(capacities: Sequence[int] = (1 << 19, 1 << 20, 1 << 21, 1 << 22, 1 << 23),
                      regen_fractions: Sequence[float] = (1.0, 0.5)):
    """(queue_capacity, regen_threshold) pairs; the threshold is a fraction of
    the capacity (refill when fewer paths than that are alive)."""
    return [dict(queue_capacity=c, regen_threshold=max(1, int(c * f)))
            for c in capacities for f in regen_fractions]


def tune_render(run: Callable[..., object], candidates: Iterable[dict] | None = None,
                timer: Timer | None = None):
    """run(**candidate) renders once with the given queue_capacity and
    regen_threshold; returns (best, [(candidate, seconds), ...])."""
    timer = timer or HipEventTimer(repeats=1, warmup=1)
    results = []
    best, best_t = None, float("inf")
    for cand in (candidates if candidates is not None else render_candidates()):
        t = float(timer(lambda: run(**cand)))
        results.append((dict(cand), t))
        if t < best_t:
            best, best_t = dict(cand), t
    if best is None:
        raise ValueError("tune_render: no candidates")
    return best, results


def tune_batch(build: Callable[[int], object], measure: Callable[[object], float],
               capacities: Sequence[int] = tuple(1 << k for k in range(20, 29)),
               release: Callable[[object], None] | None = None):
    """Generation size of a batch: build(capacity) makes the batch,
    measure(batch) returns the seconds of one run (the timer: for a captured
    graph, one replay after the capture), release(batch) frees it.  Returns
    (best capacity, [(capacity, seconds), ...]); ties keep the smaller.  The
    default candidates reach 2^28 paths in flight (about 63 GB of wavefront
    queues at 236 B per path)."""
    results = []
    best, best_t = None, float("inf")
    for cap in capacities:
        b = build(cap)
        try:
            t = float(measure(b))
        finally:
            if release:
                release(b)
        results.append((cap, t))
        if t < best_t:
            best, best_t = cap, t
    if best is None:
        raise ValueError("tune_batch: no capacities")
    return best, results
