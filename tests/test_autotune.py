"""The autotuner with a fake clock: deterministic choice, ties, the context
left on the winner."""
import importlib
import os
import sys

import pytest

sys.path.insert(0, os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))),
                                "rumi-renderer_amd"))
autotune = importlib.import_module("autotune")


class FakeCtx:
    def __init__(self):
        self.current = None
        self.history = []

    def set_trace_tuning(self, **t):
        self.current = dict(t)
        self.history.append(dict(t))


def test_trace_choice_follows_the_clock():
    ctx = FakeCtx()

    def cost(t):  # a made-up landscape: best at block 128, stack 8, refill 32, kind 1
        return (abs(t["block"] - 128) / 64 + abs(t["stack"] - 8) / 8 + abs(t["refill"] - 32) / 8
                + (5.0 if t["persistent"] < 0 else 0.0) + (0.5 if t["persistent"] == 2 else 0.0)
                + (0.1 if t["width"] == 4 else 0.0))

    timer = lambda fn: (fn(), cost(ctx.current))[1]
    calls = []
    best, table = autotune.tune_trace(ctx, lambda: calls.append(1), timer=timer)
    assert best == dict(block=128, stack=8, refill=32, persistent=1, width=2)
    assert ctx.current == best
    assert len(table) == len(autotune.trace_candidates()) == 1 + 2 * 3 * 2 + 3 * 2 * 4
    assert len(calls) == len(table)
    assert table[0][0]["persistent"] == -1


def test_trace_ties_keep_the_earlier_candidate():
    ctx = FakeCtx()
    cands = [dict(block=64, stack=16, refill=8, persistent=1),
             dict(block=256, stack=16, refill=8, persistent=1)]
    best, _ = autotune.tune_trace(ctx, lambda: None, cands, timer=lambda fn: 1.0)
    assert best == cands[0]


def test_render_choice_and_candidates():
    cands = autotune.render_candidates(capacities=(1000, 4000), regen_fractions=(1.0, 0.5))
    assert cands == [dict(queue_capacity=1000, regen_threshold=1000),
                     dict(queue_capacity=1000, regen_threshold=500),
                     dict(queue_capacity=4000, regen_threshold=4000),
                     dict(queue_capacity=4000, regen_threshold=2000)]
    seen = []

    def run(queue_capacity, regen_threshold):
        seen.append((queue_capacity, regen_threshold))

    clock = {"t": 0.0}

    def timer(fn):
        fn()
        q, r = seen[-1]
        return 1.0 / q + (0.001 if r < q else 0.0)

    best, table = autotune.tune_render(run, cands, timer=timer)
    assert best == dict(queue_capacity=4000, regen_threshold=4000)
    assert [c for c, _ in table] == cands


def test_empty_candidates_raise():
    with pytest.raises(ValueError):
        autotune.tune_trace(FakeCtx(), lambda: None, [], timer=lambda fn: 1.0)


def test_native_candidates_equal_the_python_list():
    """The C ABI's tuner (rumi_autotune_trace) times exactly these
    candidates, in this order (a host-only call: no GPU needed)."""
    import importlib
    rumi = importlib.import_module("rumi-renderer_amd")
    assert rumi.autotune_trace_candidates() == autotune.trace_candidates()


def test_batch_generation_size_choice_ties_and_release():
    """tune_batch builds, times and releases every candidate in order and
    keeps the fastest; a tie keeps the smaller generation."""
    built, released = [], []
    cost = {1 << 20: 5.0, 1 << 21: 3.0, 1 << 22: 2.0, 1 << 23: 2.0, 1 << 24: 4.0}

    def build(cap):
        built.append(cap)
        return {"cap": cap}

    best, table = autotune.tune_batch(build, lambda b: cost[b["cap"]], capacities=sorted(cost),
                                      release=lambda b: released.append(b["cap"]))
    assert best == 1 << 22
    assert built == released == sorted(cost)
    assert table == [(c, cost[c]) for c in sorted(cost)]
    with pytest.raises(ValueError):
        autotune.tune_batch(build, lambda b: 1.0, capacities=[])


def test_batch_default_candidates_reach_the_whole_config4_batch():
    """The default generation sizes run from 2^20 to 2^28 paths (the 256-scene
    workload's 170.8M samples fit one generation of 2^28)."""
    import inspect
    caps = inspect.signature(autotune.tune_batch).parameters["capacities"].default
    assert list(caps) == [1 << k for k in range(20, 29)]
    assert caps[-1] >= 170_755_624
