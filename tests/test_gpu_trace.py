"""GPU traversal parity: the HIP kernels (through the C ABI) against the oracle.

Bar: hit primitive ids and hit distances bit-exact against the oracle's
traversal of the same tree in the same visiting order; against brute force
(BVH-independent) ids equal up to ties within 1e-6 relative t.
"""
import numpy as np
import pytest

from conftest import tie_ok

pytestmark = pytest.mark.gpu


def dev(a):
    import torch
    return torch.from_numpy(np.ascontiguousarray(a)).cuda()


def closest(rumi, ctx, bvh, rays, flags=0, uv=False):
    import torch
    n = rays.shape[0]
    r = dev(rays)
    hits = torch.full((max(n, 1), 2), -7, dtype=torch.int32, device="cuda")[:n]
    uvt = torch.zeros((max(n, 1), 2), dtype=torch.float32, device="cuda")[:n] if uv else None
    rumi.trace_closest(ctx, bvh, r, hits, uvt, flags)
    torch.cuda.synchronize()
    p, t = rumi.split_hits(hits.cpu().numpy())
    if uv:
        return p, t, uvt.cpu().numpy()
    return p, t


def anyhit(rumi, ctx, bvh, rays, flags=0):
    import torch
    n = rays.shape[0]
    out = torch.full((max(n, 1),), -7, dtype=torch.int32, device="cuda")[:n]
    rumi.trace_any(ctx, bvh, dev(rays), out, flags)
    torch.cuda.synchronize()
    return out.cpu().numpy()


def counts(rumi, ctx, bvh, rays, any_hit=False, flags=0):
    import torch
    n = rays.shape[0]
    hits = torch.empty((n, 2), dtype=torch.int32, device="cuda")
    nc = torch.empty(n, dtype=torch.int32, device="cuda")
    pc = torch.empty(n, dtype=torch.int32, device="cuda")
    rumi.trace_count(ctx, bvh, dev(rays), any_hit, hits, nc, pc, flags)
    torch.cuda.synchronize()
    p, t = rumi.split_hits(hits.cpu().numpy())
    return p, t, nc.cpu().numpy(), pc.cpu().numpy()


def bits(t):
    return t.view(np.int32)


@pytest.fixture(scope="module")
def mesh(rumi):
    return rumi.make_icosphere(12, 1, 0.05)


@pytest.fixture(scope="module")
def rays(rumi):
    return rumi.make_rays_uniform(40000, 2, 2.1)


@pytest.mark.parametrize("max_leaf", [1, 4, 8])
def test_closest_sah_bitexact_vs_oracle(rumi, oracle, ctx, mesh, rays, max_leaf):
    xyz, idx = mesh
    b = rumi.Bvh(ctx, xyz, idx, max_leaf=max_leaf)
    ob = oracle.MedianBvh.from_export(b.export(), idx.shape[0])
    p, t = closest(rumi, ctx, b, rays)
    op, ot = ob.trace_mesh(xyz, idx, rays, ordered=True)
    np.testing.assert_array_equal(p, op)
    np.testing.assert_array_equal(bits(t), bits(ot))
    bp, bt = oracle.brute_mesh(xyz, idx, rays[:5000])
    assert ((p[:5000] == bp) | tie_ok(t[:5000], bt)).all()
    assert (p >= 0).sum() > 1000


def test_closest_reference_tree_and_order_bitexact(rumi, oracle, ctx, mesh, rays):
    """REF_MEDIAN6 tree + reference visiting order == the oracle's restatement
    of BVHAccel::BuildBVH + Travel, including which of two tied triangles wins."""
    xyz, idx = mesh
    b = rumi.Bvh(ctx, xyz, idx, mode=rumi.RUMI_BVH_REF_MEDIAN6)
    lo, hi = oracle.mesh_bounds(xyz, idx)
    ob = oracle.MedianBvh(lo, hi)
    p, t = closest(rumi, ctx, b, rays)
    op, ot = ob.trace_mesh(xyz, idx, rays, ordered=False)
    np.testing.assert_array_equal(p, op)
    np.testing.assert_array_equal(bits(t), bits(ot))
    pa = anyhit(rumi, ctx, b, rays)
    oa, _ = ob.trace_mesh(xyz, idx, rays, any_hit=True, ordered=False)
    np.testing.assert_array_equal(pa, oa)


@pytest.mark.parametrize("flags,ordered", [(1, False), (2, True)])
def test_order_override_flags(rumi, oracle, ctx, mesh, rays, flags, ordered):
    xyz, idx = mesh
    b = rumi.Bvh(ctx, xyz, idx)
    ob = oracle.MedianBvh.from_export(b.export(), idx.shape[0])
    p, t = closest(rumi, ctx, b, rays, flags=flags)
    op, ot = ob.trace_mesh(xyz, idx, rays, ordered=ordered)
    np.testing.assert_array_equal(p, op)
    np.testing.assert_array_equal(bits(t), bits(ot))


def test_any_hit_bitexact_and_status(rumi, oracle, ctx, mesh, rays):
    xyz, idx = mesh
    b = rumi.Bvh(ctx, xyz, idx)
    ob = oracle.MedianBvh.from_export(b.export(), idx.shape[0])
    pa = anyhit(rumi, ctx, b, rays)
    oa, _ = ob.trace_mesh(xyz, idx, rays, any_hit=True, ordered=True)
    np.testing.assert_array_equal(pa, oa)
    pc, _ = closest(rumi, ctx, b, rays)
    np.testing.assert_array_equal(pa >= 0, pc >= 0)


@pytest.mark.parametrize("any_hit", [False, True])
def test_counts_match_oracle(rumi, oracle, ctx, mesh, rays, any_hit):
    xyz, idx = mesh
    b = rumi.Bvh(ctx, xyz, idx)
    ob = oracle.MedianBvh.from_export(b.export(), idx.shape[0])
    p, t, nc, pc = counts(rumi, ctx, b, rays, any_hit=any_hit)
    op, ot, onc, opc = ob.trace_mesh(xyz, idx, rays, any_hit=any_hit, ordered=True, counts=True)
    np.testing.assert_array_equal(p, op)
    np.testing.assert_array_equal(nc, onc)
    np.testing.assert_array_equal(pc, opc)


@pytest.fixture(autouse=True)
def binary(ctx):
    """Tests compare with the oracle's binary near-first order unless they
    ask for the 4-wide tree (wide4) or set a tuning themselves."""
    ctx.set_trace_tuning(width=2)
    yield
    ctx.set_trace_tuning()


@pytest.fixture
def wide4(ctx):
    ctx.set_trace_tuning(width=4)
    yield
    ctx.set_trace_tuning()


@pytest.mark.parametrize("max_leaf", [1, 4, 8])
def test_wide_closest_any_counts_bitexact_vs_oracle(rumi, oracle, ctx, mesh, rays, max_leaf, wide4):
    """Width 4: the GPU collapses the SAH tree to 4-wide nodes and visits
    children nearest first; the oracle restates the same collapse and order,
    so ids, distances, any-hit ids and node/primitive counts agree bit for
    bit.  Brute force pins the ids up to ties."""
    xyz, idx = mesh
    b = rumi.Bvh(ctx, xyz, idx, max_leaf=max_leaf)
    ob = oracle.MedianBvh.from_export(b.export(), idx.shape[0])
    p, t = closest(rumi, ctx, b, rays)
    op, ot = ob.trace_mesh(xyz, idx, rays, ordered=2)
    np.testing.assert_array_equal(p, op)
    np.testing.assert_array_equal(bits(t), bits(ot))
    bp, bt = oracle.brute_mesh(xyz, idx, rays[:5000])
    assert ((p[:5000] == bp) | tie_ok(t[:5000], bt)).all()
    pa = anyhit(rumi, ctx, b, rays)
    oa, _ = ob.trace_mesh(xyz, idx, rays, any_hit=True, ordered=2)
    np.testing.assert_array_equal(pa, oa)
    for any_hit in (False, True):
        cp, _, nc, pc = counts(rumi, ctx, b, rays, any_hit=any_hit)
        xp, _, onc, opc = ob.trace_mesh(xyz, idx, rays, any_hit=any_hit, ordered=2, counts=True)
        np.testing.assert_array_equal(cp, xp)
        np.testing.assert_array_equal(nc, onc)
        np.testing.assert_array_equal(pc, opc)


def test_wide_edge_cases(rumi, oracle, ctx, rays, wide4):
    # a one-leaf tree (root is a leaf), a tree with one interior node, and
    # the full-reference order flag, which ignores the width
    xyz, idx = rumi.make_icosphere(1, 3, 0.0)
    for nt in (1, 3, 9):
        b = rumi.Bvh(ctx, xyz, idx[:nt], max_leaf=4)
        ob = oracle.MedianBvh.from_export(b.export(), nt)
        p, t = closest(rumi, ctx, b, rays[:4096])
        op, ot = ob.trace_mesh(xyz, idx[:nt], rays[:4096], ordered=2)
        np.testing.assert_array_equal(p, op)
        np.testing.assert_array_equal(bits(t), bits(ot))
    b = rumi.Bvh(ctx, xyz, idx)
    ob = oracle.MedianBvh.from_export(b.export(), idx.shape[0])
    p, t = closest(rumi, ctx, b, rays, flags=1)
    op, ot = ob.trace_mesh(xyz, idx, rays, ordered=0)
    np.testing.assert_array_equal(p, op)
    np.testing.assert_array_equal(bits(t), bits(ot))


def axis_rays_on_planes(boxes, n, seed):
    """Axis-aligned rays whose origins lie on face planes of real boxes in
    both perpendicular axes: slab products 0 * inf = NaN, with +0 and -0
    direction components."""
    rng = np.random.default_rng(seed)
    rays = np.zeros((n, 8), np.float32)
    axis = rng.integers(0, 3, n)
    sign = rng.choice(np.float32([-1.0, 1.0]), n)
    pick = boxes[rng.integers(0, len(boxes), n)]
    for k in range(3):
        rays[:, k] = np.where(rng.integers(0, 2, n) == 0, pick[:, k], pick[:, 3 + k])
    rays[np.arange(n), axis] = -3.0 * sign
    d = np.zeros((n, 3), np.float32)
    d[np.arange(n), axis] = sign
    d[(d == 0) & (rng.integers(0, 2, (n, 3)) == 1)] = np.float32(-0.0)
    rays[:, 4:7] = d
    rays[:, 7] = np.inf
    return rays


def seq_vs_minmax_box_tests_differ(rays, boxes):
    """Count ray/box pairs where the reference's sequential slab test
    (src/Acceleration/bbox.cpp) and a plain min/max slab test disagree."""
    with np.errstate(all="ignore"):
        o, d = rays[:, None, 0:3], rays[:, None, 4:7]
        inv = np.float32(1.0) / d
        t0 = (boxes[None, :, 0:3] - o) * inv
        t1 = (boxes[None, :, 3:6] - o) * inv
        lo = np.where(t0 > t1, t1, t0)
        hi = np.where(t0 > t1, t0, t1)
        tn, tf = lo[..., 0], hi[..., 0]
        ok = np.ones(tn.shape, bool)
        for k in (1, 2):
            ok &= ~((tn > hi[..., k]) | (lo[..., k] > tf))
            tn = np.where(tn < lo[..., k], lo[..., k], tn)
            tf = np.where(tf > hi[..., k], hi[..., k], tf)
        seq = ok & (tf >= 0) & (np.inf >= tn)
        mn, mx = np.fmin(t0, t1), np.fmax(t0, t1)
        tn2, tf2 = np.fmax.reduce(mn, axis=-1), np.fmin.reduce(mx, axis=-1)
        naive = (tn2 <= tf2) & (tf2 >= 0) & (np.inf >= tn2)
    return int((seq != naive).sum())


@pytest.mark.parametrize("width", [2, 4])
def test_axis_aligned_rays_on_box_planes(rumi, oracle, ctx, mesh, width):
    """The branch-free box test keeps the sequential test's NaN behaviour:
    rays with zero direction components starting on face planes of the
    tree's own boxes give the oracle's ids and distances bit for bit."""
    xyz, idx = mesh
    b = rumi.Bvh(ctx, xyz, idx, max_leaf=4)
    ex = b.export()
    boxes = ex["child_boxes"].reshape(-1, 6)
    ar = axis_rays_on_planes(boxes, 20000, 5)
    # the inputs reach the cases where a min/max test would decide otherwise
    assert seq_vs_minmax_box_tests_differ(ar[:400], boxes[::7]) > 0
    ob = oracle.MedianBvh.from_export(ex, idx.shape[0])
    ctx.set_trace_tuning(width=width)
    try:
        p, t = closest(rumi, ctx, b, ar)
        pa = anyhit(rumi, ctx, b, ar)
    finally:
        ctx.set_trace_tuning()
    op, ot = ob.trace_mesh(xyz, idx, ar, ordered=1 if width == 2 else 2)
    np.testing.assert_array_equal(p, op)
    np.testing.assert_array_equal(bits(t), bits(ot))
    oa, _ = ob.trace_mesh(xyz, idx, ar, any_hit=True, ordered=1 if width == 2 else 2)
    np.testing.assert_array_equal(pa, oa)
    assert (p >= 0).sum() > 1000


def test_barycentrics_reconstruct_hit_point(rumi, ctx, mesh, rays):
    xyz, idx = mesh
    b = rumi.Bvh(ctx, xyz, idx)
    p, t, uv = closest(rumi, ctx, b, rays, uv=True)
    m = p >= 0
    tri = idx[p[m]]
    v0, v1, v2 = (xyz[tri[:, k]].astype(np.float64) for k in range(3))
    u, v = uv[m, 0:1].astype(np.float64), uv[m, 1:2].astype(np.float64)
    on_tri = v0 + u * (v1 - v0) + v * (v2 - v0)
    on_ray = rays[m, :3].astype(np.float64) + t[m, None].astype(np.float64) * rays[m, 4:7]
    assert np.abs(on_tri - on_ray).max() < 1e-5
    assert (u >= 0).all() and (v >= 0).all() and (u + v <= 1 + 1e-7).all()


@pytest.mark.parametrize("width", [2, 4])
def test_edge_cases(rumi, oracle, ctx, mesh, rays, width):
    """Ragged batches, empty batches, zero directions, finite segments and a
    caller's closest-so-far, on both tree widths (width 4: the persistent
    kernel whose waves share their triangle tests, with partly filled waves
    in the ragged batches)."""
    xyz, idx = mesh
    ctx.set_trace_tuning(width=width)
    ordered = 2 if width == 4 else True
    b = rumi.Bvh(ctx, xyz, idx)
    ob = oracle.MedianBvh.from_export(b.export(), idx.shape[0])
    # ragged sizes around the block size
    for n in (1, 63, 64, 127, 128, 129, 1031):
        p, t = closest(rumi, ctx, b, rays[:n])
        op, ot = ob.trace_mesh(xyz, idx, rays[:n], ordered=ordered)
        np.testing.assert_array_equal(p, op)
        np.testing.assert_array_equal(bits(t), bits(ot))
        oa, _ = ob.trace_mesh(xyz, idx, rays[:n], any_hit=True, ordered=ordered)
        np.testing.assert_array_equal(anyhit(rumi, ctx, b, rays[:n]), oa)
    # n = 0 is a no-op
    p, t = closest(rumi, ctx, b, rays[:0])
    assert p.size == 0
    # zero direction: BVHAccel::Travel returns -1 (bvh.cpp:161-162)
    z = rays[:64].copy()
    z[:, 4:7] = 0
    p, t = closest(rumi, ctx, b, z)
    assert (p == -1).all() and (t == 0).all()
    # finite tmax (shadow segment shorter than the hit) misses; longer finds it
    p0, t0 = closest(rumi, ctx, b, rays[:4000])
    seg = rays[:4000].copy()
    hit = p0 >= 0
    seg[hit, 7] = t0[hit] * 0.5
    p1, _ = closest(rumi, ctx, b, seg)
    assert (p1[hit] == -1).all()
    a1 = anyhit(rumi, ctx, b, seg)
    assert (a1[hit] == -1).all()
    op1, _ = ob.trace_mesh(xyz, idx, seg, ordered=ordered)
    np.testing.assert_array_equal(p1, op1)
    # a caller-supplied "closest so far" (tmin > 0) is honoured
    pre = rays[:4000].copy()
    pre[hit, 3] = t0[hit] * np.float32(0.999)
    p2, t2 = closest(rumi, ctx, b, pre)
    op2, ot2 = ob.trace_mesh(xyz, idx, pre, ordered=ordered)
    np.testing.assert_array_equal(p2, op2)
    np.testing.assert_array_equal(bits(t2), bits(ot2))


@pytest.mark.parametrize("width", [2, 4])
def test_empty_and_single_triangle_meshes(rumi, oracle, ctx, rays, width):
    ctx.set_trace_tuning(width=width)
    xyz = np.array([[-1, -1, 0], [1, -1, 0], [0, 1, 0]], np.float32)
    empty = rumi.Bvh(ctx, xyz, np.zeros((0, 3), np.uint32))
    p, t = closest(rumi, ctx, empty, rays[:500])
    assert (p == -1).all()
    assert (anyhit(rumi, ctx, empty, rays[:500]) == -1).all()
    one = rumi.Bvh(ctx, xyz, np.array([[0, 1, 2]], np.uint32))
    bp, bt = oracle.brute_mesh(xyz, np.array([[0, 1, 2]], np.uint32), rays)
    p, t = closest(rumi, ctx, one, rays)
    np.testing.assert_array_equal(p, bp)
    np.testing.assert_array_equal(bits(t), bits(bt))


def test_errors_are_reported_not_crashing(rumi, ctx, mesh):
    import torch
    xyz, idx = mesh
    b = rumi.Bvh(ctx, xyz, idx)
    host_only = rumi.Bvh(None, xyz, idx)
    r = torch.zeros((8, 8), dtype=torch.float32, device="cuda")
    h = torch.zeros((8, 2), dtype=torch.int32, device="cuda")
    with pytest.raises(rumi.RumiError):
        rumi.trace_closest(ctx, host_only, r, h)            # BVH not on the device
    with pytest.raises(rumi.RumiError):
        rumi.trace_closest(ctx, b, r, h, flags=3)           # exclusive flags
    with pytest.raises(rumi.RumiError):
        rumi.trace_closest(ctx, b, r, None)                 # null output


@pytest.fixture(scope="module")
def full_config(rumi, oracle):
    """The 1,310,720-triangle noisy icosphere, 2^23 uniform rays, and the
    reference's own tree (BVHAccel's median-6 build, traversed in its order
    by the oracle) on a 65,536-ray slice."""
    xyz, idx = rumi.make_icosphere(256, 1, 0.05)
    rays = rumi.make_rays_uniform(1 << 23, 2, 2.1)
    sl = slice(1 << 22, (1 << 22) + 65536)
    rp, rt, _, _ = oracle.time_median6_mesh(xyz, idx, rays[sl])
    return xyz, idx, rays, sl, rp, rt


@pytest.mark.parametrize("width", [2, 4])
def test_full_config_icosphere_1p3m_8m_rays(rumi, oracle, ctx, full_config, width):
    """The benchmarked workload at both tree widths (4 is the benchmarked
    kernel: its LDS-only pushes and private spills are live at this depth):
    the whole batch through size-independent properties, a 65,536-ray slice
    bit-exactly against the oracle's traversal of the same tree, and that
    slice against the reference's median-6 tree: the same hit distance bit
    for bit and the same primitive except between primitives at equal
    distance (ties)."""
    import torch
    xyz, idx, rays, sl, rp, rt = full_config
    ctx.set_trace_tuning(width=width)
    print(f"full config, width {width}", flush=True)
    b = rumi.Bvh(ctx, xyz, idx)
    assert b.info().n_prims == 1310720
    p, t = closest(rumi, ctx, b, rays)
    pa = anyhit(rumi, ctx, b, rays)
    np.testing.assert_array_equal(pa >= 0, p >= 0)
    m = p >= 0
    hp = rays[m, :3].astype(np.float64) + t[m, None].astype(np.float64) * rays[m, 4:7]
    r = np.linalg.norm(hp, axis=1)
    assert r.min() > 0.95 - 1e-4 and r.max() < 1.05 + 1e-4
    # Rays starting inside the shell almost all hit; Moller-Trumbore is not
    # watertight, so a ray through a shared edge can pass between two
    # triangles: every such leak must be a miss under brute force too.
    inside = np.linalg.norm(rays[:, :3].astype(np.float64), axis=1) < 0.95
    leaks = np.flatnonzero(inside & ~m)
    assert len(leaks) <= 16
    if len(leaks):
        lp, _ = oracle.brute_mesh(xyz, idx, rays[leaks])
        assert (lp == -1).all()
    ob = oracle.MedianBvh.from_export(b.export(), idx.shape[0])
    order = 2 if width == 4 else True
    op, ot = ob.trace_mesh(xyz, idx, rays[sl], ordered=order)
    np.testing.assert_array_equal(p[sl], op)
    np.testing.assert_array_equal(bits(t[sl]), bits(ot))
    oa, _ = ob.trace_mesh(xyz, idx, rays[sl], any_hit=True, ordered=order)
    np.testing.assert_array_equal(pa[sl], oa)
    # the reference's tree: same distances, ids equal up to exact ties
    np.testing.assert_array_equal(bits(t[sl]), bits(rt))
    diff = p[sl] != rp
    assert np.all(tie_ok(t[sl][diff], rt[diff]))
    print(f"width {width}: {int(diff.sum())} of {len(rp)} ids differ from the median-6 tree (ties)")
    bp, bt = oracle.brute_mesh(xyz, idx, rays[:128])
    assert ((p[:128] == bp) | tie_ok(t[:128], bt)).all()
    del b
    torch.cuda.synchronize()


def test_every_tuning_gives_identical_hits(rumi, ctx, small_mesh):
    """Block size, LDS stack depth (with spill), refill threshold and the
    one-ray-per-lane launch change the schedule, never the result."""
    import importlib
    import os
    import sys
    import torch
    sys.path.insert(0, os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))),
                                    "rumi-renderer_amd"))
    autotune = importlib.import_module("autotune")
    xyz, idx = rumi.make_icosphere(48, 1, 0.05)
    bvh = rumi.Bvh(ctx, xyz, idx)
    rays = torch.from_numpy(rumi.make_rays_uniform(1 << 16, 9, 2.1)).cuda()
    ref_h = torch.empty((1 << 16, 2), dtype=torch.int32, device="cuda")
    ref_a = torch.empty(1 << 16, dtype=torch.int32, device="cuda")
    refs = {}
    for width in (2, 4):
        rh, ra = torch.empty_like(ref_h), torch.empty_like(ref_a)
        ctx.set_trace_tuning(persistent=-1, width=width)
        rumi.trace_closest(ctx, bvh, rays, rh)
        rumi.trace_any(ctx, bvh, rays, ra)
        torch.cuda.synchronize()
        refs[width] = (rh, ra)
    # both widths: the same hit distances and hit status
    assert torch.equal(refs[2][0][:, 1], refs[4][0][:, 1])
    assert torch.equal(refs[2][1] >= 0, refs[4][1] >= 0)
    for t in autotune.trace_candidates():
        ctx.set_trace_tuning(**t)
        assert ctx.trace_tuning() == t
        h = torch.full_like(ref_h, 7)
        a = torch.full_like(ref_a, 7)
        rumi.trace_closest(ctx, bvh, rays, h)
        rumi.trace_any(ctx, bvh, rays, a)
        torch.cuda.synchronize()
        assert torch.equal(h, refs[t["width"]][0]), t
        assert torch.equal(a, refs[t["width"]][1]), t
    best, table = autotune.tune_trace(ctx, lambda: rumi.trace_closest(ctx, bvh, rays, ref_h))
    print("tuning", best, [(c, round(s * 1e6, 1)) for c, s in table])
    ctx.set_trace_tuning()
    assert ctx.trace_tuning() == dict(block=256, stack=16, refill=16, persistent=2, width=4)


def test_native_autotuner_picks_a_candidate_and_keeps_results(rumi, ctx, small_mesh):
    """rumi_autotune_trace times every candidate on the caller's batch, leaves
    the fastest on the context and the hits are those of any tuning."""
    import torch
    xyz, idx = small_mesh
    bvh = rumi.Bvh(ctx, xyz, idx)
    rays = torch.from_numpy(rumi.make_rays_uniform(1 << 15, 4, 2.1)).cuda()
    ref = torch.empty((1 << 15, 2), dtype=torch.int32, device="cuda")
    ctx.set_trace_tuning(width=4)
    rumi.trace_closest(ctx, bvh, rays, ref)
    out = torch.empty_like(ref)
    best, table = rumi.autotune_trace(ctx, bvh, rays, out)
    cands = rumi.autotune_trace_candidates()
    assert best in cands and ctx.trace_tuning() == best
    assert [c for c, _ in table] == cands and all(t > 0 for _, t in table)
    assert min(t for _, t in table) == dict((tuple(c.items()), t) for c, t in table)[tuple(best.items())]
    h = torch.full_like(ref, 7)
    rumi.trace_closest(ctx, bvh, rays, h)
    torch.cuda.synchronize()
    assert torch.equal(h[:, 1], ref[:, 1])       # same distances for every width
    ctx.set_trace_tuning()
