"""BVH builders and traversal on the CPU.

Pins the oracle's median-6 builder (src/Acceleration/bvh.cpp:48-148) by its
structural invariants and its traversal (bvh.cpp:150-212) by brute force;
then checks the product's host builders against the oracle and against the
same invariants.  No GPU calls.
"""
import numpy as np
import pytest

from conftest import tie_ok


def check_median6_invariants(a, lo, hi, n):
    left, right, first, count, order = a["left"], a["right"], a["first"], a["count"], a["order"]
    seen = np.zeros(n, np.int32)
    ranges = []

    def walk(node, start, end, pre_dim):
        box_lo, box_hi = a["lo"][node], a["hi"][node]
        if left[node] < 0:
            assert first[node] == start and count[node] == end - start + 1
            assert 1 <= count[node] <= 6
            ids = order[start:end + 1]
            seen[ids] += 1
            np.testing.assert_array_equal(box_lo, lo[ids].min(0))
            np.testing.assert_array_equal(box_hi, hi[ids].max(0))
            ranges.append((start, end))
            return
        assert end - start + 1 > 6
        cut = (start + end) >> 1
        ids = order[start:end + 1]
        ext = hi[ids].max(0) - lo[ids].min(0)
        dim = 0
        if ext[0] < ext[1] and ext[2] < ext[1]:
            dim = 1
        if ext[0] < ext[2] and ext[1] < ext[2]:
            dim = 2
        # The range was sorted on dim when this node was made (by this node,
        # or by its parent when the axis repeats); the children only permute
        # within their halves, so the cut still separates the centres.
        c = (hi[ids, dim] + lo[ids, dim]) * np.float32(0.5)
        assert c[:cut - start + 1].max() <= c[cut - start + 1:].min()
        walk(left[node], start, cut, dim)
        walk(right[node], cut + 1, end, dim)
        np.testing.assert_array_equal(box_lo, np.minimum(a["lo"][left[node]], a["lo"][right[node]]))
        np.testing.assert_array_equal(box_hi, np.maximum(a["hi"][left[node]], a["hi"][right[node]]))

    walk(a["root"], 0, n - 1, 3)
    assert np.all(seen == 1)


@pytest.mark.parametrize("n_tris_freq", [1, 3, 12])
def test_oracle_median6_invariants(rumi, oracle, n_tris_freq):
    xyz, idx = rumi.make_icosphere(n_tris_freq, 7, 0.05)
    lo, hi = oracle.mesh_bounds(xyz, idx)
    b = oracle.MedianBvh(lo, hi)
    check_median6_invariants(b.arrays(), lo, hi, idx.shape[0])


def test_oracle_median6_sorting_ties_follow_shell_sort(oracle):
    # Equal centres: the shell sort's moves are deterministic; a box list with
    # heavy ties must still give a valid, fully covered tree.
    rng = np.random.default_rng(3)
    lo = rng.integers(0, 4, size=(257, 3)).astype(np.float32)
    hi = lo + 1
    b = oracle.MedianBvh(lo, hi)
    check_median6_invariants(b.arrays(), lo, hi, 257)


def test_oracle_traversal_matches_brute_force(rumi, oracle, small_mesh):
    xyz, idx = small_mesh
    rays = rumi.make_rays_uniform(6000, 5, 2.1)
    lo, hi = oracle.mesh_bounds(xyz, idx)
    b = oracle.MedianBvh(lo, hi)
    bp, bt = oracle.brute_mesh(xyz, idx, rays)
    for ordered in (False, True):
        p, t = b.trace_mesh(xyz, idx, rays, ordered=ordered)
        assert np.array_equal(p >= 0, bp >= 0)
        same = (p == bp) | tie_ok(t, bt)
        assert same.all()
        assert (p == bp).mean() > 0.999
    # any-hit: hit status equals closest-hit status for unbounded rays
    pa, _ = b.trace_mesh(xyz, idx, rays, any_hit=True)
    assert np.array_equal(pa >= 0, bp >= 0)


def test_oracle_timed_traversal_threads_give_the_same_hits(rumi, oracle, small_mesh):
    """The bench's CPU-baseline driver (median-6 build + Travel per ray) on one
    thread and on several: the same hits bit for bit, and those of the
    oracle's own traversal in reference order."""
    xyz, idx = small_mesh
    rays = rumi.make_rays_uniform(5000, 6, 2.1)
    p1, t1, s1, b1 = oracle.time_median6_mesh(xyz, idx, rays)
    p4, t4, s4, b4 = oracle.time_median6_mesh(xyz, idx, rays, nthreads=4)
    lo, hi = oracle.mesh_bounds(xyz, idx)
    op, ot = oracle.MedianBvh(lo, hi).trace_mesh(xyz, idx, rays, ordered=False)
    np.testing.assert_array_equal(p1, p4)
    np.testing.assert_array_equal(t1.view(np.int32), t4.view(np.int32))
    np.testing.assert_array_equal(p1, op)
    assert s1 > 0 and s4 > 0 and b1 > 0


def test_oracle_traversal_inside_sphere_all_hit(rumi, oracle, small_mesh):
    xyz, idx = small_mesh
    rays = rumi.make_rays_uniform(2000, 9, 0.5)  # origins well inside the r >= 0.95 shell
    lo, hi = oracle.mesh_bounds(xyz, idx)
    p, t = oracle.MedianBvh(lo, hi).trace_mesh(xyz, idx, rays)
    assert (p >= 0).all()
    r = np.linalg.norm(rays[:, :3] + t[:, None] * rays[:, 4:7], axis=1)
    assert ((r > 0.949) & (r < 1.051)).all()


# ------------------------------------------------------------- product host side
def test_pcg32_known_answers(rumi):
    # O'Neill's pcg32-global demo, pcg32_srandom(42, 54): first six outputs.
    expect = [0xa15c02b7, 0x7b47f409, 0xba1d3330, 0x83d2f293, 0xbfa4784b, 0xcbed606e]
    assert rumi.pcg32_sequence(42, 54, 6).tolist() == expect


@pytest.mark.parametrize("f", [1, 2, 5, 16])
def test_icosphere_is_closed_oriented_manifold(rumi, f):
    xyz, idx = rumi.make_icosphere(f, 1, 0.05)
    assert xyz.shape == (10 * f * f + 2, 3) and idx.shape == (20 * f * f, 3)
    r = np.linalg.norm(xyz.astype(np.float64), axis=1)
    assert r.min() >= 0.95 - 1e-6 and r.max() <= 1.05 + 1e-6
    # Every directed edge appears once and its reverse once: closed, consistently oriented.
    e = np.concatenate([idx[:, [0, 1]], idx[:, [1, 2]], idx[:, [2, 0]]]).astype(np.int64)
    key = e[:, 0] * (10 * f * f + 2) + e[:, 1]
    rkey = e[:, 1] * (10 * f * f + 2) + e[:, 0]
    assert len(np.unique(key)) == len(key)
    assert np.array_equal(np.sort(key), np.sort(rkey))
    # Outward orientation: each face normal points away from the origin.
    a, b, c = (xyz[idx[:, k]].astype(np.float64) for k in range(3))
    nrm = np.cross(b - a, c - a)
    assert ((nrm * (a + b + c)).sum(1) > 0).all()
    # Euler characteristic of a sphere.
    assert xyz.shape[0] - len(key) // 2 + idx.shape[0] == 2


def test_rays_uniform_properties(rumi):
    rays = rumi.make_rays_uniform(200000, 2, 2.1)
    assert np.abs(rays[:, :3]).max() <= 2.1
    d = rays[:, 4:7].astype(np.float64)
    assert np.allclose(np.linalg.norm(d, axis=1), 1.0, atol=1e-6)
    assert np.abs(d.mean(0)).max() < 0.01               # isotropic
    assert abs((d[:, 2] ** 2).mean() - 1 / 3) < 0.01    # E[z^2] = 1/3 on S^2
    assert (rays[:, 3] == 0).all() and np.isinf(rays[:, 7]).all()
    # Deterministic: stream = ray index.
    np.testing.assert_array_equal(rumi.make_rays_uniform(10, 2, 2.1), rays[:10])


def flat_from_export(oracle, ex, n):
    return oracle.MedianBvh.from_export(ex, n)


def test_product_median6_equals_oracle(rumi, oracle):
    xyz, idx = rumi.make_icosphere(10, 3, 0.05)
    b = rumi.Bvh(None, xyz, idx, mode=rumi.RUMI_BVH_REF_MEDIAN6)
    ex = b.export()
    lo, hi = oracle.mesh_bounds(xyz, idx)
    ob = oracle.MedianBvh(lo, hi)
    np.testing.assert_array_equal(ex["leaf_prims"], ob.arrays()["order"])
    rays = rumi.make_rays_uniform(4000, 11, 2.1)
    p0, t0 = ob.trace_mesh(xyz, idx, rays)
    p1, t1 = flat_from_export(oracle, ex, idx.shape[0]).trace_mesh(xyz, idx, rays)
    np.testing.assert_array_equal(p0, p1)
    np.testing.assert_array_equal(t0.view(np.int32), t1.view(np.int32))
    # Same tree => same visit counts in the reference order.
    _, _, n0, c0 = ob.trace_mesh(xyz, idx, rays, counts=True)
    _, _, n1, c1 = flat_from_export(oracle, ex, idx.shape[0]).trace_mesh(xyz, idx, rays, counts=True)
    np.testing.assert_array_equal(n0, n1)
    np.testing.assert_array_equal(c0, c1)


@pytest.mark.parametrize("max_leaf", [1, 4, 8])
def test_product_sah_invariants(rumi, oracle, max_leaf):
    xyz, idx = rumi.make_icosphere(9, 4, 0.05)
    n = idx.shape[0]
    b = rumi.Bvh(None, xyz, idx, mode=rumi.RUMI_BVH_SAH_BINNED, max_leaf=max_leaf)
    info = b.info()
    ex = b.export()
    assert sorted(ex["leaf_prims"].tolist()) == list(range(n))
    lo, hi = oracle.mesh_bounds(xyz, idx)
    refs = ex["child_refs"]
    boxes = ex["child_boxes"]
    leaves = 0
    covered = np.zeros(n, np.int32)

    def check(ref, blo, bhi, depth):
        nonlocal leaves
        if ref < 0:
            enc = ~ref
            first, count = enc >> 4, enc & 15
            assert 1 <= count <= max_leaf
            ids = ex["leaf_prims"][first:first + count]
            covered[first:first + count] += 1
            np.testing.assert_array_equal(blo, lo[ids].min(0))
            np.testing.assert_array_equal(bhi, hi[ids].max(0))
            leaves += 1
            return depth
        d = 0
        sub_lo, sub_hi = [], []
        for k in range(2):
            cb = boxes[ref, 6 * k:6 * k + 6]
            sub_lo.append(cb[:3])
            sub_hi.append(cb[3:])
            d = max(d, check(refs[ref, k], cb[:3], cb[3:], depth + 1))
        np.testing.assert_array_equal(blo, np.minimum(*sub_lo))
        np.testing.assert_array_equal(bhi, np.maximum(*sub_hi))
        return d

    depth = check(ex["root_ref"], ex["root_box"][:3], ex["root_box"][3:], 0)
    assert (covered == 1).all()
    assert leaves == info.n_leaves and info.n_inner == leaves - 1
    assert depth == info.max_depth
    # SAH beats the median split on this mesh.
    med = rumi.Bvh(None, xyz, idx, mode=rumi.RUMI_BVH_REF_MEDIAN6).info()
    if max_leaf == 4:
        assert info.sah_cost < med.sah_cost


def test_product_sah_traversal_matches_brute_force(rumi, oracle, small_mesh):
    xyz, idx = small_mesh
    b = rumi.Bvh(None, xyz, idx)
    ob = flat_from_export(oracle, b.export(), idx.shape[0])
    rays = rumi.make_rays_uniform(5000, 13, 2.1)
    bp, bt = oracle.brute_mesh(xyz, idx, rays)
    p, t = ob.trace_mesh(xyz, idx, rays, ordered=True)
    assert ((p == bp) | tie_ok(t, bt)).all()


@pytest.mark.parametrize("max_leaf", [1, 4])
def test_oracle_wide_traversal_matches_binary_and_brute_force(rumi, oracle, small_mesh, max_leaf):
    # The 4-wide form visits the same primitives' boxes in fewer node steps;
    # the closest distance is order-independent, so it equals the binary
    # traversal's bit for bit, and the ids equal brute force up to ties.
    xyz, idx = small_mesh
    b = rumi.Bvh(None, xyz, idx, max_leaf=max_leaf)
    ob = flat_from_export(oracle, b.export(), idx.shape[0])
    rays = rumi.make_rays_uniform(5000, 17, 2.1)
    bp, bt = oracle.brute_mesh(xyz, idx, rays)
    p2, t2, n2, c2 = ob.trace_mesh(xyz, idx, rays, ordered=1, counts=True)
    p4, t4, n4, c4 = ob.trace_mesh(xyz, idx, rays, ordered=2, counts=True)
    np.testing.assert_array_equal(t4.view(np.int32), t2.view(np.int32))
    assert ((p4 == bp) | tie_ok(t4, bt)).all()
    assert (p4 >= 0).sum() > 500
    assert n4.sum() < 0.7 * n2.sum()
    pa, _ = ob.trace_mesh(xyz, idx, rays, any_hit=True, ordered=2)
    np.testing.assert_array_equal(pa >= 0, bp >= 0)
    # the median tree too
    lo, hi = oracle.mesh_bounds(xyz, idx)
    pm, tm = oracle.MedianBvh(lo, hi).trace_mesh(xyz, idx, rays, ordered=2)
    np.testing.assert_array_equal(tm.view(np.int32), t2.view(np.int32))


def test_degenerate_meshes_build(rumi):
    # empty mesh, one triangle, many coincident triangles
    xyz = np.array([[0, 0, 0], [1, 0, 0], [0, 1, 0]], np.float32)
    for nt in (0, 1, 40):
        idx = np.tile(np.array([[0, 1, 2]], np.uint32), (nt, 1)).reshape(nt, 3)
        info = rumi.Bvh(None, xyz, idx).info()
        assert info.n_prims == nt
    with pytest.raises(rumi.RumiError):
        rumi.Bvh(None, xyz, np.array([[0, 1, 3]], np.uint32))   # index out of range
    with pytest.raises(rumi.RumiError):
        rumi.Bvh(None, xyz, np.array([[0, 1, 2]], np.uint32), max_leaf=16)


def test_library_exports_every_declared_symbol(rumi):
    import ctypes
    import re
    import os
    hdr = open(os.path.join(os.path.dirname(__file__), "..", "include", "rumi.h")).read()
    declared = set(re.findall(r"^\s*(?:const\s+)?\w+\*?\s+\**(rumi_\w+)\s*\(", hdr, re.M))
    assert declared, "no declarations parsed"
    handle = ctypes.CDLL(rumi.LIB_PATH)
    for name in sorted(declared):
        assert hasattr(handle, name), name
    assert declared == set(rumi.exported_symbols())
