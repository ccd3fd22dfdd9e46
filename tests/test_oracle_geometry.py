"""Pins the oracle's intersectors with closed-form cases.

Each expected value is derived by hand from the geometry (and IEEE-754 for
the infinite-reciprocal cases), not from the oracle's formula; each case is
chosen so that a plausible slip (a dropped rejection, a wrong epsilon rule,
swapped barycentrics, a sign error) fails it.  Reference rules restated:
Triangle::Intersect src/Shape/triangle.cpp:41-118, BBox::Intersect
src/Kernel/bbox.cpp:30-92, Sphere::Intersect src/Shape/sphere.cpp:51-132,
Rect::Intersect src/Shape/rect.cpp:51-122.
"""
import math

import numpy as np
import pytest

V0, V1, V2 = (0.0, 0.0, 0.0), (1.0, 0.0, 0.0), (0.0, 1.0, 0.0)
FLT_MAX = 3.4028234663852886e38


def tri(oracle, o, d, tmin=0.0, tmax=FLT_MAX, v=(V0, V1, V2)):
    r = oracle.make_ray(o, d, tmin, tmax)
    ok, u, vv, det = oracle.triangle_intersect(*v, r)
    return ok, u, vv, det, r.tmin


def test_triangle_front_hit(oracle):
    ok, u, v, det, t = tri(oracle, (0.25, 0.5, 1.0), (0, 0, -1))
    # u weights v1 (x axis), v weights v2 (y axis); det = +1 seen from +z.
    assert ok and u == 0.25 and v == 0.5 and det == 1.0 and t == 1.0


def test_triangle_back_face_is_two_sided(oracle):
    ok, u, v, det, t = tri(oracle, (0.25, 0.5, -2.0), (0, 0, 1))
    assert ok and det == -1.0 and t == 2.0 and (u, v) == (0.25, 0.5)


@pytest.mark.parametrize("xy,hit", [((0.5, 0.5), True), ((0.0, 0.0), True), ((1.0, 0.0), True),
                                    ((0.0, 1.0), True), ((0.5, 0.5000001), False),
                                    ((-1e-7, 0.5), False), ((0.5, -1e-7), False)])
def test_triangle_edges_inclusive(oracle, xy, hit):
    ok, *_ = tri(oracle, (xy[0], xy[1], 1.0), (0, 0, -1))
    assert ok == hit


def test_triangle_parallel_and_tiny_determinant(oracle):
    assert not tri(oracle, (-1, 0.25, 0), (1, 0, 0))[0]
    # A head-on hit on a triangle with |det| = 1e-10 < 1e-8 is rejected.
    s = 1e-5
    small = ((0, 0, 0), (s, 0, 0), (0, s, 0))
    ok, *_ = tri(oracle, (0.2 * s, 0.2 * s, 1.0), (0, 0, -1), v=small)
    assert not ok


def test_triangle_epsilon_rules(oracle):
    eps = np.float32(1e-4)
    # t < eps rejected, t == eps accepted
    assert not tri(oracle, (0.25, 0.25, 5e-5), (0, 0, -1))[0]
    ok, *_, t = tri(oracle, (0.25, 0.25, float(eps)), (0, 0, -1))
    assert ok and t == eps
    # tmax - t < eps rejected
    assert not tri(oracle, (0.25, 0.25, 1), (0, 0, -1), tmax=1.00005)[0]
    assert tri(oracle, (0.25, 0.25, 1), (0, 0, -1), tmax=1.0002)[0]


def test_triangle_closest_so_far_rule(oracle):
    o, d = (0.25, 0.25, 1), (0, 0, -1)
    assert not tri(oracle, o, d, tmin=0.5)[0]          # farther than the current hit
    ok, *_, t = tri(oracle, o, d, tmin=2.0)            # nearer: accepted, tmin updated
    assert ok and t == 1.0
    assert tri(oracle, o, d, tmin=1.0)[0]               # equal distance: later primitive wins
    # tmin <= eps counts as "no hit yet" (the test is tmin > eps && t > tmin)
    assert tri(oracle, o, d, tmin=5e-5)[0]


def test_bbox_axis_ray_with_zero_components(oracle):
    r = oracle.make_ray((-3, 0, 0), (1, 0, 0))
    assert oracle.bbox_intersect((-1, -1, -1), (1, 1, 1), r) == (True, 2.0, 4.0)


def test_bbox_zero_component_on_slab_plane(oracle):
    # (hi.y - o.y) * inf = 0 * inf = NaN: NaN never compares true, so the y
    # slab neither rejects nor clamps and the x interval survives.
    r = oracle.make_ray((-3, 1, 0), (1, 0, 0))
    assert oracle.bbox_intersect((-1, -1, -1), (1, 1, 1), r) == (True, 2.0, 4.0)


def test_bbox_miss_and_behind(oracle):
    r = oracle.make_ray((-3, 2, 0), (1, 0, 0))
    assert oracle.bbox_intersect((-1, -1, -1), (1, 1, 1), r)[0] is False
    # Box behind the origin still "intersects" with tfar < 0; the BVH
    # accept test (tfar >= 0) is what drops it.
    r = oracle.make_ray((3, 0, 0), (1, 0, 0))
    assert oracle.bbox_intersect((-1, -1, -1), (1, 1, 1), r) == (True, -4.0, -2.0)


def test_bbox_diagonal_inside(oracle):
    k = np.float32(1.0 / math.sqrt(3.0))
    r = oracle.make_ray((0, 0, 0), (k, k, k))
    ok, tn, tf = oracle.bbox_intersect((-1, -1, -1), (1, 1, 1), r)
    inv = np.float32(1.0) / k
    assert ok and tn == -inv and tf == inv


def test_sphere_near_far_and_inside(oracle):
    s = oracle.sphere((0, 0, 0), 1.0)
    r = oracle.make_ray((0, 0, -3), (0, 0, 1))
    assert oracle.prim_intersect(s, r) and r.tmin == 2.0      # B=-6, C=8, disc=4
    r = oracle.make_ray((0, 0, -3), (0, 0, 1))
    assert oracle.prim_intersect(s, r, in_object=True) and r.tmin == 4.0
    # From the centre, not flagged inside: near root -1 < eps, so it misses.
    r = oracle.make_ray((0, 0, 0), (0, 0, 1))
    assert not oracle.prim_intersect(s, r)
    r = oracle.make_ray((0, 0, 0), (0, 0, 1))
    assert oracle.prim_intersect(s, r, in_object=True) and r.tmin == 1.0


def test_sphere_tangent_and_limits(oracle):
    s = oracle.sphere((0, 0, 0), 1.0)
    r = oracle.make_ray((1, 0, -3), (0, 0, 1))      # disc == 0: single root t = 3
    assert oracle.prim_intersect(s, r) and r.tmin == 3.0
    r = oracle.make_ray((0, 0, -3), (0, 0, 1), tmax=2.00005)
    assert not oracle.prim_intersect(s, r)
    r = oracle.make_ray((0, 0, -3), (0, 0, 1), tmin=1.5)
    assert not oracle.prim_intersect(s, r)
    r = oracle.make_ray((0, 0, -3), (0, 0, 1), tmin=2.0)
    assert oracle.prim_intersect(s, r)


def test_rect_front_back_and_extent(oracle):
    q = oracle.rect((0, 0, 0), (1, 0, 0), (0, 1, 0))       # n = +z
    assert tuple(q.n) == (0.0, 0.0, 1.0)
    r = oracle.make_ray((0.5, 0.5, 1), (0, 0, -1))
    assert oracle.prim_intersect(q, r) and r.tmin == 1.0
    # Seen from behind exactly along the normal: cos == 1 is not in (0, 1),
    # so the back-face rule lets it through.
    r = oracle.make_ray((0.5, 0.5, -1), (0, 0, 1))
    assert oracle.prim_intersect(q, r) and r.tmin == 1.0
    # Oblique from behind: culled unless the rect emits.
    d = np.array([0.1, 0.0, 1.0], np.float32)
    d = d / np.sqrt(np.float32(d @ d), dtype=np.float32)
    r = oracle.make_ray((0.5, 0.5, -1), d)
    assert not oracle.prim_intersect(q, r)
    qe = oracle.rect((0, 0, 0), (1, 0, 0), (0, 1, 0), emissive=True)
    r = oracle.make_ray((0.5, 0.5, -1), d)
    assert oracle.prim_intersect(qe, r)
    r = oracle.make_ray((1.5, 0.5, 1), (0, 0, -1))
    assert not oracle.prim_intersect(q, r)
