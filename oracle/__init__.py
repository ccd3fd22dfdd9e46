"""ctypes binding of the oracle (liboracle.so, built from oracle/*.c).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and
bench.py's CPU-baseline leg, as the checker.  The product never imports it.
"""
from __future__ import annotations

import ctypes
import os
import subprocess
from ctypes import POINTER, Structure, byref, c_double, c_float, c_int, c_int64, c_void_p

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liboracle.so")

OR_TRI, OR_SPHERE, OR_RECT = 0, 1, 2


class OrRay(Structure):
    _fields_ = [("o", c_float * 3), ("d", c_float * 3), ("tmin", c_float), ("tmax", c_float)]


class OrPrim(Structure):
    _fields_ = [("kind", c_int), ("a", c_float * 3), ("b", c_float * 3), ("c", c_float * 3),
                ("n", c_float * 3), ("r", c_float), ("length", c_float), ("width", c_float),
                ("inv_area", c_float), ("emissive", c_int), ("le_textured", c_int)]


class OrBvh(Structure):
    _fields_ = [("n_nodes", c_int), ("root", c_int), ("n_prims", c_int),
                ("lo", POINTER(c_float)), ("hi", POINTER(c_float)),
                ("left", POINTER(c_int)), ("right", POINTER(c_int)),
                ("first", POINTER(c_int)), ("count", POINTER(c_int)),
                ("order", POINTER(c_int))]


class OrStats(Structure):
    _fields_ = [("nodes", c_int64), ("prims", c_int64)]


_lib = None


def build() -> None:
    subprocess.run(["make", "-s", "-C", HERE], check=True)


def lib() -> ctypes.CDLL:
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        L = ctypes.CDLL(LIB_PATH)
        vp = c_void_p
        L.or_triangle_intersect.argtypes = [vp, vp, vp, POINTER(OrRay), vp, vp, vp]
        L.or_sphere_intersect.argtypes = [POINTER(OrPrim), POINTER(OrRay), c_int]
        L.or_rect_intersect.argtypes = [POINTER(OrPrim), POINTER(OrRay)]
        L.or_prim_intersect.argtypes = [POINTER(OrPrim), POINTER(OrRay), c_int]
        L.or_bbox_intersect.argtypes = [vp, vp, POINTER(OrRay), vp, vp]
        L.or_prim_bound.argtypes = [POINTER(OrPrim), vp, vp]
        L.or_prim_bound.restype = None
        L.or_bvh_build_median6.argtypes = [vp, vp, c_int, POINTER(OrBvh)]
        L.or_bvh_from_childbox.argtypes = [c_int, vp, vp, vp, vp, c_int, vp, c_int, POINTER(OrBvh)]
        L.or_bvh_free.argtypes = [POINTER(OrBvh)]
        L.or_bvh_free.restype = None
        L.or_travel.argtypes = [POINTER(OrBvh), vp, POINTER(OrRay), c_int, c_int, c_int,
                                POINTER(OrStats)]
        L.or_trace_mesh.argtypes = [POINTER(OrBvh), vp, vp, c_int, vp, c_int, c_int, c_int,
                                    vp, vp, vp, vp]
        L.or_brute_mesh.argtypes = [vp, vp, c_int, vp, c_int, vp, vp]
        L.or_mesh_bounds.argtypes = [vp, vp, c_int, vp, vp]
        L.or_time_median6_mesh.argtypes = [vp, vp, c_int, vp, c_int, c_int, vp, vp,
                                           POINTER(c_double)]
        L.or_time_median6_mesh.restype = c_double
        _lib = L
    return _lib


def _p(a: np.ndarray | None) -> int | None:
    if a is None:
        return None
    assert a.flags["C_CONTIGUOUS"]
    return a.ctypes.data


def make_ray(o, d, tmin=0.0, tmax=3.4028234663852886e38) -> OrRay:
    r = OrRay()
    r.o[:] = [float(x) for x in o]
    r.d[:] = [float(x) for x in d]
    r.tmin = tmin
    r.tmax = tmax
    return r


def triangle_intersect(v0, v1, v2, ray: OrRay):
    """Returns (accepted, u, v, det); ray.tmin is updated like the reference."""
    a = np.asarray(v0, np.float32).copy()
    b = np.asarray(v1, np.float32).copy()
    c = np.asarray(v2, np.float32).copy()
    u, v, det = c_float(), c_float(), c_float()
    ok = lib().or_triangle_intersect(_p(a), _p(b), _p(c), byref(ray), ctypes.addressof(u),
                                     ctypes.addressof(v), ctypes.addressof(det))
    return bool(ok), u.value, v.value, det.value


def bbox_intersect(lo, hi, ray: OrRay):
    a = np.asarray(lo, np.float32).copy()
    b = np.asarray(hi, np.float32).copy()
    tn, tf = c_float(), c_float()
    ok = lib().or_bbox_intersect(_p(a), _p(b), byref(ray), ctypes.addressof(tn),
                                 ctypes.addressof(tf))
    return bool(ok), tn.value, tf.value


def sphere(c, r, emissive=False, le_textured=False) -> OrPrim:
    p = OrPrim()
    p.kind = OR_SPHERE
    p.a[:] = [float(x) for x in c]
    p.r = r
    p.emissive = int(emissive)
    p.le_textured = int(le_textured)
    return p


def rect(p0, l, w, emissive=False) -> OrPrim:
    """Rect(p, l, w) ctor (src/Shape/rect.cpp:6-15): n = normalize(l x w)."""
    f = np.float32
    p0, l, w = (np.asarray(x, f) for x in (p0, l, w))
    cx = f(l[1] * w[2] - w[1] * l[2])
    cy = f(l[2] * w[0] - w[2] * l[0])
    cz = f(l[0] * w[1] - w[0] * l[1])
    inv = f(1.0) / np.sqrt(f(cx * cx + cy * cy + cz * cz), dtype=f)
    n = np.array([cx * inv, cy * inv, cz * inv], f)
    length = np.sqrt(f(l[0] * l[0] + l[1] * l[1] + l[2] * l[2]), dtype=f)
    width = np.sqrt(f(w[0] * w[0] + w[1] * w[1] + w[2] * w[2]), dtype=f)
    q = OrPrim()
    q.kind = OR_RECT
    q.a[:] = [float(x) for x in p0]
    q.b[:] = [float(x) for x in l]
    q.c[:] = [float(x) for x in w]
    q.n[:] = [float(x) for x in n]
    q.length = float(length)
    q.width = float(width)
    q.inv_area = float(f(1.0) / f(length * width))
    q.emissive = int(emissive)
    return q


def prim_intersect(p: OrPrim, ray: OrRay, in_object: bool = False) -> bool:
    return bool(lib().or_prim_intersect(byref(p), byref(ray), int(in_object)))


class MedianBvh:
    """or_bvh_build_median6 over per-primitive boxes (lo, hi: (n, 3) float32)."""

    def __init__(self, lo: np.ndarray, hi: np.ndarray):
        self._lo = np.ascontiguousarray(lo, np.float32)
        self._hi = np.ascontiguousarray(hi, np.float32)
        self.b = OrBvh()
        lib().or_bvh_build_median6(_p(self._lo), _p(self._hi), self._lo.shape[0], byref(self.b))

    @classmethod
    def from_export(cls, ex: dict, n_prims: int) -> "MedianBvh":
        self = cls.__new__(cls)
        self.b = OrBvh()
        boxes = np.ascontiguousarray(ex["child_boxes"], np.float32)
        refs = np.ascontiguousarray(ex["child_refs"], np.int32)
        rb = np.ascontiguousarray(ex["root_box"], np.float32)
        leaf = np.ascontiguousarray(ex["leaf_prims"], np.int32)
        rlo, rhi = rb[:3].copy(), rb[3:].copy()
        lib().or_bvh_from_childbox(boxes.shape[0], _p(boxes), _p(refs), _p(rlo), _p(rhi),
                                   ex["root_ref"], _p(leaf), n_prims, byref(self.b))
        return self

    def arrays(self) -> dict:
        b = self.b
        n = b.n_nodes

        def arr(ptr, count, dt):
            if count == 0:
                return np.zeros(0, dt)
            return np.ctypeslib.as_array(ptr, shape=(count,)).astype(dt).copy()

        return dict(lo=arr(b.lo, 3 * n, np.float32).reshape(-1, 3),
                    hi=arr(b.hi, 3 * n, np.float32).reshape(-1, 3),
                    left=arr(b.left, n, np.int32), right=arr(b.right, n, np.int32),
                    first=arr(b.first, n, np.int32), count=arr(b.count, n, np.int32),
                    order=arr(b.order, b.n_prims, np.int32), root=b.root)

    def trace_mesh(self, xyz, idx, rays, any_hit=False, ordered=False, counts=False):
        xyz = np.ascontiguousarray(xyz, np.float32)
        idx = np.ascontiguousarray(idx, np.uint32)
        rays = np.ascontiguousarray(rays, np.float32)
        n = rays.shape[0]
        prim = np.empty(n, np.int32)
        t = np.empty(n, np.float32)
        nc = np.empty(n, np.int32) if counts else None
        pc = np.empty(n, np.int32) if counts else None
        lib().or_trace_mesh(byref(self.b), _p(xyz), _p(idx), idx.shape[0], _p(rays), n,
                            int(any_hit), int(ordered), _p(prim), _p(t), _p(nc), _p(pc))
        if counts:
            return prim, t, nc, pc
        return prim, t

    def __del__(self):
        try:
            lib().or_bvh_free(byref(self.b))
        except Exception:
            pass


def mesh_bounds(xyz, idx):
    xyz = np.ascontiguousarray(xyz, np.float32)
    idx = np.ascontiguousarray(idx, np.uint32)
    lo = np.empty((idx.shape[0], 3), np.float32)
    hi = np.empty((idx.shape[0], 3), np.float32)
    lib().or_mesh_bounds(_p(xyz), _p(idx), idx.shape[0], _p(lo), _p(hi))
    return lo, hi


def brute_mesh(xyz, idx, rays):
    xyz = np.ascontiguousarray(xyz, np.float32)
    idx = np.ascontiguousarray(idx, np.uint32)
    rays = np.ascontiguousarray(rays, np.float32)
    n = rays.shape[0]
    prim = np.empty(n, np.int32)
    t = np.empty(n, np.float32)
    lib().or_brute_mesh(_p(xyz), _p(idx), idx.shape[0], _p(rays), n, _p(prim), _p(t))
    return prim, t


def time_median6_mesh(xyz, idx, rays, any_hit=False):
    """Reference CPU path: median-6 build, then Travel per ray (1 thread).
    Returns (prim, t, traversal_seconds, build_seconds)."""
    xyz = np.ascontiguousarray(xyz, np.float32)
    idx = np.ascontiguousarray(idx, np.uint32)
    rays = np.ascontiguousarray(rays, np.float32)
    n = rays.shape[0]
    prim = np.empty(n, np.int32)
    t = np.empty(n, np.float32)
    bs = c_double()
    secs = lib().or_time_median6_mesh(_p(xyz), _p(idx), idx.shape[0], _p(rays), n,
                                      int(any_hit), _p(prim), _p(t), byref(bs))
    return prim, t, secs, bs.value
