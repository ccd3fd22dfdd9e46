"""ctypes binding of librumi (include/rumi.h) for the test harness and bench.

The product is the C ABI in librumi.so; this module only declares its
signatures and wraps handles.  Device buffers are passed as raw pointers
(``tensor.data_ptr()`` of a torch tensor on the GPU); torch supplies device
memory and streams, nothing else.  There is no fallback: if librumi.so is
missing or a call fails, an exception is raised.
"""
from __future__ import annotations

import ctypes
import os
from ctypes import (POINTER, Structure, byref, c_char_p, c_double, c_float, c_int, c_int32,
                    c_uint32, c_uint64, c_void_p)

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "librumi.so")

RUMI_OK = 0
RUMI_BVH_SAH_BINNED = 0
RUMI_BVH_REF_MEDIAN6 = 1
RUMI_TRACE_DEFAULT = 0
RUMI_TRACE_REF_ORDER = 1
RUMI_TRACE_ORDERED = 2


class RumiError(RuntimeError):
    def __init__(self, code: int, msg: str):
        super().__init__(f"rumi error {code}: {msg}")
        self.code = code


class BvhParams(Structure):
    _fields_ = [("mode", c_int32), ("max_leaf", c_int32), ("n_bins", c_int32),
                ("reserved", c_int32)]


class BvhInfo(Structure):
    _fields_ = [("mode", c_int32), ("n_prims", c_int32), ("n_inner", c_int32),
                ("n_leaves", c_int32), ("max_depth", c_int32), ("on_device", c_int32),
                ("sah_cost", c_double), ("build_seconds", c_double)]


# Every function include/rumi.h declares, with its ctypes signature.
_SIGNATURES = {
    "rumi_last_error": (c_char_p, []),
    "rumi_version": (c_char_p, []),
    "rumi_ctx_create": (c_int, [c_int, POINTER(c_void_p)]),
    "rumi_ctx_destroy": (c_int, [c_void_p]),
    "rumi_bvh_build_mesh": (c_int, [c_void_p, c_void_p, c_uint32, c_void_p, c_uint32,
                                    POINTER(BvhParams), POINTER(c_void_p)]),
    "rumi_bvh_destroy": (c_int, [c_void_p]),
    "rumi_bvh_get_info": (c_int, [c_void_p, POINTER(BvhInfo)]),
    "rumi_bvh_export": (c_int, [c_void_p, c_void_p, c_void_p, c_void_p, c_void_p, c_void_p]),
    "rumi_trace_closest": (c_int, [c_void_p, c_void_p, c_void_p, c_uint32, c_void_p, c_void_p,
                                   c_uint32, c_void_p]),
    "rumi_trace_any": (c_int, [c_void_p, c_void_p, c_void_p, c_uint32, c_void_p, c_uint32,
                               c_void_p]),
    "rumi_trace_count": (c_int, [c_void_p, c_void_p, c_void_p, c_uint32, c_int32, c_void_p,
                                 c_void_p, c_void_p, c_uint32, c_void_p]),
    "rumi_make_icosphere": (c_int, [c_uint32, c_uint64, c_float, c_void_p, c_void_p]),
    "rumi_make_rays_uniform": (c_int, [c_uint64, c_uint64, c_float, c_void_p]),
    "rumi_pcg32_sequence": (c_int, [c_uint64, c_uint64, c_uint32, c_void_p]),
}

_lib = None


def lib() -> ctypes.CDLL:
    """Loads librumi.so (built by __graft_entry__.build()); raises if absent."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"{LIB_PATH} is missing: run __graft_entry__.build() first")
        handle = ctypes.CDLL(LIB_PATH)
        for name, (res, args) in _SIGNATURES.items():
            fn = getattr(handle, name)
            fn.restype = res
            fn.argtypes = args
        _lib = handle
    return _lib


def exported_symbols() -> list[str]:
    return list(_SIGNATURES)


def check(rc: int) -> None:
    if rc != RUMI_OK:
        msg = lib().rumi_last_error()
        raise RumiError(rc, msg.decode() if msg else "")


def _ptr(a) -> int:
    """Raw address of a numpy array or torch tensor (None -> 0)."""
    if a is None:
        return 0
    if isinstance(a, np.ndarray):
        if not a.flags["C_CONTIGUOUS"]:
            raise ValueError("array must be C-contiguous")
        return a.ctypes.data
    if hasattr(a, "data_ptr"):
        if not a.is_contiguous():
            raise ValueError("tensor must be contiguous")
        return a.data_ptr()
    return int(a)


def _stream(stream) -> int:
    if stream is None:
        import torch
        return torch.cuda.current_stream().cuda_stream
    return int(getattr(stream, "cuda_stream", stream))


# --------------------------------------------------------------- workloads
def make_icosphere(freq: int, seed: int, amplitude: float):
    nv, nt = 10 * freq * freq + 2, 20 * freq * freq
    xyz = np.empty((nv, 3), np.float32)
    idx = np.empty((nt, 3), np.uint32)
    check(lib().rumi_make_icosphere(freq, seed, amplitude, _ptr(xyz), _ptr(idx)))
    return xyz, idx


def make_rays_uniform(n: int, seed: int, half_extent: float) -> np.ndarray:
    rays = np.empty((n, 8), np.float32)
    check(lib().rumi_make_rays_uniform(n, seed, half_extent, _ptr(rays)))
    return rays


def pcg32_sequence(initstate: int, initseq: int, n: int) -> np.ndarray:
    out = np.empty(n, np.uint32)
    check(lib().rumi_pcg32_sequence(initstate, initseq, n, _ptr(out)))
    return out


# --------------------------------------------------------------- handles
class Context:
    def __init__(self, device: int = 0):
        h = c_void_p()
        check(lib().rumi_ctx_create(device, byref(h)))
        self.handle = h

    def close(self):
        if self.handle:
            lib().rumi_ctx_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Bvh:
    """A BVH over an indexed triangle mesh (host build, optional upload)."""

    def __init__(self, ctx: Context | None, xyz: np.ndarray, idx: np.ndarray,
                 mode: int = RUMI_BVH_SAH_BINNED, max_leaf: int = 4, n_bins: int = 32):
        xyz = np.ascontiguousarray(xyz, np.float32)
        idx = np.ascontiguousarray(idx, np.uint32)
        p = BvhParams(mode, max_leaf, n_bins, 0)
        h = c_void_p()
        check(lib().rumi_bvh_build_mesh(ctx.handle if ctx else None, _ptr(xyz), xyz.shape[0],
                                        _ptr(idx), idx.shape[0], byref(p), byref(h)))
        self.handle = h
        self.ctx = ctx

    def info(self) -> BvhInfo:
        i = BvhInfo()
        check(lib().rumi_bvh_get_info(self.handle, byref(i)))
        return i

    def export(self):
        i = self.info()
        boxes = np.empty((max(i.n_inner, 1), 12), np.float32)
        refs = np.empty((max(i.n_inner, 1), 2), np.int32)
        root_box = np.empty(6, np.float32)
        root_ref = np.empty(1, np.int32)
        leaf = np.empty(max(i.n_prims, 1), np.int32)
        check(lib().rumi_bvh_export(self.handle, _ptr(boxes), _ptr(refs), _ptr(root_box),
                                    _ptr(root_ref), _ptr(leaf)))
        return dict(child_boxes=boxes[: i.n_inner], child_refs=refs[: i.n_inner],
                    root_box=root_box, root_ref=int(root_ref[0]), leaf_prims=leaf[: i.n_prims])

    def close(self):
        if self.handle:
            lib().rumi_bvh_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


# --------------------------------------------------------------- traversal
def trace_closest(ctx: Context, bvh: Bvh, rays, hits, uv=None, flags: int = 0, stream=None):
    """rays: device (n, 8) float32; hits: device (n, 2) int32 (prim, t bits)."""
    n = rays.shape[0]
    check(lib().rumi_trace_closest(ctx.handle, bvh.handle, _ptr(rays), n, _ptr(hits), _ptr(uv),
                                   flags, _stream(stream)))


def trace_any(ctx: Context, bvh: Bvh, rays, prim, flags: int = 0, stream=None):
    n = rays.shape[0]
    check(lib().rumi_trace_any(ctx.handle, bvh.handle, _ptr(rays), n, _ptr(prim), flags,
                               _stream(stream)))


def trace_count(ctx: Context, bvh: Bvh, rays, any_hit: bool, hits, node_cnt, prim_cnt,
                flags: int = 0, stream=None):
    n = rays.shape[0]
    check(lib().rumi_trace_count(ctx.handle, bvh.handle, _ptr(rays), n, 1 if any_hit else 0,
                                 _ptr(hits), _ptr(node_cnt), _ptr(prim_cnt), flags,
                                 _stream(stream)))


def split_hits(hits_i32: np.ndarray):
    """(n, 2) int32 hit records -> (prim int32, t float32)."""
    h = np.ascontiguousarray(hits_i32, np.int32)
    return h[:, 0].copy(), h[:, 1].copy().view(np.float32)
