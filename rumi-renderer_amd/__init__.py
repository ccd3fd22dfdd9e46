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


class Material(Structure):
    _fields_ = [("kind", c_int32), ("kd", c_float), ("ks", c_float), ("exponent", c_float),
                ("color", c_float * 3), ("eta_in", c_float), ("eta_out", c_float),
                ("absorption", c_float * 3), ("sigma_s", c_float * 3), ("sigma_a", c_float * 3),
                ("density", c_float), ("roughness", c_float), ("reserved", c_int32 * 3)]


class Shape(Structure):
    _fields_ = [("kind", c_int32), ("material", c_int32), ("is_light", c_int32),
                ("texture", c_int32), ("v", c_float * 9), ("n", c_float * 9), ("radius", c_float),
                ("xyz", c_void_p), ("nv", c_uint32), ("idx", c_void_p), ("nt", c_uint32),
                ("normals", c_void_p), ("uv", c_void_p)]


class Camera(Structure):
    _fields_ = [("position", c_float * 3), ("basis", c_float * 9), ("fov_deg", c_float),
                ("y_ratio", c_float), ("focal_distance", c_float), ("lens_radius", c_float)]


class Medium(Structure):
    _fields_ = [("bound", c_int32), ("lo", c_float * 3), ("hi", c_float * 3),
                ("center", c_float * 3), ("radius", c_float), ("sigma_s", c_float * 3),
                ("sigma_a", c_float * 3), ("g", c_float)]


class Texture(Structure):
    _fields_ = [("width", c_int32), ("height", c_int32), ("rgb", c_void_p), ("kind", c_int32),
                ("checker_w", c_float), ("checker_h", c_float), ("color_1", c_float * 3),
                ("color_2", c_float * 3)]


RUMI_TEXTURE_IMAGE, RUMI_TEXTURE_CHECKER, RUMI_TEXTURE_HDR = 0, 1, 2


class PointLight(Structure):
    _fields_ = [("position", c_float * 3), ("ls", c_float), ("color", c_float * 3)]


class SceneDesc(Structure):
    _fields_ = [("shapes", c_void_p), ("n_shapes", c_uint32), ("materials", c_void_p),
                ("n_materials", c_uint32), ("textures", c_void_p), ("n_textures", c_uint32),
                ("camera", Camera), ("medium", Medium), ("point_lights", c_void_p),
                ("n_point_lights", c_uint32)]


class RenderDesc(Structure):
    _fields_ = [("integrator", c_int32), ("width", c_int32), ("height", c_int32),
                ("spp", c_int32), ("max_depth", c_int32), ("n_light", c_int32),
                ("n_eye", c_int32), ("mis_beta", c_float), ("seed", c_uint64),
                ("queue_capacity", c_uint32), ("regen_threshold", c_uint32), ("flags", c_uint32)]


class RenderStats(Structure):
    _fields_ = [("samples", c_uint64), ("extension_rays", c_uint64), ("shadow_rays", c_uint64),
                ("iterations", c_uint64), ("milliseconds", c_double), ("ext_nodes", c_uint64),
                ("ext_prims", c_uint64), ("shadow_nodes", c_uint64), ("shadow_prims", c_uint64)]


class TraceTuning(Structure):
    _fields_ = [("block", c_int32), ("stack", c_int32), ("refill", c_int32),
                ("persistent", c_int32), ("width", c_int32)]


class AutotuneReport(Structure):
    _fields_ = [("n_candidates", c_uint32), ("best_us", c_float), ("host_candidate_us", c_void_p),
                ("candidates_capacity", c_uint32), ("repeats", c_int32)]


class BatchItem(Structure):
    _fields_ = [("scene", c_void_p), ("width", c_int32), ("height", c_int32)]


class BatchDesc(Structure):
    _fields_ = [("spp", c_int32), ("max_depth", c_int32), ("seed", c_uint64),
                ("queue_capacity", c_uint32), ("adjoint", c_int32)]


class BatchIO(Structure):
    _fields_ = [("dev_images", c_void_p), ("dev_targets", c_void_p),
                ("dev_texel_grads", c_void_p), ("dev_loss", c_void_p)]


class BatchStats(Structure):
    _fields_ = [("total_samples", c_uint64), ("total_pixels", c_uint64),
                ("total_texels", c_uint64), ("n_items", c_uint32), ("generations", c_uint32),
                ("milliseconds", c_double), ("host_item_ms", c_void_p)]


class GradBuffers(Structure):
    _fields_ = [("d_medium", c_void_p), ("d_textures", c_void_p), ("d_image", c_void_p)]


RUMI_SHAPE_TRIANGLE, RUMI_SHAPE_SPHERE, RUMI_SHAPE_RECT, RUMI_SHAPE_MESH = 0, 1, 2, 3
MATERIAL_KINDS = {"matte": 0, "glossy": 1, "specular": 2, "emissive": 3, "dielectric": 4,
                  "subsurface": 5, "ggx": 6}
RUMI_INTEGRATOR_PT, RUMI_INTEGRATOR_BDPT, RUMI_INTEGRATOR_WHITTED = 0, 1, 2
RUMI_RENDER_COUNT_TRAVERSAL = 1
RUMI_FILM_BGR_SUM, RUMI_FILM_RGB_MEAN = 0, 1

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
    "rumi_trace_set_tuning": (c_int, [c_void_p, POINTER(TraceTuning)]),
    "rumi_trace_get_tuning": (c_int, [c_void_p, POINTER(TraceTuning)]),
    "rumi_autotune_trace_candidates": (c_int, [POINTER(TraceTuning), c_uint32, POINTER(c_uint32)]),
    "rumi_autotune_trace": (c_int, [c_void_p, c_void_p, c_void_p, c_uint32, c_int32, c_uint32, c_void_p,
                                    POINTER(TraceTuning), POINTER(AutotuneReport), c_void_p]),
    "rumi_autotune_render": (c_int, [c_void_p, c_void_p, POINTER(RenderDesc), c_void_p,
                                     POINTER(AutotuneReport), c_void_p]),
    "rumi_scene_create": (c_int, [c_void_p, POINTER(SceneDesc), POINTER(c_void_p)]),
    "rumi_scene_destroy": (c_int, [c_void_p]),
    "rumi_render": (c_int, [c_void_p, c_void_p, POINTER(RenderDesc), c_void_p,
                            POINTER(RenderStats), c_void_p]),
    "rumi_render_noalloc": (c_int, [c_void_p, c_void_p, POINTER(RenderDesc), c_void_p,
                                    POINTER(RenderStats), c_void_p]),
    "rumi_adjoint": (c_int, [c_void_p, c_void_p, POINTER(RenderDesc), c_void_p,
                             POINTER(GradBuffers), POINTER(RenderStats), c_void_p]),
    "rumi_batch_create": (c_int, [c_void_p, POINTER(BatchItem), c_uint32, POINTER(BatchDesc),
                                  POINTER(c_void_p)]),
    "rumi_batch_run": (c_int, [c_void_p, POINTER(BatchIO), c_int32, POINTER(BatchStats),
                               c_void_p]),
    "rumi_batch_destroy": (c_int, [c_void_p]),
    "rumi_make_icosphere": (c_int, [c_uint32, c_uint64, c_float, c_void_p, c_void_p]),
    "rumi_make_rays_uniform": (c_int, [c_uint64, c_uint64, c_float, c_void_p]),
    "rumi_pcg32_sequence": (c_int, [c_uint64, c_uint64, c_uint32, c_void_p]),
    "rumi_obj_load": (c_int, [c_char_p, c_void_p, POINTER(c_void_p)]),
    "rumi_obj_get": (c_int, [c_void_p, POINTER(c_uint32), POINTER(c_void_p), POINTER(c_void_p),
                             c_void_p]),
    "rumi_obj_destroy": (c_int, [c_void_p]),
    "rumi_film_encode": (c_int, [c_void_p, c_void_p, c_uint32, c_uint32, c_int32, c_uint32, c_float,
                                 c_uint32, c_void_p, c_void_p]),
    "rumi_film_fireflies": (c_int, [c_void_p, c_void_p, c_uint32, c_uint32, c_uint32, c_void_p,
                                    c_void_p]),
    "rumi_bmp_header": (c_int, [c_uint32, c_uint32, c_uint32, c_void_p]),
    "rumi_bmp_texels": (c_int, [c_void_p, c_uint32, c_uint32, c_uint32, c_float, c_void_p]),
    "rumi_hdr_read": (c_int, [c_void_p, c_uint64, POINTER(c_uint32), POINTER(c_uint32), c_void_p,
                              c_uint64]),
    "rumi_bmp_read": (c_int, [c_void_p, c_uint64, POINTER(c_uint32), POINTER(c_uint32),
                              POINTER(c_uint32), c_void_p, c_uint64]),
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
def film_encode(ctx, film, width, height, layout=RUMI_FILM_RGB_MEAN, n_samples=1, gamma=2.2,
                bytes_per_pixel=4, out=None, stream=None):
    """rumi_film_encode on a device tensor (height x width x 3 float32);
    returns the uint8 device tensor of pixel bytes (height x width x bpp)."""
    import torch
    if out is None:
        out = torch.empty((height, width, bytes_per_pixel), dtype=torch.uint8, device=film.device)
    check(lib().rumi_film_encode(ctx.handle, _ptr(film), width, height, layout, n_samples, gamma,
                                 bytes_per_pixel, _ptr(out), _stream(stream)))
    return out


def film_fireflies(ctx, pixels, width, height, bytes_per_pixel=4, out=None, stream=None):
    """rumi_film_fireflies: device uint8 pixels in, a new filtered tensor out."""
    import torch
    if out is None:
        out = torch.empty_like(pixels)
    check(lib().rumi_film_fireflies(ctx.handle, _ptr(pixels), width, height, bytes_per_pixel,
                                    _ptr(out), _stream(stream)))
    return out


def bmp_bytes(pixels: np.ndarray, width: int, height: int, bytes_per_pixel: int) -> bytes:
    """SaveBmp's file: rumi_bmp_header + the pixel bytes as they are."""
    hdr = np.empty(54, np.uint8)
    check(lib().rumi_bmp_header(width, height, bytes_per_pixel, _ptr(hdr)))
    return hdr.tobytes() + np.ascontiguousarray(pixels, np.uint8).tobytes()


def bmp_read(data: bytes):
    """LoadBmp's reading rule: (pixels uint8 (h, w, bpp), width, height, bpp)."""
    buf = np.frombuffer(data, np.uint8)
    w, h, b = c_uint32(0), c_uint32(0), c_uint32(0)
    check(lib().rumi_bmp_read(_ptr(buf), buf.size, byref(w), byref(h), byref(b), None, 0))
    pix = np.empty((h.value, w.value, b.value), np.uint8)
    check(lib().rumi_bmp_read(_ptr(buf), buf.size, byref(w), byref(h), byref(b), _ptr(pix), pix.size))
    return pix, w.value, h.value, b.value


def bmp_texture(data: bytes, gamma: float = 1.0) -> np.ndarray:
    """ImageTexture(file, gamma): a BMP file's bytes -> the (h, w, 3)
    float32 texel array of a scene texture."""
    pix, w, h, b = bmp_read(data)
    rgb = np.empty((h, w, 3), np.float32)
    check(lib().rumi_bmp_texels(_ptr(pix), w, h, b, gamma, _ptr(rgb)))
    return rgb


def hdr_read(data: bytes) -> np.ndarray:
    """LoadHDR: a Radiance RGBE file's bytes -> (h, w, 3) float32, rows
    top-down (the data of a dict(kind="hdr", rgb=...) scene texture)."""
    buf = np.frombuffer(data, np.uint8)
    w, h = c_uint32(0), c_uint32(0)
    check(lib().rumi_hdr_read(_ptr(buf), buf.size, byref(w), byref(h), None, 0))
    rgb = np.empty((h.value, w.value, 3), np.float32)
    check(lib().rumi_hdr_read(_ptr(buf), buf.size, byref(w), byref(h), _ptr(rgb), rgb.size))
    return rgb


def load_obj(path, transform=None) -> dict:
    """rumi_obj_load: an OBJ model as Model(file, transform) loads it.
    Returns dict(xyz=(3n, 3) float32 vertices, three per triangle,
    normals=(3n, 3) float32 vertex normals, idx=(n, 3) uint32 = 0, 1, 2, ...,
    counts=(v, vt, vn, f) records): the fields of a "mesh" scene shape."""
    import os
    xf = None if transform is None else np.ascontiguousarray(transform, np.float32).reshape(16)
    h = c_void_p()
    check(lib().rumi_obj_load(os.fsencode(str(path)), _ptr(xf), byref(h)))
    try:
        nt = c_uint32(0)
        px, pn = c_void_p(), c_void_p()
        counts = np.zeros(4, np.uint32)
        check(lib().rumi_obj_get(h, byref(nt), byref(px), byref(pn), _ptr(counts)))
        n = nt.value
        fp = POINTER(c_float)
        xyz = np.ctypeslib.as_array(ctypes.cast(px, fp), (max(n, 1) * 9,))[:9 * n].copy()
        nrm = np.ctypeslib.as_array(ctypes.cast(pn, fp), (max(n, 1) * 9,))[:9 * n].copy()
    finally:
        lib().rumi_obj_destroy(h)
    return dict(xyz=xyz.reshape(-1, 3), normals=nrm.reshape(-1, 3),
                idx=np.arange(3 * n, dtype=np.uint32).reshape(n, 3), counts=counts)


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

    def set_trace_tuning(self, block=0, stack=0, refill=0, persistent=0, width=0):
        """Launch shape of trace_closest / trace_any (zero = default)."""
        check(lib().rumi_trace_set_tuning(self.handle, byref(TraceTuning(block, stack, refill,
                                                                         persistent, width))))

    def trace_tuning(self) -> dict:
        t = TraceTuning()
        check(lib().rumi_trace_get_tuning(self.handle, byref(t)))
        return dict(block=t.block, stack=t.stack, refill=t.refill, persistent=t.persistent,
                    width=t.width)

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


_TUNE_KEYS = ("block", "stack", "refill", "persistent", "width")


def autotune_trace_candidates() -> list[dict]:
    """rumi_autotune_trace_candidates: the traversal tunings the native
    autotuner times, in its order."""
    n = c_uint32()
    check(lib().rumi_autotune_trace_candidates(None, 0, byref(n)))
    arr = (TraceTuning * n.value)()
    check(lib().rumi_autotune_trace_candidates(arr, n.value, byref(n)))
    return [{k: int(getattr(t, k)) for k in _TUNE_KEYS} for t in arr]


def autotune_trace(ctx, bvh, rays, out, any_hit: bool = False, flags: int = 0, repeats: int = 0,
                   stream=None):
    """rumi_autotune_trace: times every candidate on this batch and leaves
    the fastest on ctx.  Returns (best tuning, [(tuning, seconds), ...])."""
    cands = autotune_trace_candidates()
    us = np.zeros(len(cands), np.float32)
    rep = AutotuneReport(0, 0.0, us.ctypes.data, len(cands), repeats)
    best = TraceTuning()
    check(lib().rumi_autotune_trace(ctx.handle, bvh.handle, _ptr(rays), int(rays.shape[0]), int(any_hit),
                                    flags, _ptr(out), byref(best), byref(rep), _stream(stream)))
    return ({k: int(getattr(best, k)) for k in _TUNE_KEYS},
            [(c, float(t) * 1e-6) for c, t in zip(cands, us[:rep.n_candidates])])


def split_hits(hits_i32: np.ndarray):
    """(n, 2) int32 hit records -> (prim int32, t float32)."""
    h = np.ascontiguousarray(hits_i32, np.int32)
    return h[:, 0].copy(), h[:, 1].copy().view(np.float32)


# --------------------------------------------------------------- scenes
def _f3(x):
    return (c_float * 3)(*[float(v) for v in x])


class Scene:
    """A device scene built from the dict form in scenes.py (shapes,
    materials, textures, camera, medium); see include/rumi.h for fields."""

    def __init__(self, ctx: Context, desc: dict):
        self.ctx = ctx
        self._keep = []
        mats = (Material * len(desc["materials"]))()
        for i, m in enumerate(desc["materials"]):
            mm = mats[i]
            mm.kind = MATERIAL_KINDS[m["kind"]]
            mm.kd = m.get("kd", 0.0)
            mm.ks = m.get("ks", 0.0)
            mm.exponent = m.get("exponent", 0.0)
            mm.color = _f3(m.get("color", (1, 1, 1)))
            mm.eta_in = m.get("eta_in", 1.0)
            mm.eta_out = m.get("eta_out", 1.0)
            mm.absorption = _f3(m.get("absorption", (0, 0, 0)))
            mm.sigma_s = _f3(m.get("sigma_s", (0, 0, 0)))
            mm.sigma_a = _f3(m.get("sigma_a", (0, 0, 0)))
            mm.density = m.get("density", 1.0)
            mm.roughness = m.get("roughness", 0.0)
        shapes = (Shape * len(desc["shapes"]))()
        for i, sd in enumerate(desc["shapes"]):
            sh = shapes[i]
            sh.material = sd["material"]
            sh.is_light = 2 if sd.get("light") == "ibl" else (1 if sd.get("light") else 0)
            sh.texture = sd.get("texture", -1)
            k = sd["kind"]
            if k == "tri":
                sh.kind = RUMI_SHAPE_TRIANGLE
                sh.v = (c_float * 9)(*np.asarray(sd["v"], np.float32).reshape(9).tolist())
                if sd.get("n") is not None:
                    sh.n = (c_float * 9)(*np.asarray(sd["n"], np.float32).reshape(9).tolist())
            elif k == "sphere":
                sh.kind = RUMI_SHAPE_SPHERE
                sh.v = (c_float * 9)(*(list(map(float, sd["center"])) + [0.0] * 6))
                sh.radius = sd["radius"]
                if sd.get("basis") is not None:  # rows of the shape's transform (texture mapping)
                    sh.n = (c_float * 9)(*np.asarray(sd["basis"], np.float32).reshape(9).tolist())
            elif k == "rect":
                sh.kind = RUMI_SHAPE_RECT
                sh.v = (c_float * 9)(*[float(x) for x in (*sd["p"], *sd["l"], *sd["w"])])
            elif k == "mesh":
                sh.kind = RUMI_SHAPE_MESH
                xyz = np.ascontiguousarray(sd["xyz"], np.float32)
                idx = np.ascontiguousarray(sd["idx"], np.uint32)
                self._keep += [xyz, idx]
                sh.xyz, sh.nv, sh.idx, sh.nt = _ptr(xyz), xyz.shape[0], _ptr(idx), idx.shape[0]
                if sd.get("normals") is not None:
                    nrm = np.ascontiguousarray(sd["normals"], np.float32)
                    self._keep.append(nrm)
                    sh.normals = _ptr(nrm)
                if sd.get("uv") is not None:
                    uv = np.ascontiguousarray(sd["uv"], np.float32)
                    self._keep.append(uv)
                    sh.uv = _ptr(uv)
            else:
                raise ValueError(f"unknown shape kind {k}")
        texs = (Texture * max(1, len(desc.get("textures", []))))()
        self.n_texels = 0
        for i, t in enumerate(desc.get("textures", [])):
            if isinstance(t, dict) and t.get("kind") == "hdr":  # dict(kind="hdr", rgb=rows top-down)
                rgb = np.ascontiguousarray(t["rgb"], np.float32)
                self.n_texels += rgb.shape[0] * rgb.shape[1]
                self._keep.append(rgb)
                texs[i].kind = RUMI_TEXTURE_HDR
                texs[i].height, texs[i].width = rgb.shape[0], rgb.shape[1]
                texs[i].rgb = _ptr(rgb)
                continue
            if isinstance(t, dict):  # dict(kind="checker", w=, h=, c1=(r, g, b), c2=(r, g, b))
                texs[i].kind = RUMI_TEXTURE_CHECKER
                texs[i].checker_w, texs[i].checker_h = t["w"], t["h"]
                texs[i].color_1 = (c_float * 3)(*t["c1"])
                texs[i].color_2 = (c_float * 3)(*t["c2"])
                continue
            rgb = np.ascontiguousarray(t, np.float32)
            self.n_texels += rgb.shape[0] * rgb.shape[1]
            self._keep.append(rgb)
            texs[i].height, texs[i].width = rgb.shape[0], rgb.shape[1]
            texs[i].rgb = _ptr(rgb)
        d = SceneDesc()
        d.shapes, d.n_shapes = ctypes.addressof(shapes), len(desc["shapes"])
        d.materials, d.n_materials = ctypes.addressof(mats), len(desc["materials"])
        d.textures, d.n_textures = ctypes.addressof(texs), len(desc.get("textures", []))
        cam = desc["camera"]
        d.camera.position = _f3(cam["position"])
        d.camera.basis = (c_float * 9)(*np.asarray(cam["basis"], np.float32).reshape(9).tolist())
        d.camera.fov_deg = cam["fov_deg"]
        d.camera.y_ratio = cam.get("y_ratio", 1.0)
        d.camera.focal_distance = cam.get("focal_distance", 0.0)
        d.camera.lens_radius = cam.get("lens_radius", 0.0)
        pls = desc.get("point_lights") or []
        pl = (PointLight * max(1, len(pls)))()
        for i, q in enumerate(pls):
            pl[i].position = _f3(q["position"])
            pl[i].ls = q.get("ls", 1.0)
            pl[i].color = _f3(q.get("color", (1.0, 1.0, 1.0)))
        d.point_lights, d.n_point_lights = ctypes.addressof(pl), len(pls)
        med = desc.get("medium")
        if med:
            d.medium.bound = {"box": 1, "sphere": 2}[med["bound"]]
            d.medium.lo = _f3(med.get("lo", (0, 0, 0)))
            d.medium.hi = _f3(med.get("hi", (0, 0, 0)))
            d.medium.center = _f3(med.get("center", (0, 0, 0)))
            d.medium.radius = med.get("radius", 0.0)
            d.medium.sigma_s = _f3(med["sigma_s"])
            d.medium.sigma_a = _f3(med["sigma_a"])
            d.medium.g = med.get("g", 0.0)
        h = c_void_p()
        check(lib().rumi_scene_create(ctx.handle, byref(d), byref(h)))
        self.handle = h
        self._keep = []

    def render(self, width, height, spp, max_depth, seed=0, out=None, queue_capacity=0,
               regen_threshold=0, integrator=RUMI_INTEGRATOR_PT, n_light=0, n_eye=0,
               mis_beta=1.0, count_traversal=False, stream=None):
        """Returns (device tensor height x width x 3 of mean radiance, RenderStats);
        count_traversal fills the stats' node and primitive totals."""
        import torch
        if out is None:
            out = torch.empty((height, width, 3), dtype=torch.float32, device="cuda")
        rd = RenderDesc(integrator, width, height, spp, max_depth, n_light, n_eye, mis_beta,
                        seed, queue_capacity, regen_threshold,
                        RUMI_RENDER_COUNT_TRAVERSAL if count_traversal else 0)
        st = RenderStats()
        check(lib().rumi_render(self.ctx.handle, self.handle, byref(rd), _ptr(out), byref(st),
                                _stream(stream)))
        return out, st

    def autotune(self, width, height, spp, max_depth, seed=0, integrator=RUMI_INTEGRATOR_PT,
                 n_light=0, n_eye=0, mis_beta=1.0, repeats=0, out=None, stream=None):
        """rumi_autotune_render: times the frame under the candidate queue
        sizes.  Returns ({queue_capacity, regen_threshold}, [seconds per
        candidate, None where skipped])."""
        import torch
        if out is None:
            out = torch.empty((height, width, 3), dtype=torch.float32, device="cuda")
        rd = RenderDesc(integrator, width, height, spp, max_depth, n_light, n_eye, mis_beta, seed, 0, 0)
        us = np.zeros(16, np.float32)
        rep = AutotuneReport(0, 0.0, us.ctypes.data, 16, repeats)
        check(lib().rumi_autotune_render(self.ctx.handle, self.handle, byref(rd), _ptr(out), byref(rep),
                                         _stream(stream)))
        return (dict(queue_capacity=int(rd.queue_capacity), regen_threshold=int(rd.regen_threshold)),
                [None if t < 0 else float(t) * 1e-6 for t in us[:rep.n_candidates]])

    def render_noalloc(self, width, height, spp, max_depth, out, seed=0, queue_capacity=0,
                       regen_threshold=0, integrator=RUMI_INTEGRATOR_PT, stream=None):
        """rumi_render_noalloc into the caller's device tensor out
        (height x width x 3); returns RenderStats."""
        rd = RenderDesc(integrator, width, height, spp, max_depth, 0, 0, 1.0, seed, queue_capacity,
                        regen_threshold)
        st = RenderStats()
        check(lib().rumi_render_noalloc(self.ctx.handle, self.handle, byref(rd), _ptr(out), byref(st),
                                        _stream(stream)))
        return st

    def adjoint(self, dloss_dimage, spp, max_depth, seed=0, textures=False, queue_capacity=0,
                regen_threshold=0, stream=None):
        """Path-replay adjoint of render(): dloss_dimage is a device tensor
        height x width x 3.  Returns (primal image, 7 medium gradients
        [sigma_a rgb, sigma_s rgb, g], texel gradients (n_texels x 3) or None,
        RenderStats), all device tensors."""
        import torch
        h, w = int(dloss_dimage.shape[0]), int(dloss_dimage.shape[1])
        if dloss_dimage.dtype != torch.float32 or not dloss_dimage.is_contiguous():
            raise ValueError("dloss_dimage must be a contiguous float32 tensor")
        img = torch.empty((h, w, 3), dtype=torch.float32, device=dloss_dimage.device)
        dmed = torch.empty(7, dtype=torch.float32, device=dloss_dimage.device)
        dtex = (torch.empty((max(1, self.n_texels), 3), dtype=torch.float32,
                            device=dloss_dimage.device) if textures else None)
        gb = GradBuffers(_ptr(dmed), _ptr(dtex) if textures else None, _ptr(img))
        rd = RenderDesc(RUMI_INTEGRATOR_PT, w, h, spp, max_depth, 0, 0, 1.0, seed,
                        queue_capacity, regen_threshold)
        st = RenderStats()
        check(lib().rumi_adjoint(self.ctx.handle, self.handle, byref(rd), _ptr(dloss_dimage),
                                 byref(gb), byref(st), _stream(stream)))
        return img, dmed, dtex, st

    def close(self):
        if getattr(self, "handle", None):
            lib().rumi_scene_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Batch:
    """Many scenes rendered as one wavefront (rumi_batch_*).  items: list of
    (Scene, width, height)."""

    def __init__(self, ctx: Context, items, spp: int, max_depth: int, seed: int = 0,
                 queue_capacity: int = 0, adjoint: bool = False):
        self.ctx = ctx
        self.items = list(items)
        arr = (BatchItem * len(self.items))()
        for i, (scene, w, h) in enumerate(self.items):
            arr[i].scene, arr[i].width, arr[i].height = scene.handle, w, h
        d = BatchDesc(spp, max_depth, seed, queue_capacity, 1 if adjoint else 0)
        h = c_void_p()
        check(lib().rumi_batch_create(ctx.handle, arr, len(self.items), byref(d), byref(h)))
        self.handle = h
        self.adjoint = adjoint
        self.sizes = [(w, h) for _, w, h in self.items]
        self.total_pixels = sum(w * h for w, h in self.sizes)
        self.n_texels = sum(scene.n_texels for scene, _, _ in self.items)

    def split(self, flat):
        """Per-item (height, width, 3) views of a concatenated image tensor."""
        out, o = [], 0
        for w, h in self.sizes:
            out.append(flat[o:o + w * h].view(h, w, 3))
            o += w * h
        return out

    def run(self, images=None, targets=None, texel_grads=None, loss=None, use_graph=True,
            item_ms=True, stream=None):
        """Returns (images (total_pixels, 3), texel_grads or None, loss or None,
        BatchStats, per-item completion ms or None)."""
        import torch
        dev = "cuda"
        if images is None:
            images = torch.empty((self.total_pixels, 3), dtype=torch.float32, device=dev)
        if self.adjoint:
            if texel_grads is None:
                texel_grads = torch.empty((max(1, self.n_texels), 3), dtype=torch.float32,
                                          device=dev)
            if loss is None:
                loss = torch.empty(len(self.items), dtype=torch.float32, device=dev)
        io = BatchIO(_ptr(images), _ptr(targets), _ptr(texel_grads) if self.adjoint else None,
                     _ptr(loss) if self.adjoint else None)
        ms = (c_double * len(self.items))() if item_ms else None
        st = BatchStats()
        st.host_item_ms = ctypes.cast(ms, c_void_p) if item_ms else None
        check(lib().rumi_batch_run(self.handle, byref(io), 1 if use_graph else 0, byref(st),
                                   _stream(stream)))
        return (images, texel_grads if self.adjoint else None, loss if self.adjoint else None, st,
                list(ms) if item_ms else None)

    def close(self):
        if getattr(self, "handle", None):
            lib().rumi_batch_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
