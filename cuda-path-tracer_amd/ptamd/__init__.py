"""ctypes binding of libptamd.so (the C-ABI in include/pt_api.h and
include/pt_scenes.h) for tests and benchmarks.

The library is loaded from ``cuda-path-tracer_amd/lib/libptamd.so``; importing
this module raises ImportError when it has not been built.  GPU entry points
raise ``PtError`` with the library's message when a call fails (for example
PT_E_NO_DEVICE on a machine without a GPU).  Arrays are numpy arrays with the
C layouts below; device-pointer entry points take integer addresses (for
instance ``torch.Tensor.data_ptr()``) and a stream handle.
"""
from __future__ import annotations

import ctypes
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(os.path.dirname(_HERE), "lib", "libptamd.so")

if not os.path.exists(LIB_PATH):
    raise ImportError(f"libptamd.so is not built ({LIB_PATH}); run __graft_entry__.build()")

_lib = ctypes.CDLL(LIB_PATH)

PT_OK = 0
PT_E_INVALID_ARG = -1
PT_E_HIP = -2
PT_E_NO_DEVICE = -3
PT_E_NO_SCENE = -4
PT_E_UNSUPPORTED = -5
PT_E_OOM = -6
PT_E_INTERNAL = -7
INVALID = 0xFFFFFFFF
PT_STATUS_STACK_OVERFLOW = 1
PT_STATUS_REFIT_OVERFLOW = 2
FLT_MAX = float(np.finfo(np.float32).max)

# pt_triangle (40 B) and pt_triangle_info (64 B).
TRI_DTYPE = np.dtype([("v0", "<f4", (3,)), ("v1", "<f4", (3,)), ("v2", "<f4", (3,)), ("id", "<u4")])
INFO_DTYPE = np.dtype([("n0", "<f4", (3,)), ("n1", "<f4", (3,)), ("n2", "<f4", (3,)),
                       ("albedo", "<f4", (3,)), ("emission", "<f4", (3,)), ("specularity", "<f4")])
assert TRI_DTYPE.itemsize == 40 and INFO_DTYPE.itemsize == 64
# pt_triangle_detail: vertex colours, uvs, tangents, bitangents, texture indices.
DETAIL_DTYPE = np.dtype([("color", "<f4", (3, 3)), ("uv", "<f4", (3, 2)), ("tangent", "<f4", (3, 3)),
                         ("bitangent", "<f4", (3, 3)), ("albedo_tex", "<u4"), ("normal_tex", "<u4"),
                         ("reserved", "<u4")])
assert DETAIL_DTYPE.itemsize == 144
NO_TEXTURE = 0xFFFFFFFF


class Texture(ctypes.Structure):
    """pt_texture: width, height, RGBA8 texels (row 0 = t = 0, the bottom row)."""
    _fields_ = [("width", ctypes.c_uint32), ("height", ctypes.c_uint32), ("rgba", ctypes.c_void_p)]


def texture_array(images):
    """(h, w, 4) uint8 images -> (ctypes array of Texture, the contiguous
    copies it points into, which must stay alive while it is used)."""
    keep = [np.ascontiguousarray(im, np.uint8) for im in images]
    for im in keep:
        if im.ndim != 3 or im.shape[2] != 4:
            raise ValueError("textures must have shape (h, w, 4)")
    arr = (Texture * max(1, len(keep)))()
    for k, im in enumerate(keep):
        arr[k] = Texture(im.shape[1], im.shape[0], im.ctypes.data)
    return arr, keep


class Settings(ctypes.Structure):
    """pt_settings: bounces, rr_start (INVALID = off), indirect_intensity, stratified."""
    _fields_ = [("bounces", ctypes.c_uint32), ("rr_start", ctypes.c_uint32),
                ("indirect_intensity", ctypes.c_float), ("stratified", ctypes.c_uint32)]


PT_CACHE_TRAIN = 1


class CacheSettings(ctypes.Structure):
    """pt_cache_settings: training paths per frame, Adam steps per frame,
    bounces of a training path past its cache vertex, flags."""
    _fields_ = [("train_samples", ctypes.c_uint32), ("train_steps", ctypes.c_uint32),
                ("extra_bounces", ctypes.c_uint32), ("flags", ctypes.c_uint32)]

    def __init__(self, train_samples=65536, train_steps=4, extra_bounces=8, train=True):
        super().__init__(train_samples, train_steps, extra_bounces, PT_CACHE_TRAIN if train else 0)


class PtError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"pt error {code}: {msg}")
        self.code = code


_vp = ctypes.c_void_p
_u32 = ctypes.c_uint32
_u64 = ctypes.c_uint64
_i = ctypes.c_int
_f = ctypes.c_float

_SIGS = {
    "pt_create": [ctypes.POINTER(_vp), _i],
    "pt_destroy": [_vp],
    "pt_last_error": [_vp, ctypes.c_char_p, ctypes.c_size_t],
    "pt_set_stream": [_vp, _vp],
    "pt_synchronize": [_vp],
    "pt_get_status": [_vp, _vp],
    "pt_upload_scene": [_vp, _vp, _u32, _vp, _u32, _vp, _u32, _vp, _vp, _vp],
    "pt_upload_scenes": [_vp, _vp, _vp, _u32, _vp, _u32, _vp, _vp],
    "pt_get_bvh_scene": [_vp, _u32, _vp, _vp, _vp, _vp, _vp, _vp],
    "pt_trace_closest_scenes_device": [_vp, _vp, _vp, _u32, _vp, _vp, _vp, _vp],
    "pt_set_settings": [_vp, _vp],
    "pt_build_bvh": [_vp, _vp],
    "pt_get_bvh": [_vp, _vp, _vp, _vp, _vp, _vp, _vp],
    "pt_get_scene_info": [_vp, _vp, _vp, _vp],
    "pt_verify_bvh": [_vp, _vp],
    "pt_trace_closest": [_vp, _vp, _u32, _vp, _vp, _vp],
    "pt_trace_closest_device": [_vp, _vp, _u32, _vp, _vp, _vp, _vp],
    "pt_trace_closest_device_noalloc": [_vp, _vp, _u32, _vp, _vp, _vp, _vp],
    "pt_trace_closest_counted": [_vp, _vp, _u32, _vp, _vp, _vp, _vp, _vp],
    "pt_set_primary_rays": [_vp, _vp, _vp],
    "pt_render_pass": [_vp, _i, _i, _u64],
    "pt_render_passes": [_vp, _i, _i, _u32, _vp, _vp, _vp],
    "pt_last_frame_rays": [_vp, _vp],
    "pt_reset_frames": [_vp],
    "pt_read_accumulator": [_vp, _vp, _vp],
    "pt_write_accumulator": [_vp, _i, _i, _vp, _u32],
    "pt_read_image": [_vp, _vp],
    "pt_upload_detail": [_vp, _vp, _u32, _vp, _u32],
    "pt_set_stop_frame": [_vp, _u32],
    "pt_generation_time": [_vp, _vp],
    "pt_read_image_device": [_vp, _vp, _u32, _u32, ctypes.c_size_t, _vp],
    "pt_set_screen_buffer": [_vp, _vp, _u32, _u32, ctypes.c_size_t],
    "pt_write_png": [ctypes.c_char_p, _vp, _u32, _u32, _u32],
    "pt_last_render_ms": [_vp, _vp],
    "pt_scene_bounds": [_vp, _u32, _vp, _vp],
    "pt_gen_soup": [_u64, _u32, _i, _u32, _vp, _vp],
    "pt_gen_sphere": [_u64, _u32, _i, _vp, _vp],
    "pt_gen_hairball": [_u64, _u32, _u32, _i, _vp, _vp],
    "pt_gen_shell_rays": [_u64, _vp, _vp, _f, _u32, _vp],
    "pt_gen_batch_counts": [_u64, _u32, _u32, _u32, _vp],
    "pt_camera_corners": [_vp, _vp, _vp, _f, _f, _f, _f, _vp],
    "pt_camera_corners_basis": [_vp, _vp, _vp, _f, _f, _f, _f, _i, _vp],
    "pt_pass_jitter": [_u64, _u32, _i, _i, _vp, _vp],
    "pt_cache_init": [_vp, _u64, _vp, _vp],
    "pt_cache_set_params": [_vp, _vp],
    "pt_cache_get_params": [_vp, _vp],
    "pt_cache_set_adam": [_vp, _f, _f, _f, _f],
    "pt_cache_infer_device": [_vp, _vp, _u32, _vp, _vp],
    "pt_cache_train_device": [_vp, _vp, _vp, _u32, _vp, _vp],
    "pt_cache_gradient_device": [_vp, _vp, _vp, _u32, _vp, _vp],
    "pt_render_cache_frame": [_vp, _i, _i, _u64, _vp],
    "pt_cache_last_batch": [_vp, _vp, _vp, _u32],
    "pt_cache_last_queries": [_vp, _vp, _vp, _vp],
    "pt_load_obj": [ctypes.c_char_p, _vp, _vp, _vp, _vp],
    "pt_asset_load_obj": [ctypes.c_char_p, _vp, _vp],
    "pt_asset_info": [_vp, _vp, _vp],
    "pt_asset_copy": [_vp, _vp, _vp, _vp],
    "pt_asset_textures": [_vp, _vp],
    "pt_asset_free": [_vp],
    "pt_load_tga": [ctypes.c_char_p, _vp, _vp, _vp],
    "pt_asset_error": [ctypes.c_char_p, ctypes.c_size_t],
    "pt_tune_set": [ctypes.c_char_p, _u32],
    "pt_device_compute_units": [_i, _vp],
    "pt_tune_get": [ctypes.c_char_p, _vp],
}
for _name, _args in _SIGS.items():
    _fn = getattr(_lib, _name)
    _fn.argtypes = _args
    _fn.restype = _i

_lib.pt_cache_param_count.argtypes = []
_lib.pt_cache_param_count.restype = _i

EXPORTED = tuple(_SIGS) + ("pt_cache_param_count",)
CACHE_PARAMS = int(_lib.pt_cache_param_count())


def _ptr(a):
    return None if a is None else a.ctypes.data


def _f32(x, n=None):
    a = np.ascontiguousarray(np.asarray(x, dtype=np.float32))
    if n is not None and a.size != n:
        raise ValueError(f"expected {n} floats")
    return a


def _check(rc, ctx=None):
    if rc < 0:
        buf = ctypes.create_string_buffer(512)
        _lib.pt_last_error(ctx, buf, 512)
        raise PtError(rc, buf.value.decode(errors="replace"))
    return rc


# ---- generators (host only) ---------------------------------------------------

def scene_bounds(tris):
    bmin = np.zeros(3, np.float32)
    bmax = np.zeros(3, np.float32)
    _check(_lib.pt_scene_bounds(_ptr(tris), len(tris), _ptr(bmin), _ptr(bmax)))
    return bmin, bmax


def gen_soup(n, seed, emitter=True, first_id=0):
    cap = n + 2
    tris = np.zeros(cap, TRI_DTYPE)
    infos = np.zeros(cap, INFO_DTYPE)
    k = _check(_lib.pt_gen_soup(seed, n, 1 if emitter else 0, first_id, _ptr(tris), _ptr(infos)))
    return tris[:k].copy(), infos[:k].copy()


def gen_sphere(bands, seed, room=True):
    cap = 2 * bands * bands + 14
    tris = np.zeros(cap, TRI_DTYPE)
    infos = np.zeros(cap, INFO_DTYPE)
    k = _check(_lib.pt_gen_sphere(seed, bands, 1 if room else 0, _ptr(tris), _ptr(infos)))
    return tris[:k].copy(), infos[:k].copy()


def gen_hairball(strands, steps, seed, light=True):
    cap = 2 * strands * steps + 2
    tris = np.zeros(cap, TRI_DTYPE)
    infos = np.zeros(cap, INFO_DTYPE)
    k = _check(_lib.pt_gen_hairball(seed, strands, steps, 1 if light else 0, _ptr(tris), _ptr(infos)))
    return tris[:k].copy(), infos[:k].copy()


def gen_shell_rays(n, seed, bmin, bmax, enlarge=1.5):
    rays = np.zeros((n, 6), np.float32)
    bmin, bmax = _f32(bmin, 3), _f32(bmax, 3)
    _check(_lib.pt_gen_shell_rays(seed, _ptr(bmin), _ptr(bmax), enlarge, n, _ptr(rays)))
    return rays


def batch_counts(nscenes, seed, lo=1024, hi=262144):
    c = np.zeros(nscenes, np.uint32)
    _check(_lib.pt_gen_batch_counts(seed, nscenes, lo, hi, _ptr(c)))
    return c


def gen_batch(nscenes, seed, lo=1024, hi=262144, counts=None):
    """Batched soups: nscenes config-0 soups (no emitter) with log-uniform sizes;
    returns (tris, infos, offsets).  Triangle ids are global."""
    if counts is None:
        counts = batch_counts(nscenes, seed, lo, hi)
    counts = np.asarray(counts, np.uint32)
    offs = np.zeros(len(counts) + 1, np.uint32)
    offs[1:] = np.cumsum(counts)
    tris = np.zeros(int(offs[-1]), TRI_DTYPE)
    infos = np.zeros(int(offs[-1]), INFO_DTYPE)
    for s, c in enumerate(counts):
        if c == 0:
            continue
        t = tris[offs[s]:offs[s + 1]]
        i = infos[offs[s]:offs[s + 1]]
        _check(_lib.pt_gen_soup(seed * 1000003 + s, int(c), 0, int(offs[s]), t.ctypes.data, i.ctypes.data))
    return tris, infos, offs


def camera_corners(eye, target, up=(0.0, 1.0, 0.0), fovy_deg=60.0, aspect=1.0, jx=0.0, jy=0.0):
    d = np.zeros(12, np.float32)
    eye, target, up = _f32(eye, 3), _f32(target, 3), _f32(up, 3)
    _check(_lib.pt_camera_corners(_ptr(eye), _ptr(target), _ptr(up), fovy_deg, aspect, jx, jy, _ptr(d)))
    return d


def camera_corners_basis(right, up, look, fovy, aspect=1.0, jx=0.0, jy=0.0, left_handed=True):
    """screenToRay's four corner directions for a camera rotation with
    columns right, up, look and glm::perspective's fovy (radians)."""
    d = np.zeros(12, np.float32)
    right, up, look = _f32(right, 3), _f32(up, 3), _f32(look, 3)
    _check(_lib.pt_camera_corners_basis(_ptr(right), _ptr(up), _ptr(look), fovy, aspect, jx, jy,
                                        1 if left_handed else 0, _ptr(d)))
    return d


def pass_jitter(seed, pass_index, w, h):
    jx = np.zeros(1, np.float32)
    jy = np.zeros(1, np.float32)
    _check(_lib.pt_pass_jitter(seed, pass_index, w, h, _ptr(jx), _ptr(jy)))
    return float(jx[0]), float(jy[0])


class ObjOptions(ctypes.Structure):
    """pt_obj_options: left_handed (the reference's import flags: z mirrored,
    winding reversed), normalize_albedo (albedo = Kd / |Kd|)."""
    _fields_ = [("left_handed", ctypes.c_uint32), ("normalize_albedo", ctypes.c_uint32)]


def load_obj(path, left_handed=True, normalize_albedo=True):
    """OBJ/MTL file -> (tris, infos) in file order, ids 0..n-1."""
    opt = ObjOptions(1 if left_handed else 0, 1 if normalize_albedo else 0)
    p = os.fsencode(path)
    n = np.zeros(1, np.uint32)
    rc = _lib.pt_load_obj(p, ctypes.byref(opt), None, None, _ptr(n))
    if rc != PT_OK:
        buf = ctypes.create_string_buffer(512)
        _lib.pt_asset_error(buf, len(buf))
        raise PtError(rc, buf.value.decode(errors="replace"))
    tris = np.zeros(int(n[0]), TRI_DTYPE)
    infos = np.zeros(int(n[0]), INFO_DTYPE)
    rc = _lib.pt_load_obj(p, ctypes.byref(opt), _ptr(tris), _ptr(infos), _ptr(n))
    if rc != PT_OK:
        buf = ctypes.create_string_buffer(512)
        _lib.pt_asset_error(buf, len(buf))
        raise PtError(rc, buf.value.decode(errors="replace"))
    return tris, infos


def _asset_error(rc):
    buf = ctypes.create_string_buffer(512)
    _lib.pt_asset_error(buf, len(buf))
    raise PtError(rc, buf.value.decode(errors="replace"))


def load_obj_asset(path, left_handed=True, normalize_albedo=True):
    """OBJ/MTL with surface detail -> (tris, infos, detail, textures), the
    textures as (h, w, 4) uint8 arrays (row 0 = bottom)."""
    opt = ObjOptions(1 if left_handed else 0, 1 if normalize_albedo else 0)
    h = ctypes.c_void_p()
    rc = _lib.pt_asset_load_obj(os.fsencode(path), ctypes.byref(opt), ctypes.byref(h))
    if rc != PT_OK:
        _asset_error(rc)
    try:
        n = np.zeros(2, np.uint32)
        _check(_lib.pt_asset_info(h, _ptr(n[:1]), n[1:].ctypes.data))
        tris = np.zeros(int(n[0]), TRI_DTYPE)
        infos = np.zeros(int(n[0]), INFO_DTYPE)
        detail = np.zeros(int(n[0]), DETAIL_DTYPE)
        _check(_lib.pt_asset_copy(h, _ptr(tris), _ptr(infos), _ptr(detail)))
        arr = (Texture * max(1, int(n[1])))()
        _check(_lib.pt_asset_textures(h, ctypes.cast(arr, ctypes.c_void_p)))
        textures = []
        for k in range(int(n[1])):
            t = arr[k]
            buf = (ctypes.c_uint8 * (4 * t.width * t.height)).from_address(t.rgba)
            textures.append(np.frombuffer(buf, np.uint8).reshape(t.height, t.width, 4).copy())
    finally:
        _lib.pt_asset_free(h)
    return tris, infos, detail, textures


def load_tga(path):
    """TGA file -> (h, w, 4) uint8 RGBA, row 0 = bottom."""
    wh = np.zeros(2, np.uint32)
    p = os.fsencode(path)
    rc = _lib.pt_load_tga(p, _ptr(wh[:1]), wh[1:].ctypes.data, None)
    if rc != PT_OK:
        _asset_error(rc)
    img = np.zeros((int(wh[1]), int(wh[0]), 4), np.uint8)
    rc = _lib.pt_load_tga(p, _ptr(wh[:1]), wh[1:].ctypes.data, _ptr(img))
    if rc != PT_OK:
        _asset_error(rc)
    return img


PT_PNG_SRGB = 1
PT_PNG_TOP_DOWN = 2


def write_png(path, rgba, srgb=False, top_down=False):
    """8-bit RGB PNG of an (h, w, 4) float32 image (pt_read_image's layout:
    row 0 at the bottom unless top_down)."""
    img = np.ascontiguousarray(rgba, np.float32)
    if img.ndim != 3 or img.shape[2] != 4:
        raise ValueError("rgba must have shape (h, w, 4)")
    flags = (PT_PNG_SRGB if srgb else 0) | (PT_PNG_TOP_DOWN if top_down else 0)
    rc = _lib.pt_write_png(os.fsencode(path), _ptr(img), img.shape[1], img.shape[0], flags)
    if rc != PT_OK:
        buf = ctypes.create_string_buffer(512)
        _lib.pt_asset_error(buf, len(buf))
        raise PtError(rc, buf.value.decode(errors="replace"))


def tune_set(name, value):
    _check(_lib.pt_tune_set(name.encode(), int(value)))


def device_compute_units(device=0):
    v = np.zeros(1, np.uint32)
    _check(_lib.pt_device_compute_units(device, _ptr(v)))
    return int(v[0])


def tune_get(name):
    v = np.zeros(1, np.uint32)
    _check(_lib.pt_tune_get(name.encode(), _ptr(v)))
    return int(v[0])


# ---- device context -------------------------------------------------------------

class Context:
    """One pt_context (one device, one scene, one accumulator)."""

    def __init__(self, device=0):
        h = _vp()
        _check(_lib.pt_create(ctypes.byref(h), device))
        self._h = h
        self.n = 0

    def close(self):
        if self._h:
            _lib.pt_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def _c(self, rc):
        return _check(rc, self._h)

    def upload_scene(self, tris, infos, bmin=None, bmax=None, settings=None, lights=None):
        tris = np.ascontiguousarray(tris, TRI_DTYPE)
        infos = np.ascontiguousarray(infos, INFO_DTYPE)
        bmin_a = None if bmin is None else _f32(bmin, 3)
        bmax_a = None if bmax is None else _f32(bmax, 3)
        st = None if settings is None else ctypes.byref(settings)
        nl = 0 if lights is None else len(lights)
        la = None if lights is None else np.ascontiguousarray(lights, np.float32)
        self._c(_lib.pt_upload_scene(self._h, _ptr(tris), len(tris), _ptr(infos), len(infos),
                                     _ptr(la), nl, _ptr(bmin_a), _ptr(bmax_a), st))
        self.n = len(tris)

    def upload_scenes(self, tris, offsets, infos, boxes=None, settings=None):
        """Batched scenes: scene s = tris[offsets[s]:offsets[s+1]]."""
        tris = np.ascontiguousarray(tris, TRI_DTYPE)
        infos = np.ascontiguousarray(infos, INFO_DTYPE)
        offs = np.ascontiguousarray(offsets, np.uint32)
        bx = None if boxes is None else np.ascontiguousarray(boxes, np.float32)
        st = None if settings is None else ctypes.byref(settings)
        self._c(_lib.pt_upload_scenes(self._h, _ptr(tris), _ptr(offs), len(offs) - 1, _ptr(infos),
                                      len(infos), _ptr(bx), st))
        self.n = int(offs[-1])
        self.offsets = offs.copy()

    def get_bvh_scene(self, scene):
        n = int(self.offsets[scene + 1] - self.offsets[scene])
        m = max(n - 1, 0)
        out = dict(morton=np.zeros(n, np.uint32), perm=np.zeros(n, np.uint32),
                   children=np.zeros(m, np.uint32), parents_leaf=np.zeros(n, np.uint32),
                   parents_internal=np.zeros(m, np.uint32), bboxes=np.zeros((m, 6), np.float32))
        self._c(_lib.pt_get_bvh_scene(self._h, scene, _ptr(out["morton"]), _ptr(out["perm"]),
                                      _ptr(out["children"]), _ptr(out["parents_leaf"]),
                                      _ptr(out["parents_internal"]), _ptr(out["bboxes"])))
        return out

    def trace_closest_scenes_device(self, d_rays, d_scene, n, d_id, d_t, d_uv=0, stream=0):
        self._c(_lib.pt_trace_closest_scenes_device(self._h, d_rays, d_scene, n, d_id or None,
                                                    d_t or None, d_uv or None, stream or None))

    def trace_closest_scenes(self, rays, scene_ids):
        from . import hip
        rays = _f32(rays).reshape(-1, 6)
        n = len(rays)
        d_r = hip.DeviceArray.from_host(rays)
        d_s = hip.DeviceArray.from_host(np.ascontiguousarray(scene_ids, np.uint32))
        d_i = hip.DeviceArray(n, np.uint32)
        d_t = hip.DeviceArray(n, np.float32)
        d_uv = hip.DeviceArray((n, 2), np.float32)
        self.trace_closest_scenes_device(d_r.ptr, d_s.ptr, n, d_i.ptr, d_t.ptr, d_uv.ptr)
        self.synchronize()
        return d_i.to_host(), d_t.to_host(), d_uv.to_host()

    def set_settings(self, settings):
        self._c(_lib.pt_set_settings(self._h, ctypes.byref(settings)))

    def build_bvh(self, timed=True):
        ms = np.zeros(1, np.float32)
        self._c(_lib.pt_build_bvh(self._h, _ptr(ms) if timed else None))
        return float(ms[0]) if timed else None

    def get_bvh(self):
        n = self.n
        m = max(n - 1, 0)
        out = dict(morton=np.zeros(n, np.uint32), perm=np.zeros(n, np.uint32),
                   children=np.zeros(m, np.uint32), parents_leaf=np.zeros(n, np.uint32),
                   parents_internal=np.zeros(m, np.uint32), bboxes=np.zeros((m, 6), np.float32))
        self._c(_lib.pt_get_bvh(self._h, _ptr(out["morton"]), _ptr(out["perm"]), _ptr(out["children"]),
                                _ptr(out["parents_leaf"]), _ptr(out["parents_internal"]),
                                _ptr(out["bboxes"])))
        return out

    def scene_info(self):
        n = np.zeros(1, np.uint32)
        bmin = np.zeros(3, np.float32)
        bmax = np.zeros(3, np.float32)
        self._c(_lib.pt_get_scene_info(self._h, _ptr(n), _ptr(bmin), _ptr(bmax)))
        return int(n[0]), bmin, bmax

    def verify_bvh(self):
        """The build self-check: raises PtError unless every scene's root box
        equals its enclosing box; returns 0 (the mismatch count) otherwise."""
        bad = np.zeros(1, np.uint32)
        self._c(_lib.pt_verify_bvh(self._h, _ptr(bad)))
        return int(bad[0])

    def trace_closest(self, rays):
        rays = _f32(rays).reshape(-1, 6)
        n = len(rays)
        ids = np.zeros(n, np.uint32)
        t = np.zeros(n, np.float32)
        uv = np.zeros((n, 2), np.float32)
        self._c(_lib.pt_trace_closest(self._h, _ptr(rays), n, _ptr(ids), _ptr(t), _ptr(uv)))
        return ids, t, uv

    def trace_closest_device(self, d_rays, n, d_id, d_t, d_uv=0, stream=0):
        self._c(_lib.pt_trace_closest_device(self._h, d_rays, n, d_id or None, d_t or None,
                                             d_uv or None, stream or None))

    def trace_closest_device_noalloc(self, d_rays, n, d_id, d_t, d_uv=0, stream=0):
        """trace_closest_device through the library's static scratch block
        (no allocation; calls overlapping on one GPU race)."""
        self._c(_lib.pt_trace_closest_device_noalloc(self._h, d_rays, n, d_id or None, d_t or None,
                                                     d_uv or None, stream or None))

    def trace_closest_counted(self, d_rays, n, d_id=0, d_t=0, d_uv=0, stream=0):
        """Node visits, box tests, triangle tests (totals) and the deepest
        deferred-node stack of any ray."""
        c = np.zeros(4, np.uint64)
        self._c(_lib.pt_trace_closest_counted(self._h, d_rays, n, d_id or None, d_t or None,
                                              d_uv or None, _ptr(c), stream or None))
        return c

    def set_primary_rays(self, dirs, eye):
        dirs, eye = _f32(dirs, 12), _f32(eye, 3)
        self._c(_lib.pt_set_primary_rays(self._h, _ptr(dirs), _ptr(eye)))

    def render_pass(self, w, h, seed):
        self._c(_lib.pt_render_pass(self._h, w, h, seed))

    def render_passes(self, w, h, seeds, dirs=None, scenes=None):
        """len(seeds) passes in one frame; dirs: (npasses, 12) corner rays or
        None; scenes: scene of each pass (batched upload) or None."""
        seeds = np.ascontiguousarray(seeds, np.uint64)
        d = None if dirs is None else _f32(dirs, 12 * len(seeds))
        sc = None if scenes is None else np.ascontiguousarray(scenes, np.uint32)
        self._c(_lib.pt_render_passes(self._h, w, h, len(seeds), _ptr(d), _ptr(seeds), _ptr(sc)))

    def last_frame_rays(self):
        v = np.zeros(1, np.uint64)
        self._c(_lib.pt_last_frame_rays(self._h, _ptr(v)))
        return int(v[0])

    def reset_frames(self):
        self._c(_lib.pt_reset_frames(self._h))

    def read_accumulator(self, w, h):
        acc = np.zeros((h, w, 3), np.float64)
        spp = np.zeros(1, np.uint32)
        self._c(_lib.pt_read_accumulator(self._h, _ptr(acc), _ptr(spp)))
        return acc, int(spp[0])

    def write_accumulator(self, acc, spp):
        acc = np.ascontiguousarray(acc, np.float64)
        h, w, _ = acc.shape
        self._c(_lib.pt_write_accumulator(self._h, w, h, _ptr(acc), spp))

    def read_image(self, w, h):
        img = np.zeros((h, w, 4), np.float32)
        self._c(_lib.pt_read_image(self._h, _ptr(img)))
        return img

    def upload_detail(self, detail, textures=()):
        """Surface detail (DETAIL_DTYPE per triangle id) and textures ((h, w, 4)
        uint8, row 0 = bottom); an empty detail removes it."""
        detail = np.ascontiguousarray(detail, DETAIL_DTYPE)
        arr, keep = texture_array(textures)
        self._c(_lib.pt_upload_detail(self._h, _ptr(detail) if len(detail) else None, len(detail),
                                      ctypes.cast(arr, ctypes.c_void_p) if keep else None, len(keep)))

    def read_image_device(self, d_rgba, w, h, row_pitch=None, stream=None):
        """Running-mean image into device memory (float4 texels)."""
        self._c(_lib.pt_read_image_device(self._h, d_rgba, w, h, row_pitch or 16 * w, stream))

    def set_screen_buffer(self, d_rgba, w=0, h=0, row_pitch=None):
        """Device RGBA32F buffer the image is written into after every frame
        (d_rgba None or 0 detaches)."""
        self._c(_lib.pt_set_screen_buffer(self._h, d_rgba or None, w, h, row_pitch or 16 * w))

    def set_stop_frame(self, frames):
        """Render calls stop once `frames` passes have accumulated (0: never)."""
        self._c(_lib.pt_set_stop_frame(self._h, int(frames)))

    def generation_time(self):
        """Seconds to reach the stop frame, or -1 while not reached."""
        v = np.zeros(1, np.float64)
        self._c(_lib.pt_generation_time(self._h, _ptr(v)))
        return float(v[0])

    def last_render_ms(self):
        ms = np.zeros(1, np.float32)
        self._c(_lib.pt_last_render_ms(self._h, _ptr(ms)))
        return float(ms[0])

    def synchronize(self):
        self._c(_lib.pt_synchronize(self._h))

    def get_status(self):
        """PT_STATUS_* flags raised since the last read (cleared by reading)."""
        v = np.zeros(1, np.uint32)
        self._c(_lib.pt_get_status(self._h, _ptr(v)))
        return int(v[0])

    def set_stream(self, stream):
        self._c(_lib.pt_set_stream(self._h, stream or None))

    # ---- learned radiance cache ------------------------------------------------
    def cache_init(self, seed, bmin, bmax):
        bmin, bmax = _f32(bmin, 3), _f32(bmax, 3)
        self._c(_lib.pt_cache_init(self._h, seed, _ptr(bmin), _ptr(bmax)))

    def cache_set_params(self, params):
        params = _f32(params, CACHE_PARAMS)
        self._c(_lib.pt_cache_set_params(self._h, _ptr(params)))

    def cache_get_params(self):
        out = np.zeros(CACHE_PARAMS, np.float32)
        self._c(_lib.pt_cache_get_params(self._h, _ptr(out)))
        return out

    def cache_set_adam(self, lr=1e-3, beta1=0.9, beta2=0.99, eps=1e-8):
        self._c(_lib.pt_cache_set_adam(self._h, lr, beta1, beta2, eps))

    def cache_infer_device(self, d_raw, n, d_out, stream=0):
        self._c(_lib.pt_cache_infer_device(self._h, d_raw or None, n, d_out or None, stream or None))

    def cache_train_device(self, d_raw, d_target, n, d_loss=0, stream=0):
        self._c(_lib.pt_cache_train_device(self._h, d_raw or None, d_target or None, n,
                                           d_loss or None, stream or None))

    def cache_gradient_device(self, d_raw, d_target, n, d_grad, stream=0):
        self._c(_lib.pt_cache_gradient_device(self._h, d_raw or None, d_target or None, n,
                                              d_grad or None, stream or None))

    def cache_infer(self, raw):
        """raw: (n, 12) samples -> (n, 3) predicted radiance."""
        from . import hip
        raw = _f32(raw).reshape(-1, 12)
        n = len(raw)
        d_raw = hip.DeviceArray.from_host(raw)
        d_out = hip.DeviceArray((n, 3), np.float32)
        self.cache_infer_device(d_raw.ptr, n, d_out.ptr)
        self.synchronize()
        return d_out.to_host()

    def cache_train(self, raw, target):
        """One Adam step on the batch; returns the per-sample losses."""
        from . import hip
        raw = _f32(raw).reshape(-1, 12)
        n = len(raw)
        d_raw = hip.DeviceArray.from_host(raw)
        d_tgt = hip.DeviceArray.from_host(_f32(target).reshape(n, 3))
        d_loss = hip.DeviceArray(n, np.float32)
        self.cache_train_device(d_raw.ptr, d_tgt.ptr, n, d_loss.ptr)
        self.synchronize()
        return d_loss.to_host()

    def cache_gradient(self, raw, target):
        from . import hip
        raw = _f32(raw).reshape(-1, 12)
        n = len(raw)
        d_raw = hip.DeviceArray.from_host(raw)
        d_tgt = hip.DeviceArray.from_host(_f32(target).reshape(n, 3))
        d_g = hip.DeviceArray(CACHE_PARAMS, np.float32)
        self.cache_gradient_device(d_raw.ptr, d_tgt.ptr, n, d_g.ptr)
        self.synchronize()
        return d_g.to_host()

    def render_cache_frame(self, w, h, seed, settings=None):
        settings = settings or CacheSettings()
        self._c(_lib.pt_render_cache_frame(self._h, w, h, seed, ctypes.byref(settings)))

    def cache_last_batch(self, n):
        raw = np.zeros((n, 12), np.float32)
        tgt = np.zeros((n, 3), np.float32)
        self._c(_lib.pt_cache_last_batch(self._h, _ptr(raw), _ptr(tgt), n))
        return raw, tgt

    def cache_last_queries(self, w, h):
        """(raw (h*w, 12), predictions (h*w, 3), valid (h*w,) bool) of the last cache frame."""
        raw = np.zeros((w * h, 12), np.float32)
        pred = np.zeros((w * h, 3), np.float32)
        valid = np.zeros(w * h, np.uint32)
        self._c(_lib.pt_cache_last_queries(self._h, _ptr(raw), _ptr(pred), _ptr(valid)))
        return raw, pred, valid.astype(bool)
