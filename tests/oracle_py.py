"""ctypes binding of oracle/liboracle.so -- the CPU restatement of the
reference hot path.  Test infrastructure: only tests/, __graft_entry__.smoke()
and bench.py's cpu_baseline leg load it."""
from __future__ import annotations

import ctypes
import os

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIB_PATH = os.path.join(ROOT, "oracle", "liboracle.so")
if not os.path.exists(LIB_PATH):
    raise ImportError(f"oracle not built ({LIB_PATH}); run __graft_entry__.build()")
_lib = ctypes.CDLL(LIB_PATH)

_vp, _u32, _u64, _i, _f = ctypes.c_void_p, ctypes.c_uint32, ctypes.c_uint64, ctypes.c_int, ctypes.c_float
_SIGS = {
    "or_morton": ([_vp, _u32, _vp, _vp, _vp], None),
    "or_encode30": ([_u32, _u32, _u32], _u32),
    "or_quantize": ([_f, _u32], _u32),
    "or_sort": ([_vp, _u32, _vp, _vp], None),
    "or_radix_tree": ([_vp, _u32, _vp, _vp, _vp], None),
    "or_refit": ([_vp, _vp, _vp, _vp, _vp, _u32, _vp], None),
    "or_build": ([_vp, _u32, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp], _i),
    "or_ray_triangle": ([_vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp], _i),
    "or_ray_aabb": ([_vp, _vp, _vp, _vp, _vp], _i),
    "or_trace": ([_vp, _vp, _vp, _vp, _u32, _vp, _u32, _vp, _vp, _vp, _vp, _i], None),
    "or_brute": ([_vp, _u32, _vp, _u32, _vp, _vp, _vp, _i], None),
    "or_render_pass": ([_vp, _vp, _vp, _vp, _u32, _vp, _vp, _i, _i, _vp, _vp, _u32, _u64, _vp, _vp, _i], None),
    "or_mlp_param_count": ([], _i),
    "or_mlp_encode": ([_vp, _vp, _vp, _vp], None),
    "or_mlp_forward": ([_vp, _vp, _u32, _vp, _vp, _vp], None),
    "or_mlp_loss_grad": ([_vp, _vp, _vp, _u32, _vp, _vp, _vp], ctypes.c_double),
    "or_adam_step": ([_vp, _vp, _vp, _vp, _u32, _u32, _f, _f, _f, _f], None),
    "or_render_cache_pass": ([_vp, _vp, _vp, _vp, _u32, _vp, _vp, _u32, _vp, _i, _i, _vp, _vp, _u32, _u64, _vp,
                              _vp, _vp, _u32, _u32, _vp, _vp, _vp, _i], None),
    "or_render_pass_lit": ([_vp, _vp, _vp, _vp, _u32, _vp, _vp, _u32, _vp, _i, _i, _vp, _vp, _u32, _u64, _vp, _vp,
                            _i], None),
    "or_pow_pos": ([_f, _f], _f),
    "or_render_pass_detail": ([_vp, _vp, _vp, _vp, _u32, _vp, _vp, _vp, _u32, _vp, _u32, _vp, _i, _i, _vp, _vp,
                               _u32, _u64, _vp, _vp, _i], None),
    "or_render_cache_pass_detail": ([_vp, _vp, _vp, _vp, _u32, _vp, _vp, _vp, _u32, _vp, _u32, _vp, _i, _i, _vp,
                                     _vp, _u32, _u64, _vp, _vp, _vp, _u32, _u32, _vp, _vp, _vp, _i], None),
    "or_tex_sample": ([_vp, _f, _f, _vp], None),
    "or_render_pass_stats": ([_vp, _vp, _vp, _vp, _u32, _vp, _vp, _i, _i, _vp, _vp, _u32, _u64, _vp, _vp, _i], None),
    "or_shade_detail": ([_vp, _vp, _vp, _u32, _f, _f, _vp, _vp], None),
    "or_pcg4d": ([_vp, _vp], None),
    "or_sincos2pi": ([_f, _vp, _vp], None),
    "or_to_hemisphere": ([_vp, _vp, _vp], None),
    "or_screen_to_ray": ([_vp, _vp, _f, _f, _f, _f, _f, _f, _f, _vp, _i, _vp], None),
}
FAST_LIB_PATH = os.path.join(ROOT, "oracle", "liboracle_fast.so")


def _bind(lib):
    for _n, (_a, _r) in _SIGS.items():
        _fn = getattr(lib, _n)
        _fn.argtypes = _a
        _fn.restype = _r
    return lib


_bind(_lib)


def fast_library_usable():
    """Whether liboracle_fast.so (AVX2 code) exists and this host reports
    the instruction sets it was compiled for."""
    if not os.path.exists(FAST_LIB_PATH):
        return False
    try:
        with open("/proc/cpuinfo") as f:
            flags = set(next(line for line in f if line.startswith("flags")).split(":", 1)[1].split())
    except (OSError, StopIteration):
        return False
    return {"avx2", "fma", "bmi2", "movbe"} <= flags


def select_library(fast):
    """Bind every function of this module to liboracle_fast.so (fast=True,
    for CPU baseline timings) or liboracle.so.  Returns the path in use."""
    global _lib
    path = FAST_LIB_PATH if fast else LIB_PATH
    _lib = _bind(ctypes.CDLL(path))
    return path

THREADS = max(1, min(16, os.cpu_count() or 1))


def _p(a):
    return None if a is None else a.ctypes.data


def _f32(x):
    return np.ascontiguousarray(np.asarray(x, np.float32))


def encode30(x, y, z):
    return int(_lib.or_encode30(x, y, z))


def quantize(x, levels=1024):
    return int(_lib.or_quantize(float(np.float32(x)), levels))


def morton(tris, bmin, bmax):
    keys = np.zeros(len(tris), np.uint32)
    bmin, bmax = _f32(bmin), _f32(bmax)
    _lib.or_morton(_p(tris), len(tris), _p(bmin), _p(bmax), _p(keys))
    return keys


def sort(keys):
    keys = np.ascontiguousarray(keys, np.uint32)
    sk = np.zeros_like(keys)
    perm = np.zeros_like(keys)
    _lib.or_sort(_p(keys), len(keys), _p(sk), _p(perm))
    return sk, perm


def radix_tree(sorted_keys):
    sk = np.ascontiguousarray(sorted_keys, np.uint32)
    n = len(sk)
    m = max(n - 1, 0)
    ch = np.zeros(m, np.uint32)
    pl = np.zeros(n, np.uint32)
    pi = np.zeros(m, np.uint32)
    _lib.or_radix_tree(_p(sk), n, _p(ch), _p(pl), _p(pi))
    return ch, pl, pi


def build(tris, bmin, bmax):
    n = len(tris)
    m = max(n - 1, 0)
    out = dict(morton=np.zeros(n, np.uint32), perm=np.zeros(n, np.uint32),
               children=np.zeros(m, np.uint32), parents_leaf=np.zeros(n, np.uint32),
               parents_internal=np.zeros(m, np.uint32), bboxes=np.zeros((m, 6), np.float32))
    bmin, bmax = _f32(bmin), _f32(bmax)
    _lib.or_build(_p(tris), n, _p(bmin), _p(bmax), _p(out["morton"]), _p(out["perm"]),
                  _p(out["children"]), _p(out["parents_leaf"]), _p(out["parents_internal"]),
                  _p(out["bboxes"]))
    return out


def ray_triangle(o, d, p0, p1, p2):
    u = np.zeros(1, np.float32)
    v = np.zeros(1, np.float32)
    t = np.zeros(1, np.float32)
    o, d, p0, p1, p2 = _f32(o), _f32(d), _f32(p0), _f32(p1), _f32(p2)
    hit = _lib.or_ray_triangle(_p(o), _p(d), _p(p0), _p(p1), _p(p2), _p(u), _p(v), _p(t))
    return bool(hit), float(u[0]), float(v[0]), float(t[0])


def ray_aabb(o, d, bmin, bmax):
    t = np.zeros(1, np.float32)
    o, d, bmin, bmax = _f32(o), _f32(d), _f32(bmin), _f32(bmax)
    hit = _lib.or_ray_aabb(_p(o), _p(d), _p(bmin), _p(bmax), _p(t))
    return bool(hit), float(t[0])


def trace(tris, bvh, rays, counters=False, threads=THREADS):
    rays = _f32(rays).reshape(-1, 6)
    n = len(rays)
    ids = np.zeros(n, np.uint32)
    t = np.zeros(n, np.float32)
    uv = np.zeros((n, 2), np.float32)
    c = np.zeros(4, np.uint64) if counters else None
    _lib.or_trace(_p(tris), _p(bvh["perm"]), _p(bvh["children"]), _p(bvh["bboxes"]), len(tris),
                  _p(rays), n, _p(ids), _p(t), _p(uv), _p(c), threads)
    return (ids, t, uv, c) if counters else (ids, t, uv)


def brute(tris, rays, threads=THREADS):
    rays = _f32(rays).reshape(-1, 6)
    n = len(rays)
    ids = np.zeros(n, np.uint32)
    t = np.zeros(n, np.float32)
    uv = np.zeros((n, 2), np.float32)
    _lib.or_brute(_p(tris), len(tris), _p(rays), n, _p(ids), _p(t), _p(uv), threads)
    return ids, t, uv


def render_pass(tris, bvh, infos, settings, w, h, dirs, eye, pass_index, seed, accum=None,
                threads=THREADS):
    if accum is None:
        accum = np.zeros((h, w, 3), np.float64)
    segs = np.zeros(1, np.uint64)
    dirs, eye = _f32(dirs), _f32(eye)
    _lib.or_render_pass(_p(tris), _p(bvh["perm"]), _p(bvh["children"]), _p(bvh["bboxes"]), len(tris),
                        _p(infos), ctypes.byref(settings), w, h, _p(dirs), _p(eye),
                        pass_index, seed, _p(accum), _p(segs), threads)
    return accum, int(segs[0])


def pcg4d(v):
    a = np.ascontiguousarray(np.asarray(v, np.uint32))
    out = np.zeros(4, np.uint32)
    _lib.or_pcg4d(_p(a), _p(out))
    return out


def sincos2pi(u):
    s = np.zeros(1, np.float32)
    c = np.zeros(1, np.float32)
    _lib.or_sincos2pi(float(np.float32(u)), _p(s), _p(c))
    return float(s[0]), float(c[0])


def to_hemisphere(v, n):
    out = np.zeros(3, np.float32)
    v, n = _f32(v), _f32(n)
    _lib.or_to_hemisphere(_p(v), _p(n), _p(out))
    return out


# ---- radiance-cache MLP (FP64) --------------------------------------------------
def mlp_param_count():
    return int(_lib.or_mlp_param_count())


def mlp_encode(raw, bmin, bmax):
    raw, bmin, bmax = _f32(raw), _f32(bmin), _f32(bmax)
    x = np.zeros(96, np.float64)
    _lib.or_mlp_encode(_p(raw), _p(bmin), _p(bmax), _p(x))
    return x


def mlp_forward(params, raw, bmin, bmax):
    params, raw, bmin, bmax = _f32(params), _f32(raw).reshape(-1, 12), _f32(bmin), _f32(bmax)
    out = np.zeros((len(raw), 3), np.float64)
    _lib.or_mlp_forward(_p(params), _p(raw), len(raw), _p(bmin), _p(bmax), _p(out))
    return out


def mlp_loss_grad(params, raw, target, bmin, bmax):
    params, raw = _f32(params), _f32(raw).reshape(-1, 12)
    target, bmin, bmax = _f32(target).reshape(-1, 3), _f32(bmin), _f32(bmax)
    g = np.zeros(mlp_param_count(), np.float64)
    loss = _lib.or_mlp_loss_grad(_p(params), _p(raw), _p(target), len(raw), _p(bmin), _p(bmax), _p(g))
    return float(loss), g


def adam_step(w, m, v, g, t, lr=1e-3, b1=0.9, b2=0.99, eps=1e-8):
    """In place on float32 w, m, v."""
    g = np.ascontiguousarray(g, np.float64)
    _lib.or_adam_step(_p(w), _p(m), _p(v), _p(g), len(w), t, lr, b1, b2, eps)


def _lights(lights):
    if lights is None or len(lights) == 0:
        return None, 0
    la = np.ascontiguousarray(np.asarray(lights, np.float32).reshape(-1, 7))
    return la, len(la)


def render_pass_lit(tris, bvh, infos, lights, settings, w, h, dirs, eye, pass_index, seed, accum=None,
                    threads=THREADS):
    """or_render_pass with directional lights: rows (color.rgb, direction.xyz, intensity)."""
    if accum is None:
        accum = np.zeros((h, w, 3), np.float64)
    la, nl = _lights(lights)
    segs = np.zeros(1, np.uint64)
    dirs, eye = _f32(dirs), _f32(eye)
    _lib.or_render_pass_lit(_p(tris), _p(bvh["perm"]), _p(bvh["children"]), _p(bvh["bboxes"]), len(tris),
                            _p(infos), _p(la), nl, ctypes.byref(settings), w, h, _p(dirs), _p(eye),
                            pass_index, seed, _p(accum), _p(segs), threads)
    return accum, int(segs[0])


def _tex(textures):
    import ptamd
    arr, keep = ptamd.texture_array(textures or [])
    return (ctypes.cast(arr, ctypes.c_void_p).value if keep else None), len(keep), (arr, keep)


def render_pass_detail(tris, bvh, infos, detail, textures, settings, w, h, dirs, eye, pass_index, seed,
                       accum=None, lights=None, threads=THREADS):
    """or_render_pass with surface detail (DETAIL_DTYPE per id, (h, w, 4) uint8 textures)."""
    if accum is None:
        accum = np.zeros((h, w, 3), np.float64)
    la, nl = _lights(lights)
    tp, nt, keep = _tex(textures)
    segs = np.zeros(1, np.uint64)
    dirs, eye = _f32(dirs), _f32(eye)
    _lib.or_render_pass_detail(_p(tris), _p(bvh["perm"]), _p(bvh["children"]), _p(bvh["bboxes"]), len(tris),
                               _p(infos), _p(detail), tp, nt, _p(la), nl, ctypes.byref(settings), w, h, _p(dirs),
                               _p(eye), pass_index, seed, _p(accum), _p(segs), threads)
    del keep
    return accum, int(segs[0])


def render_pass_stats(tris, bvh, infos, settings, w, h, dirs, eye, pass_index, seed, accum=None,
                      threads=THREADS):
    """or_render_pass with traversal statistics: (accum, {segments, visits,
    boxes, tris, hits})."""
    if accum is None:
        accum = np.zeros((h, w, 3), np.float64)
    st = np.zeros(5, np.uint64)
    dirs, eye = _f32(dirs), _f32(eye)
    _lib.or_render_pass_stats(_p(tris), _p(bvh["perm"]), _p(bvh["children"]), _p(bvh["bboxes"]), len(tris),
                              _p(infos), ctypes.byref(settings), w, h, _p(dirs), _p(eye), pass_index, seed,
                              _p(accum), _p(st), threads)
    keys = ("segments", "visits", "boxes", "tris", "hits")
    return accum, {k: int(v) for k, v in zip(keys, st)}


def tex_sample(image, s, t):
    """tex2D of one (h, w, 4) uint8 texture at (s, t): rgb float32."""
    tp, nt, keep = _tex([image])
    out = np.zeros(3, np.float32)
    _lib.or_tex_sample(tp, float(np.float32(s)), float(np.float32(t)), _p(out))
    del keep
    return out


def shade_detail(info, detail, textures, u, v):
    """(normal, albedo) of a hit at barycentrics (u, v)."""
    tp, nt, keep = _tex(textures)
    info = np.ascontiguousarray(info)
    detail = None if detail is None else np.ascontiguousarray(detail)
    n = np.zeros(3, np.float32)
    a = np.zeros(3, np.float32)
    _lib.or_shade_detail(_p(info), _p(detail), tp, nt, float(np.float32(u)), float(np.float32(v)), _p(n), _p(a))
    del keep
    return n, a


def pow_pos(x, y):
    return float(_lib.or_pow_pos(float(np.float32(x)), float(np.float32(y))))


def render_cache_pass(tris, bvh, infos, settings, w, h, dirs, eye, pass_index, seed, params, cbmin, cbmax,
                      ntrain, extra_bounces, accum=None, threads=THREADS, lights=None, detail=None,
                      textures=None):
    """One cache frame: returns (accum, train_raw (ntrain, 12), train_target (ntrain, 3))."""
    if accum is None:
        accum = np.zeros((h, w, 3), np.float64)
    raw = np.zeros((ntrain, 12), np.float32)
    tgt = np.zeros((ntrain, 3), np.float32)
    dirs, eye, params = _f32(dirs), _f32(eye), _f32(params)
    cbmin, cbmax = _f32(cbmin), _f32(cbmax)
    la, nl = _lights(lights)
    tp, nt, keep = _tex(textures)
    _lib.or_render_cache_pass_detail(_p(tris), _p(bvh["perm"]), _p(bvh["children"]), _p(bvh["bboxes"]),
                                     len(tris), _p(infos), _p(detail), tp, nt, _p(la), nl, ctypes.byref(settings),
                                     w, h, _p(dirs), _p(eye), pass_index, seed,
                                     _p(params), _p(cbmin), _p(cbmax), ntrain, extra_bounces, _p(accum),
                                     _p(raw), _p(tgt), threads)
    del keep
    return accum, raw, tgt


def screen_to_ray(rot_cols, pos, fovy, w, h, znear, zfar, jx, jy, p, left_handed=True):
    """screenToRay's direction for screen point p (x, y, z); rot_cols = the
    camera rotation's columns (right, up, look) as a (3, 3) array of rows;
    left_handed: glm's perspectiveLH, else its default perspectiveRH."""
    rot = _f32(np.asarray(rot_cols, np.float32).reshape(9))
    out = np.zeros(3, np.float32)
    _lib.or_screen_to_ray(_p(rot), _p(_f32(pos)), float(fovy), float(w), float(h), float(znear), float(zfar),
                          float(jx), float(jy), _p(_f32(p)), 1 if left_handed else 0, _p(out))
    return out
