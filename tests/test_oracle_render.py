"""Pins the oracle's progressive bounce loop (CPU only).

What pins what:
  or_pcg4d          an independent Python transcription of the published hash
                    (Jarzynski & Olano 2020) with 32-bit wrap-around
  or_sincos2pi      numpy float64 sin/cos of 2*pi*u: |error| < 3e-7 on a dense
                    grid, exact at quarter turns
  or_to_hemisphere  local +z maps to the normal; the frame is orthonormal
  or_render_pass    0 bounces: emitter pixels equal the emission exactly, all
                    else zero; floor under a large emitting ceiling: the mean
                    equals the closed form a*E*E[cos] = (2/3) a E (1 bounce) and
                    (2/3) E (a + a^2) (3 bounces), also with Russian roulette
                    (unbiased); equal seeds give equal images
"""
import math

import numpy as np
import pytest

import oracle_py as orc
import ptamd

M32 = 0xFFFFFFFF


def py_pcg4d(v):
    v = [(x * 1664525 + 1013904223) & M32 for x in v]
    v[0] = (v[0] + v[1] * v[3]) & M32
    v[1] = (v[1] + v[2] * v[0]) & M32
    v[2] = (v[2] + v[0] * v[1]) & M32
    v[3] = (v[3] + v[1] * v[2]) & M32
    v = [x ^ (x >> 16) for x in v]
    v[0] = (v[0] + v[1] * v[3]) & M32
    v[1] = (v[1] + v[2] * v[0]) & M32
    v[2] = (v[2] + v[0] * v[1]) & M32
    v[3] = (v[3] + v[1] * v[2]) & M32
    return v


def test_pcg4d_matches_published_algorithm():
    rng = np.random.default_rng(0)
    cases = [(0, 0, 0, 0), (1, 2, 3, 4), (M32, M32, M32, M32)]
    cases += [tuple(int(x) for x in rng.integers(0, 2**32, 4)) for _ in range(200)]
    for c in cases:
        assert list(orc.pcg4d(c)) == py_pcg4d(list(c))


def test_sincos2pi_accuracy():
    us = np.concatenate([np.arange(0, 1 << 16) / float(1 << 16),
                         np.random.default_rng(1).random(20000)]).astype(np.float32)
    err = 0.0
    for u in us[::7]:
        s, c = orc.sincos2pi(u)
        a = 2.0 * math.pi * float(u)
        err = max(err, abs(s - math.sin(a)), abs(c - math.cos(a)))
    assert err < 3e-7
    assert orc.sincos2pi(0.0) == (0.0, 1.0)
    assert orc.sincos2pi(0.25) == (1.0, 0.0)
    assert orc.sincos2pi(0.5) == (0.0, -1.0)
    assert orc.sincos2pi(0.75) == (-1.0, 0.0)


def test_to_hemisphere_frame():
    rng = np.random.default_rng(2)
    for _ in range(300):
        n = rng.normal(size=3)
        n /= np.linalg.norm(n)
        if rng.random() < 0.2:
            n = np.zeros(3)
            n[rng.integers(0, 3)] = 1.0 if rng.random() < 0.5 else -1.0
        n = n.astype(np.float32)
        assert np.allclose(orc.to_hemisphere((0, 0, 1), n), n, atol=1e-6)
        ex = orc.to_hemisphere((1, 0, 0), n)
        ey = orc.to_hemisphere((0, 1, 0), n)
        assert abs(np.dot(ex, n)) < 1e-6 and abs(np.dot(ey, n)) < 1e-6 and abs(np.dot(ex, ey)) < 1e-6
        assert abs(np.linalg.norm(ex) - 1) < 1e-6 and abs(np.linalg.norm(ey) - 1) < 1e-6


def quad(corners, first_id, albedo, emission):
    """Two triangles (a, b, c), (a, c, d) with flat shading records."""
    a, b, c, d = (np.asarray(p, np.float32) for p in corners)
    t = np.zeros(2, ptamd.TRI_DTYPE)
    info = np.zeros(2, ptamd.INFO_DTYPE)
    for k, (p, q, r) in enumerate(((a, b, c), (a, c, d))):
        t[k]["v0"], t[k]["v1"], t[k]["v2"], t[k]["id"] = p, q, r, first_id + k
        n = np.cross(q - p, r - p)
        n = (n / np.linalg.norm(n)).astype(np.float32)
        info[k]["n0"] = info[k]["n1"] = info[k]["n2"] = n
        info[k]["albedo"] = albedo
        info[k]["emission"] = emission
    return t, info


def floor_ceiling(a, E, S=1000.0):
    # floor y = 0 facing +y, ceiling y = 1 facing -y (emitting), both 2S wide
    f, fi = quad([(-S, 0, -S), (-S, 0, S), (S, 0, S), (S, 0, -S)], 0, (a, a, a), (0, 0, 0))
    c, ci = quad([(-S, 1, -S), (S, 1, -S), (S, 1, S), (-S, 1, S)], 2, (0.7, 0.7, 0.7), (E, E, E))
    return np.concatenate([f, c]), np.concatenate([fi, ci])


def render(tris, infos, settings, w, h, eye, target, up, passes, seed=7, fov=60.0):
    bmin, bmax = ptamd.scene_bounds(tris)
    b = orc.build(tris, bmin, bmax)
    acc = np.zeros((h, w, 3), np.float64)
    for p in range(passes):
        dirs = ptamd.camera_corners(eye, target, up, fov, w / h)
        orc.render_pass(tris, b, infos, settings, w, h, dirs, eye, p, seed, acc)
    return acc / passes


def test_zero_bounces_emission_only():
    tris, infos = ptamd.gen_soup(2000, 4)  # soup + emitter quad at y = 1.2
    st = ptamd.Settings(0, ptamd.INVALID, 1.0, 1)
    img = render(tris, infos, st, 48, 48, (0.5, -0.5, 0.5), (0.5, 1.2, 0.5), (0, 0, 1), 1)
    vals = np.unique(img.reshape(-1, 3), axis=0)
    assert all(tuple(v) in ((0.0, 0.0, 0.0), (10.0, 10.0, 10.0)) for v in vals)
    assert (img[..., 0] == 10.0).sum() > 0


@pytest.mark.parametrize("bounces,rr", [(1, ptamd.INVALID), (3, ptamd.INVALID), (3, 1)])
def test_floor_under_emitter_closed_form(bounces, rr):
    a, E = 0.5, 1.0
    tris, infos = floor_ceiling(a, E)
    st = ptamd.Settings(bounces, rr, 1.0, 1)
    img = render(tris, infos, st, 64, 64, (0, 0.5, 0), (0, 0, 0), (0, 0, 1), 8)
    expect = (2.0 / 3.0) * E * (a if bounces == 1 else a + a * a)
    mean = img[..., 0].mean()
    # per-sample std <= 0.5 E for these estimators; 32768 samples
    assert abs(mean - expect) < 4 * 0.5 / math.sqrt(64 * 64 * 8), (mean, expect)
    assert np.array_equal(img[..., 0], img[..., 1]) and np.array_equal(img[..., 0], img[..., 2])


def test_seeds_are_deterministic():
    tris, infos = ptamd.gen_soup(30000, 5)
    st = ptamd.Settings(3, ptamd.INVALID, 1.0, 1)
    eye, tgt = (0.5, 0.5, -1.5), (0.5, 0.5, 0.0)
    a = render(tris, infos, st, 32, 32, eye, tgt, (0, 1, 0), 2, seed=3)
    b = render(tris, infos, st, 32, 32, eye, tgt, (0, 1, 0), 2, seed=3)
    c = render(tris, infos, st, 32, 32, eye, tgt, (0, 1, 0), 2, seed=4)
    assert a.sum() > 0
    assert np.array_equal(a, b)
    assert not np.array_equal(a, c)


def test_render_stats_equal_trace_counts():
    """A 0-bounce pass traces exactly the primary rays: its statistics equal
    w * h segments and the closest-hit traversal's own counters over the
    same rays (built here from the corner directions as the pass builds
    them: pixel corner (x/w, y/h), glm::mix x + (y - x) a in fp32)."""
    tris, infos = ptamd.gen_soup(4096, 11)
    b = orc.build(tris, *ptamd.scene_bounds(tris))
    w, h = 24, 16
    eye = np.float32([0.5, 0.5, -1.5])
    dirs = ptamd.camera_corners(tuple(eye), (0.5, 0.5, 0.0), (0, 1, 0), 60.0, w / h).reshape(4, 3)
    st = ptamd.Settings(0, ptamd.INVALID, 1.0, 1)
    _, sts = orc.render_pass_stats(tris, b, infos, st, w, h, dirs, eye, 0, 3)
    mix = lambda x, y, a: x + (y - x) * a
    y, x = np.mgrid[0:h, 0:w]
    tx = (x.astype(np.float32) / np.float32(w))[..., None]
    ty = (y.astype(np.float32) / np.float32(h))[..., None]
    d = mix(mix(dirs[0], dirs[1], tx), mix(dirs[3], dirs[2], tx), ty).astype(np.float32).reshape(-1, 3)
    rays = np.concatenate([np.broadcast_to(eye, d.shape), d], axis=1).astype(np.float32)
    ids, _, _, c = orc.trace(tris, b, rays, counters=True)
    assert sts["segments"] == w * h
    assert (sts["visits"], sts["boxes"], sts["tris"]) == tuple(int(v) for v in c[:3])
    assert sts["hits"] == int((ids != ptamd.INVALID).sum()) > 0
