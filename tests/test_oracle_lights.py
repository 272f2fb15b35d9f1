"""Pins the oracle's directional lights (row a10; CPU only).

  or_pow_pos         float64 power within (2e-6 + 1.5e-7 |y log2 x|) relative
                     on a grid (float rounding of y log2 x grows with it); exact
                     for dyadic bases with integer exponents; 0 at x = 0
  primary shading    light straight down on a floor: a * c * I exactly; a
                     floor point under an occluder is black and the rest lit
                     (shadow rays); two identical lights give 3 a c I, the
                     reference's running light sum L (cuda_path_tracer.cu:
                     201-214) -- a plain sum over lights would give 2 a c I;
                     a specular floor matches the Blinn-Phong formula with the
                     unnormalised halfway vector evaluated in float64
  bounce light term  a floor that the light grazes (cosine 0) next to a lit
                     wall: the mean over pixels is (1/2) a b c I (half of the
                     cosine-weighted bounce directions reach the wall, whose
                     light term has cosine 1)
"""
import math

import numpy as np
import pytest

import oracle_py as orc
import ptamd
from test_oracle_render import quad


@pytest.mark.parametrize("x", [1e-6, 1e-3, 0.01, 0.3, 0.7071, 0.99, 1.0, 1.7, 3.9, 100.0, 1e5])
@pytest.mark.parametrize("y", [0.5, 1.0, 3.0, 10.0, 32.0, 128.0])
def test_pow_pos_accuracy(x, y):
    got = orc.pow_pos(x, y)
    lw = float(np.float32(y)) * math.log2(float(np.float32(x)))
    want = 2.0 ** lw if lw < 1000 else math.inf
    if want < 2.0 ** -125:
        assert got == 0.0 or abs(got - want) <= 2.0 ** -125
    elif want > 3e38:
        assert got > 1e38
    else:
        tol = 2e-6 + 1.5e-7 * abs(float(np.float32(y)) * math.log2(float(np.float32(x))))
        assert abs(got - want) <= tol * want + 1e-45, (x, y, got, want)


def test_pow_pos_exact_cases():
    assert orc.pow_pos(0.0, 5.0) == 0.0
    assert orc.pow_pos(1.0, 77.0) == 1.0
    for k in range(-10, 11):
        for n in (1, 2, 3, 7):
            assert orc.pow_pos(2.0 ** k, float(n)) == 2.0 ** (k * n)


def floor(a, spec=0.0, S=50.0, first_id=0):
    t, i = quad([(-S, 0, -S), (-S, 0, S), (S, 0, S), (S, 0, -S)], first_id, (a, a, a), (0, 0, 0))
    i["specularity"] = spec
    return t, i


def light(color, direction, intensity):
    d = np.asarray(direction, np.float64)
    d = d / np.linalg.norm(d)
    return [*color, *d.astype(np.float32), intensity]


def render_lit(tris, infos, lights, bounces, w, h, eye, target, up, passes=1, seed=3, strat=1):
    bmin, bmax = ptamd.scene_bounds(tris)
    b = orc.build(tris, bmin, bmax)
    acc = np.zeros((h, w, 3))
    st = ptamd.Settings(bounces, ptamd.INVALID, 1.0, strat)
    dirs = ptamd.camera_corners(eye, target, up, 60.0, w / h)
    for p in range(passes):
        orc.render_pass_lit(tris, b, infos, lights, st, w, h, dirs, eye, p, seed, acc)
    return acc / passes, dirs


def test_light_from_above_on_floor():
    a, c, I = 0.6, (0.5, 1.0, 0.25), 2.0
    t, i = floor(a)
    img, _ = render_lit(t, i, [light(c, (0, -1, 0), I)], 0, 16, 16, (0, 0.5, 0), (0, 0, 0), (0, 0, 1))
    want = (np.float32(c) * np.float32(I)) * np.float32(a)
    assert np.array_equal(img.astype(np.float32), np.broadcast_to(want, img.shape))


def test_two_lights_running_sum():
    a, c, I = 0.5, (1.0, 1.0, 1.0), 1.0
    t, i = floor(a)
    L = light(c, (0, -1, 0), I)
    img, _ = render_lit(t, i, [L, L], 0, 8, 8, (0, 0.5, 0), (0, 0, 0), (0, 0, 1))
    np.testing.assert_allclose(img, 3 * a * np.ones_like(img), rtol=1e-6)


def test_occluder_shadow():
    a = 0.5
    ft, fi = floor(a)
    # occluder above the camera (y = 0.6 > eye y = 0.3): only shadow rays meet it
    ot, oi = quad([(-0.1, 0.6, -0.1), (0.1, 0.6, -0.1), (0.1, 0.6, 0.1), (-0.1, 0.6, 0.1)], 2, (1, 1, 1), (0, 0, 0))
    tris, infos = np.concatenate([ft, ot]), np.concatenate([fi, oi])
    w = h = 32
    eye = (0.0, 0.3, 0.0)
    img, dirs = render_lit(tris, infos, [light((1, 1, 1), (0, -1, 0), 1.0)], 0, w, h, eye, (0, 0, 0), (0, 0, 1))
    d = np.asarray(dirs, np.float64).reshape(4, 3)
    lit = shadowed = 0
    for y in range(h):
        for x in range(w):
            tx, ty = x / w, y / h
            r = (d[0] + tx * (d[1] - d[0])) * (1 - ty) + (d[3] + tx * (d[2] - d[3])) * ty
            p = np.asarray(eye) + r * (-eye[1] / r[1])  # floor hit point
            inside = abs(p[0]) < 0.1 - 1e-4 and abs(p[2]) < 0.1 - 1e-4
            outside = abs(p[0]) > 0.1 + 1e-4 or abs(p[2]) > 0.1 + 1e-4
            if inside:
                assert img[y, x, 0] == 0.0
                shadowed += 1
            elif outside:
                assert img[y, x, 0] == pytest.approx(a, rel=1e-6)
                lit += 1
    assert shadowed > 20 and lit > 200


def test_specular_floor_matches_formula():
    a, s, I = 0.4, 16.0, 1.5
    t, i = floor(a, spec=s)
    ldir = np.array([0.3, -1.0, 0.2])
    ldir /= np.linalg.norm(ldir)
    L = light((1, 1, 1), ldir, I)
    w, h = 24, 24
    eye = np.array([0.05, 0.5, -0.1])
    img, dirs = render_lit(t, i, [L], 0, w, h, tuple(eye), (0, 0, 0), (0, 0, 1))
    d = np.asarray(dirs, np.float64).reshape(4, 3)
    lf = np.asarray(L[3:6], np.float64)
    n = np.array([0.0, 1.0, 0.0])
    for y in range(h):
        for x in range(w):
            tx, ty = x / w, y / h
            r = (d[0] + tx * (d[1] - d[0])) * (1 - ty) + (d[3] + tx * (d[2] - d[3])) * ty
            hw = 0.5 * (-r - lf)  # unnormalised halfway of the unnormalised view
            df = max(n @ -lf, 0.0)
            want = I * a * df + I * 0.5 * max(n @ hw, 0.0) ** s
            assert img[y, x, 0] == pytest.approx(want, rel=2e-5), (x, y)


def test_bounce_light_term_closed_form():
    a, b, c, I = 0.5, 0.8, 1.0, 2.0
    ft, fi = floor(a, S=1000.0)
    # wall x = 1 facing -x, 2000 high; lit by a light travelling along +x
    wt, wi = quad([(1, 0, -1000), (1, 2000, -1000), (1, 2000, 1000), (1, 0, 1000)], 2, (b, b, b), (0, 0, 0))
    if np.cross(wt[0]["v1"] - wt[0]["v0"], wt[0]["v2"] - wt[0]["v0"])[0] > 0:
        wt, wi = quad([(1, 0, -1000), (1, 0, 1000), (1, 2000, 1000), (1, 2000, -1000)], 2, (b, b, b), (0, 0, 0))
    tris, infos = np.concatenate([ft, wt]), np.concatenate([fi, wi])
    assert infos[2]["n0"][0] < 0
    # unstratified: the first-bounce stratum is shared by every pixel of a pass
    img, _ = render_lit(tris, infos, [light((c, c, c), (1, 0, 0), I)], 1, 64, 64, (0, 0.5, 0), (0, 0, 0), (0, 0, 1),
                        passes=8, strat=0)
    want = 0.5 * a * b * c * I
    # per-sample values are 0 or a b c I
    assert abs(img[..., 0].mean() - want) < 4 * a * b * c * I * 0.5 / math.sqrt(64 * 64 * 8)
