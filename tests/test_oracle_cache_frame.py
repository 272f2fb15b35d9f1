"""Pins the oracle's cache frame (CPU only).

  pixel paths     with a cache that returns 0 the frame equals or_render_pass
                  with one bounce bit for bit (same random stream, same
                  stratification, also with Russian roulette from bounce 0);
                  with a constant cache c every pixel whose path reached x2
                  adds throughput * c
  training paths  floor under an emitting ceiling: x2 lies on the ceiling
                  (emitters reflect with albedo 1), so the suffix target has
                  the closed-form mean (2/3) E a for 2-3 extra bounces and
                  (2/3) E (a + a^2) for 4-5; the recorded features are the
                  ceiling's position, normal and albedo and an upward
                  incoming direction; targets of paths that never reach x2
                  are NaN
"""
import math

import numpy as np
import pytest

import oracle_py as orc
import ptamd
from test_oracle_render import floor_ceiling

NP = orc.mlp_param_count()


def const_cache(c):
    p = np.zeros(NP, np.float32)
    p[-3:] = np.log(c) if c > 0 else -200.0
    return p


def frame(tris, infos, st, w, h, eye, tgt, up, pass_index, seed, params, ntrain, extra):
    bmin, bmax = ptamd.scene_bounds(tris)
    b = orc.build(tris, bmin, bmax)
    dirs = ptamd.camera_corners(eye, tgt, up, 60.0, w / h)
    acc, raw, t = orc.render_cache_pass(tris, b, infos, st, w, h, dirs, eye, pass_index, seed, params,
                                        bmin, bmax, ntrain, extra)
    plain = orc.render_pass(tris, b, infos, ptamd.Settings(1, st.rr_start, st.indirect_intensity,
                                                            st.stratified), w, h, dirs, eye, pass_index,
                            seed, np.zeros((h, w, 3)))[0]
    return acc, raw, t, plain


@pytest.mark.parametrize("rr,strat", [(3, 1), (0, 1), (0, 0)])
def test_zero_cache_equals_one_bounce_render(rr, strat):
    tris, infos = ptamd.gen_sphere(24, 1, room=True)
    st = ptamd.Settings(8, rr, 1.0, strat)
    total = 0.0
    for p in range(6):
        acc, _, _, plain = frame(tris, infos, st, 64, 48, (0, 0, -1.9), (0, 0, 0), (0, 1, 0), p, 55 + p,
                                 const_cache(0.0), 64, 4)
        total += plain.sum()
        np.testing.assert_array_equal(acc, plain)
    assert total > 0  # some one-bounce paths reach the ceiling emitter


def test_constant_cache_adds_throughput():
    a, E, c = 0.5, 1.0, 0.25
    tris, infos = floor_ceiling(a, E)
    st = ptamd.Settings(8, ptamd.INVALID, 1.0, 1)
    acc, _, _, plain = frame(tris, infos, st, 32, 32, (0, 0.5, 0), (0, 0, 0), (0, 0, 1), 0, 9,
                             const_cache(c), 16, 2)
    # x1 on the floor (albedo a), x2 on the ceiling (albedo 1): throughput a
    want = (plain.astype(np.float32) + np.float32(a) * np.float32(c)).astype(np.float64)
    np.testing.assert_allclose(acc, want, rtol=0, atol=1e-7)


@pytest.mark.parametrize("extra,terms", [(2, 1), (3, 1), (4, 2), (5, 2)])
def test_training_targets_closed_form(extra, terms):
    a, E = 0.5, 1.0
    tris, infos = floor_ceiling(a, E)
    st = ptamd.Settings(8, ptamd.INVALID, 1.0, 1)
    ntrain = 16384
    _, raw, t, _ = frame(tris, infos, st, 128, 128, (0, 0.5, 0), (0, 0, 0), (0, 0, 1), 5, 21,
                         const_cache(1.0), ntrain, extra)
    ok = ~np.isnan(t[:, 0])
    assert ok.sum() >= ntrain - 16  # a rare grazing bounce leaves the 2000-wide quads
    t, raw = t[ok], raw[ok]
    ntrain = len(t)
    want = (2.0 / 3.0) * E * sum(a ** k for k in range(1, terms + 1))
    # per-sample std <= E (a + a^2) here
    assert abs(t[:, 0].mean() - want) < 4 * E * (a + a * a) / math.sqrt(ntrain), (t[:, 0].mean(), want)
    assert np.array_equal(t[:, 0], t[:, 1]) and np.array_equal(t[:, 0], t[:, 2])
    np.testing.assert_allclose(raw[:, 1], 1.0, atol=1e-5)           # ceiling y = 1
    np.testing.assert_array_equal(raw[:, 3:6], np.tile([0.0, -1.0, 0.0], (ntrain, 1)).astype(np.float32))
    assert np.all(raw[:, 7] > 0)                                    # arriving upward
    np.testing.assert_array_equal(raw[:, 9:], np.ones((ntrain, 3), np.float32))


def test_training_paths_that_escape_are_masked():
    # A single floor quad: primary rays hit it, bounces escape -> no x2.
    a = 0.5
    tris, infos = floor_ceiling(a, 1.0)
    tris, infos = tris[:2].copy(), infos[:2].copy()
    st = ptamd.Settings(8, ptamd.INVALID, 1.0, 1)
    _, _, t, _ = frame(tris, infos, st, 16, 16, (0, 0.5, 0), (0, 0, 0), (0, 0, 1), 0, 3,
                       const_cache(1.0), 64, 3)
    assert np.isnan(t).all()
