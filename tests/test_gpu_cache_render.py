"""Cache frames on the GPU against the oracle's cache frame.

  training batch  bit-identical to the oracle (positions, normals, incoming
                  directions, albedos and suffix-radiance targets; NaN where
                  the path never reached x2): the path arithmetic is the plain
                  render's, which is bit-exact
  image           mean relative per-channel error <= 1e-3 against the oracle
                  with the same fp16-rounded weights (the cache's fp16
                  inference is the only difference)
  training        the in-frame Adam steps equal cache_train over the same
                  batch chunks, bit for bit
  graphs          replayed cache frames (accumulator and weights) equal eager
                  frames bit for bit
  convergence     on a closed room, cache frames reach a lower RMSE against a
                  high-spp plain render than plain 1-spp frames at equal frame
                  counts
"""
import sys

import numpy as np
import pytest

import oracle_py as orc
import ptamd
from test_gpu_cache import fp16_weights
from test_oracle_mlp import random_params

pytestmark = pytest.mark.gpu

W, H = 64, 48
EYE, TGT = (0.0, 0.0, -1.9), (0.0, 0.0, 0.0)


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def dirs_for(w, h):
    return ptamd.camera_corners(EYE, TGT, (0, 1, 0), 60.0, w / h)


@pytest.fixture(scope="module")
def room():
    tris, infos = ptamd.gen_sphere(48, 3, room=True)
    return tris, infos


def setup(ctx, room, st, w=W, h=H):
    tris, infos = room
    ctx.upload_scene(tris, infos, settings=st)
    ctx.set_primary_rays(dirs_for(w, h), EYE)
    ctx.reset_frames()
    bmin, bmax = ptamd.scene_bounds(tris)
    return bmin, bmax


def rel_err(g, o):
    den = np.maximum(np.abs(o).mean(axis=(0, 1)), 1e-12)
    return np.abs(g - o).mean(axis=(0, 1)) / den


@pytest.mark.parametrize("rr_start,strat,paths", [(3, 1, 1), (0, 0, 1), (3, 1, 0)])
def test_cache_frame_matches_oracle(ctx, room, rr_start, strat, paths):
    """Cache frames against the oracle, with the pixel paths traced
    path-persistently (render.paths 1) or as two levels (0)."""
    ptamd.tune_set("render.paths", paths)
    try:
        check_cache_frames(ctx, room, rr_start, strat)
    finally:
        ptamd.tune_set("render.paths", 1)


def check_cache_frames(ctx, room, rr_start, strat):
    st = ptamd.Settings(8, rr_start, 1.0, strat)
    bmin, bmax = setup(ctx, room, st)
    ctx.cache_init(5, bmin, bmax)
    p = random_params(21, bias=0.2)
    ctx.cache_set_params(p)
    ntrain, extra = 700, 6
    cs = ptamd.CacheSettings(ntrain, 1, extra, train=False)
    tris, infos = room
    bvh = orc.build(tris, bmin, bmax)
    acc_o = np.zeros((H, W, 3))
    for f in range(3):
        seed = 1000 + f
        say(f"[cache frame] rr_start={rr_start} frame {f}")
        ctx.render_cache_frame(W, H, seed, cs)
        raw, tgt = ctx.cache_last_batch(ntrain)
        qraw, qpred, qvalid = ctx.cache_last_queries(W, H)
        assert qvalid.any()
        np.testing.assert_allclose(qpred[qvalid], ctx.cache_infer(qraw[qvalid]), rtol=0, atol=0)
        acc_o, oraw, otgt = orc.render_cache_pass(tris, bvh, infos, st, W, H, dirs_for(W, H), EYE, f, seed,
                                                  fp16_weights(p), bmin, bmax, ntrain, extra, acc_o)
        valid = ~np.isnan(otgt[:, 0])
        assert valid.sum() > ntrain // 2
        np.testing.assert_array_equal(np.isnan(tgt[:, 0]), ~valid)
        np.testing.assert_array_equal(raw[valid], oraw[valid])
        np.testing.assert_array_equal(tgt[valid], otgt[valid])
    acc_g, spp = ctx.read_accumulator(W, H)
    assert spp == 3
    err = rel_err(acc_g, acc_o)
    say(f"  image mean relative per-channel error {err}")
    assert acc_o.sum() > 0 and max(err) <= 1e-3


def test_cache_frame_ray_count(ctx, room):
    """pt_last_frame_rays after a cache frame: without training paths it is
    the plain frame's count with one bounce at the same seed (a pixel path
    follows the plain path to x2); the training paths add between one and
    train_levels segments each (at least two for a path that reached x2);
    graph replays and eager frames count the same."""
    st = ptamd.Settings(8, 3, 1.0, 1)
    bmin, bmax = setup(ctx, room, st)
    ctx.cache_init(3, bmin, bmax)
    seed = 4242
    per_structure = []
    for paths in (1, 0):
        ptamd.tune_set("render.paths", paths)
        try:
            ctx.reset_frames()
            ctx.render_cache_frame(W, H, seed, ptamd.CacheSettings(0, 1, 8, train=False))
            per_structure.append(ctx.last_frame_rays())
        finally:
            ptamd.tune_set("render.paths", 1)
    assert per_structure[0] == per_structure[1]  # path-persistent pixel paths count as the levels do
    pixel_rays = per_structure[0]
    ctx.upload_scene(*room, settings=ptamd.Settings(1, 3, 1.0, 1))
    ctx.set_primary_rays(dirs_for(W, H), EYE)
    ctx.reset_frames()
    ctx.render_pass(W, H, seed)
    assert ctx.last_frame_rays() == pixel_rays
    ntrain, extra = 900, 8
    counts = {}
    for graphs in (1, 0):
        ptamd.tune_set("render.graphs", graphs)
        try:
            setup(ctx, room, st)
            ctx.cache_init(3, bmin, bmax)
            ctx.render_cache_frame(W, H, seed, ptamd.CacheSettings(ntrain, 1, extra, train=False))
            counts[graphs] = ctx.last_frame_rays()
        finally:
            ptamd.tune_set("render.graphs", 1)
    assert counts[0] == counts[1]
    _, tgt = ctx.cache_last_batch(ntrain)
    reached = int((~np.isnan(tgt[:, 0])).sum())
    train_rays = counts[1] - pixel_rays
    assert ntrain + reached <= train_rays <= ntrain * (2 + extra)


def test_in_frame_training_equals_cache_train(ctx, room):
    st = ptamd.Settings(8, 3, 1.0, 1)
    bmin, bmax = setup(ctx, room, st)
    ctx.cache_init(9, bmin, bmax)
    p0 = ctx.cache_get_params()
    ntrain, steps = 1024, 4
    ctx.render_cache_frame(W, H, 77, ptamd.CacheSettings(ntrain, steps, 8, train=True))
    p1 = ctx.cache_get_params()
    raw, tgt = ctx.cache_last_batch(ntrain)
    assert not np.array_equal(p0, p1)
    ctx.cache_set_params(p0)
    for k in range(steps):
        sl = slice(k * ntrain // steps, (k + 1) * ntrain // steps)
        ctx.cache_train(raw[sl], tgt[sl])
    np.testing.assert_array_equal(ctx.cache_get_params(), p1)


def test_cache_frames_graph_equals_eager(ctx, room):
    st = ptamd.Settings(8, 3, 1.0, 1)
    out = {}
    for graphs, paths in ((1, 1), (0, 1), (1, 0)):  # replays, eager frames, levels instead of paths
        ptamd.tune_set("render.graphs", graphs)
        ptamd.tune_set("render.paths", paths)
        try:
            bmin, bmax = setup(ctx, room, st)
            ctx.cache_init(13, bmin, bmax)
            for f in range(5):
                ctx.render_cache_frame(W, H, 500 + f, ptamd.CacheSettings(512, 2, 8, train=True))
            out[(graphs, paths)] = (ctx.read_accumulator(W, H)[0], ctx.cache_get_params())
        finally:
            ptamd.tune_set("render.graphs", 1)
            ptamd.tune_set("render.paths", 1)
    for key in ((0, 1), (1, 0)):
        np.testing.assert_array_equal(out[key][0], out[(1, 1)][0])
        np.testing.assert_array_equal(out[key][1], out[(1, 1)][1])


def test_cache_frame_errors(ctx, room):
    st = ptamd.Settings(8, 3, 1.0, 1)
    setup(ctx, room, st)
    with pytest.raises(ptamd.PtError):
        ctx.render_cache_frame(W, H, 1, ptamd.CacheSettings(W * H + 1, 1, 8))
    with pytest.raises(ptamd.PtError):
        ctx.render_cache_frame(W, H, 1, ptamd.CacheSettings(64, 0, 8))
    with pytest.raises(ptamd.PtError):
        ctx.render_cache_frame(W, H, 1, ptamd.CacheSettings(64, 1, 59))


def test_cache_lowers_rmse_at_equal_frames(ctx, room):
    w, h = 96, 72
    extra = 7
    st = ptamd.Settings(1 + extra, 3, 1.0, 1)
    bmin, bmax = setup(ctx, room, st, w, h)
    say("[cache rmse] plain reference 2048 spp")
    seeds = np.arange(1, 1 + 512, dtype=np.uint64)
    for k in range(4):
        ctx.render_passes(w, h, seeds + 512 * k)
    ref = ctx.read_accumulator(w, h)[0] / (4 * 512)
    frames = 64
    ctx.reset_frames()
    ctx.render_passes(w, h, np.arange(10 ** 6, 10 ** 6 + frames, dtype=np.uint64))
    plain = ctx.read_accumulator(w, h)[0] / frames
    ctx.reset_frames()
    ctx.cache_init(3, bmin, bmax)
    for f in range(frames):
        ctx.render_cache_frame(w, h, 2 * 10 ** 6 + f, ptamd.CacheSettings(2048, 4, extra, train=True))
    cached = ctx.read_accumulator(w, h)[0] / frames
    r_plain = float(np.sqrt(np.mean((plain - ref) ** 2)))
    r_cache = float(np.sqrt(np.mean((cached - ref) ** 2)))
    say(f"  RMSE after {frames} frames: plain {r_plain:.5f}, cache {r_cache:.5f}")
    assert np.isfinite(r_cache) and r_cache < r_plain


def test_cache_frame_with_lights_matches_oracle(ctx):
    from test_gpu_render import lit_scene
    tris, infos, lights = lit_scene()
    st = ptamd.Settings(6, 3, 1.0, 1)
    eye, tgt = (0.5, 0.5, -1.5), (0.5, 0.5, 0.0)
    ctx.upload_scene(tris, infos, settings=st, lights=lights)
    dirs = ptamd.camera_corners(eye, tgt, (0, 1, 0), 60.0, W / H)
    ctx.set_primary_rays(dirs, eye)
    ctx.reset_frames()
    bmin, bmax = ptamd.scene_bounds(tris)
    ctx.cache_init(5, bmin, bmax)
    p = random_params(23, bias=0.2)
    ctx.cache_set_params(p)
    ntrain = W * H  # the soup is open: few training paths reach x2
    say("[cache frame] soup with directional lights")
    ctx.render_cache_frame(W, H, 4242, ptamd.CacheSettings(ntrain, 1, 5, train=False))
    raw, tgt_g = ctx.cache_last_batch(ntrain)
    bvh = orc.build(tris, bmin, bmax)
    acc_o, oraw, otgt = orc.render_cache_pass(tris, bvh, infos, st, W, H, dirs, eye, 0, 4242, fp16_weights(p),
                                              bmin, bmax, ntrain, 5, lights=lights)
    valid = ~np.isnan(otgt[:, 0])
    assert valid.sum() > 60
    np.testing.assert_array_equal(np.isnan(tgt_g[:, 0]), ~valid)
    np.testing.assert_array_equal(raw[valid], oraw[valid])
    np.testing.assert_array_equal(tgt_g[valid], otgt[valid])
    acc_g, _ = ctx.read_accumulator(W, H)
    assert max(rel_err(acc_g, acc_o)) <= 1e-3



def test_cache_frame_with_detail_matches_oracle(ctx, room):
    """Surface detail in cache frames: the training batch (whose albedo and
    normal come from textures, vertex colours and normal maps) is bit-identical
    to the oracle's, the image within the fp16 tolerance."""
    from test_gpu_render import random_detail
    st = ptamd.Settings(8, 3, 1.0, 1)
    bmin, bmax = setup(ctx, room, st)
    tris, infos = room
    det, texs = random_detail(len(infos), 9)
    ctx.upload_detail(det, texs)
    ctx.cache_init(5, bmin, bmax)
    p = random_params(29, bias=0.2)
    ctx.cache_set_params(p)
    ntrain, extra = 900, 5
    say("[cache frame] room with surface detail")
    ctx.render_cache_frame(W, H, 777, ptamd.CacheSettings(ntrain, 1, extra, train=False))
    raw, tgt = ctx.cache_last_batch(ntrain)
    bvh = orc.build(tris, bmin, bmax)
    acc_o, oraw, otgt = orc.render_cache_pass(tris, bvh, infos, st, W, H, dirs_for(W, H), EYE, 0, 777,
                                              fp16_weights(p), bmin, bmax, ntrain, extra, detail=det,
                                              textures=texs)
    valid = ~np.isnan(otgt[:, 0])
    assert valid.sum() > ntrain // 2
    np.testing.assert_array_equal(np.isnan(tgt[:, 0]), ~valid)
    np.testing.assert_array_equal(raw[valid], oraw[valid])
    np.testing.assert_array_equal(tgt[valid], otgt[valid])
    acc_g, _ = ctx.read_accumulator(W, H)
    assert max(rel_err(acc_g, acc_o)) <= 1e-3
    ctx.upload_detail(np.zeros(0, ptamd.DETAIL_DTYPE))


def test_cache_frame_full_room_mesh(ctx):
    """A cache frame on config 4's scene (the 871,200-triangle mesh in the
    room) at 96x54 with 2,048 training paths of 8 extra bounces: training
    batch bit-identical to the oracle's, image within 1e-3 of it."""
    tris, infos = ptamd.gen_sphere(660, 2024, room=True)
    st = ptamd.Settings(9, 3, 1.0, 1)
    w, h = 96, 54
    ctx.upload_scene(tris, infos, settings=st)
    d = ptamd.camera_corners(EYE, TGT, (0, 1, 0), 60.0, w / h)
    ctx.set_primary_rays(d, EYE)
    ctx.reset_frames()
    bmin, bmax = ptamd.scene_bounds(tris)
    ctx.cache_init(5, bmin, bmax)
    p = random_params(23, bias=0.2)
    ctx.cache_set_params(p)
    ntrain, extra = 2048, 8
    cs = ptamd.CacheSettings(ntrain, 1, extra, train=False)
    bvh = orc.build(tris, bmin, bmax)
    acc_o = np.zeros((h, w, 3))
    for f in range(2):
        seed = 4000 + f
        say(f"[cache frame] full room mesh {w}x{h}, frame {f}")
        ctx.render_cache_frame(w, h, seed, cs)
        raw, tgt = ctx.cache_last_batch(ntrain)
        acc_o, oraw, otgt = orc.render_cache_pass(tris, bvh, infos, st, w, h, d, EYE, f, seed,
                                                  fp16_weights(p), bmin, bmax, ntrain, extra, acc_o)
        valid = ~np.isnan(otgt[:, 0])
        assert valid.sum() > ntrain // 2
        np.testing.assert_array_equal(np.isnan(tgt[:, 0]), ~valid)
        np.testing.assert_array_equal(raw[valid], oraw[valid])
        np.testing.assert_array_equal(tgt[valid], otgt[valid])
    acc_g, spp = ctx.read_accumulator(w, h)
    assert spp == 2
    err = rel_err(acc_g, acc_o)
    say(f"  image mean relative per-channel error {err}")
    assert acc_o.sum() > 0 and max(err) <= 1e-3
