"""The reference application's own scenes through the GPU path: the Cornell
box it renders by default (PathTracingApp.cpp:430-436: CornellBox-Original.obj
at scale 5, its camera preset, setPerspective(45)) and the Stanford bunny of
its assets, under directional lights.  Accumulators are bit-identical to the
oracle at equal seeds; the corner rays are the reference's screenToRay
restated (pt_camera_corners_basis, checked against the oracle's unprojection
in tests/test_oracle_camera.py).  The OBJ files are fixtures under
tests/golden/ (data from the reference's assets, CC BY 3.0 for the Cornell
box)."""
import os
import sys

import numpy as np
import pytest

import oracle_py as orc
import ptamd

pytestmark = pytest.mark.gpu

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def cornell_scene():
    tris, infos = ptamd.load_obj(os.path.join(GOLDEN, "CornellBox-Original.obj"))
    for k in ("v0", "v1", "v2"):
        tris[k] *= np.float32(5.0)  # SceneDesc scale 5 (PathTracingApp.cpp:433)
    return tris, infos


def preset_corners(w, h, seed, p):
    """The Cornell preset camera: position (-0.324, 4.428, -13), no rotation,
    glm::perspective(45, w/h) (45 radians, as the app passes it), the
    left-handed projection, one jitter per pass (PathTracingApp.cpp:217-242)."""
    jx, jy = ptamd.pass_jitter(seed, p, w, h)
    I = np.eye(3, dtype=np.float32)
    return ptamd.camera_corners_basis(I[:, 0], I[:, 1], I[:, 2], 45.0, w / h, jx, jy, True)


CORNELL_EYE = (-0.324, 4.428, -13.0)


@pytest.mark.parametrize("bounces,rr", [(1, ptamd.INVALID), (6, 3)])
def test_cornell_box_matches_oracle(ctx, bounces, rr):
    tris, infos = cornell_scene()
    # The reference's defaults: 1 bounce, indirect intensity 3, stratified
    # first bounce (PathTracingSettings.h:26-28); and a deeper variant.
    st = ptamd.Settings(bounces, rr, 3.0, 1)
    ctx.upload_scene(tris, infos, settings=st)
    w, h, n = 160, 90, 4
    seeds = np.arange(500, 500 + n, dtype=np.uint64)
    dirs = np.stack([preset_corners(w, h, 11, p) for p in range(n)])
    say(f"[cornell] {len(tris)} triangles, {w}x{h}, {n} passes, {bounces} bounces")
    ctx.reset_frames()
    ctx.set_primary_rays(dirs[0], CORNELL_EYE)
    ctx.render_passes(w, h, seeds, dirs)
    acc, spp = ctx.read_accumulator(w, h)
    b = orc.build(tris, *ptamd.scene_bounds(tris))
    o = np.zeros((h, w, 3))
    for p in range(n):
        orc.render_pass(tris, b, infos, st, w, h, dirs[p], CORNELL_EYE, p, int(seeds[p]), o)
    assert spp == n
    # The lamp is in view (its emission reaches the image directly) and light
    # bounces off the walls.  (Few pixels are lit after 4 passes: the first
    # bounce is stratified at stratum pass % 64 for all pixels at once, so
    # 4 passes cover 4 of the 64 strata.)
    assert o.max() >= 17.0 and (o.sum(axis=2) > 0).mean() > 0.005
    assert np.array_equal(acc, o)


def test_bunny_under_directional_lights_matches_oracle(ctx):
    from test_oracle_lights import light

    tris, infos = ptamd.load_obj(os.path.join(GOLDEN, "bunny.obj"))
    assert len(tris) == 4968
    infos["specularity"] = np.float32(16.0)
    lights = np.array([light((1.0, 0.95, 0.9), (0.4, -1.0, 0.3), 2.0),
                       light((0.3, 0.4, 1.0), (-0.6, -0.2, 1.0), 0.6)], np.float32)
    st = ptamd.Settings(3, 2, 1.0, 1)
    ctx.upload_scene(tris, infos, settings=st, lights=lights)
    bmin, bmax = ptamd.scene_bounds(tris)
    c = 0.5 * (bmin + bmax)
    eye = (float(c[0]), float(c[1]), float(c[2] - 3.0 * (bmax[1] - bmin[1])))
    w, h, n = 96, 96, 3
    seeds = np.arange(600, 600 + n, dtype=np.uint64)
    dirs = np.stack([ptamd.camera_corners(eye, tuple(c), (0, 1, 0), 40.0, 1.0, *ptamd.pass_jitter(4, p, w, h))
                     for p in range(n)])
    say(f"[bunny] {len(tris)} triangles, {w}x{h}, {n} passes, 2 lights")
    ctx.reset_frames()
    ctx.set_primary_rays(dirs[0], eye)
    ctx.render_passes(w, h, seeds, dirs)
    acc, spp = ctx.read_accumulator(w, h)
    b = orc.build(tris, bmin, bmax)
    o = np.zeros((h, w, 3))
    for p in range(n):
        orc.render_pass_lit(tris, b, infos, lights, st, w, h, dirs[p], eye, p, int(seeds[p]), o)
    assert spp == n and (o.sum(axis=2) > 0).mean() > 0.1
    assert np.array_equal(acc, o)


def test_cornell_box_bvh_and_trace_match_oracle(ctx):
    tris, infos = cornell_scene()
    ctx.upload_scene(tris, infos)
    bmin, bmax = ptamd.scene_bounds(tris)
    g = ctx.get_bvh()
    o = orc.build(tris, bmin, bmax)
    for k in g:
        assert np.array_equal(g[k].view(np.uint32), o[k].view(np.uint32)), k
    rays = ptamd.gen_shell_rays(50000, 8, bmin, bmax, 1.5)
    ids, t, uv = ctx.trace_closest(rays)
    oids, ot, ouv = orc.trace(tris, o, rays)
    assert np.array_equal(ids, oids) and np.array_equal(t.view(np.uint32), ot.view(np.uint32))
    assert (ids != ptamd.INVALID).sum() > 1000
