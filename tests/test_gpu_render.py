"""GPU render paths: multi-pass frames, hipGraph replays and launch-shape
knobs give accumulators bit-identical to sequential 1-spp passes and to the
oracle."""
import sys

import numpy as np
import pytest

import oracle_py as orc
import ptamd

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


W, H = 64, 48
EYE, TGT = (0.5, 0.5, -1.5), (0.5, 0.5, 0.0)


def pass_dirs(p, seed=21):
    jx, jy = ptamd.pass_jitter(seed, p, W, H)
    return ptamd.camera_corners(EYE, TGT, (0, 1, 0), 60.0, W / H, jx, jy)


@pytest.fixture(scope="module")
def scene():
    tris, infos = ptamd.gen_soup(65536, 0)
    return tris, infos, ptamd.Settings(4, 2, 1.0, 1)


def oracle_acc(tris, infos, st, passes, seeds):
    b = orc.build(tris, *ptamd.scene_bounds(tris))
    acc = np.zeros((H, W, 3))
    for p in range(passes):
        orc.render_pass(tris, b, infos, st, W, H, pass_dirs(p), EYE, p, int(seeds[p]), acc)
    return acc


def test_multi_pass_frame_equals_sequential(ctx, scene):
    tris, infos, st = scene
    ctx.upload_scene(tris, infos, settings=st)
    n = 6
    seeds = np.arange(100, 100 + n, dtype=np.uint64)
    dirs = np.stack([pass_dirs(p) for p in range(n)])
    o = oracle_acc(tris, infos, st, n, seeds)
    results = {}
    for graphs in (1, 0):
        ptamd.tune_set("render.graphs", graphs)
        say(f"[render] sequential passes, graphs={graphs}")
        ctx.reset_frames()
        for p in range(n):
            ctx.set_primary_rays(dirs[p], EYE)
            ctx.render_pass(W, H, int(seeds[p]))
        results[("seq", graphs)] = ctx.read_accumulator(W, H)
        say(f"[render] one {n}-pass frame, graphs={graphs}")
        ctx.reset_frames()
        ctx.render_passes(W, H, seeds, dirs)
        results[("frame", graphs)] = ctx.read_accumulator(W, H)
        say(f"[render] 2+4 pass frames, graphs={graphs}")
        ctx.reset_frames()
        ctx.render_passes(W, H, seeds[:2], dirs[:2])
        ctx.render_passes(W, H, seeds[2:], dirs[2:])
        results[("split", graphs)] = ctx.read_accumulator(W, H)
    ptamd.tune_set("render.graphs", 1)
    assert o.sum() > 0
    for k, (acc, spp) in results.items():
        assert spp == n, k
        assert np.array_equal(acc, o), k


def test_graph_replays_match_eager(ctx, scene):
    tris, infos, st = scene
    ctx.upload_scene(tris, infos, settings=st)
    frames = 5
    seeds = np.arange(7, 7 + 2 * frames, dtype=np.uint64)
    dirs = np.stack([pass_dirs(p) for p in range(2 * frames)])
    per_mode = {}
    for graphs in (0, 1):
        ptamd.tune_set("render.graphs", graphs)
        ctx.reset_frames()
        accs = []
        for f in range(frames):
            say(f"[graph replay] graphs={graphs} frame {f}")
            ctx.render_passes(W, H, seeds[2 * f:2 * f + 2], dirs[2 * f:2 * f + 2])
            ctx.synchronize()
            accs.append(ctx.read_accumulator(W, H)[0])
        per_mode[graphs] = accs
    ptamd.tune_set("render.graphs", 1)
    for f in range(frames):
        assert np.array_equal(per_mode[0][f], per_mode[1][f]), f"frame {f}"


def test_render_invariant_under_tuning(ctx, scene):
    tris, infos, st = scene
    ctx.upload_scene(tris, infos, settings=st)
    seeds = np.arange(3, dtype=np.uint64)
    dirs = np.stack([pass_dirs(p) for p in range(3)])
    o = oracle_acc(tris, infos, st, 3, seeds)
    knobs = ("trace.blocks", "trace.refill", "trace.lds_stack", "trace.pool")
    saved = {k: ptamd.tune_get(k) for k in knobs}
    try:
        for setting in ((1, 64, 8, 256), (17, 5, 32, 64), (4096, 40, 16, 4096)):
            say(f"[render tuning] {dict(zip(knobs, setting))}")
            for k, v in zip(knobs, setting):
                ptamd.tune_set(k, v)
            ctx.reset_frames()
            ctx.render_passes(W, H, seeds, dirs)
            acc, spp = ctx.read_accumulator(W, H)
            assert spp == 3 and np.array_equal(acc, o)
    finally:
        for k, v in saved.items():
            ptamd.tune_set(k, v)


def test_frame_structures_match_oracle(ctx, scene):
    """Every frame structure gives the oracle's accumulator and segment
    count: path-persistent frames (render.paths 1) at several shade batch
    sizes, and bounce levels (render.paths 0) split into 1..8 chains, each on
    its own stream, forked and joined inside the frame's graph or launched
    eagerly."""
    tris, infos, st = scene
    ctx.upload_scene(tris, infos, settings=st)
    w, h, n = 128, 64, 8  # 65,536 paths: up to 4 chains of 16,384
    seeds = np.arange(200, 200 + n, dtype=np.uint64)
    dirs = np.stack([ptamd.camera_corners(EYE, TGT, (0, 1, 0), 60.0, w / h, *ptamd.pass_jitter(5, p, w, h))
                     for p in range(n)])
    b = orc.build(tris, *ptamd.scene_bounds(tris))
    o = np.zeros((h, w, 3))
    segs = 0
    for p in range(n):
        _, stats = orc.render_pass_stats(tris, b, infos, st, w, h, dirs[p], EYE, p, int(seeds[p]), o)
        segs += stats["segments"]
    assert o.sum() > 0
    saved = {k: ptamd.tune_get(k) for k in ("render.chains", "render.graphs", "render.paths", "render.shade_min")}
    structures = [("render.paths", 1, "render.shade_min", m) for m in (1, 7, 32, 64)] + \
                 [("render.paths", 0, "render.chains", c) for c in (1, 2, 3, 4, 8)]
    try:
        for graphs in (1, 0):
            ptamd.tune_set("render.graphs", graphs)
            for k1, v1, k2, v2 in structures:
                say(f"[frame structure] {k1}={v1} {k2}={v2}, graphs={graphs}, {w}x{h}x{n} paths")
                ptamd.tune_set(k1, v1)
                ptamd.tune_set(k2, v2)
                chains = (k2, v2)
                ctx.reset_frames()
                ctx.render_passes(w, h, seeds, dirs)
                acc, spp = ctx.read_accumulator(w, h)
                assert spp == n and np.array_equal(acc, o), (chains, graphs)
                assert ctx.last_frame_rays() == segs, (chains, graphs)
            # Two frames in a row reuse the captured graph (or the streams).
            ctx.reset_frames()
            ctx.render_passes(w, h, seeds[:4], dirs[:4])
            ctx.render_passes(w, h, seeds[4:], dirs[4:])
            acc, spp = ctx.read_accumulator(w, h)
            assert spp == n and np.array_equal(acc, o), ("two frames", graphs)
    finally:
        for k, v in saved.items():
            ptamd.tune_set(k, v)


def test_batched_scene_frame(ctx):
    """One frame whose passes trace different scenes of a batch (including a
    1-triangle and an empty scene) equals the oracle pass by pass."""
    counts = [20000, 1, 0, 5000, 65536, 2]
    tris, infos, offs = ptamd.gen_batch(len(counts), 12, counts=counts)
    infos["emission"][::37] = 5.0  # scattered emitting triangles in every scene
    st = ptamd.Settings(6, 3, 1.0, 1)
    ctx.upload_scenes(tris, offs, infos, settings=st)
    npass = len(counts)
    seeds = np.arange(50, 50 + npass, dtype=np.uint64)
    dirs = np.stack([pass_dirs(p) for p in range(npass)])
    scenes = np.arange(npass, dtype=np.uint32)[::-1].copy()
    results = []
    for paths in (1, 0):
        ptamd.tune_set("render.paths", paths)
        say(f"[batched frame] {npass} scenes, paths={paths}")
        ctx.reset_frames()
        ctx.render_passes(W, H, seeds, dirs, scenes)
        results.append((*ctx.read_accumulator(W, H), ctx.last_frame_rays()))
    ptamd.tune_set("render.paths", 1)
    assert np.array_equal(results[0][0], results[1][0]) and results[0][2] == results[1][2]
    acc, spp, rays = results[0]
    assert spp == npass and rays >= W * H * npass
    o = np.zeros((H, W, 3))
    segs = 0
    for p in range(npass):
        s = int(scenes[p])
        t = tris[offs[s]:offs[s + 1]]
        bmin, bmax = ptamd.scene_bounds(t) if len(t) else (np.zeros(3), np.ones(3))
        b = orc.build(t, bmin, bmax)
        _, sg = orc.render_pass(t, b, infos, st, W, H, dirs[p], EYE, p, int(seeds[p]), o)
        segs += sg
    assert np.array_equal(acc, o)
    assert rays == segs


def lit_scene():
    from test_oracle_lights import light
    tris, infos = ptamd.gen_soup(30000, 4)
    rng = np.random.default_rng(1)
    infos["specularity"] = rng.choice(np.array([0.0, 4.0, 32.0], np.float32), len(infos))
    lights = np.array([light((1.0, 0.9, 0.8), (0.3, -1.0, 0.2), 1.5),
                       light((0.2, 0.3, 1.0), (-0.5, -0.4, 1.0), 0.7)], np.float32)
    return tris, infos, lights


@pytest.mark.parametrize("graphs,paths", [(1, 1), (0, 0), (1, 0)])
def test_render_with_lights_matches_oracle(ctx, graphs, paths):
    """Directional lights: Blinn-Phong at primary hits and the light term at
    bounce hits, each behind a shadow ray, bit-identical to the oracle."""
    tris, infos, lights = lit_scene()
    st = ptamd.Settings(3, 2, 1.0, 1)
    ctx.upload_scene(tris, infos, settings=st, lights=lights)
    n = 3
    seeds = np.arange(300, 300 + n, dtype=np.uint64)
    dirs = np.stack([pass_dirs(p) for p in range(n)])
    ptamd.tune_set("render.graphs", graphs)
    ptamd.tune_set("render.paths", paths)
    try:
        say(f"[lights] {n}-pass frame, graphs={graphs}, paths={paths}")
        ctx.reset_frames()
        ctx.render_passes(W, H, seeds, dirs)
        acc, spp = ctx.read_accumulator(W, H)
    finally:
        ptamd.tune_set("render.graphs", 1)
        ptamd.tune_set("render.paths", 1)
    b = orc.build(tris, *ptamd.scene_bounds(tris))
    o = np.zeros((H, W, 3))
    for p in range(n):
        orc.render_pass_lit(tris, b, infos, lights, st, W, H, dirs[p], EYE, p, int(seeds[p]), o)
    assert spp == n and o.sum() > 0
    assert np.array_equal(acc, o)


BOX_OBJ = """mtllib box.mtl
v -1 0 -1
v 1 0 -1
v 1 0 1
v -1 0 1
v -1 2 -1
v 1 2 -1
v 1 2 1
v -1 2 1
v -0.3 1.98 -0.3
v 0.3 1.98 -0.3
v 0.3 1.98 0.3
v -0.3 1.98 0.3
v -0.4 0 -0.2
v 0.2 0 -0.2
v -0.1 0.9 0.1
usemtl white
f 1 2 3 4
f 5 8 7 6
f 1 5 6 2
usemtl red
f 1 4 8 5
usemtl green
f 2 6 7 3
usemtl light
f 9 10 11 12
usemtl white
f 13 14 15
"""

BOX_MTL = """newmtl white
Kd 0.725 0.71 0.68
newmtl red
Kd 0.63 0.065 0.05
newmtl green
Kd 0.14 0.45 0.091
newmtl light
Kd 0.78 0.78 0.78
Ke 17 12 4
"""


def test_obj_scene_render_matches_oracle(ctx, tmp_path):
    """An OBJ box with an emitting quad, imported with the reference's import
    conventions, renders bit-identically to the oracle."""
    (tmp_path / "box.mtl").write_text(BOX_MTL)
    (tmp_path / "box.obj").write_text(BOX_OBJ)
    tris, infos = ptamd.load_obj(str(tmp_path / "box.obj"))
    assert len(tris) == 13  # six quads and one triangle
    st = ptamd.Settings(5, 3, 1.0, 1)
    ctx.upload_scene(tris, infos, settings=st)
    # the box is open towards -z once mirrored (the file's open side is +z)
    eye, tgt = (0.0, 1.0, -3.0), (0.0, 1.0, 0.0)
    n = 4
    dirs = np.stack([ptamd.camera_corners(eye, tgt, (0, 1, 0), 60.0, W / H, *ptamd.pass_jitter(3, p, W, H))
                     for p in range(n)])
    ctx.set_primary_rays(dirs[0], eye)
    ctx.reset_frames()
    seeds = np.arange(40, 40 + n, dtype=np.uint64)
    say("[obj] box scene, 4 passes")
    ctx.render_passes(W, H, seeds, dirs)
    acc, _ = ctx.read_accumulator(W, H)
    b = orc.build(tris, *ptamd.scene_bounds(tris))
    o = np.zeros((H, W, 3))
    for p in range(n):
        orc.render_pass(tris, b, infos, st, W, H, dirs[p], eye, p, int(seeds[p]), o)
    assert o.sum() > 0
    assert np.array_equal(acc, o)


def test_screen_buffer_and_device_image(ctx, scene):
    """The screen buffer written after every frame, and the device image
    readback, equal pt_read_image bit for bit; a smaller buffer gets the
    image's corner, a larger one only the image's extent."""
    from ptamd import hip
    tris, infos, st = scene
    ctx.upload_scene(tris, infos, settings=st)
    ctx.reset_frames()
    big = hip.DeviceArray.from_host(np.full((H + 3) * (W + 5) * 4, -7.0, np.float32))
    small = hip.DeviceArray(10 * 12 * 4, np.float32)
    try:
        ctx.set_screen_buffer(big.ptr, W + 5, H + 3)
        for p in range(3):
            say(f"[screen] frame {p}")
            ctx.set_primary_rays(pass_dirs(p), EYE)
            ctx.render_pass(W, H, 900 + p)
            ctx.synchronize()
            want = ctx.read_image(W, H)
            got = big.to_host().reshape(H + 3, W + 5, 4)
            np.testing.assert_array_equal(got[:H, :W], want)
            assert np.all(got[H:] == -7.0) and np.all(got[:, W:] == -7.0)
        ctx.read_image_device(small.ptr, 12, 10)
        ctx.synchronize()
        np.testing.assert_array_equal(small.to_host().reshape(10, 12, 4), ctx.read_image(W, H)[:10, :12])
        with pytest.raises(ptamd.PtError):
            ctx.set_screen_buffer(big.ptr, W, H, row_pitch=8 * W)
    finally:
        ctx.set_screen_buffer(None)
    ctx.render_pass(W, H, 999)  # detached: nothing written
    ctx.synchronize()


def random_detail(n, seed):
    rng = np.random.default_rng(seed)
    det = np.zeros(n, ptamd.DETAIL_DTYPE)
    det["color"] = rng.uniform(0.05, 1.0, (n, 3, 3))
    det["uv"] = rng.uniform(-2.0, 3.0, (n, 3, 2))
    det["tangent"] = rng.normal(size=(n, 3, 3))
    det["bitangent"] = rng.normal(size=(n, 3, 3))
    det["albedo_tex"] = np.where(rng.random(n) < 0.6, rng.integers(0, 3, n), ptamd.NO_TEXTURE)
    det["normal_tex"] = np.where(rng.random(n) < 0.5, rng.integers(0, 3, n), ptamd.NO_TEXTURE)
    texs = [rng.integers(0, 256, (8, 16, 4), dtype=np.uint8), rng.integers(0, 256, (1, 1, 4), dtype=np.uint8),
            rng.integers(0, 256, (32, 32, 4), dtype=np.uint8)]
    return det, texs


@pytest.mark.parametrize("graphs,paths", [(1, 1), (0, 0), (1, 0)])
def test_render_with_detail_matches_oracle(ctx, scene, graphs, paths):
    """Textures, vertex colours and normal maps (extractInfo / unpackNormal):
    accumulators bit-identical to the oracle, with directional lights."""
    tris, infos, st = scene
    det, texs = random_detail(len(infos), 5)
    from test_oracle_lights import light
    lights = np.array([light((1.0, 0.9, 0.8), (0.3, -1.0, 0.2), 1.5)], np.float32)
    ctx.upload_scene(tris, infos, settings=st, lights=lights)
    ctx.upload_detail(det, texs)
    ptamd.tune_set("render.graphs", graphs)
    ptamd.tune_set("render.paths", paths)
    try:
        seeds = np.arange(40, 43, dtype=np.uint64)
        dirs = np.stack([pass_dirs(p) for p in range(3)])
        ctx.reset_frames()
        say(f"[detail] 3 passes, graphs={graphs}, paths={paths}")
        ctx.render_passes(W, H, seeds, dirs)
        acc, spp = ctx.read_accumulator(W, H)
    finally:
        ptamd.tune_set("render.graphs", 1)
        ptamd.tune_set("render.paths", 1)
    b = orc.build(tris, *ptamd.scene_bounds(tris))
    o = np.zeros((H, W, 3))
    for p in range(3):
        orc.render_pass_detail(tris, b, infos, det, texs, st, W, H, pass_dirs(p), EYE, p, int(seeds[p]), o,
                               lights=lights)
    assert spp == 3 and np.array_equal(acc, o)
    plain = np.zeros((H, W, 3))
    for p in range(3):
        orc.render_pass_lit(tris, b, infos, lights, st, W, H, pass_dirs(p), EYE, p, int(seeds[p]), plain)
    assert not np.array_equal(plain, o)  # the detail changes the image
    # removing the detail restores the plain shading
    ctx.upload_detail(np.zeros(0, ptamd.DETAIL_DTYPE))
    ctx.reset_frames()
    ctx.render_passes(W, H, seeds, dirs)
    assert np.array_equal(ctx.read_accumulator(W, H)[0], plain)


def test_detail_errors(ctx, scene):
    tris, infos, st = scene
    ctx.upload_scene(tris, infos, settings=st)
    det, texs = random_detail(len(infos), 6)
    with pytest.raises(ptamd.PtError):
        ctx.upload_detail(det[:-1], texs)  # fewer records than ids
    with pytest.raises(ptamd.PtError):
        ctx.upload_detail(det, texs[:2])  # index 2 out of range
    bad = det.copy()
    bad["albedo_tex"][0] = 7
    with pytest.raises(ptamd.PtError):
        ctx.upload_detail(bad, texs)


def test_textured_obj_scene_matches_oracle(ctx, tmp_path):
    """The OBJ box with a checker-textured, normal-mapped floor (TGA files,
    vt coordinates, map_Kd and map_bump) through pt_asset_load_obj and
    pt_upload_detail renders bit-identically to the oracle with the same
    detail."""
    from test_assets_detail import tga_bytes
    j, i = np.mgrid[0:16, 0:16]
    chk = np.zeros((16, 16, 4), np.uint8)
    chk[..., :3] = np.where(((i // 4 + j // 4) % 2 == 0)[..., None], [230, 220, 200], [40, 60, 90])
    chk[..., 3] = 255
    nrm = np.zeros((8, 8, 4), np.uint8)
    nrm[..., 0] = 128 + 60 * np.sin(i[:8, :8])
    nrm[..., 1] = 128 + 60 * np.cos(j[:8, :8])
    nrm[..., 2] = 230
    nrm[..., 3] = 255
    (tmp_path / "chk.tga").write_bytes(tga_bytes(chk, 2))
    (tmp_path / "nrm.tga").write_bytes(tga_bytes(nrm, 10))
    mtl = BOX_MTL + "newmtl floor\nKd 0.7 0.7 0.7\nmap_Kd chk.tga\nmap_bump nrm.tga\n"
    obj = BOX_OBJ.replace("usemtl white\nf 1 2 3 4\n",
                          "vt 0 0\nvt 3 0\nvt 3 3\nvt 0 3\nusemtl floor\nf 1/1 2/2 3/3 4/4\nusemtl white\n")
    (tmp_path / "box.mtl").write_text(mtl)
    (tmp_path / "box.obj").write_text(obj)
    tris, infos, det, texs = ptamd.load_obj_asset(str(tmp_path / "box.obj"))
    assert len(tris) == 13 and len(texs) == 2 and det["albedo_tex"][0] == 0 and det["normal_tex"][0] == 1
    st = ptamd.Settings(4, 3, 1.0, 1)
    ctx.upload_scene(tris, infos, settings=st)
    ctx.upload_detail(det, texs)
    eye, tgt = (0.0, 1.0, -3.0), (0.0, 0.6, 0.0)
    n = 3
    dirs = np.stack([ptamd.camera_corners(eye, tgt, (0, 1, 0), 60.0, W / H, *ptamd.pass_jitter(4, p, W, H))
                     for p in range(n)])
    ctx.set_primary_rays(dirs[0], eye)
    ctx.reset_frames()
    seeds = np.arange(70, 70 + n, dtype=np.uint64)
    say("[obj] textured box scene, 3 passes")
    ctx.render_passes(W, H, seeds, dirs)
    acc, _ = ctx.read_accumulator(W, H)
    b = orc.build(tris, *ptamd.scene_bounds(tris))
    o = np.zeros((H, W, 3))
    for p in range(n):
        orc.render_pass_detail(tris, b, infos, det, texs, st, W, H, dirs[p], eye, p, int(seeds[p]), o)
    assert o.sum() > 0
    assert np.array_equal(acc, o)


def test_stop_at_frame(ctx, scene):
    """CUDAPathTracer::raytrace's stop frame: passes beyond the limit render
    nothing; a multi-pass call renders up to it; the generation time is set
    once the limit is found reached and cleared by a reset."""
    tris, infos, st = scene
    ctx.upload_scene(tris, infos, settings=st)
    seeds = np.arange(60, 66, dtype=np.uint64)
    dirs = np.stack([pass_dirs(p) for p in range(6)])
    try:
        ctx.set_primary_rays(dirs[0], EYE)  # the eye every pass starts at
        ctx.reset_frames()
        ctx.render_passes(W, H, seeds[:3], dirs[:3])
        want = ctx.read_accumulator(W, H)[0]
        ctx.set_stop_frame(3)
        ctx.reset_frames()
        for p in range(5):
            ctx.set_primary_rays(dirs[p], EYE)
            ctx.render_pass(W, H, int(seeds[p]))
        acc, spp = ctx.read_accumulator(W, H)
        assert spp == 3 and np.array_equal(acc, want)
        assert ctx.generation_time() > 0.0
        ctx.reset_frames()
        assert ctx.generation_time() == -1.0
        ctx.render_passes(W, H, seeds, dirs)  # 6 requested, 3 rendered
        acc, spp = ctx.read_accumulator(W, H)
        assert spp == 3 and np.array_equal(acc, want)
    finally:
        ctx.set_stop_frame(0)


@pytest.mark.parametrize("w,h", [(37, 21), (5, 3), (1, 1), (8, 13), (130, 9)])
def test_ragged_image_sizes_match_oracle(ctx, scene, w, h):
    """Image sizes that are not multiples of 8 (primary rays are queued in
    8x8 pixel tiles, the last band and tile of a row cut short): every
    pixel is traced exactly once, bit-identical to the oracle."""
    tris, infos, st = scene
    ctx.upload_scene(tris, infos, settings=st)
    b = orc.build(tris, *ptamd.scene_bounds(tris))
    n = 3
    seeds = np.arange(300, 300 + n, dtype=np.uint64)
    dirs = np.stack([ptamd.camera_corners(EYE, TGT, (0, 1, 0), 60.0, w / h, *ptamd.pass_jitter(4, p, w, h))
                     for p in range(n)])
    o = np.zeros((h, w, 3))
    for p in range(n):
        orc.render_pass(tris, b, infos, st, w, h, dirs[p], EYE, p, int(seeds[p]), o)
    say(f"[render ragged] {w}x{h}")
    ctx.set_primary_rays(dirs[0], EYE)
    ctx.reset_frames()
    ctx.render_passes(w, h, seeds, dirs)
    acc, spp = ctx.read_accumulator(w, h)
    assert spp == n and np.array_equal(acc, o)
