"""GPU parity: the HIP path through the C-ABI against the CPU oracle on the
same seeded inputs.

Bar (bit-exact unless stated):
  build   Morton keys, sort permutation, children, parents[0], parents[1] and
          internal node boxes (compared as bit patterns)
  trace   closest original id, t and barycentrics bit-identical to the oracle's
          reference-order traversal (the contract allows t within 1e-6
          relative; checked too), visit counters identical
  render  FP64 accumulators at equal seeds and passes: mean relative
          per-channel error <= 1e-3 (the contract), and reported exactness
"""
import sys

import numpy as np
import pytest

import oracle_py as orc
import ptamd

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def check_build(ctx, tris, infos, bmin=None, bmax=None, tag=""):
    say(f"[build parity] {tag} n={len(tris)}")
    ctx.upload_scene(tris, infos, bmin, bmax)
    g = ctx.get_bvh()
    if bmin is None:
        bmin, bmax = ptamd.scene_bounds(tris)
    o = orc.build(tris, bmin, bmax)
    for k in ("morton", "perm", "children", "parents_leaf", "parents_internal"):
        assert np.array_equal(g[k], o[k]), f"{tag}: {k} differs"
    assert np.array_equal(g["bboxes"].view(np.uint32), o["bboxes"].view(np.uint32)), f"{tag}: boxes differ"
    return o


def soup_dups(n):
    tris, infos = ptamd.gen_soup(n, 31, emitter=False)
    tris["v0"] = tris["v0"][0]
    tris["v1"] = tris["v1"][0]
    tris["v2"] = tris["v2"][0]
    return tris, infos


def soup_flat(n):
    tris, infos = ptamd.gen_soup(n, 32, emitter=False)
    for f in ("v0", "v1", "v2"):
        tris[f][:, 1] = np.float32(0.1)
    return tris, infos


BUILD_CASES = {
    "soup2": lambda: ptamd.gen_soup(2, 1, emitter=False),
    "soup3": lambda: ptamd.gen_soup(3, 1, emitter=False),
    "soup100": lambda: ptamd.gen_soup(98, 2),
    "soup4095": lambda: ptamd.gen_soup(4095, 3, emitter=False),
    "soup4096": lambda: ptamd.gen_soup(4096, 3, emitter=False),
    "soup4097": lambda: ptamd.gen_soup(4097, 3, emitter=False),
    "c0_soup": lambda: ptamd.gen_soup(65536, 0),
    "soup100k": lambda: ptamd.gen_soup(100003, 4),
    "dups": lambda: soup_dups(5000),
    "flat": lambda: soup_flat(5000),
    "sphere64": lambda: ptamd.gen_sphere(64, 5, room=True),
    "hair": lambda: ptamd.gen_hairball(200, 60, 6, light=True),
}


@pytest.mark.parametrize("case", list(BUILD_CASES))
def test_build_parity(ctx, case):
    tris, infos = BUILD_CASES[case]()
    check_build(ctx, tris, infos, tag=case)


def test_build_single_and_empty(ctx):
    tris, infos = ptamd.gen_soup(1, 9, emitter=False)
    say("[build parity] n=1")
    ctx.upload_scene(tris, infos)
    g = ctx.get_bvh()
    assert g["parents_leaf"][0] == ptamd.INVALID and g["perm"][0] == 0
    rays = ptamd.gen_shell_rays(4000, 1, *ptamd.scene_bounds(tris), 1.2)
    ids, t, uv = ctx.trace_closest(rays)
    b = orc.build(tris, *ptamd.scene_bounds(tris))
    oids, ot, ouv = orc.trace(tris, b, rays)
    assert np.array_equal(ids, oids) and np.array_equal(t, ot)
    say("[trace] n=0")
    ctx.upload_scene(tris[:0], infos[:0], (0, 0, 0), (1, 1, 1))
    ids, t, uv = ctx.trace_closest(rays)
    assert np.all(ids == ptamd.INVALID) and np.all(t == ptamd.FLT_MAX)


def test_build_given_bounds(ctx):
    # a caller box larger than the scene changes the keys; parity must hold
    tris, infos = ptamd.gen_soup(20000, 12)
    check_build(ctx, tris, infos, np.array([-1, -2, -3], np.float32),
                np.array([2, 3, 4], np.float32), tag="given-bounds")


def test_repeated_builds_identical(ctx):
    tris, infos = ptamd.gen_soup(1 << 20, 13, emitter=False)
    say("[repeat builds] n=2^20")
    ctx.upload_scene(tris, infos)
    first = ctx.get_bvh()
    for _ in range(4):
        ctx.build_bvh()
        again = ctx.get_bvh()
        for k in first:
            assert np.array_equal(first[k].view(np.uint32), again[k].view(np.uint32)), k


def interior_rays(n, seed, bmin, bmax):
    rng = np.random.default_rng(seed)
    o = bmin + (bmax - bmin) * rng.random((n, 3), dtype=np.float32)
    d = rng.normal(size=(n, 3)).astype(np.float32)
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    return np.concatenate([o, d], axis=1).astype(np.float32)


def axis_rays(n, seed, bmin, bmax):
    r = interior_rays(n, seed, bmin, bmax)
    rng = np.random.default_rng(seed + 1)
    k = rng.integers(0, 3, n)
    r[np.arange(n), 3 + k] = 0.0
    two = rng.random(n) < 0.5
    r[two, 3:] = 0.0
    r[two, 3 + k[two]] = 1.0
    return r


def check_trace(ctx, tris, infos, rays, tag):
    say(f"[trace parity] {tag} n={len(tris)} rays={len(rays)}")
    o = check_build(ctx, tris, infos, tag=tag)
    ids, t, uv = ctx.trace_closest(rays)
    oids, ot, ouv, oc = orc.trace(tris, o, rays, counters=True)
    assert np.array_equal(ids, oids), f"{tag}: {np.count_nonzero(ids != oids)} ids differ"
    hit = oids != ptamd.INVALID
    assert np.all(np.abs(t[hit] - ot[hit]) <= 1e-6 * np.abs(ot[hit]))
    assert np.array_equal(t.view(np.uint32), ot.view(np.uint32))
    assert np.array_equal(uv.view(np.uint32), ouv.view(np.uint32))
    # visit counters of the GPU traversal equal the oracle's
    from ptamd import hip
    d_rays = hip.DeviceArray.from_host(rays)
    c = ctx.trace_closest_counted(d_rays.ptr, len(rays))
    d_rays.free()
    assert np.array_equal(c, oc), (c, oc)
    return hit.mean()


TRACE_CASES = {
    "c0_soup": lambda: ptamd.gen_soup(65536, 0),
    "sphere128": lambda: ptamd.gen_sphere(128, 1, room=True),
    "hair": lambda: ptamd.gen_hairball(300, 144, 2, light=False),
    "dups": lambda: soup_dups(3000),
}


@pytest.mark.parametrize("case", list(TRACE_CASES))
def test_trace_parity(ctx, case):
    tris, infos = TRACE_CASES[case]()
    bmin, bmax = ptamd.scene_bounds(tris)
    rays = np.concatenate([ptamd.gen_shell_rays(60000, 3, bmin, bmax, 1.5),
                           interior_rays(60000, 4, bmin, bmax), axis_rays(8000, 5, bmin, bmax)])
    check_trace(ctx, tris, infos, rays, case)


def test_trace_invariant_under_tuning(ctx):
    """Launch-shape knobs (what the autotuner varies) never change results."""
    tris, infos = ptamd.gen_soup(65536, 0)
    bmin, bmax = ptamd.scene_bounds(tris)
    rays = np.concatenate([ptamd.gen_shell_rays(30000, 8, bmin, bmax, 1.5),
                           interior_rays(30000, 9, bmin, bmax)])
    ctx.upload_scene(tris, infos)
    b = orc.build(tris, bmin, bmax)
    oids, ot, ouv = orc.trace(tris, b, rays)
    knobs = ("trace.blocks", "trace.refill", "trace.lds_stack", "trace.pool", "trace.filter")
    saved = {k: ptamd.tune_get(k) for k in knobs}
    try:
        for setting in ((1, 1, 16, 256, 1), (3, 64, 8, 4096, 1), (64, 8, 32, 64, 0), (256, 32, 8, 1024, 1),
                        (100000, 48, 32, 4096, 0), (2048, 63, 16, 100, 1), (32, 48, 8, 256, 0),
                        (2048, 32, 16, 256, 0), (2048, 32, 32, 128, 1), (7, 2, 8, 64, 1)):
            say(f"[tuning invariance] {dict(zip(knobs, setting))}")
            for k, v in zip(knobs, setting):
                ptamd.tune_set(k, v)
            ids, t, uv = ctx.trace_closest(rays)
            assert np.array_equal(ids, oids) and np.array_equal(t.view(np.uint32), ot.view(np.uint32))
            assert np.array_equal(uv.view(np.uint32), ouv.view(np.uint32))
    finally:
        for k, v in saved.items():
            ptamd.tune_set(k, v)


def test_trace_parity_c1_full_mesh(ctx):
    # config 1 microbench mesh at full size: 871,200 triangles
    tris, infos = ptamd.gen_sphere(660, 2024, room=False)
    bmin, bmax = ptamd.scene_bounds(tris)
    rays = ptamd.gen_shell_rays(400000, 77, bmin, bmax, 1.5)
    rate = check_trace(ctx, tris, infos, rays, "c1-mesh")
    assert rate > 0.05


def test_build_parity_c2_full_hairball(ctx):
    tris, infos = ptamd.gen_hairball(10000, 144, 2025, light=False)
    assert len(tris) == 2_880_000
    o = check_build(ctx, tris, infos, tag="c2-hair")
    bmin, bmax = ptamd.scene_bounds(tris)
    assert np.array_equal(o["bboxes"][0], np.concatenate([bmin, bmax]))


def render_both(ctx, tris, infos, st, w, h, passes, eye, target, seed=11):
    ctx.upload_scene(tris, infos, settings=st)
    bmin, bmax = ptamd.scene_bounds(tris)
    b = orc.build(tris, bmin, bmax)
    oacc = np.zeros((h, w, 3), np.float64)
    for p in range(passes):
        jx, jy = ptamd.pass_jitter(seed, p, w, h)
        dirs = ptamd.camera_corners(eye, target, (0, 1, 0), 60.0, w / h, jx, jy)
        ctx.set_primary_rays(dirs, eye)
        ctx.render_pass(w, h, seed + p)
        orc.render_pass(tris, b, infos, st, w, h, dirs, eye, p, seed + p, oacc)
    gacc, spp = ctx.read_accumulator(w, h)
    assert spp == passes
    return gacc, oacc


def rel_err(g, o):
    return [float(np.abs(g[..., c] - o[..., c]).sum() / max(np.abs(o[..., c]).sum(), 1e-30)) for c in range(3)]


def test_render_parity_c0(ctx):
    say("[render parity] c0 soup 96x96 x4 passes, 4 bounces")
    tris, infos = ptamd.gen_soup(65536, 0)
    st = ptamd.Settings(4, ptamd.INVALID, 1.0, 1)
    g, o = render_both(ctx, tris, infos, st, 96, 96, 4, (0.5, 0.5, -1.5), (0.5, 0.5, 0.0))
    assert o.sum() > 0
    err = rel_err(g, o)
    say(f"  mean relative per-channel error {err}, exact={np.array_equal(g, o)}")
    assert max(err) <= 1e-3
    assert np.array_equal(g, o)


def test_render_parity_sphere_room_rr(ctx):
    say("[render parity] sphere room 64x48 x3 passes, 8 bounces, RR from 3")
    tris, infos = ptamd.gen_sphere(48, 3, room=True)
    st = ptamd.Settings(8, 3, 1.0, 1)
    g, o = render_both(ctx, tris, infos, st, 64, 48, 3, (0.0, 0.0, -1.9), (0.0, 0.0, 0.0))
    assert o.sum() > 0
    err = rel_err(g, o)
    say(f"  mean relative per-channel error {err}, exact={np.array_equal(g, o)}")
    assert max(err) <= 1e-3


def test_accumulator_roundtrip_and_image(ctx):
    say("[accumulator] write/read/image")
    tris, infos = ptamd.gen_soup(5000, 1)
    ctx.upload_scene(tris, infos)
    acc = np.random.default_rng(0).random((20, 30, 3))
    ctx.write_accumulator(acc, 4)
    back, spp = ctx.read_accumulator(30, 20)
    assert spp == 4 and np.array_equal(back, acc)
    img = ctx.read_image(30, 20)
    assert np.array_equal(img[..., :3], (acc / 4).astype(np.float32)) and np.all(img[..., 3] == 1)
    ctx.reset_frames()
    back, spp = ctx.read_accumulator(30, 20)
    assert spp == 0 and not back.any()


def test_errors_are_status_codes(ctx):
    say("[errors]")
    tris, infos = ptamd.gen_soup(100, 1)
    with pytest.raises(ptamd.PtError) as e:
        ctx.upload_scene(tris, infos[:10])  # missing shading records
    assert e.value.code == ptamd.PT_E_INVALID_ARG
    with pytest.raises(ptamd.PtError) as e:
        ptamd.Context(99)
    assert e.value.code == ptamd.PT_E_NO_DEVICE


def test_batched_scenes_build_and_trace(ctx):
    """Batched build: every scene's arrays equal its own single-scene oracle
    build (empty, 1-, 2-triangle and 4097-triangle scenes included), and traces
    with per-ray scene ids equal per-scene oracle traces."""
    counts = [3000, 0, 1, 2, 4097, 17, 65536, 1200, 1, 300, 70000, 5]
    tris, infos, offs = ptamd.gen_batch(len(counts), 41, counts=counts)
    say(f"[batched build] scenes={len(counts)} n={len(tris)}")
    ctx.upload_scenes(tris, offs, infos)
    rng = np.random.default_rng(3)
    rays_all, sc_all, expect = [], [], []
    for s in range(len(counts)):
        t = tris[offs[s]:offs[s + 1]]
        if len(t) == 0:
            continue
        bmin, bmax = ptamd.scene_bounds(t)
        o = orc.build(t, bmin, bmax)
        g = ctx.get_bvh_scene(s)
        for k in ("morton", "perm", "children", "parents_leaf", "parents_internal"):
            assert np.array_equal(g[k], o[k]), f"scene {s}: {k}"
        assert np.array_equal(g["bboxes"].view(np.uint32), o["bboxes"].view(np.uint32)), f"scene {s}: boxes"
        r = np.concatenate([ptamd.gen_shell_rays(3000, s, bmin, bmax, 1.3), interior_rays(3000, s, bmin, bmax)])
        rays_all.append(r)
        sc_all.append(np.full(len(r), s, np.uint32))
        expect.append(orc.trace(t, o, r))
    # rays of the empty scene and of an out-of-range scene miss
    rays_all.append(rays_all[0][:50])
    sc_all.append(np.concatenate([np.full(25, 1, np.uint32), np.full(25, 999, np.uint32)]))
    expect.append((np.full(50, ptamd.INVALID, np.uint32), np.full(50, ptamd.FLT_MAX, np.float32),
                   np.zeros((50, 2), np.float32)))
    rays = np.concatenate(rays_all)
    sc = np.concatenate(sc_all)
    perm = rng.permutation(len(rays))  # interleave scenes across waves
    ids, t, uv = ctx.trace_closest_scenes(rays[perm], sc[perm])
    eids = np.concatenate([e[0] for e in expect])[perm]
    et = np.concatenate([e[1] for e in expect])[perm]
    euv = np.concatenate([e[2] for e in expect])[perm]
    assert np.array_equal(ids, eids)
    assert np.array_equal(t.view(np.uint32), et.view(np.uint32))
    assert np.array_equal(uv.view(np.uint32), euv.view(np.uint32))
    # a single-scene upload afterwards still builds and traces as before
    check_build(ctx, *ptamd.gen_soup(5000, 2), tag="after-batch")


@pytest.mark.parametrize("staged", [1, 0])
def test_build_parity_both_refit_paths(ctx, staged):
    """The refit's upper levels (nodes whose leaves span more than one
    1024-leaf block) from node lists staged in LDS (build.staged 1) or from
    global memory (0; also the path of any list longer than 1024 nodes, such
    as the top list of the config-3 batch, ~5,600 nodes): bit-identical trees
    on a 100,003-triangle soup and on sampled scenes of the 512-scene batch
    (log-uniform 1,024..262,144 triangles, 24.0M in all)."""
    saved = ptamd.tune_get("build.staged")
    ptamd.tune_set("build.staged", staged)
    try:
        check_build(ctx, *ptamd.gen_soup(100003, 4), tag=f"staged={staged}")
        tris, infos, offs = ptamd.gen_batch(512, 3)
        say(f"[batched build] staged={staged}, 512 scenes n={len(tris)}")
        ctx.upload_scenes(tris, offs, infos)
        assert ctx.verify_bvh() == 0
        for s in list(range(0, 512, 41)) + [int(np.argmax(np.diff(offs)))]:
            t = tris[offs[s]:offs[s + 1]]
            o = orc.build(t, *ptamd.scene_bounds(t))
            g = ctx.get_bvh_scene(s)
            for k in ("children", "parents_leaf", "parents_internal"):
                assert np.array_equal(g[k], o[k]), f"scene {s}: {k}"
            assert np.array_equal(g["bboxes"].view(np.uint32), o["bboxes"].view(np.uint32)), f"scene {s}"
    finally:
        ptamd.tune_set("build.staged", saved)


def test_batched_512_scene_build_parity(ctx):
    """Config-3 shape at reduced sizes: 512 soups, log-uniform [64, 4096]."""
    tris, infos, offs = ptamd.gen_batch(512, 7, lo=64, hi=4096)
    say(f"[batched build] 512 scenes n={len(tris)}")
    ctx.upload_scenes(tris, offs, infos)
    for s in range(0, 512, 7):
        t = tris[offs[s]:offs[s + 1]]
        o = orc.build(t, *ptamd.scene_bounds(t))
        g = ctx.get_bvh_scene(s)
        for k in ("morton", "perm", "children", "parents_leaf", "parents_internal"):
            assert np.array_equal(g[k], o[k]), f"scene {s}: {k}"
        assert np.array_equal(g["bboxes"].view(np.uint32), o["bboxes"].view(np.uint32)), f"scene {s}"


def test_render_parity_full_room_mesh(ctx):
    """The 871,200-triangle mesh inside the room (deep trees: traversal stacks
    spill past the 16 LDS entries) at 160x90, 9 bounces with RR from 3:
    1-bounce and 9-bounce accumulators bit-identical to the oracle, single
    passes and a 3-pass frame."""
    tris, infos = ptamd.gen_sphere(660, 2024, room=True)
    bmin, bmax = ptamd.scene_bounds(tris)
    b = orc.build(tris, bmin, bmax)
    w, h = 160, 90
    eye, tgt = (0.0, 0.0, -1.9), (0.0, 0.0, 0.0)
    d0 = ptamd.camera_corners(eye, tgt, (0, 1, 0), 60.0, w / h)
    # (the 1-bounce frame is unstratified: a single pass's shared first-bounce
    # stratum would rarely reach the small ceiling emitter)
    for bounces, npass, strat in ((1, 1, 0), (9, 1, 1), (9, 3, 1)):
        say(f"[render parity] full room mesh {w}x{h}, {bounces} bounces, {npass} passes")
        st = ptamd.Settings(bounces, 3, 1.0, strat)
        ctx.upload_scene(tris, infos, settings=st)
        ctx.set_primary_rays(d0, eye)
        ctx.reset_frames()
        seeds = np.arange(7000, 7000 + npass, dtype=np.uint64)
        ctx.render_passes(w, h, seeds, np.stack([d0] * npass))
        g, spp = ctx.read_accumulator(w, h)
        o = np.zeros((h, w, 3))
        for p in range(npass):
            orc.render_pass(tris, b, infos, st, w, h, d0, eye, p, int(seeds[p]), o)
        assert spp == npass and o.sum() > 0
        assert np.array_equal(g, o)


def test_deep_stack_trace(ctx):
    """Trees 42 internal levels deep (tests/deep_scene.py): deferred-node
    stacks of 40+ entries, deeper than every LDS stack setting, so pushes
    spill to scratch.  Traced bit-exact against the oracle at trace.lds_stack
    8, 16 and 32; the GPU's visit counters and deepest-stack counter equal the
    oracle's; no status flag is raised."""
    from deep_scene import BOX_MAX, BOX_MIN, deep_rays, deep_scene
    from ptamd import hip

    tris, infos = deep_scene()
    o = check_build(ctx, tris, infos, BOX_MIN, BOX_MAX, tag="deep")
    rays = deep_rays(8192)
    oids, ot, ouv, oc = orc.trace(tris, o, rays, counters=True)
    assert int(oc[3]) >= 40
    saved = ptamd.tune_get("trace.lds_stack")
    d_rays = hip.DeviceArray.from_host(rays)
    try:
        for lds in (8, 16, 32):
            say(f"[deep stack] trace.lds_stack={lds}, oracle depth {int(oc[3])}")
            ptamd.tune_set("trace.lds_stack", lds)
            ids, t, uv = ctx.trace_closest(rays)
            assert np.array_equal(ids, oids)
            assert np.array_equal(t.view(np.uint32), ot.view(np.uint32))
            assert np.array_equal(uv.view(np.uint32), ouv.view(np.uint32))
            c = ctx.trace_closest_counted(d_rays.ptr, len(rays))
            assert np.array_equal(c, oc), (c, oc)
            assert int(c[3]) > lds
    finally:
        ptamd.tune_set("trace.lds_stack", saved)
        d_rays.free()
    assert ctx.get_status() == 0


def test_trace_parity_c2_full_hairball(ctx):
    """Config 2 at its full size: the 2,880,002-triangle hairball the bench
    times, 300,000 shell rays of the bench's generator plus 100,000 interior
    rays, bit-exact with equal visit counters."""
    tris, infos = ptamd.gen_hairball(10000, 144, 2025, light=True)
    assert len(tris) == 2_880_002
    bmin, bmax = ptamd.scene_bounds(tris)
    rays = np.concatenate([ptamd.gen_shell_rays(300000, 78, bmin, bmax, 1.5),
                           interior_rays(100000, 79, bmin, bmax)])
    rate = check_trace(ctx, tris, infos, rays, "c2-hair-full")
    assert rate > 0.05


def test_batched_config3_full_size(ctx):
    """Config 3 as the bench runs it: 512 soups with log-uniform sizes in
    [1,024, 262,144] (24.0M triangles) built by one launch sequence; 40
    scenes (the smallest, the largest and every 13th) checked against their
    own oracle builds, and traced through per-ray scene ids bit-exact; three
    scenes rendered through the batch (8 bounces, RR from 3) equal to their
    single-scene oracle renders."""
    tris, infos, offs = ptamd.gen_batch(512, 3)
    say(f"[batched c3] 512 scenes n={len(tris)}")
    counts = np.diff(offs)
    assert counts.min() >= 1024 and counts.max() <= 262144
    st = ptamd.Settings(8, 3, 1.0, 1)
    ctx.upload_scenes(tris, offs, infos, settings=st)
    picks = sorted(set([int(np.argmin(counts)), int(np.argmax(counts))] + list(range(0, 512, 13))))
    rays_all, sc_all, expect, builds = [], [], [], {}
    for s in picks:
        t = tris[offs[s]:offs[s + 1]]
        bmin, bmax = ptamd.scene_bounds(t)
        o = orc.build(t, bmin, bmax)
        builds[s] = o
        g = ctx.get_bvh_scene(s)
        for k in ("morton", "perm", "children", "parents_leaf", "parents_internal"):
            assert np.array_equal(g[k], o[k]), f"scene {s}: {k}"
        assert np.array_equal(g["bboxes"].view(np.uint32), o["bboxes"].view(np.uint32)), f"scene {s}"
        r = np.concatenate([ptamd.gen_shell_rays(2000, s, bmin, bmax, 1.3), interior_rays(2000, s, bmin, bmax)])
        rays_all.append(r)
        sc_all.append(np.full(len(r), s, np.uint32))
        expect.append(orc.trace(t, o, r))
    say(f"[batched c3] {len(picks)} scenes built bit-exact; tracing")
    perm = np.random.default_rng(5).permutation(sum(len(r) for r in rays_all))
    ids, t, uv = ctx.trace_closest_scenes(np.concatenate(rays_all)[perm], np.concatenate(sc_all)[perm])
    assert np.array_equal(ids, np.concatenate([e[0] for e in expect])[perm])
    assert np.array_equal(t.view(np.uint32), np.concatenate([e[1] for e in expect])[perm].view(np.uint32))
    assert np.array_equal(uv.view(np.uint32), np.concatenate([e[2] for e in expect])[perm].view(np.uint32))
    # three scenes of the batch rendered in one frame (one pass each)
    w, h = 48, 40
    eye, tgt = (0.5, 0.5, -1.5), (0.5, 0.5, 0.0)
    d0 = ptamd.camera_corners(eye, tgt, (0, 1, 0), 60.0, w / h)
    ctx.set_primary_rays(d0, eye)
    for s in picks[:3]:
        say(f"[batched c3] render scene {s}")
        ctx.reset_frames()
        ctx.render_passes(w, h, np.array([900 + s], np.uint64), None, np.array([s], np.uint32))
        g, spp = ctx.read_accumulator(w, h)
        o = np.zeros((h, w, 3))
        orc.render_pass(tris[offs[s]:offs[s + 1]], builds[s], infos, st, w, h, d0, eye, 0, 900 + s, o)
        assert spp == 1 and np.array_equal(g, o)


def test_traces_on_user_streams(ctx):
    """Device traces on two caller streams right after an asynchronous
    rebuild on the context stream: each stream has its own ray queue, so the
    two launches overlap without sharing hand-out counters, and both equal the
    oracle (outputs pre-filled with a sentinel, so a skipped ray shows)."""
    from ptamd import hip

    tris, infos = ptamd.gen_soup(200000, 21)
    ctx.upload_scene(tris, infos)
    bmin, bmax = ptamd.scene_bounds(tris)
    b = orc.build(tris, bmin, bmax)
    r1 = ptamd.gen_shell_rays(300000, 22, bmin, bmax, 1.5)
    r2 = interior_rays(300000, 23, bmin, bmax)
    o1, o2 = orc.trace(tris, b, r1), orc.trace(tris, b, r2)
    s1, s2 = hip.Stream(), hip.Stream()
    d1, d2 = hip.DeviceArray.from_host(r1), hip.DeviceArray.from_host(r2)
    i1, i2 = hip.DeviceArray(len(r1), np.uint32), hip.DeviceArray(len(r2), np.uint32)
    t1, t2 = hip.DeviceArray(len(r1), np.float32), hip.DeviceArray(len(r2), np.float32)
    try:
        for rep in range(3):
            say(f"[streams] rebuild + two concurrent traces, round {rep}")
            for a in (i1, i2, t1, t2):
                a.fill_bytes(0xAB)
            ctx.build_bvh(timed=False)
            ctx.trace_closest_device(d1.ptr, len(r1), i1.ptr, t1.ptr, 0, s1.handle)
            ctx.trace_closest_device(d2.ptr, len(r2), i2.ptr, t2.ptr, 0, s2.handle)
            s1.synchronize()
            s2.synchronize()
            assert np.array_equal(i1.to_host(), o1[0]) and np.array_equal(t1.to_host().view(np.uint32),
                                                                          o1[1].view(np.uint32))
            assert np.array_equal(i2.to_host(), o2[0]) and np.array_equal(t2.to_host().view(np.uint32),
                                                                          o2[1].view(np.uint32))
    finally:
        for a in (d1, d2, i1, i2, t1, t2):
            a.free()
        s1.destroy()
        s2.destroy()
    assert ctx.get_status() == 0


def test_verify_bvh_self_check(ctx):
    """pt_verify_bvh, the build's root-box self-check (BVHBuilder.cu:258-265,
    exact): passes for computed bounds, single and batched, fails naming the
    scene when the given enclosing box is larger than the geometry."""
    tris, infos = ptamd.gen_soup(100000, 31)
    ctx.upload_scene(tris, infos)
    assert ctx.verify_bvh() == 0
    ctx.build_bvh(timed=False)
    assert ctx.verify_bvh() == 0
    counts = [3000, 1, 0, 70000, 2]
    bt, bi, offs = ptamd.gen_batch(len(counts), 32, counts=counts)
    ctx.upload_scenes(bt, offs, bi)
    assert ctx.verify_bvh() == 0
    bmin, bmax = ptamd.scene_bounds(tris)
    ctx.upload_scene(tris, infos, bmin - 0.5, bmax + 0.5)
    with pytest.raises(ptamd.PtError) as e:
        ctx.verify_bvh()
    assert "scene 0" in str(e.value)
