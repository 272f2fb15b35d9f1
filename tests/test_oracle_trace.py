"""Pins the oracle's closest-hit query (CPU only).

What pins what:
  or_ray_triangle  known answers on axis-aligned triangles: inclusive edges,
                   t = 0 accepted, back hits rejected, parallel and zero-area
                   rejected, the reference's absolute det epsilon
                   (intersection.h:38) rejecting tiny triangles
  or_ray_aabb      known answers, including glm's NaN rule (NaN -> second
                   operand) for rays lying in a slab face with zero direction
  or_trace         brute force over all triangles with the same test and the
                   same lowest-id tie rule, on soups, a UV sphere in a room
                   and a hair ball, with shell rays, interior rays and rays
                   with zero direction components; duplicate triangles and
                   shared edges for the tie rule; visit counters consistent
"""
import numpy as np
import pytest

import oracle_py as orc
import ptamd

INF = float("inf")


def test_mt_known_answers():
    p0, p1, p2 = (0, 0, 0), (1, 0, 0), (0, 1, 0)
    hit, u, v, t = orc.ray_triangle((0.25, 0.25, -1), (0, 0, 1), p0, p1, p2)
    assert hit and (u, v, t) == (0.25, 0.25, 1.0)
    # barycentrics: u weighs v1, v weighs v2
    hit, u, v, t = orc.ray_triangle((1.0, 0.0, -2), (0, 0, 1), p0, p1, p2)
    assert hit and (u, v, t) == (1.0, 0.0, 2.0)  # vertex v1, edge inclusive
    hit, u, v, t = orc.ray_triangle((0.5, 0.5, -1), (0, 0, 1), p0, p1, p2)
    assert hit and u + v == 1.0  # hypotenuse inclusive
    assert not orc.ray_triangle((0.6, 0.5, -1), (0, 0, 1), p0, p1, p2)[0]
    assert not orc.ray_triangle((-0.01, 0.5, -1), (0, 0, 1), p0, p1, p2)[0]
    hit, u, v, t = orc.ray_triangle((0.2, 0.2, 0), (0, 0, 1), p0, p1, p2)
    assert hit and t == 0.0  # t = 0 accepted
    assert not orc.ray_triangle((0.2, 0.2, 1), (0, 0, 1), p0, p1, p2)[0]  # behind
    assert not orc.ray_triangle((0.2, 0.2, -1), (1, 0, 0), p0, p1, p2)[0]  # parallel
    assert not orc.ray_triangle((0.2, 0.2, -1), (0, 0, 1), p0, p1, p1)[0]  # zero area
    # absolute epsilon: a 1e-4-sized triangle has |det| = 1e-8 < FLT_EPSILON
    s = 1e-4
    assert not orc.ray_triangle((0.2 * s, 0.2 * s, -1), (0, 0, 1), p0, (s, 0, 0), (0, s, 0))[0]
    # ...but the same triangle hit along a longer direction vector is accepted
    assert orc.ray_triangle((0.2 * s, 0.2 * s, -1), (0, 0, 1e4), p0, (s, 0, 0), (0, s, 0))[0]


def test_slab_known_answers():
    lo, hi = (0, 0, 0), (1, 1, 1)
    assert orc.ray_aabb((-1, 0.5, 0.5), (1, 0, 0), lo, hi) == (True, 1.0)
    hit, t = orc.ray_aabb((0.5, 0.5, 0.5), (0, 1, 0), lo, hi)
    assert hit and t == -0.5  # inside: the near distance is negative
    assert not orc.ray_aabb((-1, 0.5, 0.5), (-1, 0, 0), lo, hi)[0]  # pointing away
    assert not orc.ray_aabb((-1, 2.0, 0.5), (1, 0, 0), lo, hi)[0]  # parallel, outside
    # In the y = 0 face with d.y = 0: (0 - 0) * inf = NaN; glm min/max return
    # the second operand, which makes the near distance +inf -> miss ...
    assert not orc.ray_aabb((-1, 0.0, 0.5), (1, 0, 0), lo, hi)[0]
    # ... while in the y = 1 face the NaN reaches the near distance as -inf -> hit.
    assert orc.ray_aabb((-1, 1.0, 0.5), (1, 0, 0), lo, hi) == (True, -INF)
    # box behind the origin
    assert not orc.ray_aabb((2, 0.5, 0.5), (1, 0, 0), lo, hi)[0]
    # touching at a single point: near == far is a hit
    assert orc.ray_aabb((-1, -1, 0.5), (1, 1, 0), lo, hi)[0]


def _interior_rays(n, seed, bmin, bmax):
    rng = np.random.default_rng(seed)
    o = bmin + (bmax - bmin) * rng.random((n, 3), dtype=np.float32)
    d = rng.normal(size=(n, 3)).astype(np.float32)
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    return np.concatenate([o, d], axis=1).astype(np.float32)


def _axis_rays(n, seed, bmin, bmax):
    """Rays with zero direction components (one or two)."""
    rng = np.random.default_rng(seed)
    r = _interior_rays(n, seed, bmin, bmax)
    for i in range(n):
        k = rng.integers(0, 3)
        if i % 2:
            r[i, 3:] = 0.0
            r[i, 3 + k] = 1.0 if rng.random() < 0.5 else -1.0
        else:
            r[i, 3 + k] = 0.0
    return r


SCENES = {
    "soup": lambda: ptamd.gen_soup(3000, 7),
    "sphere": lambda: ptamd.gen_sphere(24, 7, room=True),
    "hair": lambda: ptamd.gen_hairball(40, 30, 7, light=True),
}


@pytest.mark.parametrize("scene", list(SCENES))
def test_trace_equals_brute_force(scene):
    tris, _ = SCENES[scene]()
    bmin, bmax = ptamd.scene_bounds(tris)
    b = orc.build(tris, bmin, bmax)
    rays = np.concatenate([ptamd.gen_shell_rays(1500, 3, bmin, bmax, 1.5),
                           _interior_rays(1500, 4, bmin, bmax),
                           _axis_rays(600, 5, bmin, bmax)])
    ids, t, uv = orc.trace(tris, b, rays)
    bids, bt, buv = orc.brute(tris, rays)
    assert np.array_equal(ids, bids)
    assert np.array_equal(t, bt)
    assert np.array_equal(uv, buv)
    hits = ids != ptamd.INVALID
    assert hits.sum() > 100
    assert np.all(t[~hits] == ptamd.FLT_MAX)


def test_tie_rule_lowest_id():
    # the same triangle three times with ids 7, 3, 5 plus clutter
    tris, _ = ptamd.gen_soup(200, 17, emitter=False)
    base = tris[10].copy()
    dup = np.zeros(3, ptamd.TRI_DTYPE)
    for k, tid in enumerate((1000, 1001, 1002)):
        dup[k] = base
        dup[k]["id"] = tid
    dup[1]["id"] = 999  # lowest of the duplicates, placed second
    allt = np.concatenate([tris, dup])
    bmin, bmax = ptamd.scene_bounds(allt)
    b = orc.build(allt, bmin, bmax)
    c = (base["v0"] + base["v1"] + base["v2"]) / 3
    n = np.cross(base["v1"] - base["v0"], base["v2"] - base["v0"])
    n /= np.linalg.norm(n)
    ray = np.concatenate([c - 0.5 * n, n]).astype(np.float32)[None]
    ids, t, _ = orc.trace(allt, b, ray)
    bids, bt, _ = orc.brute(allt, ray)
    assert ids[0] == bids[0]
    assert ids[0] in (10, 999)
    assert ids[0] == min(10, 999)


def test_tie_on_shared_edge():
    # two triangles sharing the edge x = 0.5 in the plane z = 0; a ray through
    # the edge point hits both at the same t
    q = np.zeros(2, ptamd.TRI_DTYPE)
    q[0]["v0"], q[0]["v1"], q[0]["v2"], q[0]["id"] = (0, 0, 0), (0.5, 0, 0), (0.5, 1, 0), 4
    q[1]["v0"], q[1]["v1"], q[1]["v2"], q[1]["id"] = (0.5, 0, 0), (1, 0, 0), (0.5, 1, 0), 2
    bmin, bmax = ptamd.scene_bounds(q)
    b = orc.build(q, bmin, bmax)
    ray = np.array([[0.5, 0.25, -1, 0, 0, 1]], np.float32)
    ids, t, _ = orc.trace(q, b, ray)
    assert t[0] == 1.0
    assert ids[0] == 2


def test_counters_consistent():
    tris, _ = ptamd.gen_soup(4000, 8)
    bmin, bmax = ptamd.scene_bounds(tris)
    b = orc.build(tris, bmin, bmax)
    rays = ptamd.gen_shell_rays(2000, 9, bmin, bmax, 1.5)
    _, _, _, c = orc.trace(tris, b, rays, counters=True)
    iters, boxes, tri_tests, depth = (int(x) for x in c)
    assert iters >= len(rays)  # the root is always visited
    assert 1 <= depth <= 63
    assert boxes + tri_tests == 2 * iters
    # threads do not change the totals
    _, _, _, c1 = orc.trace(tris, b, rays, counters=True, threads=1)
    assert np.array_equal(c, c1)


def test_tiny_scenes_trace():
    tris, _ = ptamd.gen_soup(2, 1, emitter=False)
    rays = _interior_rays(200, 2, np.zeros(3, np.float32), np.ones(3, np.float32))
    for n in (0, 1, 2):
        t = tris[:n]
        bmin, bmax = ptamd.scene_bounds(t) if n else (np.zeros(3), np.ones(3))
        b = orc.build(t, bmin, bmax)
        ids, tt, _ = orc.trace(t, b, rays)
        bids, btt, _ = orc.brute(t, rays)
        assert np.array_equal(ids, bids) and np.array_equal(tt, btt)
