"""Pins the oracle's LBVH build (CPU only).

What pins what:
  or_encode30     naive bit loop (3i, 3i+1, 3i+2 interleave) on random and edge inputs
  or_quantize     integer arithmetic on exactly representable inputs; NaN/inf/negatives
  or_morton       numpy restatement of centroid -> quantize -> interleave (float32 ops
                  in the reference's order) + known answers at the box corners
  or_sort         std::stable_sort contract: sorted, a bijection, ties in input order
                  (checked against numpy's stable argsort)
  or_radix_tree   an independent top-down binary radix tree over the augmented
                  64-bit keys (key << 32 | index), plus Karras invariants
  or_refit        each internal box == min/max over the vertices of its leaf range;
                  root box == enclosing box exactly (the reference's only check,
                  BVHBuilder.cu:258-265, made exact)
"""
import numpy as np
import pytest

import oracle_py as orc
import ptamd


def naive_encode(x, y, z):
    r = 0
    for b in range(10):
        r |= ((x >> b) & 1) << (3 * b)
        r |= ((y >> b) & 1) << (3 * b + 1)
        r |= ((z >> b) & 1) << (3 * b + 2)
    return r


def test_encode30_matches_bit_loop():
    rng = np.random.default_rng(1)
    cases = [(0, 0, 0), (1023, 1023, 1023), (1, 0, 0), (0, 1, 0), (0, 0, 1), (512, 0, 0)]
    cases += [tuple(int(v) for v in rng.integers(0, 1024, 3)) for _ in range(2000)]
    for x, y, z in cases:
        assert orc.encode30(x, y, z) == naive_encode(x, y, z)
    assert orc.encode30(1023, 1023, 1023) == 0x3FFFFFFF


def test_quantize_cases():
    # exactly representable x: int(x * 1024) in integer arithmetic
    for k in range(0, 1025):
        x = k / 1024.0
        assert orc.quantize(x) == min(k, 1023)
    assert orc.quantize(-0.5) == 0
    assert orc.quantize(-1e30) == 0
    assert orc.quantize(float("nan")) == 0
    assert orc.quantize(float("inf")) == 1023
    assert orc.quantize(float("-inf")) == 0
    assert orc.quantize(2.0) == 1023
    assert orc.quantize(np.nextafter(np.float32(1 / 1024), np.float32(0))) == 0


def np_morton(tris, bmin, bmax):
    """numpy float32 restatement of morton_encoder (BVHBuilder.cu:12-32)."""
    v0, v1, v2 = (tris[f].astype(np.float32) for f in ("v0", "v1", "v2"))
    with np.errstate(all="ignore"):
        c = ((v0 + v1) + v2) / np.float32(3.0)
        inv = np.float32(1.0) / (np.asarray(bmax, np.float32) - np.asarray(bmin, np.float32))
        s = ((c - np.asarray(bmin, np.float32)) * inv) * np.float32(1024.0)
    q = np.where(np.isnan(s), 0, np.clip(np.trunc(np.nan_to_num(s, posinf=2e9, neginf=-2e9)), 0, 1023))
    q = q.astype(np.uint32)
    return np.array([naive_encode(*row) for row in q], np.uint32)


def test_morton_matches_numpy_restatement():
    tris, _ = ptamd.gen_soup(3000, 11)
    bmin, bmax = ptamd.scene_bounds(tris)
    assert np.array_equal(orc.morton(tris, bmin, bmax), np_morton(tris, bmin, bmax))


def test_morton_known_corners():
    tris = np.zeros(2, ptamd.TRI_DTYPE)
    tris[0]["v0"] = tris[0]["v1"] = tris[0]["v2"] = (0.0, 0.0, 0.0)
    tris[1]["v0"] = tris[1]["v1"] = tris[1]["v2"] = (1.0, 1.0, 1.0)
    keys = orc.morton(tris, (0, 0, 0), (1, 1, 1))
    assert keys[0] == 0 and keys[1] == 0x3FFFFFFF


def test_morton_flat_scene_is_defined():
    # zero extent on y: invScale = inf; the quantisation saturates instead of UB
    tris, _ = ptamd.gen_soup(500, 3, emitter=False)
    tris["v0"][:, 1] = tris["v1"][:, 1] = tris["v2"][:, 1] = np.float32(0.1)
    bmin, bmax = ptamd.scene_bounds(tris)
    assert bmin[1] == bmax[1]
    keys = orc.morton(tris, bmin, bmax)
    assert np.array_equal(keys, np_morton(tris, bmin, bmax))


def test_sort_is_stable_bijection():
    rng = np.random.default_rng(5)
    for n in (1, 2, 7, 4096, 10000):
        keys = rng.integers(0, 64, n).astype(np.uint32)  # many duplicates
        sk, perm = orc.sort(keys)
        ref = np.argsort(keys, kind="stable")
        assert np.array_equal(perm, ref)
        assert np.array_equal(sk, keys[ref])


def topdown_tree(keys):
    """Independent binary radix tree over augmented keys (key << 32 | index),
    split at the highest differing bit; internal node numbering by the split
    index, as Karras' layout fixes it (root 0, children gamma / gamma + 1)."""
    n = len(keys)
    aug = [(int(k) << 32) | i for i, k in enumerate(keys)]
    children = np.zeros(n - 1, np.uint32)
    pl = np.zeros(n, np.uint32)
    pi = np.zeros(n - 1, np.uint32)
    pi[0] = 0xFFFFFFFF
    stack = [(0, 0, n - 1)]
    while stack:
        node, lo, hi = stack.pop()
        bit = (aug[lo] ^ aug[hi]).bit_length() - 1
        g = lo
        while g + 1 < hi and not (aug[g + 1] >> bit) & 1:
            g += 1
        left_internal = lo != g
        right_internal = hi != g + 1
        children[node] = (g << 2) | int(left_internal) | (int(right_internal) << 1)
        if left_internal:
            pi[g] = node
            stack.append((g, lo, g))
        else:
            pl[g] = node
        if right_internal:
            pi[g + 1] = node
            stack.append((g + 1, g + 1, hi))
        else:
            pl[g + 1] = node
    return children, pl, pi


@pytest.mark.parametrize("n,dist", [(2, "rand"), (3, "rand"), (17, "rand"), (500, "rand"),
                                    (3000, "rand"), (64, "dup"), (777, "dup"), (40, "chain"),
                                    (300, "small")])
def test_radix_tree_matches_topdown(n, dist):
    rng = np.random.default_rng(n)
    if dist == "rand":
        keys = rng.integers(0, 1 << 30, n)
    elif dist == "dup":
        keys = np.full(n, 12345)
    elif dist == "chain":
        keys = np.array([(1 << (i % 30)) for i in range(n)])
    else:
        keys = rng.integers(0, 8, n)
    keys = np.sort(keys).astype(np.uint32)
    ch, pl, pi = orc.radix_tree(keys)
    tch, tpl, tpi = topdown_tree(keys)
    assert np.array_equal(ch, tch)
    assert np.array_equal(pl, tpl)
    assert np.array_equal(pi, tpi)


def test_radix_tree_invariants():
    tris, _ = ptamd.gen_soup(20000, 21)
    bmin, bmax = ptamd.scene_bounds(tris)
    b = orc.build(tris, bmin, bmax)
    n = len(tris)
    ch, pl, pi = b["children"], b["parents_leaf"], b["parents_internal"]
    left = ch >> 2
    li = (ch & 1) != 0
    ri = (ch & 2) != 0
    # every leaf and every non-root internal node appears exactly once as a child
    leaf_refs = np.concatenate([left[~li], (left + 1)[~ri]])
    int_refs = np.concatenate([left[li], (left + 1)[ri]])
    assert np.array_equal(np.sort(leaf_refs), np.arange(n))
    assert np.array_equal(np.sort(int_refs), np.arange(1, n - 1))
    assert pi[0] == 0xFFFFFFFF
    nodes = np.arange(n - 1)
    assert np.all(pl[left[~li]] == nodes[~li]) and np.all(pl[(left + 1)[~ri]] == nodes[~ri])
    assert np.all(pi[left[li]] == nodes[li]) and np.all(pi[(left + 1)[ri]] == nodes[ri])


def leaf_ranges(children, n):
    """[lo, hi] leaf range of every internal node, from the child words."""
    lo = np.zeros(n - 1, np.int64)
    hi = np.zeros(n - 1, np.int64)
    order = []
    stack = [0]
    while stack:
        v = stack.pop()
        order.append(v)
        w = int(children[v])
        for c, internal in ((w >> 2, w & 1), ((w >> 2) + 1, w & 2)):
            if internal:
                stack.append(c)
    for v in reversed(order):
        w = int(children[v])
        l, r = w >> 2, (w >> 2) + 1
        lo[v] = lo[l] if w & 1 else l
        hi[v] = hi[r] if w & 2 else r
    return lo, hi


@pytest.mark.parametrize("gen", ["soup", "sphere", "hair"])
def test_refit_boxes_are_leaf_unions(gen):
    if gen == "soup":
        tris, _ = ptamd.gen_soup(5000, 9)
    elif gen == "sphere":
        tris, _ = ptamd.gen_sphere(40, 9, room=True)
    else:
        tris, _ = ptamd.gen_hairball(30, 40, 9, light=True)
    bmin, bmax = ptamd.scene_bounds(tris)
    b = orc.build(tris, bmin, bmax)
    n = len(tris)
    st = tris[b["perm"]]
    verts = np.stack([st["v0"], st["v1"], st["v2"]], axis=1)  # (n, 3, 3)
    tmin = verts.min(axis=1)
    tmax = verts.max(axis=1)
    lo, hi = leaf_ranges(b["children"], n)
    for v in range(0, n - 1, max(1, (n - 1) // 400)):
        exp = np.concatenate([tmin[lo[v]:hi[v] + 1].min(axis=0), tmax[lo[v]:hi[v] + 1].max(axis=0)])
        assert np.array_equal(b["bboxes"][v], exp)
    # root == enclosing box, exactly
    assert np.array_equal(b["bboxes"][0], np.concatenate([bmin, bmax]))
    # Morton keys sorted, perm a permutation
    assert np.all(np.diff(b["morton"].astype(np.int64)) >= 0)
    assert np.array_equal(np.sort(b["perm"]), np.arange(n))


def test_tiny_builds():
    tris, _ = ptamd.gen_soup(3, 2, emitter=False)
    for n in (0, 1, 2, 3):
        t = tris[:n]
        bmin, bmax = ptamd.scene_bounds(t) if n else (np.zeros(3), np.ones(3))
        b = orc.build(t, bmin, bmax)
        assert len(b["children"]) == max(n - 1, 0)
        if n == 1:
            assert b["parents_leaf"][0] == 0xFFFFFFFF
        if n == 2:
            assert b["children"][0] == 0  # two leaves 0 and 1
            assert np.array_equal(b["bboxes"][0], np.concatenate([bmin, bmax]))
