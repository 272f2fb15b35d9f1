"""The deepest trees the build makes (tests/deep_scene.py): the oracle's
keys, tree depth and deferred-node stack depth on them.

The keys are pinned by construction (one set bit per spine pair); the tree
depth is measured from the children arrays; the traversal's stack-depth
counter is pinned by the spine: a ray from R along (1,1,1) enters the
spine's smaller-key child first at every level (its box starts lower on
every axis), so it pushes one sibling per spine level and one per internal
level of the equal-key subtree."""
import numpy as np

import oracle_py as orc
import ptamd
from deep_scene import BOX_MAX, BOX_MIN, deep_rays, deep_scene


def tree_depth(children):
    """Depth (internal levels) of the radix tree, root = 0."""
    m = len(children)
    depth = np.zeros(m, np.int64)
    best = 1
    stack = [0]
    depth[0] = 1
    while stack:
        i = stack.pop()
        w = int(children[i])
        l = w >> 2
        for k, internal in ((l, w & 1), (l + 1, w & 2)):
            if internal:
                depth[k] = depth[i] + 1
                best = max(best, depth[k])
                stack.append(k)
    return best


def test_deep_scene_keys_and_depth():
    group = 4096
    tris, _ = deep_scene(group)
    o = orc.build(tris, BOX_MIN, BOX_MAX)
    keys = o["morton"]
    want = np.array([0] * group + [1 << b for b in range(30) for _ in range(2)], np.uint32)
    assert np.array_equal(keys, want)
    # 30 spine levels, then the equal-key subtree (log2(group) internal levels)
    assert tree_depth(o["children"]) == 30 + 12


def test_deep_scene_stack_depth():
    group = 4096
    tris, _ = deep_scene(group)
    o = orc.build(tris, BOX_MIN, BOX_MAX)
    rays = deep_rays(512)
    ids, t, _, c = orc.trace(tris, o, rays[:1], counters=True)
    # the diagonal ray: one deferred sibling per spine level and per level of
    # the equal-key subtree that has two internal children
    assert int(c[3]) == 30 + 11
    assert int(c[0]) == len(tris) - 1  # every internal node visited
    ids, t, _, c = orc.trace(tris, o, rays, counters=True)
    assert 40 <= int(c[3]) <= 63
    bids, bt, _ = orc.brute(tris, rays)
    assert np.array_equal(ids, bids) and np.array_equal(t.view(np.uint32), bt.view(np.uint32))
    assert (ids != ptamd.INVALID).mean() > 0.5
