"""A scene whose LBVH is as deep as 30-bit Morton keys allow, plus a
balanced subtree of equal keys below it, and rays that visit every node:
the deferred-node stack of the closest-hit traversal reaches 40+ entries
(deeper than every LDS stack setting, so the scratch spill runs).

Keys: a pair of triangles at each single-bit key 1 << b (b = 0..29, Morton
bit 3m + axis = coordinate bit m of that axis) and `group` triangles at key
0, so the radix tree is a 30-level comb whose spine ends in a balanced
subtree of log2(group) levels over the equal keys.  Every triangle is large
(about 1.2 units along each axis) with its centroid in its key's cell of the
unit box, so every node's box contains the cube R = [0.34, 0.36]^3: rays
starting in R enter every box (t_near < 0 <= best distance), visit every node
and push a deferred sibling at every level of the spine.  The enclosing box
passed to the build is the unit cube, which fixes the keys."""
import numpy as np

import ptamd

BOX_MIN = np.zeros(3, np.float32)
BOX_MAX = np.ones(3, np.float32)


def _tri(c, r1, r2):
    d0 = np.array([-0.6, -0.6, r1])
    d1 = np.array([0.6, r2, -0.6])
    d2 = -(d0 + d1)
    return [c + d0, c + d1, c + d2]


def deep_scene(group=4096, seed=0):
    rng = np.random.default_rng(seed)
    cells = [np.zeros(3)] * group
    for b in range(30):
        q = np.zeros(3)
        q[b % 3] = float(1 << (b // 3))
        cells += [q, q]
    n = len(cells)
    tris = np.zeros(n, ptamd.TRI_DTYPE)
    infos = np.zeros(n, ptamd.INFO_DTYPE)
    order = rng.permutation(n)  # input order unrelated to key order
    for k, i in enumerate(order):
        c = (cells[i] + 0.5) / 1024.0
        v = _tri(c, rng.uniform(-0.3, 0.2), rng.uniform(-0.3, 0.2))
        tris[k]["v0"], tris[k]["v1"], tris[k]["v2"] = v
        tris[k]["id"] = k
        nrm = np.cross(v[1] - v[0], v[2] - v[0])
        nrm /= np.linalg.norm(nrm)
        infos[k]["n0"] = infos[k]["n1"] = infos[k]["n2"] = nrm
        infos[k]["albedo"] = 0.5
    return tris, infos


def deep_rays(n, seed=1):
    """Origins in R, directions uniform on the sphere; ray 0 along (1,1,1)."""
    rng = np.random.default_rng(seed)
    o = rng.uniform(0.34, 0.36, (n, 3))
    d = rng.normal(size=(n, 3))
    d[0] = (1.0, 1.0, 1.0)
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    return np.concatenate([o, d], axis=1).astype(np.float32)
