"""Scene generators: sizes, determinism and orientation (CPU only)."""
import numpy as np

import ptamd


def normals(tris):
    n = np.cross(tris["v1"] - tris["v0"], tris["v2"] - tris["v0"]).astype(np.float64)
    return n


def test_sizes_and_determinism():
    t1, i1 = ptamd.gen_soup(1000, 5)
    t2, _ = ptamd.gen_soup(1000, 5)
    assert len(t1) == 1002 and np.array_equal(t1, t2)
    assert np.array_equal(t1["id"], np.arange(1002))
    s, si = ptamd.gen_sphere(30, 1, room=True)
    assert len(s) == 2 * 30 * 30 + 14
    h, _ = ptamd.gen_hairball(10, 20, 1, light=True)
    assert len(h) == 2 * 10 * 20 + 2
    c = ptamd.batch_counts(512, 3)
    assert c.min() >= 1024 and c.max() <= 262144


def test_orientation():
    s, info = ptamd.gen_sphere(40, 2, room=True)
    ns = 2 * 40 * 40
    sphere, room = s[:ns], s[ns:]
    n = normals(sphere)
    c = (sphere["v0"] + sphere["v1"] + sphere["v2"]) / 3
    area = np.linalg.norm(n, axis=1)
    ok = area > 1e-12
    assert np.all((n[ok] * c[ok]).sum(1) > 0)  # outward
    assert (~ok).sum() == 2 * 40  # one zero-area triangle per pole cell
    n = normals(room[:12])
    c = (room["v0"] + room["v1"] + room["v2"])[:12] / 3
    assert np.all((n * (-c)).sum(1) > 0)  # room faces inward
    ne = normals(room[12:])
    assert np.all(ne[:, 1] < 0) and np.all(info["emission"][-2:] == 10)  # emitter faces down
    # shading records carry the unit geometric normal
    g = normals(s[ok.nonzero()[0][:50]])
    g /= np.linalg.norm(g, axis=1, keepdims=True)
    assert np.allclose(info["n0"][ok.nonzero()[0][:50]], g, atol=1e-6)


def test_soup_emitter_faces_down():
    t, info = ptamd.gen_soup(10, 1)
    assert np.all(normals(t[-2:])[:, 1] < 0)
    assert np.all(info["emission"][-2:] == 10) and np.all(info["emission"][:-2] == 0)
