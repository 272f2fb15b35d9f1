"""Pins the oracle's surface detail (textures, vertex colours, normal maps;
CPU only).

  tex2D          a constant texture reads c / 255 exactly; at texel centres of
                 a power-of-two texture the texel itself; a texture linear in
                 its texel indices is reproduced as the same linear function of
                 x = s W - 1/2, y = t H - 1/2 (what bilinear filtering is);
                 s + 1 and s - 1 read as s (wrap); row 0 is t = 0
  albedo         vertex colours r, g, b at the centroid -> (1, 1, 1) / sqrt 3
                 (computeColor normalises the interpolated colour); an albedo
                 texture replaces it
  normal map     texel (255, 255, 255) -> n = (1, 1, 1) in the tangent frame:
                 normalize(T + B + N); texel (255, 128, 128) tilts N to T within
                 the 2 * 128 / 255 - 1 offsets (closed form in float64)
"""
import numpy as np
import pytest

import oracle_py as orc
import ptamd


def info_rec(n=(0, 0, 1)):
    i = np.zeros(1, ptamd.INFO_DTYPE)
    i["n0"] = i["n1"] = i["n2"] = n
    i["albedo"] = (0.2, 0.4, 0.6)
    return i


def detail_rec(albedo_tex=ptamd.NO_TEXTURE, normal_tex=ptamd.NO_TEXTURE):
    d = np.zeros(1, ptamd.DETAIL_DTYPE)
    d["color"] = [[1, 0, 0], [0, 1, 0], [0, 0, 1]]
    d["uv"] = [[0.1, 0.2], [0.9, 0.3], [0.4, 0.8]]
    d["tangent"] = [[1, 0, 0]] * 3
    d["bitangent"] = [[0, 1, 0]] * 3
    d["albedo_tex"] = albedo_tex
    d["normal_tex"] = normal_tex
    return d


def test_constant_texture():
    img = np.zeros((4, 8, 4), np.uint8)
    img[..., :3] = (17, 200, 255)
    for s, t in ((0.0, 0.0), (0.37, 0.91), (-3.2, 7.5), (1e12, -1e12)):
        np.testing.assert_array_equal(orc.tex_sample(img, s, t),
                                      np.float32([17, 200, 255]) / np.float32(255))


def test_texel_centres_and_row_order():
    rng = np.random.default_rng(1)
    img = rng.integers(0, 256, (8, 16, 4), dtype=np.uint8)
    for j in range(8):
        for i in range(16):
            got = orc.tex_sample(img, (i + 0.5) / 16, (j + 0.5) / 8)
            np.testing.assert_array_equal(got, img[j, i, :3].astype(np.float32) / np.float32(255))


def test_bilinear_reproduces_linear_data():
    W, H = 16, 16
    j, i = np.mgrid[0:H, 0:W]
    img = np.zeros((H, W, 4), np.uint8)
    img[..., 0] = 10 * i
    img[..., 1] = 12 * j
    img[..., 2] = 5 * i + 6 * j
    rng = np.random.default_rng(2)
    for _ in range(200):
        s, t = rng.uniform(1.0 / W, 1.0 - 2.0 / W), rng.uniform(1.0 / H, 1.0 - 2.0 / H)
        x, y = np.float32(s) * W - 0.5, np.float32(t) * H - 0.5
        want = np.array([10 * x, 12 * y, 5 * x + 6 * y]) / 255.0
        np.testing.assert_allclose(orc.tex_sample(img, s, t), want, atol=2e-6)


def test_wrap():
    rng = np.random.default_rng(3)
    img = rng.integers(0, 256, (8, 8, 4), dtype=np.uint8)
    for s, t in ((0.25, 0.625), (0.0625, 0.5), (0.9375, 0.125)):
        a = orc.tex_sample(img, s, t)
        for ds, dt in ((1, 0), (-1, 0), (0, 2), (-3, -1)):
            np.testing.assert_array_equal(orc.tex_sample(img, s + ds, t + dt), a)


def test_vertex_colour_albedo():
    n, a = orc.shade_detail(info_rec(), detail_rec(), [], 1.0 / 3.0, 1.0 / 3.0)
    np.testing.assert_allclose(a, np.ones(3) / np.sqrt(3.0), rtol=1e-6)
    np.testing.assert_allclose(n, [0, 0, 1], atol=0)
    # without detail the pt_triangle_info albedo is used as given
    n, a = orc.shade_detail(info_rec(), None, [], 0.25, 0.25)
    np.testing.assert_array_equal(a, np.float32([0.2, 0.4, 0.6]))


def test_albedo_texture_replaces_colours():
    img = np.zeros((2, 2, 4), np.uint8)
    img[..., :3] = (51, 102, 204)
    _, a = orc.shade_detail(info_rec(), detail_rec(albedo_tex=0), [img], 0.2, 0.3)
    np.testing.assert_array_equal(a, np.float32([51, 102, 204]) / np.float32(255))


@pytest.mark.parametrize("texel,frame", [((255, 255, 255), (1.0, 1.0, 1.0)),
                                         ((255, 128, 128), (1.0, 2 * 128 / 255 - 1, 2 * 128 / 255 - 1))])
def test_normal_map(texel, frame):
    img = np.zeros((2, 2, 4), np.uint8)
    img[..., :3] = texel
    nrm, _ = orc.shade_detail(info_rec(), detail_rec(normal_tex=0), [img], 0.3, 0.3)
    want = np.array(frame) / np.linalg.norm(frame)  # T = x, B = y, N = z
    np.testing.assert_allclose(nrm, want, atol=1e-6)
