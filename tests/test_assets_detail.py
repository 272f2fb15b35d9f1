"""Asset import with surface detail and the TGA reader (host only, no GPU).

TGA: files written here in every supported variant (uncompressed and
run-length, true colour 24/32 bit and grey, bottom-up and top-down rows)
decode to the same RGBA image, bottom row first; the reference's Sponza
textures, when present, decode to what a direct read of their pixel bytes
gives.  OBJ: vt corners follow the reversed winding, colours come from Kd
(or the vertex's colour), each face's tangent frame satisfies
e1 = T du1 + B dv1 and e2 = T du2 + B dv2, textures are indexed once per
file and unsupported formats fail with a message.
"""
import os
import struct

import numpy as np
import pytest

import ptamd


def tga_bytes(img, itype, top_down=False):
    """img (h, w, 4) uint8 RGBA, row 0 = bottom."""
    h, w = img.shape[:2]
    grey = itype in (3, 11)
    bpp = 8 if grey else (32 if img[..., 3].min() < 255 else 24)
    rows = img[::-1] if top_down else img
    if grey:
        px = rows[..., 0].reshape(-1, 1)
    else:
        px = rows[..., [2, 1, 0, 3]] if bpp == 32 else rows[..., [2, 1, 0]]
        px = px.reshape(-1, bpp // 8)
    body = b""
    if itype in (2, 3):
        body = px.tobytes()
    else:  # run-length: alternate runs of repeats and raw packets
        k = 0
        n = len(px)
        while k < n:
            run = 1
            while k + run < n and run < 128 and np.array_equal(px[k + run], px[k]):
                run += 1
            if run >= 2:
                body += bytes([0x80 | (run - 1)]) + px[k].tobytes()
                k += run
            else:
                m = min(128, n - k)
                body += bytes([m - 1]) + px[k:k + m].tobytes()
                k += m
    hdr = struct.pack("<BBBHHBHHHHBB", 3, 0, itype, 0, 0, 0, 0, 0, w, h, bpp, 0x20 if top_down else 0)
    return hdr + b"abc" + body  # 3-byte image id field


def test_tga_variants(tmp_path):
    rng = np.random.default_rng(1)
    img = rng.integers(0, 256, (5, 7, 4), dtype=np.uint8)
    img[2:4, 1:6] = img[2, 1]  # runs for the RLE encoder
    opaque = img.copy()
    opaque[..., 3] = 255
    grey = opaque.copy()
    grey[..., 1] = grey[..., 2] = grey[..., 0]
    for itype, want in ((2, img), (2, opaque), (10, img), (10, opaque), (3, grey), (11, grey)):
        for top_down in (False, True):
            p = tmp_path / f"t{itype}_{int(top_down)}.tga"
            p.write_bytes(tga_bytes(want, itype, top_down))
            np.testing.assert_array_equal(ptamd.load_tga(p), want, err_msg=f"type {itype} top_down {top_down}")


def test_tga_errors(tmp_path):
    p = tmp_path / "bad.tga"
    p.write_bytes(b"\x00" * 10)
    with pytest.raises(ptamd.PtError, match="truncated"):
        ptamd.load_tga(p)
    img = np.zeros((2, 2, 4), np.uint8)
    p.write_bytes(tga_bytes(img, 2)[:-3])
    with pytest.raises(ptamd.PtError, match="truncated"):
        ptamd.load_tga(p)


SPONZA = "/root/reference/assets/textures/sponza_textures/textures"


@pytest.mark.skipif(not os.path.isdir(SPONZA), reason="reference textures not present")
@pytest.mark.parametrize("name", ["vase_round.tga", "chain_texture.tga", "background_ddn.tga"])
def test_reference_textures(name):
    path = os.path.join(SPONZA, name)
    raw = open(path, "rb").read()
    w, h = struct.unpack("<HH", raw[12:16])
    assert raw[2] == 2 and raw[16] == 24 and raw[17] & 0x20 == 0  # uncompressed, bottom-up
    px = np.frombuffer(raw[18 + raw[0]:18 + raw[0] + w * h * 3], np.uint8).reshape(h, w, 3)
    img = ptamd.load_tga(path)
    assert img.shape == (h, w, 4)
    np.testing.assert_array_equal(img[..., :3], px[..., ::-1])
    assert np.all(img[..., 3] == 255)


MTL = """newmtl tex
Kd 0.5 0.25 0.125
map_Kd maps\\albedo.tga
map_bump -bm 0.8 maps/nrm.tga
newmtl plain
Kd 0.1 0.2 0.3
"""

OBJ = """mtllib scene.mtl
v 0 0 0
v 2 0 0
v 2 1 0 0.9 0.8 0.7
v 0 1 0
v 0 0 -1
vt 0 0
vt 1 0
vt 1 1
vt 0 1
usemtl tex
f 1/1 2/2 3/3 4/4
usemtl plain
f 1 4 5
"""


@pytest.fixture()
def scene_files(tmp_path):
    (tmp_path / "maps").mkdir()
    rng = np.random.default_rng(4)
    a = rng.integers(0, 256, (4, 8, 4), dtype=np.uint8)
    a[..., 3] = 255
    (tmp_path / "maps" / "albedo.tga").write_bytes(tga_bytes(a, 2))
    n = np.full((2, 2, 4), 255, np.uint8)
    n[..., :3] = (128, 128, 255)
    (tmp_path / "maps" / "nrm.tga").write_bytes(tga_bytes(n, 10))
    (tmp_path / "scene.mtl").write_text(MTL)
    p = tmp_path / "scene.obj"
    p.write_text(OBJ)
    return str(p), a, n


@pytest.mark.parametrize("left_handed", [False, True])
def test_obj_detail(scene_files, left_handed):
    path, a, n = scene_files
    tris, infos, det, texs = ptamd.load_obj_asset(path, left_handed=left_handed)
    assert len(tris) == 3 and len(texs) == 2
    np.testing.assert_array_equal(texs[0], a)
    np.testing.assert_array_equal(texs[1], n)
    assert list(det["albedo_tex"]) == [0, 0, ptamd.NO_TEXTURE]
    assert list(det["normal_tex"]) == [1, 1, ptamd.NO_TEXTURE]
    # the same triangles and infos as pt_load_obj
    t2, i2 = ptamd.load_obj(path, left_handed=left_handed)
    np.testing.assert_array_equal(tris, t2)
    np.testing.assert_array_equal(infos, i2)
    # uv per corner follows the corner order (reversed with left_handed)
    quad_uv = {(0, 0, 0): (0, 0), (2, 0, 0): (1, 0), (2, 1, 0): (1, 1), (0, 1, 0): (0, 1)}
    for k in range(2):
        for j, v in enumerate(("v0", "v1", "v2")):
            key = tuple(float(c) for c in tris[k][v] * [1, 1, -1 if left_handed else 1])
            np.testing.assert_array_equal(det[k]["uv"][j], quad_uv[key])
    # colours: Kd, or the vertex's own colour (vertex 3 carries one)
    for k in range(2):
        for j, v in enumerate(("v0", "v1", "v2")):
            want = (0.9, 0.8, 0.7) if tuple(tris[k][v][:2]) == (2, 1) else (0.5, 0.25, 0.125)
            np.testing.assert_allclose(det[k]["color"][j], want, rtol=1e-7)
    np.testing.assert_allclose(det[2]["color"], [[0.1, 0.2, 0.3]] * 3, rtol=1e-7)
    # tangent frame: the edges expressed in (T, B) reproduce the uv deltas
    for k in range(2):
        p = [tris[k][v].astype(np.float64) for v in ("v0", "v1", "v2")]
        uv = det[k]["uv"].astype(np.float64)
        T, B = det[k]["tangent"][0].astype(np.float64), det[k]["bitangent"][0].astype(np.float64)
        for e in (1, 2):
            d = uv[e] - uv[0]
            np.testing.assert_allclose(p[e] - p[0], T * d[0] + B * d[1], atol=1e-6)
    # the third face has no uvs: T along its first edge, B = N x T
    p = [tris[2][v].astype(np.float64) for v in ("v0", "v1", "v2")]
    T = det[2]["tangent"][0].astype(np.float64)
    np.testing.assert_allclose(T, (p[1] - p[0]) / np.linalg.norm(p[1] - p[0]), atol=1e-7)
    N = infos[2]["n0"].astype(np.float64)
    np.testing.assert_allclose(det[2]["bitangent"][0], np.cross(N, T), atol=1e-6)


def test_unsupported_texture_format(tmp_path):
    (tmp_path / "m.mtl").write_text("newmtl x\nmap_Kd wood.png\n")
    (tmp_path / "o.obj").write_text("mtllib m.mtl\nv 0 0 0\nv 1 0 0\nv 0 1 0\nusemtl x\nf 1 2 3\n")
    with pytest.raises(ptamd.PtError, match="TGA only"):
        ptamd.load_obj_asset(tmp_path / "o.obj")
