"""OBJ/MTL import (host only, no GPU).

Small OBJ files written by the tests pin the parser: fan triangulation order,
1-based and negative indices, v//vn and v/vt/vn corners, usemtl / mtllib with
Kd, Ke and Ns, the reference's defaults for missing materials, face normals
for corners without normals, the left-handed conversion (z mirrored, winding
reversed) and the unit-length albedo of the reference's computeColor.  The
Cornell box of the reference's assets, when present, loads with its light
(Ke 17 12 4) facing down into the box.
"""
import os

import numpy as np
import pytest

import ptamd

MTL = """# test materials
newmtl red
Kd 0.6 0.0 0.0
Ns 12.5
newmtl lamp
Kd 0.78 0.78 0.78
Ke 17 12 4
"""

OBJ = """mtllib test.mtl
v 0 0 0
v 1 0 0
v 1 0 1
v 0 0 1
v 0 2 0
vn 0 1 0
vt 0 0
vt 1 0
vt 1 1
vt 0 1
vt 0.5 0
vt 0.5 1
vt 0.25 0.75
usemtl red
f 1//1 4//1 3//1 2//1
usemtl lamp
f -5 -4 -1
usemtl missing
f 1/7/1 2 5
"""


@pytest.fixture()
def files(tmp_path):
    (tmp_path / "test.mtl").write_text(MTL)
    p = tmp_path / "test.obj"
    p.write_text(OBJ)
    return str(p)


def test_counts_and_fan_order(files):
    tris, infos = ptamd.load_obj(files, left_handed=False, normalize_albedo=False)
    assert len(tris) == 4 and list(tris["id"]) == [0, 1, 2, 3]
    # quad (1,4,3,2) -> (1,4,3), (1,3,2)
    np.testing.assert_array_equal(tris[0]["v0"], [0, 0, 0])
    np.testing.assert_array_equal(tris[0]["v1"], [0, 0, 1])
    np.testing.assert_array_equal(tris[0]["v2"], [1, 0, 1])
    np.testing.assert_array_equal(tris[1]["v1"], [1, 0, 1])
    np.testing.assert_array_equal(tris[1]["v2"], [1, 0, 0])
    # negative indices: -5 -4 -1 = vertices 1, 2, 5
    np.testing.assert_array_equal(tris[2]["v2"], [0, 2, 0])


def test_materials_and_defaults(files):
    tris, infos = ptamd.load_obj(files, left_handed=False, normalize_albedo=False)
    np.testing.assert_allclose(infos[0]["albedo"], [0.6, 0, 0])
    assert infos[0]["specularity"] == pytest.approx(12.5)
    np.testing.assert_array_equal(infos[0]["emission"], [0, 0, 0])
    np.testing.assert_allclose(infos[2]["emission"], [17, 12, 4])
    np.testing.assert_allclose(infos[2]["albedo"], [0.78, 0.78, 0.78])
    # unknown material: Kd 1, Ke 0, Ns 0
    np.testing.assert_array_equal(infos[3]["albedo"], [1, 1, 1])
    np.testing.assert_array_equal(infos[3]["emission"], [0, 0, 0])
    assert infos[3]["specularity"] == 0.0


def test_normals(files):
    tris, infos = ptamd.load_obj(files, left_handed=False, normalize_albedo=False)
    for k in ("n0", "n1", "n2"):
        np.testing.assert_array_equal(infos[0][k], [0, 1, 0])  # from vn
    # face (1,2,5): partial normals -> face normal cross(v1-v0, v2-v0)
    e1 = tris[3]["v1"] - tris[3]["v0"]
    e2 = tris[3]["v2"] - tris[3]["v0"]
    n = np.cross(e1, e2)
    np.testing.assert_allclose(infos[3]["n0"], n / np.linalg.norm(n), rtol=1e-6)


def test_left_handed_and_unit_albedo(files):
    rh, ri = ptamd.load_obj(files, left_handed=False, normalize_albedo=False)
    lh, li = ptamd.load_obj(files, left_handed=True, normalize_albedo=True)
    for a, b in ((lh["v0"], rh["v2"]), (lh["v1"], rh["v1"]), (lh["v2"], rh["v0"])):
        np.testing.assert_array_equal(a[:, :2], b[:, :2])
        np.testing.assert_array_equal(a[:, 2], -b[:, 2])
    # normals from the file are mirrored too; face normals stay consistent
    np.testing.assert_array_equal(li[0]["n0"], [0, 1, 0])
    e1 = lh[3]["v1"] - lh[3]["v0"]
    e2 = lh[3]["v2"] - lh[3]["v0"]
    n = np.cross(e1, e2)
    np.testing.assert_allclose(li[3]["n0"], n / np.linalg.norm(n), rtol=1e-6)
    r = ri[3]["n0"]
    np.testing.assert_allclose(li[3]["n0"], [r[0], r[1], -r[2]], atol=1e-6)
    np.testing.assert_allclose(li[0]["albedo"], [1, 0, 0])
    np.testing.assert_allclose(np.linalg.norm(li[2]["albedo"]), 1.0, rtol=1e-6)


def test_errors(tmp_path):
    with pytest.raises(ptamd.PtError) as e:
        ptamd.load_obj(str(tmp_path / "nope.obj"))
    assert "cannot open" in str(e.value)
    bad = tmp_path / "bad.obj"
    bad.write_text("v 0 0 0\nv 1 0 0\nf 1 2 3\n")
    with pytest.raises(ptamd.PtError) as e:
        ptamd.load_obj(str(bad))
    assert "out of range" in str(e.value)


GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
CORNELL = os.path.join(GOLDEN, "CornellBox-Original.obj")  # the reference's asset, copied as a fixture


def test_reference_bunny():
    tris, infos = ptamd.load_obj(os.path.join(GOLDEN, "bunny.obj"))
    assert len(tris) == 4968  # "face count = 4968" in the file's header
    assert np.all(infos["emission"] == 0)
    # No normals in the file: every vertex normal is the face's geometric
    # normal (PathTracingApp.cpp:74-77), a unit vector.
    np.testing.assert_allclose(np.linalg.norm(infos["n0"], axis=1), 1.0, rtol=1e-5)
    assert np.all(tris["id"] == np.arange(len(tris)))


def test_reference_cornell_box():
    tris, infos = ptamd.load_obj(CORNELL)
    assert len(tris) > 20
    lamp = np.where(infos["emission"].sum(axis=1) > 0)[0]
    assert len(lamp) >= 2
    np.testing.assert_allclose(infos[lamp[0]]["emission"], [17, 12, 4])
    bmin, bmax = ptamd.scene_bounds(tris)
    # the light hangs just under the ceiling and faces down into the box
    assert np.all(tris[lamp]["v0"][:, 1] > bmax[1] - 0.05)
    assert np.all(infos[lamp]["n0"][:, 1] < -0.99)
    # every albedo has unit length (computeColor normalises)
    np.testing.assert_allclose(np.linalg.norm(infos["albedo"], axis=1), 1.0, rtol=1e-5)
