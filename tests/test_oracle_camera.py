"""Corner rays (row a14): the oracle's restatement of screenToRay
(PathTracingApp.cpp:217-242) pinned by what the matrices fix geometrically,
then the library's pt_camera_corners / pt_camera_corners_basis checked
against it.

Pinning of or_screen_to_ray (a general 4x4 inverse of translate(jitter) *
perspective * view, applied to two NDC points), for glm's left-handed and
right-handed perspective:
  - the image centre without jitter unprojects to +look (left-handed) or
    -look (right-handed): a transposed view matrix, a wrong projection sign or
    a dropped position term fails;
  - every direction lies on +-look + nx * aspect * tan(fovy/2) * right +
    ny * tan(fovy/2) * up with nx, ny the NDC point minus the jitter;
  - a jitter of 2/w (one pixel) equals the unjittered ray one pixel to the left;
  - the reference's setPerspective(45.0f) is 45 radians in glm::perspective:
    the top edge's half angle is atan(tan(22.5)) (29.15 degrees);
  - the camera position and the depth of the screen point do not change the
    direction.
The app's Cornell preset (position (-0.324, 4.428, -13), no rotation,
PathTracingApp.cpp:425-436) frames the box through the left-handed
projection and looks away from it through the right-handed one.
"""
import math

import numpy as np
import pytest

import oracle_py as orc
import ptamd


def rand_rotation(rng):
    q, r = np.linalg.qr(rng.normal(size=(3, 3)))
    q = q * np.sign(np.diag(r))
    if np.linalg.det(q) < 0:
        q[:, 0] = -q[:, 0]
    return q.astype(np.float32)  # columns: right, up, look


def ray(R, pos, fovy, w, h, jx, jy, x, y, z=0.3, zn=0.3, zf=30.0, lh=True):
    return orc.screen_to_ray(R.T, pos, fovy, w, h, zn, zf, jx, jy, (x, y, z), lh).astype(np.float64)


def unit(v):
    v = np.asarray(v, np.float64)
    return v / np.linalg.norm(v)


@pytest.mark.parametrize("lh", [True, False])
def test_centre_is_look(lh):
    rng = np.random.default_rng(1)
    for _ in range(20):
        R = rand_rotation(rng)
        pos = rng.uniform(-20, 20, 3)
        d = ray(R, pos, 1.0, 640, 480, 0.0, 0.0, 320, 240, lh=lh)
        assert np.allclose(d, (1 if lh else -1) * R[:, 2], atol=2e-6), (d, R)


@pytest.mark.parametrize("lh", [True, False])
def test_directions_follow_the_pinhole_model(lh):
    rng = np.random.default_rng(2)
    for _ in range(50):
        R = rand_rotation(rng)
        pos = rng.uniform(-20, 20, 3)
        w, h = int(rng.integers(16, 2000)), int(rng.integers(16, 2000))
        fovy = float(rng.uniform(0.3, 2.5))
        jx, jy = rng.uniform(-2.0 / w, 2.0 / w), rng.uniform(-2.0 / h, 2.0 / h)
        x, y = rng.uniform(0, w), rng.uniform(0, h)
        d = ray(R, pos, fovy, w, h, jx, jy, x, y, lh=lh)
        t = float(np.tan(np.float32(fovy) / np.float32(2)))
        nx, ny = 2 * x / w - 1 - jx, 2 * y / h - 1 - jy
        want = unit((1 if lh else -1) * R[:, 2] + nx * (w / h) * t * R[:, 0] + ny * t * R[:, 1])
        assert np.allclose(d, want, atol=3e-6)


def test_one_pixel_jitter_is_one_pixel():
    rng = np.random.default_rng(3)
    R = rand_rotation(rng)
    w, h = 200, 100
    for x, y in [(10, 10), (150, 20), (100, 50)]:
        a = ray(R, (1, 2, 3), 0.9, w, h, 2.0 / w, 0.0, x, y)
        b = ray(R, (1, 2, 3), 0.9, w, h, 0.0, 0.0, x - 1, y)
        assert np.allclose(a, b, atol=2e-6)
        a = ray(R, (1, 2, 3), 0.9, w, h, 0.0, -2.0 / h, x, y)
        b = ray(R, (1, 2, 3), 0.9, w, h, 0.0, 0.0, x, y + 1)
        assert np.allclose(a, b, atol=2e-6)


def test_reference_fov_is_45_radians():
    R = np.eye(3, dtype=np.float32)
    d = ray(R, (0, 0, 0), 45.0, 100, 100, 0.0, 0.0, 50, 100)  # top edge centre
    half = math.atan2(d[1], d[2])
    assert abs(half - math.atan(math.tan(22.5))) < 2e-6
    assert abs(math.degrees(2 * half) - 58.31) < 0.01


def test_position_and_depth_do_not_change_direction():
    rng = np.random.default_rng(4)
    R = rand_rotation(rng)
    a = ray(R, (0, 0, 0), 1.1, 64, 48, 0.01, -0.02, 5, 40, z=0.3)
    b = ray(R, (7, -3, 11), 1.1, 64, 48, 0.01, -0.02, 5, 40, z=10.0)
    assert np.allclose(a, b, atol=2e-6)


CORNERS = [(0, 0), (1, 0), (1, 1), (0, 1)]  # x, y in units of (w, h)


@pytest.mark.parametrize("lh", [True, False])
def test_basis_corners_equal_screen_to_ray(lh):
    rng = np.random.default_rng(5)
    for fovy in (45.0, 0.8, 1.6):
        R = rand_rotation(rng)
        w, h = 320, 180
        jx, jy = 0.003, -0.004
        got = ptamd.camera_corners_basis(R[:, 0], R[:, 1], R[:, 2], fovy, w / h, jx, jy, lh).reshape(4, 3)
        for k, (cx, cy) in enumerate(CORNERS):
            want = ray(R, (0.5, 1.0, -2.0), fovy, w, h, jx, jy, cx * w, cy * h, lh=lh)
            assert np.allclose(got[k], want, atol=2e-6), (fovy, k)


@pytest.mark.parametrize("eye,target,up", [((0.5, 0.5, -1.5), (0.5, 0.5, 0.0), (0, 1, 0)),
                                           ((-0.324, 4.428, -13.0), (0.0, 5.0, 0.0), (0, 1, 0)),
                                           ((3.0, 1.0, 2.0), (-1.0, 0.0, 0.5), (0, 1, 0))])
def test_look_at_corners_equal_screen_to_ray(eye, target, up):
    # Transform::lookAt's rotation (Transform.cpp:78-87)
    f = unit(np.subtract(target, eye))
    r = unit(np.cross(up, f))
    u = np.cross(f, r)
    R = np.stack([r, u, f], axis=1).astype(np.float32)
    assert abs(np.linalg.det(R) - 1.0) < 1e-5  # a proper rotation, as a Transform holds
    w, h = 256, 192
    jx, jy = ptamd.pass_jitter(9, 3, w, h)
    got = ptamd.camera_corners(eye, target, up, 60.0, w / h, jx, jy).reshape(4, 3)
    for k, (cx, cy) in enumerate(CORNERS):
        want = ray(R, eye, math.radians(60.0), w, h, jx, jy, cx * w, cy * h)
        assert np.allclose(got[k], want, atol=3e-6), k
    # Viewing +z with y up, the screen's right edge lies towards +x (a
    # left-handed frame, as the app's left-handed asset import assumes).
    d = ptamd.camera_corners((0, 0, -5), (0, 0, 0), (0, 1, 0), 60.0, 1.0).reshape(4, 3)
    assert d[0][0] < 0 < d[1][0] and d[2][1] > 0 > d[1][1]


def test_cornell_preset_frames_the_box_left_handed():
    """The app's Cornell preset camera (PathTracingApp.cpp:425-441: position
    (-0.324, 4.428, -13), Euler angles 0, setPerspective(45)) against the
    Cornell box after the left-handed import at scale 5 (z in [-4.95, 5.2],
    y in [0, 9.95]): the left-handed projection's centre ray passes through
    the box's opening; the right-handed one leaves along -z, away from it."""
    I = np.eye(3, dtype=np.float32)
    for lh, sign in ((True, 1.0), (False, -1.0)):
        d = ptamd.camera_corners_basis(I[:, 0], I[:, 1], I[:, 2], 45.0, 16 / 9, 0, 0, lh).reshape(4, 3)
        centre = unit(d.sum(axis=0))
        assert np.sign(centre[2]) == sign
    d = ptamd.camera_corners_basis(I[:, 0], I[:, 1], I[:, 2], 45.0, 16 / 9, 0, 0, True).reshape(4, 3)
    eye = np.array([-0.324, 4.428, -13.0])
    t = (-4.95 - eye[2]) / d[:, 2]  # corners reach the opening plane
    y = eye[1] + t * d[:, 1]
    assert y.min() < 0.5 and y.max() > 8.0  # the opening fills the view vertically
