"""liboracle_fast.so (the CPU baseline's -O3 AVX2 build of the oracle) gives
the same bits as liboracle.so: builds, closest hits with counters, and a
render pass."""
import numpy as np
import pytest

import oracle_py as orc
import ptamd


@pytest.mark.skipif(not orc.fast_library_usable(), reason="no AVX2 build or host support")
def test_fast_library_same_bits():
    tris, infos = ptamd.gen_sphere(96, 4, room=True)
    bmin, bmax = ptamd.scene_bounds(tris)
    rays = ptamd.gen_shell_rays(20000, 5, bmin, bmax, 1.5)
    st = ptamd.Settings(4, 2, 1.0, 1)
    eye = (0.0, 0.0, -1.9)
    dirs = ptamd.camera_corners(eye, (0, 0, 0), (0, 1, 0), 60.0, 1.0)
    out = []
    try:
        for fast in (False, True):
            orc.select_library(fast)
            b = orc.build(tris, bmin, bmax)
            tr = orc.trace(tris, b, rays, counters=True)
            acc, _ = orc.render_pass(tris, b, infos, st, 32, 32, dirs, eye, 0, 3)
            out.append((b, tr, acc))
    finally:
        orc.select_library(False)
    (b0, t0, a0), (b1, t1, a1) = out
    for k in b0:
        assert np.array_equal(b0[k].view(np.uint32), b1[k].view(np.uint32)), k
    for x, y in zip(t0, t1):
        assert np.array_equal(np.asarray(x).view(np.uint8), np.asarray(y).view(np.uint8))
    assert np.array_equal(a0, a1)
