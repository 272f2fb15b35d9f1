"""The allocation-free closest-hit entry point (pt_trace_closest_device_noalloc)
against pt_trace_closest_device: identical ids, t and uv on two calls in a
row, the second with a smaller batch, and the oversize / no-scene errors."""
import numpy as np
import pytest

import ptamd

pytestmark = pytest.mark.gpu


def say(msg):
    print(msg, flush=True)


def _trace(ctx, rays, noalloc):
    from ptamd import hip

    n = len(rays)
    d_rays = hip.DeviceArray.from_host(rays)
    d_id = hip.DeviceArray(n, np.uint32)
    d_t = hip.DeviceArray(n, np.float32)
    d_uv = hip.DeviceArray(2 * n, np.float32)
    try:
        for a in (d_id, d_t, d_uv):
            a.fill_bytes(0xAB)  # a ray left untraced shows
        f = ctx.trace_closest_device_noalloc if noalloc else ctx.trace_closest_device
        f(d_rays.ptr, n, d_id.ptr, d_t.ptr, d_uv.ptr)
        ctx.synchronize()
        return d_id.to_host(), d_t.to_host(), d_uv.to_host()
    finally:
        for a in (d_rays, d_id, d_t, d_uv):
            a.free()


def test_noalloc_trace_equals_device_trace():
    ctx = ptamd.Context(0)
    try:
        with pytest.raises(ptamd.PtError):
            ctx.trace_closest_device_noalloc(0, 0, 0, 0)  # no scene yet
        tris, infos = ptamd.gen_sphere(96, 5, room=True)
        ctx.upload_scene(tris, infos)
        bmin, bmax = ptamd.scene_bounds(tris)
        for n, seed in ((600000, 41), (70001, 42)):
            say(f"[noalloc] {len(tris)} triangles, {n} shell rays, seed {seed}")
            rays = ptamd.gen_shell_rays(n, seed, bmin, bmax, 1.5)
            a = _trace(ctx, rays, noalloc=False)
            b = _trace(ctx, rays, noalloc=True)
            assert np.array_equal(a[0], b[0])
            assert np.array_equal(a[1].view(np.uint32), b[1].view(np.uint32))
            assert np.array_equal(a[2].view(np.uint32), b[2].view(np.uint32))
            assert (a[0] != ptamd.INVALID).sum() > 0
        with pytest.raises(ptamd.PtError):
            ctx.trace_closest_device_noalloc(1, 0x80000000, 0, 0)  # beyond PT_NOALLOC_MAX_RAYS
        assert ctx.get_status() == 0
    finally:
        ctx.close()
