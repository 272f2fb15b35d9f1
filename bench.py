"""Headline benchmark: closest-hit Mrays/s on the 871,200-triangle mesh.

Workload (one step): one closest-hit launch over 16M rays, origins uniform on
the mesh AABB enlarged 1.5x about its centre, directions uniform on the unit
sphere, against the UV sphere of 660 x 660 bands radially displaced by 6-octave
fBm (871,200 fp32 triangles).  Rays, BVH and outputs are resident in HBM before
the timed region.  Beside it: LBVH build ms (2^20-triangle soup and the mesh)
and the CPU oracle on a bounded sample of the same rays.

Further legs: progressive render spp/s, the 2.88M-triangle hairball, the
512-scene batched build and frame, and the radiance cache's time-to-RMSE
against the plain tracer (1920x1080, 65,536 training paths per frame).  The
CPU leg checks the timed results against the oracle at full size (the 16M
closest hits of the mesh and of the hairball, render passes at the legs'
first seeds) and reports a "parity" object beside the roofline.

Prints one JSON line.  Options: --gpus N (only 1 is supported), --steps K,
--warmup W, --rays R, --no-cpu (skip the CPU baseline and the parity
checks), --no-batch, --no-hair, --no-cache, --cache-small (960x540 with
16,384 training paths), --profile (short run for rocprofv3: 2 warmup + 3
steps, no CPU baseline, no hairball, batched or cache legs), --profile-cache
(short eager radiance-cache workload for rocprofv3).
"""
import argparse
import hashlib
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "cuda-path-tracer_amd"))

HBM_PEAK_GBS = 8000.0  # MI355X HBM3E, published


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def render_rmse(ctx, img, w, h, eye, tgt, ref_spp, per_frame, seed):
    """RMSE over all pixels and channels of the running-mean image `img`
    (3 * w * h) against the plain tracer's FP64 accumulator at `ref_spp`
    passes, rendered as two independent halves (own seeds and camera
    jitter) whose difference estimates the reference's own variance, which
    is subtracted.  Leaves the context's accumulator holding the second
    half.  Returns (rmse, reference rmse noise)."""
    import numpy as np

    import ptamd

    halves = []
    for half in range(2):
        ctx.reset_frames()
        for b in range(0, ref_spp // 2, per_frame):
            ps = range(b, min(b + per_frame, ref_spp // 2))
            ctx.render_passes(w, h, np.array([seed + 10 ** 7 * half + p for p in ps], np.uint64),
                              np.stack([ptamd.camera_corners(eye, tgt, (0, 1, 0), 60.0, w / h,
                                                             *ptamd.pass_jitter(seed + half, p, w, h)) for p in ps]))
        acc, spp = ctx.read_accumulator(w, h)
        halves.append(acc / spp)
    ref = 0.5 * (halves[0] + halves[1])
    ref_var = float(np.mean((halves[0] - halves[1]) ** 2)) / 4.0
    mse = float(np.mean((img - ref) ** 2)) - ref_var
    return round(max(mse, 0.0) ** 0.5, 6), round(ref_var ** 0.5, 6)


def cache_leg(ctx, stream, args, seed_mesh):
    """Time-to-RMSE of cache frames against the plain tracer on the room
    scene (871,200-triangle mesh + room + ceiling emitter), 1 spp per frame,
    both against a plain FP64 reference of ref_spp passes.  The plain paths
    have 1 + 8 bounces, the length of a cache training path, so both converge
    to the same image.  RMSE is over all pixels and channels; the reference's
    own variance (from its two halves) is subtracted from every MSE."""
    import numpy as np

    import ptamd
    from ptamd import hip

    w, h = (960, 540) if args.cache_small else (1920, 1080)
    ntrain = 16384 if args.cache_small else 65536
    ref_spp = args.cache_ref_spp or 8192
    extra, steps, frames, plain_max = 8, args.cache_steps, 300, 1024
    rt, ri = ptamd.gen_sphere(660, seed_mesh, room=True)
    st = ptamd.Settings(1 + extra, 3, 1.0, 1)
    ctx.upload_scene(rt, ri, settings=st)
    bmin, bmax = ptamd.scene_bounds(rt)
    eye, tgt = (0.0, 0.0, -1.9), (0.0, 0.0, 0.0)

    def dirs(p, seed):
        return ptamd.camera_corners(eye, tgt, (0, 1, 0), 60.0, w / h, *ptamd.pass_jitter(seed, p, w, h))

    # The eye of every pass (render_passes takes per-pass corner rays only).
    ctx.set_primary_rays(dirs(0, 31), eye)

    # Reference: two halves of ref_spp / 2 passes each (16 passes per frame).
    halves = []
    for half in range(2):
        ctx.reset_frames()
        for b in range(0, ref_spp // 2, 16):
            ps = range(b, b + 16)
            ctx.render_passes(w, h, np.array([10 ** 8 * (half + 1) + p for p in ps], np.uint64),
                              np.stack([dirs(p, 31 + half) for p in ps]))
            if b % 1024 == 0:
                log(f"[bench] cache leg: reference half {half}, {b} / {ref_spp // 2} passes")
        halves.append(ctx.read_accumulator(w, h)[0] / (ref_spp // 2))
    ref = 0.5 * (halves[0] + halves[1])
    ref_var = float(np.mean((halves[0] - halves[1]) ** 2)) / 4.0
    del halves

    def mse(acc, spp):
        return float(np.mean((acc / spp - ref) ** 2)) - ref_var

    checkpoints = [1 << k for k in range(11)]
    e0, e1 = hip.Event(), hip.Event()

    def run(points, frame, reset=None, stop_mse=None, min_frames=0):
        curve, t, done = [], 0.0, 0
        ctx.reset_frames()
        frame(0)  # warm-up frame: allocation and graph capture
        stream.synchronize()
        ctx.reset_frames()
        if reset:
            reset()
        for cp in points:
            e0.record(stream)
            for f in range(done, cp):
                frame(f)
            e1.record(stream)
            e1.synchronize()
            t += e0.elapsed_ms(e1)
            done = cp
            acc, spp = ctx.read_accumulator(w, h)
            curve.append((cp, round(t, 3), mse(acc, spp)))
            if stop_mse is not None and cp >= min_frames and curve[-1][2] <= stop_mse:
                break
        return curve

    plain_seeds = np.arange(2 * 10 ** 8, 2 * 10 ** 8 + plain_max, dtype=np.uint64)
    plain_dirs = [dirs(p, 41) for p in range(plain_max)]

    def plain_frame(f):
        ctx.render_passes(w, h, plain_seeds[f:f + 1], plain_dirs[f][None])

    # Each side runs at the traversal knobs tuned on its own frames (frame
    # structure, shade batch and launch shape).
    import ptamd.autotune as autotune

    start_knobs = knob_state()
    tuning = {}
    if not args.no_tune:
        ctx.reset_frames()
        best, _ = autotune.tune_staged(lambda: plain_frame(0), stream, autotune.render_stages(autotune.device_cus()),
                                       reps=3)
        tuning["plain"] = best
    plain = run(checkpoints, plain_frame)
    set_knobs(start_knobs)
    log(f"[bench] cache leg: plain {plain_max} spp in {plain[-1][1]:.0f} ms, RMSE {max(plain[-1][2], 0) ** 0.5:.5f}")

    cs = ptamd.CacheSettings(ntrain, steps, extra, train=True)

    def cache_frame(f):
        ctx.set_primary_rays(dirs(f, 43), eye)
        ctx.render_cache_frame(w, h, 3 * 10 ** 8 + f, cs)

    # A fresh cache: its training starts with the timed frames.  The run goes
    # on past `frames` (up to 4x) until it reaches the plain tracer's
    # 1,024-spp RMSE, so the time to that RMSE is measured, not extrapolated.
    target = plain[-1][2]
    points = sorted(set([1 << k for k in range(9)] + [frames] + [frames * k // 4 for k in range(5, 17)]))
    if not args.no_tune:
        ctx.cache_init(7, bmin, bmax)
        ctx.reset_frames()
        best, _ = autotune.tune_staged(lambda: cache_frame(0), stream,
                                       [{"render.paths": [1, 0]}, {"render.shade_min": [4, 8, 16]},
                                        autotune.render_space(autotune.device_cus())], reps=3)
        tuning["cache"] = best
    ctx.cache_init(7, bmin, bmax)
    cached = run(points, cache_frame, reset=lambda: ctx.cache_init(7, bmin, bmax), stop_mse=target,
                 min_frames=frames)
    set_knobs(start_knobs)
    at_frames = next(m for c, t, m in cached if c == frames)
    log(f"[bench] cache leg: {cached[-1][0]} cache frames in {cached[-1][1]:.0f} ms, "
        f"RMSE {max(cached[-1][2], 0) ** 0.5:.5f} ({frames} frames: {max(at_frames, 0) ** 0.5:.5f})")

    def rmse(m):
        return round(max(m, 0.0) ** 0.5, 6)

    def time_to(curve, target_mse):
        prev = None
        for spp, t, m in curve:
            if m <= target_mse:
                if prev is None or prev[2] <= m:
                    return t
                # log-log interpolation between the two checkpoints
                import math
                a = (math.log(prev[2]) - math.log(target_mse)) / (math.log(prev[2]) - math.log(max(m, 1e-30)))
                return prev[1] * (t / prev[1]) ** a
            prev = (spp, t, m)
        return None

    def at_time(curve, t_target):
        # log-log interpolation between the checkpoints around t_target
        import math
        for (c0, t0, m0), (c1, t1, m1) in zip(curve, curve[1:]):
            if t0 <= t_target <= t1 and m0 > 0 and m1 > 0:
                a = (math.log(t_target) - math.log(t0)) / (math.log(t1) - math.log(t0))
                return math.exp(math.log(m0) + a * (math.log(m1) - math.log(m0)))
        return None

    t_plain = plain[-1][1]
    t_cache = time_to(cached, target)
    t_at_frames = next(t for c, t, m in cached if c == frames)
    plain_at_frames = at_time(plain, t_at_frames)
    out = {
        "resolution": [w, h], "frames": frames, "train_samples_per_frame": ntrain, "adam_steps_per_frame": steps,
        "extra_bounces": extra, "plain_bounces": 1 + extra, "rr_start": 3, "reference_spp": ref_spp,
        "reference_rmse_noise": rmse(ref_var),
        "plain_curve": [(c, t, rmse(m)) for c, t, m in plain],
        "cache_curve": [(c, t, rmse(m)) for c, t, m in cached],
        "target_rmse_plain_1024spp": rmse(target), "plain_ms_to_target": round(t_plain, 1),
        "cache_ms_to_target": None if t_cache is None else round(t_cache, 1),
        "time_ratio_cache_vs_plain": None if t_cache is None else round(t_cache / t_plain, 4),
        "cache_rmse_at_frames": rmse(at_frames),
        "equal_time_rmse": {"ms": t_at_frames, "cache": rmse(at_frames),
                            "plain": None if plain_at_frames is None else rmse(plain_at_frames)},
        "ms_per_frame": {"plain": round(t_plain / plain_max, 3), "cache": round(cached[-1][1] / cached[-1][0], 3)},
        "tuning": tuning,
    }

    # MLP kernels alone on this workload's sizes: inference over every pixel's
    # query and one Adam step over a training chunk, against the dense fp16
    # MFMA peak (2P flops per query, 6P per training sample, P = 18,624 MACs
    # per sample plus biases).
    P = 18624
    rng = np.random.default_rng(0)
    npix = w * h
    raw = np.zeros((npix, 12), np.float32)
    raw[:, :3] = rng.uniform(bmin, bmax, (npix, 3))
    raw[:, 3:9] = rng.normal(size=(npix, 6))
    raw[:, 9:] = 0.5
    d_raw = hip.DeviceArray.from_host(raw)
    d_out = hip.DeviceArray((npix, 3), np.float32)
    chunk = ntrain // steps
    d_tgt = hip.DeviceArray.from_host(rng.uniform(0, 1, (chunk, 3)).astype(np.float32))
    for _ in range(3):
        ctx.cache_infer_device(d_raw.ptr, npix, d_out.ptr, stream.handle)
        ctx.cache_train_device(d_raw.ptr, d_tgt.ptr, chunk, 0, stream.handle)
    reps = 20
    e0.record(stream)
    for _ in range(reps):
        ctx.cache_infer_device(d_raw.ptr, npix, d_out.ptr, stream.handle)
    e1.record(stream)
    e1.synchronize()
    inf_ms = e0.elapsed_ms(e1) / reps
    e0.record(stream)
    for _ in range(reps):
        ctx.cache_train_device(d_raw.ptr, d_tgt.ptr, chunk, 0, stream.handle)
    e1.record(stream)
    e1.synchronize()
    tr_ms = e0.elapsed_ms(e1) / reps
    out["mlp"] = {"infer_ms": round(inf_ms, 4), "infer_queries": npix,
                  "infer_tflops": round(2 * P * npix / (inf_ms * 1e-3) / 1e12, 2),
                  "train_step_ms": round(tr_ms, 4), "train_samples": chunk,
                  "train_tflops": round(6 * P * chunk / (tr_ms * 1e-3) / 1e12, 2),
                  "mfma_fp16_dense_peak_tflops": 2500.0}
    log(f"[bench] cache leg: target RMSE {out['target_rmse_plain_1024spp']}: plain {t_plain:.0f} ms, "
        f"cache {out['cache_ms_to_target']} ms; MLP infer {inf_ms:.3f} ms "
        f"({out['mlp']['infer_tflops']} TFLOP/s), train step {tr_ms:.3f} ms")
    return out


def profile_cache():
    """Short radiance-cache workload for rocprofv3 (--profile-cache): the
    1920x1080 room scene, 6 cache frames with 65,536 training paths and 4
    Adam steps each, launched eagerly (render.graphs 0) so the profiler sees
    each kernel launch, then 5 inferences over 2,073,600 queries and 5 Adam
    steps over 16,384 samples."""
    import ptamd
    from ptamd import hip

    w, h, ntrain = 1920, 1080, 65536
    rt, ri = ptamd.gen_sphere(660, 2024, room=True)
    ctx = ptamd.Context(0)
    stream = hip.Stream()
    ctx.set_stream(stream.handle)
    ctx.upload_scene(rt, ri, settings=ptamd.Settings(9, 3, 1.0, 1))
    bmin, bmax = ptamd.scene_bounds(rt)
    eye, tgt = (0.0, 0.0, -1.9), (0.0, 0.0, 0.0)
    ptamd.tune_set("render.graphs", 0)
    ctx.cache_init(7, bmin, bmax)
    cs = ptamd.CacheSettings(ntrain, 4, 8, train=True)
    for f in range(6):
        ctx.set_primary_rays(ptamd.camera_corners(eye, tgt, (0, 1, 0), 60.0, w / h,
                                                  *ptamd.pass_jitter(43, f, w, h)), eye)
        ctx.render_cache_frame(w, h, 3 * 10 ** 8 + f, cs)
        stream.synchronize()
        log(f"[profile-cache] cache frame {f}")
    rng = np.random.default_rng(0)
    npix = w * h
    raw = np.zeros((npix, 12), np.float32)
    raw[:, :3] = rng.uniform(bmin, bmax, (npix, 3))
    raw[:, 3:9] = rng.normal(size=(npix, 6))
    raw[:, 9:] = 0.5
    d_raw = hip.DeviceArray.from_host(raw)
    d_out = hip.DeviceArray((npix, 3), np.float32)
    chunk = ntrain // 4
    d_tgt = hip.DeviceArray.from_host(rng.uniform(0, 1, (chunk, 3)).astype(np.float32))
    for _ in range(5):
        ctx.cache_infer_device(d_raw.ptr, npix, d_out.ptr, stream.handle)
    for _ in range(5):
        ctx.cache_train_device(d_raw.ptr, d_tgt.ptr, chunk, 0, stream.handle)
    stream.synchronize()
    log("[profile-cache] 5 inferences, 5 Adam steps")
    print(json.dumps({"profile_cache": "done", "frames": 6, "resolution": [w, h], "train_samples": ntrain}))


KERNEL_SOURCES = ("cuda-path-tracer_amd/csrc/traverse.hpp", "cuda-path-tracer_amd/csrc/trace.hip",
                  "cuda-path-tracer_amd/csrc/devmath.hpp")


def kernel_source_sha():
    """SHA-256 over the closest-hit kernel's sources: ties a committed
    rocprofv3 traffic measurement to the code it was taken on."""
    h = hashlib.sha256()
    for rel in KERNEL_SOURCES:
        with open(os.path.join(ROOT, rel), "rb") as f:
            h.update(f.read())
    return h.hexdigest()


def timed(stream, fn, reps=1):
    """ms per call of fn() on `stream` (HIP events)."""
    from ptamd import hip

    e0, e1 = hip.Event(), hip.Event()
    e0.record(stream)
    for _ in range(reps):
        fn()
    e1.record(stream)
    e1.synchronize()
    ms = e0.elapsed_ms(e1) / reps
    e0.destroy()
    e1.destroy()
    return ms


def same_bits(a, b):
    return bool(np.array_equal(np.ascontiguousarray(a).view(np.uint8), np.ascontiguousarray(b).view(np.uint8)))


KNOBS = ("trace.blocks", "trace.refill", "trace.lds_stack", "trace.pool", "trace.filter", "render.paths",
         "render.chains", "render.shade_min")


def knob_state():
    import ptamd

    return {k: ptamd.tune_get(k) for k in KNOBS}


def set_knobs(state):
    import ptamd

    for k, v in state.items():
        ptamd.tune_set(k, v)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--rays", type=int, default=16 * 1024 * 1024)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--profile", action="store_true")
    ap.add_argument("--profile-cache", action="store_true",
                    help="short eager radiance-cache workload for rocprofv3, then exit")
    ap.add_argument("--cpu-seconds", type=float, default=12.0)
    ap.add_argument("--no-tune", action="store_true", help="skip the launch-shape autotuner")
    ap.add_argument("--no-batch", action="store_true", help="skip the 512-scene batched leg")
    ap.add_argument("--no-hair", action="store_true", help="skip the 2.88M-triangle hairball leg")
    ap.add_argument("--no-cache", action="store_true", help="skip the radiance-cache leg")
    ap.add_argument("--cache-small", action="store_true",
                    help="radiance-cache leg at 960x540 with 16,384 training paths per frame "
                         "(default: config 4's 1920x1080 with 65,536)")
    ap.add_argument("--cache-steps", type=int, default=4, help="Adam steps per cache frame (1..64)")
    ap.add_argument("--cache-ref-spp", type=int, default=0,
                    help="reference spp of the cache leg (default 8192, two halves of 4096)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py: only --gpus 1 is supported (this build runs on one GPU; got {args.gpus})",
              file=sys.stderr)
        sys.exit(2)
    if args.profile_cache:
        profile_cache()
        return
    if args.profile:
        args.steps, args.warmup, args.no_cpu, args.no_tune, args.no_batch = 3, 2, True, True, True
        args.no_hair = True
        args.no_cache = True

    import ptamd
    from ptamd import hip

    seed_mesh, seed_rays = 2024, 77
    t0 = time.time()
    tris, infos = ptamd.gen_sphere(660, seed_mesh, room=False)
    bmin, bmax = ptamd.scene_bounds(tris)
    rays = ptamd.gen_shell_rays(args.rays, seed_rays, bmin, bmax, 1.5)
    log(f"[bench] generated {len(tris)} triangles, {len(rays)} rays in {time.time() - t0:.1f}s")

    ctx = ptamd.Context(0)
    stream = hip.Stream()
    ctx.set_stream(stream.handle)
    # Library defaults: every leg after the closest hit starts from these (a
    # leg that tunes restores them afterwards), not from the closest hit's
    # choice.
    defaults = knob_state()
    ctx.upload_scene(tris, infos)
    d_rays = hip.DeviceArray.from_host(rays)
    n = len(rays)
    d_id = hip.DeviceArray(n, np.uint32)
    d_t = hip.DeviceArray(n, np.float32)

    # Visit counts of the reference-order traversal on exactly these rays
    # (equal to the oracle's, which the parity tests check) -> algorithmic bytes.
    cnt = ctx.trace_closest_counted(d_rays.ptr, n, stream=stream.handle)
    iters, boxes, tri_tests, max_depth = (int(x) for x in cnt)
    bytes_per_launch = 24 * n + 8 * n + 4 * iters + 24 * boxes + 36 * tri_tests
    log(f"[bench] per ray: {iters / n:.2f} node visits, {boxes / n:.2f} box tests, "
        f"{tri_tests / n:.2f} triangle tests -> {bytes_per_launch / n:.0f} B/ray; deepest stack {max_depth}")

    def step():
        ctx.trace_closest_device(d_rays.ptr, n, d_id.ptr, d_t.ptr, 0, stream.handle)

    # Autotune the launch shape on this workload (before the timed region).
    tuning = None
    if not args.no_tune:
        from ptamd import autotune

        best, table = autotune.tune_staged(step, stream, autotune.trace_stages(autotune.device_cus()), reps=4)
        tuning = {"chosen": best, "candidates": len(table),
                  "best_ms": round(min(ms for _, ms in table), 4),
                  "worst_ms": round(max(ms for _, ms in table), 4)}
        log(f"[bench] autotune: {best} ({len(table)} candidates, "
            f"{tuning['best_ms']:.3f}..{tuning['worst_ms']:.3f} ms)")
    for _ in range(args.warmup):
        step()
    stream.synchronize()
    hip.synchronize()
    e0, e1 = hip.Event(), hip.Event()
    w0 = time.time()
    e0.record(stream)
    for _ in range(args.steps):
        step()
    e1.record(stream)
    e1.synchronize()
    hip.synchronize()
    wall = time.time() - w0
    closest_hit_knobs = knob_state()
    ms = e0.elapsed_ms(e1)
    ms_per_step = ms / args.steps
    mrays = n / (ms_per_step * 1e-3) / 1e6
    achieved = bytes_per_launch / (ms_per_step * 1e-3) / 1e9
    if args.profile:
        # Calibration for the HBM byte counters: a plain 256 MiB device copy.
        a = hip.DeviceArray(256 << 20, np.uint8)
        b = hip.DeviceArray(256 << 20, np.uint8)
        a.zero()
        b.copy_from(a)
        hip.synchronize()
        a.free()
        b.free()
    gpu_ids, gpu_t = d_id.to_host(), d_t.to_host()  # the timed launches' results
    hit_rate = float((gpu_ids != ptamd.INVALID).mean())
    log(f"[bench] {mrays:.1f} Mrays/s ({ms_per_step:.3f} ms/launch, wall {wall:.2f}s), hit rate {hit_rate:.3f}")
    # The same launch through the host-buffer entry point (pageable copies
    # over PCIe both ways included): not the metric, recorded beside it.
    host_mrays = None
    if not args.profile:
        ctx.trace_closest(rays[: 1 << 20])  # staging buffers
        t1 = time.time()
        ctx.trace_closest(rays)
        host_mrays = round(n / (time.time() - t1) / 1e6, 1)
        log(f"[bench] host-buffer entry point (PCIe copies included): {host_mrays} Mrays/s")

    # LBVH build: kernel time of a rebuild on reserved buffers (HIP events),
    # and the first build of a fresh context (allocation, upload and build,
    # host wall clock), against the 320 B/triangle algorithmic bytes.
    build_ms_mesh = float(np.median([ctx.build_bvh() for _ in range(5)]))
    soup, soup_infos = ptamd.gen_soup(1 << 20, 1, emitter=False)
    fresh = ptamd.Context(0)
    t1 = time.time()
    fresh.upload_scene(soup, soup_infos)
    first_build_ms = (time.time() - t1) * 1e3
    fresh.close()
    ctx.upload_scene(soup, soup_infos)
    build_ms_1m = float(np.median([ctx.build_bvh() for _ in range(10)]))

    def build_roofline(ntri, ms_):
        gbs = 320.0 * ntri / (ms_ * 1e-3) / 1e9
        return {"bound": "hbm", "achieved": round(gbs, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                "frac": round(gbs / HBM_PEAK_GBS, 4), "bytes_per_triangle": 320}

    lbvh = {"soup_2^20": {"ms": round(build_ms_1m, 4), "roofline": build_roofline(len(soup), build_ms_1m),
                          "first_upload_and_build_ms": round(first_build_ms, 2)},
            "mesh_871200": {"ms": round(build_ms_mesh, 4), "roofline": build_roofline(len(tris), build_ms_mesh)}}
    log(f"[bench] LBVH build: {build_ms_1m:.3f} ms @ 2^20 soup ({lbvh['soup_2^20']['roofline']['frac']} of HBM; "
        f"first build with allocation and upload {first_build_ms:.1f} ms), {build_ms_mesh:.3f} ms @ 871,200 mesh")

    set_knobs(defaults)

    # Progressive render legs (spp/s): config 0 (16 passes per frame) and
    # config 1 (one pass per frame), frames replayed as captured hipGraphs.
    # Each states the RMSE of its image against a high-spp plain reference
    # (ref_spp passes, FP64 sums; outside the timed region).
    render = {}
    legs = [
        ("c0_soup_256x256_b4", ptamd.gen_soup(65536, 0), ptamd.Settings(4, ptamd.INVALID, 1.0, 1),
         (256, 256), 2, 16, (0.5, 0.5, -1.5), (0.5, 0.5, 0.0), 4096),
        ("c1_room_1024x1024_b8_rr3", ptamd.gen_sphere(660, seed_mesh, room=True), ptamd.Settings(8, 3, 1.0, 1),
         (1024, 1024), 4 if args.profile else 8, 1, (0.0, 0.0, -1.9), (0.0, 0.0, 0.0), 256)]
    if not args.profile:
        # The same passes 16 to a frame: every bounce level holds 16M paths,
        # so its tail (its longest rays) is paid once per 16 passes; the
        # accumulator equals 16 one-pass frames bit for bit.
        legs.append(("c1_room_1024x1024_b8_rr3_16pf", ptamd.gen_sphere(660, seed_mesh, room=True),
                     ptamd.Settings(8, 3, 1.0, 1), (1024, 1024), 2, 16, (0.0, 0.0, -1.9), (0.0, 0.0, 0.0), 256))
    for name, (rt, ri), st, (w, h), frames, per_frame, eye, tgt, ref_spp in legs:
        ctx.upload_scene(rt, ri, settings=st)
        ctx.set_primary_rays(ptamd.camera_corners(eye, tgt, (0, 1, 0), 60.0, w / h), eye)
        passes = frames * per_frame
        dirs = np.stack([ptamd.camera_corners(eye, tgt, (0, 1, 0), 60.0, w / h, *ptamd.pass_jitter(5, p, w, h))
                         for p in range(passes)])
        seeds = np.arange(1000, 1000 + passes, dtype=np.uint64)
        ctx.render_passes(w, h, seeds[:per_frame], dirs[:per_frame])  # warm-up frame (allocation, capture)
        # The traversal knobs tuned on this render's own frames (restored to
        # the closest-hit choice after the leg).
        rtune, saved = None, None
        if not args.no_tune:
            from ptamd import autotune

            stages = autotune.render_stages(autotune.device_cus())
            saved = {k: ptamd.tune_get(k) for st in stages for k in st}
            best, table = autotune.tune_staged(lambda: ctx.render_passes(w, h, seeds[:per_frame],
                                                                         dirs[:per_frame]),
                                               stream, stages, reps=3)
            rtune = {"chosen": best, "candidates": len(table)}
        ctx.reset_frames()
        stream.synchronize()
        rms = timed(stream, lambda: [ctx.render_passes(w, h, seeds[f * per_frame:(f + 1) * per_frame],
                                                       dirs[f * per_frame:(f + 1) * per_frame])
                                     for f in range(frames)])
        frame_rays = ctx.last_frame_rays()  # path segments of the last frame
        render[name] = {"ms_per_pass": round(rms / passes, 3), "spp_per_s": round(passes / (rms * 1e-3), 2),
                        "Msamples_per_s": round(w * h * passes / (rms * 1e-3) / 1e6, 2), "passes": passes,
                        "passes_per_frame": per_frame,
                        "segments_per_pass": int(frame_rays // per_frame),
                        "Msegments_per_s": round(frame_rays / per_frame * passes / (rms * 1e-3) / 1e6, 1)}
        if rtune:
            render[name]["tuning"] = rtune
        if not args.profile:
            acc, spp = ctx.read_accumulator(w, h)
            err, noise = render_rmse(ctx, acc / spp, w, h, eye, tgt, ref_spp, per_frame, 7 * 10 ** 8)
            render[name].update({"rmse": err, "rmse_spp": int(spp), "rmse_ref_spp": ref_spp,
                                 "rmse_ref_noise": noise})
            log(f"[bench] render {name}: RMSE {err} at {spp} spp (reference {ref_spp} spp, noise {noise})")
        log(f"[bench] render {name}: {rms / passes:.3f} ms/pass ({per_frame} passes/frame)"
            + (f", tuned {rtune['chosen']}" if rtune else ""))
        if saved:
            for k, v in saved.items():
                ptamd.tune_set(k, v)

    # Config 2: 10,000-strand hairball (2,880,000 triangles + a 2-triangle
    # emitter, heavily overlapping boxes): build, the 16M-ray closest-hit
    # microbench (same ray generator as config 1) with its algorithmic bytes,
    # and a 1920x1080 6-bounce render.
    hair, hair_parity = None, None
    if not args.no_hair:
        ht, hi = ptamd.gen_hairball(10000, 144, 2025, light=True)
        hst = ptamd.Settings(6, ptamd.INVALID, 1.0, 1)
        ctx.upload_scene(ht, hi, settings=hst)
        h_build = float(np.median([ctx.build_bvh() for _ in range(3)]))
        hb0, hb1 = ptamd.scene_bounds(ht)
        hn = n
        hrays_host = ptamd.gen_shell_rays(hn, seed_rays + 1, hb0, hb1, 1.5)
        hrays = hip.DeviceArray.from_host(hrays_host)
        hcnt = ctx.trace_closest_counted(hrays.ptr, hn, stream=stream.handle)
        hbytes = 24 * hn + 8 * hn + 4 * int(hcnt[0]) + 24 * int(hcnt[1]) + 36 * int(hcnt[2])

        def hstep():
            ctx.trace_closest_device(hrays.ptr, hn, d_id.ptr, d_t.ptr, 0, stream.handle)

        set_knobs(closest_hit_knobs)  # the launch shape tuned for the closest-hit microbench
        hstep()
        stream.synchronize()
        h_ms = timed(stream, hstep, 3)
        set_knobs(defaults)
        hair_parity = (d_id.to_host(), d_t.to_host())
        hrays.free()
        w2, h2, eye2, tgt2 = 1920, 1080, (0.0, 0.3, -3.2), (0.0, 0.0, 0.0)
        ctx.set_primary_rays(ptamd.camera_corners(eye2, tgt2, (0, 1, 0), 60.0, w2 / h2), eye2)
        hframes = 4
        hdirs = np.stack([ptamd.camera_corners(eye2, tgt2, (0, 1, 0), 60.0, w2 / h2, *ptamd.pass_jitter(6, p, w2, h2))
                          for p in range(hframes + 1)])
        hseeds = np.arange(2000, 2001 + hframes, dtype=np.uint64)
        ctx.render_passes(w2, h2, hseeds[:1], hdirs[:1])  # warm-up (allocation, capture)
        # The render's knobs tuned on its own one-pass frames, as for the
        # other render legs (restored afterwards).
        hrtune, hsaved = None, None
        if not args.no_tune:
            from ptamd import autotune

            stages = autotune.render_stages(autotune.device_cus())
            hsaved = {k: ptamd.tune_get(k) for st_ in stages for k in st_}
            hbest, htable = autotune.tune_staged(lambda: ctx.render_passes(w2, h2, hseeds[:1], hdirs[:1]),
                                                 stream, stages, reps=2)
            hrtune = {"chosen": hbest, "candidates": len(htable)}
        ctx.reset_frames()
        hr_ms = timed(stream, lambda: [ctx.render_passes(w2, h2, hseeds[p:p + 1], hdirs[p:p + 1])
                                       for p in range(1, hframes + 1)]) / hframes
        hsegs = ctx.last_frame_rays()
        if hsaved:
            for k, v in hsaved.items():
                ptamd.tune_set(k, v)
        h_err = None
        if not args.profile:
            acc, spp = ctx.read_accumulator(w2, h2)
            h_err = render_rmse(ctx, acc / spp, w2, h2, eye2, tgt2, 128, 16, 8 * 10 ** 8) + (int(spp),)
        hair = {"triangles": int(len(ht)), "build_ms": round(h_build, 3),
                "build_roofline": build_roofline(len(ht), h_build),
                "closest_hit_Mrays_per_s": round(hn / (h_ms * 1e-3) / 1e6, 1), "closest_hit_ms": round(h_ms, 3),
                "visits_per_ray": round(int(hcnt[0]) / hn, 3), "box_tests_per_ray": round(int(hcnt[1]) / hn, 3),
                "tri_tests_per_ray": round(int(hcnt[2]) / hn, 3), "deepest_stack": int(hcnt[3]),
                "algorithmic_GB_per_s": round(hbytes / (h_ms * 1e-3) / 1e9, 1),
                "hbm_frac": round(hbytes / (h_ms * 1e-3) / 1e9 / HBM_PEAK_GBS, 4),
                "render_1920x1080_b6": {"ms_per_pass": round(hr_ms, 3), "spp_per_s": round(1e3 / hr_ms, 2),
                                        "passes": hframes, "passes_per_frame": 1,
                                        "segments_per_pass": int(hsegs),
                                        "Msegments_per_s": round(hsegs / (hr_ms * 1e-3) / 1e6, 1)}}
        if hrtune:
            hair["render_1920x1080_b6"]["tuning"] = hrtune
        if h_err is not None:
            hair["render_1920x1080_b6"].update({"rmse": h_err[0], "rmse_spp": h_err[2], "rmse_ref_spp": 128,
                                                "rmse_ref_noise": h_err[1]})
        log(f"[bench] hairball ({len(ht)} tris): build {h_build:.3f} ms; closest hit {hair['closest_hit_Mrays_per_s']} "
            f"Mrays/s ({hair['visits_per_ray']} visits/ray); render 1080p b6 {hr_ms:.2f} ms/pass")

    # Config 3: 512 soups (log-uniform 1,024..262,144 triangles), one batched
    # build vs one build per scene (all 512); one frame tracing every scene
    # (a 512x512-path pass each: 2^18 camera rays plus their bounces, 8
    # bounces, RR from 3) vs one frame per scene (all 512).
    batched = None
    if not args.no_batch:
        bt, bi, boffs = ptamd.gen_batch(512, 3)
        bi["emission"][::37] = 5.0  # scattered emitting triangles in every scene
        ctx.upload_scenes(bt, boffs, bi, settings=ptamd.Settings(8, 3, 1.0, 1))
        ctx.build_bvh()
        b_ms = float(np.median([ctx.build_bvh() for _ in range(3)]))
        eye, tgt = (0.5, 0.5, -1.5), (0.5, 0.5, 0.0)
        bw = bh = 512
        cdirs = ptamd.camera_corners(eye, tgt, (0, 1, 0), 60.0, 1.0)
        ctx.set_primary_rays(cdirs, eye)
        scenes = np.arange(512, dtype=np.uint32)
        seeds = np.arange(512, dtype=np.uint64)
        ctx.render_passes(bw, bh, seeds, None, scenes)  # warm-up (allocation, capture)
        ctx.reset_frames()
        frame_ms = timed(stream, lambda: ctx.render_passes(bw, bh, seeds, None, scenes))
        rays_b = ctx.last_frame_rays()
        ctx.reset_frames()
        ctx.render_passes(bw, bh, seeds[:1], None, scenes[:1])  # one-pass frame: capture
        single_ms = timed(stream, lambda: [ctx.render_passes(bw, bh, seeds[s:s + 1], None, scenes[s:s + 1])
                                           for s in range(512)])
        sb = []
        for s in range(512):
            ctx.upload_scene(bt[boffs[s]:boffs[s + 1]], bi)
            sb.append(ctx.build_bvh())
        single_build_ms = float(np.sum(sb))
        batched = {"scenes": 512, "triangles": int(len(bt)), "camera_rays_per_scene": bw * bh,
                   "build_ms": round(b_ms, 3), "scenes_per_s_built": round(512 / (b_ms * 1e-3), 1),
                   "build_roofline": build_roofline(len(bt), b_ms),
                   "one_scene_per_launch_build_ms": round(single_build_ms, 3),
                   "frame_ms": round(frame_ms, 3), "frame_rays": rays_b,
                   "aggregate_Mrays_per_s": round(rays_b / (frame_ms * 1e-3) / 1e6, 1),
                   "one_scene_per_frame_ms": round(single_ms, 3),
                   "one_scene_per_frame_Mrays_per_s": round(rays_b / (single_ms * 1e-3) / 1e6, 1)}
        del bt, bi
        log(f"[bench] batched: build {b_ms:.2f} ms ({512 / (b_ms * 1e-3):.0f} scenes/s) vs "
            f"{single_build_ms:.2f} ms one per launch; frame {frame_ms:.2f} ms "
            f"({batched['aggregate_Mrays_per_s']} Mrays/s) vs {single_ms:.2f} ms one scene per frame")

    cache = None if args.no_cache else cache_leg(ctx, stream, args, seed_mesh)

    cpu, parity = None, None
    if not args.no_cpu:
        sys.path.insert(0, os.path.join(ROOT, "tests"))
        import oracle_py as orc

        fast = orc.fast_library_usable()
        lib = os.path.basename(orc.select_library(fast))
        threads = orc.THREADS
        parity = {}
        o = orc.build(tris, bmin, bmax)
        # One full pass over every timed ray: its results are the parity
        # check of the timed launches, its time the first sample.
        t1 = time.time()
        oids, ot, _ = orc.trace(tris, o, rays, threads=threads)
        first = time.time() - t1
        parity["c1_16M_closest_hit"] = bool(np.array_equal(oids, gpu_ids) and same_bits(ot, gpu_t))
        reps = max(1, int(round(args.cpu_seconds / max(first, 1e-6))))
        t1 = time.time()
        for _ in range(reps):
            orc.trace(tris, o, rays, threads=threads)
        dt = time.time() - t1
        cpu = {"value": round(n * reps / dt / 1e6, 3), "unit": "Mrays/s", "cores": threads, "kind": "port",
               "sample": f"all {n} rays traced {reps} times by the oracle's reference-order traversal "
                         f"({lib}: g++ -O3 -march=x86-64-v3, no FP contraction), OpenMP {threads} threads "
                         f"(the box's CPU share for one GPU), {dt:.1f} s"}
        log(f"[bench] CPU oracle: {cpu['value']} Mrays/s on {threads} threads ({reps} x {n} rays, {dt:.1f}s); "
            f"parity of the timed 16M hits: {parity['c1_16M_closest_hit']}")
        if hair_parity is not None:
            ob = orc.build(ht, hb0, hb1)
            hids, htt, _ = orc.trace(ht, ob, hrays_host, threads=threads)
            parity["c2_16M_closest_hit"] = bool(np.array_equal(hids, hair_parity[0]) and same_bits(htt, hair_parity[1]))
            log(f"[bench] parity of the hairball's 16M hits: {parity['c2_16M_closest_hit']}")

        # The other metrics' CPU restatements on the same host (bounded runs):
        # the LBVH build of the 2^20-triangle soup and the render legs' first
        # passes, each also rendered on the GPU for a bit-exact parity check.
        sb0, sb1 = ptamd.scene_bounds(soup)
        t1 = time.time()
        so = orc.build(soup, sb0, sb1)
        cpu_build_ms = (time.time() - t1) * 1e3
        ctx.upload_scene(soup, soup_infos)
        g = ctx.get_bvh()
        parity["lbvh_soup_2^20"] = all(same_bits(g[k], so[k]) for k in g)
        cpu_render = {}
        cpu_legs = [("c0_soup_256x256_b4", ptamd.gen_soup(65536, 0), ptamd.Settings(4, ptamd.INVALID, 1.0, 1),
                     (256, 256), 4, (0.5, 0.5, -1.5), (0.5, 0.5, 0.0), 5, 1000),
                    ("c1_room_1024x1024_b8_rr3", ptamd.gen_sphere(660, seed_mesh, room=True),
                     ptamd.Settings(8, 3, 1.0, 1), (1024, 1024), 1, (0.0, 0.0, -1.9), (0.0, 0.0, 0.0), 5, 1000)]
        if hair is not None:
            cpu_legs.append(("c2_hair_1920x1080_b6", (ht, hi), hst, (1920, 1080), 1, (0.0, 0.3, -3.2),
                             (0.0, 0.0, 0.0), 6, 2000))
        for name, (rt, ri), st, (w, h), passes, eye, tgt, jseed, seed0 in cpu_legs:
            ob = orc.build(rt, *ptamd.scene_bounds(rt))
            acc = np.zeros((h, w, 3))
            tot = {"segments": 0, "visits": 0, "boxes": 0, "tris": 0, "hits": 0}
            pdirs = np.stack([ptamd.camera_corners(eye, tgt, (0, 1, 0), 60.0, w / h,
                                                   *ptamd.pass_jitter(jseed, p, w, h)) for p in range(passes)])
            t1 = time.time()
            for p in range(passes):
                _, sts = orc.render_pass_stats(rt, ob, ri, st, w, h, pdirs[p], eye, p, seed0 + p, acc,
                                               threads=threads)
                for k in tot:
                    tot[k] += sts[k]
            dt = time.time() - t1
            cpu_render[name] = {"ms_per_pass": round(dt * 1e3 / passes, 1), "spp_per_s": round(passes / dt, 3),
                                "Msamples_per_s": round(w * h * passes / dt / 1e6, 3), "passes": passes}
            # The same passes on the GPU (one frame), accumulators compared.
            ctx.upload_scene(rt, ri, settings=st)
            ctx.set_primary_rays(pdirs[0], eye)
            ctx.render_passes(w, h, np.arange(seed0, seed0 + passes, dtype=np.uint64), pdirs)
            gacc, _ = ctx.read_accumulator(w, h)
            parity[f"{name}_{passes}pass"] = bool(np.array_equal(gacc, acc))
            # Algorithmic bytes of a pass (closest-hit bytes per segment with
            # the reference-order traversal's counts, FP64 accumulate per
            # pixel, shading fetch per hit) from these oracle passes; roofline
            # of the GPU legs.
            pass_bytes = (32 * tot["segments"] + 4 * tot["visits"] + 24 * tot["boxes"] + 36 * tot["tris"]
                          + 48 * w * h * passes + 24 * tot["hits"]) / passes
            targets = [render.get(name), render.get(name + "_16pf")]
            if name.startswith("c2") and hair is not None:
                targets.append(hair["render_1920x1080_b6"])
            for gleg in targets:
                if gleg:
                    gbs = pass_bytes / (gleg["ms_per_pass"] * 1e-3) / 1e9
                    gleg.update({"algorithmic_bytes_per_pass": int(pass_bytes), "algorithmic_GB_per_s": round(gbs, 1),
                                 "hbm_frac": round(gbs / HBM_PEAK_GBS, 4),
                                 "bytes_from": f"oracle traversal counts of {passes} pass(es) at the same seeds"})
        cpu["other_metrics"] = {"lbvh_build_ms_soup_2^20": round(cpu_build_ms, 1), "render": cpu_render,
                                "threads": threads}
        log(f"[bench] CPU oracle: LBVH build {cpu_build_ms:.0f} ms @ 2^20; render "
            + ", ".join(f"{k} {v['ms_per_pass']} ms/pass" for k, v in cpu_render.items()) + f"; parity {parity}")

    # HBM bytes per launch of the same kernel from the committed rocprofv3
    # FETCH_SIZE/WRITE_SIZE runs (calibrated against a 256 MiB copy), used
    # only while the kernel sources are the ones it was measured on.
    traffic, traffic_note = None, "no committed measurement for this ray count"
    tf = os.path.join(ROOT, "profiles", "closest_hit_c1_16M_hbm_bytes.json")
    if os.path.exists(tf) and n == 16 * 1024 * 1024:
        with open(tf) as f:
            rec = json.load(f)
        if rec.get("source_sha256") == kernel_source_sha():
            traffic = int(rec["bytes_per_launch"])
            traffic_note = "rocprofv3 FETCH_SIZE/WRITE_SIZE of this kernel source (profiles/)"
        else:
            traffic_note = "committed measurement is of other kernel sources: not used"

    out = {
        "metric": "closest-hit Mrays/s @871K tris",
        "value": round(mrays, 2),
        "unit": "Mrays/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_per_step, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {"workload": "closest-hit microbench: 16M shell rays (AABB x1.5, uniform S2 directions) "
                               "vs 660x660-band fBm UV sphere, 871,200 fp32 triangles",
                   "rays": n, "triangles": int(len(tris)), "seed_mesh": seed_mesh, "seed_rays": seed_rays},
        "roofline": {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": traffic, "traffic_source": traffic_note,
                     "algorithmic_bytes": int(bytes_per_launch)},
        "cpu_baseline": cpu,
        "parity": parity,
        "tuning": tuning,
        "lbvh_build": lbvh,
        "render": render,
        "batched_scenes": batched,
        "hairball": hair,
        "radiance_cache": cache,
        "host_buffers_Mrays_per_s": host_mrays,
        "traversal": {"visits_per_ray": round(iters / n, 3), "box_tests_per_ray": round(boxes / n, 3),
                      "tri_tests_per_ray": round(tri_tests / n, 3), "deepest_stack": max_depth,
                      "algorithmic_bytes_per_ray": round(bytes_per_launch / n, 1), "hit_rate": round(hit_rate, 4)},
    }
    ctx.close()
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
