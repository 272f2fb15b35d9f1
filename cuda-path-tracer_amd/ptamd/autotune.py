"""Launch-shape autotuner.

Times candidate settings of the library's tuning knobs (pt_tune_set) on the
caller's real workload with HIP events on the caller's stream and keeps the
fastest.  Results never depend on these knobs (tests/test_gpu_parity.py
checks the closest-hit knobs), so tuning is safe to run anywhere before a
timed region.
"""
from __future__ import annotations

import itertools

from . import device_compute_units, hip, tune_get, tune_set


def device_cus(device=0):
    """Compute units of the device (the launch-shape search scales with it)."""
    return device_compute_units(device)


def _time(fn, stream, reps):
    e0, e1 = hip.Event(), hip.Event()
    fn()  # warm-up at this setting
    e0.record(stream)
    for _ in range(reps):
        fn()
    e1.record(stream)
    e1.synchronize()
    ms = e0.elapsed_ms(e1) / reps
    e0.destroy()
    e1.destroy()
    return ms


def tune(run, stream, space, reps=3):
    """Grid search over `space` ({knob: [values]}).  `run()` launches one
    step of the workload on `stream`.  Leaves the best setting applied and
    returns (best: {knob: value}, table: [(setting, ms)])."""
    names = list(space)
    saved = {k: tune_get(k) for k in names}
    table = []
    best, best_ms = None, float("inf")
    try:
        for values in itertools.product(*(space[k] for k in names)):
            setting = dict(zip(names, values))
            for k, v in setting.items():
                tune_set(k, v)
            ms = _time(run, stream, reps)
            table.append((setting, ms))
            if ms < best_ms:
                best, best_ms = setting, ms
    finally:
        for k, v in (best or saved).items():
            tune_set(k, v)
    return best, table


def tune_staged(run, stream, stages, reps=3):
    """Coordinate search: grid-search each space of `stages` in turn, the
    knobs of earlier stages held at their best.  Returns (best, table) over
    all stages, like tune()."""
    best, table = {}, []
    for space in stages:
        b, t = tune(run, stream, space, reps)
        best.update(b or {})
        table.extend(t)
    return best, table


def trace_space(num_cus=256):
    """Closest-hit launch shape: persistent workgroups per CU and the
    lane-refill threshold."""
    return {"trace.blocks": [num_cus * k for k in (2, 4, 8, 12, 16)],
            "trace.refill": [16, 32, 48, 56, 64]}


def stack_space():
    """Closest-hit traversal stack and queue: LDS stack entries per lane
    (more entries, fewer spills to scratch, lower occupancy), the rays a
    wave takes from the queue at a time, and the root filter (on or off)."""
    return {"trace.lds_stack": [8, 16, 32],
            "trace.pool": [128, 256, 512, 1024],
            "trace.filter": [0, 1]}


def render_space(num_cus=256):
    """Traversal knobs for a progressive render, searched together on the
    render's own frames: its bounce rays traverse deeper than the closest-hit
    microbench's shell rays and favour other lane and queue settings."""
    return {"trace.blocks": [num_cus * 6, num_cus * 8, num_cus * 12, num_cus * 16],
            "trace.refill": [32, 48, 56, 64],
            "trace.lds_stack": [8, 16],
            "trace.pool": [128, 256]}


def render_stages(num_cus=256):
    """Staged render search: the frame's structure first (path-persistent,
    or bounce levels in 1, 2 or 4 chains), then how many waiting lanes a
    path-persistent wave shades together, then the traversal knobs, then the
    shade batch again (it trades against the lane refill threshold)."""
    return [{"render.paths": [1, 0], "render.chains": [1, 2, 4]},
            {"render.shade_min": [4, 8, 16]},
            render_space(num_cus),
            {"render.shade_min": [4, 8, 16, 32]}]


def trace_stages(num_cus=256):
    """The default staged closest-hit search: launch shape, then stack and
    queue sizes."""
    return [trace_space(num_cus), stack_space()]
