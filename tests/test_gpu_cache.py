"""GPU radiance cache (fp16 MFMA, fp32 accumulation) against the FP64 oracle.

Tolerances are those of fp16 operands: inputs, activations and deltas are
rounded to fp16 between layers, so
  inference  relative error <= 5e-3 per output (mean <= 1e-3) against the
             oracle evaluated with the same fp16-rounded weights;
  gradient   per-layer relative L2 error <= 2e-2 and cosine >= 0.999 against
             the oracle's FP64 gradient;
  Adam       the masters after one step equal the oracle's Adam step applied
             to the GPU's own gradient within 4 fp32 ulp of the weight plus
             1e-6 of the learning rate (the GPU evaluates the update in fp32,
             the oracle in fp64; near a zero crossing the weight's ulp is
             smaller than the update's rounding);
  training   deterministic (two gradient calls are bit-identical) and a few
             hundred steps cut the relative L2 loss of a smooth field 10x.
"""
import sys

import numpy as np
import pytest

import oracle_py as orc
import ptamd
from test_oracle_mlp import BMAX, BMIN, DIMS, OFF, B, W, random_params, random_raw

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def fp16_weights(p):
    q = p.copy()
    for l in range(5):
        w0, b0 = OFF[l]
        q[w0:b0] = p[w0:b0].astype(np.float16).astype(np.float32)
    return q


@pytest.fixture()
def cache(ctx):
    ctx.cache_init(1, BMIN, BMAX)
    ctx.cache_set_adam()
    return ctx


def test_param_count():
    assert ptamd.CACHE_PARAMS == orc.mlp_param_count()


def test_params_round_trip(cache):
    p = random_params(0)
    cache.cache_set_params(p)
    np.testing.assert_array_equal(cache.cache_get_params(), p)


@pytest.mark.parametrize("n", [1, 15, 16, 1000, 65536])
def test_infer_matches_oracle(cache, n):
    say(f"[cache] infer n={n}")
    p = random_params(1, bias=0.2)
    W(p, 4)[:] *= 3.0
    cache.cache_set_params(p)
    raw = random_raw(n, 2)
    got = cache.cache_infer(raw)
    want = orc.mlp_forward(fp16_weights(p), raw, BMIN, BMAX)
    rel = np.abs(got - want) / want
    assert rel.max() <= 5e-3, rel.max()
    assert rel.mean() <= 1e-3, rel.mean()


def test_infer_closed_form_path(cache):
    # Same hand-built single path as the oracle test: exp(0.75 * relu(2 n.x)).
    p = np.zeros(ptamd.CACHE_PARAMS, np.float32)
    W(p, 0)[72, 5] = 2.0
    W(p, 1)[5, 9] = 1.0
    W(p, 2)[9, 33] = 0.5
    W(p, 3)[33, 63] = 1.0
    W(p, 4)[63, 1] = 1.5
    B(p, 4)[:] = [np.log(0.5), 0.0, np.log(3.0)]
    cache.cache_set_params(p)
    raw = random_raw(200, 3)
    got = cache.cache_infer(raw)
    nx = raw[:, 3].astype(np.float16).astype(np.float64)
    np.testing.assert_allclose(got[:, 0], 0.5, rtol=1e-6)
    np.testing.assert_allclose(got[:, 2], 3.0, rtol=1e-6)
    np.testing.assert_allclose(got[:, 1], np.exp(1.5 * np.maximum(nx, 0.0)), rtol=3e-3)


@pytest.mark.parametrize("n", [16, 1000, 4099])
def test_gradient_matches_oracle(cache, n):
    say(f"[cache] gradient n={n}")
    p = random_params(4, bias=0.2)
    cache.cache_set_params(p)
    raw = random_raw(n, 5)
    tgt = np.random.default_rng(6).uniform(0.0, 2.0, (n, 3)).astype(np.float32)
    g = cache.cache_gradient(raw, tgt).astype(np.float64)
    _, want = orc.mlp_loss_grad(fp16_weights(p), raw, tgt, BMIN, BMAX)
    for l in range(5):
        w0, b0 = OFF[l]
        for lo, hi in ((w0, b0), (b0, b0 + DIMS[l + 1])):
            a, b = g[lo:hi], want[lo:hi]
            nb = np.linalg.norm(b)
            if nb == 0.0:
                assert np.linalg.norm(a) < 1e-6
                continue
            err = np.linalg.norm(a - b) / nb
            cos = float(a @ b) / (np.linalg.norm(a) * nb)
            assert err <= 2e-2 and cos >= 0.999, (l, lo, err, cos)
    # zero-padded inputs 81..95 get exactly zero weight gradient
    assert np.all(W(g.astype(np.float32), 0)[81:] == 0.0)
    g2 = cache.cache_gradient(raw, tgt)
    np.testing.assert_array_equal(g2, g.astype(np.float32))


def test_adam_step_matches_oracle(cache):
    p = random_params(7, bias=0.2)
    cache.cache_set_params(p)
    raw = random_raw(3000, 8)
    tgt = np.random.default_rng(9).uniform(0.0, 2.0, (3000, 3)).astype(np.float32)
    w = p.copy()
    m = np.zeros_like(w)
    v = np.zeros_like(w)
    for t in (1, 2):
        say(f"[cache] adam step {t}")
        g = cache.cache_gradient(raw, tgt)
        cache.cache_train(raw, tgt)
        orc.adam_step(w, m, v, g.astype(np.float64), t, lr=1e-3, b1=0.9, b2=0.99, eps=1e-8)
        got = cache.cache_get_params()
        ulp = np.spacing(np.abs(w)).astype(np.float64)
        excess = np.abs(got.astype(np.float64) - w) - (4 * ulp + 1e-6 * 1e-3)
        assert excess.max() <= 0, (t, int(np.argmax(excess)), float(excess.max()))


def test_training_reduces_loss(cache):
    rng = np.random.default_rng(12)
    n = 16384
    raw = random_raw(n, 13)
    pos = (raw[:, :3] - BMIN) / (BMAX - BMIN)
    tgt = np.stack([0.5 + 0.4 * np.sin(2 * np.pi * pos[:, 0]),
                    0.3 + 0.2 * np.cos(2 * np.pi * pos[:, 1]) * raw[:, 9],
                    0.2 + raw[:, 4] ** 2], axis=1).astype(np.float32)
    cache.cache_init(77, BMIN, BMAX)
    cache.cache_set_adam(lr=3e-3)
    try:
        first = float(np.mean(cache.cache_train(raw, tgt)))
        say(f"[cache] training, first loss {first:.4f}")
        last = first
        for it in range(300):
            idx = rng.integers(0, n, 4096)
            last = float(np.mean(cache.cache_train(raw[idx], tgt[idx])))
    finally:
        cache.cache_set_adam()
    y = cache.cache_infer(raw)
    full = float(np.mean(np.sum((y - tgt) ** 2 / (y * y + 0.01), axis=1)))
    say(f"[cache] loss {first:.4f} -> {last:.4f} (full batch {full:.4f})")
    assert np.isfinite(full)
    assert full < first / 10.0


def test_cache_errors(ctx):
    with pytest.raises(ptamd.PtError):
        ctx.cache_gradient_device(0, 0, 0, 0)
    with pytest.raises(ptamd.PtError):
        ctx.cache_infer_device(0, 5, 0)
    ctx.cache_infer_device(0, 0, 0)  # empty batch is a no-op
