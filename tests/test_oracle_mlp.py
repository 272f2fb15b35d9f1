"""Pins the FP64 oracle of the radiance-cache MLP (CPU only).

The reference has no learned component; the cache is this build's own
network, so the oracle is pinned by mathematics rather than by reference
vectors:
  or_mlp_param_count  closed form 96*64+64 + 3*(64*64+64) + 64*3+3
  or_mlp_encode       sin/cos(2^k pi p) at dyadic p with exact values
                      (0, 1/4, 1/2 ...), index order a*24 + 2k (+1 for cos),
                      raw normal/view/albedo copied to 72..80, zeros after
  or_mlp_forward      hand-built weights whose output is known in closed form:
                      bias-only network = exp(b4); a single path through
                      W0[72][5] -> W1[5][9] -> ... -> W4[j][c] gives
                      exp(w * relu(normal.x)) -- a transposed W[in][out] or a
                      dropped ReLU fails it
  or_mlp_loss_grad    loss value against the definition from the forward; the
                      gradient against central finite differences of the loss
                      with the denominator frozen at the current prediction
                      (the stop-gradient of the relative L2 loss)
  or_adam_step        Kingma & Ba closed forms: constant gradient moves every
                      weight by lr*sign(g) each step; a zero gradient after g
                      moves it by lr*(b1/(1+b1))/sqrt(b2/(1+b2))*sign(g)
"""
import math

import numpy as np
import pytest

import oracle_py as orc

DIMS = [96, 64, 64, 64, 64, 3]
BMIN = np.array([-2.0, -1.0, 0.0], np.float32)
BMAX = np.array([2.0, 3.0, 8.0], np.float32)


def offsets():
    out, o = [], 0
    for l in range(5):
        w = o
        b = o + DIMS[l] * DIMS[l + 1]
        out.append((w, b))
        o = b + DIMS[l + 1]
    return out


OFF = offsets()
NP = orc.mlp_param_count()


def W(params, l):
    w0, b0 = OFF[l]
    return params[w0:b0].reshape(DIMS[l], DIMS[l + 1])


def B(params, l):
    _, b0 = OFF[l]
    return params[b0:b0 + DIMS[l + 1]]


def random_params(seed, bias=0.1):
    rng = np.random.default_rng(seed)
    p = np.zeros(NP, np.float32)
    for l in range(5):
        a = math.sqrt(6.0 / DIMS[l]) * (1.0 if l < 4 else 0.3)
        W(p, l)[:] = rng.uniform(-a, a, (DIMS[l], DIMS[l + 1]))
        B(p, l)[:] = rng.uniform(-bias, bias, DIMS[l + 1])
    return p


def random_raw(n, seed):
    rng = np.random.default_rng(seed)
    raw = np.zeros((n, 12), np.float32)
    raw[:, :3] = rng.uniform(BMIN, BMAX, (n, 3))
    for k in (3, 6):
        v = rng.normal(size=(n, 3))
        raw[:, k:k + 3] = v / np.linalg.norm(v, axis=1, keepdims=True)
    raw[:, 9:] = rng.uniform(0.2, 0.9, (n, 3))
    return raw


def test_param_count():
    assert NP == 96 * 64 + 64 + 3 * (64 * 64 + 64) + 64 * 3 + 3 == 18883


@pytest.mark.parametrize("p,k,s,c", [
    (0.0, 0, 0.0, 1.0), (0.0, 11, 0.0, 1.0),
    (0.5, 0, 1.0, 0.0),      # sin(pi/2)
    (0.5, 1, 0.0, -1.0),     # sin(pi), cos(pi)
    (0.25, 1, 1.0, 0.0),     # 2 * pi * 0.25
    (0.25, 2, 0.0, -1.0),    # 4 * pi * 0.25
    (0.125, 3, 0.0, -1.0),   # 8 * pi / 8
    (1.0 / 6.0, 0, 0.5, math.sqrt(3) / 2),
])
def test_encode_dyadic(p, k, s, c):
    for a in range(3):
        raw = np.zeros(12, np.float32)
        raw[:3] = [0.0, 0.0, 0.0]
        raw[a] = np.float32(BMIN[a] + p * (BMAX[a] - BMIN[a]))
        x = orc.mlp_encode(raw, BMIN, BMAX)
        pf = float(np.float32((raw[a] - BMIN[a]) * np.float32(np.float32(1.0) / (BMAX[a] - BMIN[a]))))
        assert pf == pytest.approx(p, abs=2e-7)
        # exact p gives exact values; the fp32 mapping moves p by <= 1 ulp
        tol = 2e-6 * (2 ** k)
        assert x[a * 24 + 2 * k] == pytest.approx(s, abs=tol)
        assert x[a * 24 + 2 * k + 1] == pytest.approx(c, abs=tol)


def test_encode_tail_layout():
    raw = np.arange(12, dtype=np.float32) * 0.1
    raw[:3] = BMIN
    x = orc.mlp_encode(raw, BMIN, BMAX)
    for a in range(3):
        assert np.all(x[a * 24:(a + 1) * 24:2] == 0.0)   # sin(0)
        assert np.all(x[a * 24 + 1:(a + 1) * 24:2] == 1.0)  # cos(0)
    np.testing.assert_array_equal(x[72:81], raw[3:12].astype(np.float64))
    assert np.all(x[81:] == 0.0)


def test_forward_bias_only():
    p = np.zeros(NP, np.float32)
    B(p, 4)[:] = [np.log(0.5), 0.0, np.log(3.0)]
    raw = random_raw(7, 1)
    out = orc.mlp_forward(p, raw, BMIN, BMAX)
    np.testing.assert_allclose(out, np.tile([0.5, 1.0, 3.0], (7, 1)), rtol=1e-7)


@pytest.mark.parametrize("c", [0, 1, 2])
def test_forward_single_path(c):
    p = np.zeros(NP, np.float32)
    W(p, 0)[72, 5] = 2.0        # normal.x -> unit 5
    W(p, 1)[5, 9] = 1.0
    W(p, 2)[9, 33] = 0.5
    W(p, 3)[33, 63] = 1.0
    W(p, 4)[63, c] = 1.5
    raw = random_raw(64, 2)
    out = orc.mlp_forward(p, raw, BMIN, BMAX)
    h = np.maximum(2.0 * raw[:, 3].astype(np.float64), 0.0) * 0.5
    want = np.ones((64, 3))
    want[:, c] = np.exp(1.5 * h)
    np.testing.assert_allclose(out, want, rtol=1e-12)
    # negative normals are cut by the ReLU
    assert np.any(raw[:, 3] < 0) and np.all(out[raw[:, 3] < 0] == 1.0)


def test_loss_value():
    p = random_params(3)
    raw = random_raw(33, 4)
    tgt = np.random.default_rng(5).uniform(0.0, 2.0, (33, 3)).astype(np.float32)
    loss, _ = orc.mlp_loss_grad(p, raw, tgt, BMIN, BMAX)
    y = orc.mlp_forward(p, raw, BMIN, BMAX)
    want = np.mean(np.sum((y - tgt) ** 2 / (y * y + 0.01), axis=1))
    assert loss == pytest.approx(want, rel=1e-12)


def frozen_loss(p, raw, tgt, den):
    y = orc.mlp_forward(p, raw, BMIN, BMAX)
    return np.mean(np.sum((y - tgt) ** 2 / den, axis=1))


def test_gradient_finite_differences():
    p = random_params(6, bias=0.3)
    raw = random_raw(24, 7)
    tgt = np.random.default_rng(8).uniform(0.0, 2.0, (24, 3)).astype(np.float32)
    _, g = orc.mlp_loss_grad(p, raw, tgt, BMIN, BMAX)
    y0 = orc.mlp_forward(p, raw, BMIN, BMAX)
    den = y0 * y0 + 0.01
    rng = np.random.default_rng(9)
    checked = 0
    for l in range(5):
        w0, b0 = OFF[l]
        idx = list(rng.integers(w0, b0, 6)) + list(rng.integers(b0, b0 + DIMS[l + 1], 2))
        # inputs 81..95 are zero: their weights have an exactly zero gradient
        if l == 0:
            idx.append(w0 + 90 * 64 + 3)
        for k in idx:
            h = np.float32(2.0 ** -12)
            pp, pm = p.copy(), p.copy()
            pp[k] += h
            pm[k] -= h
            step = float(pp[k]) - float(pm[k])
            fd = (frozen_loss(pp, raw, tgt, den) - frozen_loss(pm, raw, tgt, den)) / step
            assert g[k] == pytest.approx(fd, rel=2e-3, abs=1e-7), (l, k)
            checked += 1
    assert checked >= 40
    assert g[OFF[0][0] + 90 * 64 + 3] == 0.0


def test_adam_constant_gradient():
    rng = np.random.default_rng(10)
    n = 257
    w = rng.normal(size=n).astype(np.float32)
    w0 = w.copy()
    m = np.zeros(n, np.float32)
    v = np.zeros(n, np.float32)
    g = rng.normal(size=n) * 0.37
    lr = 1e-3
    for t in (1, 2, 3):
        orc.adam_step(w, m, v, g, t, lr=lr, b1=0.9, b2=0.99, eps=1e-12)
        np.testing.assert_allclose(w.astype(np.float64), w0 - t * lr * np.sign(g), rtol=0, atol=3e-6)


def test_adam_zero_after_gradient():
    rng = np.random.default_rng(11)
    n = 64
    w = rng.normal(size=n).astype(np.float32)
    m = np.zeros(n, np.float32)
    v = np.zeros(n, np.float32)
    g = rng.normal(size=n)
    b1, b2, lr = 0.9, 0.99, 1e-3
    orc.adam_step(w, m, v, g, 1, lr=lr, b1=b1, b2=b2, eps=1e-12)
    w1 = w.astype(np.float64).copy()
    orc.adam_step(w, m, v, np.zeros(n), 2, lr=lr, b1=b1, b2=b2, eps=1e-12)
    want = lr * (b1 / (1 + b1)) / math.sqrt(b2 / (1 + b2))
    np.testing.assert_allclose(w1 - w.astype(np.float64), want * np.sign(g), rtol=0, atol=3e-7)
