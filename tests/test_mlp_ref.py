"""Pins of the CPU restatement of the learned start's network
(oracle/mlp_ref.py), before it is trusted as the checker of the GPU kernels:
its gradient against central differences, bfloat16 rounding against known
values, its GELU against the exact (erf) GELU, and its vectorised features
against a per-node loop."""
import math
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(os.path.dirname(HERE), "oracle"))
import mlp_ref as M  # noqa: E402


def _params(rng, scale=0.3):
    p = rng.normal(0.0, scale, M.NPARAM)
    p[M.OFF_W1:M.OFF_B1] *= 1.0 / math.sqrt(M.F)
    p[M.OFF_W2:M.OFF_B2] *= 1.0 / math.sqrt(M.H)
    return p


def test_gradient_matches_central_differences():
    rng = np.random.default_rng(3)
    p = _params(rng)
    x = rng.normal(0.0, 1.0, (40, M.F))
    t = rng.normal(0.0, 0.5, 40)
    w = rng.uniform(0.2, 2.0, 40)
    _, g = M.loss_and_grad(p, x, t, w, rounding=False)
    idx = np.concatenate([rng.choice(M.NPARAM - 1, 60, replace=False), [M.OFF_B3],
                          M.OFF_B1 + np.arange(3), M.OFF_W3 + np.arange(3)])
    h = 1e-6
    for k in idx:
        pp, pm = p.copy(), p.copy()
        pp[k] += h
        pm[k] -= h
        fd = (M.loss_and_grad(pp, x, t, w, False)[0] - M.loss_and_grad(pm, x, t, w, False)[0]) / (2 * h)
        assert abs(fd - g[k]) <= 1e-6 * max(1.0, abs(fd)), (k, fd, g[k])


def test_bf16_rounding_known_values():
    one = 1.0
    ulp = 2.0 ** -7           # bfloat16 spacing at 1
    vals = np.array([one, one + ulp / 2, one + 1.5 * ulp, one + ulp / 4, -2.0 - 2.0 ** -7,
                     3.0e38], np.float32)
    got = M.bf16(vals)
    assert got[0] == 1.0
    assert got[1] == 1.0                  # tie -> even (mantissa 0)
    assert got[2] == one + 2 * ulp        # tie -> even (upward)
    assert got[3] == 1.0
    assert got[4] == -2.0                 # spacing at 2 is 2^-6: the tie goes to even
    assert abs(got[5] / 3.0e38 - 1.0) < 2.0 ** -8


def test_gelu_is_the_erf_gelu_to_tanh_form_accuracy():
    x = np.linspace(-6.0, 6.0, 2001)
    exact = 0.5 * x * (1.0 + np.array([math.erf(v / math.sqrt(2.0)) for v in x]))
    assert np.abs(M.gelu(x) - exact).max() < 1e-3
    # its derivative against central differences
    h = 1e-6
    fd = (M.gelu(x + h) - M.gelu(x - h)) / (2 * h)
    assert np.abs(fd - M.gelu_grad(x)).max() < 1e-7


def test_input_order_is_a_bijection():
    ks = sorted(M.feature_k(h, q) for h in (0, 1) for q in range(24))
    assert ks == list(range(48))


def test_features_against_per_node_loop():
    rng = np.random.default_rng(5)
    n, mnei = 40, 9
    nneigh = rng.integers(1, mnei + 1, n)
    col = np.zeros((mnei, n), np.int64)
    for i in range(n):
        col[0, i] = i
        col[1:nneigh[i], i] = rng.choice(np.setdiff1d(np.arange(n), [i]), nneigh[i] - 1,
                                         replace=False)
    eta, es, ep = rng.normal(0, 0.5, n), rng.normal(0, 1e-3, n), rng.normal(0, 1e-3, n)
    dp, wx, wy = rng.uniform(-3, 4000, n), rng.normal(0, 3e-3, n), rng.normal(0, 3e-3, n)
    p1, p2, tk = rng.normal(10.3, 0.1, n), rng.normal(10.3, 0.1, n), rng.uniform(0, 0.01, n)
    nc = rng.integers(0, 2, n)
    x, s = M.features(n, nneigh, col.ravel(), None, eta, es, ep, dp, wx, wy, p1, p2, tk, nc, 1.0)
    for i in range(n):
        st = [i] + [int(col[j, i]) for j in range(1, 13) if j < nneigh[i]]
        st = st + [-1] * (13 - len(st))
        acc = [0.0, 0.0]
        for j, v in enumerate(st):
            if v >= 0:
                acc[0 if j < 7 else 1] += es[v] ** 2 + (es[v] - ep[v]) ** 2
        si = math.sqrt((acc[0] + acc[1]) / 13.0) + 1e-12
        assert s[i] == si
        f = np.zeros(48)
        for j, v in enumerate(st):
            if v < 0:
                continue
            h, q = (0, j) if j < 7 else (1, j - 7)
            f[M.feature_k(h, 3 * q)] = es[v] * (1.0 / si)
            f[M.feature_k(h, 3 * q + 1)] = (es[v] - ep[v]) * (1.0 / si)
            f[M.feature_k(h, 3 * q + 2)] = min(8.0, max(-8.0, (eta[v] - eta[i]) * (1.0 / si) / 64.0))
        f[M.feature_k(0, 21)] = math.tanh(dp[i] / 50.0)
        f[M.feature_k(0, 22)] = min(8.0, max(-8.0, wx[i] * 1000.0))
        f[M.feature_k(0, 23)] = min(8.0, max(-8.0, wy[i] * 1000.0))
        f[M.feature_k(1, 18)] = min(8.0, max(-8.0, (p2[i] - p1[i]) * (1.0 / si)))
        f[M.feature_k(1, 19)] = nc[i]
        f[M.feature_k(1, 20)] = np.float32(nneigh[i] - 1) / np.float32(12.0)
        f[M.feature_k(1, 21)] = math.log10(si) / 10.0
        wet = sum(nc[col[j, i]] for j in range(1, nneigh[i]))
        f[M.feature_k(1, 22)] = (np.float32(wet) / np.float32(nneigh[i] - 1)) if nneigh[i] > 1 else 0.0
        f[M.feature_k(1, 23)] = tk[i] * 1.0
        want = M.bf16(f.astype(np.float32))
        # the vectorised form rounds the same float64 values (one ulp of
        # float32 apart at most where the operation order differs)
        assert np.abs(x[i] - want).max() <= 2.0 ** -7 * np.abs(want).max(), i
