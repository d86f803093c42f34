"""CPU restatement (numpy, float64) of the learned JCG start's network.

TEST INFRASTRUCTURE ONLY: tests use it as the checker of the GPU kernels in
adcirc_amd/csrc/mlp.hip; the product never imports it.

The network has no counterpart in the reference.  It only proposes the
initial guess of the reference's JCG solve (ETAS, src/gwce.F:2002-2004); the
stopping test (src/itpackv.F:809-890) is unchanged.  Its definition is the
build's own:

  inputs   48 features per node (see features()), rounded to bfloat16;
  layer 1  Z1 = X W1 + b1, W1 [48][128] rounded to bfloat16; A1 = gelu(Z1)
           rounded to bfloat16;
  layer 2  Z2 = A1 W2 + b2, W2 [128][128] rounded to bfloat16; A2 = gelu(Z2);
  output   y = A2 . w3 + b3, w3 rounded to bfloat16;
  loss     sum_b w_b (y_b - t_b)^2 / n;
  gelu     the tanh form 0.5 x (1 + tanh(sqrt(2/pi) (x + 0.044715 x^3)))
           (Hendrycks & Gimpel 2016).

`rounding=False` drops every bfloat16 rounding: the plain float64 network,
whose gradient tests check against finite differences.
"""
from __future__ import annotations

import numpy as np

F, H, NB = 48, 128, 12
OFF_W1 = 0
OFF_B1 = F * H
OFF_W2 = OFF_B1 + H
OFF_B2 = OFF_W2 + H * H
OFF_W3 = OFF_B2 + H
OFF_B3 = OFF_W3 + H
NPARAM = OFF_B3 + 1
C0 = 0.7978845608028654
C1 = 0.044715


def bf16(x) -> np.ndarray:
    """Round float32 values to bfloat16 (nearest, ties to even); float64 out."""
    a = np.ascontiguousarray(np.asarray(x, np.float32))
    u = a.view(np.uint32).astype(np.uint64)
    r = ((u + 0x7FFF + ((u >> 16) & 1)) >> 16) << 16
    return r.astype(np.uint32).view(np.float32).astype(np.float64)


def gelu(x):
    t = np.tanh(C0 * (x + C1 * x ** 3))
    return 0.5 * x * (1.0 + t)


def gelu_grad(x):
    t = np.tanh(C0 * (x + C1 * x ** 3))
    return 0.5 * (1.0 + t) + 0.5 * x * (1.0 - t * t) * C0 * (1.0 + 3.0 * C1 * x * x)


def unpack(p):
    p = np.asarray(p, np.float64)
    return (p[OFF_W1:OFF_B1].reshape(F, H), p[OFF_B1:OFF_W2], p[OFF_W2:OFF_B2].reshape(H, H),
            p[OFF_B2:OFF_W3], p[OFF_W3:OFF_B3], p[OFF_B3])


def _r(x, rounding):
    return bf16(x) if rounding else np.asarray(x, np.float64)


def forward(p, x, rounding=True):
    """y for inputs x [n][48]; also the intermediates of the backward pass."""
    W1, b1, W2, b2, w3, b3 = unpack(p)
    if rounding:
        # the kernels keep b1, b2, b3 in float32
        b1, b2, b3 = (np.asarray(v, np.float32).astype(np.float64) for v in (b1, b2, b3))
    X = _r(x, rounding)
    Z1 = X @ _r(W1, rounding) + b1
    A1 = _r(gelu(Z1), rounding)
    Z2 = A1 @ _r(W2, rounding) + b2
    A2 = gelu(Z2)
    y = A2 @ _r(w3, rounding) + b3
    return y, (X, Z1, A1, Z2, A2)


def loss_and_grad(p, x, t, w, rounding=True):
    """Loss sum w (y - t)^2 / n and its gradient in the parameters' order.
    With rounding, dZ2 and dZ1 are rounded to bfloat16 where the kernels
    round them (the operands of the weight-gradient and dA1 products); the
    bias gradients use the unrounded values, as the kernels do."""
    W1, b1, W2, b2, w3, b3 = unpack(p)
    n = x.shape[0]
    y, (X, Z1, A1, Z2, A2) = forward(p, x, rounding)
    err = y - t
    loss = float(np.sum(w * err * err) / n)
    dy = 2.0 * w * err / n
    w3r = _r(w3, rounding)
    gw3 = A2.T @ dy
    gb3 = dy.sum()
    dZ2 = dy[:, None] * w3r[None, :] * gelu_grad(Z2)
    gb2 = dZ2.sum(0)
    dZ2r = _r(dZ2, rounding)
    gW2 = A1.T @ dZ2r
    dA1 = dZ2r @ _r(W2, rounding).T
    dZ1 = dA1 * gelu_grad(Z1)
    gb1 = dZ1.sum(0)
    gW1 = X.T @ _r(dZ1, rounding)
    g = np.concatenate([gW1.ravel(), gb1, gW2.ravel(), gb2, gw3, [gb3]])
    return loss, g


def feature_k(h, q):
    """Input index of local feature q (0..23) of half h of a lane pair."""
    return 16 * (q // 8) + 8 * h + (q % 8)


def features(np_, nneigh, col, bnd_index, eta, etas, etasp, dp, wsx2, wsy2, pr1, pr2, tk,
             nodecode, dt):
    """The 48 inputs of every node (kernel input order, bfloat16 values) and
    its scale s.  col is the 0-based neighbour table [mnei][np] (slot 0 = the
    node), nneigh its row lengths.  Half 0 of the inputs: stencil nodes 0..6
    (the node, then NEITAB neighbours 1..6), three each: ETAS/s,
    (ETAS - ETAS_prev)/s, clamp((eta_j - eta_i)/(64 s), +-8); then tanh(DP/50),
    clamp(1000 WSX2, +-8), clamp(1000 WSY2, +-8).  Half 1: stencil nodes 7..12
    likewise, then clamp((PR2 - PR1)/s, +-8), NODECODE, (NNEIGH-1)/12,
    log10(s)/10, the wet fraction of the neighbours, TK*DT.  Missing
    neighbours contribute zeros.  s = sqrt(sum over the 13 stencil slots of
    ETAS^2 + (ETAS - ETAS_prev)^2, / 13) + 1e-12."""
    nneigh = np.asarray(nneigh, np.int64)
    col = np.asarray(col, np.int64).reshape(-1, np_)
    sten = np.full((np_, NB + 1), -1, np.int64)
    sten[:, 0] = np.arange(np_)
    for j in range(1, min(col.shape[0], NB + 1)):
        sten[:, j] = np.where(j < nneigh, col[j], -1)
    m = sten >= 0
    si = np.where(m, sten, 0)
    es = np.where(m, etas[si], 0.0)
    ep = np.where(m, etasp[si], 0.0)
    d2 = es - ep
    tot = np.zeros(np_)
    for j in range(0, 7):
        tot = tot + (es[:, j] * es[:, j] + d2[:, j] * d2[:, j])
    tot2 = np.zeros(np_)
    for j in range(7, NB + 1):
        tot2 = tot2 + (es[:, j] * es[:, j] + d2[:, j] * d2[:, j])
    s = np.sqrt((tot + tot2) / 13.0) + 1e-12
    rs = 1.0 / s
    out = np.zeros((np_, F), np.float64)
    for h, (lo, hi) in enumerate(((0, 7), (7, NB + 1))):
        for q, j in enumerate(range(lo, hi)):
            a = np.where(m[:, j], es[:, j] * rs, 0.0)
            b = np.where(m[:, j], d2[:, j] * rs, 0.0)
            c = np.where(m[:, j], np.clip((eta[si[:, j]] - eta) * rs * (1.0 / 64.0), -8.0, 8.0),
                         0.0)
            out[:, feature_k(h, 3 * q)] = a
            out[:, feature_k(h, 3 * q + 1)] = b
            out[:, feature_k(h, 3 * q + 2)] = c
    out[:, feature_k(0, 21)] = np.tanh(dp / 50.0)
    out[:, feature_k(0, 22)] = np.clip(wsx2 * 1000.0, -8.0, 8.0)
    out[:, feature_k(0, 23)] = np.clip(wsy2 * 1000.0, -8.0, 8.0)
    wet = np.zeros(np_, np.int64)
    for j in range(1, col.shape[0]):
        ok = j < nneigh
        wet = wet + np.where(ok, np.asarray(nodecode)[np.where(ok, col[j], 0)], 0)
    nn1 = np.maximum(nneigh - 1, 1).astype(np.float32)
    out[:, feature_k(1, 18)] = np.clip((pr2 - pr1) * rs, -8.0, 8.0)
    out[:, feature_k(1, 19)] = nodecode
    out[:, feature_k(1, 20)] = np.asarray((nneigh - 1), np.float32) / np.float32(12.0)
    out[:, feature_k(1, 21)] = np.log10(s) / 10.0
    out[:, feature_k(1, 22)] = np.where(nneigh > 1, wet.astype(np.float32) / nn1, 0.0)
    out[:, feature_k(1, 23)] = tk * dt
    return bf16(out.astype(np.float32)), s
