"""The learned JCG start on the GPU (warm_start = 2, adcgpu_ml_*), against the
CPU restatement of its network (oracle/mlp_ref.py, pinned in
tests/test_mlp_ref.py):

- the matrix-core backward pass: loss and gradient of a batch vs float64 with
  the same bfloat16 rounding points;
- the inputs and the fused inference kernel on a live storm state;
- with a zero output layer (an untrained network) the learned mode is the
  plain path bit for bit;
- training on the model's own solves lowers the loss, and the trained
  network's effect on the JCG iteration count over the next steps.
"""
import dataclasses
import os
import sys

import numpy as np
import pytest

from adcirc_amd import meshgen
from adcirc_amd.model import Model
from gpu_helpers import case_setup, golden_tables, log

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(os.path.dirname(HERE), "oracle"))
import mlp_ref as M  # noqa: E402

pytestmark = pytest.mark.gpu


def _model(case, d, warm):
    t = golden_tables(d)
    run = dataclasses.replace(case.run, warm_start=warm)
    nb = run.nbfr
    g = Model(t, run, tau0var=d["tau0var"], evm=d["evm"], fric=d["fric"],
              corif=np.full(t["np"], run.cori), emo=np.asarray(d["emo"]).reshape(t["neta"], nb),
              efa=np.asarray(d["efa"]).reshape(t["neta"], nb))
    g.cold_start(d["eta2_0"], d["nodecode0"], d["noff0"])
    g.holland(0.0, run.holland)
    return g, run


def _random_params(seed, out_scale=0.05):
    rng = np.random.default_rng(seed)
    p = np.zeros(M.NPARAM)
    p[M.OFF_W1:M.OFF_B1] = rng.uniform(-1, 1, M.F * M.H) * np.sqrt(6.0 / (M.F + M.H))
    p[M.OFF_B1:M.OFF_W2] = rng.normal(0, 0.05, M.H)
    p[M.OFF_W2:M.OFF_B2] = rng.uniform(-1, 1, M.H * M.H) * np.sqrt(6.0 / (2 * M.H))
    p[M.OFF_B2:M.OFF_W3] = rng.normal(0, 0.05, M.H)
    p[M.OFF_W3:M.OFF_B3] = rng.normal(0, out_scale, M.H)
    p[M.OFF_B3] = 0.01
    return p.astype(np.float32)


def test_ml_gradient_matches_cpu_fp64():
    case = meshgen.beach()
    d = case_setup(case)
    g, _ = _model(case, d, 0)
    p = _random_params(11, out_scale=0.3)
    g.ml_params(p)
    rng = np.random.default_rng(12)
    n = 2048
    x = rng.normal(0.0, 1.0, (n, M.F)).astype(np.float32)
    t = rng.normal(0.0, 0.3, n).astype(np.float32)
    w = rng.uniform(0.0, 2.0, n).astype(np.float32)
    lg, gg = g.ml_grad(x, t, w)
    lo, go = M.loss_and_grad(p.astype(np.float64), x.astype(np.float64), t.astype(np.float64),
                             w.astype(np.float64), rounding=True)
    log(f"loss gpu {lg:.8e} cpu {lo:.8e}")
    assert abs(lg - lo) <= 1e-4 * abs(lo)
    blocks = {"W1": (M.OFF_W1, M.OFF_B1), "b1": (M.OFF_B1, M.OFF_W2),
              "W2": (M.OFF_W2, M.OFF_B2), "b2": (M.OFF_B2, M.OFF_W3),
              "w3": (M.OFF_W3, M.OFF_B3), "b3": (M.OFF_B3, M.NPARAM)}
    for name, (a, b) in blocks.items():
        rel = np.linalg.norm(gg[a:b] - go[a:b]) / np.linalg.norm(go[a:b])
        log(f"gradient {name}: relative difference {rel:.2e}")
        assert rel <= 5e-3, name
    g.close()


def _state_features(g, t, run):
    np_, mnei = t["np"], t["mnei"]
    col = np.asarray(t["neitab"], np.int64).reshape(mnei, np_) - 1
    bnd = np.full(np_, -1)
    bnd[np.asarray(t["nbd"], np.int64) - 1] = np.arange(t["neta"])
    st = {k: g.get(k) for k in ("eta2", "etas", "etas_prev", "wsx2", "wsy2", "pr1", "pr2", "tk")}
    nc = g.get("nodecode")
    x, s = M.features(np_, t["nneigh"], col.ravel(), bnd, st["eta2"], st["etas"],
                      st["etas_prev"], np.asarray(t["dp"]), st["wsx2"], st["wsy2"], st["pr1"],
                      st["pr2"], st["tk"], nc, run.dtdp)
    return x, s, nc, bnd


def test_ml_inputs_and_inference_match_cpu():
    case = meshgen.coastal(20000)
    d = case_setup(case)
    t = golden_tables(d)
    g, run = _model(case, d, 2)
    g.run_resident(1, 400, holland=run.holland)
    x_o, s_o, nc, bnd = _state_features(g, t, run)
    x_g, s_g = g.ml_features()
    assert np.array_equal(s_g, s_o.astype(np.float32).astype(np.float64))
    diff = x_g != x_o
    log(f"inputs: {int(diff.sum())} of {diff.size} differ (device vs host libm in tanh/log10)")
    assert diff.mean() < 1e-3
    assert np.all(np.abs(x_g - x_o) <= 2.0 ** -7 * np.maximum(np.abs(x_o), 1e-30) + 1e-30)
    p = _random_params(5)
    g.ml_params(p)
    corr_g = g.ml_correction()
    y_o, _ = M.forward(p.astype(np.float64), x_o.astype(np.float64), rounding=True)
    use = (nc == 1) & (bnd < 0)
    corr_o = np.where(use, s_o * y_o, 0.0)
    assert np.all(corr_g[~use] == 0.0)
    rel = np.linalg.norm(corr_g - corr_o) / np.linalg.norm(corr_o)
    log(f"correction: relative difference {rel:.2e} over {int(use.sum())} nodes")
    assert rel <= 1e-2
    g.close()


def test_untrained_learned_mode_is_the_plain_path():
    """Zero output layer (a new context's network): the learned start equals
    ETAS, the plain start is kept (no smaller residual), and 300 storm steps
    are bit-identical to warm_start = 0."""
    case = meshgen.coastal(20000)
    d = case_setup(case)
    a, run = _model(case, d, 0)
    b, _ = _model(case, d, 2)
    a.run_resident(1, 300, holland=run.holland)
    b.run_resident(1, 300, holland=run.holland)
    assert b.run_stats()[4] == 0
    for k in ("eta2", "etas", "uu2", "vv2", "etamax"):
        assert np.array_equal(a.get(k), b.get(k)), k
    for k in ("nodecode", "noff"):
        assert np.array_equal(a.get(k), b.get(k)), k
    a.close()
    b.close()


def test_training_lowers_loss_and_learned_start_effect():
    """Collect samples from 30 minutes of plain steps, train on the GPU, then
    run the next 20 minutes plain and with the trained network; the learned
    run keeps the JCG stopping test, so its state stays within the solver's
    tolerance of the plain one.  The iteration counts are logged."""
    case = meshgen.coastal(20000)
    d = case_setup(case)
    p_model, run = _model(case, d, 0)
    S, K = 1800, 1200
    p_model.run_resident(1, 600, holland=run.holland)
    p_model.ml_collect(512)
    p_model.run_resident(601, S - 600, holland=run.holland)
    p_model.ml_collect(0)
    ns = p_model.ml_samples()
    assert ns == 512 * (S - 600)
    losses = [p_model.ml_train(1, lr=1e-3, seed=k) for k in range(12)]
    log(f"training loss by epoch: {' '.join(f'{v:.3e}' for v in losses)}")
    assert losses[-1] < 0.8 * losses[0]
    params = p_model.ml_params()
    it_plain = p_model.run_resident(S + 1, K, holland=run.holland)
    l_model, _ = _model(case, d, 2)
    l_model.run_resident(1, S, holland=run.holland)
    l_model.ml_params(params)
    it_learned = l_model.run_resident(S + 1, K, holland=run.holland)
    used = l_model.run_stats()[4]
    log(f"JCG iterations over steps {S + 1}-{S + K}: plain {it_plain} "
        f"({it_plain / K:.3f}/step), learned {it_learned} ({it_learned / K:.3f}/step), "
        f"learned start used in {used} of {K} solves")
    de = np.abs(l_model.get("eta2") - p_model.get("eta2")).max()
    log(f"max |d eta| learned vs plain after {K} steps: {de:.2e} m")
    assert de < 1e-4
    assert it_learned <= it_plain * 1.02
    p_model.close()
    l_model.close()
