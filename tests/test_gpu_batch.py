"""Batches of contexts (adcgpu_batch): storm ensembles and uneven meshes.

The members of a batch take their steps together and their JCG solves run as
one cooperative launch, each system's blocks synchronising among themselves.
Every member must end bit for bit where its own resident run (adcgpu_run on
that member alone) ends: the batch changes the launch shape, never the
arithmetic (the reduction tree depends only on each system's layout).  The
single-member resident run is itself pinned to the oracle bit for bit
(tests/test_gpu_parity.py).
"""
import dataclasses

import numpy as np
import pytest

from adcirc_amd import _lib, meshgen
from gpu_helpers import case_setup, golden_tables, log

pytestmark = pytest.mark.gpu

FIELDS = ("eta1", "eta2", "etas", "uu2", "vv2", "qx2", "qy2", "tk", "etamax", "umax", "coef",
          "alphal", "nodecode", "noff")


def _model(case, d, warm=0, precision=None):
    from adcirc_amd.model import Model
    run = dataclasses.replace(case.run, warm_start=warm)
    t = golden_tables(d)
    nb = run.nbfr
    g = Model(t, run, tau0var=d["tau0var"], evm=d["evm"], fric=d["fric"],
              corif=np.full(t["np"], run.cori), emo=np.asarray(d["emo"]).reshape(t["neta"], nb),
              efa=np.asarray(d["efa"]).reshape(t["neta"], nb),
              precision=_lib.FP64 if precision is None else precision)
    g.cold_start(d["eta2_0"], d["nodecode0"], d["noff0"])
    return g


def _storm(case, k):
    """Member k's vortex: central pressure and radius of maximum winds varied
    (the ensemble's spread), track as the case's."""
    hol = list(case.run.holland)
    hol[0] = hol[0] + 10.0 * ((k % 5) - 2)      # Pc, hPa
    hol[2] = hol[2] * (0.8 + 0.1 * (k % 4))     # Rmax
    return tuple(hol)


def _check_members(members, singles, what):
    for k, (a, b) in enumerate(zip(members, singles)):
        for f in FIELDS:
            x, y = a.get(f), b.get(f)
            assert np.array_equal(x, y), (what, k, f, int((x != y).sum()))
        assert a.run_stats()[:3] == b.run_stats()[:3], (what, k)


def _run_pair(specs, nsteps, it0=1):
    """specs: (case, setup, warm, holland) per member.  Runs them as a batch
    and one by one; returns both lists of models."""
    from adcirc_amd.model import Batch
    members = [_model(c, d, w) for c, d, w, _ in specs]
    singles = [_model(c, d, w) for c, d, w, _ in specs]
    hols = [h for *_, h in specs]
    for m, h in zip(members + singles, hols + hols):
        m.holland(0.0, h)
    b = Batch(members)
    iters = b.run(it0, nsteps, holland=hols)
    b.close()
    for m, h in zip(singles, hols):
        m.run_resident(it0, nsteps, holland=h)
    return members, singles, iters


def test_storm_ensemble_members_equal_their_own_runs():
    """Six storms (different central pressures and radii) on the 20k-node
    coastal mesh, 400 steps as one batch: each member equals its own resident
    run bit for bit (state, COEF, flags, iteration counts)."""
    case = meshgen.coastal(20000)
    d = case_setup(case)
    specs = [(case, d, 0, _storm(case, k)) for k in range(6)]
    members, singles, iters = _run_pair(specs, 400)
    log(f"6-storm batch, 400 steps: JCG iterations per member {iters}; peak surges "
        f"{[round(float(m.get('etamax').max()), 4) for m in members]}")
    assert len(set(round(float(m.get("etamax").max()), 9) for m in members)) > 1
    _check_members(members, singles, "ensemble")


def test_uneven_batch_members_equal_their_own_runs():
    """Three meshes of different size (5k, 20k and 40k nodes) with two storms
    each, stepped as one batch: each member equals its own resident run."""
    specs = []
    for n in (5000, 20000, 40000):
        case = meshgen.coastal(n)
        d = case_setup(case)
        specs += [(case, d, 0, _storm(case, k)) for k in range(2)]
    members, singles, iters = _run_pair(specs, 200)
    log(f"uneven batch (5k, 20k, 40k x 2), 200 steps: JCG iterations {iters}")
    _check_members(members, singles, "uneven")


def test_batch_with_learned_members_and_several_launches():
    """Members with the learned start (two-start solve) and without are
    launched in separate groups; four 300k-node members do not fit in one
    cooperative launch (75 blocks each, 256 resident) and take two.  Every
    member equals its own resident run."""
    case = meshgen.coastal(20000)
    d = case_setup(case)
    specs = [(case, d, 2, _storm(case, 0)), (case, d, 0, _storm(case, 1)),
             (case, d, 2, _storm(case, 2))]
    members, singles, _ = _run_pair(specs, 200)
    _check_members(members, singles, "learned")
    big = meshgen.coastal(300000)
    db = case_setup(big)
    specs = [(big, db, 0, _storm(big, k)) for k in range(4)]
    members, singles, iters = _run_pair(specs, 30)
    log(f"4 x {members[0].np_} nodes, 30 steps: JCG iterations {iters}")
    _check_members(members, singles, "groups")
