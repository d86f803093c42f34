"""Whole-state snapshots (adcgpu_state_save / adcgpu_state_load): a run
continued from a snapshot, in a new context, equals the uninterrupted run bit
for bit (the library's hot start)."""
import numpy as np
import pytest

from adcirc_amd import meshgen
from adcirc_amd.model import Model
from gpu_helpers import case_setup, golden_tables, log

pytestmark = pytest.mark.gpu


def _model(case, d):
    t = golden_tables(d)
    run = case.run
    nb = run.nbfr
    g = Model(t, run, tau0var=d["tau0var"], evm=d["evm"], fric=d["fric"],
              corif=np.full(t["np"], run.cori), emo=np.asarray(d["emo"]).reshape(t["neta"], nb),
              efa=np.asarray(d["efa"]).reshape(t["neta"], nb))
    g.cold_start(d["eta2_0"], d["nodecode0"], d["noff0"])
    g.holland(0.0, run.holland)
    return g, run


def test_restart_from_snapshot_equals_uninterrupted_run():
    case = meshgen.coastal(20000)
    d = case_setup(case)
    a, run = _model(case, d)
    a.run_resident(1, 700, holland=run.holland)
    snap = a.save_state()
    log(f"snapshot at step 700: {snap.size / 1e6:.1f} MB")
    a.run_resident(701, 500, holland=run.holland)
    b, _ = _model(case, d)
    b.load_state(snap)
    b.run_resident(701, 500, holland=run.holland)
    for k in ("eta2", "etas", "uu2", "vv2", "tk", "etamax", "umax", "coef"):
        assert np.array_equal(a.get(k), b.get(k)), k
    for k in ("nodecode", "noff", "etamax_it", "umax_it"):
        assert np.array_equal(a.get(k), b.get(k)), k
    a.close()
    b.close()
