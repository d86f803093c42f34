"""Error paths of the GPU library, forced from the host so that every kernel
still ends normally:

- the drop-in JCG reports a failed launch (a wait bound of zero makes every
  wait inside the persistent kernel fail at once) without touching the
  caller's arrays, and the next call starts clean;
- the resident run stops on the reference's ErrorElev check
  (src/timestep.F:1242-1251, 1381-1395) at the step where the oracle's largest
  wet-node |ETA2| first exceeds it;
- a non-finite velocity (src/timestep.F:1413-1471) is reported with its step.
"""
import dataclasses

import numpy as np
import pytest

import oracle_lib
from adcirc_amd import _lib, meshgen
from adcirc_amd.model import jcg, set_wait_limit
from gpu_helpers import build_pair, log

pytestmark = pytest.mark.gpu


def _system(g, name="basin0"):
    J, Cm = g[f"{name}_in_jcoef"], g[f"{name}_in_coef"]
    itmax, zeta = g[f"{name}_prm"]
    iparm = np.zeros(12, np.int32)
    rparm = np.zeros(12)
    oracle_lib.load().adco_dfault(iparm.ctypes.data_as(oracle_lib._ip),
                                  rparm.ctypes.data_as(oracle_lib._dp))
    iparm[0], iparm[1], iparm[3] = int(itmax), 0, 16
    rparm[0] = zeta
    return [np.ascontiguousarray(np.asarray(Cm, np.float64).T.ravel()),
            np.ascontiguousarray(J.astype(np.int32).T.ravel()),
            np.array(g[f"{name}_in_rhs"], np.float64), np.array(g[f"{name}_in_u"], np.float64),
            iparm, rparm]


def _diag_system(g):
    return _system(g, "diag")


def test_dropin_failed_launch_is_reported(jcg_golden):
    g = jcg_golden
    # reference result: the diagonal system (it clears ITPACK's saved flag),
    # then the basin's GWCE system
    jcg(*_diag_system(g))
    ref = _system(g)
    ier_ref = jcg(*ref)
    assert ier_ref == 0
    jcg(*_diag_system(g))
    bad = _system(g)
    before = [x.copy() for x in bad]
    log("drop-in JCG with a wait bound of zero (every wait fails at once)")
    set_wait_limit(0.0)
    try:
        rc = jcg(*bad, raise_errors=False)
    finally:
        set_wait_limit(0.5)
    assert rc == _lib.E_HIP
    assert b"timed out" in _lib.load().adcgpu_last_error()
    for x, y in zip(bad, before):
        assert np.array_equal(x, y)   # the caller's arrays are untouched
    # the next call starts clean and solves as if the failed one never ran
    again = _system(g)
    assert jcg(*again) == ier_ref
    for x, y in zip(again, ref):
        assert np.array_equal(x, y)


def _beach_model(errorelev):
    d = dict(np.load(__import__("conftest").GOLDEN + "/beach.npz"))
    run = dataclasses.replace(meshgen.beach().run, errorelev=errorelev)
    g, o = build_pair(d, run)
    return d, run, g, o


def test_errorelev_stops_the_run_at_the_oracles_step():
    """ErrorElev lowered to 2 cm on the beach storm.  Per-step calls (host
    forcing, bit-exact with the oracle): the status turns to
    ADCGPU_E_ERRORELEV at exactly the first step whose largest wet-node |ETA2|
    exceeds 2 cm in the oracle's run (the reference stops there).  A resident
    run (device forcing) returns the same error and names that step."""
    from gpu_helpers import step_pair
    lim = 0.02
    d, run, g, o = _beach_model(lim)
    lib = _lib.load()
    first = None
    for it in range(1, 601):
        step_pair(g, o, it)
        wet = o.iarray("nodecode") == 1
        over = np.abs(o.array("eta2")[wet]).max() > lim
        rc = lib.adcgpu_last_numitr(g.h, None, None)
        assert (rc == _lib.E_ERRORELEV) == over, (it, rc)
        if over:
            first = it
            break
    assert first is not None and first > 1
    log(f"ErrorElev {lim} m: oracle and per-step path exceed it first at step {first}")
    g.close()
    d, run, g, o = _beach_model(lim)
    g.holland(0.0, run.holland)
    rc = g.lib.adcgpu_run(g.h, 1, first + 10, _table(g, 1, first + 10), _hol(run), 0)
    assert rc == _lib.E_ERRORELEV, lib.adcgpu_last_error()
    assert abs(g.run_stats()[3] - first) <= 1
    g.close()


def _table(g, it0, n):
    t = np.ascontiguousarray(g.eta_bc_table(it0, n), np.float64)
    _table.keep = t
    return t.ctypes.data_as(_lib.dp_)


def _hol(run):
    h = np.ascontiguousarray(list(run.holland) + [0.0] * (13 - len(run.holland)), np.float64)
    _hol.keep = h
    return h.ctypes.data_as(_lib.dp_)


def test_nonfinite_velocity_is_reported():
    d, run, g, o = _beach_model(0.0)
    g.holland(0.0, run.holland)
    assert g.lib.adcgpu_run(g.h, 1, 5, _table(g, 1, 5), _hol(run), 0) == 0
    u = g.get("uu2")
    wet = np.nonzero(g.get("nodecode") == 1)[0]
    u[wet[len(wet) // 2]] = np.nan
    g.set("uu2", u)
    log("NaN velocity at one wet node, one resident step")
    rc = g.lib.adcgpu_run(g.h, 6, 1, _table(g, 6, 1), _hol(run), 0)
    assert rc == _lib.E_NONFINITE, _lib.load().adcgpu_last_error()
    assert g.run_stats()[3] == 6
    g.close()


def test_graph_request_is_refused():
    d, run, g, o = _beach_model(0.0)
    rc = g.lib.adcgpu_run(g.h, 1, 1, _table(g, 1, 1), _hol(run), 1)
    assert rc == _lib.E_UNSUPPORTED
    g.close()
