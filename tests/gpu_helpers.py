"""Shared helpers of the GPU parity tests: build the product model and the
oracle (blocked reduction order) from the same reference tables and step
them side by side."""
from __future__ import annotations

import sys

import numpy as np

import oracle_lib
from adcirc_amd.model import Model

STATE_F = ("eta2", "etas", "uu2", "vv2", "tk", "etamax", "umax")
STATE_I = ("nodecode", "noff", "etamax_it", "umax_it")


def golden_tables(d):
    t = {k: d[k] for k in ("nm", "x", "y", "dp", "fdxe", "fdye", "areas", "nneigh", "neitab",
                           "csr", "nbd", "nbv", "lbcodei", "csii", "siii", "me2gw")}
    for k in ("np", "ne", "mnei", "neta", "nvel", "nvelme"):
        t[k] = int(d[k])
    if "bndlen2o3" in d:
        t["bndlen2o3"] = d["bndlen2o3"]
    return t


def build_pair(d, run, order=oracle_lib.BLOCKED):
    """Product model and oracle model on the same tables and cold start."""
    t = golden_tables(d)
    np_, nbfr = t["np"], run.nbfr
    emo = np.asarray(d["emo"]).reshape(t["neta"], nbfr) if t["neta"] else None
    efa = np.asarray(d["efa"]).reshape(t["neta"], nbfr) if t["neta"] else None
    g = Model(t, run, tau0var=d["tau0var"], evm=d["evm"], fric=d["fric"],
              corif=np.full(np_, run.cori), emo=emo, efa=efa, qnam=d.get("qnam"),
              qnph=d.get("qnph"))
    g.cold_start(d["eta2_0"], d["nodecode0"], d["noff0"])
    o = oracle_lib.OracleModel(d, run, order)
    if run.holland is not None:
        # the cold-start forcing (time 0) is the first step's WSX1/PR1
        g.set("wsx2", o.wsx); g.set("wsy2", o.wsy); g.set("pr2", o.pr)
    return g, o


def step_pair(g, o, it):
    """One step on both; the GPU gets the oracle's host forcing."""
    o.step(it)
    if o.hol is not None:
        g.timestep(it, (o.wsx, o.wsy, o.pr))
    else:
        g.timestep(it)


def compare(g, o, it, fields=STATE_F, ifields=STATE_I):
    bad = []
    for k in fields:
        a, b = g.get(k), o.array(k)
        if not np.array_equal(a, b):
            i = int(np.argmax(a != b))
            bad.append(f"{k}: {int((a != b).sum())} differ, max |d| {np.abs(a - b).max():.3e} "
                       f"(node {i}: {a[i]!r} vs {b[i]!r})")
    for k in ifields:
        a, b = g.get(k), o.iarray(k)
        if not np.array_equal(a, b):
            bad.append(f"{k}: {int((a != b).sum())} differ")
    n, ier = g.numitr()
    if n != o.numitr():
        bad.append(f"numitr {n} vs {o.numitr()}")
    return bad


def log(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def case_setup(case):
    """Oracle-style setup dictionary for a generated case, built with the
    product's own preprocessing (for meshes the reference harness did not run)."""
    from adcirc_amd import prep
    from adcirc_amd.model import cold_wetdry_state

    r = case.run
    t = prep.build_tables(case.mesh, r.nolifa, r.h0, r.anginn)
    npn = int(t["np"])
    d = dict(t)
    d["emo"] = np.asarray(case.emo, np.float64).ravel()
    d["efa"] = (np.asarray(case.efa, np.float64) * (3.141592653589793 / 180.0)).ravel()
    d["tau0var"] = prep.tau0_nodal(r.tau0, t["dp"])
    d["evm"] = np.full(npn, r.eslm)
    d["fric"] = np.full(npn, r.cf)
    e2, nc, nf = cold_wetdry_state(t, r)
    d["nodecode0"], d["noff0"], d["eta2_0"] = nc, nf, e2
    d["prbckgrnd_mh2o"] = 100.0 * 1013.0 / (1000.0 * r.g)
    return d
