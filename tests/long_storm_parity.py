"""Six simulated hours of the coastal storm case, GPU against oracle, bit for
bit (a long-horizon check run by hand; pytest does not collect this file).

The oracle (blocked reduction order) takes about half an hour of one CPU core
at 100k nodes, so the two sides run separately:

    python tests/long_storm_parity.py oracle DIR [NODES] [STEPS]   # CPU, writes DIR/oracle_*.npz
    python tests/long_storm_parity.py gpu DIR [NODES] [STEPS]      # MI355X, compares against it

The GPU side steps through the per-step C ABI (adcgpu_timestep) with the
oracle's host Holland forcing, so both sides see identical inputs; every
simulated hour it compares ETA2, UU2, VV2, TK, ETAS, ETAMAX, UMAX, NODECODE,
NOFF, the steps of the last ETAMAX/UMAX rise and NUMITR, and exits non-zero
on any difference.
"""
from __future__ import annotations

import os
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(HERE))
sys.path.insert(0, HERE)

import oracle_lib  # noqa: E402
from gpu_helpers import case_setup, golden_tables  # noqa: E402
from adcirc_amd import meshgen  # noqa: E402

F = ("eta2", "uu2", "vv2", "tk", "etas", "etamax", "umax")
I = ("nodecode", "noff", "etamax_it", "umax_it")
EVERY = 3600


def main(argv) -> int:
    mode, out_dir = argv[0], argv[1]
    nodes = int(argv[2]) if len(argv) > 2 else 100000
    nsteps = int(argv[3]) if len(argv) > 3 else 21600
    path = os.path.join(out_dir, f"oracle_{nodes}_{nsteps}.npz")
    case = meshgen.coastal(nodes)
    d = case_setup(case)
    run = case.run
    t0 = time.time()
    if mode == "oracle":
        o = oracle_lib.OracleModel(d, run, oracle_lib.BLOCKED)
        res = {}
        for it in range(1, nsteps + 1):
            o.step(it)
            if it % EVERY == 0 or it == nsteps:
                for k in F:
                    res[f"s{it}_{k}"] = o.array(k).copy()
                for k in I:
                    res[f"s{it}_{k}"] = o.iarray(k).copy()
                res[f"s{it}_numitr"] = np.array(o.numitr())
                np.savez_compressed(path, **res)
                print(f"oracle step {it}, {time.time() - t0:.0f} s", flush=True)
        return 0
    from adcirc_amd.model import Model
    ref = np.load(path)
    t = golden_tables(d)
    nb = run.nbfr
    g = Model(t, run, tau0var=d["tau0var"], evm=d["evm"], fric=d["fric"],
              corif=np.full(t["np"], run.cori), emo=np.asarray(d["emo"]).reshape(t["neta"], nb),
              efa=np.asarray(d["efa"]).reshape(t["neta"], nb))
    g.cold_start(d["eta2_0"], d["nodecode0"], d["noff0"])
    o = oracle_lib.OracleModel(d, run, oracle_lib.BLOCKED)  # host forcing only
    g.set("wsx2", o.wsx)
    g.set("wsy2", o.wsy)
    g.set("pr2", o.pr)
    bad = 0
    for it in range(1, nsteps + 1):
        g.timestep(it, o.forcing(it * run.dtdp + run.statim * 86400.0))
        if it % EVERY == 0 or it == nsteps:
            msg = []
            for k in F + I:
                a, b = g.get(k), ref[f"s{it}_{k}"]
                nd = int((a != b).sum())
                bad += nd
                msg.append(f"{k} {nd} differ")
            n, _ = g.numitr()
            bad += int(n != int(ref[f"s{it}_numitr"]))
            msg.append(f"numitr {n} vs {int(ref[f's{it}_numitr'])}")
            print(f"step {it} ({time.time() - t0:.0f} s): " + "; ".join(msg), flush=True)
    print(f"peak surge {g.get('etamax').max():.4f} m, wet nodes {int(g.get('nodecode').sum())}, "
          f"differing values {bad}", flush=True)
    g.close()
    return 1 if bad else 0


if __name__ == "__main__":
    sys.exit(main(sys.argv[1:]))
