"""Benchmark: simulated seconds per wall second on the 1M-node coastal mesh.

One step is one 2DDI timestep of the hot path (friction, met forcing, GWCE
LHS when wet/dry changed, GWCE RHS, elevation BCs, JCG to CONVCR = 1e-8,
wet/dry, momentum, peak surge) on the whole mesh, inputs resident in HBM.
Workload: the synthetic coastal mesh (Delaunay of seed-sampled points, ~1.0M
nodes / ~2.0M elements, 5 km -> 100 m), tanh shelf 4000 m -> -3 m, wet/dry
with H0 = 0.05 m, M2 + K1 on the open boundary, Holland vortex (Pc 940 hPa,
Rmax 40 km, B 1.3, 5 m/s shoreward), dt = 1 s; by default the
configuration's 6 simulated hours (21,600 steps: 30 warm-up, 21,570 timed).

    python bench.py [--gpus 1] [--steps 21570] [--warmup 30] [--nodes 1000000]

Prints one JSON line.  ``value`` is the plain path (the reference's JCG start,
results identical to the oracle's); ``warm_started`` reports the same run
again with the JCG started from the extrapolated increment 2*ETAS - ETAS_prev
(same stopping test) and its differences from the plain run (--no-warm skips
it).  The dominant kernel's roofline entry is computed from
the JCG launches timed with HIP events on the library's stream; the CPU
baseline is the oracle (serial C restatement of the reference, bit-exact to
it) timed on a bounded sample of the same run on this host.
"""
from __future__ import annotations

import argparse
import dataclasses
import json
import os
import sys
import time

import numpy as np

REPO = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, REPO)

HBM_PEAK_GBS = 8000.0


def log(msg: str) -> None:
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def build_case(nodes: int):
    from adcirc_amd import meshgen, prep

    t0 = time.time()
    case = meshgen.coastal(nodes)
    tables = prep.build_tables(case.mesh, case.run.nolifa, case.run.h0, case.run.anginn)
    log(f"mesh {tables['np']} nodes / {tables['ne']} elements in {time.time() - t0:.1f} s")
    return case, tables


def cpu_baseline(case, tables, eta2, nodecode, noff, budget_s: float):
    """The oracle (test infrastructure: the checker) timed on this host's core
    on the first steps of the same run."""
    sys.path.insert(0, os.path.join(REPO, "tests"))
    import oracle_lib
    from adcirc_amd import prep

    run = case.run
    npn = int(tables["np"])
    d = dict(tables)
    d["emo"] = np.asarray(case.emo, np.float64).ravel()
    d["efa"] = (np.asarray(case.efa, np.float64) * (3.141592653589793 / 180.0)).ravel()
    d["tau0var"] = prep.tau0_nodal(run.tau0, tables["dp"])
    d["evm"] = np.full(npn, run.eslm)
    d["fric"] = np.full(npn, run.cf)
    d["nodecode0"], d["noff0"], d["eta2_0"] = nodecode, noff, eta2
    d["prbckgrnd_mh2o"] = 100.0 * 1013.0 / (1000.0 * run.g)
    m = oracle_lib.OracleModel(d, run, oracle_lib.SEQUENTIAL)
    steps, t0 = 0, time.perf_counter()
    while True:
        m.step(steps + 1)
        steps += 1
        el = time.perf_counter() - t0
        if el >= budget_s or steps >= 200:
            break
    value = steps * run.dtdp / el
    return {"value": value, "unit": "sim-s/wall-s", "cores": 1, "kind": "port",
            "sample": f"first {steps} steps of the same {npn}-node coastal run from cold start, "
                      f"serial oracle (C restatement bit-exact to the reference build), "
                      f"{el:.1f} s wall"}


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    # the configuration's 6 simulated hours at dt = 1 s: 30 warm-up steps and
    # the remaining 21,570 timed
    ap.add_argument("--steps", type=int, default=21570)
    ap.add_argument("--warmup", type=int, default=30)
    ap.add_argument("--nodes", type=int, default=1_000_000)
    ap.add_argument("--cpu-budget", type=float, default=20.0)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-warm", action="store_true",
                    help="skip the second run with the extrapolated JCG start")
    ap.add_argument("--no-kernel-timing", action="store_true",
                    help="time the steps without HIP events around the JCG launches "
                         "(no roofline entry)")
    ap.add_argument("--dump", metavar="PATH",
                    help="write the case for the native driver adcirc_amd/adcgpu_run_case "
                         "(warmup + steps steps of forcing) and exit without using the GPU")
    a = ap.parse_args()
    if a.gpus != 1:
        print("bench.py measures one GPU: this build has no multi-GPU path "
              "(the single-mesh and ensemble sharding designs are described in DESIGN.md)",
              file=sys.stderr)
        return 2

    from adcirc_amd.model import Model, cold_wetdry_state, elevation_table, write_case

    case, tables = build_case(a.nodes)
    run = case.run
    eta2, nodecode, noff = cold_wetdry_state(tables, run)
    if a.dump:
        efa = np.asarray(case.efa, np.float64) * (3.141592653589793 / 180.0)
        table = elevation_table(run, case.emo, efa, 1, a.warmup + a.steps)
        write_case(a.dump, tables, run, eta2, nodecode, noff, table, run.holland)
        log(f"case written to {a.dump}")
        return 0
    def model(r):
        m = Model(tables, r, emo=case.emo, efa=case.efa * (3.141592653589793 / 180.0))
        m.cold_start(eta2, nodecode, noff)
        m.holland(0.0, r.holland)
        return m

    def timed(m, r, table, kernel_timing):
        log(f"warmup {a.warmup} steps")
        t0 = time.perf_counter()
        done = 0
        while done < a.warmup:
            k = min(10 if done == 0 else 50, a.warmup - done)
            m.run_resident(done + 1, k, holland=r.holland, table=table[done:done + k])
            done += k
            log(f"  warmup {done}/{a.warmup} steps, {time.perf_counter() - t0:.2f} s")
        m.synchronize()
        m.profile(kernel_timing, launches=a.steps + 16)
        log(f"timed {a.steps} steps")
        # the timed steps are issued in chunks of up to 1000 (each chunk ends
        # with a stream synchronisation inside the library) so that progress
        # is visible
        t0 = time.perf_counter()
        iters, done, rebuilds = 0, 0, 0
        while done < a.steps:
            k = min(1000, a.steps - done)
            s0 = a.warmup + done
            iters += m.run_resident(s0 + 1, k, holland=r.holland, table=table[s0:s0 + k])
            rebuilds += m.last_rebuilds
            done += k
            log(f"  {done}/{a.steps} steps, {time.perf_counter() - t0:.2f} s")
        m.synchronize()
        return time.perf_counter() - t0, iters, rebuilds

    g = model(run)
    log(f"JCG launch: {g.jcg_shape()} (blocks, resident-block limit, 2048-row chunks)")
    nsteps = a.warmup + a.steps
    log(f"elevation boundary table for {nsteps} steps")
    table = g.eta_bc_table(1, nsteps)
    wall, iters, rebuilds = timed(g, run, table, not a.no_kernel_timing)
    jcg_ms, launches = g.profile_read()
    shape = g.jcg_shape()
    g.profile(False)
    npn = int(tables["np"])
    Z = int(np.asarray(tables["nneigh"]).sum()) - npn
    # algorithmic bytes of the JCG launches: k*(12 Z + 48 NP) + (12 Z + 64 NP)
    jcg_bytes = iters * (12.0 * Z + 48.0 * npn) + launches * (12.0 * Z + 64.0 * npn)
    achieved = jcg_bytes / (jcg_ms * 1e-3) / 1e9 if jcg_ms > 0 else 0.0
    traffic = None
    tpath = os.path.join(REPO, "profiles", "k_jcg_coastal1m_traffic.json")
    if os.path.exists(tpath):
        rec = json.load(open(tpath))
        if (rec.get("nodes"), rec.get("warmup"), rec.get("steps")) == (npn, a.warmup, a.steps):
            traffic = rec["traffic_bytes_per_launch"] / 1e9
    wet = int(g.get("nodecode").sum())
    etamax = g.get("etamax")
    ms_step = wall * 1e3 / a.steps
    out = {
        "metric": "sim-s per wall-s, 1M-node mesh: plain vs warm-started JCG at same tol",
        "value": a.steps * run.dtdp / wall,
        "unit": "sim-s/wall-s",
        "n_gpus": 1,
        "steps": a.steps,
        "warmup": a.warmup,
        "ms_per_step": ms_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {
            "workload": f"coastal Delaunay mesh {npn} nodes / {int(tables['ne'])} elements, "
                        f"wet/dry H0=0.05, M2+K1, Holland vortex, dt=1 s, JCG tol 1e-8, "
                        f"plain extrapolated start, steps {a.warmup + 1}-{a.warmup + a.steps} "
                        f"({(a.warmup + a.steps) * run.dtdp / 3600.0:.2f} simulated hours)",
            "nodes": npn, "elements": int(tables["ne"]), "dt": run.dtdp, "convcr": run.convcr,
            "warm_start": "plain",
            "jcg_blocks": shape[0], "jcg_resident_cap": shape[1], "jcg_chunks": shape[2],
        },
        "timesteps_per_s": a.steps / wall,
        "jcg_iterations_per_step": iters / a.steps,
        "gwce_matrix_rebuilds": rebuilds,
        "wet_nodes_end": wet,
        "peak_surge_max_m": float(etamax.max()),
        "roofline": {
            "bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
            "frac": achieved / HBM_PEAK_GBS, "traffic": traffic, "traffic_unit": "GB/launch",
            "kernel": "k_jcg", "kernel_ms_avg": jcg_ms / max(launches, 1),
            "kernel_share_of_step": (jcg_ms / 1e3) / wall,
        },
    }
    if not a.no_warm:
        # the same run started from the extrapolated increment: same JCG
        # stopping test, results within the JCG tolerance of the plain run's
        log("warm-started JCG: initial guess 2*ETAS - ETAS_prev")
        wrun = dataclasses.replace(run, warm_start=1)
        w = model(wrun)
        wwall, witers, _ = timed(w, wrun, table, False)
        de = w.get("eta2") - g.get("eta2")
        em_g, em_w = etamax, w.get("etamax")
        both = (em_g > -99999.0) & (em_w > -99999.0)
        out["warm_started"] = {
            "method": "JCG initial guess 2*ETAS - ETAS_prev (second-order extrapolated "
                      "increment) instead of ETAS; not the learned start",
            "value": a.steps * run.dtdp / wwall,
            "ms_per_step": wwall * 1e3 / a.steps,
            "jcg_iterations_per_step": witers / a.steps,
            "speedup_vs_plain": wall / wwall,
            "vs_plain_max_abs_deta_m": float(np.abs(de).max()),
            "vs_plain_rms_deta_m": float(np.sqrt(np.mean(de * de))),
            "vs_plain_peak_surge_max_abs_diff_m": float(np.abs(em_w - em_g)[both].max()),
            "vs_plain_peak_surge_defined_mismatch": int((both != ((em_g > -99999.0) |
                                                                  (em_w > -99999.0))).sum()),
            "vs_plain_nodecode_mismatch_end": int((w.get("nodecode") != g.get("nodecode")).sum()),
            "vs_plain_noff_mismatch_end": int((w.get("noff") != g.get("noff")).sum()),
        }
        w.close()
    if not a.no_cpu:
        log("cpu baseline (oracle, 1 core)")
        out["cpu_baseline"] = cpu_baseline(case, tables, eta2, nodecode, noff, a.cpu_budget)
    g.close()
    print(json.dumps(out), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
