"""Benchmark: simulated seconds per wall second on the 1M-node coastal mesh,
the plain FP64 path against the learned JCG start and mixed precision.

One step is one 2DDI timestep of the hot path (friction, Holland met forcing,
GWCE matrix when wet/dry changed, GWCE right-hand side, elevation BCs, JCG to
CONVCR = 1e-8, wet/dry, momentum, peak surge) on the whole mesh, inputs
resident in HBM.  Workload: the synthetic coastal mesh (Delaunay of
seed-sampled points, ~1.0M nodes / ~2.0M elements, 5 km -> 100 m), tanh shelf
4000 m -> -3 m, wet/dry with H0 = 0.05 m, M2 + K1 on the open boundary,
Holland vortex (Pc 940 hPa, Rmax 40 km, B 1.3, 5 m/s shoreward), dt = 1 s.

    python bench.py [--gpus 1] [--steps 600] [--warmup 20] [--start-step 10800]

Procedure:
  1. Untimed fast-forward on the plain FP64 path from cold start to
     --start-step (default 10,800: hour 3 of the storm, the surge rising at
     the coast, so that nearly every step rebuilds the GWCE matrix and wets or
     dries nodes), collecting training samples for the learned start (node
     snapshots; the last 2^20 are kept).
  2. The learned start's network is trained on the GPU on those samples
     (--train-epochs); the training time is reported, not timed.
  3. Every leg starts a fresh context from the state snapshot at the start
     step, runs --warmup steps untimed and times the next --steps steps:
       plain          FP64, the reference's JCG start (``value``);
       extrapolated   FP64, JCG started from 2*ETAS - ETAS_prev;
       learned        FP64, JCG started from the better of the reference's
                      start and the network's (smaller initial residual);
       mixed          FP32 element passes and JCG matrix, FP64 solve vectors;
       learned_mixed  both: the combination the speed target names;
       extrapolated_mixed  the extrapolated start in mixed precision.
     Every solve runs to the same stopping test.  Each leg reports its time,
     JCG iterations (and those saved against plain), the last solve's
     stopping-test value against its bound, and its distance to the plain
     leg's state at the end of the window.
  4. C1 (basin, 1 simulated day, launch-latency bound) on its own, and C4:
     storm ensembles on meshes of 50k, 250k and the 1M nodes (eight storms
     each) stepped as one batch, against the same members one at a time.
  5. CPU baseline: the reference build (oracle/_ref/adchp_ref, the
     reference's own Fortran compiled from its sources) timed on one host
     core on the first steps of the same 1M-node run; the oracle's C
     restatement when that build is absent.

Met forcing is generated on the device (k_holland); it differs from the
oracle's host forcing by at most 1e-13 relative
(tests/test_gpu_parity.py::test_holland_forcing_vs_oracle), so the timed
trajectory is the oracle's only up to that forcing difference; bit-exactness
to the oracle is shown with host forcing in the tests.  The dominant kernel's
roofline entry is computed from the plain leg's JCG launches timed with HIP
events on the library's stream.
"""
from __future__ import annotations

import argparse
import dataclasses
import json
import math
import os
import subprocess
import sys
import tempfile
import time

import numpy as np

REPO = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, REPO)

HBM_PEAK_GBS = 8000.0
ML_CAP = 1 << 20  # training samples the library keeps


def log(msg: str) -> None:
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def build_case(nodes: int):
    from adcirc_amd import meshgen, prep

    t0 = time.time()
    case = meshgen.coastal(nodes)
    tables = prep.build_tables(case.mesh, case.run.nolifa, case.run.h0, case.run.anginn)
    log(f"mesh {tables['np']} nodes / {tables['ne']} elements in {time.time() - t0:.1f} s")
    return case, tables


def cpu_model() -> str:
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def cpu_baseline(case, tables, eta2, nodecode, noff, steps: int, budget_s: float):
    """The reference build (test infrastructure: the checker) timed on one
    host core on the first steps of the same run; the oracle's C restatement
    when the reference build is absent."""
    sys.path.insert(0, os.path.join(REPO, "tests"))
    npn = int(tables["np"])
    run = case.run
    import refdump

    if os.path.exists(refdump.REF_EXE):
        with tempfile.TemporaryDirectory() as wd:
            t0 = time.perf_counter()
            refdump.write_inputs(case, wd, steps, hash_steps=False)
            log(f"  reference inputs written in {time.perf_counter() - t0:.1f} s")
            r = subprocess.run([refdump.REF_EXE, "run.nml"], cwd=wd, stdout=subprocess.PIPE,
                               stderr=subprocess.STDOUT, text=True, timeout=600)
        loop = None
        for line in r.stdout.splitlines():
            if "step loop seconds=" in line:
                loop = float(line.split("=")[1])
        if r.returncode == 0 and loop:
            return {"value": steps * run.dtdp / loop, "unit": "sim-s/wall-s", "cores": 1,
                    "kind": "reference", "cpu_model": cpu_model(),
                    "sample": f"first {steps} steps of the same {npn}-node coastal run from "
                              f"cold start, the reference's own Fortran (flang -O3, serial) "
                              f"in oracle/_ref/adchp_ref; step loop {loop:.1f} s wall "
                              f"(mesh set-up excluded)"}
        log(f"  reference build failed (rc {r.returncode}); timing the C restatement")
    import oracle_lib
    from adcirc_amd import prep

    d = dict(tables)
    d["emo"] = np.asarray(case.emo, np.float64).ravel()
    d["efa"] = (np.asarray(case.efa, np.float64) * (math.pi / 180.0)).ravel()
    d["tau0var"] = prep.tau0_nodal(run.tau0, tables["dp"])
    d["evm"] = np.full(npn, run.eslm)
    d["fric"] = np.full(npn, run.cf)
    d["nodecode0"], d["noff0"], d["eta2_0"] = nodecode, noff, eta2
    d["prbckgrnd_mh2o"] = 100.0 * 1013.0 / (1000.0 * run.g)
    m = oracle_lib.OracleModel(d, run, oracle_lib.SEQUENTIAL)
    n, t0 = 0, time.perf_counter()
    while True:
        m.step(n + 1)
        n += 1
        el = time.perf_counter() - t0
        if el >= budget_s or n >= steps:
            break
    return {"value": n * run.dtdp / el, "unit": "sim-s/wall-s", "cores": 1, "kind": "port",
            "cpu_model": cpu_model(),
            "sample": f"first {n} steps of the same {npn}-node coastal run from cold start, "
                      f"serial oracle (C restatement bit-exact to the reference build), "
                      f"{el:.1f} s wall"}


def jcg_traffic_fit():
    """HBM bytes per JCG launch as a + b * iterations, fitted from the
    per-dispatch FETCH_SIZE + WRITE_SIZE counters of a profiled run."""
    path = os.path.join(REPO, "profiles", "k_jcg_coastal1m_traffic.json")
    if not os.path.exists(path):
        return None
    rec = json.load(open(path))
    if "bytes_per_launch" not in rec or "bytes_per_iteration" not in rec:
        return None
    return rec


def basin_day(steps_warm: int = 100):
    """C1: the 5,151-node basin for one simulated day (43,200 steps of 2 s),
    resident run; launch latency bounds it."""
    from adcirc_amd import meshgen, prep
    from adcirc_amd.model import Model, cold_wetdry_state

    case = meshgen.basin()
    run = case.run
    t = prep.build_tables(case.mesh, run.nolifa, run.h0, run.anginn)
    eta2, nodecode, noff = cold_wetdry_state(t, run)
    m = Model(t, run, emo=case.emo, efa=case.efa * (math.pi / 180.0))
    m.cold_start(eta2, nodecode, noff)
    n = run.nsteps
    table = m.eta_bc_table(1, n)
    m.run_resident(1, steps_warm, table=table[:steps_warm])
    m.synchronize()
    t0 = time.perf_counter()
    iters = m.run_resident(steps_warm + 1, n - steps_warm, table=table[steps_warm:])
    m.synchronize()
    wall = time.perf_counter() - t0
    k = n - steps_warm
    m.close()
    return {"workload": f"basin {int(t['np'])} nodes / {int(t['ne'])} elements, M2, dt "
                        f"{run.dtdp:g} s, no wet/dry, steps {steps_warm + 1}-{n} of one "
                        f"simulated day, resident run (one stream launch sequence per step)",
            "sim_s_per_wall_s": k * run.dtdp / wall, "timesteps_per_s": k / wall,
            "ms_per_step": wall * 1e3 / k, "jcg_iterations_per_step": iters / k}


def batch_leg(case1m, tables1m, steps: int, warmup: int, per_mesh: int, sizes):
    """C4: storm ensembles on coastal meshes of the given sizes (the 1M-node
    mesh of the main workload last), ``per_mesh`` storms each, stepped from
    cold start as one batch (adcgpu_batch: one JCG launch for all members
    whose blocks fit together), then the same members one by one for as many
    steps (adcgpu_run per member)."""
    from adcirc_amd import meshgen, prep
    from adcirc_amd.model import Batch, Model, cold_wetdry_state

    meshes = []
    for n in sizes:
        if n == int(tables1m["np"]) or n >= 1_000_000:
            meshes.append((case1m, tables1m))
        else:
            c = meshgen.coastal(n)
            meshes.append((c, prep.build_tables(c.mesh, c.run.nolifa, c.run.h0, c.run.anginn)))
    models, hols, tabs, desc = [], [], [], []
    for c, t in meshes:
        eta2, nc, nf = cold_wetdry_state(t, c.run)
        efa = np.asarray(c.efa, np.float64) * (math.pi / 180.0)
        tab = None
        for h in meshgen.storm_ensemble(c, per_mesh):
            m = Model(t, c.run, emo=c.emo, efa=efa)
            m.cold_start(eta2, nc, nf)
            m.holland(0.0, h)
            if tab is None:
                tab = m.eta_bc_table(1, warmup + 2 * steps)
            models.append(m)
            hols.append(h)
            tabs.append(tab)
        desc.append(f"{per_mesh} x {int(t['np'])} nodes")
    log(f"batch of {len(models)} members ({', '.join(desc)}): {warmup} warm-up steps")
    b = Batch(models)
    b.run(1, warmup, holland=hols, tables=[t[:warmup] for t in tabs])
    log(f"  batched: {steps} timed steps")
    t0 = time.perf_counter()
    its = b.run(warmup + 1, steps, holland=hols, tables=[t[warmup:warmup + steps] for t in tabs])
    wall_b = time.perf_counter() - t0
    b.close()
    log(f"  one by one: {steps} timed steps per member")
    t0 = time.perf_counter()
    its1 = 0
    s0 = warmup + steps
    for m, h, t in zip(models, hols, tabs):
        its1 += m.run_resident(s0 + 1, steps, holland=h, table=t[s0:s0 + steps])
    wall_1 = time.perf_counter() - t0
    for m in models:
        m.close()
    ms = len(models) * steps
    return {"workload": f"{', '.join(desc)} coastal meshes, storms drawn by "
                        f"meshgen.storm_ensemble (Pc 920-980 hPa, Rmax 20-60 km, landfall "
                        f"+-50 km, heading +-20 deg), from cold start: {warmup} warm-up steps, "
                        f"steps {warmup + 1}-{warmup + steps} batched, then "
                        f"{s0 + 1}-{s0 + steps} one member at a time",
            "members": len(models), "member_steps_per_s": ms / wall_b,
            "ms_per_step": wall_b * 1e3 / steps,
            "one_by_one_member_steps_per_s": ms / wall_1,
            "speedup_vs_one_by_one": wall_1 / wall_b,
            "jcg_iterations_per_member_step": sum(its) / ms,
            "one_by_one_jcg_iterations_per_member_step": its1 / ms}


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=600)
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--start-step", type=int, default=10800,
                    help="untimed fast-forward from cold start to this step")
    ap.add_argument("--nodes", type=int, default=1_000_000)
    ap.add_argument("--train-epochs", type=int, default=8)
    ap.add_argument("--legs",
                    default="plain,extrapolated,learned,mixed,learned_mixed,extrapolated_mixed")
    ap.add_argument("--cpu-steps", type=int, default=10)
    ap.add_argument("--cpu-budget", type=float, default=20.0)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-basin", action="store_true")
    ap.add_argument("--no-batch", action="store_true", help="skip the C4 batch leg")
    ap.add_argument("--batch-members", type=int, default=8, help="storms per mesh in C4")
    ap.add_argument("--batch-sizes", default="50000,250000,1000000",
                    help="node counts of the C4 meshes (1000000: the main workload's mesh)")
    ap.add_argument("--no-kernel-timing", action="store_true",
                    help="no HIP events around the JCG launches (no roofline entry)")
    ap.add_argument("--dump", metavar="PATH",
                    help="write the case for the native driver adcirc_amd/adcgpu_run_case "
                         "(warmup + steps steps of forcing from cold start) and exit "
                         "without using the GPU")
    ap.add_argument("--dump-legs", metavar="DIR",
                    help="after the fast-forward and the training, write DIR/case (forcing "
                         "for start-step + warmup + steps steps), DIR/state.bin (the snapshot "
                         "at the start step) and DIR/params.f32 (the trained network) for "
                         "adcgpu_run_case --state DIR/state.bin --it0 START --params ..., "
                         "and exit without timing the legs")
    a = ap.parse_args()
    if a.gpus != 1:
        print("bench.py measures one GPU: this build has no multi-GPU path "
              "(the single-mesh and ensemble sharding designs are described in DESIGN.md)",
              file=sys.stderr)
        return 2

    from adcirc_amd import _lib
    from adcirc_amd.model import Model, cold_wetdry_state, elevation_table, write_case

    case, tables = build_case(a.nodes)
    run = case.run
    eta2, nodecode, noff = cold_wetdry_state(tables, run)
    efa = np.asarray(case.efa, np.float64) * (math.pi / 180.0)
    if a.dump:
        table = elevation_table(run, case.emo, efa, 1, a.warmup + a.steps)
        write_case(a.dump, tables, run, eta2, nodecode, noff, table, run.holland)
        log(f"case written to {a.dump}")
        return 0
    legs = [s for s in a.legs.split(",") if s]
    if "plain" not in legs:
        legs.insert(0, "plain")
    S, W, K = a.start_step, a.warmup, a.steps
    npn, ne = int(tables["np"]), int(tables["ne"])

    def model(warm=0, precision=_lib.FP64):
        m = Model(tables, dataclasses.replace(run, warm_start=warm), emo=case.emo, efa=efa,
                  precision=precision)
        return m

    # ---- 1. fast-forward with sample collection
    f = model()
    f.cold_start(eta2, nodecode, noff)
    f.holland(0.0, run.holland)
    log(f"JCG launch: {f.jcg_shape()} (blocks, resident-block limit, 2048-row chunks)")
    log(f"elevation boundary table for {S + W + K} steps")
    table = f.eta_bc_table(1, S + W + K)
    per_step = max(1, math.ceil(ML_CAP / max(S, 1))) if S > 0 else 0
    f.ml_collect(per_step)
    t0 = time.perf_counter()
    done, ff_iters = 0, 0
    while done < S:
        k = min(2000, S - done)
        ff_iters += f.run_resident(done + 1, k, holland=run.holland, table=table[done:done + k])
        done += k
        log(f"  fast-forward {done}/{S} steps, {time.perf_counter() - t0:.1f} s")
    f.ml_collect(0)
    f.synchronize()
    ff_wall = time.perf_counter() - t0
    samples = f.ml_samples()
    # ---- 2. training
    t0 = time.perf_counter()
    losses = []
    for e in range(a.train_epochs if samples > 0 else 0):
        lr = 1e-3 if e < max(a.train_epochs - 2, 1) else 3e-4
        losses.append(f.ml_train(1, lr=lr, seed=17 + e))
        log(f"  training epoch {e + 1}/{a.train_epochs}: loss {losses[-1]:.4e}")
    f.synchronize()
    train_wall = time.perf_counter() - t0
    params = f.ml_params()
    snap = f.save_state()
    f.close()
    if a.dump_legs:
        os.makedirs(a.dump_legs, exist_ok=True)
        write_case(os.path.join(a.dump_legs, "case"), tables, run, eta2, nodecode, noff,
                   elevation_table(run, case.emo, efa, 1, S + W + K), run.holland)
        snap.tofile(os.path.join(a.dump_legs, "state.bin"))
        np.asarray(params, np.float32).tofile(os.path.join(a.dump_legs, "params.f32"))
        log(f"legs written to {a.dump_legs} (snapshot after step {S})")
        return 0

    # ---- 3. legs
    spec = {"plain": (0, _lib.FP64), "extrapolated": (1, _lib.FP64),
            "learned": (2, _lib.FP64), "mixed": (0, _lib.MIXED),
            "learned_mixed": (2, _lib.MIXED), "extrapolated_mixed": (1, _lib.MIXED)}
    res, final = {}, {}
    for name in legs:
        warm, prec = spec[name]
        m = model(warm, prec)
        m.load_state(snap)
        if warm == 2:
            m.ml_params(params)
        m.run_resident(S + 1, W, holland=run.holland, table=table[S:S + W])
        m.synchronize()
        kt = name == "plain" and not a.no_kernel_timing
        m.profile(kt, launches=K + 16)
        log(f"leg {name}: timed steps {S + W + 1}-{S + W + K}")
        t0 = time.perf_counter()
        iters, used, rebuilds, done = 0, 0, 0, 0
        while done < K:
            k = min(1000, K - done)
            s0 = S + W + done
            iters += m.run_resident(s0 + 1, k, holland=run.holland, table=table[s0:s0 + k])
            st = m.run_stats()
            rebuilds += st[1]
            used += st[4]
            done += k
        m.synchronize()
        wall = time.perf_counter() - t0
        log(f"  {K} steps in {wall:.3f} s ({wall * 1e3 / K:.3f} ms/step), "
            f"{iters / K:.2f} JCG iterations/step")
        r = {"ms_per_step": wall * 1e3 / K, "value": K * run.dtdp / wall,
             "jcg_iterations_per_step": iters / K, "gwce_matrix_rebuilds": rebuilds}
        ls = m.last_solve()
        r["last_solve_stptst"], r["last_solve_zeta"] = ls[0], ls[1]
        if warm == 2:
            r["learned_start_used"] = used
            r["learned_start_used_fraction"] = used / K
            r["last_solve_initial_residual_sq"] = {"reference_start": ls[2],
                                                   "learned_start": ls[3]}
        if kt:
            jcg_ms, launches = m.profile_read()
            r["_jcg"] = (jcg_ms, launches, iters)
            m.profile(False)
        final[name] = {k: m.get(k) for k in ("eta2", "nodecode", "noff", "etamax")}
        res[name] = r
        m.close()

    p = res["plain"]
    base = final["plain"]
    for name in legs:
        if name == "plain":
            continue
        r, fin = res[name], final[name]
        r["speedup_vs_plain"] = p["ms_per_step"] / r["ms_per_step"]
        r["jcg_iterations_saved_per_step"] = (p["jcg_iterations_per_step"]
                                              - r["jcg_iterations_per_step"])
        de = fin["eta2"] - base["eta2"]
        em, em0 = fin["etamax"], base["etamax"]
        both = (em > -99998.0) & (em0 > -99998.0)
        r["vs_plain"] = {
            "max_abs_deta_m": float(np.abs(de).max()),
            "rms_deta_m": float(np.sqrt(np.mean(de * de))),
            "peak_surge_diff_m": float(abs(em.max() - em0.max())),
            "etamax_max_abs_diff_m": float(np.abs(em - em0)[both].max()) if both.any() else 0.0,
            "nodecode_mismatch": int((fin["nodecode"] != base["nodecode"]).sum()),
            "noff_mismatch": int((fin["noff"] != base["noff"]).sum()),
        }

    # roofline of the dominant kernel (k_jcg) from the plain leg
    roof = None
    if "_jcg" in p:
        jcg_ms, launches, iters = p.pop("_jcg")
        Z = int(np.asarray(tables["nneigh"]).sum()) - npn
        # algorithmic bytes: k*(12 Z + 48 NP) per iteration + (12 Z + 64 NP) per launch
        jcg_bytes = iters * (12.0 * Z + 48.0 * npn) + launches * (12.0 * Z + 64.0 * npn)
        achieved = jcg_bytes / (jcg_ms * 1e-3) / 1e9 if jcg_ms > 0 else 0.0
        fit = jcg_traffic_fit()
        traffic = None
        if fit and launches:
            traffic = (fit["bytes_per_launch"] + fit["bytes_per_iteration"] * iters / launches) / 1e9
        roof = {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                "frac": achieved / HBM_PEAK_GBS, "traffic": traffic,
                "traffic_unit": "GB/launch", "kernel": "k_jcg",
                "algorithmic_gb_per_launch": jcg_bytes / max(launches, 1) / 1e9,
                "kernel_ms_avg": jcg_ms / max(launches, 1),
                "kernel_share_of_step": (jcg_ms / 1e3) / (p["ms_per_step"] * K / 1e3)}

    out = {
        "metric": "sim-s per wall-s, 1M-node mesh: plain vs warm-started JCG at same tol",
        "value": p["value"],
        "unit": "sim-s/wall-s",
        "n_gpus": 1,
        "steps": K,
        "warmup": W,
        "ms_per_step": p["ms_per_step"],
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {
            "workload": f"coastal Delaunay mesh {npn} nodes / {ne} elements, wet/dry H0=0.05, "
                        f"M2+K1, Holland vortex (device forcing), dt=1 s, JCG tol 1e-8, plain "
                        f"FP64 path with the reference's JCG start; fast-forward to step {S} "
                        f"({S * run.dtdp / 3600.0:.2f} h, mid-storm), {W} warm-up steps, timed "
                        f"steps {S + W + 1}-{S + W + K}",
            "nodes": npn, "elements": ne, "dt": run.dtdp, "convcr": run.convcr,
            "start_step": S,
        },
        "timesteps_per_s": K / (p["ms_per_step"] * K / 1e3),
        "jcg_iterations_per_step": p["jcg_iterations_per_step"],
        "gwce_matrix_rebuilds": p["gwce_matrix_rebuilds"],
        "last_solve_stptst": p["last_solve_stptst"],
        "last_solve_zeta": p["last_solve_zeta"],
        "wet_nodes_end": int(base["nodecode"].sum()),
        "peak_surge_max_m": float(base["etamax"].max()),
        "learned_start_training": {
            "samples": samples, "samples_per_step": per_step,
            "sample_steps": f"1-{S} (ring of the last {ML_CAP})",
            "epochs": a.train_epochs, "loss_by_epoch": losses, "train_s": train_wall,
            "fast_forward_s": ff_wall, "fast_forward_jcg_iterations_per_step": ff_iters / max(S, 1),
        },
    }
    for name in legs:
        if name != "plain":
            out[name] = res[name]
    if "learned_mixed" in res:
        out["target_learned_mixed_speedup_vs_plain"] = res["learned_mixed"]["speedup_vs_plain"]
    if roof:
        out["roofline"] = roof
    if not a.no_basin:
        log("C1: basin, one simulated day")
        out["basin_day"] = basin_day()
    if not a.no_batch:
        log("C4: storm ensembles as one batch")
        sizes = [int(v) for v in a.batch_sizes.split(",") if v]
        out["batch"] = batch_leg(case, tables, K, W, a.batch_members, sizes)
    if not a.no_cpu:
        log("cpu baseline")
        out["cpu_baseline"] = cpu_baseline(case, tables, eta2, nodecode, noff, a.cpu_steps,
                                           a.cpu_budget)
        out["gpu_vs_cpu"] = p["value"] / out["cpu_baseline"]["value"]
    print(json.dumps(out), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
