"""Does the order of the inner-product sums alone change the coastal storm's
wet/dry flags?  (A CPU study run by hand; pytest does not collect this file.)

The reference forms every inner product (the JCG dots, EP) as one sequential
sum; the GPU forms them with a fixed blocked tree.  The two differ in the last
bits.  This script runs the C oracle twice on the same case, once in each
order (everything else identical, bit for bit), and reports per simulated hour
how far the two trajectories have drifted apart: NODECODE / NOFF mismatches,
max |d eta|, max |d u|, the peak-surge difference, and, at the first step
where a flag differs, the margins of the mismatching nodes to the wet/dry
thresholds (DP + ETA2 against H0 and 1.2 H0; src/wetdry.F:234-285, 368-475).

    python tests/order_sensitivity.py [NODES] [STEPS] [OUT.json]

The record of the coastal 20k-node six-hour case is committed as
tests/golden/coastal20k_order.json.
"""
from __future__ import annotations

import hashlib
import json
import os
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(HERE))
sys.path.insert(0, HERE)

import oracle_lib  # noqa: E402
from gpu_helpers import case_setup  # noqa: E402
from adcirc_amd import meshgen  # noqa: E402

EVERY = 3600


def margins(d, run, a, b, nodes):
    """Margins of DP + ETA2 to H0 and 1.2 H0 in both runs at the given nodes."""
    dp = np.asarray(d["dp"], np.float64)
    ea, eb = a.array("eta2"), b.array("eta2")
    out = []
    for i in nodes[:16]:
        ha, hb = dp[i] + ea[i], dp[i] + eb[i]
        out.append({"node": int(i), "htot_seq": float(ha), "htot_blocked": float(hb),
                    "seq_minus_h0": float(ha - run.h0), "blocked_minus_h0": float(hb - run.h0),
                    "seq_minus_hoff": float(ha - 1.2 * run.h0),
                    "blocked_minus_hoff": float(hb - 1.2 * run.h0)})
    return out


def main(argv) -> int:
    nodes = int(argv[0]) if len(argv) > 0 else 20000
    nsteps = int(argv[1]) if len(argv) > 1 else 21600
    out_path = argv[2] if len(argv) > 2 else os.path.join(HERE, "golden",
                                                          "coastal20k_order.json")
    case = meshgen.coastal(nodes)
    run = case.run
    d = case_setup(case)
    seq = oracle_lib.OracleModel(d, run, oracle_lib.SEQUENTIAL)
    blk = oracle_lib.OracleModel(d, run, oracle_lib.BLOCKED)
    rec = {"nodes": int(d["np"]), "elements": int(d["ne"]), "steps": nsteps,
           "what": "C oracle, sequential (reference) vs blocked (GPU) inner-product order, "
                   "same case and forcing", "hours": [], "first_flag_mismatch": None,
           "first_state_difference_step": None}
    # the sequential run against the reference build's per-step state hashes
    # (tests/golden/coastal20k.npz, written by make_golden.py), when present
    ref = None
    fx = os.path.join(HERE, "golden", "coastal20k.npz")
    if nodes == 20000 and os.path.exists(fx):
        ref = np.load(fx)["step_hash"]
        rec["seq_vs_reference_hash_mismatches"] = 0
        rec["seq_vs_reference_steps_checked"] = 0
    t0 = time.time()
    prev_de = 0.0
    flag_steps = 0
    for it in range(1, nsteps + 1):
        seq.step(it)
        blk.step(it)
        if ref is not None and it <= ref.shape[0]:
            rec["seq_vs_reference_steps_checked"] += 1
            if seq.state_hash() != int(ref[it - 1]):
                rec["seq_vs_reference_hash_mismatches"] += 1
        nc_a, nc_b = seq.iarray("nodecode"), blk.iarray("nodecode")
        nf_a, nf_b = seq.iarray("noff"), blk.iarray("noff")
        if not (np.array_equal(nc_a, nc_b) and np.array_equal(nf_a, nf_b)):
            flag_steps += 1
        if rec["first_state_difference_step"] is None:
            if not np.array_equal(seq.array("eta2"), blk.array("eta2")):
                rec["first_state_difference_step"] = it
        if rec["first_flag_mismatch"] is None and (not np.array_equal(nc_a, nc_b)
                                                    or not np.array_equal(nf_a, nf_b)):
            bad = np.nonzero(nc_a != nc_b)[0]
            rec["first_flag_mismatch"] = {
                "step": it, "nodecode": int(bad.size),
                "noff": int((nf_a != nf_b).sum()),
                "max_abs_deta_step_before": prev_de,
                "margins": margins(d, run, seq, blk, bad)}
            print(f"first flag mismatch at step {it}: {bad.size} nodes", flush=True)
        if it % 60 == 0 or it == nsteps:
            prev_de = float(np.abs(seq.array("eta2") - blk.array("eta2")).max())
        if it % EVERY == 0 or it == nsteps:
            ea, eb = seq.array("eta2"), blk.array("eta2")
            ua, ub = seq.array("uu2"), blk.array("uu2")
            va, vb = seq.array("vv2"), blk.array("vv2")
            ma, mb = seq.array("etamax"), blk.array("etamax")
            both = (ma > -99999.0) & (mb > -99999.0)
            h = {"step": it,
                 # the blocked run's state (what the GPU must equal bit for bit)
                 "blocked_sha256": {k: hashlib.sha256(np.ascontiguousarray(v).tobytes())
                                    .hexdigest() for k, v in (("eta2", eb), ("uu2", ub),
                                                              ("vv2", vb), ("etamax", mb),
                                                              ("nodecode", nc_b),
                                                              ("noff", nf_b))},
                 "steps_with_flag_mismatch": flag_steps,
                 "nodecode_mismatch": int((nc_a != nc_b).sum()),
                 "noff_mismatch": int((nf_a != nf_b).sum()),
                 "max_abs_deta_m": float(np.abs(ea - eb).max()),
                 "max_abs_du_ms": float(max(np.abs(ua - ub).max(), np.abs(va - vb).max())),
                 "peak_surge_max_abs_diff_m": float(np.abs(ma - mb)[both].max())
                 if both.any() else 0.0,
                 "numitr_seq": int(seq.numitr()), "numitr_blocked": int(blk.numitr()),
                 "wet_nodes": int(nc_a.sum())}
            rec["hours"].append(h)
            print(json.dumps(h), f"({time.time() - t0:.0f} s)", flush=True)
            with open(out_path, "w") as f:
                json.dump(rec, f, indent=1)
    return 0


if __name__ == "__main__":
    sys.exit(main(sys.argv[1:]))
