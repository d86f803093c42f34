"""Per-kernel summary of rocprofv3 kernel-trace databases of a resident run
(profiling aid; reads the rocpd SQLite files rocprofv3 writes).

Each time step of ``adcgpu_run`` starts with the friction pass
(``k_friction``, or ``k_friction_holland`` with device forcing), makes exactly
one ``k_jcg`` launch and ends with ``k_mom_node``, so a dispatch belongs to
the step of the last friction pass before it.  A database whose window opens
with the run is numbered forward from its first step, one whose window closes
with the run backward from the last.  Several databases recorded over
overlapping windows of the same run (``rocprofv3 -P``) are stitched: a step
is taken from the last database that holds it completely.

    python -m adcirc_amd.rocpd_summary --anchor first:1 --anchor last:21600 \
        --first-step 31 A.db B.db

With ``--counters KERNEL`` it instead reads a counter-collection database
(``rocprofv3 --pmc``) and prints, per counter, the mean value per dispatch of
KERNEL (skipping its first ``--skip`` dispatches) and of every other kernel
recorded, e.g. a copy kernel run for calibration.
"""
from __future__ import annotations

import argparse
import sqlite3
import sys
from collections import defaultdict


def _name(full: str) -> str:
    n = full.replace("(anonymous namespace)::", "").split("(")[0].replace("adcgpu::", "")
    return n[5:] if n.startswith("void ") else n


def _base(name: str) -> str:
    """The kernel's name without its template arguments."""
    return name.split("<")[0]


def _groups(path: str):
    rows = sqlite3.connect(path).execute(
        "select name, start, end from kernels order by start").fetchall()
    # a step opens at the friction pass; dispatches before the first one
    # belong to a step cut off by the window's start
    groups, cur = [], None
    for name, t0, t1 in rows:
        name = _name(name)
        if _base(name) in ("k_friction", "k_friction_holland"):
            cur = []
            groups.append(cur)
        if cur is not None:
            cur.append((name, t1 - t0))
    # a last step cut off by the window's end (no k_jcg, or not all of the
    # step's kernels) is dropped: the step's last kernel is k_mom_node
    while groups and _base(groups[-1][-1][0]) != "k_mom_node":
        groups.pop()
    return groups


def counters(path: str, kernel: str, skip: int) -> int:
    rows = sqlite3.connect(path).execute(
        "select kernel_name, dispatch_id, counter_name, value from counters_collection "
        "order by dispatch_id").fetchall()
    per = defaultdict(lambda: defaultdict(list))  # counter -> kernel -> values
    seen = defaultdict(set)
    for name, did, cname, val in rows:
        k = _name(name)
        if _base(k) == kernel:
            seen[cname].add(did)
            if len(seen[cname]) <= skip:
                continue
        per[cname][k].append(val)
    for cname, ks in per.items():
        for k, v in sorted(ks.items()):
            print(f"{cname} {k}: {len(v)} dispatches, mean {sum(v) / len(v):.1f} per dispatch")
    return 0


def main(argv=None) -> int:
    ap = argparse.ArgumentParser(description=__doc__.split("\n\n")[0])
    ap.add_argument("db", nargs="+", help="rocpd databases, in window order")
    ap.add_argument("--counters", metavar="KERNEL",
                    help="summarise a counter-collection database for KERNEL")
    ap.add_argument("--skip", type=int, default=0,
                    help="with --counters: leading dispatches of KERNEL to skip")
    ap.add_argument("--anchor", action="append", default=[],
                    help="per database, in order: first:N (its first complete step is N) "
                         "or last:N (its last complete step is N)")
    ap.add_argument("--first-step", type=int, default=1)
    ap.add_argument("--last-step", type=int, default=None)
    a = ap.parse_args(argv)
    if a.counters:
        return counters(a.db[0], a.counters, a.skip)
    if len(a.anchor) != len(a.db):
        ap.error("one --anchor per database")
    steps = {}
    for p, anc in zip(a.db, a.anchor):
        kind, n = anc.split(":")
        g = _groups(p)
        first = int(n) if kind == "first" else int(n) - len(g) + 1
        for i, grp in enumerate(g):
            steps[first + i] = grp  # later windows take precedence
    last = a.last_step if a.last_step is not None else max(steps)
    used = sorted(s for s in steps if a.first_step <= s <= last)
    missing = [s for s in range(a.first_step, last + 1) if s not in steps]
    agg = defaultdict(list)
    for s in used:
        for name, dur in steps[s]:
            agg[name].append(dur / 1000.0)
    tot = sum(sum(v) for v in agg.values())
    print(f"steps {used[0]}-{used[-1]}: {len(used)} steps, {len(missing)} missing")
    print()
    print("| kernel | calls | total ms | avg us | min us | max us | % |")
    print("|---|---|---|---|---|---|---|")
    for k, v in sorted(agg.items(), key=lambda kv: -sum(kv[1])):
        print(f"| {k} | {len(v)} | {sum(v) / 1000:.3f} | {sum(v) / len(v):.1f} | "
              f"{min(v):.1f} | {max(v):.1f} | {100 * sum(v) / tot:.1f} |")
    print()
    print(f"Kernel time {tot / 1000:.2f} ms over {len(used)} steps = "
          f"{tot / 1000 / len(used):.4f} ms per step.")
    j = [d for k, v in agg.items() if _base(k) == "k_jcg" for d in v]
    if j:
        print(f"k_jcg: {len(j)} launches, average {sum(j) / len(j):.1f} us.")
    return 0


if __name__ == "__main__":
    sys.exit(main())
