"""Run an ADCIRC case directory (fort.14 + fort.15) on the GPU and write its
global outputs, the way the reference's serial executable runs the same
inputs: cold start (src/cstart.F), NT = floor(RNDAY*86400/DTDP) steps of the
2DDI timestep, fort.63/fort.64 records at the steps the NOUTGE/NOUTGV lines
select, and maxele.63/maxvel.63 at the end.

The steps between two output times run as one resident launch sequence
(``adcgpu_run``): state stays on the device, and only the output steps copy
ETA2, UU2, VV2 and NODECODE back.

Elevation and velocity stations (NOUTE/NOUTV with NSTAE/NSTAV stations) are
written to fort.61 and fort.62 at their own output steps.

Not handled, and refused with an error: hot starts (IHOT != 0, NHSTAR != 0),
met-station output (NOUTM with NSTAM > 0), harmonic analysis (NFREQ > 0),
output formats other than full ASCII (NOUTGE/NOUTGV/NOUTE/NOUTV other than
0, 1, -1).  fort.22 is not read: with NWS = 8 the caller passes the
parametric vortex's parameters (the 13 values ``Model.holland`` takes).

    python -m adcirc_amd.driver CASE_DIR [OUT_DIR]
"""
from __future__ import annotations

import os
import sys
import time
from typing import Callable, Optional, Sequence

import numpy as np

from . import fort14, fort15, outputs, prep
from .model import Model, cold_wetdry_state

FLUX_CODES = (2, 12, 22, 52)


def flux_entries(mesh) -> int:
    """Flow-boundary entries of fort.14's specified-flux segments: the lines
    per constituent of fort.15's NFFR block.  The segments' types fix the
    count (the sharp-angle rule turns 2 into 12, still a flux code)."""
    return sum(int(v.shape[0]) for ib, v in mesh.land_bnd if ib in FLUX_CODES)


def _check(f: fort15.Fort15, holland) -> None:
    bad = []
    if f.ihot != 0:
        bad.append(f"IHOT = {f.ihot} (hot start)")
    if f.nhstar != 0:
        bad.append(f"NHSTAR = {f.nhstar} (hot-start output)")
    if f.met_sta.nout != 0 and f.stations_m.shape[0] > 0:
        bad.append(f"NOUTM = {f.met_sta.nout} with stations (met-station output)")
    if f.nfreq != 0:
        bad.append(f"NFREQ = {f.nfreq} (harmonic analysis)")
    for o, name in ((f.elev, "NOUTGE"), (f.vel, "NOUTGV"), (f.elev_sta, "NOUTE"),
                    (f.vel_sta, "NOUTV")):
        if o.nout not in (0, 1, -1):
            bad.append(f"{name} = {o.nout} (only full ASCII output, 1 or -1)")
    if f.nws != 0 and holland is None:
        bad.append(f"NWS = {f.nws}: fort.22 is not read; pass the vortex parameters")
    if bad:
        raise fort15.Fort15Error("unsupported in this driver: " + "; ".join(bad))


def run_directory(case_dir: str, out_dir: Optional[str] = None,
                  holland: Optional[Sequence[float]] = None, device: int = 0,
                  log: Callable[[str], None] = print) -> dict:
    """Run the case in ``case_dir``; outputs go to ``out_dir`` (default the
    case directory).  Returns a summary (steps, JCG iterations, wall time,
    files written)."""
    out_dir = out_dir or case_dir
    os.makedirs(out_dir, exist_ok=True)
    mesh = fort14.read_fort14(os.path.join(case_dir, "fort.14"))
    neta = sum(int(s.shape[0]) for s in mesh.open_bnd)
    bad = sorted({ib for ib, _ in mesh.land_bnd} - set(prep.FLOW_TYPES))
    if bad:
        raise fort15.Fort15Error(f"fort.14: flow boundary types {bad} are not supported "
                                 f"(supported: {', '.join(map(str, prep.FLOW_TYPES))})")
    f = fort15.read_fort15(os.path.join(case_dir, "fort.15"), neta, flux_entries(mesh))
    _check(f, holland)
    run = f.run
    if f.nws != 0:
        run.holland = tuple(float(v) for v in holland)
    tables = prep.build_tables(mesh, run.nolifa, run.h0, run.anginn)
    eta2, nodecode, noff = cold_wetdry_state(tables, run)
    g = Model(tables, run, emo=f.emo, efa=Model.radians(f.efa), qnam=f.qnam, qnph=f.qnph)
    try:
        g.cold_start(eta2, nodecode, noff)
        if run.holland is not None:
            g.holland(0.0, run.holland)
        nt, npn = run.nsteps, int(tables["np"])
        files = []
        f63 = f64 = None
        agrid = mesh.title
        if f.elev.nout:
            p = os.path.join(out_dir, "fort.63")
            f63 = outputs.GlobalFile(p, f.rundes, f.runid, agrid, f.elev.datasets(run), npn,
                                     run.dtdp, f.elev.nspool, 1)
            files.append(p)
        if f.vel.nout:
            p = os.path.join(out_dir, "fort.64")
            f64 = outputs.GlobalFile(p, f.rundes, f.runid, agrid, f.vel.datasets(run), npn,
                                     run.dtdp, f.vel.nspool, 2)
            files.append(p)
        sta = []  # (output control, file, interpolation, fields)
        for o, xy, name, rtype, fields in ((f.elev_sta, f.stations_e, "fort.61", 1, ("eta2",)),
                                           (f.vel_sta, f.stations_v, "fort.62", 2,
                                            ("uu2", "vv2"))):
            if o.nout and xy.shape[0] > 0:
                el, fac = outputs.station_factors(tables["x"], tables["y"], tables["nm"],
                                                  tables["areas"], xy[:, 0], xy[:, 1])
                p = os.path.join(out_dir, name)
                sf = outputs.StationFile(p, f.rundes, f.runid, agrid, o.datasets(run),
                                         xy.shape[0], run.dtdp, o.nspool, rtype)
                sta.append((o, sf, el, fac, fields))
                files.append(p)
        stops = sorted({it for it in range(1, nt + 1)
                        if f.elev.due(run, it) or f.vel.due(run, it)
                        or any(o.due(run, it) for o, *_ in sta)} | {nt})
        t0 = time.time()
        it, iters = 0, 0
        for stop in stops:
            if stop > it:
                iters += g.run_resident(it + 1, stop - it, holland=run.holland)
                it = stop
            timeloc = it * run.dtdp + run.statim * 86400.0
            if f63 is not None and f.elev.due(run, it):
                f63.write(timeloc, it, g.get("nodecode"), g.get("eta2"))
            if f64 is not None and f.vel.due(run, it):
                f64.write(timeloc, it, g.get("nodecode"), g.get("uu2"), g.get("vv2"))
            for o, sf, el, fac, fields in sta:
                if o.due(run, it):
                    nc = g.get("nodecode")
                    sf.write_values(timeloc, it, *(outputs.station_values(
                        g.get(k), nc, tables["nm"], el, fac) for k in fields))
        wall = time.time() - t0
        if f.elev.nout:
            p = os.path.join(out_dir, "maxele.63")
            timeloc = nt * run.dtdp + run.statim * 86400.0
            outputs.write_maxele(p, f.rundes, f.runid, agrid, timeloc, nt, g.get("etamax"),
                                 outputs.etamax_time(g.get("etamax_it"), run.dtdp, run.statim))
            files.append(p)
        if f.vel.nout:
            p = os.path.join(out_dir, "maxvel.63")
            timeloc = nt * run.dtdp + run.statim * 86400.0
            outputs.write_maxele(p, f.rundes, f.runid, agrid, timeloc, nt, g.get("umax"),
                                 outputs.etamax_time(g.get("umax_it"), run.dtdp, run.statim))
            files.append(p)
        log(f"{nt} steps in {wall:.2f} s ({nt * run.dtdp / max(wall, 1e-9):.1f} sim-s/wall-s), "
            f"{iters} JCG iterations; wrote {', '.join(os.path.basename(p) for p in files)}")
        return {"steps": nt, "jcg_iterations": iters, "wall_s": wall, "files": files}
    finally:
        g.close()


def main(argv=None) -> int:
    argv = sys.argv[1:] if argv is None else argv
    if not argv:
        print(__doc__)
        return 2
    run_directory(argv[0], argv[1] if len(argv) > 1 else None)
    return 0


if __name__ == "__main__":
    sys.exit(main())
