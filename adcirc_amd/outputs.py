"""ASCII output files in the reference's formats: fort.63 (elevation),
fort.64 (depth-averaged velocity), maxele.63 (peak elevation and its time),
and the station files fort.61 (elevation) and fort.62 (velocity).

Formats follow the reference's serial writers:

- header line 1: ``1X,A32,2X,A24,2X,A24`` = RUNDES, RUNID, AGRID
  (src/globalio.F:1000 in header63/header64/header_max);
- header line 2: ``1X,I10,1X,I10,1X,E15.7E3,1X,I8,1X,I5,1X,'FileFmtVersion: ',I10``
  = data sets, NP, DTDP*NSPOOL, NSPOOL, record type (1 scalar, 2 vector), and
  the file-format version 1.2.0 packed as (1 << 20) | (2 << 10)
  (src/globalio.F:1011-1051, src/global.F:1295-1303);
- record header: ``2X,1PE20.10E3,5X,I10`` = TimeLoc, IT
  (src/globalio.F:1101 in writeFullFormat, src/write_output.F:2120);
- node lines: ``2X,I8,2X,1PE20.10E3`` (one value) or ``2X,I8,2X,2(1PE20.10E3)``
  (two; src/globalio.F:802-946).  Elevation and velocity at dry nodes are
  written as -99999.0 (ConsiderWetDry, src/write_output.F:487, 606, 722).

maxele.63 holds two records: ETAMAX, then ETAMAX_Time, the time of each
node's maximum (-99999.0 where it never rose; src/write_output.F:1220-1237,
4330-4400); maxvel.63 the same for the peak speed UMAX.

Station files have the same layout with an ``I5`` NSPOOL in header line 2
(src/globalio.F:963-1003) and one line per station.  A station's value is
the linear interpolation d1*f1 + d2*f2 + d3*f3 over the element that holds
it, with the reference's interpolation factors
(src/read_input.F:6246-6277), and -99999.0 when a node of that element is
dry (src/write_output.F:2000-2060, StatPartWetFix off).
"""
from __future__ import annotations

from typing import Iterable, Optional

import numpy as np

FILE_FMT_VERSION = (1 << 20) | (2 << 10) | 0
DRY_VALUE = -99999.0


def _fmt_e_scaled(v: float, w: int, d: int, e: int) -> str:
    """Fortran ``1PEw.dEe``: one digit before the point, d after."""
    if v != v:
        return "NaN".rjust(w)
    s = f"{v:.{d}E}"
    mant, ex = s.split("E")
    k = int(ex)
    out = f"{mant}E{'-' if k < 0 else '+'}{abs(k):0{e}d}"
    return out.rjust(w)


def _fmt_e(v: float, w: int, d: int, e: int) -> str:
    """Fortran ``Ew.dEe`` (no scale factor): 0.ddd...E+eee."""
    if v != v:
        return "NaN".rjust(w)
    if v == 0.0:
        mant, k = "0." + "0" * d, 0
        sign = "-" if str(v).startswith("-") else ""
    else:
        s = f"{abs(v):.{d - 1}E}"
        m, ex = s.split("E")
        mant, k = "0." + m.replace(".", ""), int(ex) + 1
        sign = "-" if v < 0 else ""
    return f"{sign}{mant}E{'-' if k < 0 else '+'}{abs(k):0{e}d}".rjust(w)


def e20(v: float) -> str:
    return _fmt_e_scaled(float(v), 20, 10, 3)


def header(rundes: str, runid: str, agrid: str, nsets: int, np_: int, dt_spool: float,
           nspool: int, rtype: int, spool_width: int = 8) -> str:
    """Header lines 1-2; ``spool_width`` is 8 for the global files, 5 for
    the station files."""
    l1 = " " + rundes[:32].ljust(32) + "  " + runid[:24].ljust(24) + "  " + agrid[:24].ljust(24)
    l2 = (f" {nsets:10d} {np_:10d} {_fmt_e(dt_spool, 15, 7, 3)} {nspool:{spool_width}d} "
          f"{rtype:5d} FileFmtVersion: {FILE_FMT_VERSION:10d}")
    return l1 + "\n" + l2 + "\n"


def record_header(timeloc: float, it: int) -> str:
    return f"  {e20(timeloc)}     {it:10d}\n"


def node_lines(*cols: np.ndarray) -> str:
    n = cols[0].shape[0]
    out = []
    for i in range(n):
        out.append(f"  {i + 1:8d}  " + "".join(e20(c[i]) for c in cols))
    return "\n".join(out) + "\n"


class GlobalFile:
    """One fort.63/fort.64-style file: the header at open, a record per call
    to ``write``."""

    def __init__(self, path: str, rundes: str, runid: str, agrid: str, nsets: int, np_: int,
                 dtdp: float, nspool: int, rtype: int, spool_width: int = 8):
        self.path, self.rtype = path, rtype
        with open(path, "w") as f:
            f.write(header(rundes, runid, agrid, nsets, np_, dtdp * nspool, nspool, rtype,
                           spool_width))

    def write(self, timeloc: float, it: int, nodecode: np.ndarray, *cols: np.ndarray) -> None:
        dry = np.asarray(nodecode) == 0
        cols = tuple(np.where(dry, DRY_VALUE, np.asarray(c, np.float64)) for c in cols)
        with open(self.path, "a") as f:
            f.write(record_header(timeloc, it))
            f.write(node_lines(*cols))


def write_maxele(path: str, rundes: str, runid: str, agrid: str, timeloc: float, it: int,
                 etamax: np.ndarray, etamax_time: np.ndarray) -> None:
    """maxele.63 (ETAMAX, ETAMAX_Time) or, with UMAX and UMAX_Time, maxvel.63
    (src/write_output.F:1205-1268): the same two-record layout."""
    np_ = etamax.shape[0]
    with open(path, "w") as f:
        f.write(header(rundes, runid, agrid, 2, np_, 1.0, 1, 1))
        f.write(record_header(timeloc, it))
        f.write(node_lines(np.asarray(etamax, np.float64)))
        f.write(record_header(timeloc, it))
        f.write(node_lines(np.asarray(etamax_time, np.float64)))


def read_global(path: str):
    """Read a file written by this module: (header fields, [(time, it,
    values[np, k])])."""
    with open(path) as f:
        lines = f.read().splitlines()
    h = lines[1].split()
    nsets, np_, rtype = int(h[0]), int(h[1]), int(h[4])
    recs = []
    p = 2
    while p < len(lines):
        t, it = lines[p].split()
        vals = np.array([[float(x) for x in ln.split()[1:]] for ln in lines[p + 1:p + 1 + np_]])
        recs.append((float(t), int(it), vals))
        p += 1 + np_
    return {"nsets": nsets, "np": np_, "rtype": rtype}, recs


def etamax_time(etamax_it: np.ndarray, dtdp: float, statim: float) -> np.ndarray:
    """ETAMAX_Time (or UMAX_Time) from the step of each node's last rise:
    TimeLoc = IT*DTDP + STATIM*86400 (src/timestep.F:257), -99999.0 where
    ETAMAX never rose."""
    it = np.asarray(etamax_it, np.int64)
    t = it * dtdp + statim * 86400.0
    return np.where(it > 0, t, DRY_VALUE)


def station_factors(x: np.ndarray, y: np.ndarray, nm: np.ndarray, areas: np.ndarray,
                    xs: np.ndarray, ys: np.ndarray):
    """For each station (xs, ys): the 0-based element that holds it (-1 if
    none) and the reference's interpolation factors f1, f2, f3
    (ComputeInterpolatingFactors, src/read_input.F:6246-6277; AREAS is twice
    the element area; NM 1-based).  Of the elements that hold a station on a
    shared edge or vertex, the one with the largest smallest factor is taken."""
    nm0 = np.asarray(nm, np.int64).reshape(-1, 3) - 1
    X1, X2, X3 = (x[nm0[:, k]] for k in range(3))
    Y1, Y2, Y3 = (y[nm0[:, k]] for k in range(3))
    ns = len(xs)
    el = np.full(ns, -1, np.int64)
    f = np.zeros((ns, 3))
    for j in range(ns):
        X4, Y4 = float(xs[j]), float(ys[j])
        f1 = ((X4 - X3) * (Y2 - Y3) + (X2 - X3) * (Y3 - Y4)) / areas
        f2 = ((X4 - X1) * (Y3 - Y1) - (Y4 - Y1) * (X3 - X1)) / areas
        f3 = (-(X4 - X1) * (Y2 - Y1) + (Y4 - Y1) * (X2 - X1)) / areas
        worst = np.minimum(np.minimum(f1, f2), f3)
        e = int(np.argmax(worst))
        if worst[e] >= -1.0e-9:
            el[j] = e
            f[j] = (f1[e], f2[e], f3[e])
    return el, f


def station_values(vals: np.ndarray, nodecode: np.ndarray, nm: np.ndarray, el: np.ndarray,
                   f: np.ndarray) -> np.ndarray:
    """d1*f1 + d2*f2 + d3*f3 at each station, -99999.0 where a node of its
    element is dry or the station lies outside the mesh (NM 1-based)."""
    nm0 = np.asarray(nm, np.int64).reshape(-1, 3) - 1
    out = np.full(el.shape[0], DRY_VALUE)
    for j, e in enumerate(el):
        if e < 0:
            continue
        n1, n2, n3 = nm0[e]
        if nodecode[n1] * nodecode[n2] * nodecode[n3] == 0:
            continue
        out[j] = vals[n1] * f[j, 0] + vals[n2] * f[j, 1] + vals[n3] * f[j, 2]
    return out


class StationFile(GlobalFile):
    """fort.61 (rtype 1) or fort.62 (rtype 2): one line per station."""

    def __init__(self, path: str, rundes: str, runid: str, agrid: str, nsets: int, nsta: int,
                 dtdp: float, nspool: int, rtype: int):
        super().__init__(path, rundes, runid, agrid, nsets, nsta, dtdp, nspool, rtype,
                         spool_width=5)

    def write_values(self, timeloc: float, it: int, *cols: np.ndarray) -> None:
        with open(self.path, "a") as f:
            f.write(record_header(timeloc, it))
            f.write(node_lines(*cols))
