"""Full-domain ASCII output files in the reference's formats: fort.63
(elevation), fort.64 (depth-averaged velocity) and maxele.63 (peak elevation
and its time).

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
4330-4400).
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
           nspool: int, rtype: int) -> str:
    l1 = " " + rundes[:32].ljust(32) + "  " + runid[:24].ljust(24) + "  " + agrid[:24].ljust(24)
    l2 = (f" {nsets:10d} {np_:10d} {_fmt_e(dt_spool, 15, 7, 3)} {nspool:8d} {rtype:5d} "
          f"FileFmtVersion: {FILE_FMT_VERSION:10d}")
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
                 dtdp: float, nspool: int, rtype: int):
        self.path, self.rtype = path, rtype
        with open(path, "w") as f:
            f.write(header(rundes, runid, agrid, nsets, np_, dtdp * nspool, nspool, rtype))

    def write(self, timeloc: float, it: int, nodecode: np.ndarray, *cols: np.ndarray) -> None:
        dry = np.asarray(nodecode) == 0
        cols = tuple(np.where(dry, DRY_VALUE, np.asarray(c, np.float64)) for c in cols)
        with open(self.path, "a") as f:
            f.write(record_header(timeloc, it))
            f.write(node_lines(*cols))


def write_maxele(path: str, rundes: str, runid: str, agrid: str, timeloc: float, it: int,
                 etamax: np.ndarray, etamax_time: np.ndarray) -> None:
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
    """ETAMAX_Time from the step of each node's last ETAMAX rise:
    TimeLoc = IT*DTDP + STATIM*86400 (src/timestep.F:257), -99999.0 where
    ETAMAX never rose."""
    it = np.asarray(etamax_it, np.int64)
    t = it * dtdp + statim * 86400.0
    return np.where(it > 0, t, DRY_VALUE)
