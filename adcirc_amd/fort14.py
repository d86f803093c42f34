"""fort.14 mesh files and the run description this package drives.

The grid file layout follows the reference reader (src/mesh.F:169-581,
docs/technical_reference/input_files/fort14.rst): a title line, ``NE NP``,
``NP`` node lines ``id x y depth``, ``NE`` element lines ``id 3 n1 n2 n3``,
then the elevation-specified (open) boundaries and the flow (land)
boundaries.  Node and element ids are 1-based as in the reference.
"""
from __future__ import annotations

import dataclasses
from typing import List, Tuple

import numpy as np


@dataclasses.dataclass
class Mesh:
    """A triangular mesh as read from / written to fort.14."""

    title: str
    x: np.ndarray            # [np] float64
    y: np.ndarray            # [np] float64
    dp: np.ndarray           # [np] float64, positive below datum
    nm: np.ndarray           # [ne, 3] int32, 1-based node ids
    open_bnd: List[np.ndarray]                  # 1-based node ids per segment
    land_bnd: List[Tuple[int, np.ndarray]]      # (ibtype, 1-based node ids)

    @property
    def np_(self) -> int:
        return int(self.x.shape[0])

    @property
    def ne(self) -> int:
        return int(self.nm.shape[0])


def write_fort14(path: str, m: Mesh) -> None:
    """Write ``m`` in the fort.14 layout the reference reads."""
    lines = [m.title[:80], f"{m.ne} {m.np_}"]
    ids = np.arange(1, m.np_ + 1)
    node_txt = np.char.mod("%d", ids)
    # 17 significant digits so that every float64 survives the round trip
    xs = np.char.mod("%.17g", m.x)
    ys = np.char.mod("%.17g", m.y)
    ds = np.char.mod("%.17g", m.dp)
    lines.extend((node_txt + " " + xs + " " + ys + " " + ds).tolist())
    eids = np.char.mod("%d", np.arange(1, m.ne + 1))
    nms = [np.char.mod("%d", m.nm[:, k]) for k in range(3)]
    lines.extend((eids + " 3 " + nms[0] + " " + nms[1] + " " + nms[2]).tolist())
    lines.append(f"{len(m.open_bnd)} ! NOPE")
    lines.append(f"{sum(len(s) for s in m.open_bnd)} ! NETA")
    for seg in m.open_bnd:
        lines.append(f"{len(seg)} 0 ! NVDLL IBTYPEE")
        lines.extend(str(int(n)) for n in seg)
    lines.append(f"{len(m.land_bnd)} ! NBOU")
    lines.append(f"{sum(len(s) for _, s in m.land_bnd)} ! NVEL")
    for ib, seg in m.land_bnd:
        lines.append(f"{len(seg)} {ib} ! NVELL IBTYPE")
        lines.extend(str(int(n)) for n in seg)
    with open(path, "w") as f:
        f.write("\n".join(lines) + "\n")


def read_fort14(path: str) -> Mesh:
    """Read the subset of fort.14 this package supports (IBTYPE 0/10/20 land,
    2/12/22 specified-flux, 52 river-radiation and 30 outward-radiating
    boundaries,
    elevation-specified open boundaries)."""
    with open(path) as f:
        tokens_lines = f.read().splitlines()
    title = tokens_lines[0]
    ne, np_ = (int(t) for t in tokens_lines[1].split()[:2])
    nodes = np.loadtxt(tokens_lines[2:2 + np_], dtype=np.float64, ndmin=2)
    x, y, dp = nodes[:, 1].copy(), nodes[:, 2].copy(), nodes[:, 3].copy()
    els = np.loadtxt(tokens_lines[2 + np_:2 + np_ + ne], dtype=np.int64, ndmin=2)
    nm = els[:, 2:5].astype(np.int32)
    p = 2 + np_ + ne
    nope = int(tokens_lines[p].split()[0]); p += 2
    open_bnd = []
    for _ in range(nope):
        n = int(tokens_lines[p].split()[0]); p += 1
        open_bnd.append(np.array([int(t.split()[0]) for t in tokens_lines[p:p + n]], np.int32))
        p += n
    nbou = int(tokens_lines[p].split()[0]); p += 2
    land = []
    for _ in range(nbou):
        hdr = tokens_lines[p].split(); p += 1
        n, ib = int(hdr[0]), int(hdr[1])
        land.append((ib, np.array([int(t.split()[0]) for t in tokens_lines[p:p + n]], np.int32)))
        p += n
    return Mesh(title, x, y, dp, nm, open_bnd, land)
