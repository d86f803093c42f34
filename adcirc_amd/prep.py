"""Mesh preprocessing: the tables the hot path consumes, built from a fort.14.

ADCIRC builds these at start-up (src/mesh.F) and, in the drop-in use, hands
them over as they are.  This module builds the same tables for the package's
own driver, element by element as the reference does, so that they are
identical to the reference's (the integer tables exactly; the float ones bit
for bit since they are formed by the same operations):

* element geometry FDXE/FDYE/AREAS, clockwise elements reordered
  (initializeMesh, src/mesh.F:1581-1677), depth floor for NOLIFA 0/1 (:1527);
* the neighbour table NEITAB/NNEIGH with neighbours sorted by angle
  (NEIGHB, src/mesh.F:3089-3318), MNEI = max elements per node + 2 (:1726-1731);
* CSR_POSITION (compute_csr_positions, src/mesh.F:5261-5282);
* elevation-boundary list NBD and flow-boundary tables NBV, LBCODEI, CSII, SIII,
  ME2GW for external boundaries of types 0/10/20
  (initializeBoundaries, src/mesh.F:1769-2922);
* the GWCE weighting TAU0 per node from fort.15's TAU0 (src/nodalattr.F:1858-1895).
"""
from __future__ import annotations

import math
from typing import Dict

import numpy as np

from .fort14 import Mesh

RAD2DEG = 180.0 / 3.141592653589793


def element_geometry(x: np.ndarray, y: np.ndarray, nm: np.ndarray):
    """FDXE, FDYE, AREAS with the reference's clockwise-element swap."""
    nm = nm.copy()
    for _ in range(2):
        X = x[nm - 1]
        Y = y[nm - 1]
        x1, x2, x3 = X[:, 0], X[:, 1], X[:, 2]
        y1, y2, y3 = Y[:, 0], Y[:, 1], Y[:, 2]
        x2mx1, x3mx2, x1mx3 = x2 - x1, x3 - x2, x1 - x3
        y2my1, y3my2, y1my3 = y2 - y1, y3 - y2, y1 - y3
        areas = x1mx3 * (-y3my2) + x3mx2 * y1my3
        neg = areas < 0.0
        if not neg.any():
            break
        nm[neg] = nm[neg][:, [0, 2, 1]]
    fdxe = np.column_stack([-y3my2, -y1my3, -y2my1])
    fdye = np.column_stack([x3mx2, x1mx3, x2mx1])
    return nm, fdxe, fdye, areas


def neighbour_table(x: np.ndarray, y: np.ndarray, nm0: np.ndarray):
    """NEITAB (0-based, -1 padded, row-major [np, mnei]) and NNEIGH."""
    npn = x.shape[0]
    cnt = np.bincount(nm0.ravel(), minlength=npn)
    mnei = int(cnt.max()) + 2
    a = np.concatenate([nm0[:, 0], nm0[:, 0], nm0[:, 1], nm0[:, 1], nm0[:, 2], nm0[:, 2]])
    b = np.concatenate([nm0[:, 1], nm0[:, 2], nm0[:, 0], nm0[:, 2], nm0[:, 0], nm0[:, 1]])
    key = np.unique(a.astype(np.int64) * npn + b)
    row = (key // npn).astype(np.int64)
    col = (key % npn).astype(np.int64)
    delx = x[col] - x[row]
    dely = y[col] - y[row]
    dist = np.sqrt(delx * delx + dely * dely)
    if np.any(dist == 0.0):
        raise ValueError("duplicate node coordinates")
    ang = np.empty_like(delx)
    nz = dely != 0.0
    ang[nz] = RAD2DEG * np.arccos(delx[nz] / dist[nz])
    pos = nz & (dely > 0.0)
    ang[pos] = 360.0 - ang[pos]
    z = ~nz
    ang[z & (delx > 0.0)] = 0.0
    ang[z & (delx < 0.0)] = 180.0
    order = np.lexsort((ang, row))
    row, col = row[order], col[order]
    nnb = np.bincount(row, minlength=npn)
    start = np.concatenate([[0], np.cumsum(nnb)[:-1]])
    slot = np.arange(row.shape[0]) - start[row] + 1
    neitab = np.full((npn, mnei), -1, dtype=np.int64)
    neitab[:, 0] = np.arange(npn)
    neitab[row, slot] = col
    nneigh = (nnb + 1).astype(np.int32)
    return neitab, nneigh, mnei


def csr_positions(neitab: np.ndarray, nneigh: np.ndarray, nm0: np.ndarray) -> np.ndarray:
    """CSR_POSITION(e, k, l) as 1-based column numbers, shape [ne, 3, 2]."""
    npn, mnei = neitab.shape
    r = np.repeat(np.arange(npn), mnei - 1)
    c = neitab[:, 1:].ravel()
    j = np.tile(np.arange(2, mnei + 1), npn)
    ok = c >= 0
    keys = r[ok].astype(np.int64) * npn + c[ok]
    jj = j[ok]
    srt = np.argsort(keys)
    keys, jj = keys[srt], jj[srt]

    def look(rows, cols):
        q = rows.astype(np.int64) * npn + cols
        ix = np.searchsorted(keys, q)
        if np.any(keys[np.minimum(ix, keys.shape[0] - 1)] != q):
            raise ValueError("element edge missing from the neighbour table")
        return jj[ix]

    n1, n2, n3 = nm0[:, 0], nm0[:, 1], nm0[:, 2]
    out = np.empty((nm0.shape[0], 3, 2), dtype=np.int32)
    out[:, 0, 0] = look(n1, n2); out[:, 0, 1] = look(n1, n3)
    out[:, 1, 0] = look(n2, n1); out[:, 1, 1] = look(n2, n3)
    out[:, 2, 0] = look(n3, n1); out[:, 2, 1] = look(n3, n2)
    return out


_ME_OVERLAP_CODES = (2, 12, 22, 3, 13, 23, 30, 32, 40, 41, 52, 64)


# land, specified-normal-flux and outward-radiation boundaries
FLOW_TYPES = (0, 10, 20, 2, 12, 22, 30, 52)
FLUX_TYPES = (2, 12, 22, 52)


def flow_boundaries(x: np.ndarray, y: np.ndarray, land, costset: float):
    """NBV, LBCODEI, CSII, SIII, ME2GW (1-based) and BNDLEN2O3 for external
    flow boundaries (src/mesh.F:2230-2420)."""
    nbv, lbc, csii, siii, me2gw, bndlen = [], [], [], [], [], []
    jgw = 0
    for ib, seg in land:
        if ib not in FLOW_TYPES:
            raise NotImplementedError(f"flow boundary type {ib}")
        v = [int(n) for n in seg]
        nvell = len(v)
        nbvv = {i + 1: v[i] for i in range(nvell)}
        nbvv[0] = nbvv[1]
        if nbvv[nvell] == nbvv[1]:
            nbvv[0] = nbvv[nvell - 1]
        nbvv[nvell + 1] = nbvv[nvell]
        theta1 = cross1 = costheta1 = None
        for i in range(1, nvell + 1):
            jgw += 1
            ni, nj, nk = nbvv[i], nbvv[i + 1], nbvv[i - 1]
            nbv.append(ni)
            xgi, xgj, xgk = x[ni - 1], x[nj - 1], x[nk - 1]
            ygi, ygj, ygk = y[ni - 1], y[nj - 1], y[nk - 1]
            # two thirds of the edge to the next node; zero for the last node
            # of a segment, whose next node is itself (src/mesh.F:2300-2303)
            delx, dely = xgj - xgi, ygj - ygi
            bndlen.append(2.0 * math.sqrt(delx * delx + dely * dely) / 3.0)
            calc = True
            if i == 1 and nbvv[1] == nbvv[0]:
                theta, cross, costheta = -9999999.0, 0.0, costset
                calc = False
            if i == nvell:
                theta, cross, costheta = theta1, cross1, costheta1
                calc = False
            if calc:
                vl1x, vl1y = xgk - xgi, ygk - ygi
                vl2x, vl2y = xgj - xgi, ygj - ygi
                dotvec = vl1x * vl2x + vl1y * vl2y
                vecnorm = math.sqrt(vl1x * vl1x + vl1y * vl1y) * math.sqrt(vl2x * vl2x + vl2y * vl2y)
                costheta = dotvec / vecnorm
                costheta = min(max(costheta, -1.0), 1.0)
                theta = RAD2DEG * math.acos(costheta)
                cross = -vl1x * vl2y + vl2x * vl1y
                if cross < 0:
                    theta = 360.0 - theta
            if i == 1:
                theta1, cross1, costheta1 = theta, cross, costheta
            code = ib
            if costheta > costset and cross > 0.0 and ib in (0, 1, 2, 3):
                code = {0: 10, 1: 11, 2: 12, 3: 13}[ib]
            lbc.append(code)
            dxl, dyl = xgk - xgj, ygk - ygj
            xl = math.sqrt(dxl * dxl + dyl * dyl)
            csii.append(1.0 * (ygj - ygk) / xl)
            siii.append((xgk - xgj) / xl)
            if i == 1:
                if jgw == 1:
                    me2gw.append(jgw)
                elif nbv[jgw - 1] != nbv[jgw - 2]:
                    me2gw.append(jgw)
                elif code in _ME_OVERLAP_CODES:
                    me2gw[-1] = jgw
            if 1 < i < nvell:
                me2gw.append(jgw)
            if i == nvell:
                if nbv[jgw - 1] != nbvv[1] and nbv[jgw - 1] != nbv[0]:
                    me2gw.append(jgw)
                if nbvv[i] == nbv[0] and code in _ME_OVERLAP_CODES:
                    me2gw[0] = jgw
    return (np.array(nbv, np.int32), np.array(lbc, np.int32), np.array(csii),
            np.array(siii), np.array(me2gw, np.int32), np.array(bndlen, np.float64))


REARTH = 6378206.4  # the reference's Rearth (src/constants.F90:79)


def lonlat_cpp(x: np.ndarray, y: np.ndarray, slam0_deg: float, sfea0_deg: float):
    """Longitude and latitude (radians) of Cartesian nodes by the inverse CPP
    projection about (SLAM0, SFEA0) degrees, as the reference computes SLAM
    and SFEA for ICS = 1 (src/mesh.F:1494-1512, INVCP src/mesh.F:3409-3420)."""
    slam0 = slam0_deg * (3.141592653589793 / 180.0)
    sfea0 = sfea0_deg * (3.141592653589793 / 180.0)
    x = np.asarray(x, np.float64)
    y = np.asarray(y, np.float64)
    slam = slam0 + x / (REARTH * math.cos(sfea0))
    sfea = y / REARTH
    return slam, sfea


def tau0_nodal(tau0: float, dp: np.ndarray) -> np.ndarray:
    """TAU0VAR from fort.15's TAU0 without a nodal-attribute file
    (src/nodalattr.F:1858-1895).  TAU0 >= 0 holds everywhere.  A negative
    value selects a depth scheme (Tau0NodalValue, src/nodalattr.F:2084-2102):
    -2 gives 0.005 at DP >= 200 m, 1/DP between 1 and 200 m and 1.0 above
    1 m; any other negative value gives 0.020 at DP <= 10 m and 0.005 deeper.
    The -2 scheme's 0.005 is a default-real literal, so it carries that
    literal's single-precision rounding; the other constants are exact or
    double."""
    dp = np.asarray(dp, np.float64)
    if tau0 >= 0.0:
        return np.full(dp.shape[0], float(tau0))
    if tau0 == -2.0:
        deep = float(np.float32(0.005))
        with np.errstate(divide="ignore"):
            mid = 1.0 / dp
        return np.where(dp >= 200.0, deep, np.where(dp >= 1.0, mid, 1.0))
    return np.where(dp <= 10.0, 0.020, 0.005)


def build_tables(mesh: Mesh, nolifa: int, h0: float, anginn: float = 110.0) -> Dict[str, np.ndarray]:
    """All tables of the reference's set-up, in the layout the C ABI takes."""
    x, y = mesh.x.astype(np.float64), mesh.y.astype(np.float64)
    dp = mesh.dp.astype(np.float64).copy()
    if nolifa in (0, 1):
        dp = np.where(dp < h0, h0, dp)
    nm, fdxe, fdye, areas = element_geometry(x, y, mesh.nm.astype(np.int64))
    nm0 = nm - 1
    neitab0, nneigh, mnei = neighbour_table(x, y, nm0)
    csr = csr_positions(neitab0, nneigh, nm0)
    nbd = np.concatenate(mesh.open_bnd).astype(np.int32) if mesh.open_bnd else np.zeros(0, np.int32)
    costset = math.cos(anginn * (3.141592653589793 / 180.0))
    nbv, lbc, csii, siii, me2gw, bndlen = flow_boundaries(x, y, mesh.land_bnd, costset)
    npn, ne = x.shape[0], nm.shape[0]
    return dict(
        np=npn, ne=ne, mnei=mnei, neta=int(nbd.shape[0]), nvel=int(nbv.shape[0]),
        nvelme=int(me2gw.shape[0]), nm=nm.astype(np.int32), x=x, y=y, dp=dp,
        fdxe=fdxe, fdye=fdye, areas=areas, nneigh=nneigh,
        # reference layout: NEITAB(NP,MNEI) column-major, 1-based, 0 padded
        neitab=(neitab0 + 1).T.astype(np.int32).ravel(),
        csr=csr, nbd=nbd, nbv=nbv, lbcodei=lbc, csii=csii, siii=siii, me2gw=me2gw,
        bndlen2o3=bndlen,
    )
