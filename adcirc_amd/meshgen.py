"""Seeded synthetic meshes and run descriptions for the test and benchmark cases.

Every generator is deterministic for a given seed (numpy PCG64), so a mesh is
regenerated identically on any host instead of being shipped as a file.

Cases
-----
``basin``      rectangular basin, alternating-diagonal quads, jittered interior
               nodes, linear depth + noise, M2 on the offshore side (the
               CPU-runnable reference case: 101 x 51 nodes, 100 x 50 km).
``beach``      small sloping beach that wets and dries (wet/dry test case).
``coastal``    graded Delaunay coastal mesh with a tanh shelf from 4000 m to
               -3 m overland, 5 km offshore -> 100 m nearshore resolution
               (1.0M nodes at full size), M2 + K1 and a Holland storm.
"""
from __future__ import annotations

import dataclasses
import math
from typing import Dict, List, Optional, Tuple

import numpy as np

from .fort14 import Mesh

M2_OMEGA = 2.0 * math.pi / (12.42 * 3600.0)
K1_OMEGA = 2.0 * math.pi / (23.93 * 3600.0)


@dataclasses.dataclass
class RunConfig:
    """The fort.15 subset the hot path consumes (names follow the reference's
    fort.15 variables, docs/technical_reference/input_files/fort15.rst)."""

    dtdp: float = 2.0
    nsteps: int = 43200
    statim: float = 0.0
    reftim: float = 0.0
    nramp: int = 1
    drampelev: float = 0.5        # days
    nolifa: int = 1               # 0 linear, 1 finite amplitude, 2 + wet/dry
    nolica: int = 1
    nolicat: int = 1
    nolibf: int = 1               # 0 linear, 1 quadratic, 2 hybrid
    cf: float = 0.0025            # CF (NOLIBF 1), CFmin (2), TAU (0)
    # hybrid friction (NOLIBF = 2): break depth (m) and the two exponents
    hbreak: float = 1.0
    ftheta: float = 10.0
    fgamma: float = 1.0 / 3.0
    tau0: float = 0.005
    eslm: float = 5.0
    cori: float = 1.0e-4
    h0: float = 0.05
    velmin: float = 0.01
    a00: float = 0.35
    b00: float = 0.30
    c00: float = 0.35
    itmax: int = 100
    convcr: float = 1.0e-8
    ilump: int = 0                # GWCE LHS: 0 consistent (JCG), 1 lumped (explicit)
    warm_start: int = 0           # JCG start: 0 previous increment, 1 extrapolated increment
    errorelev: float = 1000.0     # ErrorElev: |ETA2| above it at a wet node stops the run
    anginn: float = 110.0
    g: float = 9.80665
    # periodic elevation forcing on the open boundary
    amig: Tuple[float, ...] = (M2_OMEGA,)
    ff: Tuple[float, ...] = (1.0,)
    face: Tuple[float, ...] = (0.0,)      # degrees
    # met forcing: None, or the 13 Holland parameters of adco_holland
    holland: Optional[Tuple[float, ...]] = None
    # tidal potential (NTIP = 1): per constituent (TPK, AMIGT rad/s, ETRF, FFT,
    # FACET degrees), fort.15's NTIF block
    tipot: Tuple[Tuple[float, float, float, float, float], ...] = ()
    # centre of the CPP projection (degrees) that maps the Cartesian mesh to
    # longitude/latitude (ICS = 1)
    slam0: float = 0.0
    sfea0: float = 0.0
    # periodic specified normal flux (NFFR > 0) on IBTYPE 2/12/22 boundaries:
    # per constituent (FAMIG rad/s, FFF, FFACE degrees); the amplitudes and
    # phases per boundary node are Case.qnam / Case.qnph
    fluxf: Tuple[Tuple[float, float, float], ...] = ()

    @property
    def nbfr(self) -> int:
        return len(self.amig)

    @property
    def nws(self) -> int:
        return 0 if self.holland is None else 8

    @property
    def ntip(self) -> int:
        return 1 if self.tipot else 0

    @property
    def nffr(self) -> int:
        return len(self.fluxf)


@dataclasses.dataclass
class Case:
    mesh: Mesh
    run: RunConfig
    emo: np.ndarray   # [neta, nbfr] amplitudes (m)
    efa: np.ndarray   # [neta, nbfr] phases (degrees)
    # normal flux per node of the IBTYPE 2/12/22 segments in fort.14 order
    # (fort.15's QNAM, QNPH per constituent): [entries, nffr], m^2/s and degrees
    qnam: Optional[np.ndarray] = None
    qnph: Optional[np.ndarray] = None


def _cell_split(i: int, j: int) -> bool:
    return (i + j) % 2 == 0


def _structured(nx: int, ny: int, lx: float, ly: float, rng: np.random.Generator,
                jitter: float):
    """Nodes on an nx-by-ny grid (x-major numbering), interior jittered by
    +-jitter of the spacing; alternating-diagonal triangles, counterclockwise."""
    dx, dy = lx / (nx - 1), ly / (ny - 1)
    ii, jj = np.meshgrid(np.arange(nx), np.arange(ny), indexing="ij")
    x = ii * dx
    y = jj * dy
    interior = (ii > 0) & (ii < nx - 1) & (jj > 0) & (jj < ny - 1)
    jx = rng.uniform(-jitter, jitter, size=x.shape) * dx
    jy = rng.uniform(-jitter, jitter, size=y.shape) * dy
    x = np.where(interior, x + jx, x).ravel()
    y = np.where(interior, y + jy, y).ravel()

    def nid(i, j):
        return i * ny + j + 1

    tris = []
    for i in range(nx - 1):
        for j in range(ny - 1):
            a, b, c, d = nid(i, j), nid(i + 1, j), nid(i + 1, j + 1), nid(i, j + 1)
            if _cell_split(i, j):
                tris.append((a, b, c))
                tris.append((a, c, d))
            else:
                tris.append((a, b, d))
                tris.append((b, c, d))
    nm = np.array(tris, dtype=np.int32)
    # open boundary: east edge, south -> north (domain on the left)
    east = np.array([nid(nx - 1, j) for j in range(ny)], np.int32)
    # land: north edge east -> west, west edge north -> south, south edge west -> east
    north = [nid(i, ny - 1) for i in range(nx - 1, -1, -1)]
    west = [nid(0, j) for j in range(ny - 2, -1, -1)]
    south = [nid(i, 0) for i in range(1, nx)]
    land = np.array(north + west + south, np.int32)
    return x, y, nm, east, land


def basin(seed: int = 1, nx: int = 101, ny: int = 51, lx: float = 100.0e3,
          ly: float = 50.0e3, amp: float = 0.5) -> Case:
    """The 100 x 50 km basin: depth 5 m (x=0) -> 50 m (x=lx) + N(0, 1 m)."""
    rng = np.random.default_rng(seed)
    x, y, nm, east, land = _structured(nx, ny, lx, ly, rng, 0.2)
    dp = 5.0 + 45.0 * (x / lx) + rng.normal(0.0, 1.0, size=x.shape)
    mesh = Mesh(f"basin {nx}x{ny} seed {seed}", x, y, dp, nm, [east], [(0, land)])
    # no Coriolis: with f = 1e-4 and advection on, this jittered basin grows a
    # slow instability after ~16 h in the reference as well
    run = RunConfig(dtdp=2.0, nsteps=43200, nolifa=1, cori=0.0)
    neta = east.shape[0]
    emo = np.full((neta, 1), amp)
    efa = np.zeros((neta, 1))
    return Case(mesh, run, emo, efa)


# fort.15 tidal potential lines (TPK, AMIGT rad/s, ETRF, FFT, FACET degrees) of
# the four main constituents; FFT and FACET as for an arbitrary start date
TIPOT_M2S2K1O1 = (
    (0.242334, 0.000140518902509, 0.693, 0.9722, 212.5),
    (0.112841, 0.000145444104333, 0.693, 1.0, 0.0),
    (0.141565, 0.0000729211583579, 0.736, 1.0877, 351.8),
    (0.100514, 0.0000675977441508, 0.695, 1.1513, 119.2),
)


def basin_tide(seed: int = 1) -> Case:
    """The basin with the tidal potential of M2, S2, K1 and O1 besides its
    boundary tide; the Cartesian mesh is placed at 75 W, 35 N for the
    potential's longitude and latitude (CPP projection, ICS = 1)."""
    case = basin(seed)
    case.run.tipot = TIPOT_M2S2K1O1
    case.run.slam0, case.run.sfea0 = -75.0, 35.0
    return case


def basin_river(seed: int = 1, west_ibtype: int = 2) -> Case:
    """The basin with two rivers besides its boundary tide: 8 km of the
    shallow west edge take a specified normal flux as an essential condition
    (IBTYPE 2), 10 km of the north edge as a natural one (IBTYPE 22).  Each
    carries a steady discharge and a tidal (M2) part, positive into the
    domain; the land segments meet the rivers at shared end nodes."""
    nx, ny, lx, ly = 101, 51, 100.0e3, 50.0e3
    case = basin(seed, nx, ny, lx, ly)
    ny_ = ny

    def nid(i, j):
        return i * ny_ + j + 1

    # land runs north edge east -> west, west edge north -> south, south edge
    # west -> east (as _structured builds it); cut it at the rivers
    north = [nid(i, ny - 1) for i in range(nx - 1, -1, -1)]
    west = [nid(0, j) for j in range(ny - 2, -1, -1)]
    south = [nid(i, 0) for i in range(1, nx)]
    ring = north + west + south
    r_n0, r_n1 = ring.index(nid(40, ny - 1)), ring.index(nid(30, ny - 1))
    r_w0, r_w1 = ring.index(nid(0, 29)), ring.index(nid(0, 21))
    # west_ibtype = 52: the west river radiates too (river radiation)
    segs = [(0, ring[:r_n0 + 1]), (22, ring[r_n0:r_n1 + 1]), (0, ring[r_n1:r_w0 + 1]),
            (west_ibtype, ring[r_w0:r_w1 + 1]), (0, ring[r_w1:])]
    case.mesh.land_bnd = [(ib, np.array(v, np.int32)) for ib, v in segs]
    case.mesh.name = f"basin river {nx}x{ny} seed {seed}"
    case.run.fluxf = ((0.0, 1.0, 0.0), (M2_OMEGA, 1.0, 30.0))
    n22, n2 = r_n1 - r_n0 + 1, r_w1 - r_w0 + 1
    # m^2/s per unit width: steady 0.5 (north) and 1.2 (west), M2 part 0.2/0.4
    q = np.concatenate([np.column_stack([np.full(n22, 0.5), np.full(n22, 0.2)]),
                        np.column_stack([np.full(n2, 1.2), np.full(n2, 0.4)])])
    ph = np.concatenate([np.column_stack([np.zeros(n22), np.linspace(0.0, 20.0, n22)]),
                         np.column_stack([np.zeros(n2), np.linspace(10.0, 0.0, n2)])])
    case.qnam, case.qnph = q, ph
    return case


def basin_radiation(seed: int = 1) -> Case:
    """The basin with 30 km of its north edge open as an outward-radiating
    boundary (IBTYPE 30): the normal flux there follows the elevation
    change, c = sqrt(g H)."""
    nx, ny = 101, 51
    case = basin(seed, nx, ny)

    def nid(i, j):
        return i * ny + j + 1

    north = [nid(i, ny - 1) for i in range(nx - 1, -1, -1)]
    west = [nid(0, j) for j in range(ny - 2, -1, -1)]
    south = [nid(i, 0) for i in range(1, nx)]
    ring = north + west + south
    r0, r1 = ring.index(nid(70, ny - 1)), ring.index(nid(40, ny - 1))
    segs = [(0, ring[:r0 + 1]), (30, ring[r0:r1 + 1]), (0, ring[r1:])]
    case.mesh.land_bnd = [(ib, np.array(v, np.int32)) for ib, v in segs]
    case.mesh.name = f"basin radiation {nx}x{ny} seed {seed}"
    return case


def beach(seed: int = 3, nx: int = 41, ny: int = 21, lx: float = 8.0e3,
          ly: float = 4.0e3, storm: bool = True) -> Case:
    """A small sloping beach (6 m offshore at x=lx to -1.5 m at x=0) that wets
    and dries under an M2 tide and, optionally, a small Holland vortex."""
    rng = np.random.default_rng(seed)
    x, y, nm, east, land = _structured(nx, ny, lx, ly, rng, 0.15)
    dp = -1.5 + 7.5 * (x / lx) + rng.normal(0.0, 0.05, size=x.shape)
    mesh = Mesh(f"beach {nx}x{ny} seed {seed}", x, y, dp, nm, [east], [(0, land)])
    hol = None
    if storm:
        # Pc, Pn, Rmax, B, x0, y0, vx, vy, f, ramp days, rho_air, wfac, inflow
        hol = (990.0, 1013.0, 3.0e3, 1.3, 14.0e3, 2.0e3, -1.5, 0.0, 1.0e-4, 0.05,
               1.15, 0.8, 20.0)
    run = RunConfig(dtdp=1.0, nsteps=3600, nolifa=2, drampelev=0.02, h0=0.05,
                    velmin=0.01, holland=hol, cori=1.0e-4)
    neta = east.shape[0]
    emo = np.full((neta, 1), 0.6)
    efa = np.zeros((neta, 1))
    return Case(mesh, run, emo, efa)


# ---------------------------------------------------------------- coastal mesh
COAST_X = 152.0e3     # where the tanh shelf crosses 0 m
COAST_LX = 200.0e3


def coastal_depth(x: np.ndarray) -> np.ndarray:
    """tanh shelf: ~4000 m at x=0, 0 m at x~152 km, -3 m asymptotically inland."""
    return -3.0 + 2001.5 * (1.0 - np.tanh((x - 80.0e3) / 20.0e3))


def coastal_spacing(x: np.ndarray, band: float) -> np.ndarray:
    """100 m within +-band of the coastline, growing by 20% of the distance
    offshore up to 5 km and by 10% inland up to 600 m."""
    d = np.abs(x - COAST_X) - band
    d = np.maximum(d, 0.0)
    off = np.minimum(100.0 + 0.2 * d, 5000.0)
    on = np.minimum(100.0 + 0.1 * d, 600.0)
    return np.where(x < COAST_X, off, on)


def _columns(lx: float, band: float) -> np.ndarray:
    xs = [0.0]
    while True:
        h = float(coastal_spacing(np.array([xs[-1]]), band)[0])
        nxt = xs[-1] + h * math.sqrt(3.0) / 2.0
        if nxt >= lx - 0.3 * h:
            break
        xs.append(nxt)
    xs.append(lx)
    return np.array(xs)


def coastal_node_count(ly: float, band: float, lx: float = COAST_LX) -> int:
    xs = _columns(lx, band)
    h = coastal_spacing(xs, band)
    return int(np.sum(np.ceil(ly / h) + 1))


def coastal(target_nodes: int = 1_000_000, seed: int = 7, band: float = 10.0e3,
            lx: float = COAST_LX, storm: bool = True) -> Case:
    """Graded Delaunay coastal mesh with about ``target_nodes`` nodes."""
    from scipy.spatial import Delaunay
    from scipy.sparse import coo_matrix
    from scipy.sparse.csgraph import reverse_cuthill_mckee

    # choose the alongshore length so the node count hits the target
    per_m = coastal_node_count(100.0e3, band, lx) / 100.0e3
    ly = max(target_nodes / per_m, 2.0e3)
    rng = np.random.default_rng(seed)
    xs = _columns(lx, band)
    hs = coastal_spacing(xs, band)
    px, py, edge = [], [], []
    for k, (xc, h) in enumerate(zip(xs, hs)):
        n = int(math.ceil(ly / h)) + 1
        yy = np.linspace(0.0, ly, n)
        if 0 < k < len(xs) - 1:
            sh = (0.5 * (ly / (n - 1))) if (k % 2) else 0.0
            yy = yy + sh
            yy = yy[yy < ly - 0.3 * (ly / (n - 1))]
            yy = np.concatenate([[0.0], yy[yy > 0.3 * (ly / (n - 1))], [ly]])
        nn = yy.shape[0]
        jx = np.zeros(nn)
        jy = np.zeros(nn)
        if 0 < k < len(xs) - 1:
            jx[1:-1] = rng.uniform(-0.2, 0.2, nn - 2) * h
            jy[1:-1] = rng.uniform(-0.2, 0.2, nn - 2) * (ly / max(nn - 1, 1))
        px.append(np.full(nn, xc) + jx)
        py.append(yy + jy)
    x = np.concatenate(px)
    y = np.concatenate(py)
    tri = Delaunay(np.column_stack([x, y]), qhull_options="Qbb Qc Qz Q12 Qt")
    nm0 = tri.simplices.astype(np.int64)
    # drop degenerate triangles on the hull
    a2 = ((x[nm0[:, 1]] - x[nm0[:, 0]]) * (y[nm0[:, 2]] - y[nm0[:, 0]])
          - (x[nm0[:, 2]] - x[nm0[:, 0]]) * (y[nm0[:, 1]] - y[nm0[:, 0]]))
    keep = np.abs(a2) > 1e-6
    nm0 = nm0[keep]
    a2 = a2[keep]
    flip = a2 < 0
    nm0[flip] = nm0[flip][:, [0, 2, 1]]
    npn = x.shape[0]
    # reverse Cuthill-McKee node order for locality
    r = np.concatenate([nm0[:, 0], nm0[:, 1], nm0[:, 2]])
    c = np.concatenate([nm0[:, 1], nm0[:, 2], nm0[:, 0]])
    adj = coo_matrix((np.ones(r.shape[0], np.int8), (r, c)), shape=(npn, npn)).tocsr()
    adj = adj + adj.T
    perm = np.asarray(reverse_cuthill_mckee(adj.tocsr(), symmetric_mode=True))
    inv = np.empty(npn, np.int64)
    inv[perm] = np.arange(npn)
    x, y = x[perm], y[perm]
    nm0 = inv[nm0]
    # elements sorted by their smallest node (locality of the element sweep)
    order = np.lexsort((nm0.max(axis=1), nm0.min(axis=1)))
    nm0 = nm0[order]
    nm = (nm0 + 1).astype(np.int32)
    dp = coastal_depth(x)
    tol = 1e-6
    west = np.where(np.abs(x) < tol)[0]
    east = np.where(np.abs(x - lx) < tol)[0]
    south = np.where(np.abs(y) < tol)[0]
    north = np.where(np.abs(y - ly) < tol)[0]
    # open boundary on the offshore edge, north -> south (domain on the left)
    ob = west[np.argsort(-y[west])] + 1
    # land: south edge west -> east, east edge south -> north, north edge east -> west
    s = south[np.argsort(x[south])]
    e = east[np.argsort(y[east])]
    n_ = north[np.argsort(-x[north])]
    land = np.concatenate([s, e[1:], n_[1:]]) + 1
    mesh = Mesh(f"coastal {npn} nodes seed {seed}", x, y, dp, nm,
                [ob.astype(np.int32)], [(0, land.astype(np.int32))])
    hol = None
    if storm:
        # landfall near y = ly/2 after ~5.5 h at 5 m/s shoreward
        hol = (940.0, 1013.0, 40.0e3, 1.3, COAST_X - 100.0e3, ly / 2.0, 5.0, 0.0,
               1.0e-4, 0.1, 1.15, 0.8, 20.0)
    # TAU0 = -1: 0.020 on the shelf edge and overland (DP <= 10 m), 0.005
    # deeper; a uniform 0.005 lets the storm-driven wetting front go unstable
    run = RunConfig(dtdp=1.0, nsteps=21600, nolifa=2, drampelev=0.25, h0=0.05, tau0=-1.0,
                    velmin=0.01, holland=hol, amig=(M2_OMEGA, K1_OMEGA),
                    ff=(1.0, 1.0), face=(0.0, 0.0))
    neta = ob.shape[0]
    yb = y[ob - 1] / ly
    emo = np.column_stack([0.5 + 0.1 * yb, 0.2 + 0.05 * yb])
    efa = np.column_stack([10.0 * yb, 5.0 * yb])
    return Case(mesh, run, emo, efa)


def storm_ensemble(case: Case, n: int, seed: int = 11) -> List[Tuple[float, ...]]:
    """n Holland vortices around the case's storm, drawn from a seeded
    generator: central pressure U(920, 980) hPa, radius of maximum winds
    U(20, 60) km, landfall point shifted along the coast by U(-50, 50) km and
    heading turned by U(-20, 20) degrees at the case's forward speed."""
    base = list(case.run.holland)
    rng = np.random.default_rng(seed)
    speed = math.hypot(base[6], base[7])
    heading0 = math.atan2(base[7], base[6])
    out = []
    for _ in range(n):
        h = list(base)
        h[0] = float(rng.uniform(920.0, 980.0))
        h[2] = float(rng.uniform(20.0e3, 60.0e3))
        shift = float(rng.uniform(-50.0e3, 50.0e3))
        turn = math.radians(float(rng.uniform(-20.0, 20.0)))
        # the track still crosses the coast (x = COAST_X) at y0 + shift
        vx, vy = speed * math.cos(heading0 + turn), speed * math.sin(heading0 + turn)
        t_land = (COAST_X - base[4]) / vx
        h[5] = base[5] + shift - vy * t_land
        h[6], h[7] = vx, vy
        out.append(tuple(h))
    return out


CASES = {"basin": basin, "beach": beach, "coastal": coastal}
