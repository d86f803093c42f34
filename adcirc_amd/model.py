"""Host driver mirroring the reference's timestep interface for the GPU path.

``Model`` owns one device context (adcgpu_create) built from the reference's
mesh tables, and exposes the per-step call sites of src/timestep.F
(begin_step / friction / solve_gwce / wetdry / momentum, or ``timestep`` for
the whole 2DDI step) plus a resident multi-step ``run``.  All arithmetic runs
in libadcgpu.so; this module only moves arrays and scalars.
"""
from __future__ import annotations

import ctypes as C
import math
from typing import Dict, Optional, Sequence

import numpy as np

from . import _lib
from ._lib import check, dp_, ip_

DEG2RAD = 3.141592653589793 / 180.0
PRBCKGRND_MH2O_G = 100.0 * 1013.0 / 1000.0  # divided by g below (src/wind.F:353-354)


def _p(a, t):
    return a.ctypes.data_as(t)


class Model:
    def __init__(self, tables: Dict[str, np.ndarray], run, device: int = 0,
                 tau0var=None, evm=None, fric=None, corif=None, emo=None, efa=None,
                 precision: int = _lib.FP64):
        lib = _lib.load()
        self.lib, self.run = lib, run
        t = tables
        self.np_, self.ne = int(t["np"]), int(t["ne"])
        self.neta, self.mnei = int(t["neta"]), int(t["mnei"])
        self._keep = []

        def ia(a):
            a = np.ascontiguousarray(a, np.int32)
            self._keep.append(a)
            return _p(a, ip_)

        def da(a):
            a = np.ascontiguousarray(a, np.float64)
            self._keep.append(a)
            return _p(a, dp_)

        ne = self.ne
        nm_cm = np.asarray(t["nm"], np.int32).reshape(ne, 3).T.ravel()     # NM(NE,3)
        fdxe = np.asarray(t["fdxe"], np.float64).reshape(ne, 3).ravel()    # FDXE(3,NE)
        fdye = np.asarray(t["fdye"], np.float64).reshape(ne, 3).ravel()
        csr_cm = np.asarray(t["csr"], np.int32).reshape(ne, 3, 2).transpose(2, 1, 0).ravel()
        m = _lib.Mesh()
        m.np, m.ne, m.mnei, m.neta = self.np_, ne, self.mnei, self.neta
        m.nvel, m.nvelme = int(t["nvel"]), int(t["nvelme"])
        m.nm = ia(nm_cm); m.x = da(t["x"]); m.y = da(t["y"]); m.dp = da(t["dp"])
        m.fdxe = da(fdxe); m.fdye = da(fdye); m.areas = da(t["areas"])
        m.nneigh = ia(t["nneigh"]); m.neitab = ia(t["neitab"]); m.csr_position = ia(csr_cm)
        m.nbd = ia(t["nbd"] if self.neta else np.zeros(1)); m.nbv = ia(t["nbv"])
        m.lbcodei = ia(t["lbcodei"]); m.csii = da(t["csii"]); m.siii = da(t["siii"])
        m.me2gw = ia(t["me2gw"])
        p = _lib.Params()
        p.dtdp, p.g, p.h0, p.velmin = run.dtdp, run.g, run.h0, run.velmin
        p.a00, p.b00, p.c00, p.convcr = run.a00, run.b00, run.c00, run.convcr
        p.prbckgrnd_mh2o = PRBCKGRND_MH2O_G / (1000.0 * run.g) * 1000.0 if False else \
            100.0 * 1013.0 / (1000.0 * run.g)
        p.nolifa, p.nolica, p.nolicat, p.nolibf = run.nolifa, run.nolica, run.nolicat, run.nolibf
        p.itmax, p.nws, p.precision = run.itmax, run.nws, precision
        npn = self.np_
        p.tau0var = da(tau0var if tau0var is not None else np.full(npn, run.tau0))
        p.evm = da(evm if evm is not None else np.full(npn, run.eslm))
        p.fric = da(fric if fric is not None else np.full(npn, run.cf))
        p.corif = da(corif if corif is not None else np.full(npn, run.cori))
        self.prbckgrnd_mh2o = p.prbckgrnd_mh2o
        h = C.c_void_p()
        check(lib.adcgpu_create(C.byref(m), C.byref(p), device, C.byref(h)), "adcgpu_create")
        self.h = h
        self._keep = []  # the library copied everything
        # elevation forcing tables (radians), (NBFR, NETA) column-major = [neta][nbfr]
        self.nbfr = run.nbfr
        self.amig = np.array(run.amig, np.float64)
        self.ff = np.array(run.ff, np.float64)
        self.face = np.array([v * DEG2RAD for v in run.face], np.float64)
        if emo is not None:
            self.emo = np.ascontiguousarray(emo, np.float64).reshape(self.neta, self.nbfr)
            self.efa = np.ascontiguousarray(efa, np.float64).reshape(self.neta, self.nbfr)
        else:
            self.emo = np.zeros((self.neta, self.nbfr))
            self.efa = np.zeros((self.neta, self.nbfr))
        self._bc = np.zeros(max(self.neta, 1))

    @staticmethod
    def radians(deg: np.ndarray) -> np.ndarray:
        """Degrees -> radians exactly as the reference converts EFA/FACE."""
        return np.asarray(deg, np.float64) * DEG2RAD

    def close(self):
        if getattr(self, "h", None):
            self.lib.adcgpu_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    # ------------------------------------------------------------- state
    def cold_start(self, eta2, nodecode, noff):
        e = np.ascontiguousarray(eta2, np.float64)
        n = np.ascontiguousarray(nodecode, np.int32)
        o = np.ascontiguousarray(noff, np.int32)
        check(self.lib.adcgpu_cold_start(self.h, _p(e, dp_), _p(n, ip_), _p(o, ip_)), "cold_start")

    def get(self, name: str) -> np.ndarray:
        f = _lib.FIELDS[name]
        n = self.ne if name in ("noff", "alphal") else (self.np_ * self.mnei if name == "coef"
                                                        else self.np_)
        out = np.empty(n, np.int32 if name in _lib.INT_FIELDS else np.float64)
        check(self.lib.adcgpu_get_field(self.h, f, out.ctypes.data, n), f"get {name}")
        return out

    def set(self, name: str, a) -> None:
        dt = np.int32 if name in _lib.INT_FIELDS else np.float64
        a = np.ascontiguousarray(a, dt)
        check(self.lib.adcgpu_set_field(self.h, _lib.FIELDS[name], a.ctypes.data, a.size),
              f"set {name}")

    # ------------------------------------------------------------- forcing
    def times(self, it: int):
        r = self.run
        timeloc = it * r.dtdp + r.statim * 86400.0
        timeh = it * r.dtdp + (r.statim - r.reftim) * 86400.0
        if r.nramp == 0 or r.drampelev < 1.0e-6:
            ramp = 1.0
        else:
            ramp = math.tanh((2.0 * timeloc / 86400.0) / r.drampelev)
        return timeloc, timeh, ramp

    def eta_bc(self, it: int, out: Optional[np.ndarray] = None) -> np.ndarray:
        _, timeh, ramp = self.times(it)
        out = np.empty(max(self.neta, 1)) if out is None else out
        check(self.lib.adcgpu_elevation_bc(self.neta, self.nbfr, _p(self.amig, dp_),
                                           _p(self.ff, dp_), _p(self.face, dp_),
                                           _p(self.emo, dp_), _p(self.efa, dp_), timeh, ramp,
                                           _p(out, dp_)), "elevation_bc")
        return out

    def eta_bc_table(self, it0: int, nsteps: int) -> np.ndarray:
        tab = np.empty((nsteps, max(self.neta, 1)))
        for s in range(nsteps):
            self.eta_bc(it0 + s, tab[s])
        return tab

    # ------------------------------------------------------------- stepping
    def timestep(self, it: int, forcing: Optional[Sequence[np.ndarray]] = None) -> None:
        bc = self.eta_bc(it, self._bc)
        if forcing is not None:
            w, v, p = (np.ascontiguousarray(a, np.float64) for a in forcing)
            rc = self.lib.adcgpu_timestep(self.h, it, _p(bc, dp_), _p(w, dp_), _p(v, dp_),
                                          _p(p, dp_))
        else:
            rc = self.lib.adcgpu_timestep(self.h, it, _p(bc, dp_), None, None, None)
        check(rc, "timestep")

    def numitr(self):
        n, ier = C.c_int(), C.c_int()
        check(self.lib.adcgpu_last_numitr(self.h, C.byref(n), C.byref(ier)), "last_numitr")
        return n.value, ier.value

    def run_resident(self, it0: int, nsteps: int, holland=None, use_graph: bool = False,
                     table: Optional[np.ndarray] = None) -> int:
        tab = self.eta_bc_table(it0, nsteps) if table is None else table
        tab = np.ascontiguousarray(tab, np.float64)
        hol = None
        if holland is not None:
            hol = np.ascontiguousarray(list(holland) + [0.0] * (13 - len(holland)), np.float64)
        rc = self.lib.adcgpu_run(self.h, it0, nsteps, _p(tab, dp_),
                                 _p(hol, dp_) if hol is not None else None, int(use_graph))
        check(rc, "run")
        return int(self.lib.adcgpu_run_iterations(self.h))

    def holland(self, t: float, prm) -> None:
        """Device Holland forcing at time t into WSX2/WSY2/PR2."""
        v = np.ascontiguousarray(list(prm) + [0.0] * (13 - len(prm)), np.float64)
        check(self.lib.adcgpu_holland(self.h, t, _p(v, dp_)), "holland")

    def profile(self, enable: bool = True) -> None:
        check(self.lib.adcgpu_profile(self.h, int(enable)), "profile")

    def profile_read(self):
        ms, n = C.c_double(), C.c_int64()
        check(self.lib.adcgpu_profile_read(self.h, C.byref(ms), C.byref(n)), "profile_read")
        return ms.value, n.value

    def jcg_shape(self):
        """(blocks per JCG launch, resident-block limit, 2048-row chunks)."""
        out = np.zeros(3, np.int32)
        check(self.lib.adcgpu_jcg_shape(self.h, _p(out, ip_)), "jcg_shape")
        return tuple(int(v) for v in out)

    def set_jcg_blocks(self, blocks: int) -> None:
        check(self.lib.adcgpu_set_jcg_blocks(self.h, int(blocks)), "set_jcg_blocks")

    def synchronize(self):
        check(self.lib.adcgpu_synchronize(self.h), "synchronize")


def cold_wetdry_state(tables: Dict[str, np.ndarray], run, eta2=None):
    """Cold-start wet/dry state (src/cstart.F:1326-1429 without subgrid):
    nodes with DP + ETA2 <= H0 start dry, then nodes without a wet element are
    landlocked, and dry nodes get ETA2 = H0 - DP.  Returns (eta2, nodecode, noff)."""
    npn, ne = int(tables["np"]), int(tables["ne"])
    dp = np.asarray(tables["dp"], np.float64)
    eta2 = np.zeros(npn) if eta2 is None else np.asarray(eta2, np.float64).copy()
    nn = np.ones(npn, np.int32)
    noff = np.ones(ne, np.int32)
    if run.nolifa == 2:
        nn[dp + eta2 <= run.h0] = 0
        nm0 = np.asarray(tables["nm"], np.int64).reshape(ne, 3) - 1
        ncele = nn[nm0[:, 0]] * nn[nm0[:, 1]] * nn[nm0[:, 2]] * noff
        mju = np.bincount(nm0.ravel(), weights=np.repeat(ncele, 3), minlength=npn)
        nn[(nn == 1) & (mju == 0)] = 0
        eta2 = np.where(nn == 0, run.h0 - dp, eta2)
    return eta2, nn, noff


def jcg(coef: np.ndarray, jcoef: np.ndarray, rhs: np.ndarray, u: np.ndarray, iparm: np.ndarray,
        rparm: np.ndarray) -> int:
    """Drop-in JCG on host arrays (Fortran layout (ndim, maxnz) passed as the
    column-major flattening); arrays are updated in place like the reference."""
    lib = _lib.load()
    ndim = rhs.shape[0]
    maxnz = coef.size // ndim
    rc = lib.adcgpu_jcg(ndim, ndim, maxnz, _p(jcoef, ip_), _p(coef, dp_), _p(rhs, dp_),
                        _p(u, dp_), _p(iparm, ip_), _p(rparm, dp_))
    if rc < 0:
        check(rc, "jcg")
    return rc
