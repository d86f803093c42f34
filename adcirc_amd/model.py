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
from .prep import tau0_nodal

DEG2RAD = 3.141592653589793 / 180.0


def _p(a, t):
    return a.ctypes.data_as(t)


def c_layout(tables: Dict[str, np.ndarray], run, tau0var=None, evm=None, fric=None,
             corif=None):
    """The adcgpu_mesh / adcgpu_params contents in the C ABI's layout: a dict
    of scalars and a dict of contiguous arrays (int32 / float64), keyed by the
    struct field names."""
    t = tables
    npn, ne = int(t["np"]), int(t["ne"])
    neta = int(t["neta"])
    sc = {"np": npn, "ne": ne, "mnei": int(t["mnei"]), "neta": neta,
          "nvel": int(t["nvel"]), "nvelme": int(t["nvelme"]),
          "dtdp": run.dtdp, "g": run.g, "h0": run.h0, "velmin": run.velmin,
          "a00": run.a00, "b00": run.b00, "c00": run.c00, "convcr": run.convcr,
          # background pressure in m of water, PRBCKGRND*100/(RHOWAT0*G) (src/wind.F:353-354)
          "prbckgrnd_mh2o": 100.0 * 1013.0 / (1000.0 * run.g),
          "nolifa": run.nolifa, "nolica": run.nolica, "nolicat": run.nolicat,
          "nolibf": run.nolibf, "itmax": run.itmax, "nws": run.nws,
          "ilump": int(getattr(run, "ilump", 0)),
          "warm_start": int(getattr(run, "warm_start", 0)),
          "errorelev": float(getattr(run, "errorelev", 0.0))}
    i32 = lambda a: np.ascontiguousarray(a, np.int32)  # noqa: E731
    f64 = lambda a: np.ascontiguousarray(a, np.float64)  # noqa: E731
    ar = {
        "nm": i32(np.asarray(t["nm"], np.int32).reshape(ne, 3).T.ravel()),      # NM(NE,3)
        "x": f64(t["x"]), "y": f64(t["y"]), "dp": f64(t["dp"]),
        "fdxe": f64(np.asarray(t["fdxe"], np.float64).reshape(ne, 3).ravel()),  # FDXE(3,NE)
        "fdye": f64(np.asarray(t["fdye"], np.float64).reshape(ne, 3).ravel()),
        "areas": f64(t["areas"]), "nneigh": i32(t["nneigh"]), "neitab": i32(t["neitab"]),
        "csr_position": i32(np.asarray(t["csr"], np.int32).reshape(ne, 3, 2)
                            .transpose(2, 1, 0).ravel()),                         # (NE,3,2)
        "nbd": i32(t["nbd"] if neta else np.zeros(1)), "nbv": i32(t["nbv"]),
        "lbcodei": i32(t["lbcodei"]), "csii": f64(t["csii"]), "siii": f64(t["siii"]),
        "me2gw": i32(t["me2gw"]),
        "tau0var": f64(tau0var if tau0var is not None else tau0_nodal(run.tau0, tables["dp"])),
        "evm": f64(evm if evm is not None else np.full(npn, run.eslm)),
        "fric": f64(fric if fric is not None else np.full(npn, run.cf)),
        "corif": f64(corif if corif is not None else np.full(npn, run.cori)),
    }
    return sc, ar


_MESH_I = ("np", "ne", "mnei", "neta", "nvel", "nvelme")
_MESH_A = ("nm", "x", "y", "dp", "fdxe", "fdye", "areas", "nneigh", "neitab", "csr_position",
           "nbd", "nbv", "lbcodei", "csii", "siii", "me2gw")
_PRM_D = ("dtdp", "g", "h0", "velmin", "a00", "b00", "c00", "convcr", "prbckgrnd_mh2o",
          "errorelev")
_PRM_I = ("nolifa", "nolica", "nolicat", "nolibf", "itmax", "nws", "ilump", "warm_start")
_PRM_A = ("tau0var", "evm", "fric", "corif")


def write_case(path: str, tables, run, eta2, nodecode, noff, eta_bc_table, holland=None,
               **nodal) -> None:
    """Writes a case for the native driver (adcgpu_run_case): the C-ABI
    arrays, the cold-start state, the elevation-boundary table and the
    Holland parameters.  Format: magic "ADCGCASE", then records of
    (16-byte name, int32 kind 0=float64 1=int32, int32 0, int64 count, data)."""
    sc, ar = c_layout(tables, run, **nodal)
    recs = [("precision", np.array([_lib.FP64], np.int32))]
    recs += [(k, np.array([sc[k]], np.int32)) for k in _MESH_I + _PRM_I]
    recs += [(k, np.array([sc[k]], np.float64)) for k in _PRM_D]
    recs += [(k, ar[k]) for k in _MESH_A + _PRM_A]
    recs += [("eta2_0", np.asarray(eta2, np.float64)),
             ("nodecode_0", np.asarray(nodecode, np.int32)),
             ("noff_0", np.asarray(noff, np.int32)),
             ("eta_bc_table", np.asarray(eta_bc_table, np.float64).ravel())]
    if holland is not None:
        recs.append(("holland", np.asarray(list(holland) + [0.0] * (13 - len(holland)),
                                           np.float64)))
    with open(path, "wb") as f:
        f.write(b"ADCGCASE")
        for name, a in recs:
            a = np.ascontiguousarray(a)
            kind = 1 if a.dtype == np.int32 else 0
            assert a.dtype in (np.int32, np.float64), name
            f.write(name.encode().ljust(16, b"\0"))
            f.write(np.array([kind, 0], np.int32).tobytes())
            f.write(np.array([a.size], np.int64).tobytes())
            f.write(a.tobytes())


FLUX_CODES = (2, 12, 22, 52)


def flux_amplitudes(lbcodei, qnam, qnph_deg):
    """QNAM and QNPH (radians) laid out (NFFR, NVEL) column-major as the
    reference holds them (src/read_input.F:3562-3586): fort.15's rows, one per
    flow-boundary entry whose code takes a specified flux, in NBV order; zero
    at the other entries."""
    lbc = np.asarray(lbcodei, np.int32)
    qnam = np.asarray(qnam, np.float64)
    idx = np.flatnonzero(np.isin(lbc, FLUX_CODES))
    if qnam.ndim != 2 or idx.shape[0] != qnam.shape[0]:
        raise ValueError(f"{idx.shape[0]} flux-boundary entries, QNAM has {qnam.shape[0]} rows")
    am = np.zeros((lbc.shape[0], qnam.shape[1]))
    ph = np.zeros_like(am)
    am[idx] = qnam
    ph[idx] = np.asarray(qnph_deg, np.float64) * DEG2RAD
    return am, ph


def run_times(run, it: int):
    """(timeloc, timeh, ramp) of step it: TimeLoc (src/timestep.F:257), the
    harmonic time TIMEH of the elevation forcing, and RampElev, the tanh ramp
    on DRampElev (src/timestep.F:296-304, 350)."""
    timeloc = it * run.dtdp + run.statim * 86400.0
    timeh = it * run.dtdp + (run.statim - run.reftim) * 86400.0
    if run.nramp == 0 or run.drampelev < 1.0e-6:
        ramp = 1.0
    else:
        ramp = math.tanh((2.0 * timeloc / 86400.0) / run.drampelev)
    return timeloc, timeh, ramp


def elevation_table(run, emo, efa, it0: int, nsteps: int) -> np.ndarray:
    """Elevation-boundary values for steps it0 .. it0+nsteps-1 through
    adcgpu_elevation_bc (host computation, no device context); emo/efa are
    (NETA, NBFR), efa in radians."""
    lib = _lib.load()
    emo = np.ascontiguousarray(emo, np.float64)
    efa = np.ascontiguousarray(efa, np.float64)
    neta, nbfr = emo.shape
    amig = np.array(run.amig, np.float64)
    ff = np.array(run.ff, np.float64)
    face = np.array([v * DEG2RAD for v in run.face], np.float64)
    tab = np.empty((nsteps, max(neta, 1)))
    for s in range(nsteps):
        _, timeh, ramp = run_times(run, it0 + s)
        row = tab[s]
        check(lib.adcgpu_elevation_bc(neta, nbfr, _p(amig, dp_), _p(ff, dp_), _p(face, dp_),
                                      _p(emo, dp_), _p(efa, dp_), timeh, ramp, _p(row, dp_)),
              "elevation_bc")
    return tab


class Model:
    def __init__(self, tables: Dict[str, np.ndarray], run, device: int = 0,
                 tau0var=None, evm=None, fric=None, corif=None, emo=None, efa=None,
                 precision: int = _lib.FP64, qnam=None, qnph=None):
        lib = _lib.load()
        self.lib, self.run = lib, run
        t = tables
        self.np_, self.ne = int(t["np"]), int(t["ne"])
        self.neta, self.mnei = int(t["neta"]), int(t["mnei"])
        sc, ar = c_layout(tables, run, tau0var, evm, fric, corif)
        m = _lib.Mesh()
        for k in _MESH_I:
            setattr(m, k, sc[k])
        for k in _MESH_A:
            setattr(m, k, _p(ar[k], ip_ if ar[k].dtype == np.int32 else dp_))
        p = _lib.Params()
        for k in _PRM_D + _PRM_I:
            setattr(p, k, sc[k])
        p.precision = precision
        for k in _PRM_A:
            setattr(p, k, _p(ar[k], dp_))
        self.prbckgrnd_mh2o = p.prbckgrnd_mh2o
        self.last_rebuilds = 0
        h = C.c_void_p()
        check(lib.adcgpu_create(C.byref(m), C.byref(p), device, C.byref(h)), "adcgpu_create")
        self.h = h
        del ar  # the library copied everything
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
        if getattr(run, "tipot", ()):
            from .prep import lonlat_cpp
            slam, sfea = lonlat_cpp(t["x"], t["y"], run.slam0, run.sfea0)
            self.set_tide_potential(run.tipot, slam, sfea, run)
        if run.nolibf == 2:
            check(lib.adcgpu_set_hybrid_friction(self.h, float(run.hbreak), float(run.ftheta),
                                                 float(run.fgamma)), "set_hybrid_friction")
        self.flux = None
        if getattr(run, "fluxf", ()):
            if qnam is None or qnph is None or "bndlen2o3" not in t:
                raise ValueError("flux boundaries need QNAM/QNPH and the tables' BNDLEN2O3")
            self.set_flux_boundary(t["lbcodei"], t["bndlen2o3"], run.fluxf, qnam, qnph)
        elif int(t.get("nvel", 0)) and (np.asarray(t["lbcodei"]) == 30).any():
            # outward radiation only: the line integrals without flux forcing
            bl = np.ascontiguousarray(t["bndlen2o3"], np.float64)
            check(lib.adcgpu_set_flux_boundary(self.h, _p(bl, dp_)), "set_flux_boundary")

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
        return run_times(self.run, it)

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

    def set_tide_potential(self, constituents, slam, sfea, run) -> None:
        """Tidal potential (NTIP = 1): constituents as fort.15's (TPK, AMIGT,
        ETRF, FFT, FACET degrees), node longitudes/latitudes in radians; the
        ramp and time origin from the run."""
        c = [np.array([x[k] for x in constituents], np.float64) for k in range(5)]
        c[4] = c[4] * DEG2RAD
        sl = np.ascontiguousarray(slam, np.float64)
        sf = np.ascontiguousarray(sfea, np.float64)
        check(self.lib.adcgpu_set_tide_potential(
            self.h, len(constituents), *[_p(a, dp_) for a in c], _p(sl, dp_), _p(sf, dp_),
            int(run.nramp), float(run.drampelev), float(run.statim), float(run.reftim)),
            "set_tide_potential")

    def set_flux_boundary(self, lbcodei, bndlen2o3, constituents, qnam, qnph) -> None:
        """Specified normal flux boundaries with periodic forcing (NFFR > 0):
        constituents as fort.15's (FAMIG rad/s, FFF, FFACE degrees), QNAM and
        QNPH (degrees) per flux-boundary entry and constituent."""
        bl = np.ascontiguousarray(bndlen2o3, np.float64)
        check(self.lib.adcgpu_set_flux_boundary(self.h, _p(bl, dp_)), "set_flux_boundary")
        am, ph = flux_amplitudes(lbcodei, qnam, qnph)
        c = [np.array([x[k] for x in constituents], np.float64) for k in range(3)]
        c[2] = c[2] * DEG2RAD
        self.flux = dict(lbc=np.ascontiguousarray(lbcodei, np.int32), am=am, ph=ph,
                         famig=c[0], fff=c[1], fface=c[2])
        self._qn = np.zeros(am.shape[0])

    def flux_bc(self, it: int, out: Optional[np.ndarray] = None) -> np.ndarray:
        """QN2 of step it (adcgpu_flux_bc; RampExtFlux shares the elevation
        ramp's length)."""
        f = self.flux
        _, timeh, ramp = self.times(it)
        out = np.empty(f["lbc"].shape[0]) if out is None else out
        check(self.lib.adcgpu_flux_bc(f["lbc"].shape[0], _p(f["lbc"], ip_), f["famig"].shape[0],
                                      _p(f["famig"], dp_), _p(f["fff"], dp_), _p(f["fface"], dp_),
                                      _p(f["am"], dp_), _p(f["ph"], dp_), timeh, ramp,
                                      _p(out, dp_)), "flux_bc")
        return out

    def flux_rows(self, it0: int, nsteps: int) -> None:
        """QN2 of steps it0 .. it0+nsteps-1 for the next resident run
        (adcgpu_set_flux_table); nothing without flux boundaries."""
        if self.flux is None:
            return
        qt = np.empty((nsteps, self.flux["lbc"].shape[0]))
        for s in range(nsteps):
            self.flux_bc(it0 + s, qt[s])
        check(self.lib.adcgpu_set_flux_table(self.h, nsteps, _p(qt, dp_)), "set_flux_table")

    # ------------------------------------------------------------- stepping
    def timestep(self, it: int, forcing: Optional[Sequence[np.ndarray]] = None,
                 tip2: Optional[np.ndarray] = None) -> None:
        """One step; `tip2` replaces the step's device-formed tidal potential
        with the caller's values (the per-step calls of adcgpu_timestep with
        ADCGPU_TIP2 set after adcgpu_begin_step)."""
        bc = self.eta_bc(it, self._bc)
        if tip2 is not None or self.flux is not None:
            lib = self.lib
            check(lib.adcgpu_begin_step(self.h, it), "begin_step")
            if tip2 is not None:
                self.set("tip2", tip2)
            if self.flux is not None:
                # a fresh array each step: the copy may still read it after
                # the call returns
                self._qn = q = self.flux_bc(it)
                check(lib.adcgpu_set_flux(self.h, _p(q, dp_)), "set_flux")
            check(lib.adcgpu_friction(self.h), "friction")
            if forcing is not None and self.run.nws != 0:
                w, v, p = (np.ascontiguousarray(a, np.float64) for a in forcing)
                check(lib.adcgpu_set_forcing(self.h, _p(w, dp_), _p(v, dp_), _p(p, dp_)),
                      "set_forcing")
            check(lib.adcgpu_solve_gwce(self.h, it, _p(bc, dp_)), "solve_gwce")
            check(lib.adcgpu_wetdry(self.h, it), "wetdry")
            check(lib.adcgpu_momentum(self.h), "momentum")
            return
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
        self.flux_rows(it0, nsteps)
        hol = None
        if holland is not None:
            hol = np.ascontiguousarray(list(holland) + [0.0] * (13 - len(holland)), np.float64)
        rc = self.lib.adcgpu_run(self.h, it0, nsteps, _p(tab, dp_),
                                 _p(hol, dp_) if hol is not None else None, int(use_graph))
        self.last_run_rc = rc
        check(rc, "run")
        self.last_rebuilds = int(self.lib.adcgpu_run_rebuilds(self.h))
        return int(self.lib.adcgpu_run_iterations(self.h))

    def run_stats(self):
        """(JCG iterations, matrix rebuilds, solves that reached ITMAX, first
        step that raised a run-time error or 0, solves that started from the
        learned guess) of the last resident run."""
        out = np.zeros(5, np.int64)
        check(self.lib.adcgpu_run_stats(self.h, out.ctypes.data_as(C.POINTER(C.c_int64))),
              "run_stats")
        return tuple(int(v) for v in out)

    def last_solve(self):
        """(STPTST the last JCG solve ended with, its bound ZETA, squared
        initial pseudo-residual norms of the reference's and of the learned
        start, or 0 without the learned start)."""
        out = np.zeros(4, np.float64)
        check(self.lib.adcgpu_last_solve(self.h, out.ctypes.data_as(dp_)), "last_solve")
        return tuple(float(v) for v in out)

    # ------------------------------------------------------------- snapshot
    def save_state(self) -> np.ndarray:
        """The whole state between steps (adcgpu_state_save), as bytes."""
        n = int(self.lib.adcgpu_state_bytes(self.h))
        buf = np.empty(n, np.uint8)
        check(self.lib.adcgpu_state_save(self.h, buf.ctypes.data, n), "state_save")
        return buf

    def load_state(self, buf: np.ndarray) -> None:
        buf = np.ascontiguousarray(buf, np.uint8)
        check(self.lib.adcgpu_state_load(self.h, buf.ctypes.data, buf.size), "state_load")

    # ------------------------------------------------------------- learned start
    def ml_params(self, new=None) -> np.ndarray:
        """The network's FP32 parameters (after setting `new`, if given)."""
        fp = C.POINTER(C.c_float)
        out = np.zeros(_lib.ML_NPARAM, np.float32)
        src = None if new is None else np.ascontiguousarray(new, np.float32)
        check(self.lib.adcgpu_ml_params(self.h, None if src is None else _p(src, fp),
                                        _p(out, fp)), "ml_params")
        return out

    def ml_collect(self, per_step: int) -> None:
        check(self.lib.adcgpu_ml_collect(self.h, int(per_step)), "ml_collect")

    def ml_samples(self) -> int:
        return int(self.lib.adcgpu_ml_samples(self.h))

    def ml_train(self, epochs: int, lr: float = 1e-3, seed: int = 1) -> float:
        loss = C.c_double()
        check(self.lib.adcgpu_ml_train(self.h, int(epochs), float(lr), int(seed),
                                       C.byref(loss)), "ml_train")
        return loss.value

    def ml_grad(self, x, y, w):
        fp = C.POINTER(C.c_float)
        x = np.ascontiguousarray(x, np.float32)
        y = np.ascontiguousarray(y, np.float32)
        w = np.ascontiguousarray(w, np.float32)
        g = np.zeros(_lib.ML_NPARAM, np.float32)
        loss = C.c_double()
        check(self.lib.adcgpu_ml_grad(self.h, x.shape[0], _p(x, fp), _p(y, fp), _p(w, fp),
                                      _p(g, fp), C.byref(loss)), "ml_grad")
        return loss.value, g

    def ml_features(self):
        fp = C.POINTER(C.c_float)
        x = np.zeros((self.np_, 48), np.float32)
        s = np.zeros(self.np_)
        check(self.lib.adcgpu_ml_features(self.h, _p(x, fp), _p(s, dp_)), "ml_features")
        return x, s

    def ml_correction(self) -> np.ndarray:
        out = np.zeros(self.np_)
        check(self.lib.adcgpu_ml_correction(self.h, _p(out, dp_)), "ml_correction")
        return out

    def holland(self, t: float, prm) -> None:
        """Device Holland forcing at time t into WSX2/WSY2/PR2."""
        v = np.ascontiguousarray(list(prm) + [0.0] * (13 - len(prm)), np.float64)
        check(self.lib.adcgpu_holland(self.h, t, _p(v, dp_)), "holland")

    def profile(self, enable: bool = True, launches: int = 0) -> None:
        """Start (enable, with event pairs for `launches` JCG launches made
        ahead) or stop timing the JCG launches."""
        n = max(int(launches), 1) if enable else 0
        check(self.lib.adcgpu_profile(self.h, n), "profile")

    def profile_read(self):
        ms, n = C.c_double(), C.c_int64()
        check(self.lib.adcgpu_profile_read(self.h, C.byref(ms), C.byref(n)), "profile_read")
        return ms.value, n.value

    def jcg_shape(self):
        """(blocks per JCG launch, resident-block limit, 2048-row chunks)."""
        out = np.zeros(3, np.int32)
        check(self.lib.adcgpu_jcg_shape(self.h, _p(out, ip_)), "jcg_shape")
        return tuple(int(v) for v in out)

    def set_jcg_blocks(self, blocks: int):
        """Limits JCG launches to `blocks` blocks (0: the device's resident
        count); larger systems are then solved memory-resident."""
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
        rparm: np.ndarray, noalloc: bool = False, raise_errors: bool = True) -> int:
    """Drop-in JCG on host arrays (Fortran layout (ndim, maxnz) passed as the
    column-major flattening); arrays are updated in place like the reference.
    ``noalloc`` selects the entry point that allocates nothing
    (adcgpu_jcg_noalloc).  Returns ITPACK's ier, or the negative status code
    when ``raise_errors`` is off."""
    lib = _lib.load()
    ndim = rhs.shape[0]
    maxnz = coef.size // ndim
    fn = lib.adcgpu_jcg_noalloc if noalloc else lib.adcgpu_jcg
    rc = fn(ndim, ndim, maxnz, _p(jcoef, ip_), _p(coef, dp_), _p(rhs, dp_), _p(u, dp_),
            _p(iparm, ip_), _p(rparm, dp_))
    if rc < 0 and raise_errors:
        check(rc, "jcg")
    return rc


def set_wait_limit(seconds: float) -> None:
    """Bound of every wait inside the persistent JCG kernel (default 0.5 s;
    0 makes every wait fail at once, which forces the error path)."""
    check(_lib.load().adcgpu_set_wait_limit(float(seconds)), "set_wait_limit")


class Batch:
    """Several models stepped together (adcgpu_batch): their JCG solves run
    as one cooperative launch per step (more when their blocks do not all
    fit at once); each member's results equal its own resident run's."""

    def __init__(self, models: Sequence["Model"]):
        self.models = list(models)
        self.lib = self.models[0].lib
        arr = (C.c_void_p * len(self.models))(*[m.h for m in self.models])
        h = C.c_void_p()
        check(self.lib.adcgpu_batch_create(arr, len(self.models), C.byref(h)), "batch_create")
        self.h = h

    def run(self, it0: int, nsteps: int, holland=None, tables=None):
        """Advance every member by ``nsteps`` steps from ``it0``.  ``holland``
        and ``tables`` are per member (None: no met forcing, and each
        member's elevation-boundary table).  Returns each member's JCG
        iterations."""
        n = len(self.models)
        tabs = [np.ascontiguousarray(m.eta_bc_table(it0, nsteps) if tables is None
                                     else tables[k], np.float64)
                for k, m in enumerate(self.models)]
        tp = (dp_ * n)(*[_p(t, dp_) for t in tabs])
        for m in self.models:
            m.flux_rows(it0, nsteps)
        hp = None
        hols = []
        if holland is not None:
            for k in range(n):
                hk = holland[k]
                hols.append(None if hk is None else
                            np.ascontiguousarray(list(hk) + [0.0] * (13 - len(hk)), np.float64))
            hp = (dp_ * n)(*[_p(h, dp_) if h is not None else None for h in hols])
        check(self.lib.adcgpu_batch_run(self.h, it0, nsteps, tp, hp), "batch_run")
        return [m.run_stats()[0] for m in self.models]

    def close(self):
        if getattr(self, "h", None):
            self.lib.adcgpu_batch_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
