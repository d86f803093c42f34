"""Readers and writers for the reference harness (oracle/_ref/adchp_ref).

Test infrastructure: the harness drives the reference's own Fortran modules
and writes a binary record stream; this module writes the harness inputs for a
case and parses its output into numpy arrays.
"""
from __future__ import annotations

import os
import subprocess
from typing import Dict, List

import numpy as np

from adcirc_amd.fort14 import write_fort14

HERE = os.path.dirname(os.path.abspath(__file__))
REPO = os.path.dirname(HERE)
REF_EXE = os.path.join(REPO, "oracle", "_ref", "adchp_ref")


def write_inputs(case, workdir: str, nsteps: int, dump_every: int = 0,
                 hash_steps: bool = True) -> None:
    os.makedirs(workdir, exist_ok=True)
    write_fort14(os.path.join(workdir, "fort.14"), case.mesh)
    r = case.run
    with open(os.path.join(workdir, "bcforcing.txt"), "w") as f:
        for j in range(r.nbfr):
            for i in range(case.emo.shape[0]):
                f.write(f"{case.emo[i, j]:.17g} {case.efa[i, j]:.17g}\n")
    hol = list(r.holland) if r.holland is not None else []
    hol = hol + [0.0] * (16 - len(hol))

    def fl(v):
        return f"{v:.17g}".replace("e", "d") if "e" in f"{v:.17g}" else f"{v:.17g}d0"

    nml = {
        "nsteps": str(nsteps), "dump_every": str(dump_every),
        "bcfile": "'bcforcing.txt'", "outfile": "'ref_out.bin'",
        "dtdp": fl(r.dtdp), "statim": fl(r.statim), "reftim": fl(r.reftim),
        "nramp": str(r.nramp), "drampelev": fl(r.drampelev), "nolifa": str(r.nolifa),
        "nolica": str(r.nolica), "nolicat": str(r.nolicat), "nolibf": str(r.nolibf),
        "cf": fl(r.cf), "hbreak_in": fl(r.hbreak), "ftheta_in": fl(r.ftheta),
        "fgamma_in": fl(r.fgamma), "tau0_in": fl(r.tau0), "eslm": fl(r.eslm), "cori": fl(r.cori),
        "h0": fl(r.h0), "velmin": fl(r.velmin), "a00": fl(r.a00), "b00": fl(r.b00),
        "c00": fl(r.c00), "itmax": str(r.itmax), "convcr": fl(r.convcr),
        "nws": str(r.nws), "nbfr": str(r.nbfr), "anginn_in": fl(r.anginn),
        "amig_in": ",".join(fl(v) for v in r.amig), "ff_in": ",".join(fl(v) for v in r.ff),
        "face_in": ",".join(fl(v) for v in r.face), "hprm": ",".join(fl(v) for v in hol),
        "ilump_in": str(r.ilump), "hash_steps": "1" if hash_steps else "0",
        "slam0_in": fl(getattr(r, "slam0", 0.0)), "sfea0_in": fl(getattr(r, "sfea0", 0.0)),
    }
    fluxf = list(getattr(r, "fluxf", ()))
    if fluxf:
        nml["nffr_in"] = str(len(fluxf))
        nml["famig_in"] = ",".join(fl(c[0]) for c in fluxf)
        nml["fff_in"] = ",".join(fl(c[1]) for c in fluxf)
        nml["fface_in"] = ",".join(fl(c[2]) for c in fluxf)
        nml["fluxfile"] = "'fluxforcing.txt'"
        # fort.15 order: per constituent, every flux-boundary node in turn
        with open(os.path.join(workdir, "fluxforcing.txt"), "w") as f:
            for j in range(len(fluxf)):
                for e in range(case.qnam.shape[0]):
                    f.write(f"{case.qnam[e, j]:.17g} {case.qnph[e, j]:.17g}\n")
    tip = list(getattr(r, "tipot", ()))
    if tip:
        nml["ntip_in"] = "1"
        nml["ntif_in"] = str(len(tip))
        for k, name in enumerate(("tpk_in", "amigt_in", "etrf_in", "fft_in", "facet_in")):
            nml[name] = ",".join(fl(c[k]) for c in tip)
    with open(os.path.join(workdir, "run.nml"), "w") as f:
        f.write("&run\n")
        for k, v in nml.items():
            f.write(f"  {k} = {v},\n")
        f.write("/\n")


def run_reference(workdir: str, timeout: float = 3600.0) -> str:
    subprocess.run([REF_EXE, "run.nml"], cwd=workdir, check=True, timeout=timeout,
                   stdout=subprocess.PIPE, stderr=subprocess.STDOUT)
    return os.path.join(workdir, "ref_out.bin")


class _Stream:
    def __init__(self, path):
        self.b = open(path, "rb").read()
        self.p = 0

    def i8(self, n=1):
        v = np.frombuffer(self.b, np.int64, n, self.p)
        self.p += 8 * n
        return v

    def f8(self, n=1):
        v = np.frombuffer(self.b, np.float64, n, self.p)
        self.p += 8 * n
        return v

    def arr(self, dt):
        n = int(self.i8()[0])
        v = np.frombuffer(self.b, dt, n, self.p).copy()
        self.p += np.dtype(dt).itemsize * n
        return v

    def peek(self):
        return int(np.frombuffer(self.b, np.int64, 1, self.p)[0])


SETUP_I = ("nm", "nneigh", "neitab", "nneighele", "neitabele", "csr", "nbd", "nbv",
           "lbcodei", "me2gw", "nnodecode0", "nodecode0", "noff0", "mju0")


def read_output(path: str) -> Dict[str, np.ndarray]:
    s = _Stream(path)
    np_, ne, mnei, neta, nvel, nvelme, neimax = (int(v) for v in s.i8(7))
    d: Dict[str, np.ndarray] = dict(np=np_, ne=ne, mnei=mnei, neta=neta, nvel=nvel,
                                    nvelme=nvelme, neimax=neimax)
    d["nm"] = s.arr(np.int32).reshape(3, ne).T.copy()
    d["x"] = s.arr(np.float64); d["y"] = s.arr(np.float64); d["dp"] = s.arr(np.float64)
    d["fdxe"] = s.arr(np.float64).reshape(ne, 3)
    d["fdye"] = s.arr(np.float64).reshape(ne, 3)
    d["areas"] = s.arr(np.float64)
    d["nneigh"] = s.arr(np.int32)
    d["neitab"] = s.arr(np.int32)                      # column-major (np, mnei)
    d["nneighele"] = s.arr(np.int32)
    d["neitabele"] = s.arr(np.int32)
    d["csr"] = s.arr(np.int32).reshape(2, 3, ne).transpose(2, 1, 0).copy()
    d["nbd"] = s.arr(np.int32)[:neta]
    d["nbv"] = s.arr(np.int32)[:nvel]
    d["lbcodei"] = s.arr(np.int32)[:nvel]
    d["csii"] = s.arr(np.float64)[:nvel]
    d["siii"] = s.arr(np.float64)[:nvel]
    d["me2gw"] = s.arr(np.int32)[:nvelme]
    d["emo"] = s.arr(np.float64)                       # column-major (nbfr, neta)
    d["efa"] = s.arr(np.float64)
    d["tau0var"] = s.arr(np.float64); d["evm"] = s.arr(np.float64)
    d["fric"] = s.arr(np.float64)
    d["nnodecode0"] = s.arr(np.int32); d["nodecode0"] = s.arr(np.int32)
    d["noff0"] = s.arr(np.int32)
    d["eta2_0"] = s.arr(np.float64); d["totalarea0"] = s.arr(np.float64)
    d["mju0"] = s.arr(np.int32)
    d["prbckgrnd_mh2o"], d["costset"] = (float(v) for v in s.f8(2))
    d["bndlen2o3"] = s.arr(np.float64)[:nvel]
    steps: List[np.ndarray] = []
    snaps: Dict[int, Dict[str, np.ndarray]] = {}
    while True:
        it = s.peek()
        if it == 0:
            break
        if it < 0:
            s.i8()
            snap = {}
            for k in ("eta1", "eta2", "etas", "uu2", "vv2", "qx2", "qy2", "tk", "etamax"):
                snap[k] = s.arr(np.float64)
            snap["nodecode"] = s.arr(np.int32)
            snap["noff"] = s.arr(np.int32)
            snap["totalarea"] = s.arr(np.float64)
            snap["mju"] = s.arr(np.int32)
            snap["coef"] = s.arr(np.float64)
            snaps[-it] = snap
            continue
        steps.append(s.i8(4).copy())
    st = np.array(steps, dtype=np.int64).reshape(-1, 4)
    d["step_it"] = st[:, 0]
    d["step_numitr"] = st[:, 1]
    d["step_ncchange"] = st[:, 2]
    d["step_hash"] = st[:, 3]
    d["snaps"] = snaps
    return d
