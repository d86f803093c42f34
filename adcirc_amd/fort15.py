"""fort.15, the model parameter and periodic boundary condition file, for the
2DDI subset this package runs.

Lines are read in the order of the reference's reader, src/read_input.F:949-4566
(which differs from the list in docs/technical_reference/input_files/fort15.rst
in two places: no WTIMINC line is read when NWS = 0, and the CORI line is read
whatever NCOR is).  Each line is read like a Fortran list-directed READ: the
leading values are taken, separated by blanks or commas, ``D`` exponents
allowed, and the rest of the line (usually a ``!`` comment) is ignored.  Title
and tag lines (RUNDES, RUNID, BOUNTAG, ALPHA, TIPOTAG, NAMEFR) are read whole.

Options outside the subset raise ``Fort15Error`` naming the option and the
line: IM other than 0, 111111 (the same defaults spelled out) or 111112
(lumped GWCE); spherical coordinates (ICS = 2) and NCOR = 1; nodal attributes
(NWP > 0); the tidal potential with self-attraction (NTIP = 2, fort.24); a
tidal-potential ramp other than the elevation ramp; met forcing other than NWS = 0 or the
parametric vortex NWS = 8; TAU0 schemes other than a constant, -1 and -2;
the 3D-coupled friction of NOLIBF = 3; NRS/ice variants of NWS; normal flux read from
fort.20 (NFFR 0 or -1; the periodic flux of NFFR > 0 on IBTYPE 2, 12, 22, 52
boundaries is read); a flux-forcing ramp other than the elevation ramp;
station lists read from a separate file (NSTA < 0).
"""
from __future__ import annotations

import dataclasses
import math
import re
from typing import List, Optional, Sequence, Tuple

import numpy as np

from .meshgen import RunConfig


class Fort15Error(ValueError):
    pass


@dataclasses.dataclass
class Output:
    """One NOUT*/TOUTS*/TOUTF*/NSPOOL* output control line."""

    nout: int = 0
    touts: float = 0.0      # days after cold start
    toutf: float = 0.0
    nspool: int = 0

    def steps(self, run: RunConfig) -> Tuple[int, int]:
        """(NTCYS, NTCYF): output is written at the steps NTCYS + k*NSPOOL,
        k >= 1, up to NTCYF (src/read_input.F:4162-4164,
        src/write_output.F:1812-1822)."""
        nt = run.nsteps
        s = int((self.touts - run.statim) * (86400.0 / run.dtdp) + 0.5)
        f = int((self.toutf - run.statim) * (86400.0 / run.dtdp) + 0.5)
        return s, min(f, nt)

    def datasets(self, run: RunConfig) -> int:
        """NDSETS, the record count written in the file header
        (src/read_input.F:4168-4169)."""
        if self.nspool == 0:
            return 0
        s, f = self.steps(run)
        return int((f - s) / self.nspool)

    def due(self, run: RunConfig, it: int) -> bool:
        if self.nout == 0 or self.nspool <= 0:
            return False
        s, f = self.steps(run)
        return s < it <= f and (it - s) % self.nspool == 0


@dataclasses.dataclass
class Fort15:
    rundes: str
    runid: str
    run: RunConfig
    emo: np.ndarray                 # [neta, nbfr] amplitudes (m)
    efa: np.ndarray                 # [neta, nbfr] phases (degrees)
    bountag: List[str]
    elev: Output                    # NOUTGE ... (fort.63, maxele.63)
    vel: Output                     # NOUTGV ... (fort.64)
    nfover: int = 1
    nabout: int = 1
    nscreen: int = 100
    ihot: int = 0
    ics: int = 1
    im: int = 0
    nws: int = 0
    nramp: int = 1
    ncor: int = 0
    ntip: int = 0
    slam0: float = 0.0
    sfea0: float = 0.0
    ititer: int = 1
    isldia: int = 0
    nodedrymin: int = 0
    nodewetmin: int = 0
    vortex_line: Tuple[float, ...] = ()   # IREFYR IREFMO IREFDAY IREFHR StormNumber BLAdj
    elev_sta: Output = dataclasses.field(default_factory=Output)
    vel_sta: Output = dataclasses.field(default_factory=Output)
    met_sta: Output = dataclasses.field(default_factory=Output)
    met: Output = dataclasses.field(default_factory=Output)
    stations_e: np.ndarray = dataclasses.field(default_factory=lambda: np.zeros((0, 2)))
    stations_v: np.ndarray = dataclasses.field(default_factory=lambda: np.zeros((0, 2)))
    stations_m: np.ndarray = dataclasses.field(default_factory=lambda: np.zeros((0, 2)))
    nfreq: int = 0
    nhstar: int = 0
    nhsinc: int = 0
    # periodic normal flux (NFFR > 0): FBOUNTAG per constituent, QNAM/QNPH
    # (degrees) per flux-boundary entry and constituent
    fbountag: List[str] = dataclasses.field(default_factory=list)
    qnam: Optional[np.ndarray] = None
    qnph: Optional[np.ndarray] = None


_NUM = re.compile(r"[\s,]+")


class _Lines:
    def __init__(self, text: str, path: str):
        self.lines = text.splitlines()
        self.i = 0
        self.path = path

    def raw(self) -> str:
        if self.i >= len(self.lines):
            raise Fort15Error(f"{self.path}: ends early after line {self.i}")
        s = self.lines[self.i]
        self.i += 1
        return s

    def values(self, n: int, what: str) -> List[str]:
        line = self.raw()
        body = line.split("!")[0].split("/")[0]
        toks = [t for t in _NUM.split(body.strip()) if t]
        if len(toks) < n:
            raise Fort15Error(f"{self.path}:{self.i}: {what}: expected {n} values in {line!r}")
        return toks[:n]

    def ints(self, n: int, what: str) -> List[int]:
        out = []
        for t in self.values(n, what):
            try:
                out.append(int(t))
            except ValueError:
                v = _float(t)
                if v != int(v):
                    raise Fort15Error(f"{self.path}:{self.i}: {what}: {t!r} is not an integer")
                out.append(int(v))
        return out

    def floats(self, n: int, what: str) -> List[float]:
        return [_float(t) for t in self.values(n, what)]

    def int1(self, what: str) -> int:
        return self.ints(1, what)[0]

    def float1(self, what: str) -> float:
        return self.floats(1, what)[0]

    def fail(self, msg: str):
        raise Fort15Error(f"{self.path}:{self.i}: {msg}")


def _float(t: str) -> float:
    return float(t.replace("D", "E").replace("d", "e"))


def _output(r: _Lines, what: str) -> Output:
    v = r.values(4, what)
    return Output(int(_float(v[0])), _float(v[1]), _float(v[2]), int(_float(v[3])))


def _stations(r: _Lines, what: str) -> np.ndarray:
    n = r.int1(what)
    if n < 0:
        r.fail(f"{what} < 0 (station list in a separate file) is not supported")
    return np.array([r.floats(2, f"{what} station") for _ in range(n)]).reshape(n, 2)


def read_fort15(path: str, neta: int, flux_entries: int = 0) -> Fort15:
    """Read fort.15 for a mesh with ``neta`` elevation-boundary nodes (NETA of
    fort.14) and ``flux_entries`` flow-boundary entries of IBTYPE 2, 12 or 22
    (the NVEL entries whose QNAM/QNPH lines the NFFR block holds,
    src/read_input.F:3562-3586; 0: fort.14 has no flux boundaries and no
    NFFR block is read)."""
    with open(path) as f:
        r = _Lines(f.read(), path)
    rundes = r.raw().rstrip("\n")[:80].strip()
    runid = r.raw().rstrip("\n")[:80].strip()
    nfover = r.int1("NFOVER")
    nabout = r.int1("NABOUT")
    nscreen = r.int1("NSCREEN")
    ihot = r.int1("IHOT")
    ics = r.int1("ICS")
    if ics != 1:
        r.fail(f"ICS = {ics}: only Cartesian coordinates (ICS = 1) are supported")
    im = r.int1("IM")
    if im in (0, 111111):
        ilump = 0
    elif im == 111112:
        ilump = 1
    else:
        r.fail(f"IM = {im}: supported are 0, 111111 and 111112 (2DDI defaults, lumped GWCE)")
    nolibf = r.int1("NOLIBF")
    nolifa = r.int1("NOLIFA")
    nolica = r.int1("NOLICA")
    nolicat = r.int1("NOLICAT")
    nwp = r.int1("NWP")
    if nwp != 0:
        r.fail(f"NWP = {nwp}: nodal attributes (fort.13) are not supported")
    ncor = r.int1("NCOR")
    if ncor != 0:
        r.fail("NCOR = 1 needs spherical coordinates, which are not supported")
    ntip = r.int1("NTIP")
    if ntip not in (0, 1):
        r.fail(f"NTIP = {ntip}: supported are 0 and 1 (no self-attraction/loading)")
    nws = r.int1("NWS")
    if nws not in (0, 8, -8):
        r.fail(f"NWS = {nws}: supported are 0 and the parametric vortex 8")
    nramp = r.int1("NRAMP")
    if not 0 <= nramp <= 8:
        r.fail(f"NRAMP = {nramp}")
    g = r.float1("G")
    tau0 = r.float1("TAU0")
    if tau0 < 0 and tau0 not in (-1.0, -2.0):
        r.fail(f"TAU0 = {tau0}: supported are a constant >= 0, -1 and -2")
    dtdp = r.float1("DTDP")
    statim = r.float1("STATIM")
    reftim = r.float1("REFTIM")
    vortex: Tuple[float, ...] = ()
    if abs(nws) == 8:
        vortex = tuple(r.floats(6, "IREFYR IREFMO IREFDAY IREFHR StormNumber BLAdj"))
    rnday = r.float1("RNDAY")
    ramps = r.floats(1 if nramp < 2 else nramp + 1, "DRAMP")
    # NRAMP 0, 1: one value; 2: DRAMP DRAMPExtFlux FluxSettlingTime; 3 adds
    # DRAMPIntFlux; 4 DRAMPElev; ... (src/read_input.F:2893-2950)
    drampelev = ramps[4] if nramp >= 4 else ramps[0]
    if ntip and nramp >= 5 and ramps[5] != drampelev:
        r.fail("DRAMPTip differs from DRAMPElev: one ramp length is supported")
    if flux_entries and nramp >= 2 and (ramps[1] != drampelev or ramps[2] != 0.0):
        r.fail("DRAMPExtFlux differs from DRAMPElev, or FluxSettlingTime is not zero: "
               "one ramp length is supported")
    a00, b00, c00 = r.floats(3, "A00 B00 C00")
    if nolifa in (2, 3):
        v = r.values(4, "H0 NODEDRYMIN NODEWETMIN VELMIN")
        h0, ndry, nwet, velmin = _float(v[0]), int(_float(v[1])), int(_float(v[2])), _float(v[3])
    else:
        h0, ndry, nwet, velmin = r.float1("H0"), 0, 0, 0.01
    slam0, sfea0 = r.floats(2, "SLAM0 SFEA0")
    hbreak, ftheta, fgamma = 1.0, 10.0, 1.0 / 3.0
    if nolibf == 0:
        cf = r.float1("TAU")
    elif nolibf == 1:
        cf = r.float1("CF")
    elif nolibf == 2:
        # src/read_input.F: CF (the minimum), HBREAK, FTHETA, FGAMMA
        cf, hbreak, ftheta, fgamma = r.floats(4, "CF HBREAK FTHETA FGAMMA")
    else:
        r.fail(f"NOLIBF = {nolibf}: supported are 0, 1 and 2")
    eslm = r.float1("ESLM")
    if eslm < 0:
        r.fail("ESLM < 0 (Smagorinsky viscosity) is not supported")
    cori = r.float1("CORI")
    ntif = r.int1("NTIF")
    tipot = []
    for _ in range(ntif):
        r.raw()  # TIPOTAG
        tipot.append(tuple(r.floats(5, "TPK AMIGT ETRF FFT FACET")))
    # src/read_input.F:4818-4825
    if (ntip == 0) != (ntif == 0):
        r.fail(f"NTIP = {ntip} with NTIF = {ntif}: both zero or both non-zero")
    nbfr = r.int1("NBFR")
    tags, amig, ff, face = [], [], [], []
    for _ in range(nbfr):
        tags.append(r.raw().strip()[:5])
        a, b, c = r.floats(3, "AMIG FF FACE")
        amig.append(a); ff.append(b); face.append(c)
    emo = np.zeros((neta, nbfr))
    efa = np.zeros((neta, nbfr))
    for k in range(nbfr):
        r.raw()  # ALPHA
        for j in range(neta):
            emo[j, k], efa[j, k] = r.floats(2, "EMO EFA")
    anginn = r.float1("ANGINN")
    fluxf, fbountag, qnam, qnph = [], [], None, None
    if flux_entries:
        # src/read_input.F:3503-3588
        nffr = r.int1("NFFR")
        if nffr <= 0:
            r.fail(f"NFFR = {nffr}: normal flux from fort.20 is not supported (NFFR > 0)")
        for _ in range(nffr):
            fbountag.append(r.raw().strip()[:5])
            fluxf.append(tuple(r.floats(3, "FAMIGT FFF FFACE")))
        qnam = np.zeros((flux_entries, nffr))
        qnph = np.zeros((flux_entries, nffr))
        for k in range(nffr):
            r.raw()  # ALPHAQ
            for j in range(flux_entries):
                qnam[j, k], qnph[j, k] = r.floats(2, "QNAM QNPH")
    elev_sta = _output(r, "NOUTE TOUTSE TOUTFE NSPOOLE")
    st_e = _stations(r, "NSTAE")
    vel_sta = _output(r, "NOUTV TOUTSV TOUTFV NSPOOLV")
    st_v = _stations(r, "NSTAV")
    met_sta, st_m = Output(), np.zeros((0, 2))
    if nws != 0:
        met_sta = _output(r, "NOUTM TOUTSM TOUTFM NSPOOLM")
        st_m = _stations(r, "NSTAM")
    elev = _output(r, "NOUTGE TOUTSGE TOUTFGE NSPOOLGE")
    vel = _output(r, "NOUTGV TOUTSGV TOUTFGV NSPOOLGV")
    met = Output()
    if nws != 0:
        met = _output(r, "NOUTGW TOUTSGW TOUTFGW NSPOOLGW")
    nfreq = r.int1("NFREQ")
    for _ in range(nfreq):
        r.raw()
        r.floats(3, "HAFREQ HAFF HAFACE")
    r.values(4, "THAS THAF NHAINC FMV")
    r.values(4, "NHASE NHASV NHAGE NHAGV")
    nhstar, nhsinc = r.ints(2, "NHSTAR NHSINC")
    v = r.values(4, "ITITER ISLDIA CONVCR ITMAX")
    ititer, isldia, convcr, itmax = int(_float(v[0])), int(_float(v[1])), _float(v[2]), int(_float(v[3]))
    run = RunConfig(dtdp=dtdp, nsteps=int(math.floor(rnday * 86400.0 / dtdp)), statim=statim,
                    reftim=reftim, nramp=nramp, drampelev=drampelev, nolifa=nolifa,
                    nolica=nolica, nolicat=nolicat, nolibf=nolibf, cf=cf, tau0=tau0, eslm=eslm,
                    cori=cori, h0=h0, velmin=velmin, a00=a00, b00=b00, c00=c00, itmax=itmax,
                    convcr=convcr, ilump=ilump, anginn=anginn, g=g, amig=tuple(amig),
                    ff=tuple(ff), face=tuple(face), holland=None, tipot=tuple(tipot),
                    slam0=slam0, sfea0=sfea0, fluxf=tuple(fluxf), hbreak=hbreak,
                    ftheta=ftheta, fgamma=fgamma)
    return Fort15(rundes, runid, run, emo, efa, tags, elev, vel, nfover=nfover, nabout=nabout,
                  nscreen=nscreen, ihot=ihot, ics=ics, im=im, nws=nws, nramp=nramp, ncor=ncor,
                  ntip=ntip, slam0=slam0, sfea0=sfea0, ititer=ititer, isldia=isldia,
                  nodedrymin=ndry, nodewetmin=nwet, vortex_line=vortex, elev_sta=elev_sta,
                  vel_sta=vel_sta, met_sta=met_sta, met=met, stations_e=st_e, stations_v=st_v,
                  stations_m=st_m, nfreq=nfreq, nhstar=nhstar, nhsinc=nhsinc,
                  fbountag=fbountag, qnam=qnam, qnph=qnph)


def _g(v: float) -> str:
    return repr(float(v))


def write_fort15(path: str, f: Fort15) -> None:
    """Write ``f`` in the layout read_fort15 (and the reference) reads; every
    float is written with repr, so it reads back exactly."""
    r = f.run
    L: List[str] = [f.rundes[:80], f.runid[:80]]

    def add(vals: Sequence, name: str):
        L.append(" ".join(str(v) if isinstance(v, (int, np.integer)) else _g(v)
                          for v in vals) + f" ! {name}")

    add([f.nfover], "NFOVER")
    add([f.nabout], "NABOUT")
    add([f.nscreen], "NSCREEN")
    add([f.ihot], "IHOT")
    add([f.ics], "ICS")
    add([f.im if f.im else (111112 if r.ilump else 0)], "IM")
    add([r.nolibf], "NOLIBF")
    add([r.nolifa], "NOLIFA")
    add([r.nolica], "NOLICA")
    add([r.nolicat], "NOLICAT")
    add([0], "NWP")
    add([0], "NCOR")
    add([1 if r.tipot else 0], "NTIP")
    add([f.nws], "NWS")
    add([r.nramp], "NRAMP")
    add([r.g], "G")
    add([r.tau0], "TAU0")
    add([r.dtdp], "DTDP")
    add([r.statim], "STATIM")
    add([r.reftim], "REFTIM")
    if abs(f.nws) == 8:
        v = list(f.vortex_line) or [2000, 1, 1, 0, 1, 1.0]
        add([int(x) for x in v[:5]] + [float(v[5])],
            "IREFYR IREFMO IREFDAY IREFHR StormNumber BLAdj")
    # RNDAY such that floor(RNDAY*86400/DTDP) = nsteps
    add([(r.nsteps + 0.5) * r.dtdp / 86400.0], "RNDAY")
    if r.nramp >= 4:
        add([r.drampelev, r.drampelev, 0.0, r.drampelev, r.drampelev]
            + [r.drampelev] * (r.nramp - 4), "DRAMP ... DRAMPElev")
    elif r.nramp >= 2:
        add([r.drampelev, r.drampelev, 0.0] + [r.drampelev] * (r.nramp - 2), "DRAMP ...")
    else:
        add([r.drampelev], "DRAMP")
    add([r.a00, r.b00, r.c00], "A00 B00 C00")
    if r.nolifa in (2, 3):
        add([r.h0, f.nodedrymin, f.nodewetmin, r.velmin], "H0 NODEDRYMIN NODEWETMIN VELMIN")
    else:
        add([r.h0], "H0")
    add([f.slam0, f.sfea0], "SLAM0 SFEA0")
    if r.nolibf == 2:
        add([r.cf, r.hbreak, r.ftheta, r.fgamma], "CF HBREAK FTHETA FGAMMA")
    else:
        add([r.cf], "TAU" if r.nolibf == 0 else "CF")
    add([r.eslm], "ESLM")
    add([r.cori], "CORI")
    add([len(r.tipot)], "NTIF")
    for k, c in enumerate(r.tipot):
        L.append(f"T{k + 1}")
        add(list(c), "TPK AMIGT ETRF FFT FACET")
    add([r.nbfr], "NBFR")
    tags = list(f.bountag) + [f"F{k + 1}" for k in range(len(f.bountag), r.nbfr)]
    for k in range(r.nbfr):
        L.append(tags[k][:5])
        add([r.amig[k], r.ff[k], r.face[k]], "AMIG FF FACE")
    for k in range(r.nbfr):
        L.append(tags[k][:10])
        for j in range(f.emo.shape[0]):
            L.append(f"{_g(f.emo[j, k])} {_g(f.efa[j, k])}")
    add([r.anginn], "ANGINN")
    if r.fluxf:
        add([len(r.fluxf)], "NFFR")
        ftags = list(f.fbountag) + [f"Q{k + 1}" for k in range(len(f.fbountag), len(r.fluxf))]
        for k, c in enumerate(r.fluxf):
            L.append(ftags[k][:5])
            add(list(c), "FAMIGT FFF FFACE")
        for k in range(len(r.fluxf)):
            L.append(ftags[k][:10])
            for j in range(f.qnam.shape[0]):
                L.append(f"{_g(f.qnam[j, k])} {_g(f.qnph[j, k])}")

    def out(o: Output, name: str):
        add([o.nout, o.touts, o.toutf, o.nspool], name)

    def sta(s: np.ndarray, name: str):
        add([s.shape[0]], name)
        for x, y in s:
            L.append(f"{_g(x)} {_g(y)}")

    out(f.elev_sta, "NOUTE TOUTSE TOUTFE NSPOOLE")
    sta(f.stations_e, "NSTAE")
    out(f.vel_sta, "NOUTV TOUTSV TOUTFV NSPOOLV")
    sta(f.stations_v, "NSTAV")
    if f.nws != 0:
        out(f.met_sta, "NOUTM TOUTSM TOUTFM NSPOOLM")
        sta(f.stations_m, "NSTAM")
    out(f.elev, "NOUTGE TOUTSGE TOUTFGE NSPOOLGE")
    out(f.vel, "NOUTGV TOUTSGV TOUTFGV NSPOOLGV")
    if f.nws != 0:
        out(f.met, "NOUTGW TOUTSGW TOUTFGW NSPOOLGW")
    add([0], "NFREQ")
    add([0.0, 0.0, 0, 0.0], "THAS THAF NHAINC FMV")
    add([0, 0, 0, 0], "NHASE NHASV NHAGE NHAGV")
    add([f.nhstar, f.nhsinc], "NHSTAR NHSINC")
    add([f.ititer if f.ititer else (-1 if r.ilump else 1), f.isldia, r.convcr, r.itmax],
        "ITITER ISLDIA CONVCR ITMAX")
    with open(path, "w") as fh:
        fh.write("\n".join(L) + "\n")


def from_case(case, elev: Optional[Output] = None, vel: Optional[Output] = None,
              rundes: str = "", runid: str = "") -> Fort15:
    """The fort.15 of a generated case (its RunConfig, EMO/EFA), with the
    given global output controls."""
    r = case.run
    return Fort15(rundes or case.mesh.title, runid or "adcirc_amd", r,
                  np.asarray(case.emo, np.float64), np.asarray(case.efa, np.float64),
                  [f"F{k + 1}" for k in range(r.nbfr)], elev or Output(), vel or Output(),
                  nws=r.nws, nramp=r.nramp, ititer=-1 if r.ilump else 1,
                  im=111112 if r.ilump else 0, ntip=r.ntip, slam0=r.slam0, sfea0=r.sfea0,
                  qnam=case.qnam, qnph=case.qnph)
