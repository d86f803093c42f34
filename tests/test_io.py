"""Host-side input and output files: fort.15 reading and writing, and the
fort.63/fort.64/maxele.63 formats against the reference compiler's own
formatting of the reference's edit descriptors (tests/golden/fortran_fmt.json,
made by tests/golden/make_fmt_golden.py with ROCm flang)."""
import dataclasses
import json
import os
import struct

import numpy as np
import pytest

from adcirc_amd import fort15, meshgen, outputs

HERE = os.path.dirname(os.path.abspath(__file__))


def _double(b):
    return struct.unpack("<d", struct.pack("<q", b))[0]


def test_number_formats_match_flang():
    d = json.load(open(os.path.join(HERE, "golden", "fortran_fmt.json")))
    for b, a20, a15 in zip(d["bits"], d["e20"], d["e15"]):
        v = _double(b)
        assert "  " + outputs.e20(v) == a20, repr(v)
        assert " " + outputs._fmt_e(v, 15, 7, 3) == a15, repr(v)
    for h, line in zip(d["headers"], d["header_lines"]):
        got = outputs.header("r", "i", "g", h[0], h[1], h[2], h[3], h[4]).split("\n")[1]
        assert h[5] == outputs.FILE_FMT_VERSION
        assert got == line
    for (t, k), line in zip(d["records"], d["record_lines"]):
        assert outputs.record_header(t, k) == line + "\n"


@pytest.mark.parametrize("name", ["basin", "beach", "beach_lumped", "basin_river"])
def test_fort15_roundtrip(tmp_path, name):
    case = (meshgen.beach() if name.startswith("beach") else
            meshgen.basin_river() if name == "basin_river" else meshgen.basin())
    if name == "beach_lumped":
        case.run.ilump = 1
    f = fort15.from_case(case, fort15.Output(1, 0.0, 1.0, 300), fort15.Output(-1, 0.5, 2.0, 60))
    p = str(tmp_path / "fort.15")
    fort15.write_fort15(p, f)
    from adcirc_amd import driver
    g = fort15.read_fort15(p, case.emo.shape[0], driver.flux_entries(case.mesh))
    if case.qnam is not None:
        # the NFFR block: one line per entry of the IBTYPE 2 and 22 segments
        assert np.array_equal(g.qnam, case.qnam) and np.array_equal(g.qnph, case.qnph)
        assert g.qnam.shape[0] == 20 and len(g.run.fluxf) == 2
    a, b = dataclasses.asdict(case.run), dataclasses.asdict(g.run)
    a.pop("holland"), b.pop("holland")
    assert a == b
    assert np.array_equal(g.emo, case.emo) and np.array_equal(g.efa, case.efa)
    assert g.elev == f.elev and g.vel == f.vel
    assert g.nws == case.run.nws
    assert g.im == (111112 if case.run.ilump else 0)


DOCUMENTED = """\
Hand-written fort.15 in the documented layout            ! RUNDES
test run                                                 ! RUNID
1                       ! NFOVER
1                       ! NABOUT
1                       ! NSCREEN
0                       ! IHOT
1                       ! ICS
111112                  ! IM: lumped GWCE
1                       ! NOLIBF
2                       ! NOLIFA
1                       ! NOLICA
1                       ! NOLICAT
0                       ! NWP
0                       ! NCOR
1                       ! NTIP
0                       ! NWS
8                       ! NRAMP
9.81                    ! G
-1                      ! TAU0
0.5D0                   ! DTDP
1.5                     ! STATIM
0.25                    ! REFTIM
2.0                     ! RNDAY
1.0, 2.0, 0.0, 3.0, 0.75, 0.75, 6.0, 7.0, 8.0  ! DRAMP ... DRAMPElev DRAMPTip ... DUnRampMete
0.35 0.30 0.35          ! A00 B00 C00
0.1 0 0 0.02            ! H0 NODEDRYMIN NODEWETMIN VELMIN
-79.0 35.0              ! SLAM0 SFEA0
0.003                   ! CF
2.0                     ! ESLM
7.29E-5                 ! CORI
1                       ! NTIF
M2
0.242334 1.405189D-4 0.693 1.0 0.0  ! TPK AMIGT ETRF FFT FACET
2                       ! NBFR
M2
1.405189E-4 0.98 12.5   ! AMIG FF FACE
K1
7.292117E-5 1.02 -3.0
M2
0.5 10.0
0.6 11.0
K1
0.1 20.0
0.2 21.0
90.0                    ! ANGINN
1 1.5 2.5 60            ! NOUTE TOUTSE TOUTFE NSPOOLE
2                       ! NSTAE
100.0 200.0
300.0 400.0
0 0.0 0.0 0             ! NOUTV TOUTSV TOUTFV NSPOOLV
0                       ! NSTAV
1 1.5 3.5 120           ! NOUTGE TOUTSGE TOUTFGE NSPOOLGE
-1 2.0 3.5 240          ! NOUTGV TOUTSGV TOUTFGV NSPOOLGV
1                       ! NFREQ
M2
1.405189E-4 1.0 0.0
0.0 1.0 360 0.0         ! THAS THAF NHAINC FMV
0 0 0 0                 ! NHASE NHASV NHAGE NHAGV
0 0                     ! NHSTAR NHSINC
-1 0 1.0D-10 50         ! ITITER ISLDIA CONVCR ITMAX
"""


def test_fort15_documented_layout(tmp_path):
    """A fort.15 typed by hand in the documented layout: comments after the
    values, D exponents, comma separators, the nine-value NRAMP = 8 line,
    the tidal potential (NTIP = 1 and its constituent), stations."""
    p = tmp_path / "fort.15"
    p.write_text(DOCUMENTED)
    f = fort15.read_fort15(str(p), 2)
    r = f.run
    assert (f.rundes, f.runid) == ("Hand-written fort.15 in the documented layout            ! RUNDES"[:80].strip(),
                                   "test run                                                 ! RUNID".strip())
    assert f.im == 111112 and r.ilump == 1 and f.ititer == -1
    assert (r.nolibf, r.nolifa, r.nramp, r.g, r.tau0) == (1, 2, 8, 9.81, -1.0)
    assert (r.dtdp, r.statim, r.reftim) == (0.5, 1.5, 0.25)
    assert r.nsteps == int(np.floor(2.0 * 86400.0 / 0.5))
    assert r.drampelev == 0.75        # DRAMPElev, the fifth value
    assert (r.a00, r.b00, r.c00) == (0.35, 0.30, 0.35)
    assert (r.h0, r.velmin, r.cf, r.eslm, r.cori) == (0.1, 0.02, 0.003, 2.0, 7.29e-5)
    assert r.amig == (1.405189e-4, 7.292117e-5) and r.ff == (0.98, 1.02) and r.face == (12.5, -3.0)
    assert np.array_equal(f.emo, [[0.5, 0.1], [0.6, 0.2]])
    assert np.array_equal(f.efa, [[10.0, 20.0], [11.0, 21.0]])
    assert r.anginn == 90.0
    assert f.ntip == 1 and r.tipot == ((0.242334, 1.405189e-4, 0.693, 1.0, 0.0),)
    assert (r.slam0, r.sfea0) == (-79.0, 35.0)
    assert f.elev_sta == fort15.Output(1, 1.5, 2.5, 60)
    assert np.array_equal(f.stations_e, [[100.0, 200.0], [300.0, 400.0]])
    assert f.elev == fort15.Output(1, 1.5, 3.5, 120)
    assert f.vel == fort15.Output(-1, 2.0, 3.5, 240)
    assert f.nfreq == 1
    assert (r.convcr, r.itmax) == (1e-10, 50)
    # output steps (src/read_input.F:4162-4169): NTCYSGE = int((TOUTSGE -
    # STATIM)*86400/DTDP + 0.5) = 0, NTCYFGE = min(int(2*86400/0.5 + 0.5), NT)
    s, e = f.elev.steps(r)
    assert (s, e) == (0, r.nsteps)
    assert f.elev.datasets(r) == r.nsteps // 120
    assert [it for it in range(1, 400) if f.elev.due(r, it)] == [120, 240, 360]
    s, e = f.vel.steps(r)
    assert s == int(0.5 * 86400.0 / 0.5 + 0.5)
    assert f.vel.due(r, s + 240) and not f.vel.due(r, s) and not f.vel.due(r, s + 120)


@pytest.mark.parametrize("line,value,word", [
    (7, "2", "ICS"), (8, "1", "IM"), (8, "511112", "IM"), (13, "1", "NWP"), (15, "2", "NTIP"),
    (16, "12", "NWS"), (19, "-3", "TAU0"), (9, "3", "NOLIBF")])
def test_fort15_unsupported_options_raise(tmp_path, line, value, word):
    lines = DOCUMENTED.splitlines()
    lines[line - 1] = f"{value}   ! changed"
    p = tmp_path / "fort.15"
    p.write_text("\n".join(lines) + "\n")
    with pytest.raises(fort15.Fort15Error, match=word):
        fort15.read_fort15(str(p), 2)


def test_global_file_and_maxele(tmp_path):
    npn = 7
    eta = np.linspace(-1.0, 2.0, npn)
    u, v = eta * 0.1, -eta * 0.2
    nc = np.array([1, 1, 0, 1, 0, 1, 1], np.int32)
    f63 = outputs.GlobalFile(str(tmp_path / "fort.63"), "run", "id", "grid", 2, npn, 2.0, 300, 1)
    f63.write(600.0, 300, nc, eta)
    f63.write(1200.0, 600, nc, eta + 1.0)
    f64 = outputs.GlobalFile(str(tmp_path / "fort.64"), "run", "id", "grid", 1, npn, 2.0, 600, 2)
    f64.write(1200.0, 600, nc, u, v)
    h, recs = outputs.read_global(str(tmp_path / "fort.63"))
    assert h == {"nsets": 2, "np": npn, "rtype": 1}
    assert [(t, it) for t, it, _ in recs] == [(600.0, 300), (1200.0, 600)]
    exp = np.where(nc == 0, -99999.0, eta + 1.0)
    assert np.allclose(recs[1][2][:, 0], exp, rtol=1e-10, atol=0)
    h, recs = outputs.read_global(str(tmp_path / "fort.64"))
    assert h["rtype"] == 2 and recs[0][2].shape == (npn, 2)
    assert np.all(recs[0][2][nc == 0] == -99999.0)
    lines = open(tmp_path / "fort.63").read().splitlines()
    assert lines[0] == " " + "run".ljust(32) + "  " + "id".ljust(24) + "  " + "grid".ljust(24)
    assert lines[3] == "  " + "1".rjust(8) + "  " + outputs.e20(eta[0])  # 2X,I8,2X,1PE20.10E3
    emax = np.array([0.5, -99999.0, 1.0])
    t = outputs.etamax_time(np.array([10, 0, 3600]), 2.0, 0.5)
    assert np.array_equal(t, [20.0 + 43200.0, -99999.0, 7200.0 + 43200.0])
    outputs.write_maxele(str(tmp_path / "maxele.63"), "run", "id", "grid", 9.0, 3, emax, t)
    h, recs = outputs.read_global(str(tmp_path / "maxele.63"))
    assert h["nsets"] == 2 and len(recs) == 2
    assert np.array_equal(recs[0][2][:, 0], emax) and np.array_equal(recs[1][2][:, 0], t)


def test_driver_refuses_what_it_does_not_write(tmp_path):
    """The case driver refuses hot starts, met-station output, harmonic
    analysis, non-ASCII output and NWS = 8 without vortex parameters, before
    it touches the GPU."""
    from adcirc_amd import driver
    base = fort15.from_case(meshgen.beach(), fort15.Output(1, 0.0, 1.0, 60), fort15.Output())
    cases = {
        "IHOT": dataclasses.replace(base, ihot=67),
        "NHSTAR": dataclasses.replace(base, nhstar=1, nhsinc=100),
        "NOUTM": dataclasses.replace(base, met_sta=fort15.Output(1, 0.0, 1.0, 10),
                                     stations_m=np.array([[100.0, 200.0]])),
        "NOUTE": dataclasses.replace(base, elev_sta=fort15.Output(2, 0.0, 1.0, 10),
                                     stations_e=np.array([[100.0, 200.0]])),
        "NFREQ": dataclasses.replace(base, nfreq=1),
        "NOUTGE": dataclasses.replace(base, elev=fort15.Output(3, 0.0, 1.0, 60)),
        "fort.22": base,  # NWS = 8 (the beach storm) and no parameters passed
    }
    for word, f in cases.items():
        with pytest.raises(fort15.Fort15Error, match=word):
            driver._check(f, None)
    driver._check(base, meshgen.beach().run.holland)  # accepted


def test_station_interpolation_reproduces_linear_fields():
    """The reference's interpolation factors (src/read_input.F:6246-6277)
    reproduce a linear field exactly at any point of the element that holds
    it, sum to one, and are 1/0/0 at a vertex; a point outside the mesh gets
    -99999.  The beach mesh, with its own AREAS."""
    from adcirc_amd import prep
    case = meshgen.beach()
    t = prep.build_tables(case.mesh, case.run.nolifa, case.run.h0, case.run.anginn)
    x, y, nm = t["x"], t["y"], np.asarray(t["nm"]).reshape(-1, 3)
    rng = np.random.default_rng(2)
    xs = rng.uniform(x.min(), x.max(), 25)
    ys = rng.uniform(y.min(), y.max(), 25)
    xs = np.append(xs, [x[nm[7, 1] - 1], x.max() + 1.0e4])
    ys = np.append(ys, [y[nm[7, 1] - 1], y.max() + 1.0e4])
    el, f = outputs.station_factors(x, y, nm, t["areas"], xs, ys)
    assert (el[:-1] >= 0).all() and el[-1] == -1
    assert np.allclose(f[:-1].sum(axis=1), 1.0, atol=1e-12)
    assert np.allclose(sorted(f[-2]), [0.0, 0.0, 1.0], atol=1e-12)
    lin = 0.3 + 2.0e-4 * x - 5.0e-5 * y
    v = outputs.station_values(lin, np.ones(x.shape[0], np.int32), nm, el, f)
    assert np.allclose(v[:-1], 0.3 + 2.0e-4 * xs[:-1] - 5.0e-5 * ys[:-1], rtol=0, atol=1e-12)
    assert v[-1] == -99999.0
    dry = np.ones(x.shape[0], np.int32)
    dry[nm[el[0], 0] - 1] = 0
    assert outputs.station_values(lin, dry, nm, el, f)[0] == -99999.0
