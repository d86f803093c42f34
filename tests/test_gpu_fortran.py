"""The ISO_C_BINDING interface module (adcirc_amd/fortran/adcgpu_iface.F90)
driving the drop-in JCG from Fortran, as src/gwce.F:2002-2005 calls ITPACK."""
import os
import subprocess

import pytest

from gpu_helpers import log

EXE = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))),
                   "adcirc_amd", "fortran", "jcg_dropin_check")


@pytest.mark.gpu
def test_fortran_interface_drives_dropin_jcg():
    if not os.path.exists(EXE):
        pytest.skip("Fortran check not built (no flang at build time)")
    log("case fortran: adcgpu_jcg through the bind(C) interface, 1000-row system")
    r = subprocess.run([EXE], capture_output=True, text=True, timeout=120)
    log("  " + r.stdout.strip().splitlines()[0] if r.stdout.strip() else "  (no output)")
    assert r.returncode == 0, r.stdout + r.stderr
    assert "ier=0" in r.stdout


EXE_RES = os.path.join(os.path.dirname(EXE), "resident_check")


@pytest.mark.gpu
def test_fortran_resident_path_bitexact(beach_golden, tmp_path):
    """The per-step ABI driven from Fortran the way TIMESTEP would drive it
    (adcirc_amd/fortran/resident_check.F90): the mesh and parameters go to
    adcgpu_create through the bind(C) adcgpu_mesh / adcgpu_params types, the
    elevation boundary values come from adcgpu_elevation_bc, the met forcing
    from the oracle's host forcing.  The beach storm (wet/dry, Holland
    vortex), 60 steps: the Fortran-driven state equals the oracle's bit for
    bit."""
    import numpy as np

    import oracle_lib
    from adcirc_amd import meshgen
    from adcirc_amd.model import DEG2RAD, c_layout, run_times
    from gpu_helpers import golden_tables

    if not os.path.exists(EXE_RES):
        pytest.skip("Fortran check not built (no flang at build time)")
    d, run, nsteps = beach_golden, meshgen.beach().run, 60
    log(f"case fortran: beach, {nsteps} steps through the per-step ABI from Fortran")
    t = golden_tables(d)
    npn, ne, neta, nbfr = t["np"], t["ne"], t["neta"], run.nbfr
    sc, ar = c_layout(t, run, d["tau0var"], d["evm"], d["fric"], np.full(npn, run.cori))
    o = oracle_lib.OracleModel(d, run, oracle_lib.BLOCKED)
    forcing = [np.concatenate([o.wsx, o.wsy, o.pr])] if o.hol is not None else []
    for it in range(1, nsteps + 1):
        o.step(it)
        if o.hol is not None:
            forcing.append(np.concatenate([o.wsx, o.wsy, o.pr]))
    i32 = lambda a: np.ascontiguousarray(a, np.int32).tobytes()  # noqa: E731
    f64 = lambda a: np.ascontiguousarray(a, np.float64).tobytes()  # noqa: E731
    hdr = [npn, ne, t["mnei"], neta, t["nvel"], t["nvelme"], run.nolifa, run.nolica,
           run.nolicat, run.nolibf, run.itmax, run.nws, sc["ilump"], sc["warm_start"], nbfr,
           nsteps, 1, 0]
    hdd = [sc[k] for k in ("dtdp", "g", "h0", "velmin", "a00", "b00", "c00", "convcr",
                           "prbckgrnd_mh2o", "errorelev")]
    times = [run_times(run, it) for it in range(1, nsteps + 1)]
    parts = [i32(hdr), f64(hdd)]
    parts += [i32(ar["nm"]), f64(ar["x"]), f64(ar["y"]), f64(ar["dp"]), f64(ar["fdxe"]),
              f64(ar["fdye"]), f64(ar["areas"]), i32(ar["nneigh"]), i32(ar["neitab"]),
              i32(ar["csr_position"]), i32(ar["nbd"])]
    if t["nvel"]:
        parts += [i32(ar["nbv"]), i32(ar["lbcodei"]), f64(ar["csii"]), f64(ar["siii"])]
    if t["nvelme"]:
        parts += [i32(ar["me2gw"])]
    parts += [f64(ar["tau0var"]), f64(ar["evm"]), f64(ar["fric"]), f64(ar["corif"]),
              f64(d["eta2_0"]), i32(d["nodecode0"]), i32(d["noff0"])]
    if nbfr:
        parts += [f64(run.amig), f64(run.ff), f64([v * DEG2RAD for v in run.face])]
    if neta * nbfr:
        parts += [f64(d["emo"]), f64(d["efa"])]
    parts += [f64([x[1] for x in times]), f64([x[2] for x in times])]
    parts += [f64(f) for f in forcing]
    fin, fout = tmp_path / "case.bin", tmp_path / "out.bin"
    fin.write_bytes(b"".join(parts))
    r = subprocess.run([EXE_RES, str(fin), str(fout)], capture_output=True, text=True,
                       timeout=120)
    log("  " + (r.stdout.strip().splitlines() or ["(no output)"])[-1])
    assert r.returncode == 0, r.stdout + r.stderr
    raw = fout.read_bytes()
    numitr, ier = np.frombuffer(raw[:8], np.int32)
    fl = np.frombuffer(raw[8:8 + 32 * npn], np.float64).reshape(4, npn)
    ints = np.frombuffer(raw[8 + 32 * npn:], np.int32)
    assert ier == 0 and numitr == o.numitr()
    for k, a in zip(("eta2", "uu2", "vv2", "etamax"), fl):
        b = o.array(k)
        assert np.array_equal(a, b), f"{k}: {int((a != b).sum())} differ"
    assert np.array_equal(ints[:npn], o.iarray("nodecode"))
    assert np.array_equal(ints[npn:npn + ne], o.iarray("noff"))
