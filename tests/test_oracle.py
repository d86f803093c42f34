"""The oracle against the reference build (golden fixtures) and against
closed-form results that pin it independently of the reference.

The fixtures were written by the reference's own Fortran modules
(tests/golden/make_golden.py); the oracle must reproduce them bit for bit:
the 64-bit hash of the carried state (ETA2, UU2, VV2, ETAS, TK, NODECODE,
NOFF) after every step, the JCG iteration count of every step, and full
snapshots.
"""
import math

import numpy as np
import pytest

import oracle_lib
from adcirc_amd import meshgen

SNAP_F = ("eta1", "eta2", "etas", "uu2", "vv2", "tk", "etamax", "totalarea")


def _check_snapshot(m, d, s):
    for k in SNAP_F:
        a, b = m.array(k), d[f"s{s}_{k}"]
        assert np.array_equal(a, b), f"step {s}: {k} differs (max {np.abs(a - b).max():.3e})"
    assert np.array_equal(m.iarray("nodecode"), d[f"s{s}_nodecode"])
    assert np.array_equal(m.iarray("noff"), d[f"s{s}_noff"])


def _run(d, run, nsteps):
    m = oracle_lib.OracleModel(d, run, oracle_lib.SEQUENTIAL)
    snaps = set(int(s) for s in d["snap_steps"])
    for it in range(1, nsteps + 1):
        m.step(it)
        assert m.numitr() == d["step_numitr"][it - 1], f"step {it}: NUMITR"
        assert m.state_hash() == d["step_hash"][it - 1], f"step {it}: state hash"
        if it in snaps:
            _check_snapshot(m, d, it)
        if it == 1:
            coef = m.array("coef", int(d["np"]) * int(d["mnei"]))
            assert np.array_equal(coef, d["s1_coef"]), "COEF after step 1"
    return m


def test_beach_wetdry_storm_hour_matches_reference(beach_golden):
    """Wet/dry + Holland storm, 3600 steps: bit-identical every step."""
    _run(beach_golden, meshgen.beach().run, 3600)


def test_beach_lumped_gwce_hour_matches_reference(beach_lumped_golden):
    """The same beach with the lumped GWCE (ILump = 1: diagonal LHS, explicit
    solve, src/gwce.F:596-635, 1856-1862, 2009-2022), 3600 steps:
    bit-identical every step."""
    run = meshgen.beach().run
    run.ilump = 1
    _run(beach_lumped_golden, run, 3600)


def basin_linear_run():
    """The basin with linear friction (NOLIBF = 0, TAU = 1e-4 /s) and no
    finite-amplitude or advective terms (NOLIFA = NOLICA = NOLICAT = 0)."""
    run = meshgen.basin().run
    run.nolibf, run.cf, run.nolifa, run.nolica, run.nolicat = 0, 1.0e-4, 0, 0, 0
    return run


def test_basin_linear_terms_match_reference(basin_linear_golden):
    """Linear friction, linear continuity and no advection (the IFLINBF,
    IFNLFA, IFNLCT and IFNLCAT switches off their default), 3600 steps:
    bit-identical every step."""
    _run(basin_linear_golden, basin_linear_run(), 3600)


def test_basin_matches_reference_first_3000_steps(basin_golden):
    _run(basin_golden, meshgen.basin().run, 3000)


def test_basin_full_day_matches_reference(basin_golden):
    """The 1-day, 43,200-step basin run: every step's hash and the final state."""
    d = basin_golden
    _run(d, meshgen.basin().run, int(d["run_nsteps"]))


# ---------------------------------------------------------------- JCG
def _jcg_call(lib, st, n, J, Cm, rhs, u, itmax, zeta):
    J = np.asfortranarray(J.astype(np.int32)).T.copy()   # (maxnz, n) rows = columns
    mz = J.shape[0]
    jc = np.ascontiguousarray(J.ravel())
    co = np.ascontiguousarray(np.asarray(Cm, np.float64).T.ravel())
    r = np.ascontiguousarray(rhs, np.float64).copy()
    uu = np.ascontiguousarray(u, np.float64).copy()
    iparm = np.zeros(12, np.int32)
    rparm = np.zeros(12, np.float64)
    lib.adco_dfault(iparm.ctypes.data_as(oracle_lib._ip), rparm.ctypes.data_as(oracle_lib._dp))
    iparm[0], iparm[1], iparm[3] = itmax, 0, 16
    rparm[0] = zeta
    nw = 4 * n + 4 * itmax
    wk = np.zeros(nw)
    ier = lib.adco_jcg(st, n, n, mz, jc.ctypes.data_as(oracle_lib._ip),
                       co.ctypes.data_as(oracle_lib._dp), r.ctypes.data_as(oracle_lib._dp),
                       uu.ctypes.data_as(oracle_lib._dp), nw, wk.ctypes.data_as(oracle_lib._dp),
                       iparm.ctypes.data_as(oracle_lib._ip), rparm.ctypes.data_as(oracle_lib._dp))
    return ier, iparm, rparm, uu, r, co.reshape(mz, n).T, jc.reshape(mz, n).T


def test_jcg_sequence_matches_reference(jcg_golden):
    """Back-to-back solves in one ITPACK state (q1 and the eigenvalue history
    carry over, COEF is rescaled in place) reproduce the reference's JCG."""
    g = jcg_golden
    lib = oracle_lib.load()
    st = oracle_lib.Itpack()
    lib.adco_itpack_init(st, oracle_lib.SEQUENTIAL)
    prev = {}
    for name in [str(s) for s in g["names"]]:
        J, Cm = g[f"{name}_in_jcoef"], g[f"{name}_in_coef"]
        n = J.shape[0]
        itmax, zeta = g[f"{name}_prm"]
        ier, iparm, rparm, u, rhs, co, jc = _jcg_call(lib, st, n, J, Cm, g[f"{name}_in_rhs"],
                                                     g[f"{name}_in_u"], int(itmax), zeta)
        assert ier == int(g[f"{name}_ier"])
        assert iparm[0] == int(g[f"{name}_iparm1"]), name
        assert iparm[7] == int(g[f"{name}_iparm8"])
        assert np.array_equal(u, g[f"{name}_u"]), name
        assert np.array_equal(rhs, g[f"{name}_rhs"]), name
        assert np.array_equal(co, g[f"{name}_coef"]), name
        assert np.array_equal(jc, g[f"{name}_jcoef"]), name
        assert np.array_equal(rparm[[0, 1, 2, 10, 11]], g[f"{name}_rparm"]), name
        prev = {"coef": co}


def test_jcg_closed_forms():
    """Pins independent of the reference: a diagonal system is solved exactly
    in one iteration; a Laplacian-plus-mass system agrees with a dense direct
    solve to the stopping tolerance; the returned matrix is the input up to
    the scale/unscale rounding."""
    lib = oracle_lib.load()
    st = oracle_lib.Itpack()
    lib.adco_itpack_init(st, oracle_lib.SEQUENTIAL)
    rng = np.random.default_rng(5)
    n = 40
    J = np.zeros((n, 3), np.int32); J[:, 0] = np.arange(1, n + 1)
    C = np.zeros((n, 3)); C[:, 0] = rng.uniform(0.5, 4.0, n)
    b = rng.normal(size=n)
    ier, iparm, _, u, _, co, _ = _jcg_call(lib, st, n, J, C, b, np.zeros(n), 50, 1e-10)
    assert ier == 0 and iparm[0] <= 1
    assert np.allclose(u, b / C[:, 0], rtol=1e-14, atol=0)
    # 1-D Laplacian + mass, dense check
    n = 60
    J = np.zeros((n, 4), np.int32); C = np.zeros((n, 4))
    A = np.zeros((n, n))
    for i in range(n):
        J[i, 0] = i + 1; C[i, 0] = 2.1; A[i, i] = 2.1
        c = 1
        for j in (i - 1, i + 1):
            if 0 <= j < n:
                J[i, c] = j + 1; C[i, c] = -1.0; A[i, j] = -1.0; c += 1
    b = rng.normal(size=n)
    ier, iparm, _, u, _, co, _ = _jcg_call(lib, st, n, J, C, b, np.zeros(n), 100, 1e-12)
    x = np.linalg.solve(A, b)
    assert ier == 0
    assert np.linalg.norm(u - x) / np.linalg.norm(x) < 1e-9
    assert np.allclose(co, C, rtol=4e-16 * 8, atol=0)


def test_blocked_reduction_is_close_to_exact():
    lib = oracle_lib.load()
    rng = np.random.default_rng(1)
    for n in (1, 7, 2048, 2049, 100000):
        x = rng.normal(size=n)
        y = rng.normal(size=n)
        exact = math.fsum((x * y).tolist())
        px, py = x.ctypes.data_as(oracle_lib._dp), y.ctypes.data_as(oracle_lib._dp)
        blk = lib.adco_dot(oracle_lib.BLOCKED, n, px, py)
        seq = lib.adco_dot(oracle_lib.SEQUENTIAL, n, px, py)
        scale = float(np.sum(np.abs(x * y)))
        assert abs(blk - exact) <= 1e-14 * scale * max(1, math.log2(n))
        assert abs(seq - exact) <= 1e-15 * scale * n


# --------------------------------------------------- physics pins (no reference)
def _small_case_tables(case):
    from adcirc_amd import prep
    t = prep.build_tables(case.mesh, case.run.nolifa, case.run.h0, case.run.anginn)
    npn = t["np"]
    d = dict(t)
    d["emo"] = np.asarray(case.emo, np.float64).ravel()
    d["efa"] = (np.asarray(case.efa, np.float64) * (3.141592653589793 / 180.0)).ravel()
    d["tau0var"] = np.full(npn, case.run.tau0)
    d["evm"] = np.full(npn, case.run.eslm)
    d["fric"] = np.full(npn, case.run.cf)
    d["nodecode0"] = np.ones(npn, np.int32)
    d["noff0"] = np.ones(t["ne"], np.int32)
    d["eta2_0"] = np.zeros(npn)
    d["prbckgrnd_mh2o"] = 100.0 * 1013.0 / (1000.0 * case.run.g)
    return d


def test_lake_at_rest_stays_at_rest():
    """No forcing, flat water: every state array stays exactly zero."""
    c = meshgen.basin(nx=21, ny=11, lx=20e3, ly=10e3, amp=0.0)
    d = _small_case_tables(c)
    m = oracle_lib.OracleModel(d, c.run)
    for it in range(1, 30):
        m.step(it)
    for k in ("eta2", "uu2", "vv2", "etas"):
        assert np.all(m.array(k) == 0.0), k


def test_gwce_matrix_structure():
    """Consistent GWCE matrix before boundary rows: symmetric, and with no
    gravity term (A00 = 0) each interior row sums to 2/12+... i.e. the mass
    matrix row sum (1/dt^2 + tau0/(2 dt)) * TotalArea/3."""
    c = meshgen.basin(nx=15, ny=9, lx=14e3, ly=8e3, amp=0.0)
    c.run.a00 = 0.0
    c.run.b00 = 0.65
    d = _small_case_tables(c)
    d["nbd"] = np.zeros(0, np.int32)
    d["neta"] = 0
    m = oracle_lib.OracleModel(d, c.run)
    m.step(1)
    npn, mnei = int(d["np"]), int(d["mnei"])
    coef = m.array("coef", npn * mnei).reshape(mnei, npn).T
    nei = d["neitab"].reshape(mnei, npn).T
    A = np.zeros((npn, npn))
    for i in range(npn):
        for j in range(int(d["nneigh"][i])):
            A[i, nei[i, j] - 1] += coef[i, j]
    assert np.allclose(A, A.T, rtol=1e-12, atol=0)
    dt, tau0 = c.run.dtdp, c.run.tau0
    ta = m.array("totalarea")
    rows = A.sum(axis=1)
    expect = (1.0 / dt + tau0 / 2.0) / dt * ta / 3.0
    assert np.allclose(rows, expect, rtol=1e-12)


def test_friction_closed_form():
    """Quadratic friction: TK = min(Cf |u| / H, 2/dt) at every node."""
    c = meshgen.basin(nx=11, ny=6, lx=10e3, ly=5e3)
    d = _small_case_tables(c)
    m = oracle_lib.OracleModel(d, c.run)
    for it in range(1, 40):
        m.step(it)
    u, v = m.array("uu2"), m.array("vv2")
    eta = m.array("eta2")
    m.step(40)
    tk = m.array("tk")
    h = d["dp"] + eta
    expect = np.minimum(c.run.cf * np.sqrt(u * u + v * v) / h, 2.0 / c.run.dtdp)
    assert np.allclose(tk, expect, rtol=1e-14, atol=0)


def test_peak_surge_and_speed_tracking(beach_golden):
    """ETAMAX/UMAX and the steps of their last rise (collectMinMaxData,
    src/write_output.F:6343-6405) against running maxima formed here from
    each step's state over 120 steps of the wet/dry beach storm."""
    d = beach_golden
    m = oracle_lib.OracleModel(d, meshgen.beach().run, oracle_lib.SEQUENTIAL)
    npn = int(d["np"])
    emax = np.full(npn, -99999.0)
    umax = np.full(npn, -99999.0)
    eit = np.zeros(npn, np.int64)
    uit = np.zeros(npn, np.int64)
    for it in range(1, 121):
        m.step(it)
        wet = m.iarray("nodecode") == 1
        e2 = m.array("eta2")
        u, v = m.array("uu2"), m.array("vv2")
        sp = np.sqrt(u * u + v * v)
        up_e = wet & (e2 > emax)
        emax[up_e], eit[up_e] = e2[up_e], it
        up_u = wet & (sp > umax)
        umax[up_u], uit[up_u] = sp[up_u], it
    assert np.array_equal(m.array("etamax"), emax)
    assert np.array_equal(m.array("umax"), umax)
    assert np.array_equal(m.iarray("etamax_it"), eit)
    assert np.array_equal(m.iarray("umax_it"), uit)
    assert (uit > 0).any() and (eit > 0).any()


def test_basin_tidal_potential_matches_reference(basin_tide_golden):
    """The basin with the tidal potential of M2, S2, K1 and O1 (NTIP = 1,
    src/timestep.F:1507-1562; TIP1/TIP2 in src/gwce.F:1181-1185 and
    src/momentum.F:443-450), one hour: bit-identical every step."""
    d = basin_tide_golden
    _run(d, meshgen.basin_tide().run, int(d["run_nsteps"]))


def test_beach_hybrid_friction_matches_reference(beach_hybrid_golden):
    """Hybrid bottom friction (NOLIBF = 2, src/nodalattr.F:2400-2411) in the
    friction pass and in the wetting criterion (src/wetdry.F:415-451) on the
    wet/dry storm beach, half an hour: bit-identical every step."""
    from conftest import beach_hybrid_run
    d = beach_hybrid_golden
    _run(d, beach_hybrid_run(), int(d["run_nsteps"]))


def test_basin_rivers_match_reference(basin_river_golden):
    """The basin with two rivers, a specified normal flux as an essential
    condition (IBTYPE 2) and as a natural one (IBTYPE 22), each with a steady
    and an M2 constituent (NFFR = 2, src/timestep.F:605-612, 861-878; the
    flux line integrals of src/gwce.F:1737-1834 and the normal velocity of
    src/momentum.F:810-848), one hour: bit-identical every step."""
    d = basin_river_golden
    m = _run(d, meshgen.basin_river().run, int(d["run_nsteps"]))
    assert np.abs(m.flux["qn2"]).max() > 0.2  # the ramp is still rising


def test_basin_radiation_matches_reference(basin_radiation_golden):
    """The basin with an outward-radiating stretch of its north edge
    (IBTYPE 30: QN1 from the last velocities, src/timestep.F:929-940, and the
    radiation QFORCE of src/gwce.F:1750-1754, 1785-1789), one hour:
    bit-identical every step; the tables match the reference's."""
    from adcirc_amd import prep
    d = basin_radiation_golden
    case = meshgen.basin_radiation()
    t = prep.build_tables(case.mesh, case.run.nolifa, case.run.h0, case.run.anginn)
    for k in ("nbv", "lbcodei", "me2gw", "bndlen2o3"):
        assert np.array_equal(t[k], d[k]), k
    m = _run(d, case.run, int(d["run_nsteps"]))
    q = m.array("qn1", int(d["nvel"]))
    assert np.abs(q[np.asarray(d["lbcodei"]) == 30]).max() > 0.0


def test_river_radiation_matches_reference(basin_river52_golden):
    """The rivers with the west one as river radiation (IBTYPE 52: the
    specified flux plus Celerity*(ETAS/DT + Tau0*(ETA1 - ElevDisc)) with no
    flux-settling time, ElevDisc the cold-start elevation; src/gwce.F:
    1767-1775, src/timestep.F:906-925, src/cstart.F:1188-1192), one hour:
    bit-identical every step."""
    d = basin_river52_golden
    assert 52 in set(d["lbcodei"].tolist())
    _run(d, meshgen.basin_river(west_ibtype=52).run, int(d["run_nsteps"]))


def test_river_tables_match_reference(basin_river_golden):
    """The product's flow-boundary tables for a mesh with flux segments
    (LBCODEI 2, 22, the ME2GW precedence of flux codes at shared end nodes,
    BNDLEN2O3 with zero after each segment's last node) equal the reference's
    (src/mesh.F:2230-2420)."""
    from adcirc_amd import prep
    d = basin_river_golden
    case = meshgen.basin_river()
    t = prep.build_tables(case.mesh, case.run.nolifa, case.run.h0, case.run.anginn)
    for k in ("nbv", "lbcodei", "csii", "siii", "me2gw", "bndlen2o3"):
        assert np.array_equal(t[k], d[k]), k
    assert set(np.unique(d["lbcodei"])) >= {2, 22}
    assert (d["bndlen2o3"] == 0).sum() == len(case.mesh.land_bnd)


def test_river_flux_raises_the_water(basin_river_golden):
    """Mass balance sign: with the rivers flowing in, the basin's mean
    elevation after the hour is above the same basin's without them."""
    d = basin_river_golden
    river = d["s1800_eta2"]
    base = meshgen.basin()
    tb = _small_case_tables(base)
    mb = oracle_lib.OracleModel(tb, base.run, oracle_lib.SEQUENTIAL)
    for it in range(1, 1801):
        mb.step(it)
    dv = float(np.sum((river - mb.array("eta2")) * d["totalarea0"]))
    assert dv > 0.0


def test_tidal_potential_closed_form(basin_tide_golden):
    """TIP2 against the equilibrium tide written out independently: for each
    constituent the species factor from the latitude (long-period
    1.5 cos^2 - 1, diurnal sin 2phi, semidiurnal cos^2) times
    ramp ETRF TPK FFT cos(omega t + FACET + n lambda), the species n the
    frequency's nearest multiple of one cycle per day."""
    from adcirc_amd.prep import lonlat_cpp
    d = basin_tide_golden
    run = meshgen.basin_tide().run
    m = oracle_lib.OracleModel(d, run, oracle_lib.SEQUENTIAL)
    lam, phi = lonlat_cpp(d["x"], d["y"], run.slam0, run.sfea0)
    for it in range(1, 4):
        m.step(it)
    t = 3 * run.dtdp + (run.statim - run.reftim) * 86400.0
    ramp = math.tanh((2.0 * (3 * run.dtdp) / 86400.0) / run.drampelev)
    want = np.zeros_like(lam)
    for tpk, om, etrf, fft, facet in run.tipot:
        n = 2 if om > 1.3e-4 else (1 if om > 5e-5 else 0)
        fac = {0: 1.5 * np.cos(phi) ** 2 - 1.0, 1: np.sin(2.0 * phi), 2: np.cos(phi) ** 2}[n]
        want += ramp * etrf * tpk * fft * fac * np.cos(om * t + math.radians(facet) + n * lam)
    got = m.array("tip2")
    assert np.abs(got - want).max() <= 1e-15 * max(1.0, np.abs(want).max() * 1e3)
    assert np.abs(want).max() > 0.0
