"""GPU parity: the HIP path against the oracle on the same inputs.

The oracle runs in its blocked reduction order, which is the GPU's fixed
reduction tree, so the comparison is bit-for-bit on every state array every
step (wet/dry flags included).  Against the reference build itself (golden
fixtures) the bar is the reference's: identical wet/dry flags and
|d eta| <= 1e-6 m, |d u| <= 1e-6 m/s.
"""
import numpy as np
import pytest

from gpu_helpers import build_pair, compare, log, step_pair
from adcirc_amd import meshgen

pytestmark = pytest.mark.gpu


def _run_bitexact(d, run, nsteps, check_every=1):
    g, o = build_pair(d, run)
    for it in range(1, nsteps + 1):
        step_pair(g, o, it)
        if it % check_every == 0 or it <= 3 or it == nsteps:
            bad = compare(g, o, it)
            assert not bad, f"step {it}: " + "; ".join(bad)
    return g, o


def test_beach_first_steps_bitexact(beach_golden):
    log("case beach: 20 steps, per-step bit-exact vs oracle")
    _run_bitexact(beach_golden, meshgen.beach().run, 20)


def test_basin_first_steps_bitexact(basin_golden):
    log("case basin: 20 steps, per-step bit-exact vs oracle")
    _run_bitexact(basin_golden, meshgen.basin().run, 20)


def test_beach_hour_bitexact_and_reference(beach_golden):
    d = beach_golden
    log("case beach: 3600 steps (1 h, wet/dry + storm), bit-exact vs oracle")
    g, o = _run_bitexact(d, meshgen.beach().run, 3600, check_every=50)
    s = 3600
    assert np.array_equal(g.get("nodecode"), d[f"s{s}_nodecode"])
    assert np.array_equal(g.get("noff"), d[f"s{s}_noff"])
    assert np.abs(g.get("eta2") - d[f"s{s}_eta2"]).max() <= 1e-6
    assert np.abs(g.get("uu2") - d[f"s{s}_uu2"]).max() <= 1e-6
    assert np.abs(g.get("vv2") - d[f"s{s}_vv2"]).max() <= 1e-6


def test_basin_linear_terms_bitexact_and_reference(basin_linear_golden):
    """Linear friction (NOLIBF = 0), NOLIFA = NOLICA = NOLICAT = 0 on the
    basin, 3600 steps: bit-exact vs oracle every 50 steps, and within the
    reference's bar of the reference build's run."""
    from test_oracle import basin_linear_run
    d = basin_linear_golden
    log("case basin linear: 3600 steps, bit-exact vs oracle")
    g, o = _run_bitexact(d, basin_linear_run(), 3600, check_every=50)
    s = 3600
    assert np.array_equal(g.get("nodecode"), d[f"s{s}_nodecode"])
    assert np.abs(g.get("eta2") - d[f"s{s}_eta2"]).max() <= 1e-6
    assert np.abs(g.get("uu2") - d[f"s{s}_uu2"]).max() <= 1e-6
    assert np.abs(g.get("vv2") - d[f"s{s}_vv2"]).max() <= 1e-6


def test_coastal_mesh_bitexact():
    """The product's own preprocessing on a generated 20k-node coastal mesh
    (Delaunay, wet/dry, M2+K1, Holland storm): bit-exact vs oracle."""
    from gpu_helpers import case_setup
    case = meshgen.coastal(20000)
    d = case_setup(case)
    log(f"case coastal {int(d['np'])} nodes: 120 steps bit-exact vs oracle")
    _run_bitexact(d, case.run, 120, check_every=10)


def test_coastal_large_mesh_bitexact():
    """A 400k-node coastal mesh (about 200 JCG chunks, so every XCD range
    holds several blocks and the wetting front carries many candidates):
    bit-exact vs oracle over the first steps."""
    from gpu_helpers import case_setup
    case = meshgen.coastal(400000)
    d = case_setup(case)
    log(f"case coastal {int(d['np'])} nodes: 12 steps bit-exact vs oracle")
    _run_bitexact(d, case.run, 12, check_every=4)


def test_coastal_bench_size_bitexact():
    """The 1M-node coastal mesh the benchmark times: 245 JCG blocks, near the
    256 that are resident at once, with the pseudo-residual staged in LDS
    (bandwidth 3,873 rows): bit-exact vs oracle over the first steps."""
    from gpu_helpers import case_setup
    case = meshgen.coastal(1_000_000)
    d = case_setup(case)
    log(f"case coastal {int(d['np'])} nodes: 8 steps bit-exact vs oracle")
    g, _ = _run_bitexact(d, case.run, 8, check_every=4)
    assert g.jcg_shape()[0] == 245


def test_wetting_many_candidates_shuffled_elements():
    """A step with tens of thousands of wetting candidates whose friction
    updates chain through shared nodes, on a mesh with shuffled element
    numbering, so that a candidate's predecessor on a node usually lies
    thousands of candidates back: bit-exact vs oracle.  About a third of the
    nodes start dry and 0.5 m below their wet neighbours, so nearly every
    element with two wet nodes passes the wetting criterion."""
    from gpu_helpers import case_setup
    case = meshgen.coastal(40000)
    rng = np.random.default_rng(11)
    case.mesh.nm = case.mesh.nm[rng.permutation(case.mesh.ne)]
    d = case_setup(case)
    nc = np.asarray(d["nodecode0"], np.int32).copy()
    eta2 = np.asarray(d["eta2_0"], np.float64).copy()
    dry = (rng.random(nc.shape[0]) < 1.0 / 3.0) & (nc == 1)
    nc[dry] = 0
    eta2[dry] -= 0.5
    d["nodecode0"], d["eta2_0"] = nc, eta2
    nm0 = np.asarray(d["nm"], np.int64).reshape(-1, 3) - 1
    two_wet = int((nc[nm0].sum(axis=1) == 2).sum())
    log(f"case coastal {int(d['np'])} nodes, shuffled elements, {dry.sum()} nodes dried, "
        f"{two_wet} elements with two wet nodes: 4 steps bit-exact vs oracle")
    assert two_wet > 2 * 8192
    _run_bitexact(d, case.run, 4)


def test_holland_forcing_vs_oracle():
    """The device Holland vortex (the resident run's met forcing) against the
    oracle's host evaluation of the same formula at four times of the storm:
    they differ only by the device and host libm (relative 1e-13)."""
    from gpu_helpers import case_setup
    case = meshgen.coastal(20000)
    d = case_setup(case)
    g, o = build_pair(d, case.run)
    for t in (0.0, 3600.0, 10800.0, 21600.0):
        g.holland(t, case.run.holland)
        ref = o.forcing(t)
        for name, r in zip(("wsx2", "wsy2", "pr2"), ref):
            a = g.get(name)
            # the wind ramp is zero at t = 0: both sides exactly zero there
            err = np.abs(a - r).max() / max(np.abs(r).max(), 1e-300)
            log(f"holland t={t:.0f} s {name}: max relative difference {err:.2e}")
            assert err <= 1e-13, (t, name)


def _resident_vs_per_call(case, n, precision=None):
    """Run ``n`` steps through the resident run and the same steps one ABI
    call at a time; return the resident model's COEF rebuild count after
    checking every state array, COEF and ALPHAL identical."""
    import ctypes as C
    from gpu_helpers import case_setup, golden_tables
    from adcirc_amd import _lib
    from adcirc_amd.model import Model
    d = case_setup(case)
    run = case.run
    t = golden_tables(d)
    nb = run.nbfr
    prec = _lib.FP64 if precision is None else precision

    def model():
        g = Model(t, run, tau0var=d["tau0var"], evm=d["evm"], fric=d["fric"],
                  corif=np.full(t["np"], run.cori),
                  emo=np.asarray(d["emo"]).reshape(t["neta"], nb),
                  efa=np.asarray(d["efa"]).reshape(t["neta"], nb), precision=prec)
        g.cold_start(d["eta2_0"], d["nodecode0"], d["noff0"])
        g.holland(0.0, run.holland)
        return g

    log(f"case {t['np']} nodes, ilump {run.ilump}: {n} steps resident vs per call")
    g1, g2 = model(), model()
    g1.run_resident(1, n, holland=run.holland)
    hol = np.ascontiguousarray(list(run.holland) + [0.0] * (13 - len(run.holland)), np.float64)
    dp_ = C.POINTER(C.c_double)
    lib, h = g2.lib, g2.h
    for it in range(1, n + 1):
        bc = np.ascontiguousarray(g2.eta_bc(it), np.float64)
        assert lib.adcgpu_begin_step(h, it) == 0
        assert lib.adcgpu_friction(h) == 0
        assert lib.adcgpu_holland(h, it * run.dtdp, hol.ctypes.data_as(dp_)) == 0
        assert lib.adcgpu_solve_gwce(h, it, bc.ctypes.data_as(dp_)) == 0
        assert lib.adcgpu_wetdry(h, it) == 0
        assert lib.adcgpu_momentum(h) == 0
    log(f"  COEF rebuilt on {g1.last_rebuilds} of {n} steps")
    for k in ("eta1", "eta2", "etas", "uu2", "vv2", "qx2", "qy2", "tk", "etamax", "coef",
              "alphal", "nodecode", "noff"):
        a, b = g1.get(k), g2.get(k)
        assert np.array_equal(a, b), (k, int((a != b).sum()))
    return g1.last_rebuilds


def test_coastal_resident_run_equals_per_call_api():
    """The resident run (adcgpu_run: friction and device Holland forcing in one
    pass, the wet/dry ALPHAL pass inside the momentum element pass, COEF's
    off-diagonals left scaled between solves) against the same steps made one
    ABI call at a time (adcgpu_begin_step, _friction, _holland, _solve_gwce,
    _wetdry, _momentum) on the 20k-node coastal storm: every state array,
    COEF and ALPHAL identical after 300 steps of wetting and drying."""
    # steps that rebuild COEF and steps that restore it from its scaled form
    # (the basin's resident test covers runs that never rebuild)
    assert _resident_vs_per_call(meshgen.coastal(20000), 300) > 0


def test_beach_lumped_resident_run_equals_per_call_api():
    """The lumped GWCE (ILump = 1, no JCG) through the resident run against
    the per-call ABI on the wet/dry beach storm: identical after 600 steps."""
    case = meshgen.beach()
    case.run.ilump = 1
    _resident_vs_per_call(case, 600)


def test_beach_lumped_hour_bitexact_and_reference(beach_lumped_golden):
    """The lumped GWCE (ILump = 1: diagonal LHS, explicit solve) on the wet/dry
    beach storm, 3600 steps: bit-exact vs oracle, and within the reference's
    bar of the reference build's own run."""
    d = beach_lumped_golden
    run = meshgen.beach().run
    run.ilump = 1
    log("case beach lumped: 3600 steps (1 h, wet/dry + storm), bit-exact vs oracle")
    g, o = _run_bitexact(d, run, 3600, check_every=50)
    s = 3600
    assert g.numitr()[0] == 0
    assert np.array_equal(g.get("nodecode"), d[f"s{s}_nodecode"])
    assert np.array_equal(g.get("noff"), d[f"s{s}_noff"])
    assert np.abs(g.get("eta2") - d[f"s{s}_eta2"]).max() <= 1e-6
    assert np.abs(g.get("uu2") - d[f"s{s}_uu2"]).max() <= 1e-6
    assert np.abs(g.get("vv2") - d[f"s{s}_vv2"]).max() <= 1e-6


def test_basin_full_day_resident_vs_reference(basin_golden):
    """43,200 steps through the resident run (one launch sequence per step,
    elevation forcing from a precomputed table): final state within the
    reference's bar, and identical to stepping through the per-step API."""
    d = basin_golden
    run = meshgen.basin().run
    nsteps = int(d["run_nsteps"])
    log(f"case basin: resident run of {nsteps} steps")
    g, o = build_pair(d, run)
    iters = g.run_resident(1, nsteps)
    s = nsteps
    de = np.abs(g.get("eta2") - d[f"s{s}_eta2"]).max()
    du = max(np.abs(g.get("uu2") - d[f"s{s}_uu2"]).max(),
             np.abs(g.get("vv2") - d[f"s{s}_vv2"]).max())
    log(f"  max |d eta| {de:.3e} m, max |d u| {du:.3e} m/s, JCG iterations {iters} "
        f"(reference {int(d['step_numitr'].sum())})")
    assert de <= 1e-6 and du <= 1e-6
    assert np.array_equal(g.get("nodecode"), d[f"s{s}_nodecode"])
    # the first 200 steps again, one API call per step, against the resident run
    g2, _ = build_pair(d, run)
    g3, _ = build_pair(d, run)
    g2.run_resident(1, 200)
    for it in range(1, 201):
        g3.timestep(it)
    for k in ("eta2", "uu2", "vv2", "etas", "tk"):
        assert np.array_equal(g2.get(k), g3.get(k)), k


def test_jcg_dropin_sequence(jcg_golden):
    """adcgpu_jcg (the drop-in for CALL JCG) on the reference's fixture systems,
    solved back to back: bit-exact vs the oracle's blocked order; against the
    reference's own JCG the same iteration count (+-1) and solution to the
    stopping tolerance."""
    import oracle_lib
    from adcirc_amd.model import jcg
    from test_oracle import _jcg_call
    g = jcg_golden
    lib = oracle_lib.load()
    st = oracle_lib.Itpack()
    lib.adco_itpack_init(st, oracle_lib.BLOCKED)
    for name in [str(s) for s in g["names"]]:
        log(f"jcg system {name}")
        J, Cm = g[f"{name}_in_jcoef"], g[f"{name}_in_coef"]
        n, mz = J.shape
        itmax, zeta = g[f"{name}_prm"]
        ier_o, ip_o, rp_o, u_o, rhs_o, co_o, jc_o = _jcg_call(
            lib, st, n, J, Cm, g[f"{name}_in_rhs"], g[f"{name}_in_u"], int(itmax), zeta)
        jc = np.ascontiguousarray(J.astype(np.int32).T.ravel())
        co = np.ascontiguousarray(np.asarray(Cm, np.float64).T.ravel())
        rhs = np.array(g[f"{name}_in_rhs"], np.float64)
        u = np.array(g[f"{name}_in_u"], np.float64)
        iparm = np.zeros(12, np.int32)
        rparm = np.zeros(12)
        lib.adco_dfault(iparm.ctypes.data_as(oracle_lib._ip), rparm.ctypes.data_as(oracle_lib._dp))
        iparm[0], iparm[1], iparm[3] = int(itmax), 0, 16
        rparm[0] = zeta
        ier = jcg(co, jc, rhs, u, iparm, rparm)
        assert ier == ier_o
        assert iparm[0] == ip_o[0], name
        assert np.array_equal(u, u_o), name
        assert np.array_equal(rhs, rhs_o), name
        assert np.array_equal(co.reshape(mz, n).T, co_o), name
        assert np.array_equal(jc.reshape(mz, n).T, jc_o), name
        # against the reference JCG
        assert abs(int(iparm[0]) - int(g[f"{name}_iparm1"])) <= 1, name
        ur = g[f"{name}_u"]
        assert np.linalg.norm(u - ur) <= 100 * zeta * max(np.linalg.norm(ur), 1e-300), name


def test_warm_start_within_jcg_tolerance_of_plain():
    """The extrapolated JCG start (warm_start = 1) on the 20k-node coastal
    storm: the same stopping test, so the state stays within the solver's
    tolerance of the plain start's (no bit-exactness: a different initial
    guess ends at a different point inside the stopping ball)."""
    import dataclasses
    from gpu_helpers import case_setup, golden_tables
    from adcirc_amd.model import Model
    case = meshgen.coastal(20000)
    d = case_setup(case)
    t = golden_tables(d)
    nb = case.run.nbfr
    out = {}
    for ws in (0, 1):
        run = dataclasses.replace(case.run, warm_start=ws)
        g = Model(t, run, tau0var=d["tau0var"], evm=d["evm"], fric=d["fric"],
                  corif=np.full(t["np"], run.cori),
                  emo=np.asarray(d["emo"]).reshape(t["neta"], nb),
                  efa=np.asarray(d["efa"]).reshape(t["neta"], nb))
        g.cold_start(d["eta2_0"], d["nodecode0"], d["noff0"])
        g.holland(0.0, run.holland)
        it = g.run_resident(1, 600, holland=run.holland)
        out[ws] = (it, g.get("eta2"), g.get("nodecode"))
        g.close()
    de = np.abs(out[1][1] - out[0][1]).max()
    log(f"600 steps: JCG iterations plain {out[0][0]}, extrapolated {out[1][0]}; "
        f"max |d eta| {de:.3e} m, nodecode differences {int((out[0][2] != out[1][2]).sum())}")
    assert de <= 1e-5


def _renumbered(case, order):
    """The case with its nodes renumbered: new node k is old node order[k]."""
    import dataclasses
    m = case.mesh
    inv = np.empty_like(order)
    inv[order] = np.arange(order.shape[0])
    ren = lambda ids: (inv[np.asarray(ids, np.int64) - 1] + 1).astype(np.int32)  # noqa: E731
    mesh = dataclasses.replace(m, x=m.x[order], y=m.y[order], dp=m.dp[order], nm=ren(m.nm),
                               open_bnd=[ren(s) for s in m.open_bnd],
                               land_bnd=[(ib, ren(s)) for ib, s in m.land_bnd])
    return dataclasses.replace(case, mesh=mesh)


def _bandwidth(d):
    npn, mnei = int(d["np"]), int(d["mnei"])
    nei = np.asarray(d["neitab"], np.int64).reshape(mnei, npn)
    rows = np.broadcast_to(np.arange(npn), nei.shape)
    real = nei > 0
    return int(np.abs(nei[real] - 1 - rows[real]).max())


@pytest.mark.parametrize("numbering", ["random", "windowed"])
def test_renumbered_mesh_bitexact(numbering):
    """The 20k-node coastal storm with its nodes renumbered, so that the
    matrix bandwidth sends the JCG iterations down the paths that an
    RCM-ordered mesh never takes: random numbering (bandwidth above two row
    blocks of 4096: a second grid barrier after each update) and numbering
    shuffled within windows of 4096 nodes (a halo of two row blocks each
    side).  Bit-exact vs oracle."""
    from gpu_helpers import case_setup
    case = meshgen.coastal(20000)
    npn = case.mesh.x.shape[0]
    rng = np.random.default_rng(5)
    if numbering == "random":
        order = rng.permutation(npn)
    else:
        order = np.concatenate([w0 + rng.permutation(min(4096, npn - w0))
                                for w0 in range(0, npn, 4096)])
    case = _renumbered(case, order)
    d = case_setup(case)
    bw = _bandwidth(d)
    log(f"case coastal {npn} nodes, {numbering} numbering: bandwidth {bw}, 12 steps")
    if numbering == "random":
        assert bw > 2 * 4096
    else:
        assert 4096 < bw <= 2 * 4096
    _run_bitexact(d, case.run, 12, check_every=4)


def test_cold_start_again_equals_fresh_context():
    """A context that has run 60 steps of the wet/dry beach storm and is cold
    started again steps exactly like a fresh one (state, COEF, maxima, JCG
    iteration counts)."""
    from gpu_helpers import case_setup, golden_tables
    from adcirc_amd.model import Model
    case = meshgen.beach()
    d = case_setup(case)
    run = case.run
    t = golden_tables(d)
    nb = run.nbfr

    def fresh():
        g = Model(t, run, tau0var=d["tau0var"], evm=d["evm"], fric=d["fric"],
                  corif=np.full(t["np"], run.cori),
                  emo=np.asarray(d["emo"]).reshape(t["neta"], nb),
                  efa=np.asarray(d["efa"]).reshape(t["neta"], nb))
        g.cold_start(d["eta2_0"], d["nodecode0"], d["noff0"])
        g.holland(0.0, run.holland)
        return g

    log("case beach: 60 steps, cold start again, 60 steps vs a fresh context")
    g1 = fresh()
    g1.run_resident(1, 60, holland=run.holland)
    g1.cold_start(d["eta2_0"], d["nodecode0"], d["noff0"])
    g1.holland(0.0, run.holland)
    it1 = g1.run_resident(1, 60, holland=run.holland)
    g2 = fresh()
    it2 = g2.run_resident(1, 60, holland=run.holland)
    assert it1 == it2
    for k in ("eta1", "eta2", "etas", "uu2", "vv2", "qx2", "qy2", "tk", "etamax", "umax", "coef",
              "alphal", "nodecode", "noff", "etamax_it", "umax_it"):
        a, b = g1.get(k), g2.get(k)
        assert np.array_equal(a, b), (k, int((a != b).sum()))
