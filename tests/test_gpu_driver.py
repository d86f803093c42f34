"""The case-directory driver on the GPU: fort.14 + fort.15 in, fort.63 /
fort.64 / maxele.63 out.  The records must hold exactly the state that the
per-step C ABI produces at the same steps (the driver runs the steps between
outputs as resident launch sequences), with dry nodes written as -99999."""
import os

import numpy as np
import pytest

from gpu_helpers import log
from adcirc_amd import driver, fort14, fort15, meshgen, outputs, prep
from adcirc_amd.model import Model, cold_wetdry_state

pytestmark = pytest.mark.gpu


def _write_case(d, case, nsteps, elev, vel, sta=None, stations=None):
    case.run.nsteps = nsteps
    os.makedirs(d, exist_ok=True)
    fort14.write_fort14(os.path.join(d, "fort.14"), case.mesh)
    f = fort15.from_case(case, elev, vel)
    if sta is not None:
        f.elev_sta, f.stations_e = sta, stations
        f.vel_sta, f.stations_v = sta, stations
    fort15.write_fort15(os.path.join(d, "fort.15"), f)


def _per_step_model(d, holland):
    """The same case stepped one ABI call per step (adcgpu_timestep), set up
    from the files the driver read."""
    mesh = fort14.read_fort14(os.path.join(d, "fort.14"))
    neta = sum(s.shape[0] for s in mesh.open_bnd)
    f = fort15.read_fort15(os.path.join(d, "fort.15"), neta, driver.flux_entries(mesh))
    run = f.run
    run.holland = holland
    t = prep.build_tables(mesh, run.nolifa, run.h0, run.anginn)
    eta2, nc, noff = cold_wetdry_state(t, run)
    g = Model(t, run, emo=f.emo, efa=Model.radians(f.efa), qnam=f.qnam, qnph=f.qnph)
    g.cold_start(eta2, nc, noff)
    if holland is not None:
        g.holland(0.0, holland)
    return g, run


def _expect(vals, nodecode):
    return np.array([float(outputs.e20(v)) for v in np.where(nodecode == 0, -99999.0, vals)])


@pytest.mark.parametrize("name", ["basin", "beach_storm", "basin_river", "basin_radiation"])
def test_driver_outputs_equal_per_step_state(tmp_path, name):
    d = str(tmp_path / name)
    if name in ("basin", "basin_river", "basin_radiation"):
        case = {"basin": meshgen.basin, "basin_river": meshgen.basin_river,
                "basin_radiation": meshgen.basin_radiation}[name]()
        n, hol = 600, None
        elev, vel = fort15.Output(1, 0.0, 1.0, 200), fort15.Output(1, 0.0, 1.0, 300)
    else:
        case, n = meshgen.beach(), 900
        hol = case.run.holland
        elev, vel = fort15.Output(1, 0.0, 1.0, 150), fort15.Output(1, 0.0, 1.0, 450)
    # stations: inside the mesh, one near the shore (wets and dries on the
    # beach), one outside
    x, y = case.mesh.x, case.mesh.y
    stations = np.array([[0.5 * (x.min() + x.max()), 0.5 * (y.min() + y.max())],
                         [x.min() + 0.1 * (x.max() - x.min()), 0.37 * (y.max() - y.min())],
                         [x.max() + 1.0e4, y.max()]])
    sta = fort15.Output(1, 0.0, 1.0, 100)
    _write_case(d, case, n, elev, vel, sta, stations)
    log(f"case {name}: driver over {n} steps")
    res = driver.run_directory(d, holland=hol, log=log)
    assert res["steps"] == n
    _, e63 = outputs.read_global(os.path.join(d, "fort.63"))
    h64, e64 = outputs.read_global(os.path.join(d, "fort.64"))
    hm, emx = outputs.read_global(os.path.join(d, "maxele.63"))
    assert [r[1] for r in e63] == list(range(elev.nspool, n + 1, elev.nspool))
    assert [r[1] for r in e64] == list(range(vel.nspool, n + 1, vel.nspool))
    assert h64["rtype"] == 2 and hm["nsets"] == 2
    _, e61 = outputs.read_global(os.path.join(d, "fort.61"))
    _, e62 = outputs.read_global(os.path.join(d, "fort.62"))
    assert [r[1] for r in e61] == list(range(100, n + 1, 100))
    want61 = {it: v for _, it, v in e61}
    want62 = {it: v for _, it, v in e62}
    tabs = prep.build_tables(fort14.read_fort14(os.path.join(d, "fort.14")), case.run.nolifa,
                             case.run.h0, case.run.anginn)
    el, fac = outputs.station_factors(tabs["x"], tabs["y"], tabs["nm"], tabs["areas"],
                                      stations[:, 0], stations[:, 1])
    assert el[-1] == -1 and (el[:-1] >= 0).all()
    g, run = _per_step_model(d, hol)
    want63 = {it: v for _, it, v in e63}
    want64 = {it: v for _, it, v in e64}
    dry_seen = 0
    for it in range(1, n + 1):
        if hol is not None:
            # the resident run evaluates the vortex on the device; so does this
            _step_device_forcing(g, it, hol)
        else:
            g.timestep(it)
        nc = g.get("nodecode")
        if it in want63:
            dry_seen += int((nc == 0).sum())
            assert np.array_equal(want63[it][:, 0], _expect(g.get("eta2"), nc)), it
        if it in want64:
            assert np.array_equal(want64[it][:, 0], _expect(g.get("uu2"), nc)), it
            assert np.array_equal(want64[it][:, 1], _expect(g.get("vv2"), nc)), it
        if it in want61:
            sv = outputs.station_values(g.get("eta2"), nc, tabs["nm"], el, fac)
            assert np.array_equal(want61[it][:, 0], [float(outputs.e20(v)) for v in sv]), it
            for c, k in enumerate(("uu2", "vv2")):
                sv = outputs.station_values(g.get(k), nc, tabs["nm"], el, fac)
                assert np.array_equal(want62[it][:, c], [float(outputs.e20(v)) for v in sv]), it
    tmax = outputs.etamax_time(g.get("etamax_it"), run.dtdp, run.statim)
    assert np.array_equal(emx[0][2][:, 0], [float(outputs.e20(v)) for v in g.get("etamax")])
    assert np.array_equal(emx[1][2][:, 0], [float(outputs.e20(v)) for v in tmax])
    _, vmx = outputs.read_global(os.path.join(d, "maxvel.63"))
    umt = outputs.etamax_time(g.get("umax_it"), run.dtdp, run.statim)
    assert np.array_equal(vmx[0][2][:, 0], [float(outputs.e20(v)) for v in g.get("umax")])
    assert np.array_equal(vmx[1][2][:, 0], [float(outputs.e20(v)) for v in umt])
    if name == "beach_storm":
        assert dry_seen > 0
    log(f"  records identical; dry node-records {dry_seen}")
    g.close()


def _step_device_forcing(g, it, hol):
    """One step through the per-call ABI with the device vortex, as the
    resident run forms it (friction, then this step's forcing)."""
    import ctypes as C
    v = np.ascontiguousarray(list(hol) + [0.0] * (13 - len(hol)), np.float64)
    dp_ = C.POINTER(C.c_double)
    bc = np.ascontiguousarray(g.eta_bc(it), np.float64)
    lib, h = g.lib, g.h
    assert lib.adcgpu_begin_step(h, it) == 0
    assert lib.adcgpu_friction(h) == 0
    assert lib.adcgpu_holland(h, it * g.run.dtdp, v.ctypes.data_as(dp_)) == 0
    assert lib.adcgpu_solve_gwce(h, it, bc.ctypes.data_as(dp_)) == 0
    assert lib.adcgpu_wetdry(h, it) == 0
    assert lib.adcgpu_momentum(h) == 0
