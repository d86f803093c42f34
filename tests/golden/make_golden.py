"""Writes the golden fixtures under tests/golden/ from the reference build.

Runs oracle/_ref/adchp_ref (the reference's own Fortran hot-path modules
driven by oracle/ref/harness.F90) and oracle/_ref/jcg_ref (the reference's
JCG) on seeded synthetic cases and stores inputs and outputs as compressed
numpy archives.  Needs the reference build (``make -C oracle ref``), so it runs
in the build container only; the fixtures it writes are what the tests use
everywhere else.

    python tests/golden/make_golden.py [basin] [beach] [beach_lumped] [basin_linear] [jcg] [tau0]
                                       [coastal20k] [basin_tide] [basin_river] [beach_hybrid]
                                       [basin_radiation] [basin_river52]
"""
from __future__ import annotations

import hashlib
import os
import sys
import tempfile

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
REPO = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, REPO)
sys.path.insert(0, os.path.dirname(HERE))

from adcirc_amd import meshgen  # noqa: E402
import refdump  # noqa: E402

SETUP_KEYS = ("nm", "x", "y", "dp", "fdxe", "fdye", "areas", "nneigh", "neitab", "csr", "nbd",
              "nbv", "lbcodei", "csii", "siii", "me2gw", "emo", "efa", "tau0var", "evm", "fric",
              "nodecode0", "noff0", "eta2_0", "totalarea0", "mju0")
SNAP_KEYS = ("eta1", "eta2", "etas", "uu2", "vv2", "tk", "etamax", "nodecode", "noff", "totalarea")
# the long storm keeps the state the reference's bar is stated on
LONG_SNAP_KEYS = ("eta2", "uu2", "vv2", "etamax", "nodecode", "noff")


def run_case(name: str, case, nsteps: int, early: int, dump_every: int,
             snap_keys=SNAP_KEYS, early_snaps=(1, 2), slim_setup=False) -> str:
    out = {}
    with tempfile.TemporaryDirectory() as td:
        # early steps one by one (snapshots 1..early), then the long run
        refdump.write_inputs(case, os.path.join(td, "a"), nsteps=early, dump_every=1)
        a = refdump.read_output(refdump.run_reference(os.path.join(td, "a")))
        refdump.write_inputs(case, os.path.join(td, "b"), nsteps=nsteps, dump_every=dump_every)
        b = refdump.read_output(refdump.run_reference(os.path.join(td, "b"), timeout=7200))
    for k in ("np", "ne", "mnei", "neta", "nvel", "nvelme", "neimax", "prbckgrnd_mh2o"):
        out[k] = np.array(b[k])
    for k in SETUP_KEYS:
        if slim_setup:
            # the tables are regenerated from the seeded mesh; a digest pins them
            out[f"sha_{k}"] = np.array(hashlib.sha256(
                np.ascontiguousarray(b[k]).tobytes()).hexdigest())
        else:
            out[k] = b[k]
    out["step_numitr"] = b["step_numitr"].astype(np.int16)
    out["step_ncchange"] = np.minimum(b["step_ncchange"], 32767).astype(np.int16)
    out["step_hash"] = b["step_hash"]
    snaps = dict(a["snaps"])
    snaps.update(b["snaps"])
    keep = sorted(set(early_snaps) | ({early} if early_snaps == (1, 2) else set())
                  | set(b["snaps"].keys()))
    out["snap_steps"] = np.array(keep, np.int32)
    for s in keep:
        for k in snap_keys:
            out[f"s{s}_{k}"] = snaps[s][k]
    if snap_keys is SNAP_KEYS:
        out["s1_coef"] = snaps[1]["coef"]
    if case.qnam is not None:
        # the flux forcing per flux-boundary node
        out["qnam"], out["qnph"] = case.qnam, case.qnph
    if case.qnam is not None or any(ib == 30 for ib, _ in case.mesh.land_bnd):
        # the reference's boundary edge lengths
        out["bndlen2o3"] = b["bndlen2o3"]
    out["run_nsteps"] = np.array(nsteps)
    path = os.path.join(HERE, f"{name}.npz")
    np.savez_compressed(path, **out)
    return path


def _write_system(f, n, maxnz, itmax, zeta, jcoef, coef, rhs, u):
    f.write(np.array([n, maxnz, itmax], np.int64).tobytes())
    f.write(np.array([zeta], np.float64).tobytes())
    f.write(np.asarray(jcoef, np.int32).T.copy().tobytes())   # column-major
    f.write(np.asarray(coef, np.float64).T.copy().tobytes())
    f.write(np.asarray(rhs, np.float64).tobytes())
    f.write(np.asarray(u, np.float64).tobytes())


def jcg_systems():
    """A sequence of ELLPACK systems: the basin's GWCE matrix at step 1 with
    seeded right-hand sides (solved back to back, so ITPACK's saved state and
    the in-place rescaling drift carry between them), a 2-D Laplacian-plus-mass
    system with padding, and a diagonal system."""
    rng = np.random.default_rng(11)
    sys_ = []
    b = np.load(os.path.join(HERE, "basin.npz"))
    n, mnei = int(b["np"]), int(b["mnei"])
    coef = b["s1_coef"].reshape(mnei, n).T.copy()
    jco = b["neitab"].reshape(mnei, n).T.copy()
    for k in range(4):
        rhs = rng.normal(0.0, 1.0, n) * coef[:, 0]
        u = rng.normal(0.0, 0.1, n)
        sys_.append(("basin%d" % k, n, mnei, 100, 1.0e-8, jco, coef, rhs, u))
    # 2-D Laplacian + mass on a 37 x 29 grid, 5-point, two padding columns
    nx, ny = 37, 29
    nn = nx * ny
    J = np.zeros((nn, 7), np.int32)
    Cm = np.zeros((nn, 7))
    for i in range(nx):
        for j in range(ny):
            r = i * ny + j
            J[r, 0] = r + 1
            Cm[r, 0] = 4.0 + 0.05
            c = 1
            for di, dj in ((1, 0), (0, 1), (-1, 0), (0, -1)):
                ii, jj = i + di, j + dj
                if 0 <= ii < nx and 0 <= jj < ny:
                    J[r, c] = ii * ny + jj + 1
                    Cm[r, c] = -1.0
                    c += 1
    for k in range(2):
        rhs = rng.normal(0.0, 1.0, nn)
        u = np.zeros(nn) if k == 0 else rng.normal(0.0, 0.5, nn)
        sys_.append(("lap%d" % k, nn, 7, 100, 1.0e-10, J, Cm, rhs, u))
    # diagonal: converges at iteration 1
    nd = 50
    Jd = np.zeros((nd, 3), np.int32)
    Jd[:, 0] = np.arange(1, nd + 1)
    Cd = np.zeros((nd, 3))
    Cd[:, 0] = rng.uniform(1.0, 3.0, nd)
    sys_.append(("diag", nd, 3, 100, 1.0e-8, Jd, Cd, rng.normal(size=nd), np.zeros(nd)))
    return sys_


def make_jcg() -> str:
    systems = jcg_systems()
    out = {}
    with tempfile.TemporaryDirectory() as td:
        fin, fout = os.path.join(td, "in.bin"), os.path.join(td, "out.bin")
        with open(fin, "wb") as f:
            for (_, n, mz, itmax, zeta, J, Cm, rhs, u) in systems:
                _write_system(f, n, mz, itmax, zeta, J, Cm, rhs, u)
            f.write(np.array([0], np.int64).tobytes())
        import subprocess
        subprocess.run([os.path.join(REPO, "oracle", "_ref", "jcg_ref"), fin, fout], check=True)
        buf = open(fout, "rb").read()
    p = 0
    names = []
    for (name, n, mz, itmax, zeta, J, Cm, rhs, u) in systems:
        names.append(name)
        hdr = np.frombuffer(buf, np.int64, 3, p); p += 24
        rp = np.frombuffer(buf, np.float64, 5, p); p += 40
        uo = np.frombuffer(buf, np.float64, n, p); p += 8 * n
        ro = np.frombuffer(buf, np.float64, n, p); p += 8 * n
        co = np.frombuffer(buf, np.float64, n * mz, p).reshape(mz, n).T; p += 8 * n * mz
        jo = np.frombuffer(buf, np.int32, n * mz, p).reshape(mz, n).T; p += 4 * n * mz
        out[f"{name}_in_jcoef"] = J.astype(np.int32)
        out[f"{name}_in_coef"] = Cm
        out[f"{name}_in_rhs"] = rhs
        out[f"{name}_in_u"] = u
        out[f"{name}_prm"] = np.array([itmax, zeta])
        out[f"{name}_ier"] = np.array(hdr[0])
        out[f"{name}_iparm1"] = np.array(hdr[1])
        out[f"{name}_iparm8"] = np.array(hdr[2])
        out[f"{name}_rparm"] = rp.copy()
        out[f"{name}_u"] = uo.copy()
        out[f"{name}_rhs"] = ro.copy()
        out[f"{name}_coef"] = co.copy()
        out[f"{name}_jcoef"] = jo.copy()
    out["names"] = np.array(names)
    path = os.path.join(HERE, "jcg.npz")
    np.savez_compressed(path, **out)
    return path


def make_tau0() -> str:
    """TAU0VAR as the reference's InitNodalAttr sets it from a negative
    fort.15 TAU0 (schemes -1 and -2), on a small coastal mesh whose depths
    span 4000 m to -3 m, with three nodes placed exactly on the schemes'
    depth thresholds (1, 10 and 200 m)."""
    out = {}
    for scheme in (-1.0, -2.0):
        case = meshgen.coastal(3000, storm=False)
        case.mesh.dp[:3] = (1.0, 10.0, 200.0)
        case.run.tau0 = scheme
        with tempfile.TemporaryDirectory() as td:
            refdump.write_inputs(case, td, nsteps=1, dump_every=1)
            b = refdump.read_output(refdump.run_reference(td))
        tag = f"m{int(-scheme)}"
        out[f"{tag}_dp"] = b["dp"]
        out[f"{tag}_tau0var"] = b["tau0var"]
    path = os.path.join(HERE, "tau0.npz")
    np.savez_compressed(path, **out)
    return path


def beach_hybrid():
    case = meshgen.beach()
    r = case.run
    r.nolibf, r.cf, r.hbreak, r.ftheta, r.fgamma = 2, 0.0025, 2.0, 10.0, 1.0 / 3.0
    return case


def main(argv):
    todo = argv or ["basin", "beach", "beach_lumped", "basin_linear", "jcg", "tau0", "coastal20k"]
    if "basin" in todo:
        print(run_case("basin", meshgen.basin(), nsteps=43200, early=10, dump_every=43200))
    if "beach" in todo:
        print(run_case("beach", meshgen.beach(), nsteps=3600, early=10, dump_every=600))
    if "beach_lumped" in todo:
        # the same beach with the lumped (explicit) GWCE, ILump = 1
        case = meshgen.beach()
        case.run.ilump = 1
        print(run_case("beach_lumped", case, nsteps=3600, early=10, dump_every=600))
    if "basin_linear" in todo:
        # the basin with linear friction (NOLIBF = 0, TAU = 1e-4 /s) and the
        # finite-amplitude and advection terms off (NOLIFA = NOLICA = NOLICAT = 0)
        case = meshgen.basin()
        r = case.run
        r.nolibf, r.cf, r.nolifa, r.nolica, r.nolicat = 0, 1.0e-4, 0, 0, 0
        print(run_case("basin_linear", case, nsteps=3600, early=10, dump_every=1200))
    if "basin_tide" in todo:
        # the basin with the tidal potential of M2, S2, K1, O1 (NTIP = 1) at
        # 75 W, 35 N; one simulated hour
        print(run_case("basin_tide", meshgen.basin_tide(), nsteps=1800, early=10,
                       dump_every=600))
    if "beach_hybrid" in todo:
        # the wet/dry storm beach with hybrid friction (NOLIBF = 2: CFmin,
        # HBREAK = 2 m, FTHETA = 10, FGAMMA = 1/3), half an hour
        print(run_case("beach_hybrid", beach_hybrid(), nsteps=1800, early=10,
                       dump_every=600))
    if "basin_radiation" in todo:
        # the basin with 30 km of the north edge radiating (IBTYPE 30); one
        # simulated hour
        print(run_case("basin_radiation", meshgen.basin_radiation(), nsteps=1800, early=10,
                       dump_every=600))
    if "basin_river52" in todo:
        # the same rivers with the west one as river radiation (IBTYPE 52)
        print(run_case("basin_river52", meshgen.basin_river(west_ibtype=52), nsteps=1800, early=10,
                       dump_every=600))
    if "basin_river" in todo:
        # the basin with two rivers (IBTYPE 2 and 22), steady and M2 flux;
        # one simulated hour
        print(run_case("basin_river", meshgen.basin_river(), nsteps=1800, early=10,
                       dump_every=600))
    if "jcg" in todo:
        print(make_jcg())
    if "coastal20k" in todo:
        # the coastal storm (wet/dry, M2 + K1, Holland vortex, dt = 1 s) at
        # 20k nodes over its whole six simulated hours: per-step hash and
        # NUMITR, hourly snapshots
        print(run_case("coastal20k", meshgen.coastal(20000), nsteps=21600, early=10,
                       dump_every=3600, snap_keys=LONG_SNAP_KEYS, early_snaps=(1,),
                       slim_setup=True))
    if "tau0" in todo:
        print(make_tau0())


if __name__ == "__main__":
    main(sys.argv[1:])
