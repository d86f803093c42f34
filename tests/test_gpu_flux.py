"""Specified normal flux boundaries on the GPU: the basin with two rivers
(`meshgen.basin_river`: IBTYPE 2 and 22, a steady and an M2 discharge),
whose oracle run is pinned to the reference's own gwce/momentum modules by
tests/test_oracle.py.  The host forms QN2 (adcgpu_flux_bc, the C library's
cos), so the GPU run is bit-exact against the oracle in the GPU's summation
order."""
from __future__ import annotations

import numpy as np
import pytest

from adcirc_amd import meshgen
from gpu_helpers import build_pair, compare, log, step_pair

pytestmark = pytest.mark.gpu


def test_rivers_bitexact_vs_oracle_and_reference(basin_river_golden):
    """1,800 steps (one hour) through the per-step calls: bit-exact against
    the oracle (blocked summation order) at every checked step, and at the
    fixture's snapshots within 1e-12 of the reference's own state (its
    sequential order)."""
    d, run = basin_river_golden, meshgen.basin_river().run
    log("case basin_river: 1800 steps, per-step calls, bit-exact vs oracle")
    g, o = build_pair(d, run)
    snaps = set(int(s) for s in d["snap_steps"])
    for it in range(1, 1801):
        step_pair(g, o, it)
        if it <= 3 or it % 50 == 0:
            bad = compare(g, o, it)
            assert not bad, f"step {it}: " + "; ".join(bad)
        if it in snaps:
            # the reference sums in sequential order, the GPU (and the
            # oracle beside it) in the blocked tree: rounding-level apart
            for k in ("eta2", "uu2", "vv2"):
                dk = np.abs(g.get(k) - d[f"s{it}_{k}"]).max()
                assert dk <= 1e-12, f"step {it}: {k} differs from the reference by {dk:.3e}"
    assert np.abs(o.flux["qn2"]).max() > 0.2


def test_rivers_resident_run_equals_per_step(basin_river_golden):
    """adcgpu_run with the flux table equals the per-step calls bit for bit
    after 400 steps, run as 150 + 250 (the shift carries QN0/QN1 across
    runs)."""
    d, run = basin_river_golden, meshgen.basin_river().run
    log("case basin_river: resident run (150 + 250 steps) vs per-step calls")
    g1, _ = build_pair(d, run)
    for it in range(1, 401):
        g1.timestep(it)
    g2, _ = build_pair(d, run)
    g2.run_resident(1, 150)
    g2.run_resident(151, 250)
    for k in ("eta2", "uu2", "vv2", "etas", "tk"):
        a, b = g1.get(k), g2.get(k)
        assert np.array_equal(a, b), f"{k}: {int((a != b).sum())} differ"


def test_rivers_snapshot_restart_and_batch(basin_river_golden):
    """A snapshot taken mid-run carries the flux time levels: a new context
    loaded from it continues bit for bit; and two river contexts stepped as a
    batch equal their own resident runs."""
    from adcirc_amd.model import Batch

    d, run = basin_river_golden, meshgen.basin_river().run
    log("case basin_river: snapshot at step 200, batch of two, 200 more steps")
    a, _ = build_pair(d, run)
    a.run_resident(1, 200)
    snap = a.save_state()
    b, _ = build_pair(d, run)
    b.load_state(snap)
    c, _ = build_pair(d, run)
    c.load_state(snap)
    a.run_resident(201, 200)
    bt = Batch([b, c])
    bt.run(201, 200)
    bt.close()
    for k in ("eta2", "uu2", "vv2", "etas"):
        ref = a.get(k)
        assert np.array_equal(b.get(k), ref), k
        assert np.array_equal(c.get(k), ref), k


def test_radiation_bitexact_and_resident(basin_radiation_golden):
    """The basin with an outward-radiating stretch of its north edge
    (IBTYPE 30; QN1 from the last velocities and the radiation QFORCE formed
    at each step's start): 1,800 per-step calls bit-exact against the oracle,
    within 1e-12 of the reference at its snapshots, and a resident run (no
    flux rows needed) equal to the per-step calls."""
    d = basin_radiation_golden
    run = meshgen.basin_radiation().run
    log("case basin_radiation: 1800 steps bit-exact vs oracle; resident 600")
    g, o = build_pair(d, run)
    snaps = set(int(s) for s in d["snap_steps"])
    for it in range(1, 1801):
        step_pair(g, o, it)
        if it <= 3 or it % 50 == 0:
            bad = compare(g, o, it)
            assert not bad, f"step {it}: " + "; ".join(bad)
        if it in snaps:
            for k in ("eta2", "uu2", "vv2"):
                dk = np.abs(g.get(k) - d[f"s{it}_{k}"]).max()
                assert dk <= 1e-12, f"step {it}: {k} differs from the reference by {dk:.3e}"
        if it == 600:
            e600 = {k: g.get(k) for k in ("eta2", "uu2", "vv2", "etas")}
    r, _ = build_pair(d, run)
    r.run_resident(1, 600)
    for k, v in e600.items():
        assert np.array_equal(r.get(k), v), k


def test_river_radiation_bitexact(basin_river52_golden):
    """The rivers with the west one as river radiation (IBTYPE 52): 1,800
    per-step calls bit-exact against the oracle, within 1e-12 of the
    reference at its snapshots; a resident run of 600 steps equal to them."""
    d = basin_river52_golden
    run = meshgen.basin_river(west_ibtype=52).run
    log("case basin_river52: 1800 steps bit-exact vs oracle; resident 600")
    g, o = build_pair(d, run)
    snaps = set(int(s) for s in d["snap_steps"])
    for it in range(1, 1801):
        step_pair(g, o, it)
        if it <= 3 or it % 50 == 0:
            bad = compare(g, o, it)
            assert not bad, f"step {it}: " + "; ".join(bad)
        if it in snaps:
            for k in ("eta2", "uu2", "vv2"):
                dk = np.abs(g.get(k) - d[f"s{it}_{k}"]).max()
                assert dk <= 1e-12, f"step {it}: {k} differs from the reference by {dk:.3e}"
        if it == 600:
            e600 = {k: g.get(k) for k in ("eta2", "uu2", "vv2", "etas")}
    r, _ = build_pair(d, run)
    r.run_resident(1, 600)
    for k, v in e600.items():
        assert np.array_equal(r.get(k), v), k
