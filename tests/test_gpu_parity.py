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
