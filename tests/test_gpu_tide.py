"""Tidal potential forcing (NTIP = 1) on the GPU: the basin with M2, S2, K1
and O1 (`meshgen.basin_tide`), whose oracle run is pinned to the reference's
own gwce/momentum modules by tests/test_oracle.py.

The device forms TIP2 with its own cos, which can differ from the C
library's in the last bit; with the host's TIP2 handed in the step is
bit-exact, with the device's it stays within a few ulp of it."""
from __future__ import annotations

import numpy as np
import pytest

from adcirc_amd import meshgen
from gpu_helpers import build_pair, compare, log

pytestmark = pytest.mark.gpu


def test_tide_potential_bitexact_with_host_values(basin_tide_golden):
    """600 steps with the oracle's TIP2 (the C library's cos) written over the
    device's after each adcgpu_begin_step: bit-exact vs oracle."""
    d, run = basin_tide_golden, meshgen.basin_tide().run
    log("case basin_tide: 600 steps, host TIP2, bit-exact vs oracle")
    g, o = build_pair(d, run)
    for it in range(1, 601):
        o.step(it)
        g.timestep(it, None, tip2=o.array("tip2"))
        if it <= 3 or it % 50 == 0:
            bad = compare(g, o, it)
            assert not bad, f"step {it}: " + "; ".join(bad)
    assert np.array_equal(g.get("tip1"), o.array("tip1"))
    assert np.abs(o.array("tip2")).max() > 1e-4  # the ramp is still rising


def test_tide_potential_device_values_within_ulps(basin_tide_golden):
    """The device's TIP2 differs from the host's by at most 8 ulp of the
    largest sum its terms can reach (RampTip * sum of ETRF*TPK*FFT*1.5: both
    cos are accurate to about an ulp, the sum has four terms), and after
    1,800 steps (one hour, the ramp rising) the elevations are within 1e-9 m
    of the oracle's."""
    import math

    d, run = basin_tide_golden, meshgen.basin_tide().run
    log("case basin_tide: 1800 steps, device TIP2 vs oracle")
    g, o = build_pair(d, run)
    amp0 = 1.5 * sum(abs(etrf * tpk * fft) for tpk, _, etrf, fft, _ in run.tipot)
    worst = 0.0
    for it in range(1, 1801):
        o.step(it)
        g.timestep(it)
        if it % 100 == 0:
            ramp = math.tanh((2.0 * it * run.dtdp / 86400.0) / run.drampelev)
            amp = ramp * amp0
            a, b = g.get("tip2"), o.array("tip2")
            dmax = float(np.abs(a - b).max())
            worst = max(worst, dmax / amp)
            assert dmax <= 8 * np.finfo(np.float64).eps * amp, \
                f"step {it}: TIP2 differs by {dmax:.3e} ({dmax / amp:.2e} of {amp:.3e})"
    de = np.abs(g.get("eta2") - o.array("eta2")).max()
    log(f"  TIP2 max diff {worst:.2e} of the amplitude bound, final max |d eta2| {de:.2e} m")
    assert de <= 1e-9


def test_tide_potential_resident_run_equals_per_step(basin_tide_golden):
    """adcgpu_run with the tidal potential equals the per-step calls (both
    with the device's TIP2) bit for bit after 300 steps."""
    d, run = basin_tide_golden, meshgen.basin_tide().run
    log("case basin_tide: resident run vs per-step calls, 300 steps")
    g1, _ = build_pair(d, run)
    for it in range(1, 301):
        g1.timestep(it)
    g2, _ = build_pair(d, run)
    g2.run_resident(1, 300)
    for k in ("eta2", "uu2", "vv2", "etas", "tip1", "tip2"):
        a, b = g1.get(k), g2.get(k)
        assert np.array_equal(a, b), f"{k}: {int((a != b).sum())} differ"
