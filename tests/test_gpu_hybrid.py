"""Hybrid bottom friction (NOLIBF = 2) on the GPU: the wet/dry storm beach
whose oracle run is pinned bit for bit to the reference by
tests/test_oracle.py.  The friction law's pow runs on the device, whose pow
can differ from the C library's in the last bit; the GPU run therefore agrees
with the oracle to rounding rather than bit for bit."""
from __future__ import annotations

import numpy as np
import pytest

from conftest import beach_hybrid_run
from gpu_helpers import build_pair, log, step_pair

pytestmark = pytest.mark.gpu


def test_hybrid_friction_tracks_oracle(beach_hybrid_golden):
    """TK after step 2 (the first with flow) within 1e-13 relative of the
    oracle's; after 1,800 steps the wet/dry flags agree at all but two nodes
    and the elevations of nodes wet in both within 1e-8 m; the friction is the
    hybrid law's, not the quadratic one (above it where the water is shallower
    than HBREAK)."""
    d, run = beach_hybrid_golden, beach_hybrid_run()
    log("case beach_hybrid: 1800 steps, GPU vs oracle to rounding")
    g, o = build_pair(d, run)
    for it in range(1, 1801):
        step_pair(g, o, it)
        if it == 2:
            a, b = g.get("tk"), o.array("tk")
            rel = np.abs(a - b) / np.maximum(np.abs(b), 1e-300)
            assert rel.max() <= 1e-13, f"TK relative difference {rel.max():.3e}"
            h = d["dp"] + g.get("eta1")  # the depth the step's friction saw
            uv = np.hypot(g.get("uu1"), g.get("vv1"))
            quad = run.cf * uv / h
            shallow = (h < run.hbreak) & (uv > 0) & (b < 2.0 / run.dtdp)
            assert shallow.any() and (b[shallow] > quad[shallow]).all()
    nc_g, nc_o = g.get("nodecode"), o.iarray("nodecode")
    flips = int((nc_g != nc_o).sum())
    both = (nc_g == 1) & (nc_o == 1)
    de = float(np.abs(g.get("eta2") - o.array("eta2"))[both].max())
    log(f"  flags differing {flips}, max |d eta2| wet {de:.2e} m")
    assert flips <= 2
    assert de <= 1e-8
