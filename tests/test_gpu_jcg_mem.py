"""The memory-resident JCG (k_jcg_mem): systems with more 4,096-row tiles
than the launch's blocks, either because the caller limits the blocks
(adcgpu_set_jcg_blocks) or because the mesh is larger than the resident
blocks hold (1,044,480 rows on MI355X).  Every row's operations and the
reduction tree are those of the register-resident solve, so the results are
bit-exact against the oracle like every other configuration."""
from __future__ import annotations

import dataclasses

import numpy as np
import pytest

from adcirc_amd import meshgen
from gpu_helpers import build_pair, case_setup, compare, log, step_pair

pytestmark = pytest.mark.gpu


def _bitexact(d, run, nsteps, check_every, blocks):
    g, o = build_pair(d, run)
    if blocks is not None:
        g.set_jcg_blocks(blocks)
    for it in range(1, nsteps + 1):
        step_pair(g, o, it)
        if it % check_every == 0 or it <= 3 or it == nsteps:
            bad = compare(g, o, it)
            assert not bad, f"step {it}: " + "; ".join(bad)
    return g, o


def test_limited_blocks_coastal_bitexact():
    """40k-node coastal storm (10 row tiles) on 3 blocks: two row blocks
    work through five tiles each; 60 steps bit-exact vs oracle."""
    case = meshgen.coastal(40000)
    d = case_setup(case)
    log(f"case coastal {int(d['np'])} nodes on 3 JCG blocks: 60 steps bit-exact vs oracle")
    g, _ = _bitexact(d, case.run, 60, 10, blocks=3)
    grid, cap, nch = g.jcg_shape()
    assert grid == 3 and (nch + 1) // 2 + 1 > grid


def test_limit_lifted_returns_to_register_resident():
    """The limit can be lifted between steps: the solve goes back to the
    register-resident kernel and the run stays bit-exact."""
    case = meshgen.coastal(20000)
    d = case_setup(case)
    log(f"case coastal {int(d['np'])} nodes: 2 blocks for 10 steps, then all; bit-exact")
    g, o = build_pair(d, case.run)
    g.set_jcg_blocks(2)
    for it in range(1, 21):
        if it == 11:
            g.set_jcg_blocks(0)
        step_pair(g, o, it)
        bad = compare(g, o, it)
        assert not bad, f"step {it}: " + "; ".join(bad)
    grid, cap, nch = g.jcg_shape()
    assert grid == (nch + 1) // 2 + 1


def test_mesh_beyond_resident_limit_bitexact():
    """A 1.2M-node coastal mesh, beyond the 1,044,480 rows the resident
    blocks hold: solved memory-resident without a limit being set; the first
    steps bit-exact vs oracle."""
    case = meshgen.coastal(1_200_000)
    d = case_setup(case)
    log(f"case coastal {int(d['np'])} nodes: 4 steps bit-exact vs oracle")
    g, _ = _bitexact(d, case.run, 4, 2, blocks=None)
    grid, cap, nch = g.jcg_shape()
    assert grid == cap and (nch + 1) // 2 + 1 > cap


def test_learned_start_and_batches_refuse_memory_resident():
    """The learned start (two starts in one register-resident launch) and
    batches need the register-resident solve: limiting the blocks is refused
    for the one, and a limited context cannot join the other; both stay
    usable."""
    from adcirc_amd.model import Batch

    case = meshgen.coastal(20000)
    d = case_setup(case)
    log("learned start and batch with a 2-block JCG limit: refused")
    run2 = dataclasses.replace(case.run, warm_start=2)
    g2, _ = build_pair(d, run2)
    before = g2.jcg_shape()
    with pytest.raises(RuntimeError, match=r"\(-2\)"):
        g2.set_jcg_blocks(2)
    assert g2.jcg_shape() == before
    g, _ = build_pair(d, case.run)
    g.set_jcg_blocks(2)
    with pytest.raises(RuntimeError, match=r"\(-2\)"):
        Batch([g])
    g.set_jcg_blocks(0)
    Batch([g])
