"""The allocation-free drop-in JCG (adcgpu_jcg_noalloc) against the drop-in
that allocates (adcgpu_jcg): the same systems solved back to back through the
C ABI give bit-identical results, on two calls in a row, the second with a
smaller system; systems beyond the static scratch's limits are refused."""
import ctypes as C

import numpy as np
import pytest

import oracle_lib
from adcirc_amd import _lib
from adcirc_amd.model import jcg
from gpu_helpers import log


def _system(g, name):
    J, Cm = g[f"{name}_in_jcoef"], g[f"{name}_in_coef"]
    n, mz = J.shape
    itmax, zeta = g[f"{name}_prm"]
    jc = np.ascontiguousarray(J.astype(np.int32).T.ravel())
    co = np.ascontiguousarray(np.asarray(Cm, np.float64).T.ravel())
    rhs = np.array(g[f"{name}_in_rhs"], np.float64)
    u = np.array(g[f"{name}_in_u"], np.float64)
    iparm = np.zeros(12, np.int32)
    rparm = np.zeros(12)
    # ITPACK's defaults (dfault), then the settings the GWCE call uses
    oracle_lib.load().adco_dfault(iparm.ctypes.data_as(oracle_lib._ip),
                                  rparm.ctypes.data_as(oracle_lib._dp))
    iparm[0], iparm[1], iparm[3] = int(itmax), 0, 16
    rparm[0] = zeta
    return [co, jc, rhs, u, iparm, rparm]


@pytest.mark.gpu
def test_noalloc_matches_dropin(jcg_golden):
    g = jcg_golden
    # the diagonal system converges within five iterations, so ITPACK's
    # saved stopping-test flag ends cleared on both entry points, whatever
    # earlier calls in this process left; then a larger and a smaller system
    seq = ["diag", "basin0", "lap1"]
    sizes = []
    for name in seq:
        log(f"noalloc vs drop-in: system {name}")
        a = _system(g, name)
        b = [x.copy() for x in a]
        ier_a = jcg(*a)
        ier_b = jcg(*b, noalloc=True)
        sizes.append(a[2].shape[0])
        assert ier_a == ier_b, name
        for x, y, what in zip(a, b, ("coef", "jcoef", "rhs", "u", "iparm", "rparm")):
            assert np.array_equal(x, y), (name, what)
    assert sizes[1] > sizes[2]


def test_noalloc_refuses_oversize_without_device():
    lib = _lib.load()
    n = _lib.JCG_NOALLOC_MAX_N + 1
    one = np.zeros(1)
    i1 = np.zeros(12, np.int32)
    d1 = np.zeros(12)
    P = lambda a, t: a.ctypes.data_as(t)  # noqa: E731
    rc = lib.adcgpu_jcg_noalloc(n, n, 4, P(i1, _lib.ip_), P(one, _lib.dp_), P(one, _lib.dp_),
                                P(one, _lib.dp_), P(i1, _lib.ip_), P(d1, _lib.dp_))
    assert rc == _lib.E_ARG
    assert b"scratch" in lib.adcgpu_last_error()
    rc = lib.adcgpu_jcg_noalloc(8, 8, _lib.JCG_NOALLOC_MAX_NZ + 1, P(i1, _lib.ip_),
                                P(one, _lib.dp_), P(one, _lib.dp_), P(one, _lib.dp_),
                                P(i1, _lib.ip_), P(d1, _lib.dp_))
    assert rc == _lib.E_ARG
