"""Synthetic mesh generators: deterministic, valid, and round-trip through fort.14."""
import hashlib
import os

import numpy as np

from adcirc_amd import fort14, meshgen, prep


def _digest(m):
    h = hashlib.sha256()
    for a in (m.x, m.y, m.dp, m.nm):
        h.update(np.ascontiguousarray(a).tobytes())
    return h.hexdigest()


def test_generators_are_deterministic():
    assert _digest(meshgen.basin().mesh) == _digest(meshgen.basin().mesh)
    assert _digest(meshgen.coastal(8000).mesh) == _digest(meshgen.coastal(8000).mesh)
    assert _digest(meshgen.basin(seed=2).mesh) != _digest(meshgen.basin().mesh)


def test_basin_shape():
    c = meshgen.basin()
    assert c.mesh.np_ == 5151 and c.mesh.ne == 10000
    assert c.mesh.x.min() == 0.0 and c.mesh.x.max() == 100.0e3
    assert len(c.mesh.open_bnd[0]) == 51


def test_coastal_mesh_is_valid():
    c = meshgen.coastal(20000)
    m = c.mesh
    assert 15000 < m.np_ < 30000
    t = prep.build_tables(m, c.run.nolifa, c.run.h0)
    assert np.all(t["areas"] > 0)
    # every boundary node lies on the rectangle
    lx = m.x.max()
    ly = m.y.max()
    ob = m.open_bnd[0] - 1
    assert np.all(m.x[ob] == 0.0)
    land = m.land_bnd[0][1] - 1
    on = (m.y[land] == 0) | (m.y[land] == ly) | (m.x[land] == lx)
    assert np.all(on)
    # shelf: deep offshore, dry land inland
    assert m.dp.max() > 3900 and m.dp.min() < -2.5


def test_fort14_roundtrip(tmp_path):
    c = meshgen.beach()
    p = os.path.join(tmp_path, "fort.14")
    fort14.write_fort14(p, c.mesh)
    m = fort14.read_fort14(p)
    for a, b in ((m.x, c.mesh.x), (m.y, c.mesh.y), (m.dp, c.mesh.dp), (m.nm, c.mesh.nm)):
        assert np.array_equal(a, b)
    assert np.array_equal(m.open_bnd[0], c.mesh.open_bnd[0])
    assert m.land_bnd[0][0] == 0 and np.array_equal(m.land_bnd[0][1], c.mesh.land_bnd[0][1])
