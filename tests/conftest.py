import os
import sys

import numpy as np
import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
REPO = os.path.dirname(HERE)
for p in (REPO, HERE):
    if p not in sys.path:
        sys.path.insert(0, p)

GOLDEN = os.path.join(HERE, "golden")


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs libadcgpu.so on the device)")


def pytest_collection_modifyitems(session, config, items):
    # The Fortran interface test runs its check program as a child process.
    # A process's cooperative launch of the JCG kernel is refused while
    # another process on the same GPU has made one, so this test goes first,
    # before the test process itself launches any.
    items.sort(key=lambda it: 0 if it.nodeid.startswith("tests/test_gpu_fortran.py") else 1)


def load_golden(name):
    path = os.path.join(GOLDEN, f"{name}.npz")
    if not os.path.exists(path):
        pytest.skip(f"golden fixture {name}.npz missing (tests/golden/make_golden.py)")
    return dict(np.load(path))


@pytest.fixture(scope="session")
def beach_golden():
    return load_golden("beach")


@pytest.fixture(scope="session")
def beach_lumped_golden():
    return load_golden("beach_lumped")


@pytest.fixture(scope="session")
def basin_tide_golden():
    return load_golden("basin_tide")


@pytest.fixture(scope="session")
def basin_river_golden():
    return load_golden("basin_river")


@pytest.fixture(scope="session")
def basin_radiation_golden():
    return load_golden("basin_radiation")


@pytest.fixture(scope="session")
def basin_river52_golden():
    return load_golden("basin_river52")


@pytest.fixture(scope="session")
def beach_hybrid_golden():
    return load_golden("beach_hybrid")


def beach_hybrid_run():
    """The storm beach with hybrid friction (NOLIBF = 2: CFmin 0.0025,
    HBREAK = 2 m, FTHETA = 10, FGAMMA = 1/3), as tests/golden/make_golden.py
    runs it."""
    from adcirc_amd import meshgen
    run = meshgen.beach().run
    run.nolibf, run.cf, run.hbreak, run.ftheta, run.fgamma = 2, 0.0025, 2.0, 10.0, 1.0 / 3.0
    return run


@pytest.fixture(scope="session")
def basin_golden():
    return load_golden("basin")


@pytest.fixture(scope="session")
def basin_linear_golden():
    return load_golden("basin_linear")


@pytest.fixture(scope="session")
def jcg_golden():
    return load_golden("jcg")
