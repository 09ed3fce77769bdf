import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

GOLDEN = os.path.join(ROOT, "tests", "golden")


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP kernels)")


@pytest.fixture(scope="session")
def orc():
    import oracle
    return oracle.oracle()


@pytest.fixture(scope="session")
def ref():
    import oracle
    r = oracle.reference()
    if r is None:
        pytest.skip("genuine reference build (oracle/_ref) not present")
    return r


@pytest.fixture(scope="session")
def h2o():
    from libgrpp_amd import io
    return io.load_system(os.path.join(GOLDEN, "H2O"))


@pytest.fixture(scope="session")
def cs():
    from libgrpp_amd import io
    return io.load_system(os.path.join(GOLDEN, "Cs"))


def golden_matrix(case, name):
    from libgrpp_amd import io
    p = os.path.join(GOLDEN, case, name + ".txt")
    if not os.path.exists(p):
        p += ".gz"
    return io.read_lower_triangle(p)
