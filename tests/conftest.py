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


@pytest.fixture(scope="session")
def kcs():
    from libgrpp_amd import io
    return io.load_system(os.path.join(GOLDEN, "KCs"))


@pytest.fixture(scope="session")
def kcs_oracle(kcs):
    """oracle matrices of KCs with and without FMA contraction (the reference formulation's
    rounding spread), computed once per session"""
    import oracle
    return oracle.oracle().matrices(kcs), oracle.oracle_nofma().matrices(kcs)


def within_noise(got, want, spread, factor=10.0):
    """element-wise: the parity rule, or |d| <= 1e-10 + factor * (rounding spread of the reference
    formulation at that element)"""
    import numpy as np
    d = np.abs(got - want)
    rule = (d <= 1e-10) | ((np.abs(want) > 1.0) & (d <= 1e-9 * np.abs(want)))
    return rule | (d <= 1e-10 + factor * spread)


def golden_matrix(case, name):
    from libgrpp_amd import io
    p = os.path.join(GOLDEN, case, name + ".txt")
    if not os.path.exists(p):
        p += ".gz"
    return io.read_lower_triangle(p)
