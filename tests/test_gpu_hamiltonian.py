"""Spin-orbital Hamiltonian assembly on the device (libgrpp_batch_plan_spin_hamiltonian) against the
CPU restatement (oracle/hamiltonian.py) applied to the same plan's four matrices, Cartesian and real
spherical, for one-system and many-molecule plans; the device result is Hermitian with
Kramers-degenerate eigenvalues."""
import numpy as np
import pytest

from libgrpp_amd import configs
from oracle import hamiltonian as HM

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def G():
    from libgrpp_amd import lib
    lib.lib()
    return lib


def _check(got, mats, sysm, spherical, what):
    want = HM.spin_hamiltonian(*mats, [s.L for s in sysm.shells], spherical)
    scale = max(1.0, float(np.abs(want).max()))
    err = float(np.abs(got - want).max())
    print(f"{what} spherical={spherical}: order {got.shape[0]}, max abs {err:.3e}", flush=True)
    assert got.shape == want.shape
    assert err <= 1e-12 * scale, what
    # Hermitian up to the (anti)symmetry of the input matrices themselves (a diagonal shell block is
    # evaluated element by element and carries the formulation's rounding)
    asym = float(np.abs(mats[0] - mats[0].T).max()) + max(float(np.abs(s + s.T).max()) for s in mats[1:])
    herm = float(np.abs(got - got.conj().T).max())
    assert herm <= 1e-12 * scale + asym, (what, herm, asym, scale)


@pytest.mark.parametrize("case", ["h2o", "kcs"])
def test_single_system_plan(G, case, request):
    sysm = request.getfixturevalue(case)
    plan = G.Plan(sysm)
    plan.execute()
    mats = plan.download()
    for spherical in (False, True):
        H = plan.spin_hamiltonian(spherical)
        _check(H, mats, sysm, spherical, case)
        if case == "h2o":
            ev = np.linalg.eigvalsh(H)
            assert np.abs(ev[0::2] - ev[1::2]).max() <= 1e-9 * np.abs(ev).max()
    plan.close()


def test_molecules_plan(G):
    mols = configs.ragged_molecules(seed=3, nmol=5, atoms=(2, 4), centers=(1, 2), lmax_shell=6, nprim=(1, 3),
                                    shells_per_atom=(1, 3), alpha=(0.3, 20.0))
    plan = G.MoleculesPlan(mols)
    plan.execute()
    mats = plan.download()
    for spherical in (False, True):
        Hs = plan.spin_hamiltonian(spherical)
        assert len(Hs) == len(mols)
        for m, (H, M) in enumerate(zip(Hs, mats)):
            _check(H, M, mols[m], spherical, f"molecule {m}")
    plan.close()
