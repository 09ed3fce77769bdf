"""The CPU restatement of the spin-orbital Hamiltonian step (oracle/hamiltonian.py) against what the
mathematics fixes: the real spherical functions of a normalised libgrpp shell are orthonormal,
closed forms of the l <= 2 transforms, and the structure of the spin couplings (Hermiticity,
Kramers degeneracy of a time-reversal-symmetric operator, spin-free limit)."""
import math

import numpy as np
import pytest

from oracle import hamiltonian as HM


def _one_centre_overlap(l, a):
    """<x^r y^s z^t | x^r' y^s' z^t'> of one primitive with libgrpp's normalisation
    (2a/pi)^{3/4} (4a)^{l/2} on every component (norm_gaussian.c:16-23)"""
    def I(k):
        return 0.0 if k % 2 else math.gamma((k + 1) / 2) / (2 * a) ** ((k + 1) / 2)
    N = (2 * a / math.pi) ** 0.75 * (4 * a) ** (l / 2)
    cl = HM.cart_list(l)
    return np.array([[N * N * I(p[0] + q[0]) * I(p[1] + q[1]) * I(p[2] + q[2]) for q in cl] for p in cl])


@pytest.mark.parametrize("l", range(0, 9))
def test_spherical_functions_of_a_normalised_shell_are_orthonormal(l):
    T = HM.sph_transform(l)
    M = T @ _one_centre_overlap(l, 0.77) @ T.T
    assert np.abs(M - np.eye(2 * l + 1)).max() < 1e-13


def test_low_l_transforms_closed_forms():
    """p: (y, z, x) for m = -1, 0, 1; d: xy, yz, (2zz - xx - yy)/(2 sqrt 3), xz, (xx - yy)/2"""
    assert np.allclose(HM.sph_transform(0), [[1.0]])
    assert np.allclose(HM.sph_transform(1), [[0, 1, 0], [0, 0, 1], [1, 0, 0]])
    r3 = 1.0 / (2.0 * math.sqrt(3.0))
    want = np.array([[0, 1, 0, 0, 0, 0], [0, 0, 0, 0, 1, 0], [-r3, 0, 0, -r3, 0, 2 * r3], [0, 0, 1, 0, 0, 0],
                     [0.5, 0, 0, -0.5, 0, 0]])
    assert np.allclose(HM.sph_transform(2), want, atol=1e-15)


def _random_operator(rng, n):
    a = rng.normal(size=(n, n))
    a = a + a.T
    so = []
    for _ in range(3):
        s = rng.normal(size=(n, n))
        so.append(s - s.T)
    return a, so


@pytest.mark.parametrize("spherical", [False, True])
def test_spin_hamiltonian_structure(spherical):
    """symmetric arep and antisymmetric so_eta give a Hermitian H whose eigenvalues come in
    degenerate (Kramers) pairs, as for any time-reversal-symmetric spin-orbit operator; with the
    spin-orbit part off, H is arep in each spin block"""
    rng = np.random.default_rng(4)
    shells = [0, 1, 2, 3, 1]
    n = sum((l + 1) * (l + 2) // 2 for l in shells)
    a, so = _random_operator(rng, n)
    H = HM.spin_hamiltonian(a, *so, shells, spherical)
    assert np.abs(H - H.conj().T).max() < 1e-12
    ev = np.linalg.eigvalsh(H)
    assert np.abs(ev[0::2] - ev[1::2]).max() < 1e-9 * np.abs(ev).max()
    H0 = HM.spin_hamiltonian(a, *(np.zeros_like(s) for s in so), shells, spherical)
    T = HM.transform(shells, spherical)
    m = T.shape[0]
    assert np.allclose(H0[:m, :m], T @ a @ T.T) and np.allclose(H0[m:, m:], T @ a @ T.T)
    assert np.abs(H0[:m, m:]).max() == 0.0


def test_spin_hamiltonian_pauli_couplings():
    """one spin-orbit component at a time: H - A (x) 1 equals i so_eta (x) sigma_eta / 2 with the Pauli
    matrices written out"""
    rng = np.random.default_rng(5)
    shells = [1, 2]
    n = 9
    a, so = _random_operator(rng, n)
    sig = [np.array([[0, 1], [1, 0]], complex), np.array([[0, -1j], [1j, 0]]), np.array([[1, 0], [0, -1]], complex)]
    for eta in range(3):
        parts = [np.zeros((n, n))] * 3
        parts[eta] = so[eta]
        H = HM.spin_hamiltonian(np.zeros((n, n)), *parts, shells, False)
        # spin-major layout: block (s1, s2) = i so_eta sigma[s1, s2] / 2
        want = np.block([[1j * so[eta] * sig[eta][s1, s2] / 2 for s2 in range(2)] for s1 in range(2)])
        assert np.abs(H - want).max() < 1e-15, eta
