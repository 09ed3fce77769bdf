"""CPU restatement (TEST INFRASTRUCTURE ONLY) of the step after the GRPP integrals: the one-electron
spin-orbital Hamiltonian built from libgrpp's four matrices.

* libgrpp's spin-orbit matrices are real: <A| U^SO P l_eta |B> = i so_eta[A][B], because its L
  matrices hold Im <S_m|l_eta|S_m'> (lmatrix.c:29-74); the spin s = sigma/2 is applied by the host
  when the Hamiltonian is built (doc/manual.tex:594).  Over spin-orbitals ordered all-alpha then
  all-beta, with the Pauli matrices,
      H = A (x) 1 + sum_eta i S_eta (x) sigma_eta / 2
        = [ A + i Z/2        Y/2 + i X/2 ]
          [ -Y/2 + i X/2     A - i Z/2   ]
* Cartesian -> real spherical harmonics per shell: the coefficients libgrpp's outercore code uses,
  S_lm = sum_c u_lm,c x^r y^s z^t with u = spherical_to_cartesian_coef(l, m, r, s, t) /
  ang_norm_factor (outercore_integrals.c:171-178, 545-553), applied to both sides as
  transform_to_sph_basis_ket / _bra do (:505-542).

The coefficient function is the oracle's own restatement (orc_rsh_coef, pinned in
tests/test_oracle_math.py); this module is pinned by tests/test_oracle_hamiltonian.py (the
spherical functions of a normalised shell are orthonormal, closed forms for l <= 2, Hermiticity,
the block structure of the spin couplings).
"""
from __future__ import annotations

import ctypes as C
import math
from typing import List

import numpy as np


def _dfact(n: int) -> float:
    r = 1.0
    while n > 1:
        r *= n
        n -= 2
    return r


def cart_list(l: int):
    """libgrpp's default Cartesian order (parameters.c:94-134): x power descending, then y"""
    return [(r, s, l - r - s) for r in range(l, -1, -1) for s in range(l - r, -1, -1)]


def sph_transform(l: int) -> np.ndarray:
    """(2l+1) x ncart(l) matrix u[m + l][c] (outercore_integrals.c:171-178)"""
    import oracle
    lib = oracle.oracle().lib
    lib.orc_rsh_coef.restype = C.c_double
    lib.orc_rsh_coef.argtypes = [C.c_int] * 5
    an = 1.0 / (2.0 * math.sqrt(math.pi)) * math.sqrt(_dfact(2 * l + 1))
    cl = cart_list(l)
    u = np.zeros((2 * l + 1, len(cl)))
    for m in range(-l, l + 1):
        for c, (r, s, t) in enumerate(cl):
            u[m + l, c] = lib.orc_rsh_coef(l, m, r, s, t) / an
    return u


def transform(shell_L: List[int], spherical: bool) -> np.ndarray:
    """block-diagonal T (rows: spherical or Cartesian functions, columns: Cartesian functions)"""
    blocks = [sph_transform(l) if spherical else np.eye((l + 1) * (l + 2) // 2) for l in shell_L]
    n = sum(b.shape[0] for b in blocks)
    m = sum(b.shape[1] for b in blocks)
    T = np.zeros((n, m))
    r = c = 0
    for b in blocks:
        T[r:r + b.shape[0], c:c + b.shape[1]] = b
        r += b.shape[0]
        c += b.shape[1]
    return T


def spin_hamiltonian(arep, so_x, so_y, so_z, shell_L: List[int], spherical: bool) -> np.ndarray:
    """the complex 2n x 2n spin-orbital Hamiltonian (alpha functions first)"""
    T = transform(shell_L, spherical)
    A, X, Y, Z = (T @ M @ T.T for M in (arep, so_x, so_y, so_z))
    n = A.shape[0]
    H = np.zeros((2 * n, 2 * n), dtype=complex)
    H[:n, :n] = A + 0.5j * Z
    H[n:, n:] = A - 0.5j * Z
    H[:n, n:] = 0.5 * Y + 0.5j * X
    H[n:, :n] = -0.5 * Y + 0.5j * X
    return H
