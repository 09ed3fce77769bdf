"""Input/output formats of the libgrpp drivers, restated in Python.

* ``molecule.xyz``  -- atom count, then ``Symbol x y z`` (bohr), read token-wise
  (test_libgrpp_c/molecule.c:46-101);
* ``basis.inp``     -- ``* Z nblocks`` then per block ``nprim ncontr`` and rows
  ``alpha c_1 .. c_ncontr``; block index = L; every contracted column becomes one shell,
  normalised with the semantics of ``radial_gto_norm_factor_``
  (test_libgrpp_c/basis.c:88-163, libgrpp/fortran.c:689-702);
* ``ecp.inp``       -- ``* Z n_oc_blocks n_channels``; outercore blocks (L = block index,
  raw coefficients), then channels ``n_prim n_oc`` with rows ``n zeta d_arep d_esop d_oc..``;
  the last channel is the local U_L, SO terms exist for every L >= 1, outercore potentials
  take J = |2L -+ 1| alternately and pair with outercore shells in order
  (test_libgrpp_c/rpp.c:17-148);
* lower-triangle matrix text files ``%6d%6d%30.16E`` with |x| <= 1e-14 omitted
  (test_libgrpp_c/print_matrix.c:31-48).
"""
from __future__ import annotations

import math
from dataclasses import dataclass, field
from typing import Dict, List, Optional, Sequence, Tuple

import numpy as np

ELEMENTS = (
    "H He Li Be B C N O F Ne Na Mg Al Si P S Cl Ar K Ca Sc Ti V Cr Mn Fe Co Ni Cu Zn Ga Ge As Se Br Kr "
    "Rb Sr Y Zr Nb Mo Tc Ru Rh Pd Ag Cd In Sn Sb Te I Xe Cs Ba La Ce Pr Nd Pm Sm Eu Gd Tb Dy Ho Er Tm Yb "
    "Lu Hf Ta W Re Os Ir Pt Au Hg Tl Pb Bi Po At Rn Fr Ra Ac Th Pa U Np Pu Am Cm Bk Cf Es Fm Md No Lr Rf "
    "Db Sg Bh Hs Mt Ds Rg Cn Nh Fl Mc Lv Ts Og"
).split()


def element_charge(symbol: str) -> int:
    """Nuclear charge of a chemical symbol (test_libgrpp_c/elements.c:162)."""
    for z, s in enumerate(ELEMENTS, start=1):
        if s.lower() == symbol.lower():
            return z
    raise ValueError(f"unknown element {symbol!r}")


@dataclass
class Shell:
    """Contracted Cartesian shell: angular momentum, centre (bohr), raw coefficients, exponents."""
    L: int
    origin: Tuple[float, float, float]
    coeffs: List[float]
    alpha: List[float]

    @property
    def cart_size(self) -> int:
        return (self.L + 1) * (self.L + 2) // 2


@dataclass
class Potential:
    """Radial expansion U(r) = sum_k c_k r^(n_k-2) exp(-alpha_k r^2) of channel (L, J=2j)."""
    L: int
    J: int
    powers: List[int]
    coeffs: List[float]
    alpha: List[float]


@dataclass
class GRPP:
    """Whole operator: local U_L, semilocal AREP channels, SO channels, outercore (potential, shell) pairs."""
    local: Optional[Potential] = None
    arep: List[Potential] = field(default_factory=list)
    esop: List[Potential] = field(default_factory=list)
    oc: List[Tuple[Potential, Shell]] = field(default_factory=list)


@dataclass
class System:
    """Shell list plus GRPP centres (coordinates and operator per centre)."""
    shells: List[Shell]
    centers: List[Tuple[Tuple[float, float, float], GRPP]]

    @property
    def dim(self) -> int:
        return sum(s.cart_size for s in self.shells)

    def offsets(self) -> List[int]:
        off, o = [], 0
        for s in self.shells:
            off.append(o)
            o += s.cart_size
        return off


class _Tokens:
    """fscanf-like token stream with the drivers' skip-to-end-of-line."""

    def __init__(self, text: str):
        self.text = text
        self.pos = 0

    def next(self) -> Optional[str]:
        t, n = self.text, len(self.text)
        while self.pos < n and t[self.pos].isspace():
            self.pos += 1
        if self.pos >= n:
            return None
        start = self.pos
        while self.pos < n and not t[self.pos].isspace():
            self.pos += 1
        return t[start:self.pos]

    def skip_line(self) -> None:
        i = self.text.find("\n", self.pos)
        self.pos = len(self.text) if i < 0 else i + 1


def _num(tok: str) -> float:
    return float(tok.replace("D", "E").replace("d", "e"))


def _seek_entry(tok: _Tokens, z: int, nint: int) -> Optional[List[int]]:
    while True:
        t = tok.next()
        if t is None:
            return None
        if t == "*":
            vals = [int(tok.next()) for _ in range(nint)]
            if vals[0] == z:
                return vals


def double_factorial(n: int) -> int:
    r = 1
    while n > 1:
        r *= n
        n -= 2
    return r


def gaussian_norm_factor(L: int, a: float) -> float:
    """(2a/pi)^(3/4) (4a)^(L/2), applied by libgrpp to every primitive (norm_gaussian.c:16-23)."""
    return (2 * a / math.pi) ** 0.75 * (4 * a) ** (0.5 * L)


def contracted_norm(L: int, coeffs: Sequence[float], alpha: Sequence[float]) -> float:
    """sqrt((2L-1)!!)/sqrt(S), S the self-overlap of the x^L component of the contraction with
    libgrpp's primitive factors (fortran.c:689-702); S in closed form:
    int x^{2L} exp(-p r^2) d^3r = (2L-1)!!/(2p)^L (pi/p)^{3/2}."""
    S = 0.0
    for ci, ai in zip(coeffs, alpha):
        for cj, aj in zip(coeffs, alpha):
            p = ai + aj
            S += ci * cj * gaussian_norm_factor(L, ai) * gaussian_norm_factor(L, aj) * \
                double_factorial(2 * L - 1) / (2 * p) ** L * (math.pi / p) ** 1.5
    return math.sqrt(double_factorial(2 * L - 1)) / math.sqrt(S)


def read_molecule(path: str) -> List[Tuple[int, Tuple[float, float, float]]]:
    tok = _Tokens(open(path).read())
    n = int(tok.next())
    atoms = []
    for _ in range(n):
        sym = tok.next()
        xyz = (_num(tok.next()), _num(tok.next()), _num(tok.next()))
        atoms.append((element_charge(sym), xyz))
    return atoms


def read_basis(path: str, z: int) -> List[Tuple[int, List[float], List[float]]]:
    """Normalised contracted functions (L, coeffs, alpha) of element z."""
    tok = _Tokens(open(path).read())
    head = _seek_entry(tok, z, 2)
    if head is None:
        raise ValueError(f"basis for Z={z} not found in {path}")
    funs = []
    for L in range(head[1]):
        nprim, ncontr = int(tok.next()), int(tok.next())
        alpha, cols = [], [[] for _ in range(ncontr)]
        for _ in range(nprim):
            alpha.append(_num(tok.next()))
            for c in range(ncontr):
                cols[c].append(_num(tok.next()))
            tok.skip_line()
        for c in cols:
            f = contracted_norm(L, c, alpha)
            funs.append((L, [x * f for x in c], list(alpha)))
    return funs


def read_grpp(path: str, z: int) -> Optional[GRPP]:
    tok = _Tokens(open(path).read())
    head = _seek_entry(tok, z, 3)
    if head is None:
        return None
    _, n_blocks, n_arep = head
    g = GRPP()
    oc_shells: List[Shell] = []
    for L in range(n_blocks):
        n_alpha, n_contr = int(tok.next()), int(tok.next())
        tok.skip_line()
        alpha, cols = [], [[] for _ in range(n_contr)]
        for _ in range(n_alpha):
            alpha.append(_num(tok.next()))
            for c in range(n_contr):
                cols[c].append(_num(tok.next()))
        for c in cols:
            oc_shells.append(Shell(L, (0.0, 0.0, 0.0), list(c), list(alpha)))
    oc_pots: List[Potential] = []
    for L in range(n_arep):
        n_prim, n_oc = int(tok.next()), int(tok.next())
        tok.skip_line()
        pw, ex, arep, esop, occ = [], [], [], [], [[] for _ in range(n_oc)]
        for _ in range(n_prim):
            pw.append(int(tok.next()))
            ex.append(_num(tok.next()))
            arep.append(_num(tok.next()))
            esop.append(_num(tok.next()))
            for k in range(n_oc):
                occ[k].append(_num(tok.next()))
            tok.skip_line()
        p = Potential(L, 0, pw, arep, ex)
        if L == n_arep - 1:
            g.local = p
        else:
            g.arep.append(p)
        if L > 0:
            g.esop.append(Potential(L, 0, list(pw), esop, list(ex)))
        for k in range(n_oc):
            J = abs(2 * L + (-1 if k % 2 == 0 else +1))
            oc_pots.append(Potential(L, J, list(pw), occ[k], list(ex)))
    g.oc = list(zip(oc_pots, oc_shells[:len(oc_pots)]))
    return g


def load_system(folder: str) -> System:
    """Builds the shell list (atoms in xyz order, basis functions in file order) and the GRPP
    centres the reference driver would use (test_libgrpp_c/shell_list.c:18-51, eval_integrals.c:102-117)."""
    import os
    atoms = read_molecule(os.path.join(folder, "molecule.xyz"))
    shells, centers = [], []
    bcache: Dict[int, list] = {}
    for z, xyz in atoms:
        if z not in bcache:
            bcache[z] = read_basis(os.path.join(folder, "basis.inp"), z)
        for L, c, a in bcache[z]:
            shells.append(Shell(L, xyz, c, a))
    for z, xyz in atoms:
        g = read_grpp(os.path.join(folder, "ecp.inp"), z)
        if g is not None:
            centers.append((xyz, g))
    return System(shells, centers)


def write_lower_triangle(path: str, m: np.ndarray) -> None:
    dim = m.shape[0]
    with open(path, "w") as f:
        f.write("%6d\n" % dim)
        for i in range(dim):
            for j in range(i + 1):
                if abs(m[i, j]) > 1e-14:
                    f.write("%6d%6d%30.16E\n" % (i + 1, j + 1, m[i, j]))


def read_lower_triangle(path: str) -> np.ndarray:
    """Dense matrix holding the file's lower triangle (absent entries are 0, upper triangle 0)."""
    opener = open
    if path.endswith(".gz"):
        import gzip
        opener = gzip.open
    with opener(path, "rt") as f:
        dim = int(f.readline())
        data = np.loadtxt(f, ndmin=2) if dim else np.zeros((0, 3))
    m = np.zeros((dim, dim))
    if data.size:
        m[data[:, 0].astype(int) - 1, data[:, 1].astype(int) - 1] = data[:, 2]
    return m


def parity_error(got: np.ndarray, ref: np.ndarray) -> Tuple[float, float, bool]:
    """Parity rule: |d| <= 1e-10, or |d|/|ref| <= 1e-9 where |ref| > 1.
    Returns (max abs error, max rel error over |ref|>1, ok)."""
    d = np.abs(got - ref)
    big = np.abs(ref) > 1.0
    rel = np.zeros_like(d)
    rel[big] = d[big] / np.abs(ref[big])
    ok = bool(np.all((d <= 1e-10) | (big & (rel <= 1e-9))))
    return float(d.max(initial=0.0)), float(rel.max(initial=0.0)), ok
