"""ctypes view of libgrpp's public object layouts (libgrpp.h:42-103).

The structures are ABI-identical to the reference's, so one description serves this
project's library and any library exporting the same layout.  ``ObjectFactory`` creates
native objects through a library's own constructors (names given by a prefix map) from the
Python dataclasses of :mod:`libgrpp_amd.io`, and frees them again.
"""
from __future__ import annotations

import ctypes as C
from typing import Dict, List, Sequence

import numpy as np

from .io import GRPP, Potential, Shell

dptr = C.POINTER(C.c_double)
iptr = C.POINTER(C.c_int)


class PotentialT(C.Structure):
    _fields_ = [("L", C.c_int), ("J", C.c_int), ("num_primitives", C.c_int),
                ("powers", iptr), ("coeffs", dptr), ("alpha", dptr)]


class ShellT(C.Structure):
    _fields_ = [("L", C.c_int), ("cart_size", C.c_int), ("cart_list", iptr), ("num_primitives", C.c_int),
                ("coeffs", dptr), ("alpha", dptr), ("origin", C.c_double * 3)]


class GrppT(C.Structure):
    _fields_ = [("n_arep", C.c_int), ("n_esop", C.c_int), ("n_oc_shells", C.c_int),
                ("U_L", C.POINTER(PotentialT)), ("U_arep", C.POINTER(C.POINTER(PotentialT))),
                ("U_esop", C.POINTER(C.POINTER(PotentialT))), ("U_oc", C.POINTER(C.POINTER(PotentialT))),
                ("oc_shells", C.POINTER(C.POINTER(ShellT)))]


PSHELL = C.POINTER(ShellT)
PPOT = C.POINTER(PotentialT)
PGRPP = C.POINTER(GrppT)

LIBGRPP_NAMES = dict(new_shell="libgrpp_new_shell", delete_shell="libgrpp_delete_shell",
                     new_potential="libgrpp_new_potential", delete_potential="libgrpp_delete_potential",
                     new_grpp="libgrpp_new_grpp", set_local="libgrpp_grpp_set_local_potential",
                     add_arep="libgrpp_grpp_add_averaged_potential", add_esop="libgrpp_grpp_add_spin_orbit_potential",
                     add_oc="libgrpp_grpp_add_outercore_potential", delete_grpp="libgrpp_delete_grpp")


def darr(x: Sequence[float]):
    a = (C.c_double * max(1, len(x)))()
    for i, v in enumerate(x):
        a[i] = float(v)
    return a


def iarr(x: Sequence[int]):
    a = (C.c_int * max(1, len(x)))()
    for i, v in enumerate(x):
        a[i] = int(v)
    return a


def as_dptr(a: np.ndarray):
    assert a.dtype == np.float64 and a.flags["C_CONTIGUOUS"]
    return a.ctypes.data_as(dptr)


class ObjectFactory:
    """Builds native shells / potentials / GRPP operators through a library's constructors."""

    def __init__(self, lib: C.CDLL, names: Dict[str, str] = LIBGRPP_NAMES):
        self.lib = lib
        f = {k: getattr(lib, v) for k, v in names.items()}
        f["new_shell"].restype = PSHELL
        f["new_shell"].argtypes = [dptr, C.c_int, C.c_int, dptr, dptr]
        f["delete_shell"].argtypes = [PSHELL]
        f["new_potential"].restype = PPOT
        f["new_potential"].argtypes = [C.c_int, C.c_int, C.c_int, iptr, dptr, dptr]
        f["delete_potential"].argtypes = [PPOT]
        f["new_grpp"].restype = PGRPP
        f["new_grpp"].argtypes = []
        f["set_local"].argtypes = [PGRPP, PPOT]
        f["add_arep"].argtypes = [PGRPP, PPOT]
        f["add_esop"].argtypes = [PGRPP, PPOT]
        f["add_oc"].argtypes = [PGRPP, PPOT, PSHELL]
        f["delete_grpp"].argtypes = [PGRPP]
        self.f = f

    def shell(self, s: Shell) -> PSHELL:
        return self.f["new_shell"](darr(s.origin), s.L, len(s.alpha), darr(s.coeffs), darr(s.alpha))

    def potential(self, p: Potential) -> PPOT:
        return self.f["new_potential"](p.L, p.J, len(p.alpha), iarr(p.powers), darr(p.coeffs), darr(p.alpha))

    def grpp(self, g: GRPP) -> PGRPP:
        h = self.f["new_grpp"]()
        if g.local is not None:
            self.f["set_local"](h, self.potential(g.local))
        for p in g.arep:
            self.f["add_arep"](h, self.potential(p))
        for p in g.esop:
            self.f["add_esop"](h, self.potential(p))
        for p, s in g.oc:
            self.f["add_oc"](h, self.potential(p), self.shell(s))
        return h

    def shells(self, shells: List[Shell]):
        arr = (PSHELL * max(1, len(shells)))()
        for i, s in enumerate(shells):
            arr[i] = self.shell(s)
        return arr

    def free_shells(self, arr, n: int) -> None:
        for i in range(n):
            self.f["delete_shell"](arr[i])

    def free_grpp(self, h) -> None:
        self.f["delete_grpp"](h)
