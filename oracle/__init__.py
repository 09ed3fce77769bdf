"""CPU oracle access for tests and the CPU-baseline leg of bench.py -- TEST INFRASTRUCTURE ONLY.

* ``oracle()``    -- this project's C restatement of the libgrpp GRPP path (oracle/liborc.so);
* ``reference()`` -- the genuine libgrpp sources compiled as they lie (oracle/_ref/libgrpp_ref.so),
  present only where ``make -C oracle ref`` could run.

The product (libgrpp_amd) never imports this package.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess
from typing import Optional

import numpy as np

from libgrpp_amd.abi import (LIBGRPP_NAMES, PGRPP, PSHELL, ObjectFactory, as_dptr, dptr)
from libgrpp_amd.io import System

HERE = os.path.dirname(os.path.abspath(__file__))
ORC_SO = os.path.join(HERE, "liborc.so")
REF_SO = os.path.join(HERE, "_ref", "libgrpp_ref.so")

ORC_NAMES = dict(new_shell="orc_new_shell", delete_shell="orc_delete_shell", new_potential="orc_new_potential",
                 delete_potential="orc_delete_potential", new_grpp="orc_new_grpp", set_local="orc_grpp_set_local",
                 add_arep="orc_grpp_add_arep", add_esop="orc_grpp_add_esop", add_oc="orc_grpp_add_oc",
                 delete_grpp="orc_delete_grpp")


class RadialOpts(C.Structure):
    _fields_ = [("tol", C.c_double), ("screening", C.c_int), ("min_level", C.c_int), ("fixed_level", C.c_int)]


def build(ref: bool = True) -> None:
    """Compiles liborc.so (and, when the reference sources are present, _ref/libgrpp_ref.so)."""
    subprocess.check_call(["make", "-s", "-C", HERE])
    if ref and os.path.isdir("/root/reference/libgrpp"):
        subprocess.check_call(["make", "-s", "-C", HERE, "ref"])


class CpuLib:
    """Common driver for the oracle and the compiled reference (identical object layouts)."""

    def __init__(self, path: str, kind: str):
        self.kind = kind
        self.lib = C.CDLL(path)
        if kind == "oracle":
            self.fac = ObjectFactory(self.lib, ORC_NAMES)
            self._mat = self.lib.orc_build_matrices
            self._grad = self.lib.orc_build_gradients
            self.lib.orc_init()
            self.opts = RadialOpts.in_dll(self.lib, "orc_radial_opts")
        else:
            self.fac = ObjectFactory(self.lib, LIBGRPP_NAMES)
            self._mat = self.lib.ref_build_grpp_matrices
            self._grad = self.lib.ref_build_grpp_gradients
            self.lib.libgrpp_init()
            self.opts = None
        self._mat.restype = C.c_double
        self._mat.argtypes = [C.c_int, C.POINTER(PSHELL), C.c_int, dptr, C.POINTER(PGRPP), dptr, dptr, dptr, dptr,
                              C.c_int, C.c_int, C.c_int]
        self._grad.restype = C.c_double
        self._grad.argtypes = [C.c_int, C.POINTER(PSHELL), C.c_int, dptr, C.POINTER(PGRPP), dptr, C.c_int]

    def _objects(self, sysm: System):
        shells = self.fac.shells(sysm.shells)
        nc = len(sysm.centers)
        grpps = (PGRPP * max(1, nc))()
        xyz = np.zeros(3 * max(1, nc))
        for i, (o, g) in enumerate(sysm.centers):
            grpps[i] = self.fac.grpp(g)   # one operator object per centre (see ref_driver.c)
            xyz[3 * i:3 * i + 3] = o
        return shells, grpps, xyz

    def _free(self, sysm, shells, grpps):
        self.fac.free_shells(shells, len(sysm.shells))
        for i in range(len(sysm.centers)):
            self.fac.free_grpp(grpps[i])

    def matrices(self, sysm: System, nthreads: int = 0, pair_range=(0, -1)):
        """(arep, so_x, so_y, so_z, seconds) summed over centres for ordered pairs in pair_range."""
        shells, grpps, xyz = self._objects(sysm)
        dim = sysm.dim
        out = [np.zeros((dim, dim)) for _ in range(4)]
        t = self._mat(len(sysm.shells), shells, len(sysm.centers), as_dptr(xyz), grpps,
                      *[as_dptr(o) for o in out], nthreads, pair_range[0], pair_range[1])
        self._free(sysm, shells, grpps)
        return out[0], out[1], out[2], out[3], t

    def gradients(self, sysm: System, nthreads: int = 0):
        """grad[op, 3*atom+k] with the GRPP centres taken as the atom list."""
        shells, grpps, xyz = self._objects(sysm)
        dim, na = sysm.dim, len(sysm.centers)
        out = np.zeros((4, 3 * na, dim, dim))
        t = self._grad(len(sysm.shells), shells, na, as_dptr(xyz), grpps, as_dptr(out), nthreads)
        self._free(sysm, shells, grpps)
        return out, t


_cache = {}


def oracle() -> CpuLib:
    if "orc" not in _cache:
        if not os.path.exists(ORC_SO):
            build(ref=False)
        _cache["orc"] = CpuLib(ORC_SO, "oracle")
    return _cache["orc"]


def reference() -> Optional[CpuLib]:
    if "ref" not in _cache:
        _cache["ref"] = CpuLib(REF_SO, "reference") if os.path.exists(REF_SO) else None
    return _cache["ref"]
