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
REF_NOFMA_SO = os.path.join(HERE, "_ref", "libgrpp_ref_nofma.so")
REF_ASSOC_SO = os.path.join(HERE, "_ref", "libgrpp_ref_assoc.so")

ORC_NAMES = dict(new_shell="orc_new_shell", delete_shell="orc_delete_shell", new_potential="orc_new_potential",
                 delete_potential="orc_delete_potential", new_grpp="orc_new_grpp", set_local="orc_grpp_set_local",
                 add_arep="orc_grpp_add_arep", add_esop="orc_grpp_add_esop", add_oc="orc_grpp_add_oc",
                 delete_grpp="orc_delete_grpp")


class RadialOpts(C.Structure):
    _fields_ = [("tol", C.c_double), ("screening", C.c_int), ("min_level", C.c_int), ("fixed_level", C.c_int),
                ("engine_order", C.c_int)]


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
                              C.c_int, C.c_int, C.POINTER(C.c_int)]
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

    def matrices(self, sysm: System, nthreads: int = 0, pairs=None):
        """(arep, so_x, so_y, so_z, seconds) summed over centres; pairs = ordered-pair indices
        i*nshell + j to evaluate (None = all)."""
        shells, grpps, xyz = self._objects(sysm)
        dim = sysm.dim
        out = [np.zeros((dim, dim)) for _ in range(4)]
        if pairs is None:
            plist, npl = None, 0
        else:
            arr = np.ascontiguousarray(np.asarray(pairs, dtype=np.int32))
            plist, npl = arr.ctypes.data_as(C.POINTER(C.c_int)), len(arr)
        t = self._mat(len(sysm.shells), shells, len(sysm.centers), as_dptr(xyz), grpps,
                      *[as_dptr(o) for o in out], nthreads, npl, plist)
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


def reference_nofma() -> Optional[CpuLib]:
    """The genuine reference sources compiled with floating-point contraction off: the reference's
    own second rounding of the same algorithm.  |reference() - reference_nofma()| per element is the
    noise floor of an ill-conditioned input."""
    if "ref_nofma" not in _cache:
        _cache["ref_nofma"] = CpuLib(REF_NOFMA_SO, "reference") if os.path.exists(REF_NOFMA_SO) else None
    return _cache["ref_nofma"]


def reference_assoc() -> Optional[CpuLib]:
    """The genuine reference sources compiled with reassociation allowed (gcc then vectorises the
    quadrature point sums into four-lane partial sums): the reference's own rounding in another
    summation order.  Together with reference_nofma() it bounds how far the reference's arithmetic
    fixes an ill-conditioned input."""
    if "ref_assoc" not in _cache:
        _cache["ref_assoc"] = CpuLib(REF_ASSOC_SO, "reference") if os.path.exists(REF_ASSOC_SO) else None
    return _cache["ref_assoc"]


# ---------------------------------------------------------------- per-call entry points
_PER_CALL = {
    "oracle": dict(type1="orc_type1", type2="orc_type2", so="orc_spin_orbit", oc="orc_outercore",
                   full="orc_full_grpp", grad="orc_full_grpp_gradient"),
    "reference": dict(type1="libgrpp_type1_integrals", type2="libgrpp_type2_integrals",
                      so="libgrpp_spin_orbit_integrals", oc="libgrpp_outercore_potential_integrals",
                      full="libgrpp_full_grpp_integrals", grad="libgrpp_full_grpp_integrals_gradient",
                      oc1="libgrpp_outercore_potential_integrals_part_1",
                      oc2="libgrpp_outercore_potential_integrals_part_2"),
}


def _per_call(self, what, A, B, Cpos, arg, D=None):
    """Single-task evaluation through the library's own per-call function.  what in
    type1/type2/so (arg: Potential), oc (arg: list of (Potential, Shell)), full/grad (arg: GRPP);
    for the compiled reference also oc1 (arg: (Potential, Shell)) and oc2 (arg: two of them), its
    outercore parts 1 and 2 (outercore_integrals.c:147, :277), which add into zeroed outputs."""
    from libgrpp_amd.abi import PPOT, darr
    fn = getattr(self.lib, _PER_CALL[self.kind][what])
    fn.restype = None
    a, b = self.fac.shell(A), self.fac.shell(B)
    na, nb = A.cart_size, B.cart_size
    Cc = darr(Cpos)
    if what in ("type1", "type2"):
        p = self.fac.potential(arg)
        m = np.zeros(na * nb)
        fn(a, b, Cc, p, as_dptr(m))
        self.fac.f["delete_potential"](p)
        res = m.reshape(na, nb)
    elif what == "so":
        p = self.fac.potential(arg)
        ms = [np.zeros(na * nb) for _ in range(3)]
        fn(a, b, Cc, p, *[as_dptr(x) for x in ms])
        self.fac.f["delete_potential"](p)
        res = np.array([x.reshape(na, nb) for x in ms])
    elif what == "oc":
        n = len(arg)
        pots = (PPOT * max(1, n))(*[self.fac.potential(p) for p, _ in arg])
        shs = (PSHELL * max(1, n))(*[self.fac.shell(s) for _, s in arg])
        ms = [np.zeros(na * nb) for _ in range(4)]
        fn(a, b, Cc, n, pots, shs, *[as_dptr(x) for x in ms])
        for i in range(n):
            self.fac.f["delete_potential"](pots[i])
            self.fac.f["delete_shell"](shs[i])
        res = np.array([x.reshape(na, nb) for x in ms])
    elif what in ("oc1", "oc2"):
        entries = [arg] if what == "oc1" else list(arg)
        objs = []
        for pot, sh in entries:
            objs += [self.fac.potential(pot), self.fac.shell(sh)]
        ms = [np.zeros(na * nb) for _ in range(4)]
        fn(a, b, Cc, *objs, *[as_dptr(x) for x in ms])
        for k in range(0, len(objs), 2):
            self.fac.f["delete_potential"](objs[k])
            self.fac.f["delete_shell"](objs[k + 1])
        res = np.array([x.reshape(na, nb) for x in ms])
    elif what == "full":
        h = self.fac.grpp(arg)
        ms = [np.zeros(na * nb) for _ in range(4)]
        fn(a, b, h, Cc, *[as_dptr(x) for x in ms])
        self.fac.free_grpp(h)
        res = np.array([x.reshape(na, nb) for x in ms])
    elif what == "grad":
        h = self.fac.grpp(arg)
        arrs = [[np.zeros(na * nb) for _ in range(3)] for _ in range(4)]
        ptrs = [(dptr * 3)(*[as_dptr(x) for x in ar]) for ar in arrs]
        fn(a, b, h, Cc, darr(D), *ptrs)
        self.fac.free_grpp(h)
        res = np.array([[x.reshape(na, nb) for x in ar] for ar in arrs])
    else:
        raise ValueError(what)
    self.fac.f["delete_shell"](a)
    self.fac.f["delete_shell"](b)
    return res


CpuLib.per_call = _per_call


# nuclear models (libgrpp.h:247-254): 0 point charge (analytic), 1 charged ball (R_rms), 2 Gaussian
# (R_rms), 3 Fermi (c, a), 4 Fermi + bubble (c, a, k), 5 point charge by the numeric quadrature
def _nuclear(self, A, B, Cpos, charge, model, params=(0.0, 0.0, 0.0), tol=1e-16):
    """nuclear attraction block of shells A, B for a nucleus of charge Z at Cpos"""
    from libgrpp_amd.abi import as_dptr, darr
    a, b = self.fac.shell(A), self.fac.shell(B)
    m = np.zeros(A.cart_size * B.cart_size)
    pp = darr(list(params) + [0.0] * (4 - len(params)))
    if self.kind == "oracle":
        fn = self.lib.orc_nuclear_attraction
        fn.restype = None
        fn(a, b, darr(Cpos), C.c_int(charge), C.c_int(model), pp, C.c_double(tol), as_dptr(m))
    else:
        fn = self.lib.libgrpp_nuclear_attraction_integrals
        fn.restype = None
        fn(a, b, darr(Cpos), C.c_int(charge), C.c_int(model), pp, as_dptr(m))
    self.fac.f["delete_shell"](a)
    self.fac.f["delete_shell"](b)
    return m.reshape(A.cart_size, B.cart_size)


CpuLib.nuclear = _nuclear
