"""Python host interface of libgrpp_amd.so (ctypes), mirroring libgrpp's C API.

The library is loaded from the package folder; if it is missing or no HIP device is usable,
calls raise -- there is no CPU fallback.
"""
from __future__ import annotations

import ctypes as C
import os
from typing import List, Optional, Tuple

import numpy as np

from .abi import LIBGRPP_NAMES, PGRPP, PPOT, PSHELL, ObjectFactory, as_dptr, darr, dptr
from .io import GRPP, Potential, Shell, System

_LIB = os.path.join(os.path.dirname(os.path.abspath(__file__)), "libgrpp_amd.so")
_lib = None


class Stats(C.Structure):
    _fields_ = [("n_tasks", C.c_longlong), ("n_integrals", C.c_longlong), ("n_radial_entries", C.c_longlong),
                ("n_radial_points", C.c_longlong), ("n_fallback_tiles", C.c_longlong), ("n_tiles", C.c_longlong),
                ("fp64_flops", C.c_double), ("device_seconds", C.c_double), ("radial_seconds", C.c_double),
                ("tables_seconds", C.c_double), ("screen_seconds", C.c_double), ("canon_seconds", C.c_double),
                ("proj_seconds", C.c_double), ("type1_seconds", C.c_double), ("contract_seconds", C.c_double),
                ("screen_triples", C.c_longlong), ("screen_steps", C.c_longlong), ("screen_flops", C.c_double)]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


BATCH_LEARNED_GRID = 1
BATCH_GRADIENT = 2
BATCH_NO_SYMMETRY = 4
BATCH_RECORD_LEVELS = 8
BATCH_GRAPH = 16
NFEAT = 12


def lib() -> C.CDLL:
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB):
            raise RuntimeError(f"{_LIB} is not built: run `python -m libgrpp_amd.build` (no CPU fallback exists)")
        L = C.CDLL(_LIB)
        L.libgrpp_last_error.restype = C.c_char_p
        pp = C.POINTER(dptr)
        for n in ("libgrpp_type1_integrals", "libgrpp_type2_integrals"):
            getattr(L, n).argtypes = [PSHELL, PSHELL, dptr, PPOT, dptr]
        L.libgrpp_spin_orbit_integrals.argtypes = [PSHELL, PSHELL, dptr, PPOT, dptr, dptr, dptr]
        L.libgrpp_outercore_potential_integrals.argtypes = [PSHELL, PSHELL, dptr, C.c_int, C.POINTER(PPOT),
                                                            C.POINTER(PSHELL), dptr, dptr, dptr, dptr]
        L.libgrpp_full_grpp_integrals.argtypes = [PSHELL, PSHELL, PGRPP, dptr, dptr, dptr, dptr, dptr]
        L.libgrpp_full_grpp_integrals_gradient.argtypes = [PSHELL, PSHELL, PGRPP, dptr, dptr, pp, pp, pp, pp]
        L.libgrpp_type1_integrals_gradient.argtypes = [PSHELL, PSHELL, dptr, PPOT, dptr, pp]
        L.libgrpp_type2_integrals_gradient.argtypes = [PSHELL, PSHELL, dptr, PPOT, dptr, pp]
        L.libgrpp_spin_orbit_integrals_gradient.argtypes = [PSHELL, PSHELL, dptr, PPOT, dptr, pp, pp, pp]
        L.libgrpp_outercore_potential_integrals_gradient.argtypes = [PSHELL, PSHELL, dptr, C.c_int, C.POINTER(PPOT),
                                                                     C.POINTER(PSHELL), dptr, pp, pp, pp, pp]
        L.libgrpp_overlap_integrals.argtypes = [PSHELL, PSHELL, dptr]
        L.libgrpp_nuclear_attraction_integrals.argtypes = [PSHELL, PSHELL, dptr, C.c_int, C.c_int, dptr, dptr]
        L.libgrpp_batch_nuclear_attraction_integrals.argtypes = [C.c_int, C.POINTER(PSHELL), C.c_int, dptr,
                                                                 C.POINTER(C.c_int), C.c_int, dptr, dptr]
        L.libgrpp_batch_full_grpp_integrals.restype = C.c_int
        L.libgrpp_batch_full_grpp_integrals.argtypes = [C.c_int, C.POINTER(PSHELL), C.c_int, dptr, C.POINTER(PGRPP),
                                                        C.c_int, dptr, dptr, dptr, dptr, C.POINTER(Stats)]
        L.libgrpp_batch_full_grpp_integrals_noalloc.restype = C.c_int
        L.libgrpp_batch_full_grpp_integrals_noalloc.argtypes = [C.c_int, C.POINTER(PSHELL), C.c_int, dptr,
                                                                C.POINTER(PGRPP), C.c_int, dptr, dptr, dptr, dptr,
                                                                C.c_void_p, C.POINTER(Stats)]
        L.libgrpp_batch_full_grpp_gradients.restype = C.c_int
        L.libgrpp_batch_full_grpp_gradients.argtypes = [C.c_int, C.POINTER(PSHELL), C.c_int, dptr, C.POINTER(PGRPP),
                                                        C.c_int, dptr, C.POINTER(Stats)]
        L.libgrpp_batch_plan_create.restype = C.c_int
        L.libgrpp_batch_plan_create.argtypes = [C.c_int, C.POINTER(PSHELL), C.c_int, dptr, C.POINTER(PGRPP), C.c_int,
                                                C.POINTER(C.c_void_p)]
        L.libgrpp_batch_plan_execute.restype = C.c_int
        L.libgrpp_batch_plan_execute.argtypes = [C.c_void_p, C.c_void_p]
        L.libgrpp_batch_plan_download.restype = C.c_int
        L.libgrpp_batch_plan_download.argtypes = [C.c_void_p, dptr, dptr, dptr, dptr]
        L.libgrpp_batch_plan_dim.restype = C.c_int
        L.libgrpp_batch_plan_dim.argtypes = [C.c_void_p]
        L.libgrpp_batch_plan_stats.argtypes = [C.c_void_p, C.POINTER(Stats)]
        L.libgrpp_batch_plan_destroy.argtypes = [C.c_void_p]
        L.libgrpp_batch_plan_features.argtypes = [C.c_void_p, C.POINTER(C.c_int), C.POINTER(C.c_float)]
        L.libgrpp_batch_plan_set_predictions.argtypes = [C.c_void_p, C.c_int, C.POINTER(C.c_float)]
        L.libgrpp_batch_gradients_plan_create.argtypes = L.libgrpp_batch_plan_create.argtypes
        L.libgrpp_batch_plan_autotune.argtypes = [C.c_void_p, C.c_int, C.c_void_p]
        L.libgrpp_batch_plan_set_tuning.argtypes = [C.c_void_p, C.c_void_p]
        L.libgrpp_batch_plan_gradient_len.restype = C.c_longlong
        L.libgrpp_batch_plan_gradient_len.argtypes = [C.c_void_p]
        L.libgrpp_batch_plan_download_gradients.argtypes = [C.c_void_p, dptr]
        L.libgrpp_batch_plan_download_gradients_atoms.argtypes = [C.c_void_p, C.c_int, C.POINTER(C.c_int), dptr]
        L.libgrpp_init()
        _lib = L
    return _lib


def last_error() -> str:
    return lib().libgrpp_last_error().decode()


def factory() -> ObjectFactory:
    return ObjectFactory(lib(), LIBGRPP_NAMES)


def _check(rc: int, what: str):
    if rc != 0:
        raise RuntimeError(f"{what} failed ({rc}): {last_error()}")


class _Objects:
    """Native shells and per-centre GRPP objects of a System (freed on close)."""

    def __init__(self, sysm: System):
        self.f = factory()
        self.sysm = sysm
        self.shells = self.f.shells(sysm.shells)
        nc = len(sysm.centers)
        self.grpps = (PGRPP * max(1, nc))()
        self.xyz = np.zeros(3 * max(1, nc))
        for i, (o, g) in enumerate(sysm.centers):
            self.grpps[i] = self.f.grpp(g) if g is not None else None   # None: a centre without GRPP
            self.xyz[3 * i:3 * i + 3] = o

    def close(self):
        self.f.free_shells(self.shells, len(self.sysm.shells))
        for i, (_, g) in enumerate(self.sysm.centers):
            if g is not None:
                self.f.free_grpp(self.grpps[i])


def batch_matrices(sysm: System, flags: int = 0):
    """(arep, so_x, so_y, so_z, stats) for all ordered shell pairs and centres of sysm."""
    o = _Objects(sysm)
    dim = sysm.dim
    out = [np.zeros((dim, dim)) for _ in range(4)]
    st = Stats()
    try:
        rc = lib().libgrpp_batch_full_grpp_integrals(len(sysm.shells), o.shells, len(sysm.centers), as_dptr(o.xyz),
                                                     o.grpps, flags, *[as_dptr(m) for m in out], C.byref(st))
        _check(rc, "libgrpp_batch_full_grpp_integrals")
    finally:
        o.close()
    return out[0], out[1], out[2], out[3], st.as_dict()


def batch_matrices_noalloc(sysm: System, flags: int = 0, stream=None):
    """libgrpp_batch_full_grpp_integrals_noalloc: the same matrices with no device allocation
    (buffers in the library's per-device scratch array; inputs too large for it raise)."""
    o = _Objects(sysm)
    dim = sysm.dim
    out = [np.zeros((dim, dim)) for _ in range(4)]
    st = Stats()
    try:
        rc = lib().libgrpp_batch_full_grpp_integrals_noalloc(len(sysm.shells), o.shells, len(sysm.centers),
                                                             as_dptr(o.xyz), o.grpps, flags,
                                                             *[as_dptr(m) for m in out], stream, C.byref(st))
        _check(rc, "libgrpp_batch_full_grpp_integrals_noalloc")
    finally:
        o.close()
    return out[0], out[1], out[2], out[3], st.as_dict()


NUCLEAR_MODEL_POINT_CHARGE, NUCLEAR_MODEL_CHARGED_BALL, NUCLEAR_MODEL_GAUSSIAN = 0, 1, 2
NUCLEAR_MODEL_FERMI, NUCLEAR_MODEL_FERMI_BUBBLE, NUCLEAR_MODEL_POINT_CHARGE_NUMERICAL = 3, 4, 5


def batch_nuclear_attraction(shells: List[Shell], nuclei, charges, model: int, params=None) -> np.ndarray:
    """libgrpp_batch_nuclear_attraction_integrals: the dim x dim nuclear attraction matrix of a shell
    list summed over the nuclei (xyz in bohr); params[n] = (R_rms,) or (c, a[, k]) per nucleus."""
    f = factory()
    sh = f.shells(shells)
    nn = len(nuclei)
    xyz = np.ascontiguousarray(np.asarray(nuclei, dtype=np.float64).reshape(-1))
    z = (C.c_int * max(1, nn))(*[int(q) for q in charges])
    p3 = np.zeros(3 * max(1, nn))
    if params is not None:
        for n, pr in enumerate(params):
            p3[3 * n:3 * n + len(pr)] = pr
    dim = sum(s.cart_size for s in shells)
    V = np.zeros((dim, dim))
    try:
        rc = lib().libgrpp_batch_nuclear_attraction_integrals(len(shells), sh, nn, as_dptr(xyz), z, model,
                                                              as_dptr(p3), as_dptr(V))
        _check(rc, "libgrpp_batch_nuclear_attraction_integrals")
    finally:
        f.free_shells(sh, len(shells))
    return V


def batch_gradients(sysm: System, flags: int = 0):
    """grad[op, 3*centre+k] (op = arep, so_x, so_y, so_z) with the centre list as atoms."""
    o = _Objects(sysm)
    dim, na = sysm.dim, len(sysm.centers)
    out = np.zeros((4, 3 * na, dim, dim))
    st = Stats()
    try:
        rc = lib().libgrpp_batch_full_grpp_gradients(len(sysm.shells), o.shells, na, as_dptr(o.xyz), o.grpps, flags,
                                                     as_dptr(out), C.byref(st))
        _check(rc, "libgrpp_batch_full_grpp_gradients")
    finally:
        o.close()
    return out, st.as_dict()


class Tuning(C.Structure):
    """libgrpp_batch_tuning_t"""
    _fields_ = [("screen_probe", C.c_int), ("contract_threads", C.c_int), ("type1_threads", C.c_int),
                ("stream_layout", C.c_int), ("screen_split", C.c_int), ("seconds_before", C.c_double),
                ("seconds_after", C.c_double)]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


class _LevelHooks:
    """Radial-grid level predictor hooks and the autotuner of a plan handle ``self.h``."""

    def autotune(self, repeats: int = 2) -> dict:
        t = Tuning()
        _check(lib().libgrpp_batch_plan_autotune(self.h, repeats, C.byref(t)), "libgrpp_batch_plan_autotune")
        return t.as_dict()

    def set_tuning(self, screen_probe: int = 1, contract_threads: int = 128, type1_threads: int = 256,
                   stream_layout: int = 0, screen_split: int = 0):
        """fixed launch shapes (screen_split 0 keeps the plan's current split)"""
        t = Tuning(screen_probe, contract_threads, type1_threads, stream_layout, screen_split, 0.0, 0.0)
        _check(lib().libgrpp_batch_plan_set_tuning(self.h, C.byref(t)), "libgrpp_batch_plan_set_tuning")

    def features(self) -> np.ndarray:
        n = C.c_int()
        lib().libgrpp_batch_plan_features(self.h, C.byref(n), None)
        f = np.zeros((n.value, NFEAT), dtype=np.float32)
        _check(lib().libgrpp_batch_plan_features(self.h, C.byref(n), f.ctypes.data_as(C.POINTER(C.c_float))),
               "libgrpp_batch_plan_features")
        return f

    def row_levels(self) -> np.ndarray:
        n = C.c_int()
        lib().libgrpp_batch_plan_row_levels(self.h, C.byref(n), None)
        v = np.zeros((n.value, 3), dtype=np.float32)
        _check(lib().libgrpp_batch_plan_row_levels(self.h, C.byref(n), v.ctypes.data_as(C.POINTER(C.c_float))),
               "libgrpp_batch_plan_row_levels")
        return v

    def set_predictions(self, pred: np.ndarray):
        pred = np.ascontiguousarray(pred, dtype=np.float32).reshape(-1, 3)
        _check(lib().libgrpp_batch_plan_set_predictions(self.h, pred.shape[0],
                                                        pred.ctypes.data_as(C.POINTER(C.c_float))),
               "libgrpp_batch_plan_set_predictions")

def _spin_hamiltonian(h, systems, spherical: bool):
    """libgrpp_batch_plan_spin_hamiltonian of a matrix plan: one complex (2n x 2n) matrix per molecule"""
    L = lib()
    L.libgrpp_batch_plan_spin_hamiltonian_len.restype = C.c_longlong
    L.libgrpp_batch_plan_spin_hamiltonian_len.argtypes = [C.c_void_p, C.c_int]
    L.libgrpp_batch_plan_spin_hamiltonian.restype = C.c_int
    L.libgrpp_batch_plan_spin_hamiltonian.argtypes = [C.c_void_p, C.c_int, dptr]
    n = L.libgrpp_batch_plan_spin_hamiltonian_len(h, int(spherical))
    flat = np.zeros(2 * max(n, 1))
    _check(L.libgrpp_batch_plan_spin_hamiltonian(h, int(spherical), as_dptr(flat)), "libgrpp_batch_plan_spin_hamiltonian")
    z = flat[0::2] + 1j * flat[1::2]
    out, o = [], 0
    for m in systems:
        d = sum((2 * s.L + 1) if spherical else s.cart_size for s in m.shells)
        out.append(z[o:o + 4 * d * d].reshape(2 * d, 2 * d))
        o += 4 * d * d
    return out


class Plan(_LevelHooks):
    """Packed, device-resident batched build: execute() runs the kernels only."""

    def __init__(self, sysm: System, flags: int = 0):
        self.sysm = sysm
        self.o = _Objects(sysm)
        self.h = C.c_void_p()
        rc = lib().libgrpp_batch_plan_create(len(sysm.shells), self.o.shells, len(sysm.centers),
                                             as_dptr(self.o.xyz), self.o.grpps, flags, C.byref(self.h))
        _check(rc, "libgrpp_batch_plan_create")
        self.dim = lib().libgrpp_batch_plan_dim(self.h)

    def execute(self, stream: int = 0):
        _check(lib().libgrpp_batch_plan_execute(self.h, C.c_void_p(stream)), "libgrpp_batch_plan_execute")

    def download(self):
        out = [np.zeros((self.dim, self.dim)) for _ in range(4)]
        _check(lib().libgrpp_batch_plan_download(self.h, *[as_dptr(m) for m in out]), "libgrpp_batch_plan_download")
        return out

    def spin_hamiltonian(self, spherical: bool = True):
        """complex spin-orbital Hamiltonian (2n x 2n, alpha functions first) of the last execute"""
        return _spin_hamiltonian(self.h, [self.sysm], spherical)[0]

    def stats(self):
        st = Stats()
        lib().libgrpp_batch_plan_stats(self.h, C.byref(st))
        return st.as_dict()

    def close(self):
        if self.h:
            lib().libgrpp_batch_plan_destroy(self.h)
            self.h = C.c_void_p()
        self.o.close()


class MoleculesPlan(_LevelHooks):
    """Many independent molecules in one device-resident plan (libgrpp_batch_molecules_plan_create)."""

    def __init__(self, systems, flags: int = 0):
        from .io import System
        self.systems = list(systems)
        shells = [s for m in self.systems for s in m.shells]
        centers = [c for m in self.systems for c in m.centers]
        self.whole = System(shells, centers)
        self.o = _Objects(self.whole)
        so = np.cumsum([0] + [len(m.shells) for m in self.systems]).astype(np.int32)
        co = np.cumsum([0] + [len(m.centers) for m in self.systems]).astype(np.int32)
        L = lib()
        L.libgrpp_batch_plan_output_len.restype = C.c_longlong
        self.h = C.c_void_p()
        rc = L.libgrpp_batch_molecules_plan_create(len(self.systems), so.ctypes.data_as(C.POINTER(C.c_int)),
                                                   co.ctypes.data_as(C.POINTER(C.c_int)), self.o.shells,
                                                   as_dptr(self.o.xyz), self.o.grpps, flags, C.byref(self.h))
        _check(rc, "libgrpp_batch_molecules_plan_create")
        self.out_len = L.libgrpp_batch_plan_output_len(self.h)

    def execute(self, stream: int = 0):
        _check(lib().libgrpp_batch_plan_execute(self.h, C.c_void_p(stream)), "libgrpp_batch_plan_execute")

    def download(self):
        """per molecule: [arep, so_x, so_y, so_z], each dim_m x dim_m"""
        flat = np.zeros(4 * self.out_len)
        _check(lib().libgrpp_batch_plan_download_flat(self.h, as_dptr(flat)), "libgrpp_batch_plan_download_flat")
        flat = flat.reshape(4, self.out_len)
        out, o = [], 0
        for m in self.systems:
            d = m.dim
            out.append([flat[k, o:o + d * d].reshape(d, d).copy() for k in range(4)])
            o += d * d
        return out

    def spin_hamiltonian(self, spherical: bool = True):
        """one complex spin-orbital Hamiltonian per molecule (see Plan.spin_hamiltonian)"""
        return _spin_hamiltonian(self.h, self.systems, spherical)

    def stats(self):
        st = Stats()
        lib().libgrpp_batch_plan_stats(self.h, C.byref(st))
        return st.as_dict()

    def close(self):
        if self.h:
            lib().libgrpp_batch_plan_destroy(self.h)
            self.h = C.c_void_p()
        self.o.close()


class GradientPlan(_LevelHooks):
    """Device-resident batched gradients (libgrpp_batch_gradients_plan_create): execute() runs the
    kernels, download() scatters grad[op, 3*centre+k] as batch_gradients returns it."""

    def __init__(self, sysm: System, flags: int = 0):
        self.sysm = sysm
        self.o = _Objects(sysm)
        self.h = C.c_void_p()
        rc = lib().libgrpp_batch_gradients_plan_create(len(sysm.shells), self.o.shells, len(sysm.centers),
                                                       as_dptr(self.o.xyz), self.o.grpps, flags, C.byref(self.h))
        _check(rc, "libgrpp_batch_gradients_plan_create")

    def execute(self, stream: int = 0):
        _check(lib().libgrpp_batch_plan_execute(self.h, C.c_void_p(stream)), "libgrpp_batch_plan_execute")

    def download(self) -> np.ndarray:
        dim, na = self.sysm.dim, len(self.sysm.centers)
        out = np.zeros((4, 3 * na, dim, dim))
        assert lib().libgrpp_batch_plan_gradient_len(self.h) == out.size
        _check(lib().libgrpp_batch_plan_download_gradients(self.h, as_dptr(out)), "libgrpp_batch_plan_download_gradients")
        return out

    def download_atoms(self, atoms) -> np.ndarray:
        """the gradients of the listed atoms only: grad[op, 3*a+k] for the a-th atom of the list"""
        dim = self.sysm.dim
        at = (C.c_int * len(atoms))(*[int(a) for a in atoms])
        out = np.zeros((4, 3 * len(atoms), dim, dim))
        _check(lib().libgrpp_batch_plan_download_gradients_atoms(self.h, len(atoms), at, as_dptr(out)),
               "libgrpp_batch_plan_download_gradients_atoms")
        return out

    def stats(self):
        st = Stats()
        lib().libgrpp_batch_plan_stats(self.h, C.byref(st))
        return st.as_dict()

    def close(self):
        if self.h:
            lib().libgrpp_batch_plan_destroy(self.h)
            self.h = C.c_void_p()
        self.o.close()


# ---------------------------------------------------------------- per-call API (libgrpp.h:123-231)
def _grad_arrays(na: int, nb: int):
    arrs = [np.zeros(na * nb) for _ in range(3)]
    pp = (dptr * 3)(*[as_dptr(a) for a in arrs])
    return arrs, pp


class Api:
    """Thin object-level mirror of the reference calls, for tests and examples."""

    def __init__(self):
        self.f = factory()
        self.L = lib()

    def type1(self, A: Shell, B: Shell, C: Tuple[float, float, float], pot: Optional[Potential]) -> np.ndarray:
        a, b = self.f.shell(A), self.f.shell(B)
        p = self.f.potential(pot) if pot is not None else None
        m = np.zeros(A.cart_size * B.cart_size)
        self.L.libgrpp_type1_integrals(a, b, darr(C), p, as_dptr(m))
        self._free(a, b, p)
        return m.reshape(A.cart_size, B.cart_size)

    def type2(self, A, B, C, pot):
        a, b, p = self.f.shell(A), self.f.shell(B), self.f.potential(pot)
        m = np.zeros(A.cart_size * B.cart_size)
        self.L.libgrpp_type2_integrals(a, b, darr(C), p, as_dptr(m))
        self._free(a, b, p)
        return m.reshape(A.cart_size, B.cart_size)

    def spin_orbit(self, A, B, C, pot):
        a, b, p = self.f.shell(A), self.f.shell(B), self.f.potential(pot)
        ms = [np.zeros(A.cart_size * B.cart_size) for _ in range(3)]
        self.L.libgrpp_spin_orbit_integrals(a, b, darr(C), p, *[as_dptr(m) for m in ms])
        self._free(a, b, p)
        return [m.reshape(A.cart_size, B.cart_size) for m in ms]

    def outercore(self, A, B, C, oc: List[Tuple[Potential, Shell]]):
        a, b = self.f.shell(A), self.f.shell(B)
        n = len(oc)
        pots = (PPOT * max(1, n))(*[self.f.potential(p) for p, _ in oc])
        shs = (PSHELL * max(1, n))(*[self.f.shell(s) for _, s in oc])
        ms = [np.zeros(A.cart_size * B.cart_size) for _ in range(4)]
        self.L.libgrpp_outercore_potential_integrals(a, b, darr(C), n, pots, shs, *[as_dptr(m) for m in ms])
        for i in range(n):
            self.f.f["delete_potential"](pots[i])
            self.f.f["delete_shell"](shs[i])
        self._free(a, b, None)
        return [m.reshape(A.cart_size, B.cart_size) for m in ms]

    def nuclear(self, A, B, C, charge: int, model: int, params=()):
        """libgrpp_nuclear_attraction_integrals (nuclear_attraction.c:54)"""
        a, b = self.f.shell(A), self.f.shell(B)
        m = np.zeros(A.cart_size * B.cart_size)
        self.L.libgrpp_nuclear_attraction_integrals(a, b, darr(C), int(charge), int(model), darr(list(params) or [0.0]),
                                                    as_dptr(m))
        self._free(a, b, None)
        return m.reshape(A.cart_size, B.cart_size)

    def full(self, A, B, C, g: GRPP):
        a, b, h = self.f.shell(A), self.f.shell(B), self.f.grpp(g)
        ms = [np.zeros(A.cart_size * B.cart_size) for _ in range(4)]
        self.L.libgrpp_full_grpp_integrals(a, b, h, darr(C), *[as_dptr(m) for m in ms])
        self.f.free_grpp(h)
        self._free(a, b, None)
        return [m.reshape(A.cart_size, B.cart_size) for m in ms]

    def full_gradient(self, A, B, C, g: GRPP, D):
        a, b, h = self.f.shell(A), self.f.shell(B), self.f.grpp(g)
        na, nb = A.cart_size, B.cart_size
        outs = [_grad_arrays(na, nb) for _ in range(4)]
        self.L.libgrpp_full_grpp_integrals_gradient(a, b, h, darr(C), darr(D), *[pp for _, pp in outs])
        self.f.free_grpp(h)
        self._free(a, b, None)
        return np.array([[x.reshape(na, nb) for x in arrs] for arrs, _ in outs])

    def _free(self, a, b, p):
        self.f.f["delete_shell"](a)
        self.f.f["delete_shell"](b)
        if p is not None:
            self.f.f["delete_potential"](p)
