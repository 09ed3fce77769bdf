"""Host-side checks that need no GPU: the C-ABI library loads and exports every symbol the
public headers declare; input parsers, fixture I/O and the seeded generators."""
import ctypes as C
import os
import re

import numpy as np
import pytest

from conftest import GOLDEN, ROOT
from libgrpp_amd import configs, io

LIB = os.path.join(ROOT, "libgrpp_amd", "libgrpp_amd.so")


def declared_functions():
    names = set()
    for h in ("libgrpp.h", "libgrpp_amd.h", "libgrpp_fortran.h"):
        text = open(os.path.join(ROOT, "include", h)).read()
        text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
        for m in re.finditer(r"\b(libgrpp_\w+|evaluate_\w+_|radial_gto_\w+_)\s*\(", text):
            names.add(m.group(1))
    return sorted(names)


def test_library_exports_every_declared_symbol():
    if not os.path.exists(LIB):
        pytest.skip("libgrpp_amd.so not built")
    lib = C.CDLL(LIB)
    missing = [n for n in declared_functions() if not hasattr(lib, n)]
    assert not missing, missing
    assert hasattr(lib, "libgrpp_params")
    assert len(declared_functions()) >= 45


def test_library_objects_without_gpu():
    """object constructors are host code: usable with no device present"""
    if not os.path.exists(LIB):
        pytest.skip("libgrpp_amd.so not built")
    from libgrpp_amd.abi import LIBGRPP_NAMES, ObjectFactory
    lib = C.CDLL(LIB)
    f = ObjectFactory(lib, LIBGRPP_NAMES)
    s = f.shell(io.Shell(3, (0.1, 0.2, 0.3), [1.0, 0.5], [2.0, 0.3]))
    assert s.contents.cart_size == 10
    got = [tuple(s.contents.cart_list[3 * i:3 * i + 3]) for i in range(10)]
    assert got == [(3, 0, 0), (2, 1, 0), (2, 0, 1), (1, 2, 0), (1, 1, 1), (1, 0, 2), (0, 3, 0), (0, 2, 1),
                   (0, 1, 2), (0, 0, 3)]
    p = f.potential(io.Potential(2, 0, [2, 2, 1], [1.0, 0.0, -2.0], [1.0, 2.0, 3.0]))
    assert p.contents.num_primitives == 2        # zero coefficient dropped (potential.c:34)
    g = f.grpp(io.GRPP(local=io.Potential(1, 0, [2], [1.0], [1.0])))
    f.free_grpp(g)
    f.f["delete_potential"](p)
    f.f["delete_shell"](s)


def test_parsers_on_reference_inputs(h2o, cs):
    assert h2o.dim == 65 and len(h2o.shells) == 22 and len(h2o.centers) == 3
    assert cs.dim == 183 and len(cs.centers) == 1
    _, g = h2o.centers[0]
    assert g.local is not None and len(g.arep) == 1 and len(g.esop) == 1 and len(g.oc) == 1
    assert g.oc[0][0].J == 1 and g.oc[0][1].L == 0


def test_lower_triangle_roundtrip(tmp_path):
    m = np.random.default_rng(0).standard_normal((7, 7))
    m[2, 1] = 1e-15
    p = str(tmp_path / "m.txt")
    io.write_lower_triangle(p, m)
    r = io.read_lower_triangle(p)
    t = np.tril(m)
    t[2, 1] = 0.0
    assert np.allclose(r, t, rtol=1e-15, atol=0)


def test_parity_rule():
    ref = np.array([0.5, 2.0, 3e5])
    assert io.parity_error(ref + [9e-11, 1.9e-9, 2.9e-4], ref)[2]
    assert not io.parity_error(ref + [2e-10, 0, 0], ref)[2]
    assert not io.parity_error(ref + [0, 0, 4e-4], ref)[2]


def test_generators_deterministic():
    a, b = configs.small_pair(0), configs.small_pair(0)
    assert a == b and configs.small_pair(1) != a
    assert a.dim == 28 and configs.n_tasks(a) == 100
    h = configs.heavy_cluster()
    assert len(h.shells) == 112 and h.dim == 552 and configs.n_tasks(h) == 100352
    xyz = np.array([c for c, _ in h.centers])
    d = np.linalg.norm(xyz[:, None] - xyz[None], axis=-1) + 1e9 * np.eye(8)
    assert d.min() >= 3.0 and xyz.min() >= 0 and xyz.max() <= 12.0


def test_contracted_normalisation():
    """normalised contractions have unit self-overlap of the x^L component / (2L-1)!!"""
    for L in range(5):
        c, a = [0.3, 0.9, 0.4], [5.0, 0.9, 0.11]
        f = io.contracted_norm(L, c, a)
        assert io.contracted_norm(L, [x * f for x in c], a) == pytest.approx(1.0, rel=1e-13)


def test_batch_plans_refuse_reserved_flags():
    """LIBGRPP_BATCH_LEARNED_GRID and LIBGRPP_BATCH_GRADIENT are refused with an error before any
    device work (host code: no GPU needed), for all three plan constructors"""
    if not os.path.exists(LIB):
        pytest.skip("libgrpp_amd.so not built")
    from libgrpp_amd.abi import LIBGRPP_NAMES, PGRPP, PSHELL, ObjectFactory
    lib = C.CDLL(LIB)
    lib.libgrpp_last_error.restype = C.c_char_p
    f = ObjectFactory(lib, LIBGRPP_NAMES)
    sh = (PSHELL * 1)(f.shell(io.Shell(0, (0.0, 0.0, 0.0), [1.0], [1.0])))
    g = (PGRPP * 1)(f.grpp(io.GRPP(local=io.Potential(0, 0, [2], [1.0], [1.0]))))
    xyz = (C.c_double * 3)(0.0, 0.0, 0.0)
    plan = C.c_void_p()
    for flags in (1, 2, 3, 1 | 16):
        rc = lib.libgrpp_batch_plan_create(1, sh, 1, xyz, g, flags, C.byref(plan))
        assert rc != 0 and not plan.value
        assert b"unsupported flags" in lib.libgrpp_last_error()
        rc = lib.libgrpp_batch_gradients_plan_create(1, sh, 1, xyz, g, flags, C.byref(plan))
        assert rc != 0 and not plan.value
        so, co = (C.c_int * 2)(0, 1), (C.c_int * 2)(0, 1)
        rc = lib.libgrpp_batch_molecules_plan_create(1, so, co, sh, xyz, g, flags, C.byref(plan))
        assert rc != 0 and not plan.value
    f.free_grpp(g[0])
    f.f["delete_shell"](sh[0])
