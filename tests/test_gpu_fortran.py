"""The Fortran-77 ABI (include/libgrpp_fortran.h, the reference's fortran.c) and the two outercore
parts it exposes.  The wrappers must give bitwise what the C entry points give for the same
objects; the outercore parts are checked against the compiled reference's own part functions
(outercore_integrals.c:147, :277; parity rule) and must sum to the whole outercore operator; the
overlap helpers against closed forms."""
import ctypes as C
import math

import numpy as np
import pytest

from libgrpp_amd import configs, io

pytestmark = pytest.mark.gpu

I32 = C.POINTER(C.c_int32)
D = C.POINTER(C.c_double)


def _d(x):
    a = np.ascontiguousarray(np.asarray(x, dtype=np.float64))
    return a, a.ctypes.data_as(D)


def _i(x):
    a = np.ascontiguousarray(np.asarray(x, dtype=np.int32).reshape(-1))
    return a, a.ctypes.data_as(I32)


class F:
    """flat Fortran arguments of a shell / potential, kept alive with their numpy buffers"""

    def __init__(self):
        self.keep = []

    def d(self, x):
        a, p = _d(x)
        self.keep.append(a)
        return p

    def i(self, x):
        a, p = _i(x)
        self.keep.append(a)
        return p

    def shell(self, s):
        return [self.d(s.origin), self.i(s.L), self.i(len(s.alpha)), self.d(s.coeffs), self.d(s.alpha)]

    def pot(self, p, with_L=True, with_J=False):
        head = ([self.i(p.L)] if with_L else []) + ([self.i(p.J)] if with_J else [])
        return head + [self.i(len(p.alpha)), self.i(p.powers), self.d(p.coeffs), self.d(p.alpha)]


@pytest.fixture(scope="module")
def env():
    from libgrpp_amd import lib as G
    L = G.lib()
    rng = np.random.default_rng(11)
    C0 = (0.1, -0.2, 0.3)
    A = configs._shell(rng, 2, (0.7, -0.5, 1.0), 3, (0.2, 15.0))
    B = configs._shell(rng, 3, (-0.6, 0.9, -0.4), 3, (0.2, 15.0))
    g = io.GRPP(local=configs._channel(rng, 4, 0, 4), arep=[configs._channel(rng, L_, 0, 4) for L_ in range(4)],
                esop=[configs._channel(rng, L_, 0, 4) for L_ in range(1, 4)])
    oc = [(configs._channel(rng, L_, J, 4, zeta=(0.3, 20.0)), configs._shell(rng, L_, (0, 0, 0), 3, (0.5, 30.0)))
          for L_, J in ((0, 1), (1, 1), (1, 3), (1, 3), (2, 5))]
    return G, L, A, B, C0, g, oc


def test_fortran_integrals_equal_c_entry_points(env):
    G, L, A, B, C0, g, oc = env
    api = G.Api()
    f = F()
    na, nb = A.cart_size, B.cart_size
    m = np.zeros(na * nb)
    L.libgrpp_type1_integrals_(*f.shell(A), *f.shell(B), f.d(C0), *f.pot(g.local, with_L=False), m.ctypes.data_as(D))
    assert np.array_equal(m.reshape(na, nb), api.type1(A, B, C0, g.local))
    for p in g.arep:
        L.libgrpp_type2_integrals_(*f.shell(A), *f.shell(B), f.d(C0), *f.pot(p), m.ctypes.data_as(D))
        assert np.array_equal(m.reshape(na, nb), api.type2(A, B, C0, p)), p.L
    for p in g.esop:
        so = [np.zeros(na * nb) for _ in range(3)]
        L.libgrpp_spin_orbit_integrals_(*f.shell(A), *f.shell(B), f.d(C0), *f.pot(p), *[x.ctypes.data_as(D) for x in so])
        want = api.spin_orbit(A, B, C0, p)
        for k in range(3):
            assert np.array_equal(so[k].reshape(na, nb), want[k]), (p.L, k)


def test_fortran_gradients_equal_c_entry_points(env):
    G, L, A, B, C0, g, oc = env
    fac = G.factory()
    from libgrpp_amd.abi import darr
    na, nb = A.cart_size, B.cart_size
    Dpt = A.origin
    f = F()
    a, b = fac.shell(A), fac.shell(B)
    for what, pot in (("type1", g.local), ("type2", g.arep[2])):
        got = [np.zeros(na * nb) for _ in range(3)]
        args = f.pot(pot, with_L=(what == "type2"))
        getattr(L, f"libgrpp_{what}_integrals_gradient_")(*f.shell(A), *f.shell(B), f.d(C0), *args, f.d(Dpt),
                                                          *[x.ctypes.data_as(D) for x in got])
        want = [np.zeros(na * nb) for _ in range(3)]
        pp = (D * 3)(*[x.ctypes.data_as(D) for x in want])
        p = fac.potential(pot)
        getattr(L, f"libgrpp_{what}_integrals_gradient")(a, b, darr(C0), p, darr(Dpt), pp)
        fac.f["delete_potential"](p)
        for k in range(3):
            assert np.array_equal(got[k], want[k]), (what, k)
    pot = g.esop[1]
    got = [np.zeros(na * nb) for _ in range(9)]
    L.libgrpp_spin_orbit_integrals_gradient_(*f.shell(A), *f.shell(B), f.d(C0), *f.pot(pot), f.d(Dpt),
                                             *[x.ctypes.data_as(D) for x in got])
    want = [np.zeros(na * nb) for _ in range(9)]
    pps = [(D * 3)(*[want[3 * e + k].ctypes.data_as(D) for k in range(3)]) for e in range(3)]
    p = fac.potential(pot)
    L.libgrpp_spin_orbit_integrals_gradient(a, b, darr(C0), p, darr(Dpt), *pps)
    fac.f["delete_potential"](p)
    for k in range(9):
        assert np.array_equal(got[k], want[k]), k
    fac.f["delete_shell"](a)
    fac.f["delete_shell"](b)


def _parts_c(G, L, A, B, C0, entries):
    """the C part functions of this library, summed into zeroed outputs"""
    fac = G.factory()
    from libgrpp_amd.abi import darr
    na, nb = A.cart_size, B.cart_size
    out = [np.zeros(na * nb) for _ in range(4)]
    a, b = fac.shell(A), fac.shell(B)
    objs = []
    for pot, sh in entries:
        objs += [fac.potential(pot), fac.shell(sh)]
    fn = L.libgrpp_outercore_potential_integrals_part_1 if len(entries) == 1 else \
        L.libgrpp_outercore_potential_integrals_part_2
    fn(a, b, darr(C0), *objs, *[x.ctypes.data_as(D) for x in out])
    for k in range(0, len(objs), 2):
        fac.f["delete_potential"](objs[k])
        fac.f["delete_shell"](objs[k + 1])
    fac.f["delete_shell"](a)
    fac.f["delete_shell"](b)
    return np.array([x.reshape(na, nb) for x in out])


def test_outercore_parts_match_reference_and_sum_to_the_operator(env):
    G, L, A, B, C0, g, oc = env
    import oracle
    ref = oracle.reference()
    total = 0.0
    for i in range(len(oc)):
        p1 = _parts_c(G, L, A, B, C0, [oc[i]])
        total = total + p1
        if ref is not None:
            want = ref.per_call("oc1", A, B, C0, oc[i])
            err, rel, ok = io.parity_error(p1, want)
            assert ok, ("part 1", i, err, rel)
        for j in range(len(oc)):
            p2 = _parts_c(G, L, A, B, C0, [oc[i], oc[j]])
            total = total + p2
            if oc[i][0].L != oc[j][0].L or oc[i][0].J != oc[j][0].J:
                assert not p2.any(), (i, j)   # part 2 vanishes unless (L, J) agree
            if ref is not None:
                want = ref.per_call("oc2", A, B, C0, (oc[i], oc[j]))
                err, rel, ok = io.parity_error(p2, want)
                assert ok, ("part 2", i, j, err, rel)
    whole = np.asarray(G.Api().outercore(A, B, C0, oc))
    assert np.abs(total - whole).max() <= 1e-12 * max(1.0, np.abs(whole).max())


def test_fortran_outercore_parts_equal_c_parts(env):
    G, L, A, B, C0, g, oc = env
    f = F()
    na, nb = A.cart_size, B.cart_size
    ocs = lambda s: [f.i(len(s.alpha)), f.d(s.coeffs), f.d(s.alpha)]
    for i in (1, 2):
        pot, sh = oc[i]
        got = [np.zeros(na * nb) for _ in range(4)]
        L.libgrpp_outercore_potential_integrals_part_1_(*f.shell(A), *f.shell(B), f.d(C0), *f.pot(pot, with_J=True),
                                                        *ocs(sh), *[x.ctypes.data_as(D) for x in got])
        assert np.array_equal(np.array([x.reshape(na, nb) for x in got]), _parts_c(G, L, A, B, C0, [oc[i]]))
    (p1, s1), (p2, s2) = oc[2], oc[3]
    got = [np.zeros(na * nb) for _ in range(4)]
    L.libgrpp_outercore_potential_integrals_part_2_(*f.shell(A), *f.shell(B), f.d(C0), *f.pot(p1, with_J=True), *ocs(s1),
                                                    *f.pot(p2, with_J=True), *ocs(s2), *[x.ctypes.data_as(D) for x in got])
    want = _parts_c(G, L, A, B, C0, [oc[2], oc[3]])
    assert want.any()
    assert np.array_equal(np.array([x.reshape(na, nb) for x in got]), want)


def test_fortran_overlap_helpers(env):
    G, L, A, B, C0, g, oc = env
    f = F()
    # radial_gto_norm_factor_: sqrt((2L-1)!!)/sqrt(<x^L|x^L>) has the closed form of io.contracted_norm
    rng = np.random.default_rng(4)
    for Lq in range(0, 7):
        a = np.exp(rng.uniform(np.log(0.1), np.log(30.0), 4))
        c = rng.uniform(0.1, 1.0, 4)
        out = np.zeros(1)
        L.radial_gto_norm_factor_(f.i(Lq), f.i(4), f.d(c), f.d(a), out.ctypes.data_as(D))
        want = io.contracted_norm(Lq, c, a)
        assert abs(out[0] - want) <= 1e-12 * want, Lq
    # evaluate_overlap_integral_contracted_: one element of the block of libgrpp_overlap_integrals_
    na, nb = A.cart_size, B.cart_size
    S = np.zeros(na * nb)
    L.libgrpp_overlap_integrals_(*f.shell(A), *f.shell(B), S.ctypes.data_as(D))
    S = S.reshape(na, nb)
    comps = lambda l: [(x, y, l - x - y) for x in range(l, -1, -1) for y in range(l - x, -1, -1)]
    for ia, (n1, l1, m1) in enumerate(comps(A.L)):
        for ib, (n2, l2, m2) in enumerate(comps(B.L)):
            v = np.zeros(1)
            L.evaluate_overlap_integral_contracted_(f.d(A.origin), f.i(n1), f.i(l1), f.i(m1), f.i(len(A.alpha)),
                                                    f.d(A.coeffs), f.d(A.alpha), f.d(B.origin), f.i(n2), f.i(l2),
                                                    f.i(m2), f.i(len(B.alpha)), f.d(B.coeffs), f.d(B.alpha),
                                                    v.ctypes.data_as(D))
            assert v[0] == S[ia, ib], (ia, ib)
    # s-s overlap of two primitives in closed form: (pi/p)^1.5 exp(-ab/p R^2) times the norm factors
    a1, a2, R = 0.7, 1.9, np.array([0.3, -0.4, 1.2])
    v = np.zeros(1)
    L.evaluate_overlap_integral_contracted_(f.d([0, 0, 0]), f.i(0), f.i(0), f.i(0), f.i(1), f.d([1.0]), f.d([a1]),
                                            f.d(R), f.i(0), f.i(0), f.i(0), f.i(1), f.d([1.0]), f.d([a2]),
                                            v.ctypes.data_as(D))
    p = a1 + a2
    want = io.gaussian_norm_factor(0, a1) * io.gaussian_norm_factor(0, a2) * (math.pi / p) ** 1.5 * \
        math.exp(-a1 * a2 / p * float(R @ R))
    assert abs(v[0] - want) <= 1e-14 * want
