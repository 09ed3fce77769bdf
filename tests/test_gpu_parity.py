"""HIP engine (through the C ABI) against the oracle, the genuine libgrpp vectors and the
reference-held goldens.  Tolerance: the parity rule, |d| <= 1e-10 or |d|/|x| <= 1e-9 for |x| > 1
(libgrpp_amd.io.parity_error)."""
import os

import numpy as np
import pytest

from conftest import GOLDEN, block_floor_matrix, error_summary, golden_matrix, within_floor
from libgrpp_amd import configs, io

pytestmark = pytest.mark.gpu
OPS = ["arep", "so_x", "so_y", "so_z"]


@pytest.fixture(scope="module")
def G():
    from libgrpp_amd import lib
    lib.lib()
    return lib


def _check(got, want, what):
    err, rel, ok = io.parity_error(got, want)
    assert ok, f"{what}: max abs {err:.3e}, max rel(|x|>1) {rel:.3e}"


@pytest.mark.parametrize("case", ["H2O", "Cs"])
def test_batched_matches_oracle_and_goldens(G, orc, case, request):
    sysm = request.getfixturevalue(case.lower())
    print(f"case {case}", flush=True)
    M = G.batch_matrices(sysm)
    R = orc.matrices(sysm)
    for k, nm in enumerate(OPS):
        _check(M[k], R[k], f"{case} {nm} vs oracle")
        _check(np.tril(M[k]), golden_matrix(case, nm), f"{case} {nm} vs golden")


def test_batched_kcs_matches_goldens(G, kcs):
    """KCs (304 functions, l up to 4 on Cs) against the reference's goldens: per element the parity
    rule, or within 3 noise floors of its shell-pair block (the reference's own moves under
    contraction off / reassociation, and its own distance to the golden: 6.6e-10 on arep)"""
    M = G.batch_matrices(kcs)
    fl = np.load(os.path.join(GOLDEN, "KCs", "block_floor.npz"))["floor"]
    low = np.tril(np.ones((kcs.dim, kcs.dim), bool))
    for k, nm in enumerate(OPS):
        want, got, F = golden_matrix("KCs", nm)[low], np.tril(M[k])[low], block_floor_matrix(kcs, fl[k])[low]
        a, r, q = error_summary(got, want, F)
        print(f"KCs {nm}: max abs {a:.3e}, max rel {r:.3e}, worst error/floor {q:.2f}", flush=True)
        assert within_floor(got, want, F).all(), nm


def test_batched_kcs_gradients_match_goldens(G, kcs):
    """two KCs gradient goldens of the reference's test (d/dx and d/dy of the Cs atom), per element
    the parity rule or within 3 block floors of the reference's gradients"""
    fl = np.load(os.path.join(GOLDEN, "KCs", "gradient_block_floor.npz"))
    Gr, _ = G.batch_gradients(kcs)
    low = np.tril(np.ones((kcs.dim, kcs.dim), bool))
    for nm, (op, at, c) in {"arep_grad_1x": (0, 1, 0), "so_z_grad_1y": (3, 1, 1)}.items():
        got, want = np.tril(Gr[op, 3 * at + c])[low], golden_matrix("KCs", nm)[low]
        F = block_floor_matrix(kcs, fl[nm])[low]
        a, r, q = error_summary(got, want, F)
        print(f"KCs {nm}: max abs {a:.3e}, max rel {r:.3e}, worst error/floor {q:.2f}", flush=True)
        assert within_floor(got, want, F).all(), nm


def test_batched_small_pair_matches_reference_vectors(G):
    d = np.load(os.path.join(GOLDEN, "small_pair_reference.npz"))
    M = G.batch_matrices(configs.small_pair())
    for k, nm in enumerate(OPS):
        _check(M[k], d[nm], f"small_pair {nm}")


def test_per_call_integrators_match_reference_vectors(G):
    """type-1, sum of type-2, sum of 2/(2L+1) SO blocks per ordered pair (libgrpp_*_integrals)"""
    d = np.load(os.path.join(GOLDEN, "small_pair_reference.npz"))
    sysm = configs.small_pair()
    api = G.Api()
    origin, g = sysm.centers[0]
    off = sysm.offsets()
    T1, T2, SO = np.zeros_like(d["type1"]), np.zeros_like(d["type2"]), np.zeros_like(d["so"])
    for i, A in enumerate(sysm.shells):
        for j, B in enumerate(sysm.shells):
            sl = (slice(off[i], off[i] + A.cart_size), slice(off[j], off[j] + B.cart_size))
            T1[sl] = api.type1(A, B, origin, g.local)
            for p in g.arep:
                T2[sl] += api.type2(A, B, origin, p)
            for p in g.esop:
                so = api.spin_orbit(A, B, origin, p)
                for k in range(3):
                    SO[(k,) + sl] += 2.0 / (2 * p.L + 1) * so[k]
    _check(T1, d["type1"], "type1")
    _check(T2, d["type2"], "type2")
    for k in range(3):
        _check(SO[k], d["so"][k], f"so[{k}]")


def test_per_call_outercore_and_full_match_oracle(G, orc, h2o):
    api = G.Api()
    for a, b, c in [(0, 14, 2), (3, 19, 1), (10, 11, 0), (16, 17, 1), (13, 21, 2), (7, 7, 0)]:
        A, B = h2o.shells[a], h2o.shells[b]
        C, g = h2o.centers[c]
        _check(np.array(api.outercore(A, B, C, g.oc)), orc.per_call("oc", A, B, C, g.oc), f"oc {a},{b},{c}")
        _check(np.array(api.full(A, B, C, g)), orc.per_call("full", A, B, C, g), f"full {a},{b},{c}")


def test_per_call_type1_all_powers(G, orc):
    """type-1 with r^-2, r^-1 and r^0 Gaussian terms (n = 0, 1, 2: the G_n / Dawson, Boys and plain
    Gaussian branches of type1_mcmurchie_davidson.c), small and large |Q - C| (Dawson beyond
    q|QC|^2 = 12), l up to 4, against the oracle"""
    rng = np.random.default_rng(5)
    api = G.Api()
    pot = io.Potential(0, 0, [0, 1, 2, 0, 1], [1.7, -2.3, 0.9, -0.4, 3.1], [0.35, 1.2, 4.0, 12.0, 0.08])
    C = (0.0, 0.0, 0.0)
    for la, lb, ra, rb in [(0, 0, (0.1, 0, 0), (0, 0.2, 0)), (2, 1, (1.5, -0.5, 0.3), (-2.0, 1.0, 0.5)),
                           (4, 3, (3.0, 2.0, -1.0), (2.5, -3.0, 1.5)), (3, 3, (6.0, 0, 0), (0, 6.5, 0))]:
        A = configs._shell(rng, la, ra, 3, (0.2, 15.0))
        B = configs._shell(rng, lb, rb, 3, (0.2, 15.0))
        _check(np.asarray(api.type1(A, B, C, pot)), np.asarray(orc.per_call("type1", A, B, C, pot)), f"type1 ({la},{lb})")


def test_per_call_high_angular_momentum(G):
    """shells up to l = 8 and channels up to L = 5 (the sizes of the largest workloads) near the GRPP
    centre, per-call type-1 / type-2 / spin-orbit against the genuine reference's blocks.  Blocks of
    l = 8 reach ~1e3 and their small elements come out of cancelling sums: per element the parity
    rule or within 3 noise floors of the call (the reference's own moves under contraction off and
    reassociation)"""
    from golden.make_golden import high_l_calls
    d = np.load(os.path.join(GOLDEN, "high_l_percall_reference.npz"))
    api = G.Api()
    for name, A, B, C, what, arg in high_l_calls():
        f = {"type1": api.type1, "type2": api.type2, "so": api.spin_orbit}[what]
        got, want, fl = np.asarray(f(A, B, C, arg)), d[name], float(d[name + "_floor"])
        a, r, q = error_summary(got, want, fl)
        print(f"{name}: max abs {a:.3e}, max rel {r:.3e}, floor {fl:.3e}, worst error/floor {q:.2f}", flush=True)
        assert within_floor(got, want, fl).all(), name


def test_per_call_gradient_branches_match_reference(G, h2o):
    """libgrpp_full_grpp_integrals_gradient over the ACB / ACA / ACC / CCB / AAA cases"""
    d = np.load(os.path.join(GOLDEN, "h2o_pergrad_reference.npz"))
    api = G.Api()
    atoms = [c for c, _ in h2o.centers]
    pos = 0
    for (a, b, c, dd), n in zip(d["tasks"], d["lengths"]):
        got = api.full_gradient(h2o.shells[a], h2o.shells[b], atoms[c], h2o.centers[c][1], atoms[dd]).ravel()
        _check(got, d["values"][pos:pos + n], f"gradient task {(a, b, c, dd)}")
        pos += n


def test_per_term_gradient_entry_points_match_reference(G, h2o):
    """libgrpp_type1 / type2 / spin_orbit / outercore_potential_integrals_gradient against the
    genuine reference's vectors on H2O tasks (the A/B/C coincidences of the differentiated point;
    spin-orbit includes the wrapper's (2L+1)/2 rescaling, grpp_integrals_gradient.c:226-237)"""
    from golden.make_golden import PER_TERM_TASKS, per_term_gradient, per_term_operand
    d = np.load(os.path.join(GOLDEN, "per_term_gradient_reference.npz"))
    atoms = [c for c, _ in h2o.centers]
    lib, fac = G.lib(), G.factory()
    pos = 0
    for (kind, a, b, c, dd, ch), n in zip(PER_TERM_TASKS, d["lengths"]):
        got = per_term_gradient(lib, fac, kind, h2o.shells[a], h2o.shells[b], atoms[c],
                                per_term_operand(h2o, kind, c, ch), atoms[dd]).ravel()
        _check(got, d["values"][pos:pos + n], f"{kind} gradient task {(a, b, c, dd, ch)}")
        pos += n


def test_batched_gradients_match_goldens(G, h2o):
    """all 36 H2O gradient matrices of the reference's test (outercore included)"""
    Gr, _ = G.batch_gradients(h2o)
    for op, nm in enumerate(OPS):
        for at in range(3):
            for k, c in enumerate("xyz"):
                _check(np.tril(Gr[op, 3 * at + k]), golden_matrix("H2O", f"{nm}_grad_{at}{c}"), f"{nm}_grad_{at}{c}")


def test_gradients_translational_invariance(G, h2o):
    """sum over all atoms of d/dR_atom vanishes (every shell and centre sits on an atom)"""
    Gr, _ = G.batch_gradients(h2o)
    for op in range(4):
        for k in range(3):
            s = Gr[op, k::3].sum(axis=0)
            assert np.abs(s).max() < 1e-11, (op, k)


def test_full_size_properties(G):
    """heavy_cluster at full size (100,352 tasks): repeated execution is bit-identical, results are
    finite, and evaluating every ordered pair agrees with the mirrored (A <= B) evaluation within the
    formulation's conditioning noise (relative 1e-4 of the largest element)"""
    sysm = configs.heavy_cluster()
    plan = G.Plan(sysm)
    plan.execute()
    a = plan.download()
    plan.execute()
    b = plan.download()
    plan.close()
    for k in range(4):
        assert np.array_equal(a[k], b[k])
        assert np.isfinite(a[k]).all()
    full = G.batch_matrices(sysm, flags=G.BATCH_NO_SYMMETRY)
    for k, nm in enumerate(OPS):
        scale = np.abs(a[k]).max()
        assert np.abs(full[k] - a[k]).max() <= 1e-4 * scale, nm
        sgn = 1 if k == 0 else -1
        assert np.abs(a[k] - sgn * a[k].T).max() <= 1e-4 * scale, nm


def test_heavy_cluster_gradients_translational_invariance(G):
    """full heavy_cluster (8 atoms, l <= 4 plus the l + 1 shells of the derivatives): the sum over
    all atoms of d/dR_atom vanishes to rounding, relative to the largest gradient element"""
    Gr, st = G.batch_gradients(configs.heavy_cluster())
    for op in range(4):
        scale = np.abs(Gr[op]).max()
        for k in range(3):
            assert np.abs(Gr[op, k::3].sum(axis=0)).max() <= 1e-13 * scale, (op, k)


def test_ragged_molecules_in_one_plan(G):
    """uneven batch of six small molecules (2-5 atoms, 1-3 GRPP centres with L up to 5, shells up to
    l = 4) in one plan: every molecule's matrices equal those of its own plan"""
    mols = configs.ragged_molecules(seed=3, nmol=6, atoms=(2, 5), centers=(1, 3), lmax_shell=4, nprim=(1, 3),
                                    shells_per_atom=(1, 3), alpha=(0.3, 20.0))
    plan = G.MoleculesPlan(mols)
    plan.execute()
    both = plan.download()
    st = plan.stats()
    plan.close()
    assert st["n_tasks"] == sum(len(m.shells) ** 2 * len(m.centers) for m in mols)
    for m, got in zip(mols, both):
        alone = G.batch_matrices(m)
        for k in range(4):
            assert np.abs(got[k] - alone[k]).max() <= 1e-12 * max(1.0, np.abs(alone[k]).max()), k


def test_invalid_input_reports_error(G):
    """shells beyond the supported angular momentum are rejected on the host with a message"""
    bad = io.System([io.Shell(11, (0, 0, 0), [1.0], [1.0])],
                    [((0, 0, 0), io.GRPP(local=io.Potential(0, 0, [2], [1.0], [1.0])))])
    with pytest.raises(RuntimeError, match="supported range"):
        G.batch_matrices(bad)
