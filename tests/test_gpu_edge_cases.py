"""Edge cases of the batched and per-call paths: centres without a GRPP (NULL), operators with only
some of their parts, no centres at all, a shell sitting on the GRPP centre together with one that
does not, the largest supported angular momentum, zero-coefficient potential terms, and a molecule
without centres inside a many-molecule plan.  Each is compared with the oracle (parity rule) or
with an equivalent input whose answer is known."""
import os

import numpy as np
import pytest

from conftest import GOLDEN
from libgrpp_amd import configs, io

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def G():
    from libgrpp_amd import lib
    lib.lib()
    return lib


def _ok(got, want, what):
    err, rel, ok = io.parity_error(np.asarray(got), np.asarray(want))
    assert ok, f"{what}: max abs {err:.3e}, rel {rel:.3e}"


def test_null_centre_is_skipped(G):
    s = configs.small_pair()
    with_null = io.System(s.shells, s.centers + [((1.0, -1.0, 0.5), None)])
    a, b = G.batch_matrices(s), G.batch_matrices(with_null)
    for k in range(4):
        assert np.array_equal(a[k], b[k]), k


def test_no_centres_gives_zero_matrices(G):
    s = configs.small_pair()
    M = G.batch_matrices(io.System(s.shells, []))
    for k in range(4):
        assert M[k].shape == (s.dim, s.dim) and not M[k].any(), k


def test_partial_operators_match_oracle(G, orc):
    """local only, semilocal only, spin-orbit only, outercore only (H2O's operator split apart)"""
    h2o = io.load_system(os.path.join(GOLDEN, "H2O"))
    parts = {"local": lambda g: io.GRPP(local=g.local), "arep": lambda g: io.GRPP(arep=g.arep),
             "so": lambda g: io.GRPP(esop=g.esop), "oc": lambda g: io.GRPP(oc=g.oc)}
    for name, f in parts.items():
        s = io.System(h2o.shells, [(p, f(g)) for p, g in h2o.centers])
        M, R = G.batch_matrices(s), orc.matrices(s)
        for k in range(4):
            _ok(M[k], R[k], f"{name} op {k}")
        if name in ("local", "arep"):
            assert not any(M[k].any() for k in (1, 2, 3)), name   # no spin-orbit part


def test_shell_on_centre_and_off_centre(G, orc):
    """one-centre blocks (A = B = C), mixed blocks (A = C != B) and two-centre blocks in one plan"""
    rng = np.random.default_rng(21)
    C0 = (0.2, 0.1, -0.3)
    g = io.GRPP(local=configs._channel(rng, 3, 0, 3), arep=[configs._channel(rng, L, 0, 3) for L in range(3)],
                esop=[configs._channel(rng, L, 0, 3) for L in range(1, 3)])
    shells = [configs._shell(rng, l, C0, 3, (0.3, 10.0)) for l in (0, 1, 2)] + \
             [configs._shell(rng, l, (1.1, -0.4, 0.9), 3, (0.3, 10.0)) for l in (0, 2)]
    s = io.System(shells, [(C0, g)])
    M, R = G.batch_matrices(s), orc.matrices(s)
    for k in range(4):
        _ok(M[k], R[k], f"op {k}")


def test_largest_angular_momentum_type1(G, orc):
    """l = 10 (the largest basis shell the engine takes) against an s shell, local term"""
    rng = np.random.default_rng(8)
    pot = configs._channel(rng, 3, 0, 2)
    A = configs._shell(rng, 10, (0.3, -0.2, 0.4), 2, (0.5, 5.0))
    B = configs._shell(rng, 0, (-0.4, 0.5, 0.1), 2, (0.5, 5.0))
    C0 = (0.0, 0.0, 0.0)
    got = np.asarray(G.Api().type1(A, B, C0, pot))
    want = np.asarray(orc.per_call("type1", A, B, C0, pot))
    assert got.shape == (66, 1)
    _ok(got, want, "type1 l=10")


def test_zero_coefficient_terms(G, orc):
    """a channel with a zero coefficient term equals the channel without it"""
    s = configs.small_pair()
    (C0, g) = s.centers[0]
    p = g.arep[1]
    p0 = io.Potential(p.L, p.J, p.powers + [2], p.coeffs + [0.0], p.alpha + [3.0])
    A, B = s.shells[2], s.shells[9]
    api = G.Api()
    a, b = np.asarray(api.type2(A, B, C0, p)), np.asarray(api.type2(A, B, C0, p0))
    _ok(b, a, "zero-coefficient term")
    _ok(b, np.asarray(orc.per_call("type2", A, B, C0, p0)), "vs oracle")


def test_molecule_without_centres_in_a_molecules_plan(G):
    mols = configs.ragged_molecules(seed=5, nmol=3, atoms=(2, 3), centers=(1, 2), lmax_shell=2, nprim=(1, 3))
    bare = io.System(mols[1].shells, [])
    plan = G.MoleculesPlan([mols[0], bare, mols[2]])
    plan.execute()
    out = plan.download()
    plan.close()
    assert not any(m.any() for m in out[1])
    for i in (0, 2):
        alone = G.batch_matrices(mols[i])
        for k in range(4):
            assert np.abs(out[i][k] - alone[k]).max() <= 1e-12 * max(1.0, np.abs(alone[k]).max())


def test_concurrent_per_call_integrators_match_sequential(G):
    """host threads calling libgrpp_full_grpp_integrals at once (the reference driver's OpenMP
    pattern, eval_integrals.c:78-123) on shell pairs of very different LDS needs (l = 0..6 against a
    GRPP with L <= 5, spin-orbit and outercore s/p/d): every result bitwise equal to the same call made
    alone; the per-kernel dynamic-LDS limits are set once per device, not per call"""
    from concurrent.futures import ThreadPoolExecutor
    s = configs.high_l_cluster(seed=1, natom=2, nprim=3, alpha=(0.5, 20.0), lmax_shell=6)
    C, g = s.centers[0]
    pairs = [(a, b) for a in range(len(s.shells)) for b in range(len(s.shells)) if (a + 3 * b) % 5 == 0]
    api = G.Api()
    seq = [api.full(s.shells[a], s.shells[b], C, g) for a, b in pairs]
    jobs = [k for k in range(len(pairs)) for _ in range(3)]
    with ThreadPoolExecutor(16) as ex:
        res = list(ex.map(lambda k: (k, api.full(s.shells[pairs[k][0]], s.shells[pairs[k][1]], C, g)), jobs))
    print(f"{len(jobs)} concurrent calls over {len(pairs)} pairs, l up to {max(x.L for x in s.shells)}", flush=True)
    for k, got in res:
        for op in range(4):
            assert np.array_equal(got[op], seq[k][op]), (pairs[k], op)


@pytest.mark.parametrize("workload", ["small_pair", "heavy_cluster", "ThO"])
def test_wave_per_part_contraction_equals_block_contraction(G, workload, monkeypatch):
    """the contraction of parts of <= 64 Cartesian pairs by one wave each (four per workgroup), forced
    on inputs below its size threshold, gives bitwise the matrices of the block-per-part kernel.  On
    ThO (outercore s/p/d/f, shells to l = 6) the wave's LDS does not hold the outercore L-products or
    vectors of the larger ket shells, so the outercore terms also take their unstaged path there"""
    if workload == "ThO":
        sysm = io.load_system(os.path.join(GOLDEN, "ThO"))
    else:
        sysm = configs.WORKLOADS[workload]()
    monkeypatch.setenv("GRPP_CTW_MIN_PARTS", "1000000000")
    block = G.batch_matrices(sysm)
    monkeypatch.setenv("GRPP_CTW_MIN_PARTS", "0")
    wave = G.batch_matrices(sysm)
    for k in range(4):
        assert np.array_equal(block[k], wave[k]), k
