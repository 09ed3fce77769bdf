"""Nuclear attraction on the device (libgrpp_nuclear_attraction_integrals and the batched
libgrpp_batch_nuclear_attraction_integrals) against the genuine reference's vectors
(tests/golden/nuclear_reference.npz, six nuclear models, l <= 6) and the CPU restatement."""
import os

import numpy as np
import pytest

from conftest import GOLDEN, error_summary, within_floor

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def api():
    from libgrpp_amd.lib import Api
    return Api()   # loads the library, which calls libgrpp_init


def test_per_call_matches_reference(api):
    from golden.make_golden import nuclear_calls
    d = np.load(os.path.join(GOLDEN, "nuclear_reference.npz"))
    worst = 0.0
    for name, A, B, Cp, Z, model, prm in nuclear_calls():
        print("nuclear", name, flush=True)
        got = api.nuclear(A, B, Cp, Z, model, prm)
        want, fl = d[name], float(d[name + "_floor"])
        ok = within_floor(got, want, fl)
        assert ok.all(), (name, float(np.abs(got - want).max()), fl)
        worst = max(worst, error_summary(got, want, fl)[2])
    print("worst (|d| - 1e-10) / floor outside the parity rule", worst)


def test_per_call_transpose(api):
    """<A|V|B> = <B|V|A>^T for every model, on pairs the expansion about the nucleus leaves well
    conditioned (the reference's own builds agree to 1e-13 relative there)"""
    from golden.make_golden import nuclear_calls
    for name, A, B, Cp, Z, model, prm in nuclear_calls():
        if not name.endswith(("_21", "_44")):
            continue
        ab = api.nuclear(A, B, Cp, Z, model, prm)
        ba = api.nuclear(B, A, Cp, Z, model, prm)
        assert np.abs(ab - ba.T).max() <= 1e-12 * np.abs(ab).max(), name


@pytest.mark.parametrize("model", range(6))
def test_batch_matches_reference(model, api):
    """the dim x dim matrix of a three-nucleus molecule (l <= 4) against the reference's per-call
    blocks summed over the nuclei, within the parity rule or 3 noise floors per block; blocks
    i < j mirrored exactly; the same blocks as the per-call entry point's"""
    from golden.make_golden import NUC_MOLECULE_PARAMS, nuclear_molecule
    from libgrpp_amd.lib import batch_nuclear_attraction
    shells, nuclei, Z = nuclear_molecule()
    d = np.load(os.path.join(GOLDEN, "nuclear_molecule_reference.npz"))
    prm = NUC_MOLECULE_PARAMS[model]
    print("batch nuclear model", model, flush=True)
    V = batch_nuclear_attraction(shells, nuclei, Z, model, prm)
    W, F = d[f"V{model}"], d[f"floor{model}"]
    ns = len(shells)
    off = np.cumsum([0] + [s.cart_size for s in shells])
    Fd = np.zeros_like(W)
    for i in range(ns):
        for j in range(ns):
            Fd[off[i]:off[i + 1], off[j]:off[j + 1]] = max(F[i, j], F[j, i])
    ok = within_floor(V, W, Fd)
    assert ok.all(), (model, float(np.abs(V - W).max()), error_summary(V, W, Fd))
    print("worst (|d| - 1e-10) / floor outside the parity rule", error_summary(V, W, Fd)[2])
    # blocks i < j are written once and mirrored; a diagonal block is one evaluation of all its
    # elements, symmetric to rounding
    sc = np.abs(W).max()
    for i in range(ns):
        for j in range(ns):
            blk, mir = V[off[i]:off[i + 1], off[j]:off[j + 1]], V[off[j]:off[j + 1], off[i]:off[i + 1]]
            if i != j:
                assert (blk == mir.T).all(), (i, j)
            else:
                assert np.abs(blk - blk.T).max() <= 1e-13 * sc + 3 * F[i, i], i
    # the per-call entry point gives the same blocks up to the order of the sum over nuclei
    for i, j in ((0, 4), (5, 8), (9, 11), (4, 11)):
        blk = sum(api.nuclear(shells[i], shells[j], c, Z[n], model, prm[n] if prm else ())
                  for n, c in enumerate(nuclei))
        assert np.abs(blk - V[off[i]:off[i + 1], off[j]:off[j + 1]]).max() <= 1e-12 * sc, (i, j)


def test_errors(api):
    """shells beyond l = 6 in a numeric model, and unknown models, are refused"""
    from libgrpp_amd import configs
    from libgrpp_amd.lib import batch_nuclear_attraction
    rng = np.random.default_rng(0)
    s7 = configs._shell(rng, 7, (0.0, 0.0, 0.0), 1, (1.0, 2.0))
    s0 = configs._shell(rng, 0, (0.0, 0.0, 1.0), 1, (1.0, 2.0))
    with pytest.raises(RuntimeError, match="l <= 6"):
        batch_nuclear_attraction([s7, s0], [(0.0, 0.0, 0.0)], [10], 3, [(0.4, 0.06)])
    with pytest.raises(RuntimeError, match="unknown nuclear model"):
        batch_nuclear_attraction([s0], [(0.0, 0.0, 0.0)], [10], 9, [(0.4,)])
    # the point charge runs the analytic kernels, which take l = 7
    V = batch_nuclear_attraction([s7, s0], [(0.0, 0.0, 0.0)], [10], 0)
    assert np.isfinite(V).all() and np.abs(V).max() > 0
