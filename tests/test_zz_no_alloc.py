"""The allocation-free entry point (libgrpp_batch_full_grpp_integrals_noalloc) through the C ABI
against the allocating one (libgrpp_batch_full_grpp_integrals) on the same inputs: two calls in a
row, the second with a smaller input, then an input too large for the scratch array (refused with
a message before any kernel runs) and a further call that is still right."""
import numpy as np
import pytest

from conftest import parity_rule
from libgrpp_amd import configs

pytestmark = pytest.mark.gpu


def _same(got, want, what):
    for k in range(4):
        ok = parity_rule(got[k], want[k])
        print(f"{what} op {k}: max abs {np.abs(got[k] - want[k]).max():.3e}, bitwise {np.array_equal(got[k], want[k])}",
              flush=True)
        assert ok.all(), (what, k)


def test_noalloc_matches_allocating_entry_point(h2o):
    from libgrpp_amd import lib as G
    G.lib()
    small = configs.small_pair()
    assert small.dim < h2o.dim
    for sysm, what in [(h2o, "H2O"), (small, "small_pair"), (h2o, "H2O again")]:
        want = G.batch_matrices(sysm)
        got = G.batch_matrices_noalloc(sysm)
        _same(got[:4], want[:4], what)
        assert got[4]["n_tasks"] == want[4]["n_tasks"]
    mirrored = G.batch_matrices_noalloc(small, flags=G.BATCH_NO_SYMMETRY)
    _same(mirrored[:4], G.batch_matrices(small, flags=G.BATCH_NO_SYMMETRY)[:4], "small_pair all pairs")


def test_noalloc_refuses_inputs_beyond_the_scratch_array():
    from libgrpp_amd import lib as G
    G.lib()
    with pytest.raises(RuntimeError, match="too large for the allocation-free entry point"):
        G.batch_matrices_noalloc(configs.heavy_cluster())
    small = configs.small_pair()
    _same(G.batch_matrices_noalloc(small)[:4], G.batch_matrices(small)[:4], "small_pair after a refusal")
