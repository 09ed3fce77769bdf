"""A plan created with LIBGRPP_BATCH_GRAPH captures its whole build (both streams) into one hipGraph
on the first execute and replays it afterwards.  Every replay must give bitwise the matrices of the
eager build, also with other GPU work and a new row order of the radial tables between replays."""
import numpy as np
import pytest

from libgrpp_amd import configs

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("workload", ["small_pair", "heavy_cluster", "ragged"])
def test_graph_replays_equal_eager_build(workload):
    from libgrpp_amd import lib as G
    if workload == "small_pair":
        sysm = configs.small_pair()
        make = lambda flags: G.Plan(sysm, flags)
    elif workload == "heavy_cluster":
        sysm = configs.heavy_cluster()
        make = lambda flags: G.Plan(sysm, flags)
    else:
        mols = configs.ragged_molecules(seed=0, nmol=200)[150:156]
        make = lambda flags: G.MoleculesPlan(mols, flags)
    flat = lambda ms: [m for x in ms for m in (x if isinstance(x, list) else [x])]
    eager = make(0)
    eager.execute()
    want = flat(eager.download())
    graph = make(G.BATCH_GRAPH)
    n = graph.features().shape[0]
    rng = np.random.default_rng(3)
    for rep in range(4):
        graph.execute()
        got = flat(graph.download())
        assert all(np.array_equal(a, b) for a, b in zip(want, got)), rep
        st = graph.stats()
        assert st["device_seconds"] > 0 and st["n_radial_points"] > 0
        eager.execute()              # the program's own eager work between replays
        if rep == 1:                 # a new row order is read by the next replay
            graph.set_predictions(rng.random((n, 3)).astype(np.float32))
    eager.close()
    graph.close()
