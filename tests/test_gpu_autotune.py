"""The autotuner changes launch shapes only: every shape it may pick gives bitwise the matrices of
the default shapes (screening probe width 1/2/4 and 4/8/16 waves per tile, contraction and type-1
workgroups of 64/128/256 threads, any of the three stream layouts), and a tuned plan, eager or graph-replayed, still does."""
import numpy as np
import pytest

from libgrpp_amd import configs

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("workload", ["small_pair", "heavy_cluster"])
def test_every_launch_shape_gives_the_same_matrices(workload):
    from libgrpp_amd import lib as G
    sysm = getattr(configs, workload)()
    plan = G.Plan(sysm)
    plan.execute()
    want = plan.download()
    for shape in ((4, 64, 64, 1, 16), (2, 256, 128, 0, 8), (1, 128, 256, 1, 4), (4, 256, 64, 0, 16),
                  (2, 64, 256, 1, 8), (1, 128, 128, 0, 16), (1, 128, 128, 0, 4), (1, 128, 128, 0, 0),
                  (1, 128, 256, 2, 4), (2, 64, 128, 2, 16)):
        plan.set_tuning(*shape)
        plan.execute()
        got = plan.download()
        assert all(np.array_equal(a, b) for a, b in zip(want, got)), shape
    plan.close()


def test_autotuned_graph_plan_keeps_results():
    from libgrpp_amd import lib as G
    sysm = configs.heavy_cluster()
    eager = G.Plan(sysm)
    eager.execute()
    want = eager.download()
    eager.close()
    plan = G.Plan(sysm, G.BATCH_GRAPH)
    plan.execute()                       # captures the default shapes
    t = plan.autotune(repeats=1)         # eager timing runs; the graph is recaptured afterwards
    print("tuned:", t, flush=True)
    assert t["screen_probe"] in (1, 2, 4) and t["contract_threads"] in (64, 128, 256)
    assert t["screen_split"] in (4, 8, 16)
    assert t["seconds_after"] > 0.0
    for _ in range(2):
        plan.execute()
        assert all(np.array_equal(a, b) for a, b in zip(want, plan.download()))
    with pytest.raises(RuntimeError, match="tuning"):
        plan.set_tuning(5, 128, 256, 0)
    with pytest.raises(RuntimeError, match="tuning"):
        plan.set_tuning(1, 128, 256, 0, 2)
    with pytest.raises(RuntimeError, match="tuning"):
        plan.set_tuning(1, 128, 256, 0, 2)
    plan.close()
