"""Captured single-problem steps (kmig_set_graph): the first resident run of a
plan is enqueued eagerly and captured; later runs of the same plan replay the
graph (trace preparation on a parallel branch, tuple preparation, the migrate
launch and its partial sums).  Every replay must give the eager run's CIG bit
for bit; replays are interleaved with other eager work and synchronisation,
and a change of shot range or data runs eagerly and captures again."""
import numpy as np
import pytest

from oracle import oracle as O

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def small():
    from kirch_datuming_amd import configs
    return configs.small(threads=8)


def _bits(a):
    return a.view(np.uint32)


@pytest.mark.parametrize("cmp", [0, 1])
def test_graph_replays_match_eager_runs(small, monkeypatch, cmp):
    from kirch_datuming_amd import kmig
    from tests.test_gpu_cmp1 import _vz_cmp1
    monkeypatch.delenv("KMIG_SPLIT", raising=False)
    if cmp:
        kv, args = _vz_cmp1()
        data, dax, st_, sd_, sax, rt_, rd_, rax = args
    else:
        kv = dict(small.kv, direct_mute=1, direct_mute_slope=2.0, direct_mute_taper=0.05)
        data, dax, st_, sd_, sax, rt_, rd_, rax = small.args()
    p = kmig.parse_params(dict(kv, variant="resident"), int(dax[1][0]), dax[1][1], dax[1][2])
    eager = kmig.Migrator(p)
    graph = kmig.Migrator(p)
    for m in (eager, graph):
        m.set_tables(0, st_, sd_, sax)
        m.set_tables(1, rt_, rd_, rax)
        m.set_data(data, dax)
    graph.set_graph(True)
    eager.run()
    eager.sync()
    ref = eager.cig()
    flags = []
    for rep in range(4):
        print(f"graph cmp={cmp}: run {rep}", flush=True)
        graph.run()
        graph.sync()
        flags.append(graph.stats()["graph"])
        assert np.array_equal(_bits(graph.cig()), _bits(ref)), f"run {rep}"
        eager.run()  # other eager work between the replays
        eager.sync()
    assert flags == [0, 1, 1, 1]
    # a shot range is another plan: eager, captured, then replayed
    ns = int(dax[2][0])
    eager.run((1, ns - 1))
    eager.sync()
    sub = eager.cig()
    for rep in range(2):
        graph.run((1, ns - 1))
        graph.sync()
        assert graph.stats()["graph"] == rep
        assert np.array_equal(_bits(graph.cig()), _bits(sub))
    # new data: eager again, then replays of the new plan
    data2 = (0.5 * data).astype(np.float32)
    for m in (eager, graph):
        m.set_data(data2, dax)
    eager.run()
    eager.sync()
    ref2 = eager.cig()
    for rep in range(2):
        graph.run()
        graph.sync()
        assert graph.stats()["graph"] == rep
        assert np.array_equal(_bits(graph.cig()), _bits(ref2))
    assert O.rel_l2(ref2, 0.5 * ref) < 1e-6
    eager.close()
    graph.close()
