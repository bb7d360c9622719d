"""Host producers: generators and the slicer keep the invariants the
contraction relies on (they stand in for Graphs.jl, TreeSA and slice_bfs)."""
import numpy as np
import pytest


def test_random_regular(tb):
    g = tb.random_regular_graph(300, 3, seed=2)
    deg = np.bincount(g.edges.reshape(-1), minlength=g.nv)
    assert g.nv == 300 and g.ne() == 450 and np.all(deg == 3)
    assert np.array_equal(tb.random_regular_graph(300, 3, seed=2).edges, g.edges)
    assert not np.array_equal(tb.random_regular_graph(300, 3, seed=3).edges, g.edges)


def test_ksg_and_er(tb):
    g = tb.random_ksg(10, 10, 0.8, seed=3)
    assert g.nv == 80
    assert g.ne() > 0
    e = tb.erdos_renyi(50, 0.1, seed=1)
    assert e.nv == 50 and 0 < e.ne() < 50 * 49 // 2


@pytest.mark.parametrize("sc_target", [6, 9])
def test_slices_are_valid(tb, orc, sc_target):
    g = tb.random_regular_graph(60, 3, seed=8)
    br = tb.slice_graph(g, None, tb.ContractionTreeSlicer(sc_target=sc_target, order_trials=8))
    assert len(br) >= 1
    for b in br:
        if b.code is None:
            assert b.p.g.nv == 0
            continue
        assert tb.complexity(b).sc <= sc_target
        # leaves are the vertex tensors then the edge tensors of the subgraph
        ixs = b.code.ixs
        assert ixs[:b.p.g.nv] == [[v] for v in range(b.p.g.nv)]
        assert sorted(map(tuple, ixs[b.p.g.nv:])) == sorted(map(tuple, b.p.g.edges.tolist()))
    assert orc.contract_slices(br, np.float32).max() == orc.mis_exact(g)


def test_weighted_slices_keep_integer_offsets(tb, orc):
    g = tb.random_ksg(7, 7, 0.8, seed=3)
    w = np.random.default_rng(3).integers(1, 101, g.nv).astype(np.int32)
    br = tb.slice_graph(g, w, tb.ContractionTreeSlicer(sc_target=6, order_trials=8))
    assert all(float(b.r) == int(b.r) for b in br)
    assert orc.contract_slices(br, np.int32).max() == int(orc.mis_bruteforce(g, w.astype(np.float64)))


def test_usecuda_false_raises(tb):
    g = tb.random_regular_graph(10, 3, seed=0)
    br = [tb.SlicedBranch(tb.MISProblem(g), tb.initialize_code(g, 2), 0)]
    with pytest.raises(NotImplementedError):
        tb.contract_slices(br, np.float32, False)
