"""Branching tables of regions (branching_table(p, TensorNetworkSolver(),
region), the table step of optimal_branches, src/branch.jl:79): every
region's boundary configurations contracted in one batched submit, then
pruned (mis_compactify!) and enumerated on the GPU (tbh_region_tables).
Checked against brute-force enumeration of every region configuration
(oracle.branching_table_bruteforce, itself pinned by hand-worked tables):
the alpha tensor, the reduced tensor and every row, bit for bit.  Exact
parity with OptimalBranchingMIS, which is not available here, is unpinned."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _check(tb, orc, tables, items, dtype):
    for t, (p, vs) in zip(tables, items):
        ov, alpha, red, rows = orc.branching_table_bruteforce(p.g, p.weights, vs)
        assert t.open_vertices == ov
        got = np.asarray(t.alpha, np.float64)
        if dtype == np.int32:
            got[t.alpha <= tb.TBH_I32_NEG_INF] = -np.inf
        assert got.tolist() == red.tolist(), (vs, got, red)
        assert t.table == rows, vs


def test_hand_worked_tables(tb, orc, backend):
    g = tb.SimpleGraph(5, np.array([[0, 1], [1, 2], [2, 3], [3, 4]], np.int32))
    t = tb.branching_table(tb.MISProblem(g), [1, 2, 3], backend=backend)
    assert t.open_vertices == [1, 3]
    assert t.alpha.tolist() == [1.0, -np.inf, -np.inf, 2.0]
    assert t.table == [[0b010], [0b101]]
    g = tb.SimpleGraph(3, np.array([[0, 1], [1, 2]], np.int32))
    w = np.array([1, 5, 1], np.int32)
    t = tb.branching_table(tb.MISProblem(g, w), [0, 1], np.int32, backend=backend)
    assert t.alpha.tolist() == [1, 5] and t.table == [[0b01], [0b10]]


@pytest.mark.parametrize("dtype,weighted", [(np.float32, False), (np.float32, True),
                                            (np.int32, True)])
def test_regions_against_bruteforce(tb, orc, backend, dtype, weighted):
    """Neighbourhood regions of 8 to 20 vertices in random 3-regular and
    Erdos-Renyi graphs, all in one call."""
    rng = np.random.default_rng(17)
    items = []
    for k in range(24):
        n = int(rng.integers(40, 200)) // 2 * 2
        g = tb.random_regular_graph(n, 3, seed=k) if k % 2 == 0 else \
            tb.erdos_renyi(n, 3.0 / (n - 1), seed=k)
        w = None
        if weighted:
            w = rng.integers(1, 20, g.nv).astype(dtype)
        n_max = int(rng.integers(8, 21))
        vs = tb.neighbor_region(g, [int(rng.integers(g.nv))], n_max)
        if len(tb.open_vertices(g, vs)) > 12:
            vs = vs[:12]
        items.append((tb.MISProblem(g, w), vs))
    print(f"[branching] {len(items)} regions, sizes {[len(v) for _, v in items]}", flush=True)
    tables = tb.branching_tables(items, dtype, backend=backend)
    _check(tb, orc, tables, items, dtype)
    # the same regions one call each give the same tables
    for t, it in zip(tables[:4], items[:4]):
        one = tb.branching_table(it[0], it[1], dtype, backend=backend)
        assert one.table == t.table and np.array_equal(one.alpha, t.alpha)


def test_region_errors(tb, backend):
    g = tb.random_regular_graph(60, 3, seed=1)
    with pytest.raises(ValueError):
        tb.branching_table(tb.MISProblem(g), list(range(25)), backend=backend)
    w = np.full(g.nv, 0.5, np.float32)
    with pytest.raises(tb.TbhError):
        tb.branching_table(tb.MISProblem(g, w), tb.neighbor_region(g, [0], 8),
                           backend=backend)
