"""Parity at the full size of the benchmarked workload: a 3-regular graph
with n=300 (seed 2) sliced to sc_target 25, 766 slices of up to 2^36 work
each.  The oracle cannot contract all of them in test time, so the whole
solve is checked through properties that hold at any size, and the slices of
least work are checked bit-exact against the oracle:
- two slicings of the same graph (sc_target 25 and 22: different trees and
  slice sets) give the same maximum, which is the slicing invariant of
  test/slice.jl:31-33 at full size;
- BATCHED equals PLAIN slice by slice (two executors, different plans);
- Int32 equals Float32 on unit weights;
- one-stream launches equal two-stream launches;
- the maximum lies between a greedy independent set and n/2 (a bound that
  holds for every regular graph)."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

N, SEED = 300, 2


@pytest.fixture(scope="module")
def c3(tb):
    g = tb.random_regular_graph(N, 3, seed=SEED)
    br = tb.slice_graph(g, None, tb.ContractionTreeSlicer(sc_target=25, order_trials=200),
                        seed=SEED)
    return g, br


def _greedy_mis(g):
    """Minimum-degree greedy independent set (a lower bound of the MIS)."""
    adj = [set() for _ in range(g.nv)]
    for u, v in g.edges:
        adj[u].add(v)
        adj[v].add(u)
    alive = set(range(g.nv))
    size = 0
    while alive:
        v = min(alive, key=lambda x: (len(adj[x] & alive), x))
        alive -= adj[v] | {v}
        size += 1
    return size


def _least_work(tb, br, k):
    tcs = np.array([tb.complexity(b).tc if b.code is not None else np.inf for b in br])
    return np.argsort(tcs, kind="stable")[:k]


def test_full_size_solve_properties(tb, orc, backend, c3):
    g, br = c3
    assert len(br) > 500
    assert max(tb.sc(b) for b in br if b.code is not None) <= 25
    batched = backend.contract_slices(br, np.float32, "batched")
    mis = batched.max()
    assert _greedy_mis(g) <= mis <= N // 2
    assert np.array_equal(backend.contract_slices(br, np.float32, "plain"), batched)
    got32 = backend.contract_slices(br, np.int32, "batched")
    assert got32.dtype == np.int32
    assert np.array_equal(got32.astype(np.float32), batched)
    t0 = backend.tuning()
    try:
        backend.set_tuning(**dict(t0, concurrent=0))
        assert np.array_equal(backend.contract_slices(br, np.float32, "batched"), batched)
    finally:
        backend.set_tuning(**t0)
    # the slices of least work, bit-exact against the oracle
    idx = _least_work(tb, br, 16)
    ref = orc.contract_slices([br[k] for k in idx], np.float32, nthreads=16)
    assert np.array_equal(batched[idx], ref)


def test_full_size_slicing_invariant(tb, orc, backend, c3):
    g, br25 = c3
    br22 = tb.slice_graph(g, None, tb.ContractionTreeSlicer(sc_target=22, order_trials=200),
                          seed=SEED)
    assert len(br22) > len(br25)
    assert max(tb.sc(b) for b in br22 if b.code is not None) <= 22
    r22 = backend.contract_slices(br22, np.float32, "batched")
    assert r22.max() == backend.contract_max(br25, np.float32, "batched")
    idx = _least_work(tb, br22, 32)
    ref = orc.contract_slices([br22[k] for k in idx], np.float32, nthreads=16)
    assert np.array_equal(r22[idx], ref)
