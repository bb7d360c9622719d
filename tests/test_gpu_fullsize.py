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
  holds for every regular graph).
Every one of the 766 per-slice optima is also pinned: the oracle computed
them once (tests/golden/regular3_n300_sc25.json, about 450 s on 8 CPU threads,
written by make_fixtures.py --n300 with the slice set's digest), and the test
regenerates the slice set, checks that its digest matches, and compares the
GPU's results with the stored optima bit for bit."""
import importlib.util
import json
import os

import numpy as np
import pytest

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def _fixtures_module():
    spec = importlib.util.spec_from_file_location("make_fixtures",
                                                  os.path.join(GOLDEN, "make_fixtures.py"))
    mod = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(mod)
    return mod

pytestmark = pytest.mark.gpu

N, SEED = 300, 2


@pytest.fixture(scope="module")
def n300(tb):
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


def test_full_size_solve_properties(tb, orc, backend, n300):
    g, br = n300
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


def test_full_size_slicing_invariant(tb, orc, backend, n300):
    g, br25 = n300
    br22 = tb.slice_graph(g, None, tb.ContractionTreeSlicer(sc_target=22, order_trials=200),
                          seed=SEED)
    assert len(br22) > len(br25)
    assert max(tb.sc(b) for b in br22 if b.code is not None) <= 22
    r22 = backend.contract_slices(br22, np.float32, "batched")
    assert r22.max() == backend.contract_max(br25, np.float32, "batched")
    idx = _least_work(tb, br22, 32)
    ref = orc.contract_slices([br22[k] for k in idx], np.float32, nthreads=16)
    assert np.array_equal(r22[idx], ref)


@pytest.mark.parametrize("mode", ["batched", "plain"])
def test_full_size_every_slice_against_oracle(tb, backend, n300, mode):
    """All 766 slices of the n=300 workload, large ones included (2^25-entry
    tensors, split-K GEMMs, LDS-staged GEMM outputs, outer-reduced EW
    steps), bit-exact against the oracle's stored per-slice optima."""
    g, br = n300
    fx = json.load(open(os.path.join(GOLDEN, "regular3_n300_sc25.json")))
    assert fx["params"] == {"n": N, "seed": SEED, "sc_target": 25, "order_trials": 200}
    assert len(br) == fx["slices"]
    assert _fixtures_module().sliceset_digest(br) == fx["digest"]
    want = np.array(fx["per_slice"], np.float32)
    got = backend.contract_slices(br, np.float32, mode)
    bad = np.flatnonzero(got != want)
    assert bad.size == 0, [(int(k), float(got[k]), float(want[k]), fx["tc"][k]) for k in bad[:10]]
    assert got.max() == fx["mis"]
