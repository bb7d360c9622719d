"""contract_slices on the GPU (PLAIN and BATCHED) against the oracle and
independent MIS solvers, bit-exact."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _slices(tb, g, sc_target, weights=None, seed=0, trials=20):
    return tb.slice_graph(g, weights, tb.ContractionTreeSlicer(sc_target=sc_target,
                                                               order_trials=trials), seed=seed)


@pytest.mark.parametrize("mode", ["batched", "plain"])
def test_regular3_n100_unweighted(tb, orc, backend, mode):
    """3-regular n=100 seed 0, sliced to sc_target 10: per-slice results equal
    the oracle's and the maximum equals the exact MIS."""
    g = tb.random_regular_graph(100, 3, seed=0)
    br = _slices(tb, g, 10)
    got = backend.contract_slices(br, np.float32, mode)
    ref = orc.contract_slices(br, np.float32)
    assert np.array_equal(got, ref)
    assert int(got.max()) == orc.mis_exact(g)


@pytest.mark.parametrize("mode", ["batched", "plain"])
def test_weighted_int32(tb, orc, backend, mode):
    g = tb.random_ksg(8, 8, 0.8, seed=3)
    rng = np.random.default_rng(3)
    w = rng.integers(1, 101, g.nv).astype(np.int32)
    br = _slices(tb, g, 8, weights=w)
    got = backend.contract_slices(br, np.int32, mode)
    ref = orc.contract_slices(br, np.int32)
    assert got.dtype == np.int32
    assert np.array_equal(got, ref)
    if g.nv <= 60:
        assert int(got.max()) == int(orc.mis_bruteforce(g, w.astype(np.float64)))


def test_int32_equals_float32(tb, backend):
    g = tb.random_regular_graph(60, 3, seed=4)
    br = _slices(tb, g, 8)
    a = backend.contract_slices(br, np.float32)
    b = backend.contract_slices(br, np.int32)
    assert np.array_equal(a.astype(np.int64), b.astype(np.int64))


def test_unsliced_whole_graph(tb, orc, backend):
    g = tb.random_regular_graph(60, 3, seed=1)
    code = tb.initialize_code(g, order_trials=20, seed=1)
    br = [tb.SlicedBranch(tb.MISProblem(g), code, 0)]
    for mode in ("batched", "plain"):
        assert backend.solve_slice(br[0], np.float32, mode) == orc.mis_exact(g)


def test_empty_and_tiny_slices(tb, orc, backend):
    empty = tb.SlicedBranch(tb.MISProblem(tb.SimpleGraph(0)), None, 7)
    one = tb.SlicedBranch(tb.MISProblem(tb.SimpleGraph(1), np.array([5], np.float32)),
                          tb.CompressedEinsum([[0]], [], 0), 2)
    g2 = tb.SimpleGraph(3, [[0, 1]])  # an edge plus an isolated vertex
    code2 = tb.CompressedEinsum([[0], [1], [2], [0, 1]], [],
                                tb.ContractionTree(tb.ContractionTree(0, 3),
                                                   tb.ContractionTree(1, 2)))
    two = tb.SlicedBranch(tb.MISProblem(g2), code2, 1)
    br = [empty, one, two]
    for mode in ("batched", "plain"):
        got = backend.contract_slices(br, np.float32, mode)
        assert got.tolist() == [7.0, 7.0, 3.0]
        assert np.array_equal(got, orc.contract_slices(br, np.float32))


@pytest.mark.parametrize("mode", ["batched", "plain"])
def test_closed_forms(tb, backend, mode):
    cases = []
    for n in (5, 12, 31):  # cycles: floor(n/2)
        cases.append((tb.SimpleGraph(n, [[i, (i + 1) % n] for i in range(n)]), n // 2))
    for n in (6, 17):  # paths: ceil(n/2)
        cases.append((tb.SimpleGraph(n, [[i, i + 1] for i in range(n - 1)]), (n + 1) // 2))
    a, b = 4, 7  # complete bipartite: max(a, b)
    cases.append((tb.SimpleGraph(a + b, [[i, a + j] for i in range(a) for j in range(b)]), 7))
    pet = [[i, (i + 1) % 5] for i in range(5)] + [[i, i + 5] for i in range(5)] + \
          [[5 + i, 5 + (i + 2) % 5] for i in range(5)]
    cases.append((tb.SimpleGraph(10, pet), 4))  # Petersen
    L = 7  # grid L x L: ceil(L^2 / 2)
    grid = [[i * L + j, i * L + j + 1] for i in range(L) for j in range(L - 1)] + \
           [[i * L + j, (i + 1) * L + j] for i in range(L - 1) for j in range(L)]
    cases.append((tb.SimpleGraph(L * L, grid), (L * L + 1) // 2))
    for g, expect in cases:
        code = tb.initialize_code(g, order_trials=10)
        br = tb.SlicedBranch(tb.MISProblem(g), code, 0)
        assert backend.solve_slice(br, np.float32, mode) == expect


def test_order_invariance(tb, orc, backend):
    """Different trees over the same leaves give identical bits."""
    g = tb.random_regular_graph(50, 3, seed=9)
    vals = set()
    for seed in range(4):
        code = tb.initialize_code(g, order_trials=4 + seed, seed=seed)
        br = tb.SlicedBranch(tb.MISProblem(g), code, 0)
        vals.add(float(backend.solve_slice(br, np.float32)))
    assert vals == {float(orc.mis_exact(g))}


def test_slicing_invariant(tb, orc, backend):
    """max_i(slice_i + r_i) equals the unsliced contraction."""
    g = tb.random_regular_graph(80, 3, seed=5)
    w = np.random.default_rng(5).integers(1, 50, g.nv).astype(np.float32)
    br = _slices(tb, g, 9, weights=w)
    assert len(br) > 1
    whole = tb.SlicedBranch(tb.MISProblem(g, w), tb.initialize_code(g, 20), 0)
    assert backend.contract_slices(br).max() == backend.solve_slice(whole) == \
        orc.solve_slice(whole)


@pytest.mark.parametrize("dtype", [np.float32, np.int32])
def test_tile_variants_bit_exact(tb, orc, backend, dtype):
    """A workload whose batched plan has outer-product tiles of every
    layout (2 x 8, 4 x 4, 8 x 2 per thread), tiles
    with outer reduced labels (looped outside LDS), GEMM, direct and tiled
    steps: every slice's optimum equals the oracle's in both modes."""
    g = tb.random_regular_graph(160, 3, seed=5)
    w = None
    if dtype == np.int32:
        w = np.random.default_rng(5).integers(1, 101, g.nv).astype(np.int32)
    br = _slices(tb, g, 18, weights=w, seed=5, trials=30)
    steps = [s for b in br if b.code is not None for s in tb.plan_steps(b, dtype)]
    ew = [s for s in steps if s["kind"] == 0]
    assert {s["outer_product"] for s in ew} == {0, 1, 2, 3}  # every tile layout
    assert any(s["outer_reduced"] > 0 for s in ew)
    assert any(s["direct"] for s in ew) and any(not s["direct"] for s in ew)
    ref = orc.contract_slices(br, dtype)
    for mode in ("batched", "plain"):
        assert np.array_equal(backend.contract_slices(br, dtype, mode), ref), mode
    # direct wave tiles with reductions (every eligible step direct), no
    # direct steps at all, every launch on one stream (GEMMs not beside the
    # EW steps of their level), tiled EW steps prefetched into registers
    # (alone and with every EW step tiled), persistent grids, both GEMM tile
    # rules besides the default, chain rotations under the rank caps alone
    # (no cost test), and claims of one tile (no sub-tensor reuse within a
    # claim) and of sixteen
    t0 = backend.tuning()
    try:
        for kw in ({"direct_max_reduce": 12}, {"direct_max_reduce": -1}, {"concurrent": 0},
                   {"ew_prefetch": 1}, {"ew_prefetch": 1, "direct_max_reduce": -1},
                   {"ew_grid_mult": 1, "gemm_grid_mult": 1}, {"gemm_tiles": 0},
                   {"gemm_tiles": 2}, {"chain_cost": 0}, {"ew_chunk": 1}, {"ew_chunk": 16}):
            backend.set_tuning(**dict(t0, **kw))
            assert np.array_equal(backend.contract_slices(br, dtype, "batched"), ref), kw
    finally:
        backend.set_tuning(**t0)


def test_fractional_float32_weights(tb, orc, backend):
    """Float32 weights that are not integers (outside the integer-valued
    contract).  Each tree node adds two values once per term and takes maxima,
    and max commutes with monotone rounding, so every lowering that keeps the
    slice's tree forms the oracle's sums bit for bit: PLAIN, and BATCHED with
    chain re-association off (chain_big = 30).  Re-association regroups the
    sums of a chain, (x + s1) + s2 against x + (s1 + s2); with positive
    weights each configuration's sum of at most nv terms then moves by at
    most 2 (nv - 1) 2^-24 sum(w), and so does the maximum over them."""
    g = tb.random_regular_graph(160, 3, seed=5)
    w = (np.random.default_rng(6).random(g.nv) * 10).astype(np.float32)
    br = _slices(tb, g, 18, weights=w, seed=5, trials=30)
    ref = orc.contract_slices(br, np.float32)
    assert np.any(ref != np.round(ref))
    assert np.array_equal(backend.contract_slices(br, np.float32, "plain"), ref)
    t0 = backend.tuning()
    try:
        backend.set_tuning(**dict(t0, chain_big=30))
        assert np.array_equal(backend.contract_slices(br, np.float32, "batched"), ref)
    finally:
        backend.set_tuning(**t0)
    got = backend.contract_slices(br, np.float32, "batched")
    tol = 2.0 * (g.nv - 1) * 2.0 ** -24 * float(w.astype(np.float64).sum())
    assert np.max(np.abs(got.astype(np.float64) - ref)) <= tol


@pytest.mark.parametrize("dtype", [np.float32, np.int32])
def test_gemm_output_stores_bit_exact(tb, orc, backend, dtype):
    """GEMM nodes that write their output in the consumer's label order (row
    quads, or an LDS-staged tile read back along the output's lowest
    positions, in one or two passes): every slice's optimum equals the
    oracle's."""
    g = tb.random_regular_graph(200, 3, seed=3)
    w = None
    if dtype == np.int32:
        w = np.random.default_rng(3).integers(1, 101, g.nv).astype(np.int32)
    br = tb.slice_graph(g, w, tb.ContractionTreeSlicer(sc_target=22, order_trials=30), seed=3)
    stores = {s["gemm_store"] for b in br if b.code is not None
              for s in tb.plan_steps(b, dtype) if s["kind"] == 1}
    assert stores == {0, 1, 2}
    ref = orc.contract_slices(br, dtype)
    assert np.array_equal(backend.contract_slices(br, dtype, "batched"), ref)


def test_slice_directory_on_gpu(tb, orc, backend, tmp_path):
    """A slice directory in the reference's on-disk format (src/io.jl),
    loaded with load_all_finished, contracts on the GPU to the oracle's
    per-slice optima and the exact MIS (as test/slice.jl:22-35)."""
    g = tb.random_regular_graph(100, 3, seed=0)
    br = _slices(tb, g, 10)
    tb.save_slices(str(tmp_path), br)
    back = tb.load_all_finished(str(tmp_path))
    ref = orc.contract_slices(back, np.float32)
    for mode in ("batched", "plain"):
        assert np.array_equal(backend.contract_slices(back, np.float32, mode), ref)
    assert int(ref.max()) == orc.mis_exact(g)


def test_autotune_keeps_results(tb, orc):
    """The autotuner installs a tuning under which the results are unchanged;
    out-of-range tunings are refused."""
    be = tb.HIPBackend(0, pool_bytes=2 << 30)
    try:
        g = tb.random_regular_graph(100, 3, seed=0)
        br = _slices(tb, g, 10)
        t0 = be.tuning()
        best, ms = be.autotune(br, np.float32, max_trials=6)
        assert ms > 0 and best == be.tuning()
        assert best["chain_side"] < best["chain_big"]
        assert np.array_equal(be.contract_slices(br, np.float32), orc.contract_slices(br, np.float32))
        for bad in ({"ew_chunk": 0}, {"ew_per_cu": 9}, {"direct_max_reduce": 13},
                    {"concurrent": 2}, {"first_batch": 0.0}, {"gemm_tiles": 3},
                    {"ew_grid_mult": 0}, {"ew_prefetch": 2}, {"chain_cost": -1}):
            with pytest.raises(tb.TbhError):
                be.set_tuning(**bad)
        be.set_tuning(**t0)
        assert be.tuning() == t0
    finally:
        be.close()


def test_autotune_cache(tb, orc, tmp_path):
    """A cached autotune result is installed without timing: the second call
    returns the first call's tuning and time, on a fresh context too; another
    slice set misses and adds its own entry."""
    path = str(tmp_path / "tuning.json")
    g = tb.random_regular_graph(100, 3, seed=0)
    br = _slices(tb, g, 10)
    be = tb.HIPBackend(0, pool_bytes=2 << 30)
    try:
        best, ms = be.autotune(br, np.float32, max_trials=5, cache=path)
    finally:
        be.close()
    be = tb.HIPBackend(0, pool_bytes=2 << 30)
    try:
        be.set_tuning(**dict(be.tuning(), ew_chunk=1))
        best2, ms2 = be.autotune(br, np.float32, max_trials=5, cache=path)
        assert best2 == best and ms2 == ms and be.tuning() == best
        assert np.array_equal(be.contract_slices(br, np.float32), orc.contract_slices(br, np.float32))
        other = _slices(tb, tb.random_regular_graph(80, 3, seed=3), 8)
        be.autotune(other, np.float32, max_trials=3, cache=path)
        import json
        assert len(json.load(open(path))) == 2
    finally:
        be.close()


def test_autotune_fractional_float32_weights(tb, orc):
    """Non-integer Float32 weights: chain re-association could move a result
    by a rounding step, so the autotuner leaves the chain parameters where
    the caller set them and every candidate reproduces the first run's bits.
    Started with re-association off, the installed tuning still has it off
    and the results equal the oracle's bit for bit."""
    be = tb.HIPBackend(0, pool_bytes=2 << 30)
    try:
        g = tb.random_regular_graph(120, 3, seed=7)
        w = (np.random.default_rng(7).random(g.nv) * 10).astype(np.float32)
        br = _slices(tb, g, 14, weights=w, seed=7, trials=20)
        ref = orc.contract_slices(br, np.float32)
        assert np.any(ref != np.round(ref))
        t0 = be.tuning()
        be.set_tuning(**dict(t0, chain_big=30))
        best, ms = be.autotune(br, np.float32, max_trials=8)
        assert ms > 0 and best == be.tuning()
        assert (best["chain_big"], best["chain_side"], best["chain_reduce"], best["chain_cost"]) == \
            (30, t0["chain_side"], t0["chain_reduce"], t0["chain_cost"])
        assert np.array_equal(be.contract_slices(br, np.float32, "batched"), ref)
    finally:
        be.close()
