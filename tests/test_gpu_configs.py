"""Parity at the full sizes of the benchmark workloads that the oracle
finishes in seconds: one unsliced contraction of a 3-regular n=200 graph
(sc 26), max-plus GEMMs over that contraction's node shape range
({2^4 .. 2^12}^3), and a 40 x 40 King's subgraph (filling 0.8, Int32 weights
in [1, 100], sc_target 25).  Bit-exact against the oracle, per slice."""
import numpy as np
import pytest

from conftest import rand_tropical

pytestmark = pytest.mark.gpu


def test_unsliced_regular3_n200(tb, orc, backend):
    g = tb.random_regular_graph(200, 3, seed=1)
    code = tb.initialize_code(g, order_trials=100, seed=1)
    br = [tb.SlicedBranch(tb.MISProblem(g), code, 0)]
    assert tb.sc(br[0]) >= 24
    ref = orc.contract_slices(br, np.float32)
    for mode in ("batched", "plain"):
        assert np.array_equal(backend.contract_slices(br, np.float32, mode), ref), mode


def test_gemm_shape_range_2e4_2e12(backend, orc):
    """M, N, K drawn from {2^4 .. 2^12} (the unsliced n=200 nodes' shapes), with the
    CPU reference's work bounded to 2^27 MACs per product."""
    rng = np.random.default_rng(2)
    done = 0
    while done < 24:
        M, N, K = (int(2 ** rng.integers(4, 13)) for _ in range(3))
        if M * N * K > 1 << 27:
            continue
        A = rand_tropical(rng, (M, K), np.float32)
        B = rand_tropical(rng, (K, N), np.float32)
        assert np.array_equal(backend.maxplus_gemm(A, B), orc.maxplus_gemm(A, B)), (M, N, K)
        done += 1


def test_ksg40_int32(tb, orc, backend):
    g = tb.random_ksg(40, 40, 0.8, seed=3)
    w = np.random.default_rng(3).integers(1, 101, g.nv).astype(np.int32)
    br = tb.slice_graph(g, w, tb.ContractionTreeSlicer(sc_target=25, order_trials=50), seed=3)
    assert max(tb.sc(b) for b in br if b.code is not None) <= 25
    ref = orc.contract_slices(br, np.int32)
    got = backend.contract_slices(br, np.int32, "batched")
    assert got.dtype == np.int32
    assert np.array_equal(got, ref)
    # Float32 on the same integer weights gives the same values (< 2^24)
    got32 = backend.contract_slices(br, np.float32, "batched")
    assert np.array_equal(got32.astype(np.int64), got.astype(np.int64))
