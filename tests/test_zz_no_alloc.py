"""tbh_contract_slices_noalloc: contract_slices in PLAIN mode with every
device byte in the library's static per-GPU scratch array (no allocation).
Called through the C ABI and checked against tbh_contract_slices(PLAIN) and
the oracle, bit for bit, on two calls in a row (the second smaller, over the
first call's leftovers in the array), and refusing a slice too large for the
array before any work."""
import ctypes as C

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _noalloc(be, branches, npdt):
    from tensorbranching_jl_amd import _lib
    from tensorbranching_jl_amd.api import _Marshal, _dtype_code
    code, npdt = _dtype_code(npdt)
    m = _Marshal(branches, npdt)
    out = np.zeros(max(1, len(branches)), npdt)
    rc = _lib.tbh().tbh_contract_slices_noalloc(be.handle, code, m.arr, m.n, out.ctypes.data)
    return rc, out[:len(branches)]


def test_noalloc_equals_plain_two_calls(tb, orc):
    from tensorbranching_jl_amd import _lib
    be = tb.HIPBackend(0, pool_bytes=1 << 30)
    try:
        # first call: 3-regular n=100 seed 0 sliced to sc_target 10, Float32
        g1 = tb.random_regular_graph(100, 3, seed=0)
        br1 = tb.slice_graph(g1, None, tb.ContractionTreeSlicer(sc_target=10, order_trials=20),
                             seed=0)
        print(f"[noalloc] call 1: {len(br1)} slices (n=100, sc 10, f32)", flush=True)
        rc, got1 = _noalloc(be, br1, np.float32)
        assert rc == 0, _lib.tbh().tbh_last_error().decode()
        ref1 = be.contract_slices(br1, np.float32, "plain")
        assert np.array_equal(got1, ref1)
        assert np.array_equal(got1, orc.contract_slices(br1, np.float32))
        assert int(got1.max()) == orc.mis_exact(g1)
        # second call, smaller: 6 slices of a weighted Int32 graph
        g2 = tb.random_regular_graph(60, 3, seed=4)
        w2 = np.random.default_rng(4).integers(1, 101, g2.nv).astype(np.int32)
        br2 = tb.slice_graph(g2, w2, tb.ContractionTreeSlicer(sc_target=7, order_trials=20),
                             seed=4)[:6]
        assert len(br2) < len(br1)
        print(f"[noalloc] call 2: {len(br2)} slices (n=60, sc 7, i32)", flush=True)
        rc, got2 = _noalloc(be, br2, np.int32)
        assert rc == 0, _lib.tbh().tbh_last_error().decode()
        assert got2.dtype == np.int32
        assert np.array_equal(got2, be.contract_slices(br2, np.int32, "plain"))
        assert np.array_equal(got2, orc.contract_slices(br2, np.int32))
        # the pool was never used by the no-allocation calls
        be.reset_stats()
        rc, got1b = _noalloc(be, br1, np.float32)
        assert rc == 0 and np.array_equal(got1b, ref1)
        assert be.stats()["pool_peak_bytes"] == 0
    finally:
        be.close()


def test_noalloc_refuses_slice_beyond_scratch(tb):
    from tensorbranching_jl_amd import _lib
    be = tb.HIPBackend(0, pool_bytes=1 << 30)
    try:
        g = tb.random_regular_graph(200, 3, seed=8)  # unsliced: sc about 24
        code = tb.initialize_code(g, order_trials=10, seed=8)
        big = [tb.SlicedBranch(tb.MISProblem(g), code, 0)]
        assert tb.slice_footprint(big[0]) > (64 << 20)
        print("[noalloc] oversized slice (unsliced n=200)", flush=True)
        rc, _ = _noalloc(be, big, np.float32)
        assert rc == _lib.TBH_ENOMEM
        assert "64 MiB" in _lib.tbh().tbh_last_error().decode()
        # an empty request and an empty-graph slice need no device memory
        rc, _ = _noalloc(be, [], np.float32)
        assert rc == 0
        empty = tb.SlicedBranch(tb.MISProblem(tb.SimpleGraph(0)), None, 3)
        rc, got = _noalloc(be, [empty], np.float32)
        assert rc == 0 and got[0] == 3
    finally:
        be.close()
