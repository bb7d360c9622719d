"""A kernel launch that fails inside a batched contraction, forced from the
host: TBH_INJECT_LAUNCH_FAILURE=N (read when a context is created) lets the
context's first N batched launches through and makes the next one report a
launch failure without launching, once.  Every kernel that was launched runs
to its end.  The call reports the failure, whichever batch it hit (the first,
found when the next batch is enqueued; a middle one; the last batch's
epilogue, found when the run is flushed), and the context then contracts the
same slices to the oracle's values; a served round that fails is run again
and its requests get their oracle values."""
import os

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

POOL = 32 << 20  # a small pool: batches bounded by their arenas and descriptors


def _backend(tb, inject=None):
    key = "TBH_INJECT_LAUNCH_FAILURE"
    old = os.environ.get(key)
    if inject is not None:
        os.environ[key] = str(inject)
    try:
        return tb.HIPBackend(0, pool_bytes=POOL)
    finally:
        if old is None:
            os.environ.pop(key, None)
        else:
            os.environ[key] = old


def _slices(tb):
    g = tb.random_regular_graph(160, 3, seed=3)
    return tb.slice_graph(g, None, tb.ContractionTreeSlicer(sc_target=14, order_trials=10),
                          seed=3)


def test_failed_launch_is_reported_and_the_context_recovers(tb, orc):
    br = _slices(tb)
    ref = orc.contract_slices(br, np.float32)
    be = _backend(tb)
    try:
        be.reset_stats()
        assert np.array_equal(be.contract_slices(br, np.float32, "batched"), ref)
        total = be.stats()["launches"]
    finally:
        be.close()
    assert total > 30
    for at in (0, total // 3, total - 1):
        be = _backend(tb, at)
        try:
            with pytest.raises(tb.TbhError) as ex:
                be.contract_slices(br, np.float32, "batched")
            assert ex.value.code == -3 and "launch" in str(ex.value), (at, str(ex.value))
            # the failure was used up: the same context contracts correctly
            assert np.array_equal(be.contract_slices(br, np.float32, "batched"), ref), at
            assert np.array_equal(be.contract_slices(br, np.int32, "batched"),
                                  ref.astype(np.int32)), at
        finally:
            be.close()


def test_failed_launch_in_a_served_round_is_run_again(tb, orc):
    reqs = []
    for n, seed, sc in [(40, 11, 6), (60, 12, 8), (100, 0, 10)]:
        g = tb.random_regular_graph(n, 3, seed=seed)
        br = tb.slice_graph(g, None, tb.ContractionTreeSlicer(sc_target=sc, order_trials=20),
                            seed=seed)
        reqs.append((br, orc.mis_exact(g)))
    be = _backend(tb, 5)
    try:
        with tb.Server(be) as srv:
            tickets = [srv.submit(br) for br, _ in reqs]
            for t, (_, want) in zip(tickets, reqs):
                assert srv.wait(t) == want
            st = srv.stats()
            assert st["retried_rounds"] >= 1 and st["failed_requests"] == 0, st
    finally:
        be.close()
