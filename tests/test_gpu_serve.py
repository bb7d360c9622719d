"""Continuous-batching scheduler (tbh_serve_start / tbh_submit / tbh_poll /
tbh_wait / tbh_cancel): request values against the oracle and exact MIS, the
same value served alone and mixed with other requests, per-request errors."""
import time

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _sliced(tb, g, sc_target, weights=None, seed=0):
    return tb.slice_graph(g, weights, tb.ContractionTreeSlicer(sc_target=sc_target,
                                                               order_trials=20), seed=seed)


def _requests(tb, orc):
    """(branches, dtype, expected MIS) for a mix of graphs and sizes."""
    out = []
    for n, seed, sc in [(40, 11, 6), (60, 12, 8), (100, 0, 10), (80, 13, 12)]:
        g = tb.random_regular_graph(n, 3, seed=seed)
        out.append((_sliced(tb, g, sc, seed=seed), np.float32, orc.mis_exact(g)))
    g = tb.random_ksg(7, 7, 0.8, seed=3)
    w = np.random.default_rng(3).integers(1, 101, g.nv).astype(np.int32)
    br = _sliced(tb, g, 8, weights=w)
    out.append((br, np.int32, int(orc.contract_slices(br, np.int32).max())))
    return out


@pytest.fixture
def be(tb):
    b = tb.HIPBackend(0, pool_bytes=2 << 30)
    yield b
    b.close()


def test_requests_match_oracle(tb, orc, be):
    reqs = _requests(tb, orc)
    with tb.Server(be) as srv:
        tickets = [srv.submit(br, dt) for br, dt, _ in reqs]
        for t, (br, dt, want) in zip(tickets, reqs):
            got = srv.wait(t)
            assert got == want, (t, got, want)
            assert np.asarray(got).dtype == dt


@pytest.mark.parametrize("cap", [0, 7, 64])
def test_alone_equals_mixed(tb, orc, be, cap):
    """A request's value is identical served alone and mixed into rounds with
    other requests (small caps split its slices over several rounds)."""
    reqs = _requests(tb, orc)
    alone = []
    with tb.Server(be, max_resident_slices=cap) as srv:
        for br, dt, _ in reqs:
            alone.append(srv.wait(srv.submit(br, dt)))
        # burst: every request in flight at once, twice over
        tickets = [srv.submit(br, dt) for _ in range(2) for br, dt, _ in reqs]
        mixed = [srv.wait(t) for t in tickets]
    for i, (br, dt, want) in enumerate(reqs):
        assert alone[i] == want
        assert mixed[i] == alone[i] and mixed[i + len(reqs)] == alone[i]
    # and equal to the synchronous batched call
    for (br, dt, _), a in zip(reqs, alone):
        assert be.contract_slices(br, dt, "batched").max() == a


def test_poll_reports_once_with_times(tb, orc, be):
    g = tb.random_regular_graph(60, 3, seed=5)
    br = _sliced(tb, g, 8, seed=5)
    with tb.Server(be) as srv:
        t = srv.submit(br)
        for _ in range(2000):
            st, val, t0, t1 = srv.poll(t)
            if st in ("done", "failed"):
                break
            assert st in ("queued", "running")
            time.sleep(0.005)
        assert st == "done" and val == orc.mis_exact(g)
        assert t1 >= t0 > 0
        with pytest.raises(tb.TbhError):
            srv.poll(t)  # a finished request is reported once


def test_cancel_queued(tb, be):
    """cancel drops a request that has not been admitted; an admitted one
    refuses with TBH_EBUSY."""
    from tensorbranching_jl_amd import _lib
    g = tb.random_regular_graph(100, 3, seed=6)
    big = _sliced(tb, g, 10, seed=6)
    small = _sliced(tb, tb.random_regular_graph(30, 3, seed=7), 6, seed=7)
    with tb.Server(be, max_resident_slices=1) as srv:
        t_big = srv.submit(big)
        t_small = srv.submit(small)
        try:
            srv.cancel(t_small)
            cancelled = True
        except tb.TbhError as e:
            assert e.code == _lib.TBH_EBUSY
            cancelled = False
        if cancelled:
            with pytest.raises(tb.TbhError):
                srv.poll(t_small)
        else:
            srv.wait(t_small)
        srv.wait(t_big)


def test_failed_request_does_not_block_others(tb, orc):
    """A request with a slice whose arena exceeds the pool's region is
    refused alone, also when it arrives in one burst with good requests
    (no round cap, a 64 MiB pool): the good ones, before and after it in
    the queue, finish with their oracle values."""
    be = tb.HIPBackend(0, pool_bytes=64 << 20)
    try:
        g_big = tb.random_regular_graph(200, 3, seed=8)  # unsliced: sc about 24
        code = tb.initialize_code(g_big, order_trials=10, seed=8)
        too_big = [tb.SlicedBranch(tb.MISProblem(g_big), code, 0)]
        assert tb.slice_footprint(too_big[0]) > (64 << 20)
        good = []
        for n, seed in [(40, 9), (60, 10), (50, 11)]:
            g = tb.random_regular_graph(n, 3, seed=seed)
            good.append((_sliced(tb, g, 6, seed=seed), orc.mis_exact(g)))
        with tb.Server(be) as srv:
            t0 = srv.submit(good[0][0])
            t_bad = srv.submit(too_big)
            t1 = srv.submit(good[1][0])
            t2 = srv.submit(good[2][0])
            with pytest.raises(tb.TbhError, match="admission"):
                srv.wait(t_bad)
            for t, (_, want) in zip((t0, t1, t2), good):
                assert srv.wait(t) == want
            st = srv.stats()
            assert st["failed_requests"] == 1
            # and a good request after the failure
            assert srv.wait(srv.submit(good[1][0])) == good[1][1]
    finally:
        be.close()


def test_round_work_bound_splits_requests(tb, orc, be):
    """A small round work bound splits each request over many pipelined
    rounds (a round always takes at least one slice); every value is
    unchanged and equals the oracle's."""
    reqs = _requests(tb, orc)
    with tb.Server(be, round_log2_macs=12) as srv:
        tickets = [srv.submit(br, dt) for br, dt, _ in reqs]
        got = [srv.wait(t) for t in tickets]
        rounds = srv.stats()["rounds"]
    assert got == [want for _, _, want in reqs]
    assert rounds > len(reqs)


def test_submit_errors(tb, be):
    from tensorbranching_jl_amd import _lib
    with tb.Server(be) as srv:
        with pytest.raises(tb.TbhError):
            srv.submit([])
        with pytest.raises(tb.TbhError):
            srv.poll(12345)
    # after stop the serving entry points refuse
    with pytest.raises(tb.TbhError):
        _lib.check(_lib.tbh().tbh_cancel(be.handle, 1))


def test_latency_lane(tb, orc, be):
    """A latency-lane request (one slice, as slice_dfs_lp submits,
    src/slice.jl:39) submitted while a large request is being served and a
    queue of others waits behind it: it does not wait for that request or
    the queue (rounds are bounded by work), it finishes before most of the
    queued requests, well within the large request's own service time, and
    its value equals the oracle's."""
    g_big = tb.random_regular_graph(220, 3, seed=21)
    big = tb.slice_graph(g_big, None, tb.ContractionTreeSlicer(sc_target=20, order_trials=30),
                         seed=21)
    queue = []
    for seed in range(6):
        g = tb.random_regular_graph(100, 3, seed=20 + seed)
        queue.append(_sliced(tb, g, 8, seed=seed))
    g1 = tb.random_regular_graph(40, 3, seed=30)
    one = _sliced(tb, g1, 12, seed=30)[:1]
    want = orc.contract_slices(one, np.float32)[0]
    with tb.Server(be) as srv:
        # the large request alone: its service time
        t = srv.submit(big)
        while True:
            st, v_big, t0, t1 = srv.poll(t)
            if st in ("done", "failed"):
                break
            time.sleep(0.001)
        assert st == "done"
        alone_s = t1 - t0
        # again, with the queue behind it and the priority slice meanwhile
        t_big = srv.submit(big)
        normal = [srv.submit(b) for b in queue]
        time.sleep(min(0.25 * alone_s, 0.05))
        fast = srv.submit(one, priority=1)
        done = {}
        for tk in [t_big] + normal + [fast]:
            while True:
                st, v, t0, t1 = srv.poll(tk)
                if st in ("done", "failed"):
                    break
                time.sleep(0.001)
            assert st == "done"
            done[tk] = (v, t0, t1)
    assert done[fast][0] == want
    assert done[t_big][0] == v_big
    later = sum(1 for tk in normal if done[tk][2] > done[fast][2])
    assert later >= len(normal) // 2, (later, done)
    lat_fast = done[fast][2] - done[fast][1]
    print(f"big request alone {1e3 * alone_s:.1f} ms; priority slice {1e3 * lat_fast:.1f} ms")
    if alone_s > 0.05:
        assert lat_fast < 0.6 * alone_s, (lat_fast, alone_s)


def test_many_producer_threads(tb, orc, be):
    """tbh_submit and tbh_poll from eight Python threads at once (ctypes
    releases the GIL inside the calls): every request gets its own value,
    reported once."""
    import threading
    reqs = _requests(tb, orc)
    results, errors = {}, []
    lock = threading.Lock()

    def producer(k):
        try:
            mine = []
            for r in range(3):
                br, dt, want = reqs[(k + r) % len(reqs)]
                mine.append((srv.submit(br, dt), want))
            for t, want in mine:
                while True:
                    st, v, _, _ = srv.poll(t)
                    if st == "done":
                        break
                    assert st in ("queued", "running"), st
                    time.sleep(0.001)
                with lock:
                    results[t] = (v, want)
        except Exception as e:  # surfaced below
            with lock:
                errors.append(repr(e))

    with tb.Server(be) as srv:
        threads = [threading.Thread(target=producer, args=(k,)) for k in range(8)]
        for th in threads:
            th.start()
        for th in threads:
            th.join(timeout=120)
        assert not any(th.is_alive() for th in threads)
    assert not errors, errors
    assert len(results) == 24
    for t, (v, want) in results.items():
        assert v == want, (t, v, want)
