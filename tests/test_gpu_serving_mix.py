"""The serving configuration: a seeded Poisson stream of MIS requests drawn
from its ranges -- 3-regular graphs with n ~ U[100,400], Erdos-Renyi graphs
with n ~ U[100,250] at mean degree 3, sc_target ~ U[15,25], Tropical Float32
(unit weights) and Int32 (integer weights in [1,100]) -- served by the
continuous-batching scheduler over a 200 GB pool.  Draws beyond the host
slicer's reach are refused by serving_requests (the counts are part of the
fixture), as are requests above 2^36 MACs, which keeps the oracle's work for
the fixture to about a minute.  Every served value is checked against the
oracle's optimum stored in tests/golden/serving_mix.json (written by
make_fixtures.py --serving with each request's slice-set digest), alone and in
the mixed stream (maximum over slices, src/dynamic_ob.jl:24-27)."""
import importlib.util
import json
import os
import time

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def _fixtures_module():
    spec = importlib.util.spec_from_file_location("make_fixtures",
                                                  os.path.join(GOLDEN, "make_fixtures.py"))
    mod = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(mod)
    return mod


@pytest.fixture(scope="module")
def mix(tb):
    mf = _fixtures_module()
    fx = json.load(open(os.path.join(GOLDEN, "serving_mix.json")))
    assert fx["params"] == mf.SERVING_DRAW
    t = time.time()
    reqs, refused = mf.serving_draw(tb)
    print(f"[mix] {len(reqs)} requests drawn in {time.time() - t:.1f}s, refused {refused}",
          flush=True)
    assert refused == fx["refused"]
    assert len(reqs) == len(fx["requests"])
    for q, f in zip(reqs, fx["requests"]):
        assert q.desc == f["desc"]
        assert mf.sliceset_digest(q.branches) == f["digest"], q.desc
    want = [np.float32(f["optimum"]) if f["dtype"] == "float32" else np.int32(f["optimum"])
            for f in fx["requests"]]
    return reqs, want


@pytest.mark.timeout(900)
def test_serving_mix_covers_the_ranges(mix):
    reqs, _ = mix
    kinds = {q.kind for q in reqs}
    dtypes = {np.dtype(q.dtype) for q in reqs}
    assert kinds == {"3-regular", "ER"} and dtypes == {np.dtype(np.float32), np.dtype(np.int32)}
    assert all(15 <= q.sc_target <= 25 for q in reqs)
    assert all(100 <= q.n <= 400 if q.kind == "3-regular" else 100 <= q.n <= 250 for q in reqs)
    assert max(len(q.branches) for q in reqs) > 100


@pytest.mark.timeout(900)
def test_serving_mix_alone_and_poisson_against_oracle(tb, mix):
    """Each request served alone, then the whole mix twice over as Poisson
    arrivals (at about twice the rate the requests are served alone, so
    rounds mix requests), over a 200 GB pool: every value equals the
    oracle's, in its own element type."""
    reqs, want = mix
    be = tb.HIPBackend(0, pool_bytes=200 << 30)
    try:
        with tb.Server(be) as srv:
            t0 = time.perf_counter()
            alone = []
            for q in reqs:
                alone.append(srv.wait(srv.submit(q.branches, q.dtype)))
            t_alone = time.perf_counter() - t0
            for q, a, w in zip(reqs, alone, want):
                assert a == w and np.asarray(a).dtype == q.dtype, (q.desc, a, w)
            rate = 2.0 * len(reqs) / t_alone
            rng = np.random.default_rng(55)
            order = list(range(len(reqs))) * 2
            gaps = rng.exponential(1.0 / rate, len(order))
            tickets = []
            start = time.perf_counter()
            due = 0.0
            for k, gap in zip(order, gaps):
                due += gap
                dt = start + due - time.perf_counter()
                if dt > 0:
                    time.sleep(dt)
                tickets.append((srv.submit(reqs[k].branches, reqs[k].dtype), k))
            lat = []
            for t, k in tickets:
                while True:
                    st, v, ts, td = srv.poll(t)
                    if st in ("done", "failed"):
                        break
                    time.sleep(0.0005)
                assert st == "done", reqs[k].desc
                assert v == want[k], (reqs[k].desc, v, want[k])
                lat.append(td - ts)
            st = srv.stats()
        lat = np.sort(lat)
        print(f"[mix] alone {1e3 * t_alone:.1f} ms for {len(reqs)} requests; Poisson "
              f"{rate:.0f} req/s x {len(order)}: p50 {1e3 * np.percentile(lat, 50):.2f} ms, "
              f"p99 {1e3 * np.percentile(lat, 99):.2f} ms, rounds {st['rounds']}", flush=True)
        assert st["failed_requests"] == 0 and st["retried_rounds"] == 0
    finally:
        be.close()
