"""Committed fixtures (tests/golden/, written by make_fixtures.py and
checked there against independent MIS solvers): the oracle reproduces them
on the CPU, the HIP backend on the GPU, in both modes, bit for bit."""
import json
import os

import numpy as np
import pytest

HERE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
CASES = ["regular3_n100_sc10.json", "ksg6_int32.json", "tree30_f32.json"]


def load_case(tb, name):
    d = json.load(open(os.path.join(HERE, name)))
    dt = np.float32 if d["dtype"] == "float32" else np.int32
    br = []
    for s in d["slices"]:
        g = tb.SimpleGraph(s["nv"], np.array(s["edges"], np.int32).reshape(-1, 2))
        w = None if s["weights"] is None else np.array(s["weights"], dt)
        code = None
        if "ix_ptr" in s:
            code = tb.CompressedEinsum.from_arrays(np.array(s["ix_ptr"], np.int32),
                                                  np.array(s["ix"], np.int32),
                                                  np.array(s["tree"], np.int32))
        r = s["r"] if dt == np.float32 else int(s["r"])
        br.append(tb.SlicedBranch(tb.MISProblem(g, w), code, r))
    return d, dt, br


@pytest.mark.parametrize("name", CASES)
def test_oracle_reproduces_fixture(tb, orc, name):
    d, dt, br = load_case(tb, name)
    res = orc.contract_slices(br, dt)
    assert res.tolist() == d["per_slice"]
    assert res.max() == d["mis"]


def _dec(M):
    return np.asfortranarray(np.array([[-np.inf if x is None else x for x in r] for r in M],
                                      np.float32))


def test_oracle_gemm_fixture(orc):
    d = json.load(open(os.path.join(HERE, "gemm_5x7x3.json")))
    assert np.array_equal(orc.maxplus_gemm(_dec(d["A"]), _dec(d["B"])), _dec(d["C"]))


@pytest.mark.gpu
@pytest.mark.parametrize("mode", ["batched", "plain"])
@pytest.mark.parametrize("name", CASES)
def test_gpu_reproduces_fixture(tb, backend, name, mode):
    d, dt, br = load_case(tb, name)
    res = backend.contract_slices(br, dt, mode)
    assert res.dtype == dt
    assert res.tolist() == d["per_slice"]


@pytest.mark.gpu
def test_gpu_gemm_fixture(backend):
    d = json.load(open(os.path.join(HERE, "gemm_5x7x3.json")))
    assert np.array_equal(backend.maxplus_gemm(_dec(d["A"]), _dec(d["B"])), _dec(d["C"]))


def _make_fixtures():
    import importlib.util
    spec = importlib.util.spec_from_file_location("make_fixtures",
                                                  os.path.join(HERE, "make_fixtures.py"))
    mod = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(mod)
    return mod


def test_c3_fixture_slices_regenerate(tb, orc):
    """The full-size fixture's slice set is what the host slicer produces
    (same digest), and the oracle reproduces its least-work slices."""
    mf = _make_fixtures()
    d = json.load(open(os.path.join(HERE, "regular3_n300_sc25.json")))
    g, br = mf.n300_slices(tb)
    assert len(br) == d["slices"] == 766
    assert mf.sliceset_digest(br) == d["digest"]
    tcs = np.array([np.inf if t is None else t for t in d["tc"]])
    idx = np.argsort(tcs, kind="stable")[:12]
    res = orc.contract_slices([br[k] for k in idx], np.float32)
    assert res.tolist() == [d["per_slice"][k] for k in idx]
    assert max(d["per_slice"]) == d["mis"]


def test_c5_fixture_requests_regenerate(tb, orc):
    """The serving-mix fixture's requests are what serving_requests draws
    (same descriptions, refusals and digests), and the oracle reproduces the
    stored optima of the smaller requests."""
    mf = _make_fixtures()
    d = json.load(open(os.path.join(HERE, "serving_mix.json")))
    reqs, refused = mf.serving_draw(tb)
    assert refused == d["refused"]
    for q, f in zip(reqs, d["requests"]):
        assert q.desc == f["desc"] and mf.sliceset_digest(q.branches) == f["digest"]
        if f["log2_macs"] <= 28:
            assert orc.contract_slices(q.branches, q.dtype).max() == f["optimum"], q.desc
