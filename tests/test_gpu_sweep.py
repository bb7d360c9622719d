"""A seeded sweep over the graph families and weight types the reference's
workloads use (random regular graphs, Erdos-Renyi graphs, King's subgraphs;
unit, integer Float32 and Int32 weights) at slicing targets from 8 to 18.
Each case's slices reach the planner with their own shapes (tensor ranks,
label orders, GEMM vs EW nodes, tile layouts), and every slice's optimum
must equal the oracle's, bit for bit, in both executors.  Small graphs also
meet the exact MIS / forest MWIS of the independent solvers."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

# (family, parameters, weights, sc_target, seed)
CASES = [
    ("regular", (120, 3), None, 12, 11),
    ("regular", (90, 4), None, 14, 12),
    ("regular", (140, 3), "f32", 14, 13),
    ("regular", (100, 5), "i32", 16, 14),
    ("er", (150, 0.03), None, 16, 15),
    ("er", (110, 0.04), "f32", 14, 16),
    ("er", (200, 0.02), "i32", 18, 17),
    ("er", (80, 0.06), None, 10, 18),
    ("ksg", (20, 20, 0.9), None, 10, 19),
    ("ksg", (16, 16, 0.7), "i32", 7, 20),
    ("ksg", (24, 24, 0.8), "f32", 12, 21),
    ("ksg", (30, 30, 0.8), None, 14, 22),
]


def _graph(tb, family, params, seed):
    if family == "regular":
        return tb.random_regular_graph(*params, seed=seed)
    if family == "er":
        return tb.erdos_renyi(*params, seed=seed)
    return tb.random_ksg(*params, seed=seed)


@pytest.mark.parametrize("case", CASES, ids=lambda c: f"{c[0]}{c[1]}-{c[2]}-sc{c[3]}")
def test_family_sweep_bit_exact(tb, orc, backend, case):
    family, params, wt, sc_target, seed = case
    g = _graph(tb, family, params, seed)
    w, dtype = None, np.float32
    if wt is not None:
        dtype = np.int32 if wt == "i32" else np.float32
        w = np.random.default_rng(seed).integers(1, 101, g.nv).astype(dtype)
    br = tb.slice_graph(g, w, tb.ContractionTreeSlicer(sc_target=sc_target, order_trials=20),
                        seed=seed)
    assert max(tb.sc(b) for b in br if b.code is not None) <= sc_target
    ref = orc.contract_slices(br, dtype, nthreads=16)
    for mode in ("batched", "plain"):
        got = backend.contract_slices(br, dtype, mode)
        assert got.dtype == dtype
        assert np.array_equal(got, ref), mode
    if g.nv <= 130 and w is None:
        assert ref.max() == orc.mis_exact(g)
