"""Writes the committed fixtures under tests/golden/.

The reference has no stored vectors (its tests check invariants only), so
these are produced by this package's generators and slicer and checked,
before they are written, against solvers independent of the contraction:
the exact branch-and-reduce MIS, brute force (small graphs) and the forest
DP.  Run from the repository root:  python tests/golden/make_fixtures.py

The full-size fixtures take the oracle minutes to hours of CPU time and are
written only on request:
  --n300    the 766 slices of the 3-regular n=300 (seed 2) graph sliced to
          sc_target 25: every per-slice optimum, written with the slice
          set's digest so that a test can regenerate the slices (the host
          slicer is deterministic) and check that they are the same ones;
  --serving    the per-request optima of a seeded serving mix drawn from the
          serving configuration's ranges (serving_requests(24, seed=5),
          requests up to 2^36 MACs), each with its slice set's digest.
"""
import hashlib
import json
import os
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)


def dump_branches(br):
    out = []
    for b in br:
        item = {"nv": b.p.g.nv, "edges": b.p.g.edges.reshape(-1).tolist(), "r": float(b.r),
                "weights": None if b.p.weights is None else np.asarray(b.p.weights).tolist()}
        if b.code is not None:
            item.update(ix_ptr=b.code.ix_ptr.tolist(), ix=b.code.ix.tolist(),
                        tree=b.code.tree.tolist())
        out.append(item)
    return out


def sliceset_digest(br):
    """SHA-256 over every slice's graph, weights, r and tree arrays: two
    slice sets with the same digest hold the same contraction inputs."""
    h = hashlib.sha256()
    for b in br:
        h.update(np.int64([b.p.g.nv, b.p.g.ne()]).tobytes())
        h.update(np.ascontiguousarray(b.p.g.edges, np.int32).tobytes())
        if b.p.weights is not None:
            h.update(np.ascontiguousarray(np.asarray(b.p.weights), np.float64).tobytes())
        h.update(np.float64([float(b.r)]).tobytes())
        if b.code is not None:
            for a in (b.code.ix_ptr, b.code.ix, b.code.tree):
                h.update(np.ascontiguousarray(a, np.int32).tobytes())
    return h.hexdigest()


N300 = dict(n=300, seed=2, sc_target=25, order_trials=200)


def n300_slices(tb):
    g = tb.random_regular_graph(N300["n"], 3, seed=N300["seed"])
    br = tb.slice_graph(g, None, tb.ContractionTreeSlicer(sc_target=N300["sc_target"],
                                                          order_trials=N300["order_trials"]),
                        seed=N300["seed"])
    return g, br


def write_n300(tb, oracle, nthreads=0):
    """Every per-slice optimum of the n=300 workload, by the oracle."""
    g, br = n300_slices(tb)
    nthreads = nthreads or os.cpu_count()
    t = time.time()
    res = oracle.contract_slices(br, np.float32, nthreads=nthreads)
    wall = time.time() - t
    tcs = [None if b.code is None else round(tb.complexity(b).tc, 4) for b in br]
    json.dump({"workload": "random 3-regular n=300 seed 2, sliced to sc_target 25 "
                           "(order_trials 200), unit weights, Tropical Float32",
               "params": N300, "slices": len(br), "digest": sliceset_digest(br),
               "per_slice": res.tolist(), "mis": float(res.max()), "tc": tcs,
               "oracle_wall_s": round(wall, 1), "oracle_threads": nthreads},
              open(os.path.join(HERE, "regular3_n300_sc25.json"), "w"))
    print(f"n300: {len(br)} slices, MIS {res.max()}, oracle {wall:.0f}s on {nthreads} threads")


SERVING_DRAW = dict(count=24, seed=5, order_trials=50, max_slices=4096, max_log2_macs=36,
          max_width_gap=16)


def serving_draw(tb):
    refused = {}
    reqs = tb.serving_requests(SERVING_DRAW["count"], seed=SERVING_DRAW["seed"], order_trials=SERVING_DRAW["order_trials"],
                               max_slices=SERVING_DRAW["max_slices"], max_log2_macs=SERVING_DRAW["max_log2_macs"],
                               max_width_gap=SERVING_DRAW["max_width_gap"], refused=refused)
    return reqs, refused


def write_serving_draw(tb, oracle):
    """Per-request optima of the seeded serving mix, by the oracle; requests
    with unit weights and n <= 120 are cross-checked against the exact
    branch-and-reduce MIS."""
    reqs, refused = serving_draw(tb)
    out = []
    for q in reqs:
        res = oracle.contract_slices(q.branches, q.dtype, nthreads=os.cpu_count())
        opt = res.max()
        if q.weights is None and q.n <= 120:
            assert int(opt) == oracle.mis_exact(q.graph), q.desc
        out.append({"desc": q.desc, "dtype": "float32" if q.dtype == np.float32 else "int32",
                    "slices": len(q.branches), "digest": sliceset_digest(q.branches),
                    "log2_macs": round(q.log2_macs, 3),
                    "optimum": float(opt) if q.dtype == np.float32 else int(opt)})
    json.dump({"mix": tb.SERVING_MIX, "params": SERVING_DRAW, "refused": refused, "requests": out},
              open(os.path.join(HERE, "serving_mix.json"), "w"), indent=0)
    print(f"serving mix: {len(out)} requests, refused {refused}")


def main():
    import __graft_entry__
    from oracle import oracle
    tb = __graft_entry__.package()

    # 3-regular n=100, seed 0, sliced to sc_target 10, unit weights
    g = tb.random_regular_graph(100, 3, seed=0)
    br = tb.slice_graph(g, None, tb.ContractionTreeSlicer(sc_target=10, order_trials=20), seed=0)
    res = oracle.contract_slices(br, np.float32)
    mis = oracle.mis_exact(g)
    assert res.max() == mis
    json.dump({"graph": {"nv": g.nv, "edges": g.edges.reshape(-1).tolist()},
               "dtype": "float32", "slices": dump_branches(br),
               "per_slice": res.tolist(), "mis": mis},
              open(os.path.join(HERE, "regular3_n100_sc10.json"), "w"))

    # weighted Int32: 6x6 King's subgraph, filling 0.8, weights in [1, 100]
    g = tb.random_ksg(6, 6, 0.8, seed=3)
    w = np.random.default_rng(3).integers(1, 101, g.nv).astype(np.int32)
    br = tb.slice_graph(g, w, tb.ContractionTreeSlicer(sc_target=5, order_trials=10), seed=3)
    res = oracle.contract_slices(br, np.int32)
    mwis = oracle.mis_bruteforce(g, w.astype(np.float64))
    assert int(res.max()) == int(mwis)
    json.dump({"graph": {"nv": g.nv, "edges": g.edges.reshape(-1).tolist()},
               "weights": w.tolist(), "dtype": "int32", "slices": dump_branches(br),
               "per_slice": res.tolist(), "mis": int(mwis)},
              open(os.path.join(HERE, "ksg6_int32.json"), "w"))

    # a weighted tree, pinned by the forest DP
    rng = np.random.default_rng(5)
    n = 30
    g = tb.SimpleGraph(n, [[int(rng.integers(0, i)), i] for i in range(1, n)])
    w = rng.integers(1, 50, n).astype(np.float32)
    code = tb.initialize_code(g, 4, 5)
    br = [tb.SlicedBranch(tb.MISProblem(g, w), code, 0)]
    res = oracle.contract_slices(br, np.float32)
    assert res[0] == oracle.mwis_forest(g, w.astype(np.float64))
    json.dump({"graph": {"nv": g.nv, "edges": g.edges.reshape(-1).tolist()},
               "weights": w.tolist(), "dtype": "float32", "slices": dump_branches(br),
               "per_slice": res.tolist(), "mis": float(res[0])},
              open(os.path.join(HERE, "tree30_f32.json"), "w"))

    # max-plus GEMM: 5x7 by 7x3, integers with -inf, C checked by enumeration
    A = rng.integers(0, 20, (5, 7)).astype(np.float32)
    B = rng.integers(0, 20, (7, 3)).astype(np.float32)
    A[rng.random(A.shape) < 0.2] = -np.inf
    C = np.full((5, 3), -np.inf, np.float32)
    for i in range(5):
        for j in range(3):
            for k in range(7):
                C[i, j] = max(C[i, j], A[i, k] + B[k, j])
    assert np.array_equal(C, oracle.maxplus_gemm(np.asfortranarray(A), np.asfortranarray(B)))
    enc = lambda M: [[None if np.isneginf(x) else float(x) for x in row] for row in M]
    json.dump({"A": enc(A), "B": enc(B), "C": enc(C)},
              open(os.path.join(HERE, "gemm_5x7x3.json"), "w"))
    print("fixtures written")


if __name__ == "__main__":
    if "--n300" in sys.argv or "--serving" in sys.argv:
        import __graft_entry__
        from oracle import oracle
        if "--n300" in sys.argv:
            write_n300(__graft_entry__.package(), oracle)
        if "--serving" in sys.argv:
            write_serving_draw(__graft_entry__.package(), oracle)
    else:
        main()
