"""The oracle (CPU restatement) pinned by checks that do not restate it:
tropical identities, associativity and transposition of the max-plus GEMM,
longest paths in a DAG, brute-force enumeration of a binary contraction,
brute-force / closed-form / forest-DP / exact MIS values, and the slicing
and order invariants of the reference tests (test/slice.jl:31-33,
test/decompose.jl:28)."""
import itertools

import numpy as np
import pytest

from conftest import rand_tropical


# ------------------------------------------------------------ GEMM ------
@pytest.mark.parametrize("dtype", [np.float32, np.int32])
def test_gemm_identity(orc, dtype):
    rng = np.random.default_rng(1)
    A = rand_tropical(rng, (37, 23), dtype)
    neg = -np.inf if dtype == np.float32 else -(1 << 30)
    I = np.full((23, 23), neg, dtype, order="F")
    np.fill_diagonal(I, 0)
    assert np.array_equal(orc.maxplus_gemm(A, I), A)
    I2 = np.full((37, 37), neg, dtype, order="F")
    np.fill_diagonal(I2, 0)
    assert np.array_equal(orc.maxplus_gemm(I2, A), A)


@pytest.mark.parametrize("dtype", [np.float32, np.int32])
def test_gemm_associative_and_transpose(orc, dtype):
    rng = np.random.default_rng(2)
    A = rand_tropical(rng, (13, 7), dtype)
    B = rand_tropical(rng, (7, 11), dtype)
    C = rand_tropical(rng, (11, 5), dtype)
    AB_C = orc.maxplus_gemm(orc.maxplus_gemm(A, B), C)
    A_BC = orc.maxplus_gemm(A, orc.maxplus_gemm(B, C))
    assert np.array_equal(AB_C, A_BC)
    # (AB)^T = B^T A^T: catches a transposed operand
    AB = orc.maxplus_gemm(A, B)
    BtAt = orc.maxplus_gemm(np.asfortranarray(B.T), np.asfortranarray(A.T))
    assert np.array_equal(AB.T, BtAt)


def test_gemm_longest_path(orc):
    """Max-plus powers of a DAG's weight matrix give longest paths with a
    fixed number of edges; compared with a dynamic program over the DAG."""
    rng = np.random.default_rng(3)
    n = 12
    W = np.full((n, n), -np.inf, np.float32)
    for i in range(n):
        for j in range(i + 1, n):
            if rng.random() < 0.4:
                W[i, j] = float(rng.integers(1, 20))
    W = np.asfortranarray(W)
    P = W.copy(order="F")
    for hops in range(2, 5):
        P = orc.maxplus_gemm(P, W)
        # DP: best[h][i][j] over exactly `hops` edges
        best = W.astype(np.float64)
        for _ in range(hops - 1):
            nb = np.full((n, n), -np.inf)
            for i in range(n):
                for j in range(n):
                    for k in range(n):
                        nb[i, j] = max(nb[i, j], best[i, k] + W[k, j])
            best = nb
        assert np.array_equal(P.astype(np.float64), best)


def test_gemm_batched_and_int32_agree(orc):
    rng = np.random.default_rng(4)
    A = rand_tropical(rng, (9, 6, 3), np.float32, neg_frac=0.0)
    B = rand_tropical(rng, (6, 4, 3), np.float32, neg_frac=0.0)
    Cf = orc.maxplus_gemm(A, B)
    Ci = orc.maxplus_gemm(np.asfortranarray(A.astype(np.int32)),
                          np.asfortranarray(B.astype(np.int32)))
    assert np.array_equal(Cf.astype(np.int64), Ci.astype(np.int64))
    for t in range(3):
        assert np.array_equal(Cf[:, :, t], orc.maxplus_gemm(np.asfortranarray(A[:, :, t]),
                                                            np.asfortranarray(B[:, :, t])))


def test_int32_sentinel_canonical(orc):
    neg = -(1 << 30)
    A = np.asfortranarray(np.array([[neg, 5], [neg, neg]], np.int32))
    B = np.asfortranarray(np.array([[neg, 3], [7, neg]], np.int32))
    C = orc.maxplus_gemm(A, B)
    # -inf + finite and -inf + -inf both read back as the sentinel
    assert C.tolist() == [[12, neg], [neg, neg]]


# ---------------------------------------------------------- einsum2 -----
def _brute_einsum(ixa, A, ixb, B, ixc):
    labels = sorted(set(ixa) | set(ixb))
    out = np.full(1 << len(ixc), -np.inf)
    for vals in itertools.product((0, 1), repeat=len(labels)):
        asg = dict(zip(labels, vals))
        ia = sum(asg[l] << i for i, l in enumerate(ixa))
        ib = sum(asg[l] << i for i, l in enumerate(ixb))
        ic = sum(asg[l] << i for i, l in enumerate(ixc))
        out[ic] = max(out[ic], A[ia] + B[ib])
    return out


@pytest.mark.parametrize("case", range(40))
def test_einsum2_bruteforce(orc, case):
    rng = np.random.default_rng(100 + case)
    pool = list(range(7))
    ixa = list(rng.choice(pool, rng.integers(1, 6), replace=False))
    ixb = list(rng.choice(pool, rng.integers(1, 6), replace=False))
    union = list(dict.fromkeys(ixa + ixb))
    ixc = [l for l in union if rng.random() < 0.5]
    rng.shuffle(ixc)
    A = rng.integers(0, 50, 1 << len(ixa)).astype(np.float32)
    B = rng.integers(0, 50, 1 << len(ixb)).astype(np.float32)
    A[rng.random(A.shape) < 0.2] = -np.inf
    got = orc.einsum2(ixa, A, ixb, B, ixc)
    assert np.array_equal(got, _brute_einsum(ixa, A, ixb, B, ixc).astype(np.float32))


# --------------------------------------------------------- solve_slice --
def _graph_code(tb, g, trials=6, seed=0):
    return tb.SlicedBranch(tb.MISProblem(g), tb.initialize_code(g, trials, seed), 0)


@pytest.mark.parametrize("seed", range(12))
def test_solve_matches_bruteforce(tb, orc, seed):
    rng = np.random.default_rng(seed)
    n = int(rng.integers(6, 22))
    g = tb.erdos_renyi(n, 0.25, seed=seed)
    w = rng.integers(1, 30, n).astype(np.float32)
    code = tb.initialize_code(g, 4, seed)
    br = tb.SlicedBranch(tb.MISProblem(g, w), code, 0)
    assert orc.solve_slice(br, np.float32) == orc.mis_bruteforce(g, w.astype(np.float64))
    br_i = tb.SlicedBranch(tb.MISProblem(g, w.astype(np.int32)), code, 0)
    assert orc.solve_slice(br_i, np.int32) == int(orc.mis_bruteforce(g, w.astype(np.float64)))


def test_closed_forms(tb, orc):
    cases = []
    for n in (3, 8, 21):
        cases.append((tb.SimpleGraph(n, [[i, (i + 1) % n] for i in range(n)]), n // 2))
    for n in (2, 9):
        cases.append((tb.SimpleGraph(n, [[i, i + 1] for i in range(n - 1)]), (n + 1) // 2))
    cases.append((tb.SimpleGraph(9, [[i, 3 + j] for i in range(3) for j in range(6)]), 6))
    pet = [[i, (i + 1) % 5] for i in range(5)] + [[i, i + 5] for i in range(5)] + \
          [[5 + i, 5 + (i + 2) % 5] for i in range(5)]
    cases.append((tb.SimpleGraph(10, pet), 4))
    L = 6
    grid = [[i * L + j, i * L + j + 1] for i in range(L) for j in range(L - 1)] + \
           [[i * L + j, (i + 1) * L + j] for i in range(L - 1) for j in range(L)]
    cases.append((tb.SimpleGraph(L * L, grid), L * L // 2))
    for g, expect in cases:
        assert orc.solve_slice(_graph_code(tb, g), np.float32) == expect
        assert orc.mis_exact(g) == expect


def test_weighted_forest_dp(tb, orc):
    rng = np.random.default_rng(7)
    n = 40
    parent = [int(rng.integers(0, i)) for i in range(1, n)]
    g = tb.SimpleGraph(n, [[p, i + 1] for i, p in enumerate(parent)])
    w = rng.integers(1, 100, n).astype(np.float32)
    br = tb.SlicedBranch(tb.MISProblem(g, w), tb.initialize_code(g, 4), 0)
    assert orc.solve_slice(br, np.float32) == orc.mwis_forest(g, w.astype(np.float64))


def test_exact_solver_matches_bruteforce(tb, orc):
    for seed in range(6):
        g = tb.random_regular_graph(30, 3, seed=seed)
        assert orc.mis_exact(g) == orc.mis_bruteforce(g)


def test_order_invariance_and_slicing_invariant(tb, orc):
    g = tb.random_regular_graph(60, 3, seed=3)
    vals = {float(orc.solve_slice(_graph_code(tb, g, trials=2 + s, seed=s), np.float32))
            for s in range(3)}
    assert vals == {float(orc.mis_exact(g))}
    br = tb.slice_graph(g, None, tb.ContractionTreeSlicer(sc_target=7, order_trials=8))
    assert len(br) > 1
    assert orc.contract_slices(br, np.float32).max() == orc.mis_exact(g)


def test_regular3_n100_mis(tb, orc):
    """3-regular n=100 seed 0 sliced to sc_target 10: max over slices (+r)
    equals the exact MIS (README.md:38-43 shows 44 for an unseeded graph)."""
    g = tb.random_regular_graph(100, 3, seed=0)
    br = tb.slice_graph(g, None, tb.ContractionTreeSlicer(sc_target=10, order_trials=20))
    mis = orc.mis_exact(g)
    assert 40 <= mis <= 47
    assert orc.contract_slices(br, np.float32).max() == mis


# ------------------------------------------------------ memory model ----
def test_peak_and_all_memory_hand_computed(tb, orc):
    """Path a-b-c: leaves [a],[b],[c],[a,b],[b,c] (ranks 1,1,1,2,2); tree
    ((0,3),(1,(2,4))).  Node labels: (0,3)->[a,b]? a occurs only in leaves
    0 and 3, so it is dropped: [b].  (2,4)->[b]; (1,(2,4))->[b]; root [].
    Start: 2+2+2+4+4 = 14.  Post-order: n0=(0,3) +2 -> 16; n1=(2,4) +2 -> 18;
    n2=(1,n1) +2 - (2+4) -> 14; root=(n0,n2) +1 - (2+4) - (2+2) -> 5.
    peak = log2 18; all = log2(2+2+2+1)."""
    code = tb.CompressedEinsum([[0], [1], [2], [0, 1], [1, 2]], [],
                               tb.ContractionTree(tb.ContractionTree(0, 3),
                                                  tb.ContractionTree(1, tb.ContractionTree(2, 4))))
    assert orc.node_labels(code) == [[1], [1], [1], []]
    assert orc.peak_memory(code) == pytest.approx(np.log2(18))
    assert orc.all_memory(code) == pytest.approx(np.log2(7))
    tc, sc, rwc = orc.complexity(code)
    # node unions: {a,b}=2, {b,c}=2, {b}=1, {b}=1 -> 4+4+2+2
    assert tc == pytest.approx(np.log2(12))
    assert sc == 1
    # rw: (2+4+2) + (2+4+2) + (2+2+2) + (2+2+1)
    assert rwc == pytest.approx(np.log2(27))


def test_branching_table_oracle_hand_worked(tb, orc):
    """The brute-force table oracle on tables worked out by hand.
    Path 0-1-2, region [0, 1]: vertex 1 is open (neighbour 2 outside).
    Unit weights: alpha = [1 ({0}), 1 ({1})]; taking 1 is no better than not
    (the empty boundary configuration reaches 1 too), so only b = 0 stays,
    with the one configuration {0} = 0b01."""
    g = tb.SimpleGraph(3, np.array([[0, 1], [1, 2]], np.int32))
    ov, alpha, red, rows = orc.branching_table_bruteforce(g, None, [0, 1])
    assert ov == [1]
    assert alpha.tolist() == [1.0, 1.0] and red.tolist() == [1.0, -np.inf]
    assert rows == [[0b01]]
    # weight 5 on the open vertex: taking it is better, both rows stay
    ov, alpha, red, rows = orc.branching_table_bruteforce(g, [1.0, 5.0, 1.0], [0, 1])
    assert alpha.tolist() == [1.0, 5.0] and red.tolist() == [1.0, 5.0]
    assert rows == [[0b01], [0b10]]
    # star: centre 0 (open, neighbour 4 outside) with leaves 1, 2, 3 inside:
    # b = 0 reaches 3 with the three leaves, b = 1 only 1 -> pruned
    g = tb.SimpleGraph(5, np.array([[0, 1], [0, 2], [0, 3], [0, 4]], np.int32))
    ov, alpha, red, rows = orc.branching_table_bruteforce(g, None, [0, 1, 2, 3])
    assert ov == [0] and alpha.tolist() == [3.0, 1.0] and red.tolist() == [3.0, -np.inf]
    assert rows == [[0b1110]]
    # two open ends of a path 0-1-2-3 inside a longer path (-1 and 4 are
    # outside): region [1, 2, 3] with 1 and 3 open; configurations:
    # b=00: {2} -> 1; b=01 (1 taken): {1, 3}? 3 is open and not taken, so
    # {1} -> 1, pruned by b=00; b=10 (3 taken): {3} -> 1, pruned; b=11:
    # {1, 3} -> 2, kept (its proper subsets reach only 1); rows: [{2}],
    # [{1, 3}] = [0b010], [0b101]
    g = tb.SimpleGraph(5, np.array([[0, 1], [1, 2], [2, 3], [3, 4]], np.int32))
    ov, alpha, red, rows = orc.branching_table_bruteforce(g, None, [1, 2, 3])
    assert ov == [1, 3]
    assert alpha.tolist() == [1.0, 1.0, 1.0, 2.0]
    assert red.tolist() == [1.0, -np.inf, -np.inf, 2.0]
    assert rows == [[0b010], [0b101]]
