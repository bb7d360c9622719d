"""ctypes wrapper of the CPU restatement (oracle/_build/liboracle.so).

TEST INFRASTRUCTURE ONLY: used by tests/, __graft_entry__.smoke() and
bench.py's cpu_baseline leg as the checker.  The product never imports it.
"""
import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB = os.path.join(HERE, "_build", "liboracle.so")

F32, I32 = 0, 1
I32_NEG = -(1 << 30)


class OrcSlice(C.Structure):
    _fields_ = [("nv", C.c_int32), ("weights", C.c_void_p), ("nleaves", C.c_int32),
                ("ix_ptr", C.c_void_p), ("ix", C.c_void_p), ("tree", C.c_void_p),
                ("r", C.c_double)]


_SIG = {
    "orc_node_labels": (C.c_int, [C.c_int32, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p,
                                  C.c_void_p, C.c_int32]),
    "orc_complexity": (C.c_int, [C.c_int32, C.c_void_p, C.c_void_p, C.c_void_p,
                                 C.POINTER(C.c_double), C.POINTER(C.c_double),
                                 C.POINTER(C.c_double)]),
    "orc_peak_memory": (C.c_double, [C.c_int32, C.c_void_p, C.c_void_p, C.c_void_p]),
    "orc_all_memory": (C.c_double, [C.c_int32, C.c_void_p, C.c_void_p, C.c_void_p]),
    "orc_maxplus_gemm_f32": (C.c_int, [C.c_int64] * 4 + [C.c_void_p, C.c_int64, C.c_int64] * 3),
    "orc_maxplus_gemm_i32": (C.c_int, [C.c_int64] * 4 + [C.c_void_p, C.c_int64, C.c_int64] * 3),
    "orc_einsum2": (C.c_int, [C.c_int, C.c_int, C.c_void_p, C.c_void_p, C.c_int, C.c_void_p,
                              C.c_void_p, C.c_int, C.c_void_p, C.c_void_p]),
    "orc_solve_slice": (C.c_int, [C.c_int, C.POINTER(OrcSlice), C.c_void_p]),
    "orc_contract_slices": (C.c_int, [C.c_int, C.POINTER(OrcSlice), C.c_int64, C.c_void_p,
                                      C.c_int]),
    "orc_mis_bruteforce": (C.c_double, [C.c_int32, C.c_int32, C.c_void_p, C.c_void_p]),
    "orc_mis_exact": (C.c_int64, [C.c_int32, C.c_int32, C.c_void_p]),
    "orc_mwis_forest": (C.c_double, [C.c_int32, C.c_int32, C.c_void_p, C.c_void_p]),
    "orc_last_error": (C.c_char_p, []),
}

_lib = None


def build():
    subprocess.run(["make", "-s", "-C", HERE], check=True)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB):
            build()
        _lib = C.CDLL(LIB)
        for k, (res, args) in _SIG.items():
            f = getattr(_lib, k)
            f.restype = res
            f.argtypes = args
    return _lib


def _err():
    return lib().orc_last_error().decode()


def _dt(npdt):
    return F32 if np.dtype(npdt) == np.float32 else I32


def maxplus_gemm(A, B):
    """Column-major (Fortran) A (M,K[,b]) and B (K,N[,b]) -> C."""
    A = np.asfortranarray(A)
    B = np.asfortranarray(B)
    batched = A.ndim == 3
    M, K = A.shape[:2]
    N = B.shape[1]
    nb = A.shape[2] if batched else 1
    Cm = np.empty((M, N, nb) if batched else (M, N), A.dtype, order="F")
    f = lib().orc_maxplus_gemm_f32 if A.dtype == np.float32 else lib().orc_maxplus_gemm_i32
    rc = f(M, N, K, nb, A.ctypes.data, M, M * K, B.ctypes.data, K, K * N, Cm.ctypes.data, M, M * N)
    if rc:
        raise RuntimeError(_err())
    return Cm


def einsum2(ixa, A, ixb, B, ixc):
    """Binary tropical einsum over labelled 2^r tensors (flat, column-major)."""
    A = np.ascontiguousarray(A)
    B = np.ascontiguousarray(B)
    la = np.asarray(ixa, np.int32)
    lb = np.asarray(ixb, np.int32)
    lc = np.asarray(ixc, np.int32)
    out = np.empty(1 << len(ixc), A.dtype)
    rc = lib().orc_einsum2(_dt(A.dtype), len(la), la.ctypes.data, A.ctypes.data, len(lb),
                           lb.ctypes.data, B.ctypes.data, len(lc), lc.ctypes.data,
                           out.ctypes.data)
    if rc:
        raise RuntimeError(_err())
    return out


class _Slices:
    def __init__(self, branches, npdt):
        self.keep = []
        self.arr = (OrcSlice * max(1, len(branches)))()
        for i, b in enumerate(branches):
            s = self.arr[i]
            s.nv = b.p.g.nv
            s.r = float(b.r)
            if b.p.weights is not None:
                w = np.ascontiguousarray(np.asarray(b.p.weights).astype(npdt))
                self.keep.append(w)
                s.weights = w.ctypes.data
            if b.code is None or b.p.g.nv == 0:
                s.nleaves = 0
                continue
            c = b.code
            self.keep.append(c)
            s.nleaves = c.nleaves
            s.ix_ptr = c.ix_ptr.ctypes.data
            s.ix = c.ix.ctypes.data
            s.tree = c.tree.ctypes.data if len(c.tree) else None


def contract_slices(branches, element_type=np.float32, nthreads=0):
    """CPU restatement of contract_slices (src/dynamic_ob.jl:36-48)."""
    m = _Slices(branches, element_type)
    out = np.zeros(max(1, len(branches)), element_type)
    rc = lib().orc_contract_slices(_dt(element_type), C.cast(m.arr, C.POINTER(OrcSlice)),
                                   len(branches), out.ctypes.data, nthreads)
    if rc:
        raise RuntimeError(_err())
    return out[:len(branches)]


def solve_slice(branch, element_type=np.float32):
    m = _Slices([branch], element_type)
    out = np.zeros(1, element_type)
    rc = lib().orc_solve_slice(_dt(element_type), C.cast(m.arr, C.POINTER(OrcSlice)),
                               out.ctypes.data)
    if rc:
        raise RuntimeError(_err())
    return out[0]


def complexity(code):
    tc, sc, rw = C.c_double(), C.c_double(), C.c_double()
    rc = lib().orc_complexity(code.nleaves, code.ix_ptr.ctypes.data, code.ix.ctypes.data,
                              code.tree.ctypes.data if len(code.tree) else None,
                              C.byref(tc), C.byref(sc), C.byref(rw))
    if rc:
        raise RuntimeError(_err())
    return tc.value, sc.value, rw.value


def peak_memory(code):
    return lib().orc_peak_memory(code.nleaves, code.ix_ptr.ctypes.data, code.ix.ctypes.data,
                                 code.tree.ctypes.data if len(code.tree) else None)


def all_memory(code):
    return lib().orc_all_memory(code.nleaves, code.ix_ptr.ctypes.data, code.ix.ctypes.data,
                                code.tree.ctypes.data if len(code.tree) else None)


def node_labels(code):
    cap = int(code.ix_ptr[-1]) * 4 + 1024 + 64 * code.nleaves
    ptr = np.zeros(max(1, code.nleaves), np.int32)
    lab = np.zeros(cap, np.int32)
    n = lib().orc_node_labels(code.nleaves, code.ix_ptr.ctypes.data, code.ix.ctypes.data,
                              code.tree.ctypes.data if len(code.tree) else None,
                              ptr.ctypes.data, lab.ctypes.data, cap)
    if n < 0:
        raise RuntimeError(_err())
    return [lab[ptr[j]:ptr[j + 1]].tolist() for j in range(code.nleaves - 1)]


def _edges(g):
    return np.ascontiguousarray(np.asarray(g.edges, np.int32).reshape(-1))


def mis_bruteforce(g, weights=None):
    e = _edges(g)
    w = None if weights is None else np.ascontiguousarray(weights, np.float64)
    return lib().orc_mis_bruteforce(g.nv, g.ne(), e.ctypes.data, None if w is None else w.ctypes.data)


def mis_exact(g):
    e = _edges(g)
    return int(lib().orc_mis_exact(g.nv, g.ne(), e.ctypes.data))


def mwis_forest(g, weights=None):
    e = _edges(g)
    w = None if weights is None else np.ascontiguousarray(weights, np.float64)
    return lib().orc_mwis_forest(g.nv, g.ne(), e.ctypes.data, None if w is None else w.ctypes.data)


def branching_table_bruteforce(g, weights, region):
    """Branching table of a region by enumerating every configuration of its
    vertices (bit i = region[i]): the open vertices (a neighbour outside the
    region, in region order), the alpha tensor (per boundary configuration b,
    bit k = k-th open vertex: the best weight of an independent set of the
    region whose boundary part is b), the reduced tensor (an entry that a
    proper subset of b matches or beats becomes -inf, the rule of
    GenericTensorNetworks' mis_compactify! used by OptimalBranchingMIS's
    reduced_alpha_configs for branching_table, src/branch.jl:79) and the
    rows: per surviving b, ascending, the configurations reaching it.
    Returns (open_vertices, alpha, reduced, rows) with -inf for zeros."""
    region = [int(v) for v in region]
    pos = {v: i for i, v in enumerate(region)}
    nv = len(region)
    adj = [0] * nv
    nbrs = {}
    for u, v in np.asarray(g.edges).reshape(-1, 2):
        u, v = int(u), int(v)
        nbrs.setdefault(u, set()).add(v)
        nbrs.setdefault(v, set()).add(u)
        if u in pos and v in pos:
            adj[pos[u]] |= 1 << pos[v]
            adj[pos[v]] |= 1 << pos[u]
    ov = [v for v in region if any(x not in pos for x in nbrs.get(v, ()))]
    opos = [pos[v] for v in ov]
    w = np.ones(nv) if weights is None else np.asarray(weights, np.float64)[region]
    x = np.arange(1 << nv, dtype=np.int64)
    valid = np.ones(x.shape, bool)
    wt = np.zeros(x.shape, np.float64)
    for v in range(nv):
        bit = (x >> v) & 1
        valid &= ~((bit == 1) & ((x & adj[v]) != 0))
        wt += bit * w[v]
    b = np.zeros(x.shape, np.int64)
    for k, p in enumerate(opos):
        b |= ((x >> p) & 1) << k
    nb = 1 << len(opos)
    alpha = np.full(nb, -np.inf)
    np.maximum.at(alpha, b[valid], wt[valid])
    reduced = alpha.copy()
    for a in range(nb):
        for s in range(nb):
            if s != a and (s & a) == s and alpha[s] >= alpha[a]:
                reduced[a] = -np.inf
                break
    rows = []
    for bb in range(nb):
        if reduced[bb] == -np.inf:
            continue
        sel = valid & (b == bb) & (wt == reduced[bb])
        rows.append(sorted(int(v) for v in x[sel]))
    return ov, alpha, reduced, rows
