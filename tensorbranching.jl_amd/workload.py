"""Host producers of sliced MIS networks (libtbh_workload.so) and the
dynamic_ob_mis driver twin.

These stand in for the reference's host algorithms that need Julia:
graph generators (Graphs.random_regular_graph, erdos_renyi, random_ksg at
src/utils.jl:191-195), the initial TreeSA code (initialize_code,
src/dynamic_ob.jl:50-54) and branching-based slicing (slice / slice_bfs,
src/slice.jl:4-11,68-116).  Seeds index this package's own PRNG.
"""
import ctypes as C

import numpy as np

from . import _lib
from .api import (CompressedEinsum, MISProblem, SimpleGraph, SlicedBranch, UnitWeight,
                  contract_slices)


def _tbw_err():
    return _lib.tbw().tbw_last_error().decode(errors="replace")


def _graph_from_handle(h):
    if not h:
        raise ValueError(_tbw_err())
    L = _lib.tbw()
    try:
        nv = L.tbw_graph_nv(h)
        ne = L.tbw_graph_ne(h)
        e = np.ctypeslib.as_array(L.tbw_graph_edges(h), (2 * ne,)).copy() if ne else np.zeros(0)
        return SimpleGraph(nv, e.reshape(-1, 2))
    finally:
        L.tbw_graph_free(h)


def random_regular_graph(n, d, seed=0):
    return _graph_from_handle(_lib.tbw().tbw_random_regular(n, d, seed))


def erdos_renyi(n, p, seed=0):
    return _graph_from_handle(_lib.tbw().tbw_erdos_renyi(n, float(p), seed))


def random_ksg(m, n, rho, seed=0):
    """m x n King's subgraph with round(rho*m*n) occupied sites."""
    return _graph_from_handle(_lib.tbw().tbw_ksg(m, n, float(rho), seed))


def _handle_of(g):
    e = np.ascontiguousarray(g.edges, np.int32).reshape(-1)
    h = _lib.tbw().tbw_graph_from_edges(g.nv, g.ne(), e.ctypes.data)
    if not h:
        raise ValueError(_tbw_err())
    return h


def _read_sliceset(ss, weighted, wdtype):
    L = _lib.tbw()
    out = []
    v = _lib.TbwSliceView()
    try:
        for i in range(L.tbw_sliceset_count(ss)):
            if L.tbw_sliceset_get(ss, i, C.byref(v)):
                raise RuntimeError(_tbw_err())
            nv, ne = v.nv, v.ne
            edges = np.ctypeslib.as_array(v.edges, (2 * ne,)).copy().reshape(-1, 2) if ne else \
                np.zeros((0, 2), np.int32)
            g = SimpleGraph(nv, edges)
            w = None
            if weighted and nv:
                w = np.ctypeslib.as_array(v.weights, (nv,)).copy().astype(wdtype)
            if nv == 0:
                out.append(SlicedBranch(MISProblem(g, w), None, v.r))
                continue
            nl = v.nleaves
            ix_ptr = np.ctypeslib.as_array(v.ix_ptr, (nl + 1,)).copy()
            ix = np.ctypeslib.as_array(v.ix, (int(ix_ptr[-1]),)).copy()
            tree = np.ctypeslib.as_array(v.tree, (2 * (nl - 1),)).copy() if nl > 1 else \
                np.zeros(0, np.int32)
            code = CompressedEinsum.from_arrays(ix_ptr, ix, tree)
            r = v.r
            if weighted and np.issubdtype(np.dtype(wdtype), np.integer):
                r = int(round(r))
            out.append(SlicedBranch(MISProblem(g, w), code, r))
    finally:
        L.tbw_sliceset_free(ss)
    return out


def _weights_arg(g, weights):
    if weights is None or isinstance(weights, UnitWeight):
        return None, None, np.float32
    w = np.asarray(weights)
    if w.shape != (g.nv,):
        raise ValueError("one weight per vertex")
    wd = np.ascontiguousarray(w, np.float64)
    return wd, wd.ctypes.data, w.dtype


def initialize_code(g, order_trials=100, seed=0):
    """A contraction tree for the whole graph (TreeSA-style search)."""
    h = _handle_of(g)
    try:
        ss = _lib.tbw().tbw_whole(h, None, order_trials, seed)
    finally:
        _lib.tbw().tbw_graph_free(h)
    if not ss:
        raise RuntimeError(_tbw_err())
    return _read_sliceset(ss, False, np.float32)[0].code


class ContractionTreeSlicer:
    """Slicer config (src/types.jl:43-49); only sc_target is used by the
    backend, as the memory bound of every slice."""

    def __init__(self, sc_target=30, order_trials=100, kernelize=True):
        self.sc_target = int(sc_target)
        self.order_trials = int(order_trials)
        self.kernelize = bool(kernelize)


def slice_graph(g, weights=None, slicer=None, seed=0, nthreads=0):
    """Kernelize, order and slice g until every slice has sc <= sc_target;
    returns the SlicedBranch list whose maximum (+r) is the MIS."""
    slicer = slicer or ContractionTreeSlicer()
    wd, wp, wdt = _weights_arg(g, weights)
    h = _handle_of(g)
    try:
        ss = _lib.tbw().tbw_slice(h, wp, slicer.sc_target, slicer.order_trials, seed, nthreads,
                                  1 if slicer.kernelize else 0)
    finally:
        _lib.tbw().tbw_graph_free(h)
    if not ss:
        raise RuntimeError(_tbw_err())
    return _read_sliceset(ss, wd is not None, wdt)


def dynamic_ob_mis(g, weights=None, *, slicer=None, usecuda=True, element_type=np.float32,
                   mode="batched", backend=None, seed=0):
    """Twin of dynamic_ob_mis (src/dynamic_ob.jl:6-28): host kernelization,
    ordering and slicing, then the GPU contract_slices and the maximum."""
    branches = slice_graph(g, weights, slicer, seed=seed)
    res = contract_slices(branches, element_type, usecuda, mode=mode, backend=backend)
    return res.max()


def serving_requests(count, seed=0, order_trials=50):
    """A mixed stream of MIS requests for the serving scheduler: random
    3-regular graphs with even n ~ U[100, 260] and Erdos-Renyi graphs with n ~
    U[100, 250], p = 3 / (n - 1), each sliced to sc_target ~ U[lo(n), 25]
    with lo(n) = max(15, n // 10 - 3) (lower targets multiply the slice count
    of large graphs into the 10^5 range).  Returns a list of
    (description, graph, branches)."""
    rng = np.random.default_rng(seed)
    out = []
    for _ in range(count):
        gseed = int(rng.integers(1 << 31))
        if rng.random() < 0.5:
            n = 2 * int(rng.integers(50, 131))  # n * 3 must be even
            g, kind = random_regular_graph(n, 3, seed=gseed), "3-regular"
        else:
            n = int(rng.integers(100, 251))
            g, kind = erdos_renyi(n, 3.0 / (n - 1), seed=gseed), "ER"
        lo = max(15, n // 10 - 3)
        sc_target = int(rng.integers(lo, 26))
        br = slice_graph(g, None, ContractionTreeSlicer(sc_target=sc_target,
                                                        order_trials=order_trials), seed=gseed)
        out.append((f"{kind} n={n} seed={gseed} sc_target={sc_target}", g, br))
    return out
