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
                  complexity, contract_slices)


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


def slice_graph(g, weights=None, slicer=None, seed=0, nthreads=0, max_slices=1 << 18):
    """Kernelize, order and slice g until every slice has sc <= sc_target;
    returns the SlicedBranch list whose maximum (+r) is the MIS.  Raises
    RuntimeError as soon as the slices would outnumber max_slices."""
    slicer = slicer or ContractionTreeSlicer()
    wd, wp, wdt = _weights_arg(g, weights)
    h = _handle_of(g)
    try:
        ss = _lib.tbw().tbw_slice_limited(h, wp, slicer.sc_target, slicer.order_trials, seed,
                                          nthreads, 1 if slicer.kernelize else 0,
                                          int(max_slices))
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


#: the serving configuration's request ranges
SERVING_MIX = ("3-regular n ~ U[100,400] (even) and Erdos-Renyi n ~ U[100,250] with "
               "p = 3/(n-1), half each; sc_target ~ U[15,25]; half Tropical Float32 with unit "
               "weights, half Tropical Int32 with integer weights ~ U[1,100]")


class ServingRequest:
    """One drawn MIS request: its graph, weights, element type and slices."""

    def __init__(self, kind, n, gseed, sc_target, dtype, graph, weights, branches):
        self.kind, self.n, self.gseed, self.sc_target = kind, n, gseed, sc_target
        self.dtype, self.graph, self.weights, self.branches = dtype, graph, weights, branches
        tcs = [complexity(b).tc for b in branches if b.code is not None]
        self.log2_macs = float(np.log2(np.sum(np.exp2(tcs)))) if tcs else 0.0
        self.desc = (f"{kind} n={n} seed={gseed} sc_target={sc_target} "
                     f"{'f32' if dtype == np.float32 else 'i32'}")

    def __iter__(self):  # (description, graph, branches)
        return iter((self.desc, self.graph, self.branches))


def serving_requests(count, seed=0, order_trials=50, max_slices=4096, max_log2_macs=None,
                     max_width_gap=16, refused=None):
    """A seeded stream of `count` MIS requests drawn from the serving
    configuration's ranges (SERVING_MIX).  A draw the host slicer cannot
    serve is refused and the next one drawn: (a) when the graph's unsliced
    contraction width (a 20-trial tree of the unweighted graph) exceeds
    sc_target by more than max_width_gap -- a cheap test that refuses most
    3-regular graphs with n above about 300, whose slicing to sc_target <= 25
    goes beyond 2^13 slices (n = 400 beyond 2^18); (b) when slicing yields
    more than max_slices slices; (c) when max_log2_macs is given and the
    request's work (log2 of the summed 2^tc of its slices) exceeds it.
    `refused`, when a dict, receives the counts per reason.  Returns a list
    of ServingRequest (iterable as (description, graph, branches))."""
    rng = np.random.default_rng(seed)
    out = []
    counts = {"width": 0, "slices": 0, "work": 0}
    while len(out) < count:
        gseed = int(rng.integers(1 << 31))
        regular = rng.random() < 0.5
        n = 2 * int(rng.integers(50, 201)) if regular else int(rng.integers(100, 251))
        sc_target = int(rng.integers(15, 26))
        f32 = rng.random() < 0.5
        wseed = int(rng.integers(1 << 31))
        if regular:
            g, kind = random_regular_graph(n, 3, seed=gseed), "3-regular"
        else:
            g, kind = erdos_renyi(n, 3.0 / (n - 1), seed=gseed), "ER"
        dtype = np.float32 if f32 else np.int32
        w = None if f32 else np.random.default_rng(wseed).integers(1, 101, g.nv).astype(np.int32)
        width = complexity(SlicedBranch(MISProblem(g), initialize_code(g, 20, gseed), 0)).sc
        if width - sc_target > max_width_gap:
            counts["width"] += 1
            continue
        try:
            br = slice_graph(g, w, ContractionTreeSlicer(sc_target=sc_target,
                                                         order_trials=order_trials),
                             seed=gseed, max_slices=max_slices)
        except RuntimeError:
            counts["slices"] += 1
            continue
        req = ServingRequest(kind, n, gseed, sc_target, dtype, g, w, br)
        if max_log2_macs is not None and req.log2_macs > max_log2_macs:
            counts["work"] += 1
            continue
        out.append(req)
    if refused is not None:
        refused.update(counts)
    return out
