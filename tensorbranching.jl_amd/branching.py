"""Branching tables of regions: the table step of optimal_branches
(src/branch.jl:79, branching_table(p, TensorNetworkSolver(), region) from
OptimalBranchingMIS), batched over many regions.

For a region (a vertex list of a MIS problem) the boundary ("open")
vertices are those with a neighbour outside it.  Its alpha tensor holds, per
boundary configuration b (bit k = the k-th open vertex in region order),
the best weight of an independent set of the region whose boundary part is
b: the SizeMax contraction of the region's tensor network with its open
labels kept.  It is computed here as one scalar contraction per b -- the
open vertices fixed to b (those in b taken, with r = their weight, and
their neighbours removed; the others removed), the rest of the region
contracted -- for every region and b in ONE batched tbh_contract_slices
submit.  tbh_region_tables then prunes each tensor as mis_compactify! does
(an entry that a proper subset of its boundary configuration matches or
beats becomes the tropical zero) and enumerates, for each surviving b,
every region configuration that reaches it: the rows of the table.  Both
steps run on the GPU.

Exact parity with OptimalBranchingMIS (not vendored here) is unpinned; the
tests pin the tables against brute-force enumeration of every region
configuration and against hand-worked examples.
"""
import ctypes as C

import numpy as np

from . import _lib
from ._lib import check
from .api import MISProblem, SimpleGraph, SlicedBranch, _dtype_code, default_backend
from .workload import initialize_code


def open_vertices(g, vs):
    """The vertices of vs with a neighbour outside vs, in vs order."""
    inside = set(int(v) for v in vs)
    nb = {}
    for u, v in np.asarray(g.edges).reshape(-1, 2):
        nb.setdefault(int(u), []).append(int(v))
        nb.setdefault(int(v), []).append(int(u))
    return [int(v) for v in vs if any(x not in inside for x in nb.get(int(v), ()))]


def neighbor_region(g, seeds, n_max):
    """A region grown from `seeds` by breadth-first layers (lowest vertex id
    first within a layer) until it holds n_max vertices."""
    nb = [[] for _ in range(g.nv)]
    for u, v in np.asarray(g.edges).reshape(-1, 2):
        nb[int(u)].append(int(v))
        nb[int(v)].append(int(u))
    region = []
    seen = set()
    layer = sorted(set(int(s) for s in seeds))
    while layer and len(region) < n_max:
        nxt = set()
        for v in layer:
            if v in seen:
                continue
            if len(region) == n_max:
                break
            seen.add(v)
            region.append(v)
            nxt.update(nb[v])
        layer = sorted(nxt - seen)
    return region


class BranchingTable:
    """The table of one region: `vertices` (the region), `open_vertices`,
    `alpha` (the reduced alpha tensor, one value per boundary configuration,
    the tropical zero where pruned) and `table` (one row per surviving
    boundary configuration, ascending; a row lists the region
    configurations reaching it as bitmasks, bit i = vertices[i])."""

    def __init__(self, vertices, open_vertices, alpha, table):
        self.vertices = list(vertices)
        self.open_vertices = list(open_vertices)
        self.bit_length = len(self.vertices)
        self.alpha = alpha
        self.table = table

    def __repr__(self):
        return (f"BranchingTable(bit_length={self.bit_length}, open={len(self.open_vertices)}, "
                f"rows={len(self.table)})")


def _region(p, vs, npdt):
    vs = [int(v) for v in vs]
    pos = {v: i for i, v in enumerate(vs)}
    if len(pos) != len(vs):
        raise ValueError("a region lists a vertex twice")
    nv = len(vs)
    adj = np.zeros(max(1, nv), np.uint32)
    for u, v in np.asarray(p.g.edges).reshape(-1, 2):
        u, v = int(u), int(v)
        if u in pos and v in pos:
            adj[pos[u]] |= np.uint32(1 << pos[v])
            adj[pos[v]] |= np.uint32(1 << pos[u])
    ov = open_vertices(p.g, vs)
    opos = [pos[v] for v in ov]
    if p.weights is None:
        w = np.ones(nv, npdt)
    else:
        w = np.asarray(p.weights)[vs].astype(npdt)
    return vs, adj[:nv] if nv else adj[:0], ov, opos, w


def _inner_slice(adj, w, keep, r, weighted, cache, trials, seed):
    """The slice of the region's vertices in `keep` (fixed part worth r)."""
    key = tuple(keep)
    if key not in cache:
        idx = {v: i for i, v in enumerate(keep)}
        edges = [[idx[v], idx[u]] for v in keep for u in keep
                 if u > v and (int(adj[v]) >> u) & 1]
        g = SimpleGraph(len(keep), np.array(edges, np.int32).reshape(-1, 2))
        code = initialize_code(g, trials, seed) if len(keep) else None
        cache[key] = (g, code)
    g, code = cache[key]
    wk = np.asarray(w)[list(keep)] if weighted and len(keep) else None
    return SlicedBranch(MISProblem(g, wk), code, r)


def branching_tables(items, element_type=np.float32, backend=None, order_trials=4, seed=0):
    """Branching tables of many regions at once: items = [(MISProblem,
    region vertex list), ...] -> [BranchingTable, ...].  Every region's
    boundary configurations are contracted in one batched submit, then
    pruned and enumerated on the GPU (tbh_region_tables)."""
    be = backend or default_backend()
    code, npdt = _dtype_code(element_type)
    zero = -np.inf if npdt == np.float32 else _lib.TBH_I32_NEG_INF
    regions, slices, where = [], [], []
    for it, (p, vs) in enumerate(items):
        vs, adj, ov, opos, w = _region(p, vs, npdt)
        if len(vs) > 24 or len(ov) > 16:
            raise ValueError("a region needs at most 24 vertices and 16 open vertices")
        open_mask = sum(1 << i for i in opos)
        inner = [i for i in range(len(vs)) if not (open_mask >> i) & 1]
        cache = {}
        nb = 1 << len(opos)
        for b in range(nb):
            chosen = [opos[k] for k in range(len(opos)) if (b >> k) & 1]
            cmask = sum(1 << c for c in chosen)
            if any(int(adj[c]) & cmask for c in chosen):
                continue  # not independent: alpha[b] is the tropical zero
            removed = 0
            for c in chosen:
                removed |= int(adj[c])
            keep = [i for i in inner if not (removed >> i) & 1]
            r = sum(w[c] for c in chosen)
            r = float(r) if npdt == np.float32 else int(r)
            slices.append(_inner_slice(adj, w, keep, r, p.weights is not None, cache,
                                       order_trials, seed))
            where.append((it, b))
        regions.append((vs, adj, ov, opos, w, open_mask, nb))
    # one batched submit for every region's boundary configurations
    vals = be.contract_slices(slices, npdt, "batched") if slices else np.zeros(0, npdt)
    a0 = np.cumsum([0] + [R[6] for R in regions])
    alpha = np.full(int(a0[-1]), zero, npdt)
    for (it, b), v in zip(where, vals):
        alpha[a0[it] + b] = v
    # prune and enumerate on the GPU
    arr = (_lib.TbhRegion * max(1, len(regions)))()
    keep = []
    for i, (vs, adj, ov, opos, w, open_mask, nb) in enumerate(regions):
        a = np.ascontiguousarray(adj, np.uint32)
        keep.append(a)
        arr[i].nv = len(vs)
        arr[i].adj = a.ctypes.data if len(vs) else None
        if items[i][0].weights is not None:
            ww = np.ascontiguousarray(w, npdt)
            keep.append(ww)
            arr[i].weights = ww.ctypes.data
        else:
            arr[i].weights = None
        arr[i].open = open_mask
    gptr = np.zeros(len(regions) + 1, np.uint64)
    total = C.c_size_t()
    raw = alpha.copy()
    check(_lib.tbh().tbh_region_tables(be.handle, code, arr, len(regions), raw.ctypes.data,
                                       gptr.ctypes.data, None, 0, C.byref(total)))
    cfg = np.zeros(max(1, total.value), np.uint32)
    reduced = alpha.copy()
    check(_lib.tbh().tbh_region_tables(be.handle, code, arr, len(regions), reduced.ctypes.data,
                                       gptr.ctypes.data, cfg.ctypes.data, total.value,
                                       C.byref(total)))
    out = []
    for i, (vs, adj, ov, opos, w, open_mask, nb) in enumerate(regions):
        red = reduced[a0[i]:a0[i + 1]]
        rows = {}
        for x in cfg[int(gptr[i]):int(gptr[i + 1])]:
            x = int(x)
            b = sum(((x >> opos[k]) & 1) << k for k in range(len(opos)))
            rows.setdefault(b, []).append(x)
        table = [rows[b] for b in sorted(rows)]
        out.append(BranchingTable(vs, ov, red.copy(), table))
    return out


def branching_table(p, region, element_type=np.float32, backend=None, **kw):
    """branching_table(p, TensorNetworkSolver(), region) for one region."""
    return branching_tables([(p, region)], element_type, backend, **kw)[0]
