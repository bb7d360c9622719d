"""The on-disk slice format of the reference's disk-backed slicer
(src/io.jl:1-120, written by slice_bfs_rw, src/slice.jl:148-232), so that
slice sets produced in Julia feed this backend and slice sets produced here
can be read by the reference.

A directory holds, per finished slice `id`:
  graph_<id>.dot      the graph in Graphs.jl's LG text format (savegraph's
                      default format, despite the extension): a header
                      `nv,ne,u,name[,eltype,simplegraph]`, then one 1-based
                      edge `u,v` per line;
  eincode_<id>.json   the uncompressed contraction tree as OMEinsum's
                      writejson stores a NestedEinsum: {"label-type",
                      "inputs": leaf label lists, "output": [], "tree":
                      nested {"isleaf", "tensorindex" (1-based) | "args",
                      "eins": {"ixs", "iy"}}}, or the text `nothing`;
  weights_<id>.txt    `UnitWeight` and n, or the element type name and one
                      weight per line as Julia's string(w) prints it (`1.5`
                      for a Float32, no `f0` suffix: the reader parses each
                      line with parse(T, line), which rejects the suffix);
and slices.csv with columns id,sc,tc,r.

Labels and tensor indices are 1-based on disk (Julia) and 0-based here.
"""
import csv
import json
import os
import sys
from concurrent.futures import ThreadPoolExecutor

import numpy as np

from .api import (CompressedEinsum, ContractionTree, MISProblem, SimpleGraph, SlicedBranch,
                  UnitWeight, complexity)

# contraction trees of sliced MIS networks are chains several hundred nodes
# deep; the JSON codec and the tree walks below recurse once per level
_DEPTH = 20000


def _deep():
    if sys.getrecursionlimit() < _DEPTH:
        sys.setrecursionlimit(_DEPTH)


_WEIGHT_TYPES = {"Float64": np.float64, "Float32": np.float32, "Int64": np.int64,
                 "Int32": np.int32}


def _parse_julia_number(text, dtype):
    t = text.strip()
    if np.issubdtype(dtype, np.floating):
        # also accept Julia's Float32 literal syntax (1.5f0, 2.0f-3), which
        # repr / show print, although string(w) does not
        if "f" in t and not t.lower().startswith(("inf", "-inf", "nan")):
            t = t.replace("f", "e")
        return dtype(float(t))
    return dtype(int(t))


def _julia_string(x):
    """A weight as Julia's string(w) prints it (src/io.jl:17): the shortest
    decimal that reads back to the same value in the weight's own type, with
    no type suffix, so that parse(T, line) (src/io.jl:40) accepts it."""
    if isinstance(x, np.floating):
        t = str(x)  # shortest round-trip digits in x's own precision
        return t if any(c in t for c in ".eEn") else t + ".0"
    return repr(x.item())


# ---------------------------------------------------------------- weights --
def save_weights(filename, weights):
    """src/io.jl:1-21."""
    if weights is None:
        raise TypeError("unit weights need the vertex count: pass UnitWeight(n)")
    with open(filename, "w") as f:
        if isinstance(weights, UnitWeight):
            f.write("UnitWeight\n%d\n" % weights.n)
            return
        w = np.asarray(weights)
        names = {np.dtype(v): k for k, v in _WEIGHT_TYPES.items()}
        if w.dtype not in names:
            raise TypeError(f"weights of type {w.dtype} have no Julia counterpart")
        f.write(names[w.dtype] + "\n")
        for x in w:
            f.write(_julia_string(x))
            f.write("\n")


def load_weights(filename):
    """src/io.jl:23-44: UnitWeight(n) or a typed vector."""
    with open(filename) as f:
        line1 = f.readline().strip()
        if line1 == "UnitWeight":
            return UnitWeight(int(f.readline()))
        if line1 not in _WEIGHT_TYPES:
            raise ValueError(f"Unknown weight type: {line1}")
        dt = _WEIGHT_TYPES[line1]
        return np.asarray([_parse_julia_number(l, dt) for l in f if l.strip()], dtype=dt)


# ------------------------------------------------------------------ graph --
def savegraph(filename, g, gname="graph"):
    """Graphs.jl LG format (savegraph's default)."""
    with open(filename, "w") as f:
        f.write("%d,%d,u,%s,Int64,simplegraph\n" % (g.nv, g.ne(), gname))
        for u, v in sorted((min(a, b), max(a, b)) for a, b in g.edges):
            f.write("%d,%d\n" % (u + 1, v + 1))


def loadgraph(filename):
    """Reads one LG-format graph (header `nv,ne,u|d,name[,...]`, `#`
    comments ignored)."""
    with open(filename) as f:
        lines = [l.strip() for l in f if l.strip() and not l.lstrip().startswith("#")]
    if not lines:
        raise ValueError(f"{filename}: empty graph file")
    head = [x.strip() for x in lines[0].split(",")]
    if len(head) < 4:
        raise ValueError(f"{filename}: LG header needs nv,ne,u|d,name")
    nv, ne, kind = int(head[0]), int(head[1]), head[2]
    if kind != "u":
        raise ValueError(f"{filename}: only undirected graphs (u) describe MIS problems")
    edges = []
    for l in lines[1:1 + ne]:
        a, b = (int(x) for x in l.split(",")[:2])
        if not (1 <= a <= nv and 1 <= b <= nv):
            raise ValueError(f"{filename}: edge {a},{b} outside 1..{nv}")
        edges.append((a - 1, b - 1))
    if len(edges) != ne:
        raise ValueError(f"{filename}: header says {ne} edges, found {len(edges)}")
    return SimpleGraph(nv, edges)


# ------------------------------------------------------------------- code --
def _node_labels(ixs, ct):
    """Labels of every internal node by the reference rule (a subtree keeps a
    label iff it also occurs outside it), post-order, first-appearance order."""
    total = {}
    for x in ixs:
        for l in x:
            total[l] = total.get(l, 0) + 1
    out = {}

    def walk(node):
        if isinstance(node, (int, np.integer)):
            return list(ixs[int(node)]), {l: 1 for l in ixs[int(node)]}
        la, ca = walk(node.left)
        lb, cb = walk(node.right)
        cnt = dict(ca)
        for l, c in cb.items():
            cnt[l] = cnt.get(l, 0) + c
        labels = []
        for l in la + lb:
            if l not in labels and cnt[l] < total[l]:
                labels.append(l)
        out[id(node)] = (la, lb, labels)
        # only the labels that leave the subtree carry on (linear in the tree)
        return labels, {l: cnt[l] for l in labels}

    walk(ct)
    return out


def _code_to_dict(code):
    ixs = code.ixs
    ct = code.ct
    labs = _node_labels(ixs, ct) if not isinstance(ct, (int, np.integer)) else {}

    def one(l):
        return [int(x) + 1 for x in l]

    def todict(node):
        if isinstance(node, (int, np.integer)):
            return {"isleaf": True, "tensorindex": int(node) + 1}
        la, lb, iy = labs[id(node)]
        return {"isleaf": False, "args": [todict(node.left), todict(node.right)],
                "eins": {"ixs": [one(la), one(lb)], "iy": one(iy)}}

    return {"label-type": "Int64", "inputs": [one(x) for x in ixs], "output": [],
            "tree": todict(ct)}


def save_code(filename, code):
    """src/io.jl:58-70: the tree as JSON, or the text `nothing`."""
    _deep()
    with open(filename, "w") as f:
        if code is None:
            f.write("nothing")
        else:
            # dumps (the C encoder); dump's generator encoder is quadratic in depth
            f.write(json.dumps(_code_to_dict(code)))


def load_code(filename):
    """src/io.jl:72-79: a CompressedEinsum (0-based labels) or None.  Every
    internal node's `iy`, when present, must hold the labels the reference
    rule gives that node."""
    _deep()
    with open(filename) as f:
        text = f.read()
    if text.strip() == "nothing":
        return None
    d = json.loads(text)
    ixs = [[int(l) - 1 for l in x] for x in d["inputs"]]
    if d.get("output"):
        raise ValueError(f"{filename}: MIS networks have an empty output")
    n = len(ixs)
    seen = set()

    def build(node):
        if node.get("isleaf"):
            k = int(node["tensorindex"]) - 1
            if not 0 <= k < n or k in seen:
                raise ValueError(f"{filename}: leaf tensorindex {k + 1} invalid or repeated")
            seen.add(k)
            return k
        args = node["args"]
        if len(args) != 2:
            raise ValueError(f"{filename}: only binary trees are supported "
                             "(src/decompose.jl:103 asserts is_binary)")
        return ContractionTree(build(args[0]), build(args[1]))

    ct = build(d["tree"])
    if len(seen) != n:
        raise ValueError(f"{filename}: the tree does not use every input tensor")
    code = CompressedEinsum(ixs, [], ct)
    if not isinstance(ct, (int, np.integer)):
        labs = _node_labels(ixs, ct)

        def check(node, nd):
            if isinstance(node, (int, np.integer)):
                return
            want = set(labs[id(node)][2])
            iy = nd.get("eins", {}).get("iy")
            if iy is not None and {int(l) - 1 for l in iy} != want:
                raise ValueError(f"{filename}: a node's iy differs from its labels by the "
                                 "intermediate-label rule")
            check(node.left, nd["args"][0])
            check(node.right, nd["args"][1])

        check(ct, d["tree"])
    return code


# ------------------------------------------------------------------ slices --
def _paths(dirname, id_):
    return (os.path.join(dirname, f"graph_{id_}.dot"),
            os.path.join(dirname, f"eincode_{id_}.json"),
            os.path.join(dirname, f"weights_{id_}.txt"))


def save_finished(dirname, branch, id_):
    """src/io.jl:81-89."""
    gp, cp, wp = _paths(dirname, id_)
    savegraph(gp, branch.p.g)
    save_code(cp, branch.code)
    w = branch.p.weights
    save_weights(wp, UnitWeight(branch.p.g.nv) if w is None else w)


def load_finished(dirname, id_):
    """src/io.jl:91-99: (graph, code, weights)."""
    gp, cp, wp = _paths(dirname, id_)
    return loadgraph(gp), load_code(cp), load_weights(wp)


def _parse_r(text):
    t = text.strip()
    try:
        return int(t)
    except ValueError:
        return _parse_julia_number(t, np.float64).item()


def load_all_finished(dirname):
    """src/io.jl:113-120: every slice listed in slices.csv."""
    out = []
    with open(os.path.join(dirname, "slices.csv")) as f:
        for row in csv.DictReader(f):
            id_ = int(row["id"])
            g, code, w = load_finished(dirname, id_)
            out.append(SlicedBranch(MISProblem(g, w), code, _parse_r(row["r"])))
    return out


def threaded_saveslices(dirname, slices, ids, threads=8):
    """src/io.jl:46-56 (slices without a code are skipped, as there)."""
    if len(slices) != len(ids):
        raise ValueError("one id per slice")
    with ThreadPoolExecutor(max_workers=max(1, threads)) as ex:
        list(ex.map(lambda a: save_finished(dirname, *a) if a[0].code is not None else None,
                    zip(slices, ids)))


def save_slices(dirname, slices, first_id=1):
    """A directory as slice_bfs_rw leaves it (src/slice.jl:148-232): the
    finished slices' files and slices.csv (id, sc, tc, r).  Empty-graph
    slices are written too (code `nothing`), so that the maximum over the
    directory is the maximum over `slices`."""
    os.makedirs(dirname, exist_ok=True)
    ids = list(range(first_id, first_id + len(slices)))
    with ThreadPoolExecutor(max_workers=8) as ex:
        list(ex.map(lambda a: save_finished(dirname, *a), zip(slices, ids)))
    with open(os.path.join(dirname, "slices.csv"), "w", newline="") as f:
        w = csv.writer(f)
        w.writerow(["id", "sc", "tc", "r"])
        for id_, b in zip(ids, slices):
            if b.code is None:
                sc_, tc_ = 0.0, 0.0
            else:
                cc = complexity(b)
                sc_, tc_ = cc.sc, cc.tc
            w.writerow([id_, sc_, tc_, b.r])
    return ids
