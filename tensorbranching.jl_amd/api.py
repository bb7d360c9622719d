"""Host-side mirror of TensorBranching.jl's contraction interface over the
MI355X backend (libtbh.so).

Names, argument meaning and error behaviour follow the reference:
  SlicedBranch, MISProblem           src/types.jl:85-103
  CompressedEinsum, ContractionTree  src/types.jl:51-58 (ct leaves index ixs)
  contract_slices, solve_slice       src/dynamic_ob.jl:30-48
  complexity / tc / sc               src/types.jl:115-121
  contraction_peak_memory /
  contraction_all_memory             src/utils.jl:197-229
`usecuda` selects the device path as in the reference; this backend has no
CPU contraction, so usecuda=False raises instead of silently computing on
the host.  Labels and leaf indices are 0-based here (Julia's are 1-based).
"""
import ctypes as C
import hashlib
import json
import os
from collections import namedtuple

import numpy as np

from . import _lib
from ._lib import check

ContractionComplexity = namedtuple("ContractionComplexity", ["tc", "sc", "rwc"])


# --------------------------------------------------------------- graphs ---
class SimpleGraph:
    """Undirected simple graph: nv vertices 0..nv-1 and an (ne, 2) edge array."""

    def __init__(self, nv, edges=()):
        self.nv = int(nv)
        e = np.asarray(edges, dtype=np.int32).reshape(-1, 2)
        if e.size and (e.min() < 0 or e.max() >= self.nv or np.any(e[:, 0] == e[:, 1])):
            raise ValueError("edges must join distinct vertices in [0, nv)")
        e = np.sort(e, axis=1)
        self.edges = np.unique(e, axis=0) if e.size else e.reshape(0, 2)

    def ne(self):
        return int(self.edges.shape[0])

    def __repr__(self):
        return f"{{{self.nv}, {self.ne()}}} undirected simple graph"


# ----------------------------------------------------------- tree IR ------
class ContractionTree:
    """Binary contraction tree; leaves are 0-based indices into ixs."""

    __slots__ = ("left", "right")

    def __init__(self, left, right):
        self.left = left
        self.right = right


def _ct_to_postorder(ct, nleaves):
    """Post-order child list: node j has children tree[2j], tree[2j+1]; ids
    below nleaves are leaves, others internal node id - nleaves."""
    if isinstance(ct, (int, np.integer)):
        return np.zeros(0, np.int32)
    out = []
    stack = [(ct, False)]
    ids = []
    while stack:
        node, done = stack.pop()
        if isinstance(node, (int, np.integer)):
            if not 0 <= int(node) < nleaves:
                raise ValueError(f"tree leaf {node} outside 0..{nleaves - 1}")
            ids.append(int(node))
            continue
        if done:
            r = ids.pop()
            l = ids.pop()
            out.extend((l, r))
            ids.append(nleaves + len(out) // 2 - 1)
        else:
            stack.append((node, True))
            stack.append((node.right, False))
            stack.append((node.left, False))
    return np.asarray(out, dtype=np.int32)


def _postorder_to_ct(tree, nleaves):
    if nleaves == 1 and len(tree) == 0:
        return 0
    nodes = []
    for j in range(len(tree) // 2):
        kids = []
        for c in (int(tree[2 * j]), int(tree[2 * j + 1])):
            kids.append(c if c < nleaves else nodes[c - nleaves])
        nodes.append(ContractionTree(*kids))
    return nodes[-1]


class CompressedEinsum:
    """Leaf label lists `ixs`, output labels `iy` (empty for MIS) and the
    binary tree `ct` (src/types.jl:51-58).  Stored flat for the C ABI."""

    def __init__(self, ixs, iy=(), ct=None, *, _arrays=None):
        if _arrays is not None:
            self.ix_ptr, self.ix, self.tree = _arrays
        else:
            ixs = [list(map(int, x)) for x in ixs]
            self.ix_ptr = np.zeros(len(ixs) + 1, np.int32)
            self.ix_ptr[1:] = np.cumsum([len(x) for x in ixs])
            self.ix = np.asarray([l for x in ixs for l in x], dtype=np.int32)
            if ct is None:
                raise ValueError("a contraction tree is required")
            self.tree = _ct_to_postorder(ct, len(ixs))
            if len(ixs) > 1 and len(self.tree) != 2 * (len(ixs) - 1):
                raise ValueError("the tree must contain every leaf exactly once")
        self.iy = list(iy)
        if self.iy:
            raise ValueError("MIS networks have an empty output (iy = [])")
        self.ix_ptr = np.ascontiguousarray(self.ix_ptr, np.int32)
        self.ix = np.ascontiguousarray(self.ix, np.int32)
        self.tree = np.ascontiguousarray(self.tree, np.int32)

    @classmethod
    def from_arrays(cls, ix_ptr, ix, tree):
        return cls(None, _arrays=(np.asarray(ix_ptr), np.asarray(ix), np.asarray(tree)))

    @property
    def nleaves(self):
        return len(self.ix_ptr) - 1

    def _c_args(self):
        """(nleaves, ix_ptr, ix, tree) for a tbh_slice, the arrays as
        addresses; kept while the three arrays are the same objects."""
        c = self.__dict__.get("_c")
        if c is None or c[0] is not self.ix_ptr or c[1] is not self.ix or c[2] is not self.tree:
            c = (self.ix_ptr, self.ix, self.tree,
                 (len(self.ix_ptr) - 1, self.ix_ptr.ctypes.data, self.ix.ctypes.data,
                  self.tree.ctypes.data if len(self.tree) else None))
            self._c = c
        return c[3]

    @property
    def ixs(self):
        return [self.ix[self.ix_ptr[i]:self.ix_ptr[i + 1]].tolist() for i in range(self.nleaves)]

    @property
    def ct(self):
        return _postorder_to_ct(self.tree, self.nleaves)


class UnitWeight:
    def __init__(self, n):
        self.n = n


class MISProblem:
    def __init__(self, g, weights=None):
        self.g = g
        if isinstance(weights, UnitWeight):
            weights = None
        if weights is not None:
            weights = np.asarray(weights)
            if weights.shape != (g.nv,):
                raise ValueError("one weight per vertex")
        self.weights = weights


class SlicedBranch:
    """One slice: MIS problem, compressed code (None for an empty graph) and
    the fixed weight r (src/types.jl:85-103)."""

    def __init__(self, p, code, r=0):
        if not isinstance(p, MISProblem):
            raise TypeError("p must be a MISProblem")
        if code is None and p.g.nv != 0:
            raise ValueError("a non-empty graph needs a contraction code")
        self.p = p
        self.code = code
        self.r = r

    def __repr__(self):
        kind = "simple graph" if self.p.weights is None else "weighted graph"
        cc = complexity(self)
        return (f"SlicedBranch: graph: {{{self.p.g.nv}, {self.p.g.ne()}}} {kind}; "
                f"code complexity: sc: {cc.sc}, tc: {cc.tc}; fixed weight: {self.r}")


# ------------------------------------------------------------- dtypes -----
def _dtype_code(element_type):
    dt = np.dtype(element_type)
    if dt == np.float32:
        return _lib.TBH_F32, np.float32
    if dt == np.int32:
        return _lib.TBH_I32, np.int32
    raise TypeError(f"element_type must be Float32 or Int32 (got {dt}); "
                    "tropical values are integer-valued and exact in 32 bits")


def workload_key(branches, element_type=np.float32):
    """A digest of a slice set as the backend sees it (element type, and per
    slice nv, weights, r, leaf label lists and tree): the key of the
    autotuner's cache."""
    _, npdt = _dtype_code(element_type)
    h = hashlib.sha256(np.dtype(npdt).name.encode())
    for b in branches:
        h.update(np.array([b.p.g.nv], np.int64).tobytes())
        h.update(np.array([float(b.r)], np.float64).tobytes())
        if b.p.weights is not None:
            h.update(b"w" + np.ascontiguousarray(np.asarray(b.p.weights).astype(npdt)).tobytes())
        if b.code is not None:
            c = b.code
            for a in (c.ix_ptr, c.ix, c.tree):
                h.update(b"|" + np.ascontiguousarray(a, dtype=np.int32).tobytes())
    return h.hexdigest()


def _read_tuning_cache(path):
    try:
        with open(path) as f:
            return json.load(f)
    except FileNotFoundError:
        return {}


class _Marshal:
    """tbh_slice array over the branches, keeping its buffers alive."""

    def __init__(self, branches, npdt):
        self.keep = []
        n = len(branches)
        self.arr = (_lib.TbhSlice * max(1, n))()
        for i, b in enumerate(branches):
            s = self.arr[i]
            s.nv = b.p.g.nv
            s.r = float(b.r)
            if b.p.weights is not None:
                w = np.ascontiguousarray(b.p.weights)
                if npdt == np.int32:
                    if not np.all(np.asarray(w) == np.round(w)):
                        raise ValueError("Int32 contraction needs integer weights")
                w = np.ascontiguousarray(w.astype(npdt))
                self.keep.append(w)
                s.weights = w.ctypes.data
            else:
                s.weights = None
            if b.code is None or b.p.g.nv == 0:
                s.nleaves = 0
                continue
            c = b.code
            s.nleaves, s.ix_ptr, s.ix, s.tree = c._c_args()
            self.keep.append(c)
        self.n = n

    @property
    def ptr(self):
        return C.cast(self.arr, C.POINTER(_lib.TbhSlice))


# ------------------------------------------------------------ backend -----
class HIPBackend:
    """One MI355X context (device memory pool + stream), the replacement of
    the implicit CUDA.jl state behind usecuda=true."""

    def __init__(self, device=0, pool_bytes=0):
        h = C.c_void_p()
        check(_lib.tbh().tbh_create(int(device), int(pool_bytes), C.byref(h)))
        self._h = h
        self.device = device

    @property
    def handle(self):
        if self._h is None:
            raise RuntimeError("backend closed")
        return self._h

    def close(self):
        if getattr(self, "_h", None) is not None:
            _lib.tbh().tbh_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def contract_slices(self, branches, element_type=np.float32, mode="batched"):
        code, npdt = _dtype_code(element_type)
        m = _Marshal(branches, npdt)
        out = np.zeros(max(1, m.n), npdt)
        check(_lib.tbh().tbh_contract_slices(self.handle, code, m.ptr, m.n, out.ctypes.data,
                                             _mode(mode)))
        return out[:m.n]

    def solve_slice(self, branch, element_type=np.float32, mode="batched"):
        code, npdt = _dtype_code(element_type)
        m = _Marshal([branch], npdt)
        out = np.zeros(1, npdt)
        check(_lib.tbh().tbh_solve_slice(self.handle, code, m.ptr, out.ctypes.data, _mode(mode)))
        return out[0]

    def contract_max(self, branches, element_type=np.float32, mode="batched"):
        code, npdt = _dtype_code(element_type)
        m = _Marshal(branches, npdt)
        out = np.zeros(1, npdt)
        check(_lib.tbh().tbh_contract_max(self.handle, code, m.ptr, m.n, out.ctypes.data,
                                          _mode(mode)))
        return out[0]

    def tuning(self):
        """The context's planner / launch parameters (tbh_tuning) as a dict."""
        t = _lib.TbhTuning()
        check(_lib.tbh().tbh_get_tuning(self.handle, C.byref(t)))
        return {k: getattr(t, k) for k, _ in t._fields_}

    def set_tuning(self, **kw):
        t = _lib.TbhTuning()
        check(_lib.tbh().tbh_get_tuning(self.handle, C.byref(t)))
        for k, v in kw.items():
            if k not in dict(t._fields_):
                raise KeyError(k)
            setattr(t, k, v)
        check(_lib.tbh().tbh_set_tuning(self.handle, C.byref(t)))

    def autotune(self, branches, element_type=np.float32, max_trials=16, cache=None):
        """Times BATCHED contractions of `branches` under candidate tunings,
        installs the fastest; returns (tuning dict, ms of one contraction).

        cache: path of a JSON file of earlier results on this machine, keyed
        by the workload (element type, trial budget and a digest of every
        slice): a hit installs the stored tuning without timing anything; a
        miss tunes and adds the result to the file."""
        key = None
        if cache is not None:
            key = f"{workload_key(branches, element_type)}:{int(max_trials)}"
            stored = _read_tuning_cache(cache).get(key)
            if stored is not None:
                self.set_tuning(**stored["tuning"])
                return self.tuning(), float(stored["ms"])
        code, npdt = _dtype_code(element_type)
        m = _Marshal(branches, npdt)
        t = _lib.TbhTuning()
        ms = C.c_double()
        check(_lib.tbh().tbh_autotune(self.handle, code, m.ptr, m.n, int(max_trials), C.byref(t),
                                      C.byref(ms)))
        best = {k: getattr(t, k) for k, _ in t._fields_}
        if cache is not None:
            table = _read_tuning_cache(cache)
            table[key] = {"tuning": best, "ms": ms.value}
            tmp = f"{cache}.tmp{os.getpid()}"
            with open(tmp, "w") as f:
                json.dump(table, f, indent=1, sort_keys=True)
            os.replace(tmp, cache)
        return best, ms.value

    def valu_peak(self, element_type=np.float32):
        """Measured tropical ops/s of the GEMM's VALU instruction mix with
        register operands (the roofline ceiling of compute-bound steps)."""
        code, _ = _dtype_code(element_type)
        v = C.c_double()
        check(_lib.tbh().tbh_valu_peak(self.handle, code, C.byref(v)))
        return v.value

    def maxplus_gemm(self, A, B, C_out=None):
        """C = A (x) B in the tropical semiring; A (M,K[,batch]) and B
        (K,N[,batch]) are column-major (Fortran-ordered) numpy arrays, or
        torch tensors on this device (then C is a new device tensor)."""
        try:
            import torch
        except ImportError:  # pragma: no cover
            torch = None
        if torch is not None and isinstance(A, torch.Tensor):
            return _torch_gemm(self, A, B, C_out)
        A = np.asfortranarray(A)
        B = np.asfortranarray(B)
        if A.dtype != B.dtype:
            raise TypeError("A and B must have the same dtype")
        code, npdt = _dtype_code(A.dtype)
        batched = A.ndim == 3
        if batched != (B.ndim == 3):
            raise ValueError("A and B must both be matrices or both batches")
        M, K = A.shape[:2]
        K2, N = B.shape[:2]
        if K != K2:
            raise ValueError(f"inner dimensions differ: {K} vs {K2}")
        nb = A.shape[2] if batched else 1
        if batched and B.shape[2] != nb:
            raise ValueError("batch sizes differ")
        Cm = np.empty((M, N, nb) if batched else (M, N), dtype=npdt, order="F")
        check(_lib.tbh().tbh_maxplus_gemm(self.handle, code, M, N, K, nb, A.ctypes.data, M, M * K,
                                          B.ctypes.data, K, K * N, Cm.ctypes.data, M, M * N,
                                          _lib.TBH_HOST))
        return Cm

    def stats(self):
        """Counters; kind_* lists are [small EW, tiled EW, GEMM] (times from
        HIP events while profiling is on)."""
        s = _lib.TbhStats()
        check(_lib.tbh().tbh_get_stats(self.handle, C.byref(s)))
        out = {}
        for k, _ in s._fields_:
            v = getattr(s, k)
            out[k] = list(v) if hasattr(v, "__len__") else v
        return out

    def set_profiling(self, on=True):
        check(_lib.tbh().tbh_set_profiling(self.handle, 1 if on else 0))

    def reset_stats(self):
        check(_lib.tbh().tbh_reset_stats(self.handle))


def _torch_gemm(be, A, B, C_out):
    import torch
    if A.dtype not in (torch.float32, torch.int32) or A.dtype != B.dtype:
        raise TypeError("torch operands must both be float32 or int32")
    code = _lib.TBH_F32 if A.dtype == torch.float32 else _lib.TBH_I32
    # column-major (M,K) == row-major (K,M) transposed: take .T views
    batched = A.dim() == 3
    At = A if batched else A.unsqueeze(0)
    Bt = B if batched else B.unsqueeze(0)
    # expect (batch, K, M) row-major storage == column-major (M, K) per batch
    nb, K, M = At.shape
    nb2, N, K2 = Bt.shape
    if K != K2 or nb != nb2:
        raise ValueError("shape mismatch: A is (batch, K, M), B is (batch, N, K) row-major")
    At = At.contiguous()
    Bt = Bt.contiguous()
    Ct = torch.empty((nb, N, M), dtype=A.dtype, device=A.device) if C_out is None else C_out
    check(_lib.tbh().tbh_maxplus_gemm(be.handle, code, M, N, K, nb, At.data_ptr(), M, M * K,
                                      Bt.data_ptr(), K, K * N, Ct.data_ptr(), M, M * N,
                                      _lib.TBH_DEVICE))
    return Ct if batched else Ct[0]


def _mode(mode):
    if mode in ("batched", _lib.TBH_BATCHED):
        return _lib.TBH_BATCHED
    if mode in ("plain", _lib.TBH_PLAIN):
        return _lib.TBH_PLAIN
    raise ValueError("mode must be 'batched' or 'plain'")


_default = None


def default_backend():
    global _default
    if _default is None:
        _default = HIPBackend(0)
    return _default


def _require_cuda(usecuda):
    if not usecuda:
        raise NotImplementedError(
            "this backend contracts on the MI355X only (usecuda=true); there is no CPU path")


# ------------------------------------------------ reference-named API -----
def contract_slices(branches, element_type=np.float32, usecuda=True, *, mode="batched",
                    backend=None):
    """contract_slices(branches, T, usecuda) -> per-slice results (+r)."""
    _require_cuda(usecuda)
    return (backend or default_backend()).contract_slices(branches, element_type, mode)


def solve_slice(branch, element_type=np.float32, usecuda=True, *, mode="batched", backend=None):
    """solve_slice(branch, T, usecuda): the slice's contraction without r."""
    _require_cuda(usecuda)
    return (backend or default_backend()).solve_slice(branch, element_type, mode)


def _slice_struct(branch):
    m = _Marshal([branch], np.float32)
    return m


def complexity(branch):
    """tc, sc, rwc with every label of size 2 (zeros for an empty slice)."""
    if branch.code is None:
        return ContractionComplexity(0.0, 0.0, 0.0)
    return code_complexity(branch.code, branch.p.g.nv)


def code_complexity(code, nv=None):
    nv = int(code.ix.max()) + 1 if nv is None and len(code.ix) else (nv or 0)
    s = _lib.TbhSlice(nv, None, code.nleaves, code.ix_ptr.ctypes.data, code.ix.ctypes.data,
                      code.tree.ctypes.data if len(code.tree) else None, 0.0)
    tc, sc, rw = C.c_double(), C.c_double(), C.c_double()
    check(_lib.tbh().tbh_complexity(C.byref(s), C.byref(tc), C.byref(sc), C.byref(rw)))
    return ContractionComplexity(tc.value, sc.value, rw.value)


def tc(branch):
    return complexity(branch).tc


def sc(branch):
    return complexity(branch).sc


def _mem(fn, code, nv):
    nv = int(code.ix.max()) + 1 if nv is None and len(code.ix) else (nv or 0)
    s = _lib.TbhSlice(nv, None, code.nleaves, code.ix_ptr.ctypes.data, code.ix.ctypes.data,
                      code.tree.ctypes.data if len(code.tree) else None, 0.0)
    v = C.c_double()
    check(fn(C.byref(s), C.byref(v)))
    return v.value


def contraction_peak_memory(code, nv=None):
    """log2 of the peak live element count (src/utils.jl:197-219)."""
    return _mem(_lib.tbh().tbh_peak_memory, code, nv)


def contraction_all_memory(code, nv=None):
    """log2 of all elements allocated by internal nodes (src/utils.jl:221-229)."""
    return _mem(_lib.tbh().tbh_all_memory, code, nv)


def plan_steps(branch, element_type=np.float32, mode="batched"):
    """The steps the executor runs for one slice, as a list of dicts."""
    code, npdt = _dtype_code(element_type)
    m = _Marshal([branch], npdt)
    n = _lib.tbh().tbh_plan_steps(m.ptr, code, _mode(mode), None, 0)
    if n < 0:
        check(n)
    arr = (_lib.TbhStepInfo * max(1, n))()
    check(min(0, _lib.tbh().tbh_plan_steps(m.ptr, code, _mode(mode), arr, n)))
    return [{k: getattr(arr[i], k) for k, _ in _lib.TbhStepInfo._fields_} for i in range(n)]


def slice_footprint(branch, element_type=np.float32):
    """Device bytes the batched executor reserves for the slice."""
    code, npdt = _dtype_code(element_type)
    m = _Marshal([branch], npdt)
    v = C.c_size_t()
    check(_lib.tbh().tbh_slice_footprint(m.ptr, code, C.byref(v)))
    return v.value
