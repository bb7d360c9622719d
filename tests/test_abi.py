"""The C-ABI libraries load without a GPU and export every function their
headers declare; host-only entry points behave (no device calls here)."""
import ctypes as C
import os
import re

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _declared(header, prefix):
    text = open(os.path.join(ROOT, "include", header)).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(" + prefix + r"_[a-z0-9_]+)\s*\(", text)))


def test_tbh_exports_everything_declared(tb):
    from tensorbranching_jl_amd import _lib
    lib = _lib.tbh()
    names = _declared("tbh.h", "tbh")
    assert "tbh_contract_slices" in names and "tbh_maxplus_gemm" in names
    for n in names:
        assert hasattr(lib, n), n
    assert set(names) == set(_lib.TBH_SIGNATURES), "ctypes table and header differ"


def test_workload_exports_everything_declared(tb):
    from tensorbranching_jl_amd import _lib
    lib = _lib.tbw()
    names = _declared("tbh_workload.h", "tbw")
    for n in names:
        assert hasattr(lib, n), n
    assert set(names) == set(_lib.TBW_SIGNATURES)


def _slice_struct(tb, br):
    from tensorbranching_jl_amd.api import _Marshal
    return _Marshal([br], np.float32)


def test_host_complexity_and_memory_match_oracle(tb, orc):
    g = tb.random_regular_graph(80, 3, seed=11)
    for b in tb.slice_graph(g, None, tb.ContractionTreeSlicer(sc_target=9, order_trials=8))[:20]:
        if b.code is None:
            continue
        cc = tb.complexity(b)
        tc, sc, rw = orc.complexity(b.code)
        assert (cc.tc, cc.sc, cc.rwc) == pytest.approx((tc, sc, rw))
        assert tb.contraction_peak_memory(b.code, b.p.g.nv) == pytest.approx(orc.peak_memory(b.code))
        assert tb.contraction_all_memory(b.code, b.p.g.nv) == pytest.approx(orc.all_memory(b.code))


def test_malformed_slices_are_rejected(tb):
    from tensorbranching_jl_amd import _lib
    lib = _lib.tbh()
    ix_ptr = np.array([0, 1, 2, 4], np.int32)
    ix = np.array([0, 1, 0, 1], np.int32)

    def call(tree, nv=2, ixp=ix_ptr, ixx=ix):
        s = _lib.TbhSlice(nv, None, len(ixp) - 1, ixp.ctypes.data, ixx.ctypes.data,
                          tree.ctypes.data, 0.0)
        tc, sc, rw = C.c_double(), C.c_double(), C.c_double()
        return lib.tbh_complexity(C.byref(s), C.byref(tc), C.byref(sc), C.byref(rw))

    assert call(np.array([0, 2, 3, 1], np.int32)) == 0          # ((0,2),1)... valid
    assert call(np.array([0, 0, 3, 1], np.int32)) == _lib.TBH_EINVAL  # leaf used twice
    assert b"malformed" in lib.tbh_last_error()
    assert call(np.array([0, 4, 3, 1], np.int32)) == _lib.TBH_EINVAL  # forward reference
    bad_ix = np.array([0, 1, 0, 5], np.int32)
    assert call(np.array([0, 2, 3, 1], np.int32), ixx=bad_ix) == _lib.TBH_EINVAL
    assert b"out of range" in lib.tbh_last_error()
    rank3 = np.array([0, 1, 2, 5], np.int32)
    assert call(np.array([0, 2, 3, 1], np.int32), ixp=rank3,
                ixx=np.array([0, 1, 0, 1, 1], np.int32)) == _lib.TBH_EINVAL


def test_int32_weight_bound(tb):
    """An Int32 slice needs sum |w| + |r| < 2^28 (every finite value of its
    contraction is a subset sum of the weights plus r, and entries below
    -2^29 are the tropical zero): just inside the bound plans, at it the
    slice is refused before anything runs."""
    from tensorbranching_jl_amd import _lib
    g = tb.random_regular_graph(16, 3, seed=1)
    code = tb.initialize_code(g, 10)

    def plan(w, r=0):
        b = tb.SlicedBranch(tb.MISProblem(g, np.asarray(w, np.int32)), code, r)
        return tb.plan_steps(b, np.int32)

    lim = 1 << 28
    w = np.full(16, lim // 16 - 1, np.int64)
    assert plan(w, r=15)                              # sum + |r| = 2^28 - 1
    with pytest.raises(_lib.TbhError, match="sum"):
        plan(w, r=16)                                 # = 2^28
    with pytest.raises(_lib.TbhError, match="sum"):
        plan(-w, r=-16)                               # negative weights count too
    with pytest.raises(_lib.TbhError):
        plan(np.full(16, lim, np.int64))              # one weight beyond 2^28


def test_plan_steps_cover_every_node(tb):
    g = tb.random_regular_graph(60, 3, seed=2)
    b = tb.SlicedBranch(tb.MISProblem(g), tb.initialize_code(g, 8), 0)
    nodes = b.code.nleaves - 1
    plain = tb.plan_steps(b, mode="plain")
    batched = tb.plan_steps(b, mode="batched")
    # every tree node is at least one step; batched never adds OMEinsum permutes
    assert len(plain) >= nodes and len(batched) >= nodes
    assert len(batched) <= len(plain)
    assert max(s["rank_c"] for s in batched) <= tb.complexity(b).sc + 1
    levels = [s["level"] for s in batched]
    assert min(levels) >= 1
    assert [s["level"] for s in plain] == list(range(1, len(plain) + 1))


def test_gemm_output_store_modes(tb):
    """Batched plans write a GEMM node's output straight into the label order
    its consumer asks for (as row quads, or staged through LDS) instead of a
    permute step after it; plain plans (OMEinsum's lowering) never do."""
    g = tb.random_regular_graph(200, 3, seed=3)
    br = tb.slice_graph(g, None, tb.ContractionTreeSlicer(sc_target=22, order_trials=30), seed=3)
    stores = {"batched": set(), "plain": set()}
    for b in br:
        if b.code is None:
            continue
        for mode in stores:
            stores[mode] |= {s["gemm_store"] for s in tb.plan_steps(b, mode=mode) if s["kind"] == 1}
    assert stores["batched"] == {0, 1, 2}
    assert stores["plain"] == {0}


def test_footprint_bounds_reference_model(tb):
    """The batched arena holds at least the largest tensor and at most a
    small multiple of the reference's peak-memory model (src/utils.jl:197-219,
    x 4 bytes): the arena also holds every leaf tensor and the re-associated
    side tensors, and blocks are reused only within a size class.  On the 766
    slices of the n=300 workload the ratio is 0.59-2.05 (median 1.43), on the
    n=100 sc_target 10 slices 1.23-2.34; the bound is 3."""
    g = tb.random_regular_graph(70, 3, seed=6)
    b = tb.SlicedBranch(tb.MISProblem(g), tb.initialize_code(g, 8), 0)
    fp = tb.slice_footprint(b)
    peak = 2 ** tb.contraction_peak_memory(b.code, g.nv) * 4
    assert fp >= 4 * 2 ** tb.complexity(b).sc
    assert fp <= 3 * peak
    for n, seed, sc in ((100, 0, 10), (160, 4, 14)):
        g = tb.random_regular_graph(n, 3, seed=seed)
        br = tb.slice_graph(g, None, tb.ContractionTreeSlicer(sc_target=sc, order_trials=10),
                            seed=seed)
        for b in br:
            if b.code is not None:
                peak = 2 ** tb.contraction_peak_memory(b.code, b.p.g.nv) * 4
                assert 4 * 2 ** tb.sc(b) <= tb.slice_footprint(b) <= 3 * peak, n


def test_serving_entry_points_reject_null_context(tb):
    """Serving calls validate their arguments before touching a device."""
    from tensorbranching_jl_amd import _lib
    lib = _lib.tbh()
    t = C.c_uint64()
    st = C.c_int()
    assert lib.tbh_serve_start(None, 0) == _lib.TBH_EINVAL
    assert lib.tbh_serve_stop(None) == 0
    assert lib.tbh_submit(None, None, C.byref(t)) == _lib.TBH_EINVAL
    assert lib.tbh_poll(None, 1, None, C.byref(st), None, None) == _lib.TBH_EINVAL
    assert lib.tbh_wait(None, 1, None, C.byref(st)) == _lib.TBH_EINVAL
    assert lib.tbh_cancel(None, 1) == _lib.TBH_EINVAL
    assert b"context" in lib.tbh_last_error()


def test_tuning_struct_and_null_context(tb):
    from tensorbranching_jl_amd import _lib
    lib = _lib.tbh()
    assert C.sizeof(_lib.TbhTuning) == 72  # 10 int32, a double, 5 int32, padding to 8
    t = _lib.TbhTuning()
    assert lib.tbh_get_tuning(None, C.byref(t)) == _lib.TBH_EINVAL
    assert lib.tbh_set_tuning(None, C.byref(t)) == _lib.TBH_EINVAL
    ms = C.c_double()
    assert lib.tbh_autotune(None, 0, None, 0, 4, C.byref(t), C.byref(ms)) == _lib.TBH_EINVAL


def test_workload_key_tracks_every_slice_field(tb):
    """The autotuner cache key: equal for the same slice set, different when
    a weight, r, the element type or the slice order changes."""
    g = tb.random_regular_graph(40, 3, seed=4)
    br = tb.slice_graph(g, None, tb.ContractionTreeSlicer(sc_target=6))
    assert len(br) >= 2
    k = tb.workload_key(br, np.float32)
    assert k == tb.workload_key(list(br), np.float32)
    assert k != tb.workload_key(br, np.int32)
    assert k != tb.workload_key(br[::-1], np.float32)
    b0 = br[0]
    shifted = tb.SlicedBranch(b0.p, b0.code, b0.r + 1)
    assert k != tb.workload_key([shifted] + list(br[1:]), np.float32)
    w = np.ones(b0.p.g.nv, np.float32)
    w[0] = 2.0
    weighted = tb.SlicedBranch(tb.MISProblem(b0.p.g, w), b0.code, b0.r)
    assert k != tb.workload_key([weighted] + list(br[1:]), np.float32)
