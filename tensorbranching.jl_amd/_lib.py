"""ctypes bindings of libtbh.so (include/tbh.h) and libtbh_workload.so
(include/tbh_workload.h).

The product path is libtbh.so only: loading fails loudly when it is missing,
and nothing here computes a contraction on the CPU.
"""
import ctypes as C
import os

HERE = os.path.dirname(os.path.abspath(__file__))
LIBDIR = os.path.join(HERE, "lib")

TBH_F32, TBH_I32 = 0, 1
TBH_PLAIN, TBH_BATCHED = 0, 1
TBH_HOST, TBH_DEVICE = 0, 1
TBH_I32_NEG_INF = -(1 << 30)
TBH_EINVAL, TBH_ENOMEM, TBH_EDEVICE, TBH_EBUSY, TBH_ENOTFOUND = -1, -2, -3, -4, -5

STATUS = {0: "queued", 1: "running", 2: "done", 3: "failed"}


class TbhSlice(C.Structure):
    _fields_ = [("nv", C.c_int32), ("weights", C.c_void_p), ("nleaves", C.c_int32),
                ("ix_ptr", C.c_void_p), ("ix", C.c_void_p), ("tree", C.c_void_p),
                ("r", C.c_double)]


class TbhServeConfig(C.Structure):
    _fields_ = [("max_resident_slices", C.c_size_t), ("round_log2_macs", C.c_double)]


class TbhRegion(C.Structure):
    _fields_ = [("nv", C.c_int32), ("adj", C.c_void_p), ("weights", C.c_void_p),
                ("open", C.c_uint32)]


class TbhRequest(C.Structure):
    _fields_ = [("dtype", C.c_int), ("slices", C.POINTER(TbhSlice)), ("n", C.c_size_t),
                ("priority", C.c_int32)]


class TbhStats(C.Structure):
    _fields_ = [("launches", C.c_uint64), ("waves", C.c_uint64), ("ew_ops", C.c_uint64),
                ("gemm_ops", C.c_uint64), ("leaf_tensors", C.c_uint64),
                ("ew_ops_algo_bytes", C.c_double), ("gemm_ops_algo_ops", C.c_double),
                ("total_algo_ops", C.c_double), ("pool_bytes", C.c_size_t),
                ("pool_peak_bytes", C.c_size_t), ("kind_ms", C.c_double * 3),
                ("kind_bytes", C.c_double * 3), ("kind_macs", C.c_double * 3),
                ("kind_launches", C.c_uint64 * 3), ("host_ms", C.c_double * 4)]


class TbhTuning(C.Structure):
    _fields_ = [("chain_big", C.c_int32), ("chain_side", C.c_int32), ("chain_reduce", C.c_int32),
                ("split_tiles_bits", C.c_int32), ("split_min_k_bits", C.c_int32),
                ("ew_chunk", C.c_int32), ("direct_chunk", C.c_int32), ("ew_per_cu", C.c_int32),
                ("direct_max_reduce", C.c_int32), ("concurrent", C.c_int32),
                ("first_batch", C.c_double), ("ew_prefetch", C.c_int32),
                ("ew_grid_mult", C.c_int32), ("gemm_grid_mult", C.c_int32),
                ("gemm_tiles", C.c_int32), ("chain_cost", C.c_int32)]


class TbhStepInfo(C.Structure):
    _fields_ = [("kind", C.c_int32), ("level", C.c_int32), ("rank_a", C.c_int32),
                ("rank_b", C.c_int32), ("rank_c", C.c_int32), ("n_reduced", C.c_int32),
                ("M", C.c_int32), ("N", C.c_int32), ("K", C.c_int32), ("batch", C.c_int32),
                ("lanes_a", C.c_int32), ("lanes_b", C.c_int32), ("tile_bits", C.c_int32),
                ("outer_reduced", C.c_int32), ("direct", C.c_int32), ("tile_a_only", C.c_int32),
                ("tile_b_only", C.c_int32), ("outer_product", C.c_int32),
                ("gemm_store", C.c_int32), ("macs", C.c_double), ("bytes", C.c_double)]


class TbwSliceView(C.Structure):
    _fields_ = [("nv", C.c_int32), ("ne", C.c_int32), ("edges", C.POINTER(C.c_int32)),
                ("weights", C.POINTER(C.c_double)), ("nleaves", C.c_int32),
                ("ix_ptr", C.POINTER(C.c_int32)), ("ix", C.POINTER(C.c_int32)),
                ("tree", C.POINTER(C.c_int32)), ("r", C.c_double), ("sc", C.c_double),
                ("tc", C.c_double)]


# Every function the C headers declare, with its ctypes signature.
TBH_SIGNATURES = {
    "tbh_create": (C.c_int, [C.c_int, C.c_size_t, C.POINTER(C.c_void_p)]),
    "tbh_destroy": (C.c_int, [C.c_void_p]),
    "tbh_last_error": (C.c_char_p, []),
    "tbh_contract_slices": (C.c_int, [C.c_void_p, C.c_int, C.POINTER(TbhSlice), C.c_size_t,
                                      C.c_void_p, C.c_int]),
    "tbh_solve_slice": (C.c_int, [C.c_void_p, C.c_int, C.POINTER(TbhSlice), C.c_void_p, C.c_int]),
    "tbh_contract_max": (C.c_int, [C.c_void_p, C.c_int, C.POINTER(TbhSlice), C.c_size_t,
                                   C.c_void_p, C.c_int]),
    "tbh_maxplus_gemm": (C.c_int, [C.c_void_p, C.c_int, C.c_int64, C.c_int64, C.c_int64, C.c_int64,
                                   C.c_void_p, C.c_int64, C.c_int64, C.c_void_p, C.c_int64,
                                   C.c_int64, C.c_void_p, C.c_int64, C.c_int64, C.c_int]),
    "tbh_complexity": (C.c_int, [C.POINTER(TbhSlice), C.POINTER(C.c_double),
                                 C.POINTER(C.c_double), C.POINTER(C.c_double)]),
    "tbh_peak_memory": (C.c_int, [C.POINTER(TbhSlice), C.POINTER(C.c_double)]),
    "tbh_all_memory": (C.c_int, [C.POINTER(TbhSlice), C.POINTER(C.c_double)]),
    "tbh_slice_footprint": (C.c_int, [C.POINTER(TbhSlice), C.c_int, C.POINTER(C.c_size_t)]),
    "tbh_valu_peak": (C.c_int, [C.c_void_p, C.c_int, C.POINTER(C.c_double)]),
    "tbh_get_tuning": (C.c_int, [C.c_void_p, C.POINTER(TbhTuning)]),
    "tbh_set_tuning": (C.c_int, [C.c_void_p, C.POINTER(TbhTuning)]),
    "tbh_autotune": (C.c_int, [C.c_void_p, C.c_int, C.POINTER(TbhSlice), C.c_size_t, C.c_int,
                               C.POINTER(TbhTuning), C.POINTER(C.c_double)]),
    "tbh_serve_start": (C.c_int, [C.c_void_p, C.c_size_t]),
    "tbh_serve_stop": (C.c_int, [C.c_void_p]),
    "tbh_submit": (C.c_int, [C.c_void_p, C.POINTER(TbhRequest), C.POINTER(C.c_uint64)]),
    "tbh_poll": (C.c_int, [C.c_void_p, C.c_uint64, C.c_void_p, C.POINTER(C.c_int),
                           C.POINTER(C.c_double), C.POINTER(C.c_double)]),
    "tbh_wait": (C.c_int, [C.c_void_p, C.c_uint64, C.c_void_p, C.POINTER(C.c_int)]),
    "tbh_cancel": (C.c_int, [C.c_void_p, C.c_uint64]),
    "tbh_get_stats": (C.c_int, [C.c_void_p, C.POINTER(TbhStats)]),
    "tbh_plan_steps": (C.c_int, [C.POINTER(TbhSlice), C.c_int, C.c_int, C.POINTER(TbhStepInfo),
                                 C.c_int]),
    "tbh_reset_stats": (C.c_int, [C.c_void_p]),
    "tbh_serve_start_ex": (C.c_int, [C.c_void_p, C.POINTER(TbhServeConfig)]),
    "tbh_serve_stats": (C.c_int, [C.c_void_p, C.POINTER(C.c_uint64), C.POINTER(C.c_uint64),
                                  C.POINTER(C.c_uint64)]),
    "tbh_region_tables": (C.c_int, [C.c_void_p, C.c_int, C.POINTER(TbhRegion), C.c_size_t,
                                    C.c_void_p, C.c_void_p, C.c_void_p, C.c_size_t,
                                    C.POINTER(C.c_size_t)]),
    "tbh_contract_slices_noalloc": (C.c_int, [C.c_void_p, C.c_int, C.POINTER(TbhSlice),
                                              C.c_size_t, C.c_void_p]),
    "tbh_set_profiling": (C.c_int, [C.c_void_p, C.c_int]),
}

TBW_SIGNATURES = {
    "tbw_last_error": (C.c_char_p, []),
    "tbw_random_regular": (C.c_void_p, [C.c_int32, C.c_int32, C.c_uint64]),
    "tbw_erdos_renyi": (C.c_void_p, [C.c_int32, C.c_double, C.c_uint64]),
    "tbw_ksg": (C.c_void_p, [C.c_int32, C.c_int32, C.c_double, C.c_uint64]),
    "tbw_graph_from_edges": (C.c_void_p, [C.c_int32, C.c_int32, C.c_void_p]),
    "tbw_graph_nv": (C.c_int32, [C.c_void_p]),
    "tbw_graph_ne": (C.c_int32, [C.c_void_p]),
    "tbw_graph_edges": (C.POINTER(C.c_int32), [C.c_void_p]),
    "tbw_graph_free": (None, [C.c_void_p]),
    "tbw_elimination_order": (C.c_int, [C.c_void_p, C.c_int32, C.c_uint64, C.c_void_p]),
    "tbw_slice": (C.c_void_p, [C.c_void_p, C.c_void_p, C.c_int32, C.c_int32, C.c_uint64,
                               C.c_int32, C.c_int32]),
    "tbw_slice_limited": (C.c_void_p, [C.c_void_p, C.c_void_p, C.c_int32, C.c_int32, C.c_uint64,
                                       C.c_int32, C.c_int32, C.c_int64]),
    "tbw_whole": (C.c_void_p, [C.c_void_p, C.c_void_p, C.c_int32, C.c_uint64]),
    "tbw_sliceset_count": (C.c_int64, [C.c_void_p]),
    "tbw_sliceset_get": (C.c_int, [C.c_void_p, C.c_int64, C.POINTER(TbwSliceView)]),
    "tbw_sliceset_free": (None, [C.c_void_p]),
}


def _load(name, sigs):
    path = os.path.join(LIBDIR, name)
    if not os.path.exists(path):
        raise RuntimeError(
            f"{path} is not built: run tensorbranching.jl_amd/build.py (or __graft_entry__.build())")
    lib = C.CDLL(path)
    for fn, (res, args) in sigs.items():
        f = getattr(lib, fn)
        f.restype = res
        f.argtypes = args
    return lib


_tbh = None
_tbw = None


def tbh():
    """libtbh.so (the HIP backend); raises if it is not built."""
    global _tbh
    if _tbh is None:
        _tbh = _load("libtbh.so", TBH_SIGNATURES)
    return _tbh


def tbw():
    """libtbh_workload.so (host producers of slices)."""
    global _tbw
    if _tbw is None:
        _tbw = _load("libtbh_workload.so", TBW_SIGNATURES)
    return _tbw


class TbhError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"tbh error {code}: {msg}")
        self.code = code


def check(rc):
    if rc != 0:
        raise TbhError(rc, tbh().tbh_last_error().decode(errors="replace"))
    return rc
