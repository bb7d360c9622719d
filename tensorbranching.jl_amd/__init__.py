"""MI355X-native tropical contraction backend for TensorBranching.jl.

The contraction step behind contract_slices / solve_slice
(src/dynamic_ob.jl:30-48) runs in libtbh.so: hand-written gfx950 HIP
kernels (VALU max-plus GEMM, bit-permuted tropical contraction steps, leaf
builder) driven by a host planner and a wave-batched executor.  This
package is the host-side mirror of the reference interface over its C ABI
(include/tbh.h); the directory is loaded under the module name
``tensorbranching_jl_amd``.
"""
from ._lib import TbhError, TBH_I32_NEG_INF  # noqa: F401
from .api import (  # noqa: F401
    CompressedEinsum, ContractionComplexity, ContractionTree, HIPBackend, MISProblem,
    SimpleGraph, SlicedBranch, UnitWeight, code_complexity, complexity, contract_slices,
    contraction_all_memory, contraction_peak_memory, default_backend, plan_steps, sc,
    slice_footprint,
    solve_slice, tc, workload_key)
from .workload import (  # noqa: F401
    ContractionTreeSlicer, dynamic_ob_mis, erdos_renyi, initialize_code, random_ksg,
    random_regular_graph, serving_requests, slice_graph, SERVING_MIX, ServingRequest)
from .serve import Server  # noqa: F401
from .branching import (  # noqa: F401
    BranchingTable, branching_table, branching_tables, neighbor_region, open_vertices)
from .io import (  # noqa: F401
    load_all_finished, load_code, load_finished, load_weights, loadgraph, save_code,
    save_finished, save_slices, save_weights, savegraph, threaded_saveslices)
