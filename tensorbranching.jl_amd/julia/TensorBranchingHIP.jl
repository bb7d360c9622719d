# TensorBranchingHIP.jl -- ccall shim binding libtbh.so (include/tbh.h) into
# TensorBranching.jl's contraction step.
#
# Drop-in for contract_slices / solve_slice (src/dynamic_ob.jl:30-48): the
# same names and arguments, with the backend object in place of usecuda.
# Slices are marshalled straight from SlicedBranch / CompressedEinsum
# (src/types.jl:51-58, 85-103) without uncompress; labels and leaf indices
# become 0-based.  Errors come back as Julia exceptions carrying
# tbh_last_error().
module TensorBranchingHIP

using TensorBranching
using TensorBranching: SlicedBranch, CompressedEinsum
using TensorBranching.OMEinsum.OMEinsumContractionOrders: ContractionTree
using TensorBranching.Graphs: nv
using TensorBranching.OptimalBranching.OptimalBranchingMIS: UnitWeight

const libtbh = get(ENV, "TBH_LIB", joinpath(@__DIR__, "..", "lib", "libtbh.so"))

const TBH_F32 = Cint(0)
const TBH_I32 = Cint(1)
const TBH_PLAIN = Cint(0)
const TBH_BATCHED = Cint(1)

struct TbhSlice            # tbh_slice
    nv::Int32
    weights::Ptr{Cvoid}
    nleaves::Int32
    ix_ptr::Ptr{Int32}
    ix::Ptr{Int32}
    tree::Ptr{Int32}
    r::Float64
end

mutable struct HIPBackend
    ctx::Ptr{Cvoid}
    mode::Cint
    function HIPBackend(; device::Integer = 0, pool_bytes::Integer = 0, batched::Bool = true)
        out = Ref{Ptr{Cvoid}}(C_NULL)
        rc = ccall((:tbh_create, libtbh), Cint, (Cint, Csize_t, Ptr{Ptr{Cvoid}}),
                   device, pool_bytes, out)
        rc == 0 || error(unsafe_string(ccall((:tbh_last_error, libtbh), Cstring, ())))
        b = new(out[], batched ? TBH_BATCHED : TBH_PLAIN)
        finalizer(b) do x
            x.ctx == C_NULL || ccall((:tbh_destroy, libtbh), Cint, (Ptr{Cvoid},), x.ctx)
            x.ctx = C_NULL
        end
        return b
    end
end

dtype_code(::Type{Float32}) = TBH_F32
dtype_code(::Type{Int32}) = TBH_I32
dtype_code(T::Type) = error("element_type must be Float32 or Int32, got $T")

# post-order child list of a ContractionTree whose leaves index ixs (1-based)
function postorder(ct, nleaves::Int)
    out = Int32[]
    function walk(node)
        node isa Integer && return Int32(node - 1)
        l = walk(node.left)
        r = walk(node.right)
        push!(out, l, r)
        return Int32(nleaves + length(out) ÷ 2 - 1)
    end
    walk(ct)
    return out
end

# flat buffers of one branch; kept alive by the caller around the ccall
function marshal(b::SlicedBranch, ::Type{T}) where {T}
    g = b.p.g
    w = b.p.weights isa UnitWeight ? nothing : T.(b.p.weights)
    if nv(g) == 0 || b.code === nothing
        return (w, Int32[], Int32[], Int32[], 0)
    end
    ixs = b.code.ixs
    ix_ptr = Int32[0; cumsum(Int32.(length.(ixs)))]
    ix = Int32[l - 1 for x in ixs for l in x]
    tree = postorder(b.code.ct, length(ixs))
    return (w, ix_ptr, ix, tree, length(ixs))
end

function contract_slices(branches::Vector{<:SlicedBranch}, element_type::Type, be::HIPBackend)
    bufs = [marshal(b, element_type) for b in branches]
    res = Vector{element_type}(undef, length(branches))
    GC.@preserve bufs res begin
        slices = [TbhSlice(nv(b.p.g),
                           buf[1] === nothing ? C_NULL : pointer(buf[1]),
                           Int32(buf[5]), pointer(buf[2]), pointer(buf[3]), pointer(buf[4]),
                           Float64(b.r)) for (b, buf) in zip(branches, bufs)]
        GC.@preserve slices begin
            rc = ccall((:tbh_contract_slices, libtbh), Cint,
                       (Ptr{Cvoid}, Cint, Ptr{TbhSlice}, Csize_t, Ptr{Cvoid}, Cint),
                       be.ctx, dtype_code(element_type), slices, length(slices), res, be.mode)
        end
    end
    rc == 0 || error(unsafe_string(ccall((:tbh_last_error, libtbh), Cstring, ())))
    return res
end

function solve_slice(b::SlicedBranch, element_type::Type, be::HIPBackend)
    buf = marshal(b, element_type)
    res = Ref{element_type}()
    GC.@preserve buf begin
        s = Ref(TbhSlice(nv(b.p.g), buf[1] === nothing ? C_NULL : pointer(buf[1]),
                         Int32(buf[5]), pointer(buf[2]), pointer(buf[3]), pointer(buf[4]),
                         Float64(b.r)))
        rc = ccall((:tbh_solve_slice, libtbh), Cint,
                   (Ptr{Cvoid}, Cint, Ptr{TbhSlice}, Ptr{Cvoid}, Cint),
                   be.ctx, dtype_code(element_type), s, res, be.mode)
    end
    rc == 0 || error(unsafe_string(ccall((:tbh_last_error, libtbh), Cstring, ())))
    return res[]
end

# ---- serving: continuous batching (tbh_serve_start / tbh_submit / tbh_wait)
struct TbhRequest          # tbh_request
    dtype::Cint
    slices::Ptr{TbhSlice}
    n::Csize_t
    priority::Int32
end

struct TbhServeConfig       # tbh_serve_config
    max_resident_slices::Csize_t   # slices per round, 0: no cap
    round_log2_macs::Float64       # round work bound (log2 MACs), 0: default (38)
end

start_serving(be::HIPBackend; max_resident_slices::Integer = 0, round_log2_macs::Real = 0) =
    check(ccall((:tbh_serve_start_ex, libtbh), Cint, (Ptr{Cvoid}, Ref{TbhServeConfig}), be.ctx,
                TbhServeConfig(max_resident_slices, round_log2_macs)))
stop_serving(be::HIPBackend) = check(ccall((:tbh_serve_stop, libtbh), Cint, (Ptr{Cvoid},), be.ctx))

# (rounds, rounds run again request by request, requests that ended FAILED)
function serve_stats(be::HIPBackend)
    r, rr, f = Ref{UInt64}(0), Ref{UInt64}(0), Ref{UInt64}(0)
    check(ccall((:tbh_serve_stats, libtbh), Cint,
                (Ptr{Cvoid}, Ptr{UInt64}, Ptr{UInt64}, Ptr{UInt64}), be.ctx, r, rr, f))
    return (rounds = r[], retried_rounds = rr[], failed_requests = f[])
end

function check(rc::Cint)
    rc == 0 || error(unsafe_string(ccall((:tbh_last_error, libtbh), Cstring, ())))
    return nothing
end

# submits maximum(contract_slices(branches, T)) as one request; priority = 1
# is the latency lane (a DFS caller blocked on one slice, src/slice.jl:39)
function submit(be::HIPBackend, branches::Vector{<:SlicedBranch}, element_type::Type;
                priority::Integer = 0)
    bufs = [marshal(b, element_type) for b in branches]
    ticket = Ref{UInt64}(0)
    GC.@preserve bufs begin
        slices = [TbhSlice(nv(b.p.g),
                           buf[1] === nothing ? C_NULL : pointer(buf[1]),
                           Int32(buf[5]), pointer(buf[2]), pointer(buf[3]), pointer(buf[4]),
                           Float64(b.r)) for (b, buf) in zip(branches, bufs)]
        GC.@preserve slices begin   # copied by tbh_submit
            req = Ref(TbhRequest(dtype_code(element_type), pointer(slices), length(slices),
                                 Int32(priority)))
            check(ccall((:tbh_submit, libtbh), Cint, (Ptr{Cvoid}, Ptr{TbhRequest}, Ptr{UInt64}),
                        be.ctx, req, ticket))
        end
    end
    return ticket[]
end

function wait_result(be::HIPBackend, ticket::UInt64, element_type::Type)
    res = Ref{element_type}()
    status = Ref{Cint}(0)
    check(ccall((:tbh_wait, libtbh), Cint, (Ptr{Cvoid}, UInt64, Ptr{Cvoid}, Ptr{Cint}),
                be.ctx, ticket, res, status))
    status[] == 2 || error("request $ticket failed: " *
                           unsafe_string(ccall((:tbh_last_error, libtbh), Cstring, ())))
    return res[]
end

# ---- planner / launch parameters (tbh_tuning; results never depend on them)
mutable struct TbhTuning
    chain_big::Int32; chain_side::Int32; chain_reduce::Int32
    split_tiles_bits::Int32; split_min_k_bits::Int32
    ew_chunk::Int32; direct_chunk::Int32
    ew_per_cu::Int32
    direct_max_reduce::Int32
    concurrent::Int32
    first_batch::Float64
    ew_prefetch::Int32
    ew_grid_mult::Int32; gemm_grid_mult::Int32
    gemm_tiles::Int32
    chain_cost::Int32
    TbhTuning() = new()
end

function get_tuning(be::HIPBackend)
    t = TbhTuning()
    check(ccall((:tbh_get_tuning, libtbh), Cint, (Ptr{Cvoid}, Ref{TbhTuning}), be.ctx, t))
    return t
end
set_tuning!(be::HIPBackend, t::TbhTuning) =
    check(ccall((:tbh_set_tuning, libtbh), Cint, (Ptr{Cvoid}, Ref{TbhTuning}), be.ctx, t))

# ---- branching tables of regions (the table step of optimal_branches,
# src/branch.jl:79): alpha holds, per region, the SizeMax value of each
# boundary configuration (from contract_slices, one slice per configuration);
# returns the reduced alpha and, per region, the table's configurations
struct TbhRegion           # tbh_region
    nv::Int32
    adj::Ptr{UInt32}
    weights::Ptr{Cvoid}
    open::UInt32
end

function region_tables(be::HIPBackend, element_type::Type, adj::Vector{Vector{UInt32}},
                       weights::Vector, open::Vector{UInt32}, alpha::Vector)
    n = length(adj)
    w = [isnothing(x) ? nothing : convert(Vector{element_type}, x) for x in weights]
    a = convert(Vector{element_type}, copy(alpha))
    group_ptr = zeros(Csize_t, n + 1)
    total = Ref{Csize_t}(0)
    GC.@preserve adj w begin
        regs = [TbhRegion(Int32(length(adj[i])), pointer(adj[i]),
                          isnothing(w[i]) ? C_NULL : pointer(w[i]), open[i]) for i in 1:n]
        # sizing pass, then the configurations
        check(ccall((:tbh_region_tables, libtbh), Cint,
                    (Ptr{Cvoid}, Cint, Ptr{TbhRegion}, Csize_t, Ptr{Cvoid}, Ptr{Csize_t},
                     Ptr{UInt32}, Csize_t, Ptr{Csize_t}),
                    be.ctx, dtype_code(element_type), regs, n, copy(a), group_ptr, C_NULL, 0,
                    total))
        configs = zeros(UInt32, total[])
        check(ccall((:tbh_region_tables, libtbh), Cint,
                    (Ptr{Cvoid}, Cint, Ptr{TbhRegion}, Csize_t, Ptr{Cvoid}, Ptr{Csize_t},
                     Ptr{UInt32}, Csize_t, Ptr{Csize_t}),
                    be.ctx, dtype_code(element_type), regs, n, a, group_ptr, configs,
                    length(configs), total))
    end
    return a, [configs[group_ptr[i]+1:group_ptr[i+1]] for i in 1:n]
end

# dynamic_ob_mis with the contraction step on the MI355X (src/dynamic_ob.jl:6-28)
function dynamic_ob_mis(g, weights, be::HIPBackend; reducer = TensorBranching.TensorNetworkReducer(),
                        slicer = TensorBranching.ContractionTreeSlicer(),
                        optimizer = TensorBranching.TreeSA(), element_type::Type = Float32,
                        verbose::Int = 0)
    res = TensorBranching.kernelize(g, weights, reducer, verbose = verbose)
    nv(res.g) == 0 && return res.r
    code = TensorBranching.initialize_code(res.g, optimizer)
    branches = TensorBranching.slice(TensorBranching.MISProblem(res.g, res.weights), code, res.r,
                                     slicer, reducer, verbose = verbose)
    return maximum(contract_slices(branches, element_type, be))
end

end # module
