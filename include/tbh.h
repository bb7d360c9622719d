/*
 * tbh.h -- C ABI of the MI355X tropical contraction backend (libtbh.so).
 *
 * Drop-in for the contraction step of TensorBranching.jl: what the Julia
 * side binds with ccall in place of GenericTensorNetworks.solve(...,
 * SizeMax(); T, usecuda) over OMEinsum + TropicalGEMM/CuTropicalGEMM.
 * Plain pointers and sizes only.  Every call returns 0 on success or a
 * negative TBH_E* code; tbh_last_error() (thread-local) says why.
 *
 * Data conventions
 *   - A slice is what SlicedBranch holds (src/types.jl:85-103): the graph's
 *     vertex count and weights, the CompressedEinsum leaf label lists `ixs`
 *     and binary ContractionTree `ct` (src/types.jl:51-58), and the fixed
 *     weight r.  Labels are 0-based vertex ids.  Leaf content follows from
 *     its rank, as GenericTensorNetworks builds IndependentSet tensors for
 *     SizeMax: [v] -> [0, w_v]; [u,v] -> [0 0; 0 -inf].
 *   - tree: the internal nodes in post-order; node j has children
 *     tree[2j], tree[2j+1]; a child id c < nleaves is leaf c, otherwise the
 *     internal node c - nleaves (< j).  The last node is the root.
 *   - Intermediate labels follow parse_eincode: a label of a subtree is kept
 *     iff it occurs outside it; the output labels of the network are empty.
 *   - Tensors are column-major: label position i has stride 2^i.
 *   - Tropical Float32: (+) = max, (x) = +, zero = -inf.  Tropical Int32:
 *     zero = TBH_I32_NEG_INF (-2^30); every stored entry below -2^29 is
 *     rewritten to it, so finite values must satisfy |x| < 2^28: an Int32
 *     slice needs sum |w| + |r| < 2^28, which bounds every finite value of
 *     its contraction (TBH_EINVAL otherwise).
 *   - Results are bit-exact for integer-valued tensors (Float32 sums below
 *     2^24): every sum is exact, so any evaluation order gives identical
 *     bits.  For non-integer Float32 values PLAIN keeps the slice's tree
 *     (max is exact and rounding monotone, so it matches any evaluation of
 *     that tree bit for bit); BATCHED may re-associate chains of the tree,
 *     which can move a result by a rounding step.
 */
#ifndef TBH_H
#define TBH_H
#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define TBH_I32_NEG_INF (-(1 << 30))

typedef enum { TBH_F32 = 0, TBH_I32 = 1 } tbh_dtype;

/* PLAIN: one slice at a time, one launch per step of the OMEinsum lowering
 * (permute, GEMM, permute), synchronous per slice -- the per-slice GPU loop
 * of contract_slices(..., usecuda=true).  BATCHED: all slices of a batch run
 * wave by wave, one persistent launch per wave and kernel kind. */
typedef enum { TBH_PLAIN = 0, TBH_BATCHED = 1 } tbh_mode;

typedef enum { TBH_HOST = 0, TBH_DEVICE = 1 } tbh_ptr_kind;

enum {
  TBH_OK = 0,
  TBH_EINVAL = -1,   /* bad argument or malformed slice */
  TBH_ENOMEM = -2,   /* does not fit the pool */
  TBH_EDEVICE = -3,  /* HIP runtime error */
  TBH_EBUSY = -4,    /* serving queue full */
  TBH_ENOTFOUND = -5 /* unknown ticket */
};

typedef struct tbh_ctx tbh_ctx;

/* One SlicedBranch (src/types.jl:85-103). */
typedef struct {
  int32_t nv;             /* vertices; 0 -> the slice's result is r          */
  const void* weights;    /* nv values of the dtype; NULL = UnitWeight       */
  int32_t nleaves;        /* length(code.ixs)                                */
  const int32_t* ix_ptr;  /* nleaves + 1 offsets into ix                     */
  const int32_t* ix;      /* leaf labels (0-based vertex ids)                */
  const int32_t* tree;    /* 2*(nleaves-1) child ids in post-order (code.ct) */
  double r;               /* branch.r                                        */
} tbh_slice;

/* Context: one GPU, one device memory pool of pool_bytes (0 = 80% of the
 * free memory), three streams at the default priority -- the main stream,
 * a second one on which BATCHED runs a wave's GEMM launches beside its EW
 * launches, and a copy stream for BATCHED descriptor uploads (event-ordered;
 * every call returns with all of them idle) -- and a host thread that issues
 * BATCHED launch sequences while the caller packs the next batch.  Replaces
 * the implicit CUDA.jl state used by solve(..., usecuda=true)
 * (src/dynamic_ob.jl:32). */
int tbh_create(int device, size_t pool_bytes, tbh_ctx** out);
int tbh_destroy(tbh_ctx* ctx);
const char* tbh_last_error(void);

/* contract_slices(branches, T, usecuda) (src/dynamic_ob.jl:36-48): results[i]
 * = solve_slice(slice i) + T(r_i), or T(r_i) when nv == 0.  results is a
 * host array of n dtype values. */
int tbh_contract_slices(tbh_ctx* ctx, tbh_dtype dtype, const tbh_slice* slices,
                        size_t n, void* results, tbh_mode mode);

/* No counterpart in the reference: tbh_contract_slices(..., TBH_PLAIN) for
 * callers that cannot allocate device memory (code under HIP graph capture,
 * tight serving loops, harnesses that forbid hipMalloc).  It allocates
 * nothing: each slice's arena and step descriptors live in a static device
 * array of TBH_NOALLOC_SCRATCH_BYTES (64 MiB) that every GPU the library runs
 * on holds until the process ends; the context's pool is not touched.  A
 * slice needing more than the array fails the call with TBH_ENOMEM before
 * any work.  All calls on one GPU share its array, so the
 * library runs them one at a time (a per-GPU lock; unserialized they would
 * race on it); results equal tbh_contract_slices(..., TBH_PLAIN) bit for
 * bit. */
#define TBH_NOALLOC_SCRATCH_BYTES (64u << 20)
int tbh_contract_slices_noalloc(tbh_ctx* ctx, tbh_dtype dtype,
                                const tbh_slice* slices, size_t n,
                                void* results);

/* solve_slice(branch, T, usecuda) (src/dynamic_ob.jl:30-34): the contraction
 * of one slice WITHOUT r (the tropical one, 0, for a slice with no leaves). */
int tbh_solve_slice(tbh_ctx* ctx, tbh_dtype dtype, const tbh_slice* slice,
                    void* result, tbh_mode mode);

/* maximum(contract_slices(...)) (src/dynamic_ob.jl:24-27) in one call. */
int tbh_contract_max(tbh_ctx* ctx, tbh_dtype dtype, const tbh_slice* slices,
                     size_t n, void* result, tbh_mode mode);

/* Max-plus GEMM, the lower seam mul!(C, A, B) on (Cu)Matrix{Tropical{T}}
 * that TropicalGEMM / CuTropicalGEMM supply (README.md:14-22):
 *   C[:,:,t] = A[:,:,t] (x) B[:,:,t],  C[m,n] = max_k A[m,k] + B[k,n]
 * column-major with leading dimensions and batch strides (in elements).
 * ptr_kind TBH_HOST: host buffers (copied through the pool; entries of C
 * outside the M x N blocks -- rows M..ldc-1, gaps between batch entries --
 * come back unchanged); TBH_DEVICE: device pointers, run on the context's
 * stream, synchronous. */
int tbh_maxplus_gemm(tbh_ctx* ctx, tbh_dtype dtype, int64_t M, int64_t N,
                     int64_t K, int64_t batch, const void* A, int64_t lda,
                     int64_t strideA, const void* B, int64_t ldb,
                     int64_t strideB, void* C, int64_t ldc, int64_t strideC,
                     tbh_ptr_kind ptr_kind);

/* Complexity with every label of size 2 (complexity(branch),
 * src/types.jl:115-121): tc, sc, rwc as log2. */
int tbh_complexity(const tbh_slice* slice, double* tc, double* sc,
                   double* rwc);

/* contraction_peak_memory / contraction_all_memory (src/utils.jl:197-229),
 * log2 element counts of the reference memory model. */
int tbh_peak_memory(const tbh_slice* slice, double* log2_elems);
int tbh_all_memory(const tbh_slice* slice, double* log2_elems);

/* Device bytes the BATCHED executor reserves for this slice (its arena, the
 * admission footprint used by the pool and the serving scheduler). */
int tbh_slice_footprint(const tbh_slice* slice, tbh_dtype dtype,
                        size_t* bytes);

/* ---- serving: continuous batching of MIS requests -------------------- */
/* A request is one MIS solve: maximum over its slices of contract_slices. */
typedef struct {
  tbh_dtype dtype;
  const tbh_slice* slices;  /* copied at submit; caller may free afterwards */
  size_t n;
  int32_t priority;         /* 0: normal; 1: latency lane (a caller blocked on
                             * one slice, as slice_dfs_lp, src/slice.jl:39):
                             * its slices go into the next round formed,
                             * ahead of every normal request's */
} tbh_request;

typedef enum {
  TBH_QUEUED = 0,
  TBH_RUNNING = 1,
  TBH_DONE = 2,
  TBH_FAILED = 3
} tbh_status;

/* Start the planner and scheduler threads.  A request is refused alone
 * (FAILED) when one of its slices' arenas (tbh_slice_footprint) exceeds the
 * region an idle pool gives a batched run, or its descriptors the pool's
 * share for them.  The scheduler runs rounds: a round admits slices --
 * latency-lane requests first, then arrival order -- while their summed
 * arenas fit that region, their descriptors that share, their summed
 * estimated work (the MACs of the slices' trees) stays within
 * 2^round_log2_macs and at most max_resident_slices (0: no cap) are in it.
 * Rounds are pipelined on the stream (the next one is enqueued while one
 * runs, and formed once the one before it is issued), so a new request
 * waits for at most the round running and the one queued.  A round whose
 * execution fails is run again request by request, and only the requests
 * that fail alone end FAILED.  While requests are in flight the scheduler
 * holds the context: synchronous calls on it wait until serving is idle. */
#define TBH_SERVE_ROUND_LOG2_MACS 38
typedef struct {
  size_t max_resident_slices; /* slices per round, 0: no cap                */
  double round_log2_macs;     /* round work bound (log2 MACs), 0: default   */
} tbh_serve_config;
int tbh_serve_start_ex(tbh_ctx* ctx, const tbh_serve_config* cfg);
/* tbh_serve_start_ex with the default round work bound. */
int tbh_serve_start(tbh_ctx* ctx, size_t max_resident_slices);
/* Rounds run, rounds run again request by request after a failed
 * execution, and requests that ended FAILED, since serving started. */
int tbh_serve_stats(tbh_ctx* ctx, uint64_t* rounds, uint64_t* retried_rounds,
                    uint64_t* failed_requests);
int tbh_serve_stop(tbh_ctx* ctx);
/* Thread-safe (many producers). */
int tbh_submit(tbh_ctx* ctx, const tbh_request* req, uint64_t* ticket);
/* Non-blocking; result (one dtype value, the MIS) valid when DONE.  Times
 * in seconds on the scheduler's monotonic clock. */
int tbh_poll(tbh_ctx* ctx, uint64_t ticket, void* result, tbh_status* status,
             double* t_submit, double* t_done);
/* Blocks until the request is DONE or FAILED. */
int tbh_wait(tbh_ctx* ctx, uint64_t ticket, void* result, tbh_status* status);
/* Drops a queued (not yet admitted) request. */
int tbh_cancel(tbh_ctx* ctx, uint64_t ticket);

/* ---- branching tables of regions --------------------------------------- */
/* A region of a MIS problem for the table step of optimal_branches
 * (branching_table(p, TensorNetworkSolver(), region), src/branch.jl:79):
 * its induced subgraph (vertex i = the region's i-th vertex), weights and
 * boundary (open) vertices -- those with a neighbour outside the region.  A
 * boundary configuration b has bit k for the region's k-th open vertex in
 * ascending vertex order. */
typedef struct {
  int32_t nv;              /* region vertices (<= 24)                          */
  const uint32_t* adj;     /* nv bitmasks: bit u of adj[v] set iff v ~ u       */
  const void* weights;     /* nv dtype values (Float32: integers below 2^20;
                            * Int32: |w| < 2^24); NULL = unit weights        */
  uint32_t open;           /* bitmask of the open vertices (<= 16 of them)     */
} tbh_region;
/* alpha: per region, 2^popcount(open) values (regions concatenated): on
 * entry the region's best weight with its open vertices fixed to each b (the
 * tropical zero where b is not independent) -- the SizeMax contraction of
 * the region's network with its boundary open, which callers obtain from
 * tbh_contract_slices (one slice per b); on return the reduced alpha tensor
 * (mis_compactify!: an entry is set to the tropical zero when a proper
 * subset of its boundary configuration reaches at least its value).
 * configs: every region configuration (bitmask over the region's vertices)
 * that is independent and whose weight equals the reduced alpha of its
 * boundary bits, i.e. the rows of the BranchingTable; region i's
 * configurations are configs[group_ptr[i] .. group_ptr[i+1]), ordered by
 * boundary configuration, then ascending.  With configs == NULL only
 * group_ptr and *total (the number of configurations) are filled; a total
 * above cap fails with TBH_ENOMEM (group_ptr and *total filled). */
int tbh_region_tables(tbh_ctx* ctx, tbh_dtype dtype, const tbh_region* regions,
                      size_t n, void* alpha, size_t* group_ptr,
                      uint32_t* configs, size_t cap, size_t* total);

/* ---- introspection ---------------------------------------------------- */
/* One planned step of a slice (no reference counterpart: used by tests and
 * the bench to account algorithmic work per kernel). */
typedef struct {
  int32_t kind;             /* 0 EW step, 1 GEMM step                       */
  int32_t level;            /* wave (BATCHED) or sequence index (PLAIN)     */
  int32_t rank_a, rank_b, rank_c;  /* rank_b = -1: unary step               */
  int32_t n_reduced;        /* EW: reduced labels                           */
  int32_t M, N, K, batch;   /* GEMM extents                                 */
  int32_t lanes_a, lanes_b; /* EW: contiguous A/B positions under lane bits */
  int32_t tile_bits;        /* EW: outputs per tile (log2)                  */
  int32_t outer_reduced;    /* EW: reduced labels looped outside LDS        */
  int32_t direct;           /* EW: 1 = direct (wave) step, 0 = LDS-tiled    */
  int32_t tile_a_only;      /* EW: tile labels of A and not B               */
  int32_t tile_b_only;      /* EW: tile labels of B and not A               */
  int32_t outer_product;    /* EW: 1 = 4 x 4 outer-product tile layout      */
  int32_t gemm_store;       /* GEMM: 0 column-major C; C in its consumer's
                               label order, 1 stored as row quads, 2 staged
                               through LDS                                  */
  double macs, bytes;       /* executed configurations, algorithmic bytes   */
} tbh_step_info;
/* Plans one slice and writes up to cap step records; returns the number of
 * steps (>= 0) or a negative error code. */
int tbh_plan_steps(const tbh_slice* slice, tbh_dtype dtype, tbh_mode mode,
                   tbh_step_info* out, int cap);

/* ---- tuning ------------------------------------------------------------ */
/* Planner and launch parameters of the BATCHED executor (no reference
 * counterpart).  chain_*: spine-chain re-association (spine rank >=
 * chain_big absorbs side tensors of rank <= chain_side, <= chain_reduce
 * labels reduced per step); split_*: GEMM split-K (up to 2^split_tiles_bits
 * tiles per node, K parts >= 2^split_min_k_bits); ew_chunk / direct_chunk:
 * tiles per claim of the tiled / direct EW kernels (a tiled claim above 4
 * is halved, down to 4, on launches with too few tiles to give every
 * workgroup of the grid such a claim); ew_per_cu: tiled EW
 * workgroups per CU in a launch's grid (at most what is resident);
 * direct_max_reduce: EW steps reducing more labels than this run LDS-tiled,
 * not as direct wave tiles (-1: all tiled); concurrent: 1 runs a level's
 * GEMM launches on a second stream beside its EW launches; first_batch:
 * fraction of the arena region the first batch may use; ew_prefetch: 1 runs
 * LDS-tiled EW steps with the next tile's sub-tensors loaded into registers
 * while the current tile computes, 0 stages each tile by DMA after the
 * previous one; ew_grid_mult / gemm_grid_mult: a tiled EW / GEMM launch's
 * grid holds this many times the workgroups that stay resident (1: a
 * persistent grid; more: later workgroups start as earlier ones finish,
 * which evens out uneven tiles); gemm_tiles: GEMM tile shapes (0: 128 x 128
 * only; 1: 256 x 64 / 256 x 32 tiles for outputs of M >= 256 rows and N <= 64
 * / 32 columns; 2: 256 x 64 up to N = 128 too); chain_cost > 0: a chain
 * rotation is taken only when it lowers the estimated cost of the two steps
 * it replaces, bytes moved + MACs * chain_cost / 64, in place of the
 * chain_reduce cap (0: the caps alone).  Results never depend on them. */
typedef struct {
  int32_t chain_big, chain_side, chain_reduce;
  int32_t split_tiles_bits, split_min_k_bits;
  int32_t ew_chunk, direct_chunk;
  int32_t ew_per_cu;
  int32_t direct_max_reduce;
  int32_t concurrent;
  double first_batch;
  int32_t ew_prefetch;
  int32_t ew_grid_mult, gemm_grid_mult;
  int32_t gemm_tiles;
  int32_t chain_cost;
} tbh_tuning;
int tbh_get_tuning(tbh_ctx* ctx, tbh_tuning* out);
int tbh_set_tuning(tbh_ctx* ctx, const tbh_tuning* t);
/* Autotuner: times BATCHED contractions of the given slices under candidate
 * tunings (coordinate descent from the current one over at most
 * max_trials - 2 candidates, each timed as the best of two contractions
 * whose results must equal the first's; a candidate replaces the current
 * tuning only when 2% faster), installs the fastest on the context and
 * returns it with its wall time.  For Float32 slices whose sums are not all
 * exact (a non-integer weight or r, or sum |w| + |r| >= 2^24) the chain_*
 * parameters are not searched: re-association may move such results by a
 * rounding step, and every candidate must reproduce the first run's bits. */
int tbh_autotune(tbh_ctx* ctx, tbh_dtype dtype, const tbh_slice* slices, size_t n,
                 int max_trials, tbh_tuning* best, double* best_ms);

/* Measured VALU ceiling of the max-plus instruction mix (no reference
 * counterpart: the denominator of the GEMM's roofline).  Runs a
 * register-only kernel with the GEMM inner loop's mix -- 2 adds and 1 max3
 * per 2 tropical MACs, 8 x 8 register blocks -- on every CU and returns
 * tropical ops/s (2 ops per MAC). */
int tbh_valu_peak(tbh_ctx* ctx, tbh_dtype dtype, double* tropical_ops_per_s);

typedef struct {
  uint64_t launches;        /* kernel launches issued                     */
  uint64_t waves;           /* batched waves executed                     */
  uint64_t ew_ops, gemm_ops, leaf_tensors;
  double ew_ops_algo_bytes; /* algorithmic bytes of EW ops (|A|+|B|+|C|)   */
  double gemm_ops_algo_ops; /* 2 * MACs of GEMM ops                        */
  double total_algo_ops;    /* 2 * MACs over all ops                       */
  size_t pool_bytes, pool_peak_bytes;
  /* kernel time per kind from HIP events on the context's stream, filled
   * while profiling is on (tbh_set_profiling): [0] small EW steps,
   * [1] tiled EW steps, [2] GEMM steps */
  double kind_ms[3];
  double kind_bytes[3];     /* algorithmic bytes of the steps timed         */
  double kind_macs[3];      /* executed (output x reduced) configurations   */
  uint64_t kind_launches[3];
  /* host wall time of contract calls (ms): [0] converting and planning the
   * slices, [1] building and packing step descriptors, [2] issuing the
   * uploads and launches, [3] waiting for the stream */
  double host_ms[4];
} tbh_stats;
/* Time every launch with HIP events (adds a little host overhead). */
int tbh_set_profiling(tbh_ctx* ctx, int on);
int tbh_get_stats(tbh_ctx* ctx, tbh_stats* out);
int tbh_reset_stats(tbh_ctx* ctx);

#ifdef __cplusplus
}
#endif
#endif
