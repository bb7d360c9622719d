/*
 * tbh_oracle.h -- CPU restatement of the reference's tropical contraction path.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the checker for the HIP backend:
 * only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may load
 * it.  The product (libtbh.so) never links, calls or falls back to it.
 *
 * What it restates (reference paths relative to the TensorBranching.jl tree):
 *   - contract_slices / solve_slice           src/dynamic_ob.jl:30-48
 *   - leaf tensors of IndependentSet, SizeMax GenericTensorNetworks generate_tensors
 *     (vertex [0, w_v], edge [[0,0],[0,-inf]]; used at test/utils.jl:24)
 *   - per-node evaluation as OMEinsum does it (reduce labels private to one
 *     input, permute to (a,k,t)/(k,b,t), batched max-plus GEMM, permute to iy),
 *     under the intermediate-label rule of parse_eincode (src/types.jl:75-79)
 *   - max-plus GEMM, the role of TropicalGEMM.mul! (test/slice.jl:5)
 *   - contraction_peak_memory / contraction_all_memory   src/utils.jl:197-229
 *   - contraction complexity tc/sc/rwc, every label of size 2
 *                                                  src/types.jl:115-121
 * plus independent MIS solvers used to pin it (brute force, exact
 * branch-and-reduce in the role of EliminateGraphs.mis2, forest DP).
 *
 * Conventions shared with the product:
 *   - labels are 0-based vertex ids; a tensor of rank r holds 2^r entries,
 *     label position i has stride 2^i (column-major, as Julia);
 *   - tree: internal nodes listed in post-order, node j has children
 *     tree[2j], tree[2j+1]; a child id c < nleaves is leaf c, otherwise it is
 *     internal node c - nleaves (which precedes j); the root is the last node;
 *   - Tropical Float32: zero element -inf; Tropical Int32: zero element is the
 *     canonical sentinel ORC_I32_NEG = -2^30 and every stored entry below
 *     -2^29 is rewritten to the sentinel, so finite values (|x| < 2^28) add
 *     without overflow and -inf-like entries stay canonical.
 */
#ifndef TBH_ORACLE_H
#define TBH_ORACLE_H
#include <stdint.h>
#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

#define ORC_F32 0
#define ORC_I32 1
#define ORC_I32_NEG (-(1 << 30))
#define ORC_I32_NEG_THRESH (-(1 << 29))

typedef struct {
  int32_t nv;             /* number of vertices; 0 = empty graph -> result r */
  const void* weights;    /* nv values of the dtype, NULL = unit weights     */
  int32_t nleaves;
  const int32_t* ix_ptr;  /* nleaves + 1 offsets                            */
  const int32_t* ix;      /* leaf labels                                     */
  const int32_t* tree;    /* 2 * (nleaves - 1) child ids, post-order         */
  double r;               /* fixed-weight offset                             */
} orc_slice;

/* Labels of every internal node (CSR), by the rule: a label of the subtree is
 * kept iff it also occurs in a leaf outside the subtree.  Output order: first
 * appearance in (ix_left, ix_right).  Returns total label count, or <0. */
int orc_node_labels(int32_t nleaves, const int32_t* ix_ptr, const int32_t* ix,
                    const int32_t* tree, int32_t* out_ptr, int32_t* out_lab,
                    int32_t out_cap);

/* tc, sc, rwc (log2) with every label of size 2. */
int orc_complexity(int32_t nleaves, const int32_t* ix_ptr, const int32_t* ix,
                   const int32_t* tree, double* tc, double* sc, double* rwc);

/* src/utils.jl:197-219 and :221-229, log2 of element counts. */
double orc_peak_memory(int32_t nleaves, const int32_t* ix_ptr,
                       const int32_t* ix, const int32_t* tree);
double orc_all_memory(int32_t nleaves, const int32_t* ix_ptr,
                      const int32_t* ix, const int32_t* tree);

/* C[:,:,t] = max_k A[:,k,t] + B[k,:,t]; column-major with leading dims and
 * batch strides, as mul! on Matrix{Tropical{T}}. */
int orc_maxplus_gemm_f32(int64_t M, int64_t N, int64_t K, int64_t batch,
                         const float* A, int64_t lda, int64_t sA,
                         const float* B, int64_t ldb, int64_t sB, float* C,
                         int64_t ldc, int64_t sC);
int orc_maxplus_gemm_i32(int64_t M, int64_t N, int64_t K, int64_t batch,
                         const int32_t* A, int64_t lda, int64_t sA,
                         const int32_t* B, int64_t ldb, int64_t sB, int32_t* C,
                         int64_t ldc, int64_t sC);

/* Binary tropical einsum exactly as OMEinsum lowers it. Labels are arbitrary
 * ints; out labels must be a subset of ixa U ixb. Values of dtype. */
int orc_einsum2(int dtype, int na, const int32_t* ixa, const void* A, int nb,
                const int32_t* ixb, const void* B, int nc, const int32_t* ixc,
                void* C);

/* solve_slice: contraction of one slice's tree, WITHOUT r. */
int orc_solve_slice(int dtype, const orc_slice* s, void* result);
/* contract_slices: per slice result (+r; nv==0 -> r). nthreads<=0: all. */
int orc_contract_slices(int dtype, const orc_slice* s, int64_t n,
                        void* results, int nthreads);

/* Independent MIS solvers (double weights; NULL = unit). */
double orc_mis_bruteforce(int32_t nv, int32_t ne, const int32_t* edges,
                          const double* w);
int64_t orc_mis_exact(int32_t nv, int32_t ne, const int32_t* edges);
double orc_mwis_forest(int32_t nv, int32_t ne, const int32_t* edges,
                       const double* w);

const char* orc_last_error(void);

#ifdef __cplusplus
}
#endif
#endif
