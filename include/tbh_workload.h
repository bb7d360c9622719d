/*
 * tbh_workload.h -- host-side producers of the sliced MIS tensor networks
 * that the contraction backend (tbh.h) consumes.
 *
 * These are not on the hot path.  They stand in for the reference's host
 * algorithms, which need Julia packages absent here, so that the backend has
 * realistic inputs:
 *   tbw_random_regular   Graphs.random_regular_graph (test/slice.jl:23, README.md:30)
 *   tbw_erdos_renyi      Graphs.erdos_renyi
 *   tbw_ksg              random_ksg / random_diagonal_coupled_graph
 *                        (src/utils.jl:191-195)
 *   tbw_order_tree       initialize_code (src/dynamic_ob.jl:50-54) plus
 *                        order2eincode (src/decompose.jl:64-94): a greedy
 *                        min-fill elimination order turned into a binary tree
 *                        over rawcode(IndependentSet(g)) leaves
 *   tbw_slice            slice / slice_bfs (src/slice.jl:4-11,68-90): branch
 *                        on a vertex of the largest tensor (exclude it, or
 *                        include it and drop its neighbours), simple
 *                        kernelization of each branch (src/kernelize.jl:4-17
 *                        with a simplicial-vertex rule) and the old order
 *                        restricted to the surviving vertices (update_order,
 *                        src/decompose.jl:49-56), until sc <= sc_target
 * Graphs are 0-based edge lists.  The PRNG is SplitMix64-seeded
 * xoshiro256**; seeds index this generator, not Julia's.
 */
#ifndef TBH_WORKLOAD_H
#define TBH_WORKLOAD_H
#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct tbw_graph tbw_graph;     /* owned edge list */
typedef struct tbw_sliceset tbw_sliceset;

/* One finished slice, pointing into storage owned by its tbw_sliceset. */
typedef struct {
  int32_t nv;
  int32_t ne;
  const int32_t* edges;   /* 2*ne, 0-based (relabelled subgraph)          */
  const double* weights;  /* nv weights                                   */
  int32_t nleaves;        /* nv + ne leaves: vertex tensors, then edges   */
  const int32_t* ix_ptr;  /* nleaves + 1                                  */
  const int32_t* ix;      /* leaf labels                                  */
  const int32_t* tree;    /* 2*(nleaves-1) children, post-order           */
  double r;               /* weight fixed by branching/kernelization      */
  double sc, tc;          /* complexity of the tree, every label size 2   */
} tbw_slice_view;

const char* tbw_last_error(void);

/* generators: return a graph or NULL (see tbw_last_error) */
tbw_graph* tbw_random_regular(int32_t n, int32_t d, uint64_t seed);
tbw_graph* tbw_erdos_renyi(int32_t n, double p, uint64_t seed);
/* m x n King's subgraph: round(rho*m*n) sites drawn without replacement,
 * edges between sites at Chebyshev distance 1. */
tbw_graph* tbw_ksg(int32_t m, int32_t n, double rho, uint64_t seed);
tbw_graph* tbw_graph_from_edges(int32_t nv, int32_t ne, const int32_t* edges);
int32_t tbw_graph_nv(const tbw_graph*);
int32_t tbw_graph_ne(const tbw_graph*);
const int32_t* tbw_graph_edges(const tbw_graph*);
void tbw_graph_free(tbw_graph*);

/* Contraction order for the whole graph: best of `trials` randomized
 * min-fill orders.  Writes the elimination order (nv entries). */
int tbw_elimination_order(const tbw_graph* g, int32_t trials, uint64_t seed,
                          int32_t* order_out);

/* Slice the graph to sc <= sc_target (one slice if it already fits).
 * weights: nv values or NULL (unit); integer weights stay integers. */
tbw_sliceset* tbw_slice(const tbw_graph* g, const double* weights,
                        int32_t sc_target, int32_t order_trials, uint64_t seed,
                        int32_t nthreads, int32_t kernelize);
/* tbw_slice that gives up (NULL, "more than max_slices slices") as soon as
 * the finished and open branches outnumber max_slices; tbw_slice uses
 * 2^18. */
tbw_sliceset* tbw_slice_limited(const tbw_graph* g, const double* weights,
                                int32_t sc_target, int32_t order_trials,
                                uint64_t seed, int32_t nthreads, int32_t kernelize,
                                int64_t max_slices);
/* The unsliced network of g with its tree (a single slice, r = 0). */
tbw_sliceset* tbw_whole(const tbw_graph* g, const double* weights,
                        int32_t order_trials, uint64_t seed);
int64_t tbw_sliceset_count(const tbw_sliceset*);
int tbw_sliceset_get(const tbw_sliceset*, int64_t i, tbw_slice_view* out);
void tbw_sliceset_free(tbw_sliceset*);

#ifdef __cplusplus
}
#endif
#endif
