/*
 * tbh_oracle.c -- CPU restatement of the reference's tropical contraction path.
 *
 * TEST INFRASTRUCTURE ONLY (the checker, never the product): loaded by tests/,
 * __graft_entry__.smoke() and bench.py's cpu_baseline leg.  See tbh_oracle.h
 * for the conventions and the list of reference files restated here.
 *
 * Every function cites the reference behaviour it follows.  Parity pinning:
 *   orc_maxplus_gemm_*  pinned by tropical identities/associativity/transpose
 *                       and DAG longest-path checks (tests/test_oracle.py);
 *   orc_einsum2         pinned by brute-force enumeration over label configs
 *                       on small ranks and by GEMM special cases;
 *   orc_solve_slice /
 *   orc_contract_slices pinned by brute-force MIS, closed forms (cycles,
 *                       paths, complete bipartite, Petersen, grids), the
 *                       forest DP, the exact branch-and-reduce MIS, order
 *                       invariance and the slicing invariant;
 *   orc_complexity /
 *   orc_peak_memory /
 *   orc_all_memory      pinned by hand-computed small trees.
 */
#include "tbh_oracle.h"

#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

static __thread char g_err[512];
const char* orc_last_error(void) { return g_err; }

/* Threads inside one contraction (GEMM columns, permutes): set by
 * orc_contract_slices while it contracts a large slice on its own; 1 while
 * slices run one per thread.  Only loops whose iterations write disjoint
 * outputs are split, so results do not depend on it. */
static int g_inner = 1;
#define ORC_INNER_MIN ((int64_t)1 << 16) /* work below which a loop stays serial */
#define FAIL(...)                                    \
  do {                                               \
    snprintf(g_err, sizeof g_err, __VA_ARGS__);      \
    return -1;                                       \
  } while (0)

/* ------------------------------------------------------------------------ */
/* Intermediate labels: parse_eincode keeps a label of a subtree iff it is    */
/* connected to a tensor outside the subtree (OMEinsumContractionOrders,     */
/* used by uncompress, src/types.jl:75-79).  The MIS network's output iy is   */
/* empty (rawcode(IndependentSet(g))), so the root's labels are empty.        */
/* ------------------------------------------------------------------------ */
typedef struct {
  int n;
  int32_t* lab;
  int32_t* cnt; /* occurrences of lab[i] among the leaves of the subtree */
} lset;

static int max_label(int32_t nleaves, const int32_t* ix_ptr, const int32_t* ix) {
  int m = -1;
  for (int32_t i = 0; i < ix_ptr[nleaves]; i++)
    if (ix[i] > m) m = ix[i];
  return m;
}

static int check_tree(int32_t nleaves, const int32_t* tree) {
  /* every leaf and every internal node (but the root) used exactly once */
  if (nleaves < 2) return 0;
  int32_t nn = nleaves - 1, tot = nleaves + nn;
  char* used = (char*)calloc((size_t)tot, 1);
  for (int32_t j = 0; j < nn; j++) {
    for (int s = 0; s < 2; s++) {
      int32_t c = tree[2 * j + s];
      if (c < 0 || c >= tot || (c >= nleaves && c - nleaves >= j) || used[c]) {
        free(used);
        FAIL("bad tree: node %d child %d", j, c);
      }
      used[c] = 1;
    }
  }
  free(used);
  return 0;
}

/* builds label sets for leaves and internal nodes; returns array of size
 * nleaves + nleaves - 1 (caller frees with free_lsets). */
static lset* build_lsets(int32_t nleaves, const int32_t* ix_ptr,
                         const int32_t* ix, const int32_t* tree) {
  int L = max_label(nleaves, ix_ptr, ix) + 1;
  int32_t* total = (int32_t*)calloc((size_t)(L > 0 ? L : 1), sizeof(int32_t));
  for (int32_t i = 0; i < ix_ptr[nleaves]; i++) total[ix[i]]++;
  int32_t nn = nleaves > 0 ? nleaves - 1 : 0;
  lset* s = (lset*)calloc((size_t)(nleaves + nn), sizeof(lset));
  for (int32_t i = 0; i < nleaves; i++) {
    int n = ix_ptr[i + 1] - ix_ptr[i];
    s[i].n = n;
    s[i].lab = (int32_t*)malloc(sizeof(int32_t) * (n + 1));
    s[i].cnt = (int32_t*)malloc(sizeof(int32_t) * (n + 1));
    for (int k = 0; k < n; k++) {
      s[i].lab[k] = ix[ix_ptr[i] + k];
      s[i].cnt[k] = 1;
    }
  }
  int32_t* cnt = (int32_t*)calloc((size_t)(L > 0 ? L : 1), sizeof(int32_t));
  for (int32_t j = 0; j < nn; j++) {
    lset* a = &s[tree[2 * j]];
    lset* b = &s[tree[2 * j + 1]];
    lset* o = &s[nleaves + j];
    o->lab = (int32_t*)malloc(sizeof(int32_t) * (a->n + b->n + 1));
    o->cnt = (int32_t*)malloc(sizeof(int32_t) * (a->n + b->n + 1));
    for (int k = 0; k < a->n; k++) cnt[a->lab[k]] += a->cnt[k];
    for (int k = 0; k < b->n; k++) cnt[b->lab[k]] += b->cnt[k];
    o->n = 0;
    for (int side = 0; side < 2; side++) {
      lset* x = side ? b : a;
      for (int k = 0; k < x->n; k++) {
        int l = x->lab[k];
        if (cnt[l] > 0) {
          if (cnt[l] < total[l]) {
            o->lab[o->n] = l;
            o->cnt[o->n] = cnt[l];
            o->n++;
          }
          cnt[l] = 0; /* mark as emitted */
        }
      }
    }
  }
  free(cnt);
  free(total);
  return s;
}

static void free_lsets(lset* s, int32_t count) {
  for (int32_t i = 0; i < count; i++) {
    free(s[i].lab);
    free(s[i].cnt);
  }
  free(s);
}

int orc_node_labels(int32_t nleaves, const int32_t* ix_ptr, const int32_t* ix,
                    const int32_t* tree, int32_t* out_ptr, int32_t* out_lab,
                    int32_t out_cap) {
  if (check_tree(nleaves, tree)) return -1;
  int32_t nn = nleaves > 0 ? nleaves - 1 : 0;
  lset* s = build_lsets(nleaves, ix_ptr, ix, tree);
  int32_t pos = 0;
  out_ptr[0] = 0;
  for (int32_t j = 0; j < nn; j++) {
    lset* o = &s[nleaves + j];
    for (int k = 0; k < o->n; k++) {
      if (pos >= out_cap) {
        free_lsets(s, nleaves + nn);
        FAIL("label buffer too small");
      }
      out_lab[pos++] = o->lab[k];
    }
    out_ptr[j + 1] = pos;
  }
  free_lsets(s, nleaves + nn);
  return pos;
}

/* ------------------------------------------------------------------------ */
/* Complexity with uniform size 2 (OMEinsumContractionOrders               */
/* contraction_complexity as used by complexity(), src/types.jl:115-121):   */
/* tc = log2 sum_nodes 2^|ixa U ixb|, sc = max_nodes |iy|,                    */
/* rwc = log2 sum_nodes (2^|ixa| + 2^|ixb| + 2^|iy|).                         */
/* A tree that is a single leaf has sc = its rank and tc = rwc = -inf.       */
/* ------------------------------------------------------------------------ */
static int union_size(const lset* a, const lset* b) {
  int u = a->n;
  for (int k = 0; k < b->n; k++) {
    int found = 0;
    for (int q = 0; q < a->n; q++)
      if (a->lab[q] == b->lab[k]) {
        found = 1;
        break;
      }
    if (!found) u++;
  }
  return u;
}

int orc_complexity(int32_t nleaves, const int32_t* ix_ptr, const int32_t* ix,
                   const int32_t* tree, double* tc, double* sc, double* rwc) {
  if (check_tree(nleaves, tree)) return -1;
  if (nleaves == 0) {
    *tc = -INFINITY;
    *sc = 0;
    *rwc = -INFINITY;
    return 0;
  }
  if (nleaves == 1) {
    *tc = -INFINITY;
    *sc = ix_ptr[1] - ix_ptr[0];
    *rwc = -INFINITY;
    return 0;
  }
  int32_t nn = nleaves - 1;
  lset* s = build_lsets(nleaves, ix_ptr, ix, tree);
  double t = 0, rw = 0;
  int scm = 0;
  for (int32_t j = 0; j < nn; j++) {
    const lset* a = &s[tree[2 * j]];
    const lset* b = &s[tree[2 * j + 1]];
    const lset* o = &s[nleaves + j];
    t += ldexp(1.0, union_size(a, b));
    rw += ldexp(1.0, a->n) + ldexp(1.0, b->n) + ldexp(1.0, o->n);
    if (o->n > scm) scm = o->n;
  }
  *tc = log2(t);
  *sc = scm;
  *rwc = log2(rw);
  free_lsets(s, nleaves + nn);
  return 0;
}

/* contraction_peak_memory (src/utils.jl:197-219): all leaves are live at the
 * start; in post-order each node allocates its output and frees the inputs
 * of its internal children (the child call returns the sizes of its own
 * inputs, which the parent frees).  Result: log2 of the maximum. */
double orc_peak_memory(int32_t nleaves, const int32_t* ix_ptr,
                       const int32_t* ix, const int32_t* tree) {
  if (nleaves == 0) return -INFINITY;
  if (check_tree(nleaves, tree)) return NAN;
  int32_t nn = nleaves - 1;
  lset* s = build_lsets(nleaves, ix_ptr, ix, tree);
  double cur = 0;
  for (int32_t i = 0; i < nleaves; i++) cur += ldexp(1.0, s[i].n);
  double peak = cur;
  double* ffs = (double*)calloc((size_t)(nn > 0 ? nn : 1), sizeof(double));
  for (int32_t j = 0; j < nn; j++) {
    double freed = 0;
    for (int q = 0; q < 2; q++) {
      int32_t c = tree[2 * j + q];
      if (c >= nleaves) freed += ffs[c - nleaves];
      ffs[j] += ldexp(1.0, s[c].n);
    }
    cur += ldexp(1.0, s[nleaves + j].n) - freed;
    if (cur > peak) peak = cur;
  }
  free(ffs);
  free_lsets(s, nleaves + nn);
  return log2(peak);
}

/* contraction_all_memory (src/utils.jl:221-229): log2 of the total number of
 * elements allocated by internal nodes (leaves count 0). */
double orc_all_memory(int32_t nleaves, const int32_t* ix_ptr,
                      const int32_t* ix, const int32_t* tree) {
  if (nleaves < 2) return -INFINITY;
  if (check_tree(nleaves, tree)) return NAN;
  int32_t nn = nleaves - 1;
  lset* s = build_lsets(nleaves, ix_ptr, ix, tree);
  double tot = 0;
  for (int32_t j = 0; j < nn; j++) tot += ldexp(1.0, s[nleaves + j].n);
  free_lsets(s, nleaves + nn);
  return log2(tot);
}

/* ------------------------------------------------------------------------ */
/* Max-plus GEMM: the role of TropicalGEMM's mul! on Matrix{Tropical{T}}     */
/* (test/slice.jl:5): C = A (x) B with (+) = max, (x) = +.                   */
/* Column-major; loop order n, k, m so the inner loop is contiguous.        */
/* ------------------------------------------------------------------------ */
int orc_maxplus_gemm_f32(int64_t M, int64_t N, int64_t K, int64_t batch,
                         const float* A, int64_t lda, int64_t sA,
                         const float* B, int64_t ldb, int64_t sB, float* C,
                         int64_t ldc, int64_t sC) {
  if (M < 0 || N < 0 || K < 0 || batch < 0) FAIL("negative size");
  if (lda < M || ldb < K || ldc < M) FAIL("leading dimension too small");
  /* each (t, n) column of C is one iteration: disjoint outputs */
#ifdef _OPENMP
#pragma omp parallel for collapse(2) schedule(static) \
    if (g_inner > 1 && batch * N * M * (K + 1) >= ORC_INNER_MIN) num_threads(g_inner)
#endif
  for (int64_t t = 0; t < batch; t++) {
    for (int64_t n = 0; n < N; n++) {
      const float* a = A + t * sA;
      const float* b = B + t * sB;
      float* cc = C + t * sC + n * ldc;
      for (int64_t m = 0; m < M; m++) cc[m] = -INFINITY;
      for (int64_t k = 0; k < K; k++) {
        const float bk = b[k + n * ldb];
        const float* ak = a + k * lda;
        for (int64_t m = 0; m < M; m++) {
          float v = ak[m] + bk;
          cc[m] = v > cc[m] ? v : cc[m];
        }
      }
    }
  }
  return 0;
}

static inline int32_t canon_i32(int32_t v) {
  return v < ORC_I32_NEG_THRESH ? ORC_I32_NEG : v;
}

int orc_maxplus_gemm_i32(int64_t M, int64_t N, int64_t K, int64_t batch,
                         const int32_t* A, int64_t lda, int64_t sA,
                         const int32_t* B, int64_t ldb, int64_t sB, int32_t* C,
                         int64_t ldc, int64_t sC) {
  if (M < 0 || N < 0 || K < 0 || batch < 0) FAIL("negative size");
  if (lda < M || ldb < K || ldc < M) FAIL("leading dimension too small");
  /* each (t, n) column of C is one iteration: disjoint outputs */
#ifdef _OPENMP
#pragma omp parallel for collapse(2) schedule(static) \
    if (g_inner > 1 && batch * N * M * (K + 1) >= ORC_INNER_MIN) num_threads(g_inner)
#endif
  for (int64_t t = 0; t < batch; t++) {
    for (int64_t n = 0; n < N; n++) {
      const int32_t* a = A + t * sA;
      const int32_t* b = B + t * sB;
      int32_t* cc = C + t * sC + n * ldc;
      for (int64_t m = 0; m < M; m++) cc[m] = INT32_MIN;
      for (int64_t k = 0; k < K; k++) {
        const int32_t bk = b[k + n * ldb];
        const int32_t* ak = a + k * lda;
        for (int64_t m = 0; m < M; m++) {
          /* inputs are canonical (>= -2^30), so the sum cannot overflow */
          int32_t v = ak[m] + bk;
          cc[m] = v > cc[m] ? v : cc[m];
        }
      }
      for (int64_t m = 0; m < M; m++) cc[m] = canon_i32(cc[m]);
    }
  }
  return 0;
}

/* ------------------------------------------------------------------------ */
/* Permute with max-reduction: dst[labels dl] = max over the src labels that  */
/* dl drops.  dl must be a subset of sl.                                      */
/* ------------------------------------------------------------------------ */
static int pos_of(int n, const int32_t* l, int32_t x) {
  for (int i = 0; i < n; i++)
    if (l[i] == x) return i;
  return -1;
}

static int permute_reduce(int dtype, int ns, const int32_t* sl, const void* src,
                          int nd, const int32_t* dl, void* dst) {
  if (ns > 40 || nd > ns) FAIL("permute: rank");
  /* bit i of the src index goes to bit dpos[i] of dst (or nowhere) */
  int dpos[64];
  for (int i = 0; i < ns; i++) dpos[i] = pos_of(nd, dl, sl[i]);
  for (int i = 0; i < nd; i++)
    if (pos_of(ns, sl, dl[i]) < 0) FAIL("permute: dst label not in src");
  int lo = ns < 12 ? ns : 12;
  int64_t nlo = (int64_t)1 << lo, nhi = (int64_t)1 << (ns - lo);
  int64_t* lot = (int64_t*)malloc(sizeof(int64_t) * nlo);
  int64_t* hit = (int64_t*)malloc(sizeof(int64_t) * nhi);
  for (int64_t i = 0; i < nlo; i++) {
    int64_t j = 0;
    for (int b = 0; b < lo; b++)
      if ((i >> b & 1) && dpos[b] >= 0) j |= (int64_t)1 << dpos[b];
    lot[i] = j;
  }
  for (int64_t i = 0; i < nhi; i++) {
    int64_t j = 0;
    for (int b = 0; b < ns - lo; b++)
      if ((i >> b & 1) && dpos[lo + b] >= 0) j |= (int64_t)1 << dpos[lo + b];
    hit[i] = j;
  }
  int64_t ndst = (int64_t)1 << nd;
  /* distinct h reach disjoint dst entries when every hi bit is kept */
  int hi_kept = 1;
  for (int b = lo; b < ns; b++)
    if (dpos[b] < 0) hi_kept = 0;
  const int par = g_inner > 1 && hi_kept && ((int64_t)1 << ns) >= ORC_INNER_MIN;
  (void)par;
  if (dtype == ORC_F32) {
    const float* s = (const float*)src;
    float* d = (float*)dst;
#ifdef _OPENMP
#pragma omp parallel for schedule(static) if (par) num_threads(g_inner)
#endif
    for (int64_t j = 0; j < ndst; j++) d[j] = -INFINITY;
#ifdef _OPENMP
#pragma omp parallel for schedule(static) if (par) num_threads(g_inner)
#endif
    for (int64_t h = 0; h < nhi; h++)
      for (int64_t l = 0; l < nlo; l++) {
        int64_t j = hit[h] | lot[l];
        float v = s[(h << lo) | l];
        d[j] = v > d[j] ? v : d[j];
      }
  } else {
    const int32_t* s = (const int32_t*)src;
    int32_t* d = (int32_t*)dst;
#ifdef _OPENMP
#pragma omp parallel for schedule(static) if (par) num_threads(g_inner)
#endif
    for (int64_t j = 0; j < ndst; j++) d[j] = INT32_MIN;
#ifdef _OPENMP
#pragma omp parallel for schedule(static) if (par) num_threads(g_inner)
#endif
    for (int64_t h = 0; h < nhi; h++)
      for (int64_t l = 0; l < nlo; l++) {
        int64_t j = hit[h] | lot[l];
        int32_t v = s[(h << lo) | l];
        d[j] = v > d[j] ? v : d[j];
      }
  }
  free(lot);
  free(hit);
  return 0;
}

/* ------------------------------------------------------------------------ */
/* Binary tropical einsum lowered the way OMEinsum does it for one node of a */
/* DynamicNestedEinsum: labels private to one input and absent from the     */
/* output are max-reduced first; inputs are permuted to (a,k,t) and (k,b,t); */
/* one batched GEMM gives (a,b,t); the result is permuted to the output.    */
/* Group orders: a, b, t follow the output, k follows input A.               */
/* ------------------------------------------------------------------------ */
static void* xmalloc_t(int dtype, int rank) {
  return malloc((size_t)4 << rank);
}

int orc_einsum2(int dtype, int na, const int32_t* ixa, const void* A, int nb,
                const int32_t* ixb, const void* B, int nc, const int32_t* ixc,
                void* C) {
  if (dtype != ORC_F32 && dtype != ORC_I32) FAIL("dtype");
  for (int i = 0; i < nc; i++)
    if (pos_of(na, ixa, ixc[i]) < 0 && pos_of(nb, ixb, ixc[i]) < 0)
      FAIL("output label %d not in inputs", ixc[i]);
  /* 1. reduce private labels */
  int32_t a1[64], b1[64];
  int na1 = 0, nb1 = 0;
  for (int i = 0; i < na; i++)
    if (pos_of(nb, ixb, ixa[i]) >= 0 || pos_of(nc, ixc, ixa[i]) >= 0)
      a1[na1++] = ixa[i];
  for (int i = 0; i < nb; i++)
    if (pos_of(na, ixa, ixb[i]) >= 0 || pos_of(nc, ixc, ixb[i]) >= 0)
      b1[nb1++] = ixb[i];
  /* 2. groups */
  int32_t ga[64], gb[64], gk[64], gt[64];
  int nga = 0, ngb = 0, ngk = 0, ngt = 0;
  for (int i = 0; i < nc; i++) {
    int ia = pos_of(na1, a1, ixc[i]) >= 0, ib = pos_of(nb1, b1, ixc[i]) >= 0;
    if (ia && ib)
      gt[ngt++] = ixc[i];
    else if (ia)
      ga[nga++] = ixc[i];
    else
      gb[ngb++] = ixc[i];
  }
  for (int i = 0; i < na1; i++)
    if (pos_of(nb1, b1, a1[i]) >= 0 && pos_of(nc, ixc, a1[i]) < 0)
      gk[ngk++] = a1[i];
  int32_t pa[64], pb[64], pc[64];
  int npa = 0, npb = 0, npc = 0;
  for (int i = 0; i < nga; i++) pa[npa++] = ga[i];
  for (int i = 0; i < ngk; i++) pa[npa++] = gk[i];
  for (int i = 0; i < ngt; i++) pa[npa++] = gt[i];
  for (int i = 0; i < ngk; i++) pb[npb++] = gk[i];
  for (int i = 0; i < ngb; i++) pb[npb++] = gb[i];
  for (int i = 0; i < ngt; i++) pb[npb++] = gt[i];
  for (int i = 0; i < nga; i++) pc[npc++] = ga[i];
  for (int i = 0; i < ngb; i++) pc[npc++] = gb[i];
  for (int i = 0; i < ngt; i++) pc[npc++] = gt[i];
  if (npa != na1 || npb != nb1) FAIL("einsum2: inconsistent labels");
  /* 3. permute (with the private reduction fused, as a unary op) */
  void* Ap = xmalloc_t(dtype, npa);
  void* Bp = xmalloc_t(dtype, npb);
  void* Cp = xmalloc_t(dtype, npc);
  int rc = permute_reduce(dtype, na, ixa, A, npa, pa, Ap);
  if (!rc) rc = permute_reduce(dtype, nb, ixb, B, npb, pb, Bp);
  /* 4. batched GEMM */
  if (!rc) {
    int64_t M = (int64_t)1 << nga, N = (int64_t)1 << ngb,
            K = (int64_t)1 << ngk, T = (int64_t)1 << ngt;
    if (dtype == ORC_F32)
      rc = orc_maxplus_gemm_f32(M, N, K, T, (const float*)Ap, M, M * K,
                                (const float*)Bp, K, K * N, (float*)Cp, M,
                                M * N);
    else
      rc = orc_maxplus_gemm_i32(M, N, K, T, (const int32_t*)Ap, M, M * K,
                                (const int32_t*)Bp, K, K * N, (int32_t*)Cp, M,
                                M * N);
  }
  /* 5. permute to the output order */
  if (!rc) rc = permute_reduce(dtype, npc, pc, Cp, nc, ixc, C);
  free(Ap);
  free(Bp);
  free(Cp);
  return rc;
}

/* ------------------------------------------------------------------------ */
/* solve_slice (src/dynamic_ob.jl:30-34): leaf tensors of                    */
/* IndependentSet(g, w) under SizeMax, contracted along the tree.            */
/* Vertex leaf [v] = [0, w_v]; edge leaf [u,v] = [0 0; 0 -inf] (entry        */
/* (i_u, i_v) at i_u + 2 i_v).                                               */
/* ------------------------------------------------------------------------ */
static int leaf_tensor(int dtype, const orc_slice* s, int32_t leaf, void* out) {
  int n = s->ix_ptr[leaf + 1] - s->ix_ptr[leaf];
  const int32_t* l = s->ix + s->ix_ptr[leaf];
  if (n == 1) {
    int32_t v = l[0];
    if (v < 0 || v >= s->nv) FAIL("leaf label %d out of range", v);
    if (dtype == ORC_F32) {
      float* o = (float*)out;
      o[0] = 0.0f;
      o[1] = s->weights ? ((const float*)s->weights)[v] : 1.0f;
    } else {
      int32_t* o = (int32_t*)out;
      o[0] = 0;
      o[1] = s->weights ? ((const int32_t*)s->weights)[v] : 1;
    }
    return 0;
  }
  if (n == 2) {
    if (l[0] == l[1] || l[0] < 0 || l[1] < 0 || l[0] >= s->nv || l[1] >= s->nv)
      FAIL("bad edge leaf (%d,%d)", l[0], l[1]);
    if (dtype == ORC_F32) {
      float* o = (float*)out;
      o[0] = o[1] = o[2] = 0.0f;
      o[3] = -INFINITY;
    } else {
      int32_t* o = (int32_t*)out;
      o[0] = o[1] = o[2] = 0;
      o[3] = ORC_I32_NEG;
    }
    return 0;
  }
  FAIL("leaf %d has rank %d (MIS leaves have rank 1 or 2)", leaf, n);
}

int orc_solve_slice(int dtype, const orc_slice* s, void* result) {
  if (dtype != ORC_F32 && dtype != ORC_I32) FAIL("dtype");
  int32_t nl = s->nleaves;
  if (nl <= 0) {
    if (dtype == ORC_F32)
      *(float*)result = 0.0f;
    else
      *(int32_t*)result = 0;
    return 0;
  }
  if (check_tree(nl, s->tree)) return -1;
  int32_t nn = nl - 1;
  lset* ls = build_lsets(nl, s->ix_ptr, s->ix, s->tree);
  void** val = (void**)calloc((size_t)(nl + nn), sizeof(void*));
  int rc = 0;
  for (int32_t i = 0; i < nl && !rc; i++) {
    val[i] = malloc(16);
    rc = leaf_tensor(dtype, s, i, val[i]);
  }
  for (int32_t j = 0; j < nn && !rc; j++) {
    int32_t ca = s->tree[2 * j], cb = s->tree[2 * j + 1];
    lset* o = &ls[nl + j];
    val[nl + j] = xmalloc_t(dtype, o->n);
    rc = orc_einsum2(dtype, ls[ca].n, ls[ca].lab, val[ca], ls[cb].n,
                     ls[cb].lab, val[cb], o->n, o->lab, val[nl + j]);
    free(val[ca]);
    free(val[cb]);
    val[ca] = val[cb] = NULL;
  }
  if (!rc) {
    /* root (or single leaf): reduce whatever labels remain */
    int32_t root = nn > 0 ? nl + nn - 1 : 0;
    rc = permute_reduce(dtype, ls[root].n, ls[root].lab, val[root], 0, NULL,
                        result);
  }
  for (int32_t i = 0; i < nl + nn; i++) free(val[i]);
  free(val);
  free_lsets(ls, nl + nn);
  return rc;
}

/* contract_slices (src/dynamic_ob.jl:36-48): empty graph -> r, otherwise
 * solve_slice + r in the element type. */
static int contract_one(int dtype, const orc_slice* si, void* result) {
  if (si->nv == 0) {
    if (dtype == ORC_F32)
      *(float*)result = (float)si->r;
    else
      *(int32_t*)result = (int32_t)si->r;
    return 0;
  }
  union {
    float f;
    int32_t i;
  } t;
  if (orc_solve_slice(dtype, si, &t)) return -1;
  if (dtype == ORC_F32)
    *(float*)result = t.f + (float)si->r;
  else
    *(int32_t*)result = t.i <= ORC_I32_NEG_THRESH ? ORC_I32_NEG : t.i + (int32_t)si->r;
  return 0;
}

/* Slices of at least 2^ORC_BIG_TC configurations are contracted one at a
 * time with the threads inside each contraction, the rest one per thread. */
#define ORC_BIG_TC 33.0

int orc_contract_slices(int dtype, const orc_slice* s, int64_t n,
                        void* results, int nthreads) {
  int err = 0;
#ifdef _OPENMP
  if (nthreads <= 0) nthreads = omp_get_max_threads();
#endif
  char* big = (char*)calloc(n > 0 ? (size_t)n : 1, 1);
  if (nthreads > 1)
    for (int64_t i = 0; i < n; i++) {
      double tc, sc, rw;
      if (s[i].nv > 0 && s[i].nleaves > 1 &&
          orc_complexity(s[i].nleaves, s[i].ix_ptr, s[i].ix, s[i].tree, &tc, &sc, &rw) == 0 &&
          tc >= ORC_BIG_TC)
        big[i] = 1;
    }
  g_inner = nthreads;
  for (int64_t i = 0; i < n; i++)
    if (big[i] && contract_one(dtype, &s[i], (char*)results + 4 * i)) err = 1;
  g_inner = 1;
#ifdef _OPENMP
#pragma omp parallel for schedule(dynamic, 1) num_threads(nthreads)
#endif
  for (int64_t i = 0; i < n; i++) {
    if (big[i]) continue;
    if (s[i].nv == 0) {
      if (dtype == ORC_F32)
        ((float*)results)[i] = (float)s[i].r;
      else
        ((int32_t*)results)[i] = (int32_t)s[i].r;
      continue;
    }
    union {
      float f;
      int32_t i;
    } t;
    if (orc_solve_slice(dtype, &s[i], &t)) {
#ifdef _OPENMP
#pragma omp atomic write
#endif
      err = 1;
      continue;
    }
    if (dtype == ORC_F32)
      ((float*)results)[i] = t.f + (float)s[i].r;
    else
      ((int32_t*)results)[i] =
          t.i <= ORC_I32_NEG_THRESH ? ORC_I32_NEG : t.i + (int32_t)s[i].r;
  }
  free(big);
  if (err) FAIL("contract_slices: a slice failed");
  return 0;
}

/* ------------------------------------------------------------------------ */
/* Independent MIS solvers (used to pin the contraction).                    */
/* ------------------------------------------------------------------------ */

/* Exhaustive search over independent sets, n <= 64 (bitmask recursion:     */
/* every independent set is reached through the include/exclude choice).    */
static double bf_rec(uint64_t P, const uint64_t* adj, const double* w) {
  if (!P) return 0;
  int v = __builtin_ctzll(P);
  uint64_t rest = P & ~(1ULL << v);
  double ex = bf_rec(rest, adj, w);
  double in = w[v] + bf_rec(rest & ~adj[v], adj, w);
  return in > ex ? in : ex;
}

double orc_mis_bruteforce(int32_t nv, int32_t ne, const int32_t* edges,
                          const double* w) {
  if (nv > 64 || nv < 0) return NAN;
  uint64_t adj[64] = {0};
  double ww[64];
  for (int i = 0; i < nv; i++) ww[i] = w ? w[i] : 1.0;
  for (int e = 0; e < ne; e++) {
    int u = edges[2 * e], v = edges[2 * e + 1];
    if (u < 0 || v < 0 || u >= nv || v >= nv || u == v) return NAN;
    adj[u] |= 1ULL << v;
    adj[v] |= 1ULL << u;
  }
  uint64_t P = nv == 64 ? ~0ULL : ((1ULL << nv) - 1);
  return bf_rec(P, adj, ww);
}

/* Exact unweighted MIS by branch and reduce, in the role of
 * EliminateGraphs.mis2 (the test oracle of test/utils.jl:34): degree <= 1
 * and triangle degree-2 vertices are taken; a degree-2 vertex v with
 * non-adjacent neighbours u1,u2 branches into {v} or {u1,u2}; otherwise a
 * maximum-degree vertex is branched; components are solved separately and
 * graphs of maximum degree 2 use the path/cycle closed forms. */
#define XW 16 /* up to 1024 vertices */
typedef struct {
  int n, words;
  uint64_t (*adj)[XW];
} xgraph;

static inline int xpop(const uint64_t* s, int words) {
  int c = 0;
  for (int i = 0; i < words; i++) c += __builtin_popcountll(s[i]);
  return c;
}
static inline int xdeg(const xgraph* g, const uint64_t* P, int v) {
  int c = 0;
  for (int i = 0; i < g->words; i++) c += __builtin_popcountll(g->adj[v][i] & P[i]);
  return c;
}
static inline int xhas(const uint64_t* s, int v) { return (s[v >> 6] >> (v & 63)) & 1; }
static inline void xclr(uint64_t* s, int v) { s[v >> 6] &= ~(1ULL << (v & 63)); }
static inline void xset(uint64_t* s, int v) { s[v >> 6] |= 1ULL << (v & 63); }
static void xremove_closed(const xgraph* g, uint64_t* P, int v) {
  for (int i = 0; i < g->words; i++) P[i] &= ~g->adj[v][i];
  xclr(P, v);
}

static int64_t xsolve(const xgraph* g, const uint64_t* P0);

static int64_t xsolve_component(const xgraph* g, uint64_t* P) {
  int64_t acc = 0;
  int words = g->words;
  for (;;) {
    int changed = 0;
    for (int wi = 0; wi < words; wi++) {
      uint64_t m = P[wi];
      while (m) {
        int v = wi * 64 + __builtin_ctzll(m);
        m &= m - 1;
        if (!xhas(P, v)) continue;
        int d = xdeg(g, P, v);
        int take = d <= 1;
        if (d == 2) {
          int nb[2], k = 0;
          for (int i = 0; i < words && k < 2; i++) {
            uint64_t x = g->adj[v][i] & P[i];
            while (x && k < 2) {
              nb[k++] = i * 64 + __builtin_ctzll(x);
              x &= x - 1;
            }
          }
          if (xhas(g->adj[nb[0]], nb[1])) take = 1;
        }
        if (take) {
          acc++;
          xremove_closed(g, P, v);
          changed = 1;
        }
      }
    }
    if (!changed) break;
  }
  if (!xpop(P, words)) return acc;
  /* split into components */
  uint64_t comp[XW], frontier[XW], rest[XW];
  memcpy(rest, P, sizeof(uint64_t) * words);
  int first = -1;
  for (int i = 0; i < words; i++)
    if (rest[i]) {
      first = i * 64 + __builtin_ctzll(rest[i]);
      break;
    }
  memset(comp, 0, sizeof comp);
  memset(frontier, 0, sizeof frontier);
  xset(comp, first);
  xset(frontier, first);
  for (;;) {
    uint64_t nf[XW];
    memset(nf, 0, sizeof nf);
    int any = 0;
    for (int wi = 0; wi < words; wi++) {
      uint64_t m = frontier[wi];
      while (m) {
        int v = wi * 64 + __builtin_ctzll(m);
        m &= m - 1;
        for (int i = 0; i < words; i++) nf[i] |= g->adj[v][i];
      }
    }
    for (int i = 0; i < words; i++) {
      nf[i] &= P[i] & ~comp[i];
      comp[i] |= nf[i];
      frontier[i] = nf[i];
      if (nf[i]) any = 1;
    }
    if (!any) break;
  }
  int whole = 1;
  for (int i = 0; i < words; i++)
    if (comp[i] != P[i]) whole = 0;
  if (!whole) {
    for (int i = 0; i < words; i++) rest[i] = P[i] & ~comp[i];
    return acc + xsolve(g, comp) + xsolve(g, rest);
  }
  /* connected, min degree >= 2 (triangle-free deg-2 vertices remain) */
  int vmax = -1, dmax = -1, v2 = -1;
  for (int wi = 0; wi < words; wi++) {
    uint64_t m = P[wi];
    while (m) {
      int v = wi * 64 + __builtin_ctzll(m);
      m &= m - 1;
      int d = xdeg(g, P, v);
      if (d > dmax) {
        dmax = d;
        vmax = v;
      }
      if (d == 2 && v2 < 0) v2 = v;
    }
  }
  if (dmax <= 2) return acc + xpop(P, words) / 2; /* a cycle */
  uint64_t Q[XW];
  if (v2 >= 0) {
    int nb[2], k = 0;
    for (int i = 0; i < words && k < 2; i++) {
      uint64_t x = g->adj[v2][i] & P[i];
      while (x && k < 2) {
        nb[k++] = i * 64 + __builtin_ctzll(x);
        x &= x - 1;
      }
    }
    memcpy(Q, P, sizeof(uint64_t) * words);
    xremove_closed(g, Q, v2);
    int64_t a = 1 + xsolve(g, Q);
    memcpy(Q, P, sizeof(uint64_t) * words);
    xremove_closed(g, Q, nb[0]);
    xremove_closed(g, Q, nb[1]);
    int64_t b = 2 + xsolve(g, Q);
    return acc + (a > b ? a : b);
  }
  memcpy(Q, P, sizeof(uint64_t) * words);
  xremove_closed(g, Q, vmax);
  int64_t a = 1 + xsolve(g, Q);
  memcpy(Q, P, sizeof(uint64_t) * words);
  xclr(Q, vmax);
  int64_t b = xsolve(g, Q);
  return acc + (a > b ? a : b);
}

static int64_t xsolve(const xgraph* g, const uint64_t* P0) {
  uint64_t P[XW];
  memcpy(P, P0, sizeof(uint64_t) * g->words);
  return xsolve_component(g, P);
}

int64_t orc_mis_exact(int32_t nv, int32_t ne, const int32_t* edges) {
  if (nv < 0 || nv > 64 * XW) return -1;
  if (nv == 0) return 0;
  xgraph g;
  g.n = nv;
  g.words = (nv + 63) / 64;
  g.adj = calloc((size_t)nv, sizeof(*g.adj));
  for (int e = 0; e < ne; e++) {
    int u = edges[2 * e], v = edges[2 * e + 1];
    if (u < 0 || v < 0 || u >= nv || v >= nv || u == v) {
      free(g.adj);
      return -1;
    }
    xset(g.adj[u], v);
    xset(g.adj[v], u);
  }
  uint64_t P[XW];
  memset(P, 0, sizeof P);
  for (int v = 0; v < nv; v++) xset(P, v);
  int64_t r = xsolve(&g, P);
  free(g.adj);
  return r;
}

/* Maximum weight independent set on a forest: textbook tree DP
 * (in[v] = w_v + sum out[c], out[v] = sum max(in[c], out[c])).
 * Returns NAN if the graph has a cycle. */
double orc_mwis_forest(int32_t nv, int32_t ne, const int32_t* edges,
                       const double* w) {
  if (ne >= nv && nv > 0) return NAN;
  int32_t* deg = calloc((size_t)nv + 1, sizeof(int32_t));
  int32_t* off = calloc((size_t)nv + 1, sizeof(int32_t));
  int32_t* nbr = malloc(sizeof(int32_t) * (2 * (size_t)ne + 1));
  for (int e = 0; e < ne; e++) {
    deg[edges[2 * e]]++;
    deg[edges[2 * e + 1]]++;
  }
  for (int v = 0; v < nv; v++) off[v + 1] = off[v] + deg[v];
  int32_t* fill = calloc((size_t)nv + 1, sizeof(int32_t));
  for (int e = 0; e < ne; e++) {
    int u = edges[2 * e], v = edges[2 * e + 1];
    nbr[off[u] + fill[u]++] = v;
    nbr[off[v] + fill[v]++] = u;
  }
  double* in = calloc((size_t)nv + 1, sizeof(double));
  double* out = calloc((size_t)nv + 1, sizeof(double));
  int32_t* parent = malloc(sizeof(int32_t) * ((size_t)nv + 1));
  int32_t* order = malloc(sizeof(int32_t) * ((size_t)nv + 1));
  char* seen = calloc((size_t)nv + 1, 1);
  double total = 0;
  int cyc = 0;
  for (int root = 0; root < nv && !cyc; root++) {
    if (seen[root]) continue;
    int head = 0, tail = 0;
    order[tail++] = root;
    parent[root] = -1;
    seen[root] = 1;
    while (head < tail) {
      int v = order[head++];
      for (int k = off[v]; k < off[v + 1]; k++) {
        int u = nbr[k];
        if (u == parent[v]) continue;
        if (seen[u]) {
          cyc = 1;
          break;
        }
        seen[u] = 1;
        parent[u] = v;
        order[tail++] = u;
      }
    }
    for (int i = tail - 1; i >= 0; i--) {
      int v = order[i];
      in[v] += w ? w[v] : 1.0;
      if (parent[v] >= 0) {
        in[parent[v]] += out[v];
        out[parent[v]] += in[v] > out[v] ? in[v] : out[v];
      }
    }
    total += in[root] > out[root] ? in[root] : out[root];
  }
  free(deg);
  free(off);
  free(nbr);
  free(fill);
  free(in);
  free(out);
  free(parent);
  free(order);
  free(seen);
  return cyc ? NAN : total;
}
