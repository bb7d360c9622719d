// tbh_plan.hpp -- host planner: lowers one slice's contraction tree to the
// steps the gfx950 kernels run, with tensor layouts and an arena layout.
#pragma once
#include <stdint.h>

#include <string>
#include <vector>

#include "tbh_internal.hpp"

namespace tbh {

// One slice as handed over through the C ABI (tbh_slice), weights already
// converted to the dtype's 32-bit pattern (nullptr = unit weights).
struct SliceIn {
  int32_t nv = 0;
  const uint32_t* wbits = nullptr;
  uint32_t unit_wbits = 0;   // bit pattern of weight 1 in the dtype
  int32_t nleaves = 0;
  const int32_t* ix_ptr = nullptr;
  const int32_t* ix = nullptr;
  const int32_t* tree = nullptr;
  double r = 0;
};

struct PTensor {
  int rank = 0;
  int64_t bytes = 0;
  int64_t off = 0;   // arena byte offset
  int birth = 0, death = 0;
};

enum OpKind { kEw = 0, kGemm = 1 };

struct POp {
  int kind = kEw;
  int a = -1, b = -1, c = -1;  // tensor ids (b = -1: unary)
  int level = 0;               // wave (batched) or sequence index (plain)
  // EW: the step's descriptor with device addresses left 0 (set at launch)
  int nP = 0, nR = 0;
  int64_t tiles = 0;
  bool direct = false;  // wave tiles of <= 2^10 outputs read straight from memory
  EwDesc ew{};
  // GEMM
  int32_t M = 0, N = 0, K = 0, batch = 0;
  int shape = 0;                 // tile shape (gemm_shape under the plan's tile rule)
  int scatter = 0;               // GemmDesc::scatter: C in its own label order, GEMM
  int8_t pc[kGemmScatter] = {};  // output bit q (m, then n, then batch bits) at pc[q]
  int8_t ep[16] = {};            // scatter 2: a pass's tile bits in ascending C position
  int ep_n = 0, ep_split = 0;
  double macs = 0;     // executed (output x reduced) configurations
  double bytes = 0;    // algorithmic bytes: |A| + |B| + |C| elements x 4
};

struct PLeaf {
  int tensor;
  int kind;        // 1 vertex, 2 edge
  uint32_t wbits;
};

struct SlicePlan {
  bool empty = true;     // no leaves: the contraction is the tropical one
  std::vector<PTensor> tensors;
  std::vector<POp> ops;  // in dependency order
  std::vector<PLeaf> leaves;
  int root = -1;         // tensor holding the scalar result
  int nlevels = 0;       // ops use levels 1..nlevels
  int64_t arena_bytes = 0;
  double node_macs = 0;  // sum over tree nodes of 2^|union| (= 2^tc)
  double exec_macs = 0, exec_bytes = 0, gemm_macs = 0, ew_bytes = 0;
  int n_ew = 0, n_gemm = 0;
};

struct PlanOptions {
  bool batched = true;   // wave levels + layout planning; false: OMEinsum lowering
  // GEMM-shaped nodes: M, N >= 64 (K >= gemm_min_k_mn), or min(M,128) x
  // min(N,128) >= gemm_min_fill with K >= gemm_min_k
  int gemm_min_k_mn = 16, gemm_min_fill = 4096, gemm_min_k = 64;
  // split-K of GEMM nodes (batched only): up to 2^split_tiles_bits tiles per
  // node, K parts of >= 2^split_min_k_bits, temporaries of rank <= split_max_rank
  bool split_k = true;
  int split_tiles_bits = 9, split_min_k_bits = 11, split_max_rank = 27;
  // spine-chain re-association (batched only): spine tensors of rank >=
  // chain_big absorbing side tensors of rank <= chain_side, combined side
  // tensor <= chain_side, at most chain_reduce labels reduced per step
  bool reassociate = true;
  int chain_big = 17, chain_side = 16, chain_reduce = 7;
  // chain_cost > 0: a rotation is also taken only when it lowers the two
  // steps' estimated cost, bytes moved + MACs * chain_cost / 64 (chain_cost
  // / 64 bytes per MAC), and the chain_reduce cap is lifted; 0: caps only.
  // On the 3-regular n=300 slices (sc 25) 12 measured 272 ms per solve
  // against 282 ms with the caps alone (8: 277, 16: 274, 24: 279)
  int chain_cost = 12;
  // a GEMM node writes its output straight in the layout its consumer asks
  // for (instead of a permute step after it) when the output's lowest
  // gemm_scatter_bits positions (at least 2) are row or column bits of one
  // tile, so each tile writes whole runs of 2^gemm_scatter_bits elements
  // (< 0: never).  2 (16-byte runs) measured best on the 3-regular n=300
  // slices: 0.328 s against 0.331 s for whole 128-byte lines (5).
  int gemm_scatter_bits = 2;
  // direct (wave-tile) EW steps reduce at most this many labels (-1: none
  // are direct); steps with reductions run faster LDS-tiled (DMA staging)
  int direct_max_reduce = 0;
  // GEMM tile rule (gemm_shape): 0 128 x 128 only, 1 tall-skinny tiles for
  // N <= 64, 2 also 256 x 64 up to N = 128
  int gemm_tiles = 1;
  // tiled EW steps: outer labels absent from the larger sub-tensor first,
  // so that consecutive tiles reuse it from LDS; false: C's label order
  bool reuse_order = true;
};

// Returns 0 or sets *err and returns -1.
int plan_slice(const SliceIn& s, const PlanOptions& opt, SlicePlan* out, std::string* err);

// Reference-model helpers (every label of size 2).
int slice_labels(const SliceIn& s, std::vector<std::vector<int32_t>>* labels, std::string* err);
int slice_complexity(const SliceIn& s, double* tc, double* sc, double* rwc, std::string* err);
int slice_peak_memory(const SliceIn& s, double* peak, double* all, std::string* err);

}  // namespace tbh
