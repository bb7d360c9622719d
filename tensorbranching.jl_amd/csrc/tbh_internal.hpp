// tbh_internal.hpp -- descriptors shared by the host planner and the gfx950
// kernels of libtbh.so.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace tbh {

constexpr int32_t kI32NegInf = -(1 << 30);
constexpr int32_t kI32NegThresh = -(1 << 29);
constexpr int kMaxRank = 30;     // largest tensor rank (2^30 entries)
constexpr int kMaxReduce = 12;   // reduced labels of one EW step
constexpr int kEwBlock = 256;    // threads per EW workgroup
constexpr int kTileBits = 12;    // a tile has at most 2^12 outputs
constexpr int kGemmScatter = 40; // GEMM output bit -> position table (rank + tile bits)
constexpr int kEpiElems = 8192;  // outputs per pass of a GEMM's LDS-staged epilogue
constexpr int kSubBits = 12;     // an LDS sub-tensor has at most 2^12 entries
// staged sub-tensors hold one pad entry per 2^TBH_LDS_PAD_SHIFT entries (0:
// none), which moves entries a multiple of that apart into different banks
#ifndef TBH_LDS_PAD_SHIFT
#define TBH_LDS_PAD_SHIFT 0
#endif
constexpr int kLdsElems =
    (1 << kSubBits) + (TBH_LDS_PAD_SHIFT ? ((1 << kSubBits) >> TBH_LDS_PAD_SHIFT) : 0);

// GEMM tile shape for an M x N output under a tile rule: 1 (default):
// tall-skinny outputs (N <= 64, M >= 256) take a 256-row tile so that a
// narrow N still fills it; 0: 128 x 128 tiles only; 2: 256 x 64 tiles up to
// N = 128 as well.
__host__ __device__ inline int gemm_shape(int64_t M, int64_t N, int rule = 1) {
  if (rule <= 0 || M < 256) return 0;
  if (N <= 32) return 2;
  if (N <= (rule >= 2 ? 128 : 64)) return 1;
  return 0;
}
__host__ __device__ inline int gemm_bm(int shape) { return shape ? 256 : 128; }
__host__ __device__ inline int gemm_bn(int shape) { return shape == 2 ? 32 : shape == 1 ? 64 : 128; }

// One EW step C = max over the reduced labels of A + B, tiled through LDS.
// The output labels split into tile labels (nT bits, ordered by C position)
// and outer labels (nO bits, one tile per configuration).  A tile first loads
// the sub-tensors of A and B it needs -- their tile and reduced labels, nSA /
// nSB bits in ascending operand position, so consecutive threads read
// consecutive addresses along the operand's lowest positions -- into LDS,
// then each of its outputs maxes A + B over the 2^nR reduced configurations
// and is written back contiguously along C's lowest positions.  Reduced
// labels beyond the sub-tensor budget are outer reduced labels (nRo): the
// tile repeats the load + max for each of their 2^nRo configurations,
// keeping its maxima in registers, before it writes.
// All positions are label positions (stride 2^p); -1 marks an absent label.
struct EwDesc {
  uint64_t a, b, c;        // device addresses; b == 0: unary step (B = tropical 1)
  int32_t nT, nO, nR;      // tile, outer and (sub-tensor) reduced bits
  int32_t nSA, nSB;        // sub-tensor bits of A and B
  int32_t tiles;           // 2^nO
  int32_t nRo;             // outer reduced bits (tiled steps only)
  int32_t mlay;            // MA in 1..3: tile bits 8..8+MA-1 are A-only labels and
                           // the rest of 8..11 B-only (a thread's 16 outputs are a
                           // 2^MA x 2^(4-MA) outer product); 0: no such layout
  int8_t qA[kMaxReduce];   // outer reduced bit -> A / B position
  int8_t qB[kMaxReduce];
  int8_t tC[kTileBits];    // tile bit -> C position
  int8_t tA[kTileBits];    // tile bit -> bit of A's sub-tensor
  int8_t tB[kTileBits];
  int8_t rA[kMaxReduce];   // reduced bit -> bit of A's / B's sub-tensor
  int8_t rB[kMaxReduce];
  int8_t sA[kSubBits];     // sub-tensor bit -> A position
  int8_t sB[kSubBits];
  int8_t oC[32], oA[32], oB[32];  // outer bit -> C / A / B position
};

__host__ __device__ inline uint32_t pos_stride(int8_t p) { return p >= 0 ? (1u << p) : 0u; }
__host__ __device__ inline uint32_t lds_pad(uint32_t i) {
  return TBH_LDS_PAD_SHIFT ? i + (i >> TBH_LDS_PAD_SHIFT) : i;
}

// One batched max-plus GEMM step (column-major, element strides).
struct GemmDesc {
  uint64_t a, b, c;         // device addresses
  int32_t M, N, K, batch;
  int64_t lda, ldb, ldc;
  int64_t sA, sB, sC;
  int32_t tiles_m, tiles_n; // tiles per batch entry
  int32_t shape;            // tile 0: 128 x 128, 1: 256 x 64, 2: 256 x 32
  // scatter: 0 C column-major with ldc; otherwise C has its own label order
  // (M, N, batch powers of two): bit q of (m, n, batch) -- m bits first,
  // then n, then batch -- sits at C position pc[q].  1: rows 0..4 sit at
  // positions 0..4, the tile stores its rows directly; 2: the tile is staged
  // in LDS and read back in ascending C position (ep), one pass of 2^ep_n
  // outputs (two passes split on the top column bit of the tile if ep_split)
  int32_t scatter;
  int8_t pc[kGemmScatter];  // -1 past the rank
  int8_t ep[16];            // tile bit b < 32: row bit b; 32 + b: column bit b
  int32_t ep_n, ep_split;
};

// Leaf tensor writes (vertex [0,w] or edge [0,0,0,-inf]).
struct LeafDesc {
  uint64_t p;      // device address
  int32_t kind;    // 1 = vertex, 2 = edge
  int32_t wbits;   // weight bit pattern of the dtype
};

// Per-slice epilogue: result = root + r (or r for an empty slice).
struct EpiDesc {
  uint64_t root;   // device address of the root scalar, 0: empty slice
  int32_t rbits;   // r as the dtype's bit pattern
  int32_t pad_;
};

}  // namespace tbh
