// tbh_internal.hpp -- descriptors shared by the host planner and the gfx950
// kernels of libtbh.so.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace tbh {

constexpr int32_t kI32NegInf = -(1 << 30);
constexpr int32_t kI32NegThresh = -(1 << 29);
constexpr int kMaxRank = 30;     // largest tensor rank (2^30 entries)
constexpr int kMaxReduce = 15;   // reduced labels of one EW step
constexpr int kWavesPerBlock = 4;
constexpr int kEwBlock = 64 * kWavesPerBlock;

// One EW step: C[c] = max over r of A[dep_a(c) + rdep_a(r)] + B[...], where
// dep_a(c) = sum of 2^pa[i] over the set bits i of c (pa[i]: position in A
// of the label at C position i, -1 if A lacks it) and likewise for the
// reduced labels.  Lanes take output bits 0..5, a lane's 2^ept elements the next
// bits, and the wave tile index the rest.
struct EwDesc {
  uint64_t a, b, c;  // device addresses; b == 0: unary step (B = tropical 1)
  int32_t nP;        // output rank
  int32_t nR;        // reduced rank
  int32_t ept;       // log2 elements per lane (0..4)
  int32_t tiles;     // wave tiles of this step
  int8_t pa[32], pb[32];    // output bit i -> position in A / B (-1: absent)
  int8_t ra[16], rb[16];    // reduced bit i -> position in A / B (-1: absent)
};

__host__ __device__ inline uint32_t pos_stride(int8_t p) { return p >= 0 ? (1u << p) : 0u; }

// One batched max-plus GEMM step (column-major, element strides).
struct GemmDesc {
  uint64_t a, b, c;         // device addresses
  int32_t M, N, K, batch;
  int64_t lda, ldb, ldc;
  int64_t sA, sB, sC;
  int32_t tiles_m, tiles_n; // tiles per batch entry
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
