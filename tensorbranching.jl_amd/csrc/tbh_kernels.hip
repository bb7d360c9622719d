// tbh_kernels.hip -- gfx950 kernels of the tropical (max-plus) contraction
// backend.  All arithmetic is VALU: max-plus is not a multiply-accumulate, so
// MFMA has no role.  Tropical Float32 uses -inf as zero; Tropical Int32 uses
// the canonical sentinel -2^30 and rewrites stored values below -2^29 to it.
//
//   K1 tbh_gemm_kernel    batched / grouped max-plus GEMM: 128x128, 256x64 or
//                         256x32 tiles, 16-deep LDS slabs, 8x8 / 8x4 register
//                         blocking, add + max3 pairs; the output column-major
//                         or in the consumer's label order (row quads, or
//                         staged through LDS).
//   K2 tbh_ew_kernel      one wave-batched EW step: a binary tropical
//                         contraction over bit-permuted 2^r tensors (labels of
//                         size 2), reducing the contracted labels; also the
//                         permute / reduce of a single tensor.  Sub-tensors
//                         are staged HBM -> LDS by DMA.
//      tbh_ew_direct_kernel  the same for small steps without LDS (wave tiles
//                         reading straight from memory).
//   K3 tbh_leaf_kernel    leaf tensors of the MIS network from weights.
//   K5 tbh_epi_kernel     per-slice root + r.
#include <hip/hip_runtime.h>
#include <stddef.h>
#include <stdint.h>

#include "tbh_internal.hpp"

// minimum waves per SIMD the EW kernels are register-limited to
#ifndef TBH_EW_WAVES
#define TBH_EW_WAVES 4
#endif
#ifndef TBH_DIRECT_WAVES
#define TBH_DIRECT_WAVES 4
#endif
#ifndef TBH_GEMM_WAVES
#define TBH_GEMM_WAVES 3
#endif

namespace tbh {

// Address space of the step descriptors and tile prefixes (uploaded before a
// launch, never written by a kernel).  Generic by default: their fields are
// then read with vector loads.  Reading them through the constant address
// space (-DTBH_KCONST='__attribute__((address_space(4)))') turns the reads
// into scalar loads, which measured 3x slower for the direct kernel, whose
// waves re-read a descriptor (and search the prefixes) every few tiles.
#ifndef TBH_KCONST
#define TBH_KCONST
#endif
using KEw = const TBH_KCONST EwDesc*;
// An int8 position array of a descriptor, read through its containing
// dwords (one load per four positions when the compiler merges them).
struct KI8 {
  KEw d;
  uint32_t off;
  __device__ __forceinline__ int operator[](uint32_t i) const {
    const uint32_t o = off + i;
    const uint32_t w = reinterpret_cast<const TBH_KCONST uint32_t*>(d)[o >> 2];
    return (int)(int8_t)(w >> ((o & 3) * 8));
  }
};
// The same through the constant address space: scalar loads, counted by
// lgkmcnt only (the prefetching EW kernel keeps vmcnt for its staging loads).
using KEwS = const __attribute__((address_space(4))) EwDesc*;
struct KI8S {
  KEwS d;
  uint32_t off;
  __device__ __forceinline__ int operator[](uint32_t i) const {
    const uint32_t o = off + i;
    const uint32_t w = reinterpret_cast<const __attribute__((address_space(4))) uint32_t*>(d)[o >> 2];
    return (int)(int8_t)(w >> ((o & 3) * 8));
  }
};
template <class D>
struct KAcc;
template <>
struct KAcc<KEw> {
  using type = KI8;
};
template <>
struct KAcc<KEwS> {
  using type = KI8S;
};
#define TBH_KF(d, f) (typename KAcc<decltype(d)>::type{(d), (uint32_t)offsetof(EwDesc, f)})
// Tensors are addressed in the global address space: their loads are
// global_load (counted by vmcnt only), not flat loads, which also count in
// lgkmcnt and so would make every wait for an LDS or scalar access wait for
// the data loads in flight too.
#define TBH_GLOBAL __attribute__((address_space(1)))
#define TBH_LDS __attribute__((address_space(3)))
// Sub-tensors are staged with global_load_lds_dword (HBM -> LDS without a
// register round trip, so every load of a sub-tensor is in flight at once);
// 0 stages through registers in groups of 8.
#ifndef TBH_LDS_DMA
#define TBH_LDS_DMA 1
#endif
// 16-byte DMA (global_load_lds_dwordx4) where a sub-tensor's lowest bits
// are contiguous in the operand
#ifndef TBH_LDS_X4
#define TBH_LDS_X4 1
#endif
// tiled EW outputs written with nontemporal stores (measured: tiled EW 151.6
// against 153 ms per n=300 solve; the same for the GEMM's output quads made
// the GEMM 113.5 against 106.7 ms, its outputs being read soon after)
#ifndef TBH_EW_NT_STORE
#define TBH_EW_NT_STORE 1
#endif
#if TBH_LDS_DMA && TBH_LDS_PAD_SHIFT && TBH_LDS_PAD_SHIFT < 6
#error "LDS DMA writes 64 consecutive entries per wave: pad at most once per 64"
#endif

template <typename T>
struct Trop;
template <>
struct Trop<float> {
  static __device__ __forceinline__ float zero() { return -__builtin_inff(); }
  static __device__ __forceinline__ float one() { return 0.0f; }
  static __device__ __forceinline__ float mul(float a, float b) { return a + b; }
  static __device__ __forceinline__ float add(float a, float b) { return __builtin_fmaxf(a, b); }
  static __device__ __forceinline__ float canon(float x) { return x; }
};
template <>
struct Trop<int32_t> {
  // accumulator start; sums of canonical values are >= -2^31
  static __device__ __forceinline__ int32_t zero() { return INT32_MIN; }
  static __device__ __forceinline__ int32_t one() { return 0; }
  static __device__ __forceinline__ int32_t mul(int32_t a, int32_t b) {
    return (int32_t)((uint32_t)a + (uint32_t)b);
  }
  static __device__ __forceinline__ int32_t add(int32_t a, int32_t b) { return a > b ? a : b; }
  static __device__ __forceinline__ int32_t canon(int32_t x) {
    return x < kI32NegThresh ? kI32NegInf : x;
  }
};
// max(a, b, c) as one v_max3 (the compiler forms it for only part of the
// GEMM's reductions).  Register operands only; inputs are never NaN.
__device__ __forceinline__ float max3(float a, float b, float c) {
  float r;
  asm("v_max3_f32 %0, %1, %2, %3" : "=v"(r) : "v"(a), "v"(b), "v"(c));
  return r;
}
__device__ __forceinline__ int32_t max3(int32_t a, int32_t b, int32_t c) {
  int32_t r;
  asm("v_max3_i32 %0, %1, %2, %3" : "=v"(r) : "v"(a), "v"(b), "v"(c));
  return r;
}

// zero used to pad GEMM operands (adds without overflow for Int32)
template <typename T>
__device__ __forceinline__ T pad_zero();
template <>
__device__ __forceinline__ float pad_zero<float>() { return -__builtin_inff(); }
template <>
__device__ __forceinline__ int32_t pad_zero<int32_t>() { return kI32NegInf; }

// ---------------------------------------------------------------- K2 -----
// The 4 high bits (8..11) of a per-thread index: strides of those bits, and
// the sum over the bits set in a compile-time m.
template <class K>
__device__ __forceinline__ void hi_strides(uint32_t (&h)[4], K pos) {
#pragma unroll
  for (int b = 0; b < 4; b++) h[b] = pos_stride(pos[8 + b]);
}
template <int M>
__device__ __forceinline__ uint32_t hi_at(const uint32_t (&h)[4]) {
  return ((M & 1) ? h[0] : 0u) + ((M & 2) ? h[1] : 0u) + ((M & 4) ? h[2] : 0u) +
         ((M & 8) ? h[3] : 0u);
}
__device__ __forceinline__ uint32_t hi_at(const uint32_t (&h)[4], int m) {
  return ((m & 1) ? h[0] : 0u) + ((m & 2) ? h[1] : 0u) + ((m & 4) ? h[2] : 0u) +
         ((m & 8) ? h[3] : 0u);
}

// the low 8 bits of the thread index mapped through pos[0..7]
template <class K>
__device__ __forceinline__ uint32_t lo_dep(uint32_t tid, K pos) {
  uint32_t v = 0;
#pragma unroll
  for (int b = 0; b < 8; b++) {
    const uint32_t st = pos_stride(pos[b]);  // the same in every lane: read it in all
    v += ((tid >> b) & 1) ? st : 0u;
  }
  return v;
}

struct EwThread {  // per-step constants of one thread
  uint32_t loSA, loSB, loTA, loTB, loTC;
  uint32_t hSA[4], hSB[4], hTA[4], hTB[4], hTC[4];
  // 16-byte staging (sub-tensor bits 0, 1 are the operand's positions 0, 1,
  // at least 2^10 entries): the thread's offset of entry 4 tid
  uint32_t lo4SA, lo4SB;
  bool x4A, x4B;
};

template <class D>
__device__ __forceinline__ void ew_thread_setup(D d, EwThread& e) {
  const uint32_t tid = threadIdx.x;
  e.loSA = lo_dep(tid, TBH_KF(d, sA));
  e.loSB = lo_dep(tid, TBH_KF(d, sB));
  e.loTA = lo_dep(tid, TBH_KF(d, tA));
  e.loTB = lo_dep(tid, TBH_KF(d, tB));
  e.loTC = lo_dep(tid, TBH_KF(d, tC));
  hi_strides(e.hSA, TBH_KF(d, sA));
  hi_strides(e.hSB, TBH_KF(d, sB));
  hi_strides(e.hTA, TBH_KF(d, tA));
  hi_strides(e.hTB, TBH_KF(d, tB));
  hi_strides(e.hTC, TBH_KF(d, tC));
  using K = typename KAcc<D>::type;
  const K sA = TBH_KF(d, sA), sB = TBH_KF(d, sB);
  constexpr bool x4 = TBH_LDS_X4 && TBH_LDS_DMA && TBH_LDS_PAD_SHIFT == 0;
  e.x4A = x4 && d->nSA >= 10 && sA[0] == 0 && sA[1] == 1;
  e.x4B = x4 && d->nSB >= 10 && sB[0] == 0 && sB[1] == 1;
  e.lo4SA = lo_dep(tid, K{d, sA.off + 2});
  e.lo4SB = lo_dep(tid, K{d, sB.off + 2});
}

// The step index and its tile range are the same in every lane of a wave;
// saying so keeps them in scalar registers and the step loop's control flow
// uniform.
__device__ __forceinline__ int64_t uniform64(int64_t x) {
  const uint32_t lo = __builtin_amdgcn_readfirstlane((uint32_t)(uint64_t)x);
  const uint32_t hi = __builtin_amdgcn_readfirstlane((uint32_t)((uint64_t)x >> 32));
  return (int64_t)(((uint64_t)hi << 32) | lo);
}

// The step holding tile t, searched forward from step op (prefix[op] <= t;
// a wave's or workgroup's chunks only move forward): galloping, then
// bisection, so a search that stays in or near the current step reads one or
// two prefix entries instead of log2(nops).  prefix[nops] = total > t.
__device__ __forceinline__ int find_step(const int64_t* __restrict__ prefix, int nops, int op,
                                         int64_t t) {
  int lo = op, hi = op + 1, step = 1;
  while (hi < nops && prefix[hi] <= t) {
    lo = hi;
    step <<= 1;
    hi = lo + step;
  }
  if (hi > nops) hi = nops;
  while (hi - lo > 1) {
    const int mid = (lo + hi) >> 1;
    if (prefix[mid] <= t)
      lo = mid;
    else
      hi = mid;
  }
  return __builtin_amdgcn_readfirstlane(lo);
}

// outer offsets of tile o (sum of the strides of its set bits)
struct EwOuter {
  uint32_t a, b, c;
};
template <class D>
__device__ __forceinline__ EwOuter outer_offsets(D d, uint32_t o) {
  EwOuter r{0, 0, 0};
  const int nO = d->nO;
#pragma unroll
  for (int k = 0; k < 32; k++)
    if (k < nO && ((o >> k) & 1)) {
      r.a += pos_stride(TBH_KF(d, oA)[k]);
      r.b += pos_stride(TBH_KF(d, oB)[k]);
      r.c += pos_stride(TBH_KF(d, oC)[k]);
    }
  return r;
}
// from tile o-1 to tile o: only the bits that flip change
template <class D>
__device__ __forceinline__ void outer_step(D d, uint32_t o, EwOuter& r) {
  uint32_t flip = o ^ (o - 1);
  while (flip) {
    const int k = __builtin_ctz(flip);
    flip &= flip - 1;
    const uint32_t sa = pos_stride(TBH_KF(d, oA)[k]), sb = pos_stride(TBH_KF(d, oB)[k]), sc = pos_stride(TBH_KF(d, oC)[k]);
    if ((o >> k) & 1) {
      r.a += sa;
      r.b += sb;
      r.c += sc;
    } else {
      r.a -= sa;
      r.b -= sb;
      r.c -= sc;
    }
  }
}

// the sub-tensor of one operand (2^nS elements) into LDS; lds_pad(tid + 256m)
// = lds_pad(tid) + lds_pad(256) m
template <typename T>
__device__ __forceinline__ void ew_stage(const TBH_GLOBAL T* __restrict__ X, int nS, uint32_t lo,
                                         const uint32_t (&hS)[4], uint32_t lo4, bool x4,
                                         T* __restrict__ lX) {
  const int tid = threadIdx.x;
  const int n = 1 << nS;
  // recomputed per call: addresses are not held across the caller's loop
  asm volatile("" : "+v"(lo), "+v"(lo4));
#if TBH_LDS_DMA
  static_assert(sizeof(T) == 4, "dword LDS DMA");
  if (x4) {
    // entries 4 l .. 4 l + 3 of lane l are 16 contiguous bytes of the
    // operand (sub-tensor bits 0, 1 are its positions 0, 1; every other bit
    // sits at a position >= 2, so the address is 16-byte aligned): one
    // dwordx4 DMA per 256 entries, written at M0 + 16 x lane
    TBH_LDS T* w = (TBH_LDS T*)lX + 4 * (tid & ~63);
    const int cnt = n >> 10;  // 1..4, wave-uniform
#pragma unroll
    for (int m = 0; m < 4; m++) {
      if (m >= cnt) break;
      const uint32_t hi = ((m & 1) ? hS[2] : 0u) + ((m & 2) ? hS[3] : 0u);
#if defined(__HIP_DEVICE_COMPILE__)  // the host pass does not know gfx950's 16-byte form
      __builtin_amdgcn_global_load_lds(X + lo4 + hi, w + 1024 * m, 16, 0, 0);
#endif
    }
    return;
  }
  // lane l of the wave writes entry (its wave's first tid) + l + 256 m: the
  // LDS address is wave-uniform (M0) plus 4 x lane, and a run of 64 entries
  // starting at a multiple of 64 stays contiguous under the padding
  TBH_LDS T* w = (TBH_LDS T*)lX + lds_pad(tid & ~63);
  if (n <= 256) {
    if (tid < n) __builtin_amdgcn_global_load_lds(X + lo, w, 4, 0, 0);
    return;
  }
  const int cnt = n >> 8;  // 2..16, wave-uniform
#pragma unroll
  for (int m = 0; m < 16; m++) {
    if (m >= cnt) break;
    __builtin_amdgcn_global_load_lds(X + lo + hi_at(hS, m), w + lds_pad(256) * m, 4, 0, 0);
  }
#else
  const uint32_t pt = lds_pad(tid);
  if (n <= 256) {
    if (tid < n) lX[pt] = X[lo];
    return;
  }
  const int cnt = n >> 8;  // 2..16, wave-uniform
  // in groups of 8 loads in flight (few registers: the tile's maxima stay
  // live across the staging)
#pragma unroll
  for (int g = 0; g < 16; g += 8) {
    if (g >= cnt) break;
    T v[8];
#pragma unroll
    for (int m = 0; m < 8; m++)
      if (g + m < cnt) v[m] = X[lo + hi_at(hS, g + m)];
#pragma unroll
    for (int m = 0; m < 8; m++)
      if (g + m < cnt) lX[pt + lds_pad(256) * (g + m)] = v[m];
  }
#endif
}

// Max A + B over the sub-tensor reduced configurations of the staged unit
// into acc (one thread's CNT outputs).  The lowest (up to 3) sub-tensor
// reduced bits are unrolled in pairs (add + max3), the rest walked in
// Gray-code order.  D is the descriptor pointer (generic or constant address
// space: the prefetching kernel reads descriptors with scalar loads only).
template <typename T, bool HASB, int CNT, int MA, class D>
__device__ __forceinline__ void ew_compute(D d, const EwThread& e, const T* __restrict__ lA,
                                           const T* __restrict__ lB, T (&acc)[CNT]) {
  const int nR = d->nR;
  const int nlo = nR < 3 ? nR : 3;
  uint32_t lsA[4] = {0, 0, 0, 0}, lsB[4] = {0, 0, 0, 0};
#pragma unroll
  for (int b = 0; b < 3; b++)
    if (b < nlo) {
      lsA[b] = pos_stride(TBH_KF(d, rA)[b]);
      lsB[b] = pos_stride(TBH_KF(d, rB)[b]);
    }
  const int nq = 1 << nlo;
  const uint32_t nh = 1u << (nR - nlo);
  if (MA > 0) {
    constexpr int NA = 1 << MA, NB = 16 >> MA;
    uint32_t ltA = e.loTA, ltB = e.loTB;
    asm volatile("" : "+v"(ltA), "+v"(ltB));
    uint32_t pA[NA], pB[NB];
#pragma unroll
    for (int j = 0; j < NA; j++) pA[j] = lds_pad(ltA + hi_at(e.hTA, j));
#pragma unroll
    for (int j = 0; j < NB; j++) pB[j] = lds_pad(ltB + hi_at(e.hTB, j << MA));
    uint32_t rA = 0, rB = 0;
    for (uint32_t h = 0; h < nh; h++) {
      if (h) {
        const int i = __builtin_ctz(h);
        const uint32_t g = h ^ (h >> 1);
        const uint32_t sa = pos_stride(TBH_KF(d, rA)[nlo + i]), sb = pos_stride(TBH_KF(d, rB)[nlo + i]);
        if ((g >> i) & 1) {
          rA += sa;
          rB += sb;
        } else {
          rA -= sa;
          rB -= sb;
        }
      }
#pragma unroll
      for (int q = 0; q < 8; q += 2) {
        if (q >= nq) break;
        const uint32_t qa0 = lds_pad(rA + hi_at(lsA, q)), qb0 = lds_pad(rB + hi_at(lsB, q));
        T a0[NA], b0[NB];
#pragma unroll
        for (int j = 0; j < NA; j++) a0[j] = lA[pA[j] + qa0];
#pragma unroll
        for (int j = 0; j < NB; j++) b0[j] = lB[pB[j] + qb0];
        if (nq > 1) {
          const uint32_t qa1 = lds_pad(rA + hi_at(lsA, q + 1));
          const uint32_t qb1 = lds_pad(rB + hi_at(lsB, q + 1));
          T a1[NA], b1[NB];
#pragma unroll
          for (int j = 0; j < NA; j++) a1[j] = lA[pA[j] + qa1];
#pragma unroll
          for (int j = 0; j < NB; j++) b1[j] = lB[pB[j] + qb1];
#pragma unroll
          for (int mb = 0; mb < NB; mb++)
#pragma unroll
            for (int ma = 0; ma < NA; ma++)
              acc[ma + NA * mb] = max3(acc[ma + NA * mb], Trop<T>::mul(a0[ma], b0[mb]),
                                       Trop<T>::mul(a1[ma], b1[mb]));
        } else {
#pragma unroll
          for (int mb = 0; mb < NB; mb++)
#pragma unroll
            for (int ma = 0; ma < NA; ma++)
              acc[ma + NA * mb] = Trop<T>::add(acc[ma + NA * mb], Trop<T>::mul(a0[ma], b0[mb]));
        }
      }
    }
    return;
  }
  {
    // per-thread LDS addresses, recomputed per chunk: the empty asm hides
    // that they are loop-invariant, so they are not held in registers
    // across the staging
    uint32_t ltA = e.loTA, ltB = e.loTB;
    asm volatile("" : "+v"(ltA), "+v"(ltB));
    uint32_t pA[CNT], pB[CNT];
#pragma unroll
    for (int m = 0; m < CNT; m++) {
      pA[m] = lds_pad(ltA + hi_at(e.hTA, m));
      if (HASB) pB[m] = lds_pad(ltB + hi_at(e.hTB, m));
    }
    uint32_t rA = 0, rB = 0;
    for (uint32_t h = 0; h < nh; h++) {
      if (h) {
        const int i = __builtin_ctz(h);
        const uint32_t g = h ^ (h >> 1);
        const uint32_t sa = pos_stride(TBH_KF(d, rA)[nlo + i]), sb = pos_stride(TBH_KF(d, rB)[nlo + i]);
        if ((g >> i) & 1) {
          rA += sa;
          rB += sb;
        } else {
          rA -= sa;
          rB -= sb;
        }
      }
#pragma unroll
      for (int q = 0; q < 8; q += 2) {
        if (q >= nq) break;
        const uint32_t qa0 = lds_pad(rA + hi_at(lsA, q)), qb0 = lds_pad(rB + hi_at(lsB, q));
        if (nq > 1) {
          const uint32_t qa1 = lds_pad(rA + hi_at(lsA, q + 1));
          const uint32_t qb1 = lds_pad(rB + hi_at(lsB, q + 1));
#pragma unroll
          for (int m = 0; m < CNT; m++) {
            T x0 = lA[pA[m] + qa0], x1 = lA[pA[m] + qa1];
            if (HASB) {
              x0 = Trop<T>::mul(x0, lB[pB[m] + qb0]);
              x1 = Trop<T>::mul(x1, lB[pB[m] + qb1]);
            }
            acc[m] = max3(acc[m], x0, x1);
          }
        } else {
#pragma unroll
          for (int m = 0; m < CNT; m++) {
            T x0 = lA[pA[m] + qa0];
            if (HASB) x0 = Trop<T>::mul(x0, lB[pB[m] + qb0]);
            acc[m] = Trop<T>::add(acc[m], x0);
          }
        }
      }
    }
  }
}

// One tile of step d at outer offsets off, CNT = max(1, 2^nT / 256) outputs
// per thread.  For each configuration of the outer reduced labels (Gray-code
// order, one stride change each) the tile stages its sub-tensors of A and B
// in LDS and maxes A + B over the sub-tensor reduced labels into registers;
// the outputs are written once at the end.
// A sub-tensor index is a sum of distinct powers of two (tile bits and
// reduced bits sit at different positions), and lds_pad(x + y) =
// lds_pad(x) + lds_pad(y) when x and y share no bit, so the padded address of
// (output m, reduced configuration) is a per-thread part plus a wave-uniform
// part: one add per LDS read.
// MA > 0 (CNT = 16, binary, descriptor mlay = MA): tile bits 8 .. 8+MA-1
// are A-only and the rest of 8..11 B-only labels, so output m = ma + 2^MA mb
// needs A value ma and B value mb: 2^MA + 2^(4-MA) LDS reads feed 16 maxima
// per reduced configuration.
// stA / stB: the addresses the sub-tensors now in LDS were staged from (this
// step; 0: nothing valid).  A sub-tensor whose address repeats -- the next
// tile differs only in labels the operand lacks -- is not staged again.
template <typename T, bool HASB, int CNT, int MA = 0>
__device__ __forceinline__ void ew_tile(KEw d, const EwOuter& off,
                                        const EwThread& e, T* __restrict__ lA,
                                        T* __restrict__ lB, uint64_t& stA,
                                        uint64_t& stB) {
  const int tid = threadIdx.x;
  const TBH_GLOBAL T* __restrict__ A = reinterpret_cast<const TBH_GLOBAL T*>(d->a) + off.a;
  const TBH_GLOBAL T* __restrict__ B = reinterpret_cast<const TBH_GLOBAL T*>(d->b) + off.b;
  TBH_GLOBAL T* __restrict__ C = reinterpret_cast<TBH_GLOBAL T*>(d->c) + off.c;
  const bool active = CNT > 1 || tid < (1 << d->nT);
  T acc[CNT];
#pragma unroll
  for (int m = 0; m < CNT; m++) acc[m] = Trop<T>::zero();
  const uint32_t nc = 1u << d->nRo;
  uint32_t cA = 0, cB = 0;
  for (uint32_t c = 0; c < nc; c++) {
    if (c) {
      const int i = __builtin_ctz(c);
      const uint32_t g = c ^ (c >> 1);
      const uint32_t sa = pos_stride(TBH_KF(d, qA)[i]), sb = pos_stride(TBH_KF(d, qB)[i]);
      if ((g >> i) & 1) {
        cA += sa;
        cB += sb;
      } else {
        cA -= sa;
        cB -= sb;
      }
    }
    // the staging and the stores are issued at a raised wave priority, so
    // that a workgroup's memory requests leave ahead of other waves' compute
    // (tiled EW 151.0 against 151.9 ms per n=300 solve)
    __builtin_amdgcn_s_setprio(3);
    const uint64_t pa = (uint64_t)(A + cA);
    if (__builtin_amdgcn_readfirstlane((uint32_t)(pa != stA)))
      ew_stage<T>(A + cA, d->nSA, e.loSA, e.hSA, e.lo4SA, e.x4A, lA);
    stA = pa;
    if (HASB) {
      const uint64_t pb = (uint64_t)(B + cB);
      if (__builtin_amdgcn_readfirstlane((uint32_t)(pb != stB)))
        ew_stage<T>(B + cB, d->nSB, e.loSB, e.hSB, e.lo4SB, e.x4B, lB);
      stB = pb;
    }
    __builtin_amdgcn_s_setprio(0);
#if TBH_LDS_DMA
    // the DMA writes land in LDS when the loads retire (vmcnt), which the
    // barrier alone does not wait for
    __builtin_amdgcn_s_waitcnt(0);
#endif
    __syncthreads();
    if (active) ew_compute<T, HASB, CNT, MA>(d, e, lA, lB, acc);
    __syncthreads();
  }
  // contiguous writes along C's lowest positions
  __builtin_amdgcn_s_setprio(3);
  if (active) {
    uint32_t lc = e.loTC;
    asm volatile("" : "+v"(lc));
#pragma unroll
    for (int m = 0; m < CNT; m++) {
#if TBH_EW_NT_STORE
      // streamed past the caches: C is read again only by a later level
      __builtin_nontemporal_store(Trop<T>::canon(acc[m]), &C[lc + hi_at(e.hTC, m)]);
#else
      C[lc + hi_at(e.hTC, m)] = Trop<T>::canon(acc[m]);
#endif
    }
  }
  __builtin_amdgcn_s_setprio(0);
}

template <typename T, bool HASB>
__device__ __forceinline__ void ew_tile_dispatch(KEw d, const EwOuter& off,
                                                 const EwThread& e, T* __restrict__ lA,
                                                 T* __restrict__ lB, uint64_t& stA,
                                                 uint64_t& stB) {
  switch (d->nT <= 8 ? 0 : d->nT - 8) {
    case 0: ew_tile<T, HASB, 1>(d, off, e, lA, lB, stA, stB); break;
    case 1: ew_tile<T, HASB, 2>(d, off, e, lA, lB, stA, stB); break;
    case 2: ew_tile<T, HASB, 4>(d, off, e, lA, lB, stA, stB); break;
    case 3: ew_tile<T, HASB, 8>(d, off, e, lA, lB, stA, stB); break;
    default:
      if constexpr (HASB) {
        switch (d->mlay) {
          case 1: ew_tile<T, true, 16, 1>(d, off, e, lA, lB, stA, stB); return;
          case 2: ew_tile<T, true, 16, 2>(d, off, e, lA, lB, stA, stB); return;
          case 3: ew_tile<T, true, 16, 3>(d, off, e, lA, lB, stA, stB); return;
          default: break;
        }
      }
      ew_tile<T, HASB, 16>(d, off, e, lA, lB, stA, stB);
      break;
  }
}

// Direct steps: wave-level tiles of at most 2^10 outputs (lane = tile bits
// 0..5, a lane's 16 elements = bits 6..9), operands read straight from
// global memory -- used when every operand is either small (cache resident)
// or aligned, i.e. its lowest positions are tile lane bits or reduced
// labels, so the loads are coalesced without LDS and without barriers.  The
// operand position of tile bit k is sA[tA[k]] (sB[tB[k]]), of reduced bit i
// sA[rA[i]], of outer bit k oA[k].
__device__ __forceinline__ uint32_t sub_stride(KI8 s, int t) {
  return t >= 0 ? pos_stride(s[t]) : 0u;
}

struct EwLane {  // per-step constants of one lane
  uint32_t la, lb, lc;
  uint32_t hA[4], hB[4], hC[4];
};

__device__ __forceinline__ void ew_lane_setup(KEw d, int lane, EwLane& L) {
  const int nT = d->nT;
  L.la = L.lb = L.lc = 0;
#pragma unroll
  for (int k = 0; k < 6; k++)
    if (k < nT && ((lane >> k) & 1)) {
      L.la += sub_stride(TBH_KF(d, sA), TBH_KF(d, tA)[k]);
      L.lb += sub_stride(TBH_KF(d, sB), TBH_KF(d, tB)[k]);
      L.lc += pos_stride(TBH_KF(d, tC)[k]);
    }
#pragma unroll
  for (int b = 0; b < 4; b++) {
    L.hA[b] = 6 + b < nT ? sub_stride(TBH_KF(d, sA), TBH_KF(d, tA)[6 + b]) : 0u;
    L.hB[b] = 6 + b < nT ? sub_stride(TBH_KF(d, sB), TBH_KF(d, tB)[6 + b]) : 0u;
    L.hC[b] = 6 + b < nT ? pos_stride(TBH_KF(d, tC)[6 + b]) : 0u;
  }
}

// CNT = max(1, 2^nT / 64) outputs per lane (nT < 6: the low 2^nT lanes).
template <typename T, bool HASB, int CNT>
__device__ __forceinline__ void ew_direct_tile(KEw d, const EwLane& L,
                                               const EwOuter& off, int lane) {
  if (CNT == 1 && lane >= (1 << d->nT)) return;
  const TBH_GLOBAL T* __restrict__ A = reinterpret_cast<const TBH_GLOBAL T*>(d->a) + off.a;
  const TBH_GLOBAL T* __restrict__ B = reinterpret_cast<const TBH_GLOBAL T*>(d->b) + off.b;
  TBH_GLOBAL T* __restrict__ C = reinterpret_cast<TBH_GLOBAL T*>(d->c) + off.c;
  const int nR = d->nR;
  uint32_t ia[CNT], ib[CNT];
  T acc[CNT];
#pragma unroll
  for (int m = 0; m < CNT; m++) {
    ia[m] = L.la + hi_at(L.hA, m);
    ib[m] = L.lb + hi_at(L.hB, m);
  }
  if (nR <= 1) {
    // the common spine step: both reduced configurations' loads in flight
    // before any is consumed
    const uint32_t sA = nR ? sub_stride(TBH_KF(d, sA), TBH_KF(d, rA)[0]) : 0u;
    const uint32_t sB = nR ? sub_stride(TBH_KF(d, sB), TBH_KF(d, rB)[0]) : 0u;
    T x0[CNT], x1[CNT];
#pragma unroll
    for (int m = 0; m < CNT; m++) {
      x0[m] = A[ia[m]];
      if (nR) x1[m] = A[ia[m] + sA];
      if (HASB) {
        x0[m] = Trop<T>::mul(x0[m], B[ib[m]]);
        if (nR) x1[m] = Trop<T>::mul(x1[m], B[ib[m] + sB]);
      }
    }
#pragma unroll
    for (int m = 0; m < CNT; m++) acc[m] = nR ? Trop<T>::add(x0[m], x1[m]) : x0[m];
  } else {
#pragma unroll
    for (int m = 0; m < CNT; m++) acc[m] = Trop<T>::zero();
    uint32_t rA = 0, rB = 0;
    const uint32_t nr = 1u << nR;
    for (uint32_t r = 0; r < nr; r++) {
      if (r) {
        const int i = __builtin_ctz(r);
        const uint32_t g = r ^ (r >> 1);
        const uint32_t sa = sub_stride(TBH_KF(d, sA), TBH_KF(d, rA)[i]), sb = sub_stride(TBH_KF(d, sB), TBH_KF(d, rB)[i]);
        if ((g >> i) & 1) {
          rA += sa;
          rB += sb;
        } else {
          rA -= sa;
          rB -= sb;
        }
      }
#pragma unroll
      for (int m = 0; m < CNT; m++) {
        T x = A[ia[m] + rA];
        if (HASB) x = Trop<T>::mul(x, B[ib[m] + rB]);
        acc[m] = Trop<T>::add(acc[m], x);
      }
    }
  }
#pragma unroll
  for (int m = 0; m < CNT; m++) C[L.lc + hi_at(L.hC, m)] = Trop<T>::canon(acc[m]);
}

template <typename T, bool HASB>
__device__ __forceinline__ void ew_direct_dispatch(KEw d, const EwLane& L,
                                                   const EwOuter& off, int lane) {
  switch (d->nT <= 6 ? 0 : d->nT - 6) {
    case 0: ew_direct_tile<T, HASB, 1>(d, L, off, lane); break;
    case 1: ew_direct_tile<T, HASB, 2>(d, L, off, lane); break;
    case 2: ew_direct_tile<T, HASB, 4>(d, L, off, lane); break;
    case 3: ew_direct_tile<T, HASB, 8>(d, L, off, lane); break;
    default: ew_direct_tile<T, HASB, 16>(d, L, off, lane); break;
  }
}

// Persistent over wave tiles: prefix[i] is the first tile of step i.
template <typename T>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(TBH_DIRECT_WAVES))) void tbh_ew_direct_kernel(const EwDesc* __restrict__ ops,
                                                            const int64_t* __restrict__ prefix,
                                                            int nops, int64_t total, int chunk) {
  const int lane = threadIdx.x & 63;
  const int wid = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  const int64_t nw = (int64_t)gridDim.x * 4;
  const int64_t gw = (int64_t)blockIdx.x * 4 + wid;
  EwLane L;
  EwOuter off{0, 0, 0};
  int op = 0;
  int64_t ob = 0, oe = 0;  // oe = 0: the first chunk searches
  for (int64_t c0 = gw * chunk; c0 < total; c0 += nw * chunk) {
    const int64_t c1 = c0 + chunk < total ? c0 + chunk : total;
    // a wave's chunks only move forward: search from the current step
    if (c0 >= oe) {
      op = find_step(prefix, nops, op, c0);
      ob = uniform64(((const TBH_KCONST int64_t*)prefix)[op]);
      oe = uniform64(((const TBH_KCONST int64_t*)prefix)[op + 1]);
      ew_lane_setup((KEw)(ops + op), lane, L);
    }
    for (int64_t t = c0; t < c1; t++) {
      bool fresh = t == c0;
      while (t >= oe) {
        op++;
        ob = oe;
        oe = uniform64(((const TBH_KCONST int64_t*)prefix)[op + 1]);
        fresh = true;
        if (t < oe) ew_lane_setup((KEw)(ops + op), lane, L);
      }
      KEw d = (KEw)(ops + op);
      const uint32_t o = (uint32_t)(t - ob);
      if (fresh)
        off = outer_offsets(d, o);
      else
        outer_step(d, o, off);
      if (d->b != 0)
        ew_direct_dispatch<T, true>(d, L, off, lane);
      else
        ew_direct_dispatch<T, false>(d, L, off, lane);
    }
  }
}

// Persistent: each workgroup walks chunks of the global tile space; prefix[i]
// is the first tile of step i (prefix[nops] = total).
template <typename T>
__global__ __launch_bounds__(kEwBlock) __attribute__((amdgpu_waves_per_eu(TBH_EW_WAVES))) void tbh_ew_kernel(const EwDesc* __restrict__ ops,
                                                          const int64_t* __restrict__ prefix,
                                                          int nops, int64_t total, int chunk) {
  __shared__ T lA[kLdsElems];
  __shared__ T lB[kLdsElems];
  EwThread e;
  int op = 0;
  int64_t ob = 0, oe = 0;  // oe = 0: the first chunk searches
  uint64_t stA = 0, stB = 0;  // sources of the LDS sub-tensors
  for (int64_t c0 = (int64_t)blockIdx.x * chunk; c0 < total; c0 += (int64_t)gridDim.x * chunk) {
    const int64_t c1 = c0 + chunk < total ? c0 + chunk : total;
    // a workgroup's chunks only move forward: search from the current step
    if (c0 >= oe) {
      op = find_step(prefix, nops, op, c0);
      ob = uniform64(((const TBH_KCONST int64_t*)prefix)[op]);
      oe = uniform64(((const TBH_KCONST int64_t*)prefix)[op + 1]);
      ew_thread_setup((KEw)(ops + op), e);
      stA = stB = 0;
    }
    EwOuter off{0, 0, 0};
    for (int64_t t = c0; t < c1; t++) {
      bool fresh = t == c0;
      while (t >= oe) {
        op++;
        ob = oe;
        oe = uniform64(((const TBH_KCONST int64_t*)prefix)[op + 1]);
        fresh = true;
        stA = stB = 0;
        if (t < oe) ew_thread_setup((KEw)(ops + op), e);
      }
      KEw d = (KEw)(ops + op);
      const uint32_t o = (uint32_t)(t - ob);
      if (fresh)
        off = outer_offsets(d, o);
      else
        outer_step(d, o, off);
      if (d->b != 0)
        ew_tile_dispatch<T, true>(d, off, e, lA, lB, stA, stB);
      else
        ew_tile_dispatch<T, false>(d, off, e, lA, lB, stA, stB);
    }
  }
}

// ------------------------------------------------- K2, prefetching ----
// tbh_ew_pf_kernel: the LDS-tiled EW step with the next unit's sub-tensors
// loaded into registers while the current unit is computed.  A unit is one
// staging of a tile (one configuration of its outer reduced labels).  One
// LDS buffer pair, as tbh_ew_kernel, so four workgroups stay resident per CU
// (two buffer pairs, with two workgroups per CU, measured 1.7x slower: too
// few waves hide the LDS reads of the compute).  Per unit:
//   registers -> LDS; the previous tile's outputs stored; barrier;
//   the next unit's global loads issued into the registers;
//   compute from LDS; barrier.
// The loads stay in flight through the compute because nothing there waits
// on vmcnt: descriptors are read with scalar loads (constant address space),
// the barriers are a bare s_barrier after an LDS wait (a fence would wait for
// every memory operation), and a tile's outputs are stored only after the
// next unit's registers have landed.  Only units of the same step are
// prefetched: steps of >= 256 tiles hold nearly all of the work, and a
// workgroup meets a step boundary at most once per step.
// Selected by the ew_prefetch tuning (default off): on the 3-regular n=300
// slices it runs as fast as tbh_ew_kernel alone (144 vs 140 ms of tiled EW
// per solve) and slower beside the GEMMs of a level (120 VGPRs against 83
// leave them less room), so staging, not its latency, bounds the step.

__device__ __forceinline__ void lds_barrier() {
  asm volatile("s_waitcnt lgkmcnt(0)\n\ts_barrier" ::: "memory");
}

// One operand's sub-tensor (2^nS entries) into registers: entry tid + 256 m
// in v[m].  Every register is loaded, whatever nS: the strides of sub-tensor
// bits at or above nS are 0, so the loads past the sub-tensor repeat
// addresses inside it (L1 hits).  Loads under a guard would make each
// register a merge of its old and new value, which the compiler resolves by
// waiting for the load and copying -- the wait this kernel exists to avoid;
// for the same reason there is one load form (dword: a 16-byte form beside
// it made the compiler's wait analysis merge both forms' pending registers
// and wait for every load before issuing the next unit's).
template <typename T>
__device__ __forceinline__ void pf_load(const TBH_GLOBAL T* __restrict__ X, uint32_t lo,
                                        const uint32_t (&hS)[4], T (&v)[16]) {
  // a uniform base plus a 32-bit byte offset per lane (operands hold at most
  // 2^30 entries): global loads with an SGPR base, one VGPR of address each
  const TBH_GLOBAL char* Xb = reinterpret_cast<const TBH_GLOBAL char*>(X);
  // recomputed per call: the per-load offsets are not held across the unit loop
  asm volatile("" : "+v"(lo));
#pragma unroll
  for (int m = 0; m < 16; m++)
    v[m] = *reinterpret_cast<const TBH_GLOBAL T*>(Xb + (uint32_t)((lo + hi_at(hS, m)) * sizeof(T)));
}

// the registers of pf_load into the LDS (the layout of ew_stage; entries past
// the sub-tensor land in the unused rest of the 2^12-entry buffer)
template <typename T>
__device__ __forceinline__ void pf_store(T* __restrict__ lX, const T (&v)[16]) {
  uint32_t pt = lds_pad(threadIdx.x);
  asm volatile("" : "+v"(pt));
#pragma unroll
  for (int m = 0; m < 16; m++) lX[pt + lds_pad(256) * m] = v[m];
}

// A workgroup's tiles: claims of `chunk` tiles, gridDim.x * chunk apart.
// Every field is the same in all lanes; readfirstlane says so, which keeps
// them (and the descriptor reads they index) scalar.
struct PfWalk {
  int64_t t, s, e, total;  // current tile, its claim's start and end
  int64_t stride;
  int chunk;
  __device__ __forceinline__ void next() {
    if (++t >= e) {
      s += stride;
      t = s;
      e = s + chunk < total ? s + chunk : total;
    }
    t = uniform64(t);
    s = uniform64(s);
    e = uniform64(e);
  }
};

__device__ __forceinline__ uint32_t uniform32(uint32_t x) {
  return __builtin_amdgcn_readfirstlane(x);
}

// Every unit of the workgroup's tiles in step d (tiles [ob, oe)), from the
// walk's current tile; returns with the walk at its first tile past the step.
// The first unit is loaded before the loop, so every iteration starts with
// its registers in flight from the previous one (a loop entry without them
// would make the compiler wait on them where the next loads are issued).
template <typename T, bool HASB, int CNT, int MA>
__device__ __forceinline__ void pf_run(KEwS d, int64_t ob, int64_t oe, const EwThread& e,
                                       PfWalk& w, T* __restrict__ lA, T* __restrict__ lB) {
  const int tid = threadIdx.x;
  const bool active = CNT > 1 || tid < (1 << d->nT);
  const TBH_GLOBAL T* A0 = reinterpret_cast<const TBH_GLOBAL T*>(d->a);
  const TBH_GLOBAL T* B0 = reinterpret_cast<const TBH_GLOBAL T*>(d->b);
  TBH_GLOBAL T* C0 = reinterpret_cast<TBH_GLOBAL T*>(d->c);
  const uint32_t nc = 1u << d->nRo;
  // the unit in the registers: tile t at outer offsets off, configuration c
  // of its outer reduced labels
  int64_t t = w.t;
  EwOuter off = outer_offsets(d, uniform32((uint32_t)(t - ob)));
  uint32_t c = 0, cA = 0, cB = 0;
  T va[16], vb[16];
  pf_load<T>(A0 + off.a, e.loSA, e.hSA, va);
  if (HASB) pf_load<T>(B0 + off.b, e.loSB, e.hSB, vb);
  T acc[CNT];
#pragma unroll
  for (int m = 0; m < CNT; m++) acc[m] = Trop<T>::zero();
  bool pend = false;  // acc holds a finished tile (outputs at pC)
  uint32_t pC = 0;
  for (;;) {
    pf_store<T>(lA, va);
    if (HASB) pf_store<T>(lB, vb);
    if (pend) {
      if (active) {
        uint32_t lc = e.loTC;
        asm volatile("" : "+v"(lc));
#pragma unroll
        for (int m = 0; m < CNT; m++) C0[pC + lc + hi_at(e.hTC, m)] = Trop<T>::canon(acc[m]);
      }
#pragma unroll
      for (int m = 0; m < CNT; m++) acc[m] = Trop<T>::zero();
      pend = false;
    }
    lds_barrier();
    // the next unit: the tile's next configuration, or the next tile
    int64_t tn = t;
    uint32_t cn = 0, cAn = 0, cBn = 0;
    EwOuter offn = off;
    bool more = true;
    if (c + 1 < nc) {
      cn = c + 1;
      const int i = __builtin_ctz(cn);
      const uint32_t g = cn ^ (cn >> 1);
      const uint32_t sa = pos_stride(TBH_KF(d, qA)[i]), sb = pos_stride(TBH_KF(d, qB)[i]);
      cAn = ((g >> i) & 1) ? cA + sa : cA - sa;
      cBn = ((g >> i) & 1) ? cB + sb : cB - sb;
    } else {
      w.next();
      tn = w.t;
      more = tn < oe;
      if (more) {
        const uint32_t o = uniform32((uint32_t)(tn - ob));
        if (tn == t + 1)
          outer_step(d, o, offn);
        else
          offn = outer_offsets(d, o);
      }
    }
    if (more) {
      pf_load<T>(A0 + offn.a + cAn, e.loSA, e.hSA, va);
      if (HASB) pf_load<T>(B0 + offn.b + cBn, e.loSB, e.hSB, vb);
    }
    if (active) ew_compute<T, HASB, CNT, MA>(d, e, lA, lB, acc);
    if (c + 1 == nc) {
      pend = true;
      pC = off.c;
    }
    lds_barrier();
    if (!more) break;
    t = tn;
    c = cn;
    cA = cAn;
    cB = cBn;
    off = offn;
  }
  if (active) {  // the last unit always ends a tile
    uint32_t lc = e.loTC;
    asm volatile("" : "+v"(lc));
#pragma unroll
    for (int m = 0; m < CNT; m++) C0[pC + lc + hi_at(e.hTC, m)] = Trop<T>::canon(acc[m]);
  }
}

template <typename T, bool HASB>
__device__ __forceinline__ void pf_dispatch(KEwS d, int64_t ob, int64_t oe, const EwThread& e,
                                            PfWalk& w, T* __restrict__ lA, T* __restrict__ lB) {
  switch (d->nT <= 8 ? 0 : d->nT - 8) {
    case 0: pf_run<T, HASB, 1, 0>(d, ob, oe, e, w, lA, lB); break;
    case 1: pf_run<T, HASB, 2, 0>(d, ob, oe, e, w, lA, lB); break;
    case 2: pf_run<T, HASB, 4, 0>(d, ob, oe, e, w, lA, lB); break;
    case 3: pf_run<T, HASB, 8, 0>(d, ob, oe, e, w, lA, lB); break;
    default:
      if constexpr (HASB) {
        switch (d->mlay) {
          case 1: pf_run<T, true, 16, 1>(d, ob, oe, e, w, lA, lB); return;
          case 2: pf_run<T, true, 16, 2>(d, ob, oe, e, w, lA, lB); return;
          case 3: pf_run<T, true, 16, 3>(d, ob, oe, e, w, lA, lB); return;
          default: break;
        }
      }
      pf_run<T, HASB, 16, 0>(d, ob, oe, e, w, lA, lB);
      break;
  }
}

template <typename T>
__global__ __launch_bounds__(kEwBlock) __attribute__((amdgpu_waves_per_eu(TBH_EW_WAVES))) void
tbh_ew_pf_kernel(const EwDesc* __restrict__ ops_g, const int64_t* __restrict__ prefix_g, int nops,
                 int64_t total, int chunk) {
  __shared__ __align__(16) T lA[kLdsElems];
  __shared__ __align__(16) T lB[kLdsElems];
  const KEwS ops = (KEwS)ops_g;
  const __attribute__((address_space(4))) int64_t* prefix =
      (const __attribute__((address_space(4))) int64_t*)prefix_g;
  PfWalk w;
  w.total = total;
  w.chunk = chunk;
  w.stride = (int64_t)gridDim.x * chunk;
  w.s = (int64_t)blockIdx.x * chunk;
  w.t = w.s;
  w.e = w.s + chunk < total ? w.s + chunk : total;
  int op = 0;
  EwThread e;
  while (w.t < total) {
    // the step holding tile w.t (forward from the current one: galloping,
    // then bisection; prefix[nops] = total > w.t)
    int lo = op, hi = op + 1, step = 1;
    while (hi < nops && prefix[hi] <= w.t) {
      lo = hi;
      step <<= 1;
      hi = lo + step;
    }
    if (hi > nops) hi = nops;
    while (hi - lo > 1) {
      const int mid = (lo + hi) >> 1;
      if (prefix[mid] <= w.t)
        lo = mid;
      else
        hi = mid;
    }
    op = __builtin_amdgcn_readfirstlane(lo);  // uniform: descriptors by scalar loads
    const KEwS d = ops + op;
    ew_thread_setup(d, e);
    if (d->b != 0)
      pf_dispatch<T, true>(d, prefix[op], prefix[op + 1], e, w, lA, lB);
    else
      pf_dispatch<T, false>(d, prefix[op], prefix[op + 1], e, w, lA, lB);
  }
}

// ---------------------------------------------------------------- K1 -----
constexpr int BK = 16, BPAD = 4;

template <typename T>
struct Vec4;
// clang vector types (they copy across address spaces, unlike HIP's float4)
template <>
struct Vec4<float> {
  using type = float __attribute__((ext_vector_type(4)));
};
template <>
struct Vec4<int32_t> {
  using type = int32_t __attribute__((ext_vector_type(4)));
};

// Tile shapes (GemmDesc::shape): 128 x 128, and for tall-skinny steps
// (N <= 64 or <= 32) 256 x 64 and 256 x 32, so that narrow outputs still
// fill the tile.  Each of the 256 threads owns 8 rows x TN columns: rows
// {tx*4 .. +3, BM/2 + tx*4 .. +3} (tx < BM/8), columns {ty*4 .. +3,
// BN/2 + ty*4 .. +3} (TN = 8) or {ty*4 .. +3} (TN = 4).
template <int SHAPE>
struct GemmShape;
template <>
struct GemmShape<0> {
  static constexpr int BM = 128, BN = 128, TN = 8;
};
template <>
struct GemmShape<1> {
  static constexpr int BM = 256, BN = 64, TN = 8;
};
template <>
struct GemmShape<2> {
  static constexpr int BM = 256, BN = 32, TN = 4;
};

// Offset of the output bits x (bit b is GEMM output bit q0 + b) in a C whose
// bit q sits at position pc[q] (-1: no position; only rows / columns past M /
// N, which are not written, reach those).
__device__ __forceinline__ int64_t scatter_bits(const int8_t* __restrict__ pc, int q0, uint32_t x) {
  int64_t r = 0;
  for (int b = 0; x != 0 && q0 + b < kGemmScatter; b++, x >>= 1) {
    const int p = pc[q0 + b];
    if ((x & 1) && p >= 0) r += (int64_t)1 << p;
  }
  return r;
}

// Output tile of a GEMM whose C has its own label order (scatter 2): the
// tile goes to LDS, then the threads read it back in ascending C position
// (enumeration bit k of output e = tid + 256 r is tile bit ep[k]), so that
// consecutive lanes store consecutive addresses.  In LDS, output (m, n) of a
// pass (n without the split bit) sits at (m ^ 4 (n & s)) + 2^vm n: quads of
// rows stay contiguous and 16-byte aligned, and the XOR spreads the columns
// of one row quad over the banks.  Every output of a pass is a valid one: M
// and N are powers of two, so min(M, BM) rows and min(N, BN) columns of the
// tile exist.  P is the pass: with a split, pass P holds the columns whose
// top tile bit is P (the thread's maxima j >= 4 when P = 1), else P = 0
// holds the whole tile.
template <typename T, int SHAPE, int P>
__device__ __forceinline__ void gemm_epilogue_pass(TBH_GLOBAL T* __restrict__ C,
                                                   const T (&acc)[8][GemmShape<SHAPE>::TN],
                                                   int m0, int n0, int M, int N, int64_t base,
                                                   const int8_t* __restrict__ pc, int ma,
                                                   const GemmDesc* __restrict__ d,
                                                   T* __restrict__ st) {
  constexpr int BM = GemmShape<SHAPE>::BM, BN = GemmShape<SHAPE>::BN, TN = GemmShape<SHAPE>::TN;
  constexpr int LBM = BM == 256 ? 8 : 7;
  using V = typename Vec4<T>::type;
  const int tid = threadIdx.x;
  const int tx = tid % (BM / 8), ty = tid / (BM / 8);
  const int vm = ma < LBM ? ma : LBM;
  const bool split = d->ep_split != 0;
  const int nbits = d->ep_n;  // 12 .. 13
  // the XOR mask: row bits 2 .. vm-1 (at most 5 of them)
  const uint32_t smask = (1u << ((vm - 2) < 5 ? (vm - 2) : 5)) - 1;
#pragma unroll
  for (int j = 0; j < TN; j++) {
    const int nl = j < 4 ? ty * 4 + j : BN / 2 + ty * 4 + (j - 4);
    const bool mine = split ? ((j >= 4) == (P == 1)) : true;
    if (mine && n0 + nl < N) {
      const uint32_t ns = split && j >= 4 ? nl - BN / 2 : nl;
#pragma unroll
      for (int h = 0; h < 2; h++) {
        const uint32_t ml = h * (BM / 2) + tx * 4;
        if (m0 + (int)ml < M) {
          const uint32_t phys = (ml ^ ((ns & smask) << 2)) + (ns << vm);
          T x0 = Trop<T>::canon(acc[4 * h][j]), x1 = Trop<T>::canon(acc[4 * h + 1][j]),
            x2 = Trop<T>::canon(acc[4 * h + 2][j]), x3 = Trop<T>::canon(acc[4 * h + 3][j]);
          // opaque: a quad gathered straight from the maxima array gets
          // vectorised into a load from it, which keeps the whole array out
          // of registers (in scratch) through the k loop
          asm volatile("" : "+v"(x0), "+v"(x1), "+v"(x2), "+v"(x3));
          *reinterpret_cast<V*>(st + phys) = V{x0, x1, x2, x3};
        }
      }
    }
  }
  __syncthreads();
  // enumeration bit k as its share of the LDS index (xk: XORed part, ak:
  // added part; row bit b: xk = 2^b; column bit b: xk = 4 (2^b & s), ak =
  // 2^(vm+b)) and of the C offset (gk), decoded once (uniform)
  uint32_t xk[13], ak[13], gk[13];
#pragma unroll
  for (int k = 0; k < 13; k++) {
    xk[k] = ak[k] = gk[k] = 0;
    if (k < nbits) {
      const int b = d->ep[k];
      if (b < 32) {
        xk[k] = 1u << b;
        gk[k] = 1u << pc[b];
      } else {
        xk[k] = ((1u << (b - 32)) & smask) << 2;
        ak[k] = (1u << (b - 32)) << vm;
        gk[k] = 1u << pc[ma + b - 32];
      }
    }
  }
  // enumeration bits 0 and 1 are C positions 0 and 1 (ep lists the pass's
  // bits by C position, and C's lowest positions are pass bits): a thread
  // owns 4 consecutive outputs, one aligned 16-byte store; bits 2..9 are
  // the thread, bits 10.. the iteration r
  uint32_t xt = 0, at = 0, gt = 0;
#pragma unroll
  for (int k = 0; k < 8; k++) {
    const bool on = (tid >> k) & 1;
    xt ^= on ? xk[2 + k] : 0u;
    at += on ? ak[2 + k] : 0u;
    gt += on ? gk[2 + k] : 0u;
  }
  uint32_t li[4];
#pragma unroll
  for (int i = 0; i < 4; i++)
    li[i] = ((i & 1) ? xk[0] : 0u) ^ ((i & 2) ? xk[1] : 0u);
  const uint32_t a1 = ak[0], a2 = ak[1];
  TBH_GLOBAL T* __restrict__ Ct = C + base;
  const int R = 1 << (nbits - 10);  // 4 .. 8 per pass
#pragma unroll
  for (int r = 0; r < 8; r++) {
    if (r >= R) break;
    uint32_t xr = xt, ar = at, gr = gt;
#pragma unroll
    for (int k = 0; k < 3; k++)
      if ((r >> k) & 1) {
        xr ^= xk[10 + k];
        ar += ak[10 + k];
        gr += gk[10 + k];
      }
    V q;
    q.x = st[(xr ^ li[0]) + ar];
    q.y = st[(xr ^ li[1]) + ar + a1];
    q.z = st[(xr ^ li[2]) + ar + a2];
    q.w = st[(xr ^ li[3]) + ar + a1 + a2];
    *reinterpret_cast<TBH_GLOBAL V*>(Ct + gr) = q;
  }
  __syncthreads();
}

template <typename T, int SHAPE>
__device__ __forceinline__ void gemm_epilogue_lds(TBH_GLOBAL T* __restrict__ C,
                                                  const T (&acc)[8][GemmShape<SHAPE>::TN], int m0,
                                                  int n0, int M, int N,
                                                  const int8_t* __restrict__ pc, int ma,
                                                  const GemmDesc* __restrict__ d,
                                                  T* __restrict__ st) {
  constexpr int BN = GemmShape<SHAPE>::BN;
  constexpr int LBN = BN == 128 ? 7 : BN == 64 ? 6 : 5;
  const int64_t base = scatter_bits(pc, 0, m0) + scatter_bits(pc, ma, n0);
  gemm_epilogue_pass<T, SHAPE, 0>(C, acc, m0, n0, M, N, base, pc, ma, d, st);
  if (GemmShape<SHAPE>::TN == 8 && d->ep_split)
    gemm_epilogue_pass<T, SHAPE, 1>(C, acc, m0, n0, M, N,
                                    base + scatter_bits(pc, ma + LBN - 1, 1), pc, ma, d, st);
}

// One BM x BN tile of one batch entry.  VEC: A, B 16-byte aligned with lda,
// ldb, M, K multiples of 4.  pc != nullptr: C is written in its own label
// order (scatter_bits; M a power of two, ma = log2 M), else column-major with
// ldc.
template <typename T, bool VEC, int SHAPE>
__device__ __forceinline__ void gemm_tile(
    const TBH_GLOBAL T* __restrict__ A, const TBH_GLOBAL T* __restrict__ B,
    TBH_GLOBAL T* __restrict__ C, int M, int N, int K, int64_t lda, int64_t ldb, int64_t ldc,
    int m0, int n0, const int8_t* __restrict__ pc, int ma, int nb, const GemmDesc* __restrict__ d,
    T (*As)[BK][GemmShape<SHAPE>::BM], T (*Bs)[BK][GemmShape<SHAPE>::BN + BPAD],
    T* __restrict__ smem) {
  constexpr int BM = GemmShape<SHAPE>::BM, BN = GemmShape<SHAPE>::BN, TN = GemmShape<SHAPE>::TN;
  constexpr int TX = BM / 8;                    // threads along M
  constexpr int NVA = BM * BK / 4 / 256;        // 4-vectors of A per thread
  constexpr int NVBT = BN * BK / 4;             // 4-vectors of B per tile
  constexpr int NVB = (NVBT + 255) / 256;       // ... per thread (some idle)
  constexpr int NEA = BM * BK / 256;            // scalar loads per thread
  constexpr int NEB = (BN * BK + 255) / 256;
  using V = typename Vec4<T>::type;
  const int tid = threadIdx.x;
  const int tx = tid % TX, ty = tid / TX;
  T acc[8][TN];
#pragma unroll
  for (int i = 0; i < 8; i++)
#pragma unroll
    for (int j = 0; j < TN; j++) acc[i][j] = Trop<T>::zero();
  const T pz = pad_zero<T>();
  T ra[VEC ? 4 * NVA : NEA], rb[VEC ? 4 * NVB : NEB];
  auto load = [&](int k0) {
    if (VEC) {
      // A: vector v = tid + 256e covers m = (v % (BM/4))*4 .. +3, k = v / (BM/4)
#pragma unroll
      for (int e = 0; e < NVA; e++) {
        const int v = tid + 256 * e;
        const int m = m0 + (v % (BM / 4)) * 4, k = k0 + v / (BM / 4);
        V x;
        if (m < M && k < K)
          x = *reinterpret_cast<const TBH_GLOBAL V*>(A + (int64_t)k * lda + m);
        else
          x = V{pz, pz, pz, pz};
        ra[4 * e + 0] = x.x;
        ra[4 * e + 1] = x.y;
        ra[4 * e + 2] = x.z;
        ra[4 * e + 3] = x.w;
      }
      // B: vector v covers k = (v & 3)*4 .. +3, n = v >> 2
#pragma unroll
      for (int e = 0; e < NVB; e++) {
        const int v = tid + 256 * e;
        const int k = k0 + (v & 3) * 4, n = n0 + (v >> 2);
        V x;
        if (v < NVBT && n < N && k < K)
          x = *reinterpret_cast<const TBH_GLOBAL V*>(B + (int64_t)n * ldb + k);
        else
          x = V{pz, pz, pz, pz};
        rb[4 * e + 0] = x.x;
        rb[4 * e + 1] = x.y;
        rb[4 * e + 2] = x.z;
        rb[4 * e + 3] = x.w;
      }
    } else {
#pragma unroll
      for (int e = 0; e < NEA; e++) {
        const int idx = tid + 256 * e;
        const int m = m0 + idx % BM, k = k0 + idx / BM;
        ra[e] = (m < M && k < K) ? A[(int64_t)k * lda + m] : pz;
      }
#pragma unroll
      for (int e = 0; e < NEB; e++) {
        const int idx = tid + 256 * e;
        const int k = k0 + idx % BK, n = n0 + idx / BK;
        rb[e] = (idx < BN * BK && n < N && k < K) ? B[(int64_t)n * ldb + k] : pz;
      }
    }
  };
  auto store = [&](int buf) {
    if (VEC) {
#pragma unroll
      for (int e = 0; e < NVA; e++) {
        const int v = tid + 256 * e;
        *reinterpret_cast<V*>(&As[buf][v / (BM / 4)][(v % (BM / 4)) * 4]) =
            V{ra[4 * e], ra[4 * e + 1], ra[4 * e + 2], ra[4 * e + 3]};
      }
#pragma unroll
      for (int e = 0; e < NVB; e++) {
        const int v = tid + 256 * e;
        if (v < NVBT)
#pragma unroll
          for (int q = 0; q < 4; q++) Bs[buf][(v & 3) * 4 + q][v >> 2] = rb[4 * e + q];
      }
    } else {
#pragma unroll
      for (int e = 0; e < NEA; e++) {
        const int idx = tid + 256 * e;
        As[buf][idx / BM][idx % BM] = ra[e];
      }
#pragma unroll
      for (int e = 0; e < NEB; e++) {
        const int idx = tid + 256 * e;
        if (idx < BN * BK) Bs[buf][idx % BK][idx / BK] = rb[e];
      }
    }
  };
  load(0);
  store(0);
  __syncthreads();
  int buf = 0;
  for (int k0 = 0; k0 < K; k0 += BK) {
    const bool more = k0 + BK < K;
    if (more) load(k0 + BK);
#pragma unroll
    for (int kk = 0; kk < BK; kk += 2) {
      T a0[8], b0[TN], a1[8], b1[TN];
      V v;
#pragma unroll
      for (int h = 0; h < 2; h++) {
        v = *reinterpret_cast<const V*>(&As[buf][kk][h * (BM / 2) + tx * 4]);
        a0[4 * h] = v.x; a0[4 * h + 1] = v.y; a0[4 * h + 2] = v.z; a0[4 * h + 3] = v.w;
        v = *reinterpret_cast<const V*>(&As[buf][kk + 1][h * (BM / 2) + tx * 4]);
        a1[4 * h] = v.x; a1[4 * h + 1] = v.y; a1[4 * h + 2] = v.z; a1[4 * h + 3] = v.w;
      }
#pragma unroll
      for (int h = 0; h < TN / 4; h++) {
        v = *reinterpret_cast<const V*>(&Bs[buf][kk][h * (BN / 2) + ty * 4]);
        b0[4 * h] = v.x; b0[4 * h + 1] = v.y; b0[4 * h + 2] = v.z; b0[4 * h + 3] = v.w;
        v = *reinterpret_cast<const V*>(&Bs[buf][kk + 1][h * (BN / 2) + ty * 4]);
        b1[4 * h] = v.x; b1[4 * h + 1] = v.y; b1[4 * h + 2] = v.z; b1[4 * h + 3] = v.w;
      }
#pragma unroll
      for (int i = 0; i < 8; i++)
#pragma unroll
        for (int j = 0; j < TN; j++)
          acc[i][j] = max3(acc[i][j], Trop<T>::mul(a0[i], b0[j]), Trop<T>::mul(a1[i], b1[j]));
      // one k pair at a time: keeps the LDS values of later pairs from being
      // hoisted into registers
      __builtin_amdgcn_sched_barrier(0);
    }
    if (more) store(buf ^ 1);
    __syncthreads();
    buf ^= 1;
  }
  if (pc && d->scatter == 2) {
    gemm_epilogue_lds<T, SHAPE>(C, acc, m0, n0, M, N, pc, ma, d, smem);
    return;
  }
  if (pc) {
    // row m0 + h BM/2 + tx 4 + i: its bits are disjoint parts (m0 a multiple
    // of BM, tx 4 < BM/2, i < 4), so its offset is a sum of per-part offsets;
    // the same for columns.  Row bits 0 and 1 sit at C positions 0 and 1 (pc
    // maps them so), every other bit at a position >= 2: the four rows
    // tx 4 .. +3 are one aligned quad.
    const int64_t rb = scatter_bits(pc, 0, m0) + scatter_bits(pc, 0, tx * 4);
    const int64_t rh = scatter_bits(pc, 0, BM / 2);
    const int64_t cb = scatter_bits(pc, ma, n0) + scatter_bits(pc, ma, ty * 4);
    const int64_t ch = scatter_bits(pc, ma, BN / 2), c1 = scatter_bits(pc, ma, 1),
                  c2 = scatter_bits(pc, ma, 2);
#pragma unroll
    for (int j = 0; j < TN; j++) {
      const int n = n0 + (j < 4 ? ty * 4 + j : BN / 2 + ty * 4 + (j - 4));
      if (n >= N) continue;
      const int64_t co = cb + (j >= 4 ? ch : 0) + ((j & 1) ? c1 : 0) + ((j & 2) ? c2 : 0);
#pragma unroll
      for (int h = 0; h < 2; h++) {
        const int m = m0 + h * (BM / 2) + tx * 4;
        if (m < M)
          *reinterpret_cast<TBH_GLOBAL V*>(C + co + rb + (h ? rh : 0)) =
              V{Trop<T>::canon(acc[4 * h][j]), Trop<T>::canon(acc[4 * h + 1][j]),
                Trop<T>::canon(acc[4 * h + 2][j]), Trop<T>::canon(acc[4 * h + 3][j])};
      }
    }
    return;
  }
#pragma unroll
  for (int j = 0; j < TN; j++) {
    const int n = n0 + (j < 4 ? ty * 4 + j : BN / 2 + ty * 4 + (j - 4));
    if (n >= N) continue;
    if (VEC) {  // M and ldc are multiples of 4: rows come in aligned quads
#pragma unroll
      for (int h = 0; h < 2; h++) {
        const int m = m0 + h * (BM / 2) + tx * 4;
        if (m < M)
          *reinterpret_cast<TBH_GLOBAL V*>(C + (int64_t)n * ldc + m) =
              V{Trop<T>::canon(acc[4 * h][j]), Trop<T>::canon(acc[4 * h + 1][j]),
                Trop<T>::canon(acc[4 * h + 2][j]), Trop<T>::canon(acc[4 * h + 3][j])};
      }
    } else {
#pragma unroll
      for (int i = 0; i < 8; i++) {
        const int m = m0 + (i < 4 ? tx * 4 + i : BM / 2 + tx * 4 + (i - 4));
        if (m < M) C[(int64_t)n * ldc + m] = Trop<T>::canon(acc[i][j]);
      }
    }
  }
}

// Grouped: prefix[i] is the first tile of GEMM step i; a block walks tiles.
// Every step of one launch has the tile shape SHAPE.
// VEC kernels (every batched and plain launch) fit three resident workgroups
// per CU without spills (<= 168 VGPRs); the scalar-load variants keep two.
template <typename T, bool VEC, int SHAPE>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(VEC ? TBH_GEMM_WAVES : 2))) void
tbh_gemm_kernel(const GemmDesc* __restrict__ ops, const int64_t* __restrict__ prefix, int nops,
                int64_t total) {
  constexpr int BM = GemmShape<SHAPE>::BM, BN = GemmShape<SHAPE>::BN;
  // A and B tiles (double-buffered), and after the k loop the output tile
  // of a scattered epilogue (up to kEpiElems entries)
  constexpr int NA = 2 * BK * BM, NB = 2 * BK * (BN + BPAD);
  __shared__ __align__(16) T smem[NA + NB > kEpiElems ? NA + NB : kEpiElems];
  auto As = reinterpret_cast<T(*)[BK][BM]>(smem);
  auto Bs = reinterpret_cast<T(*)[BK][BN + BPAD]>(smem + NA);
  for (int64_t tile = blockIdx.x; tile < total; tile += gridDim.x) {
    int lo = 0, hi = nops;
    while (hi - lo > 1) {
      const int mid = (lo + hi) >> 1;
      if (prefix[mid] <= tile)
        lo = mid;
      else
        hi = mid;
    }
    const GemmDesc* d = ops + lo;
    const int64_t local = tile - prefix[lo];
    const int64_t per = (int64_t)d->tiles_m * d->tiles_n;
    const int64_t t = local / per;
    const int64_t rem = local - t * per;
    const int tm = (int)(rem % d->tiles_m), tn = (int)(rem / d->tiles_m);
    const TBH_GLOBAL T* A = reinterpret_cast<const TBH_GLOBAL T*>(d->a) + t * d->sA;
    const TBH_GLOBAL T* B = reinterpret_cast<const TBH_GLOBAL T*>(d->b) + t * d->sB;
    const int8_t* pc = d->scatter ? d->pc : nullptr;
    const int ma = pc ? __builtin_ctz((uint32_t)d->M) : 0;
    const int nb = pc ? __builtin_ctz((uint32_t)d->N) : 0;
    TBH_GLOBAL T* C = reinterpret_cast<TBH_GLOBAL T*>(d->c) +
                      (pc ? scatter_bits(pc, ma + nb, (uint32_t)t) : t * d->sC);
    gemm_tile<T, VEC, SHAPE>(A, B, C, d->M, d->N, d->K, d->lda, d->ldb, d->ldc, tm * BM, tn * BN,
                             pc, ma, nb, d, As, Bs, smem);
    __syncthreads();
  }
}

// ------------------------------------------------------ VALU probe -----
// The GEMM inner loop's instruction mix with register operands only: per
// iteration an 8 x 8 block of maxima takes two k values (128 adds, 64 max3
// for 128 tropical MACs).  The empty asm marks a and b as changed so the
// adds are not hoisted out of the loop; one store per thread keeps the
// maxima live.
template <typename T>
__global__ __launch_bounds__(256) void tbh_valu_probe_kernel(T* __restrict__ out, int iters) {
  const int tid = blockIdx.x * blockDim.x + threadIdx.x;
  T a0[8], a1[8], b0[8], b1[8], acc[8][8];
#pragma unroll
  for (int i = 0; i < 8; i++) {
    a0[i] = (T)(tid & 7) + (T)i;
    a1[i] = (T)((tid >> 3) & 7) + (T)i;
    b0[i] = (T)(2 * i);
    b1[i] = (T)(3 * i);
#pragma unroll
    for (int j = 0; j < 8; j++) acc[i][j] = Trop<T>::zero();
  }
  for (int it = 0; it < iters; it++) {
#pragma unroll
    for (int i = 0; i < 8; i++)
#pragma unroll
      for (int j = 0; j < 8; j++)
        acc[i][j] = max3(acc[i][j], Trop<T>::mul(a0[i], b0[j]), Trop<T>::mul(a1[i], b1[j]));
#pragma unroll
    for (int i = 0; i < 8; i++) asm volatile("" : "+v"(a0[i]), "+v"(a1[i]), "+v"(b0[i]), "+v"(b1[i]));
  }
  T m = Trop<T>::zero();
#pragma unroll
  for (int i = 0; i < 8; i++)
#pragma unroll
    for (int j = 0; j < 8; j++) m = Trop<T>::add(m, acc[i][j]);
  out[tid] = m;
}

// ---------------------------------------------------------------- K3 -----
template <typename T>
__global__ void tbh_leaf_kernel(const LeafDesc* __restrict__ leaves, int64_t n) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const LeafDesc L = leaves[i];
  T* p = reinterpret_cast<T*>(L.p);
  if (L.kind == 1) {
    T w;
    __builtin_memcpy(&w, &L.wbits, sizeof(T));
    p[0] = Trop<T>::one();
    p[1] = w;
  } else {
    p[0] = Trop<T>::one();
    p[1] = Trop<T>::one();
    p[2] = Trop<T>::one();
    p[3] = pad_zero<T>();
  }
}

// ---------------------------------------------------------------- K5 -----
template <typename T>
__global__ void tbh_epi_kernel(const EpiDesc* __restrict__ e, int64_t n, T* __restrict__ out) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const EpiDesc E = e[i];
  T r;
  __builtin_memcpy(&r, &E.rbits, sizeof(T));
  if (!E.root) {
    out[i] = r;
    return;
  }
  const T v = *reinterpret_cast<const T*>(E.root);
  out[i] = Trop<T>::canon(Trop<T>::mul(v, r));
}

// -------------------------------------------------------- launchers -----
hipError_t launch_ew(int dtype, const EwDesc* ops, const int64_t* prefix, int nops, int64_t total,
                     int grid, int chunk, hipStream_t s) {
  if (total <= 0) return hipSuccess;
  if (dtype == 0)
    tbh_ew_kernel<float><<<grid, kEwBlock, 0, s>>>(ops, prefix, nops, total, chunk);
  else
    tbh_ew_kernel<int32_t><<<grid, kEwBlock, 0, s>>>(ops, prefix, nops, total, chunk);
  return hipGetLastError();
}

int ew_blocks_per_cu(int dtype) {
  int n = 0;
  const hipError_t err =
      dtype == 0 ? hipOccupancyMaxActiveBlocksPerMultiprocessor(&n, tbh_ew_kernel<float>, kEwBlock, 0)
                 : hipOccupancyMaxActiveBlocksPerMultiprocessor(&n, tbh_ew_kernel<int32_t>, kEwBlock, 0);
  return err == hipSuccess && n > 0 ? n : 4;
}

hipError_t launch_ew_pf(int dtype, const EwDesc* ops, const int64_t* prefix, int nops,
                        int64_t total, int grid, int chunk, hipStream_t s) {
  if (total <= 0) return hipSuccess;
  if (dtype == 0)
    tbh_ew_pf_kernel<float><<<grid, kEwBlock, 0, s>>>(ops, prefix, nops, total, chunk);
  else
    tbh_ew_pf_kernel<int32_t><<<grid, kEwBlock, 0, s>>>(ops, prefix, nops, total, chunk);
  return hipGetLastError();
}

int ew_pf_blocks_per_cu(int dtype) {
  int n = 0;
  const hipError_t err =
      dtype == 0
          ? hipOccupancyMaxActiveBlocksPerMultiprocessor(&n, tbh_ew_pf_kernel<float>, kEwBlock, 0)
          : hipOccupancyMaxActiveBlocksPerMultiprocessor(&n, tbh_ew_pf_kernel<int32_t>, kEwBlock, 0);
  return err == hipSuccess && n > 0 ? n : 4;
}

hipError_t launch_ew_direct(int dtype, const EwDesc* ops, const int64_t* prefix, int nops,
                            int64_t total, int grid, int chunk, hipStream_t s) {
  if (total <= 0) return hipSuccess;
  if (dtype == 0)
    tbh_ew_direct_kernel<float><<<grid, 256, 0, s>>>(ops, prefix, nops, total, chunk);
  else
    tbh_ew_direct_kernel<int32_t><<<grid, 256, 0, s>>>(ops, prefix, nops, total, chunk);
  return hipGetLastError();
}

hipError_t launch_gemm(int dtype, bool vec, int shape, const GemmDesc* ops, const int64_t* prefix,
                       int nops, int64_t total, int grid, hipStream_t s) {
  if (total <= 0) return hipSuccess;
#define TBH_GEMM_LAUNCH(T, V, S) tbh_gemm_kernel<T, V, S><<<grid, 256, 0, s>>>(ops, prefix, nops, total)
#define TBH_GEMM_SHAPES(T, V)      \
  if (shape == 1)                  \
    TBH_GEMM_LAUNCH(T, V, 1);      \
  else if (shape == 2)             \
    TBH_GEMM_LAUNCH(T, V, 2);      \
  else                             \
    TBH_GEMM_LAUNCH(T, V, 0);
  if (dtype == 0) {
    if (vec) {
      TBH_GEMM_SHAPES(float, true)
    } else {
      TBH_GEMM_SHAPES(float, false)
    }
  } else {
    if (vec) {
      TBH_GEMM_SHAPES(int32_t, true)
    } else {
      TBH_GEMM_SHAPES(int32_t, false)
    }
  }
#undef TBH_GEMM_SHAPES
#undef TBH_GEMM_LAUNCH
  return hipGetLastError();
}

int gemm_blocks_per_cu(int dtype, bool vec, int shape) {
  int n = 0;
  hipError_t err = hipSuccess;
#define TBH_GEMM_OCC(T, V, S) \
  err = hipOccupancyMaxActiveBlocksPerMultiprocessor(&n, tbh_gemm_kernel<T, V, S>, 256, 0)
#define TBH_GEMM_OCC_SHAPES(T, V) \
  if (shape == 1)                 \
    TBH_GEMM_OCC(T, V, 1);        \
  else if (shape == 2)            \
    TBH_GEMM_OCC(T, V, 2);        \
  else                            \
    TBH_GEMM_OCC(T, V, 0);
  if (dtype == 0) {
    if (vec) {
      TBH_GEMM_OCC_SHAPES(float, true)
    } else {
      TBH_GEMM_OCC_SHAPES(float, false)
    }
  } else {
    if (vec) {
      TBH_GEMM_OCC_SHAPES(int32_t, true)
    } else {
      TBH_GEMM_OCC_SHAPES(int32_t, false)
    }
  }
#undef TBH_GEMM_OCC_SHAPES
#undef TBH_GEMM_OCC
  return err == hipSuccess && n > 0 ? n : 2;
}

hipError_t launch_leaves(int dtype, const LeafDesc* leaves, int64_t n, hipStream_t s) {
  if (n <= 0) return hipSuccess;
  const int grid = (int)((n + 255) / 256);
  if (dtype == 0)
    tbh_leaf_kernel<float><<<grid, 256, 0, s>>>(leaves, n);
  else
    tbh_leaf_kernel<int32_t><<<grid, 256, 0, s>>>(leaves, n);
  return hipGetLastError();
}

hipError_t launch_epilogue(int dtype, const EpiDesc* e, int64_t n, void* out, hipStream_t s) {
  if (n <= 0) return hipSuccess;
  const int grid = (int)((n + 255) / 256);
  if (dtype == 0)
    tbh_epi_kernel<float><<<grid, 256, 0, s>>>(e, n, reinterpret_cast<float*>(out));
  else
    tbh_epi_kernel<int32_t><<<grid, 256, 0, s>>>(e, n, reinterpret_cast<int32_t*>(out));
  return hipGetLastError();
}

hipError_t launch_valu_probe(int dtype, void* out, int blocks, int iters, hipStream_t s) {
  if (dtype == 0)
    tbh_valu_probe_kernel<float><<<blocks, 256, 0, s>>>(static_cast<float*>(out), iters);
  else
    tbh_valu_probe_kernel<int32_t><<<blocks, 256, 0, s>>>(static_cast<int32_t*>(out), iters);
  return hipGetLastError();
}

}  // namespace tbh
