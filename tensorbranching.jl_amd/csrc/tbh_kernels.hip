// tbh_kernels.hip -- gfx950 kernels of the tropical (max-plus) contraction
// backend.  All arithmetic is VALU: max-plus is not a multiply-accumulate, so
// MFMA has no role.  Tropical Float32 uses -inf as zero; Tropical Int32 uses
// the canonical sentinel -2^30 and rewrites stored values below -2^29 to it.
//
//   K1 tbh_gemm_kernel    batched / grouped max-plus GEMM: 128x128x16 LDS
//                         tiles, 8x8 register blocking, add + max3 pairs.
//   K2 tbh_ew_kernel      one wave-batched EW step: a binary tropical
//                         contraction over bit-permuted 2^r tensors (labels of
//                         size 2), reducing the contracted labels; also the
//                         permute / reduce of a single tensor.
//   K3 tbh_leaf_kernel    leaf tensors of the MIS network from weights.
//   K5 tbh_epi_kernel     per-slice root + r.
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "tbh_internal.hpp"

namespace tbh {

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
// zero used to pad GEMM operands (adds without overflow for Int32)
template <typename T>
__device__ __forceinline__ T pad_zero();
template <>
__device__ __forceinline__ float pad_zero<float>() { return -__builtin_inff(); }
template <>
__device__ __forceinline__ int32_t pad_zero<int32_t>() { return kI32NegInf; }

// ---------------------------------------------------------------- K2 -----
// Reduced-label loop: the lowest NLO (<= 2) reduced bits are unrolled, the
// rest walk a Gray code so each step changes one stride.
template <typename T, int EPT, int NLO, bool HASB>
__device__ __forceinline__ void ew_tile(const T* __restrict__ A, const T* __restrict__ B,
                                        T* __restrict__ C, uint32_t baseA, uint32_t baseB,
                                        const uint32_t (&jA)[EPT], const uint32_t (&jB)[EPT],
                                        const EwDesc* __restrict__ d, int nR, uint32_t cbase,
                                        int lane, bool active) {
  T acc[EPT];
#pragma unroll
  for (int j = 0; j < EPT; j++) acc[j] = Trop<T>::zero();
  uint32_t loA[4] = {0, 0, 0, 0}, loB[4] = {0, 0, 0, 0};
  if (NLO >= 1) {
    loA[1] = pos_stride(d->ra[0]);
    loB[1] = pos_stride(d->rb[0]);
  }
  if (NLO >= 2) {
    loA[2] = pos_stride(d->ra[1]);
    loB[2] = pos_stride(d->rb[1]);
    loA[3] = loA[1] + loA[2];
    loB[3] = loB[1] + loB[2];
  }
  const int nhi = nR - NLO;
  const uint32_t nh = 1u << nhi;
  uint32_t hA = 0, hB = 0;
  for (uint32_t h = 0; h < nh; h++) {
    if (h) {
      const int i = __builtin_ctz(h);
      const uint32_t g = h ^ (h >> 1);
      const uint32_t sa = pos_stride(d->ra[NLO + i]), sb = pos_stride(d->rb[NLO + i]);
      if ((g >> i) & 1) {
        hA += sa;
        hB += sb;
      } else {
        hA -= sa;
        hB -= sb;
      }
    }
#pragma unroll
    for (int q = 0; q < (1 << NLO); q += (NLO ? 2 : 1)) {
#pragma unroll
      for (int j = 0; j < EPT; j++) {
        const uint32_t oa = baseA + jA[j] + hA;
        const uint32_t ob = baseB + jB[j] + hB;
        T x0 = A[oa + loA[q]];
        if (HASB) x0 = Trop<T>::mul(x0, B[ob + loB[q]]);
        if (NLO) {
          T x1 = A[oa + loA[q + 1]];
          if (HASB) x1 = Trop<T>::mul(x1, B[ob + loB[q + 1]]);
          acc[j] = Trop<T>::add(acc[j], Trop<T>::add(x0, x1));
        } else {
          acc[j] = Trop<T>::add(acc[j], x0);
        }
      }
    }
  }
  if (active) {
#pragma unroll
    for (int j = 0; j < EPT; j++) C[cbase + ((uint32_t)j << 6) + lane] = Trop<T>::canon(acc[j]);
  }
}

template <typename T, int EPT, int NLO, bool HASB>
__device__ void ew_range(const EwDesc* __restrict__ d, int64_t l0, int64_t l1, int lane) {
  const int nP = d->nP, nR = d->nR;
  const T* A = reinterpret_cast<const T*>(d->a);
  const T* B = reinterpret_cast<const T*>(d->b);
  T* C = reinterpret_cast<T*>(d->c);
  constexpr int E = EPT == 1 ? 0 : EPT == 2 ? 1 : EPT == 4 ? 2 : EPT == 8 ? 3 : 4;
  uint32_t laneA = 0, laneB = 0;
#pragma unroll
  for (int i = 0; i < 6; i++)
    if (i < nP && ((lane >> i) & 1)) {
      laneA += pos_stride(d->pa[i]);
      laneB += pos_stride(d->pb[i]);
    }
  const bool active = nP >= 6 || lane < (1 << nP);
  uint32_t jA[EPT], jB[EPT];
  jA[0] = 0;
  jB[0] = 0;
#pragma unroll
  for (int b = 0; b < E; b++) {
    const uint32_t sa = pos_stride(d->pa[6 + b]), sb = pos_stride(d->pb[6 + b]);
#pragma unroll
    for (int j = 0; j < (1 << b); j++) {
      jA[j + (1 << b)] = jA[j] + sa;
      jB[j + (1 << b)] = jB[j] + sb;
    }
  }
  const int tb = 6 + E;
  uint32_t tA = 0, tB = 0;
  for (int i = tb; i < nP; i++)
    if ((l0 >> (i - tb)) & 1) {
      tA += pos_stride(d->pa[i]);
      tB += pos_stride(d->pb[i]);
    }
  for (int64_t l = l0; l < l1; l++) {
    if (l != l0) {
      uint64_t flip = (uint64_t)l ^ (uint64_t)(l - 1);
      while (flip) {
        const int i = __builtin_ctzll(flip);
        flip &= flip - 1;
        const uint32_t sa = pos_stride(d->pa[tb + i]), sb = pos_stride(d->pb[tb + i]);
        if (((uint64_t)l >> i) & 1) {
          tA += sa;
          tB += sb;
        } else {
          tA -= sa;
          tB -= sb;
        }
      }
    }
    ew_tile<T, EPT, NLO, HASB>(A, B, C, tA + laneA, tB + laneB, jA, jB, d, nR,
                               (uint32_t)(l << tb), lane, active);
  }
}

template <typename T, int EPT>
__device__ __forceinline__ void ew_dispatch_nlo(const EwDesc* d, int64_t l0, int64_t l1,
                                                int lane) {
  const int nR = d->nR;
  const bool hasB = d->b != 0;
  if (nR == 0) {
    if (hasB)
      ew_range<T, EPT, 0, true>(d, l0, l1, lane);
    else
      ew_range<T, EPT, 0, false>(d, l0, l1, lane);
  } else if (nR == 1) {
    if (hasB)
      ew_range<T, EPT, 1, true>(d, l0, l1, lane);
    else
      ew_range<T, EPT, 1, false>(d, l0, l1, lane);
  } else {
    if (hasB)
      ew_range<T, EPT, 2, true>(d, l0, l1, lane);
    else
      ew_range<T, EPT, 2, false>(d, l0, l1, lane);
  }
}

// Persistent: every wave walks contiguous chunks of the global tile space;
// prefix[i] is the first tile of step i (prefix[nops] = total).
template <typename T>
__global__ __launch_bounds__(kEwBlock) void tbh_ew_kernel(const EwDesc* __restrict__ ops,
                                                          const int64_t* __restrict__ prefix,
                                                          int nops, int64_t total, int chunk) {
  const int lane = threadIdx.x & 63;
  const int wid = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  const int64_t nwaves = (int64_t)gridDim.x * kWavesPerBlock;
  const int64_t gw = (int64_t)blockIdx.x * kWavesPerBlock + wid;
  int op = 0;
  int64_t ob = 0, oe = 0;
  for (int64_t c0 = gw * chunk; c0 < total; c0 += nwaves * chunk) {
    const int64_t c1 = c0 + chunk < total ? c0 + chunk : total;
    int64_t t = c0;
    while (t < c1) {
      if (t < ob || t >= oe) {
        int lo = 0, hi = nops;
        while (hi - lo > 1) {
          const int mid = (lo + hi) >> 1;
          if (prefix[mid] <= t)
            lo = mid;
          else
            hi = mid;
        }
        op = lo;
        ob = prefix[op];
        oe = prefix[op + 1];
      }
      const int64_t te = c1 < oe ? c1 : oe;
      const EwDesc* d = ops + op;
      switch (d->ept) {
        case 0: ew_dispatch_nlo<T, 1>(d, t - ob, te - ob, lane); break;
        case 1: ew_dispatch_nlo<T, 2>(d, t - ob, te - ob, lane); break;
        case 2: ew_dispatch_nlo<T, 4>(d, t - ob, te - ob, lane); break;
        case 3: ew_dispatch_nlo<T, 8>(d, t - ob, te - ob, lane); break;
        default: ew_dispatch_nlo<T, 16>(d, t - ob, te - ob, lane); break;
      }
      t = te;
    }
  }
}

// ---------------------------------------------------------------- K1 -----
constexpr int BM = 128, BN = 128, BK = 16, BPAD = 4;

template <typename T>
struct Vec4;
template <>
struct Vec4<float> {
  using type = float4;
};
template <>
struct Vec4<int32_t> {
  using type = int4;
};

// Block of 256 threads computes a 128x128 tile of one batch entry; thread
// (tx, ty) owns rows {tx*4..+3, 64+tx*4..+3} and the same pattern of columns.
// VEC: A, B 16-byte aligned with lda, ldb, M, K multiples of 4.
template <typename T, bool VEC>
__device__ __forceinline__ void gemm_tile(const T* __restrict__ A, const T* __restrict__ B,
                                          T* __restrict__ C, int M, int N, int K, int64_t lda,
                                          int64_t ldb, int64_t ldc, int m0, int n0,
                                          T (*As)[BK][BM], T (*Bs)[BK][BN + BPAD]) {
  using V = typename Vec4<T>::type;
  const int tid = threadIdx.x;
  const int tx = tid & 15, ty = tid >> 4;
  T acc[8][8];
#pragma unroll
  for (int i = 0; i < 8; i++)
#pragma unroll
    for (int j = 0; j < 8; j++) acc[i][j] = Trop<T>::zero();
  const T pz = pad_zero<T>();
  T ra[8], rb[8];
  auto load = [&](int k0) {
    if (VEC) {
      // A: 2 x 4-vector per thread: m = (tid&31)*4, k = (tid>>5) + 8e
#pragma unroll
      for (int e = 0; e < 2; e++) {
        const int m = m0 + (tid & 31) * 4, k = k0 + (tid >> 5) + 8 * e;
        V v;
        if (m < M && k < K)
          v = *reinterpret_cast<const V*>(A + (int64_t)k * lda + m);
        else
          v = V{pz, pz, pz, pz};
        ra[4 * e + 0] = v.x;
        ra[4 * e + 1] = v.y;
        ra[4 * e + 2] = v.z;
        ra[4 * e + 3] = v.w;
      }
      // B: k = (tid&3)*4, n = (tid>>2) + 64e
#pragma unroll
      for (int e = 0; e < 2; e++) {
        const int k = k0 + (tid & 3) * 4, n = n0 + (tid >> 2) + 64 * e;
        V v;
        if (n < N && k < K)
          v = *reinterpret_cast<const V*>(B + (int64_t)n * ldb + k);
        else
          v = V{pz, pz, pz, pz};
        rb[4 * e + 0] = v.x;
        rb[4 * e + 1] = v.y;
        rb[4 * e + 2] = v.z;
        rb[4 * e + 3] = v.w;
      }
    } else {
#pragma unroll
      for (int e = 0; e < 8; e++) {
        const int m = m0 + (tid & 127), k = k0 + (tid >> 7) + 2 * e;
        ra[e] = (m < M && k < K) ? A[(int64_t)k * lda + m] : pz;
      }
#pragma unroll
      for (int e = 0; e < 8; e++) {
        const int k = k0 + (tid & 15), n = n0 + (tid >> 4) + 16 * e;
        rb[e] = (n < N && k < K) ? B[(int64_t)n * ldb + k] : pz;
      }
    }
  };
  auto store = [&](int buf) {
    if (VEC) {
#pragma unroll
      for (int e = 0; e < 2; e++) {
        *reinterpret_cast<V*>(&As[buf][(tid >> 5) + 8 * e][(tid & 31) * 4]) =
            V{ra[4 * e], ra[4 * e + 1], ra[4 * e + 2], ra[4 * e + 3]};
      }
#pragma unroll
      for (int e = 0; e < 2; e++)
#pragma unroll
        for (int q = 0; q < 4; q++) Bs[buf][(tid & 3) * 4 + q][(tid >> 2) + 64 * e] = rb[4 * e + q];
    } else {
#pragma unroll
      for (int e = 0; e < 8; e++) As[buf][(tid >> 7) + 2 * e][tid & 127] = ra[e];
#pragma unroll
      for (int e = 0; e < 8; e++) Bs[buf][tid & 15][(tid >> 4) + 16 * e] = rb[e];
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
      T a0[8], b0[8], a1[8], b1[8];
      V v;
      v = *reinterpret_cast<const V*>(&As[buf][kk][tx * 4]);
      a0[0] = v.x; a0[1] = v.y; a0[2] = v.z; a0[3] = v.w;
      v = *reinterpret_cast<const V*>(&As[buf][kk][64 + tx * 4]);
      a0[4] = v.x; a0[5] = v.y; a0[6] = v.z; a0[7] = v.w;
      v = *reinterpret_cast<const V*>(&Bs[buf][kk][ty * 4]);
      b0[0] = v.x; b0[1] = v.y; b0[2] = v.z; b0[3] = v.w;
      v = *reinterpret_cast<const V*>(&Bs[buf][kk][64 + ty * 4]);
      b0[4] = v.x; b0[5] = v.y; b0[6] = v.z; b0[7] = v.w;
      v = *reinterpret_cast<const V*>(&As[buf][kk + 1][tx * 4]);
      a1[0] = v.x; a1[1] = v.y; a1[2] = v.z; a1[3] = v.w;
      v = *reinterpret_cast<const V*>(&As[buf][kk + 1][64 + tx * 4]);
      a1[4] = v.x; a1[5] = v.y; a1[6] = v.z; a1[7] = v.w;
      v = *reinterpret_cast<const V*>(&Bs[buf][kk + 1][ty * 4]);
      b1[0] = v.x; b1[1] = v.y; b1[2] = v.z; b1[3] = v.w;
      v = *reinterpret_cast<const V*>(&Bs[buf][kk + 1][64 + ty * 4]);
      b1[4] = v.x; b1[5] = v.y; b1[6] = v.z; b1[7] = v.w;
#pragma unroll
      for (int i = 0; i < 8; i++)
#pragma unroll
        for (int j = 0; j < 8; j++)
          acc[i][j] = Trop<T>::add(acc[i][j], Trop<T>::add(Trop<T>::mul(a0[i], b0[j]),
                                                           Trop<T>::mul(a1[i], b1[j])));
    }
    if (more) store(buf ^ 1);
    __syncthreads();
    buf ^= 1;
  }
#pragma unroll
  for (int j = 0; j < 8; j++) {
    const int n = n0 + (j < 4 ? ty * 4 + j : 64 + ty * 4 + (j - 4));
    if (n >= N) continue;
    if (VEC) {  // M and ldc are multiples of 4: rows come in aligned quads
#pragma unroll
      for (int h = 0; h < 2; h++) {
        const int m = m0 + h * 64 + tx * 4;
        if (m < M)
          *reinterpret_cast<V*>(C + (int64_t)n * ldc + m) =
              V{Trop<T>::canon(acc[4 * h][j]), Trop<T>::canon(acc[4 * h + 1][j]),
                Trop<T>::canon(acc[4 * h + 2][j]), Trop<T>::canon(acc[4 * h + 3][j])};
      }
    } else {
#pragma unroll
      for (int i = 0; i < 8; i++) {
        const int m = m0 + (i < 4 ? tx * 4 + i : 64 + tx * 4 + (i - 4));
        if (m < M) C[(int64_t)n * ldc + m] = Trop<T>::canon(acc[i][j]);
      }
    }
  }
}

// Grouped: prefix[i] is the first tile of GEMM step i; a block walks tiles.
template <typename T, bool VEC>
__global__ __launch_bounds__(256) void tbh_gemm_kernel(const GemmDesc* __restrict__ ops,
                                                       const int64_t* __restrict__ prefix,
                                                       int nops, int64_t total) {
  __shared__ T As[2][BK][BM];
  __shared__ T Bs[2][BK][BN + BPAD];
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
    const T* A = reinterpret_cast<const T*>(d->a) + t * d->sA;
    const T* B = reinterpret_cast<const T*>(d->b) + t * d->sB;
    T* C = reinterpret_cast<T*>(d->c) + t * d->sC;
    gemm_tile<T, VEC>(A, B, C, d->M, d->N, d->K, d->lda, d->ldb, d->ldc, tm * BM, tn * BN, As, Bs);
    __syncthreads();
  }
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

hipError_t launch_gemm(int dtype, bool vec, const GemmDesc* ops, const int64_t* prefix, int nops,
                       int64_t total, int grid, hipStream_t s) {
  if (total <= 0) return hipSuccess;
  if (dtype == 0) {
    if (vec)
      tbh_gemm_kernel<float, true><<<grid, 256, 0, s>>>(ops, prefix, nops, total);
    else
      tbh_gemm_kernel<float, false><<<grid, 256, 0, s>>>(ops, prefix, nops, total);
  } else {
    if (vec)
      tbh_gemm_kernel<int32_t, true><<<grid, 256, 0, s>>>(ops, prefix, nops, total);
    else
      tbh_gemm_kernel<int32_t, false><<<grid, 256, 0, s>>>(ops, prefix, nops, total);
  }
  return hipGetLastError();
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

}  // namespace tbh
