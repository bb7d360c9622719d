// tbh_regions.hip -- branching tables of regions (tbh_region_tables), the
// table step of optimal_branches (src/branch.jl:79): for a region of a MIS
// problem and its boundary ("open") vertices, the reduced alpha tensor --
// per boundary configuration b the region's best weight with its boundary
// fixed to b, with the entries that a configuration using fewer boundary
// vertices matches or beats set to the tropical zero (GenericTensorNetworks'
// mis_compactify!) -- and, per surviving b, every region configuration that
// reaches it (the table rows of a BranchingTable).
//
// The alpha tensors themselves are tropical contractions (SizeMax of the
// region's network with its boundary fixed, one scalar slice per b) and
// come in through tbh_contract_slices, batched over every region and b; this
// file prunes them and enumerates the configurations, both on the GPU:
//   tbh_compactify_kernel   one thread per (region, a): a survives unless a
//                           proper subset b of a has alpha[b] >= alpha[a];
//   tbh_configs_kernel      one thread per (region, x in {0,1}^nv): x is kept
//                           when it is independent in the region and its
//                           weight equals the pruned alpha of its boundary
//                           bits; pass 0 counts per (region, b), pass 1
//                           writes into each group's slots.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>

#include "tbh_runtime.hpp"

namespace tbh {

namespace {

constexpr int kRegionMaxNv = 24;    // 2^24 configurations per region at most
constexpr int kRegionMaxOpen = 16;  // 2^16 boundary configurations

struct RegionDesc {
  int32_t nv, nopen;
  uint32_t open;       // boundary vertices (bitmask over the region)
  int64_t x0;          // first configuration index of this region (prefix)
  int64_t a0;          // first alpha entry of this region (prefix)
  uint32_t adj[kRegionMaxNv];
  uint32_t wbits[kRegionMaxNv];  // dtype bit patterns (unit weights expanded)
};

// bits of x at the set positions of m, packed (x86 pext)
__device__ __forceinline__ uint32_t pext32(uint32_t x, uint32_t m) {
  uint32_t r = 0, k = 0;
  while (m) {
    const uint32_t low = m & (~m + 1u);
    if (x & low) r |= 1u << k;
    k++;
    m &= m - 1;
  }
  return r;
}

template <typename T>
__device__ __forceinline__ T tzero();
template <>
__device__ __forceinline__ float tzero<float>() { return -__builtin_inff(); }
template <>
__device__ __forceinline__ int32_t tzero<int32_t>() { return kI32NegInf; }

template <typename T>
__device__ __forceinline__ bool is_zero(T v) { return v <= tzero<T>(); }

// region of global index g: prefix[i] <= g < prefix[i + 1]
__device__ __forceinline__ int find_region(const int64_t* __restrict__ prefix, int n, int64_t g) {
  int lo = 0, hi = n;
  while (hi - lo > 1) {
    const int mid = (lo + hi) >> 1;
    if (prefix[mid] <= g)
      lo = mid;
    else
      hi = mid;
  }
  return lo;
}

template <typename T>
__global__ __launch_bounds__(256) void tbh_compactify_kernel(const RegionDesc* __restrict__ R,
                                                              const int64_t* __restrict__ aprefix,
                                                              int n, int64_t total,
                                                              const T* __restrict__ in,
                                                              T* __restrict__ out) {
  const int64_t g = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (g >= total) return;
  const int r = find_region(aprefix, n, g);
  const int64_t base = aprefix[r];
  const uint32_t a = (uint32_t)(g - base);
  const T va = in[g];
  bool dominated = false;
  // proper subsets b of a, from a - 1 down (b = (b - 1) & a)
  for (uint32_t b = (a - 1) & a; a != 0; b = (b - 1) & a) {
    if (in[base + b] >= va) {
      dominated = true;
      break;
    }
    if (b == 0) break;
  }
  out[g] = dominated ? tzero<T>() : va;
}

// PASS 0: counts[alpha entry] += 1 per matching configuration; PASS 1: writes
// x at slot[entry]++ within the entry's group (group offsets in `offs`).
template <typename T, int PASS>
__global__ __launch_bounds__(256) void tbh_configs_kernel(const RegionDesc* __restrict__ R,
                                                           const int64_t* __restrict__ xprefix,
                                                           int n, int64_t total,
                                                           const T* __restrict__ alpha,
                                                           uint32_t* __restrict__ counts,
                                                           const int64_t* __restrict__ offs,
                                                           uint32_t* __restrict__ configs) {
  const int64_t g = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (g >= total) return;
  const int r = find_region(xprefix, n, g);
  const RegionDesc& d = R[r];
  const uint32_t x = (uint32_t)(g - xprefix[r]);
  T w = (T)0;
  for (uint32_t m = x; m; m &= m - 1) {
    const int v = __builtin_ctz(m);
    if (d.adj[v] & x) return;  // not independent
    T wv;
    __builtin_memcpy(&wv, &d.wbits[v], sizeof(T));
    w += wv;
  }
  const int64_t e = d.a0 + pext32(x, d.open);
  const T va = alpha[e];
  if (is_zero(va) || w != va) return;
  if (PASS == 0) {
    atomicAdd(&counts[e], 1u);
  } else {
    const uint32_t slot = atomicAdd(&counts[e], 1u);
    configs[offs[e] + slot] = x;
  }
}

}  // namespace
}  // namespace tbh

using namespace tbh;

extern "C" int tbh_region_tables(tbh_ctx* ctx, tbh_dtype dtype, const tbh_region* regions,
                                 size_t n, void* alpha, size_t* group_ptr, uint32_t* configs,
                                 size_t cap, size_t* total) {
  if (!ctx || (n && (!regions || !alpha || !group_ptr || !total))) {
    set_error("null argument");
    return TBH_EINVAL;
  }
  if (dtype != TBH_F32 && dtype != TBH_I32) {
    set_error("unknown dtype");
    return TBH_EINVAL;
  }
  *total = 0;
  if (n == 0) return 0;
  // descriptors and prefixes
  std::vector<RegionDesc> desc(n);
  std::vector<int64_t> xprefix(n + 1, 0), aprefix(n + 1, 0);
  for (size_t i = 0; i < n; i++) {
    const tbh_region& q = regions[i];
    if (q.nv < 0 || q.nv > kRegionMaxNv || (q.nv > 0 && !q.adj)) {
      set_error("region " + std::to_string(i) + ": needs 0 <= nv <= 24 and adjacency masks");
      return TBH_EINVAL;
    }
    const uint32_t all = q.nv == 32 ? ~0u : ((1u << q.nv) - 1u);
    if (q.open & ~all) {
      set_error("region " + std::to_string(i) + ": open vertex outside the region");
      return TBH_EINVAL;
    }
    RegionDesc& d = desc[i];
    std::memset(&d, 0, sizeof d);
    d.nv = q.nv;
    d.open = q.open;
    d.nopen = __builtin_popcount(q.open);
    if (d.nopen > kRegionMaxOpen) {
      set_error("region " + std::to_string(i) + ": more than 16 open vertices");
      return TBH_EINVAL;
    }
    for (int v = 0; v < q.nv; v++) {
      const uint32_t a = q.adj[v];
      if ((a & ~all) || ((a >> v) & 1u)) {
        set_error("region " + std::to_string(i) + ": adjacency of vertex " + std::to_string(v) +
                  " outside the region or a self loop");
        return TBH_EINVAL;
      }
      for (int u = 0; u < q.nv; u++)
        if (((a >> u) & 1u) != ((q.adj[u] >> v) & 1u)) {
          set_error("region " + std::to_string(i) + ": adjacency is not symmetric");
          return TBH_EINVAL;
        }
      d.adj[v] = a;
      if (!q.weights) {
        if (dtype == TBH_F32) {
          const float one = 1.0f;
          std::memcpy(&d.wbits[v], &one, 4);
        } else {
          d.wbits[v] = 1u;
        }
      } else {
        std::memcpy(&d.wbits[v], static_cast<const char*>(q.weights) + 4 * v, 4);
        if (dtype == TBH_F32) {
          float f;
          std::memcpy(&f, &d.wbits[v], 4);
          // configurations are matched by their weight: sums must be exact
          if (f != std::floor(f) || std::fabs(f) >= (float)(1 << 20)) {
            set_error("region " + std::to_string(i) +
                      ": Float32 table weights must be integers below 2^20");
            return TBH_EINVAL;
          }
        } else {
          const int32_t w = (int32_t)d.wbits[v];
          if (w <= -(1 << 24) || w >= (1 << 24)) {
            set_error("region " + std::to_string(i) + ": Int32 weights must satisfy |w| < 2^24");
            return TBH_EINVAL;
          }
        }
      }
    }
    d.x0 = xprefix[i];
    d.a0 = aprefix[i];
    xprefix[i + 1] = xprefix[i] + ((int64_t)1 << q.nv);
    aprefix[i + 1] = aprefix[i] + ((int64_t)1 << d.nopen);
  }
  const int64_t X = xprefix[n], A = aprefix[n];
  std::lock_guard<std::mutex> g(ctx->mu);
  hipError_t e = hipSetDevice(ctx->device);
  if (e != hipSuccess) {
    set_error(std::string("hipSetDevice: ") + hipGetErrorString(e));
    return TBH_EDEVICE;
  }
  // device layout: descriptors, prefixes, alpha in / out, counts, offsets
  size_t off = 0;
  auto take = [&](size_t bytes) {
    const size_t o = (off + 255) / 256 * 256;
    off = o + bytes;
    return o;
  };
  const size_t o_desc = take(n * sizeof(RegionDesc));
  const size_t o_xp = take((n + 1) * sizeof(int64_t));
  const size_t o_ap = take((n + 1) * sizeof(int64_t));
  const size_t o_in = take(4 * A);
  const size_t o_out = take(4 * A);
  const size_t o_cnt = take(4 * A);
  const size_t o_offs = take(8 * (A + 1));
  const size_t head = off;
  const int64_t doff = ctx->pool.alloc(head);
  if (doff < 0) {
    set_error("pool exhausted (region tables)");
    return TBH_ENOMEM;
  }
  char* dev = ctx->base + doff;
  std::vector<int64_t> offs(A + 1, 0);
  std::vector<uint32_t> cnt(A, 0);
  int64_t cfg_off = -1;
  auto done = [&](int rc) {
    (void)hipStreamSynchronize(ctx->stream);
    ctx->pool.free(doff);
    if (cfg_off >= 0) ctx->pool.free(cfg_off);
    return rc;
  };
  auto fail = [&](hipError_t err, const char* what) {
    set_error(std::string("region tables (") + what + "): " + hipGetErrorString(err));
    return done(TBH_EDEVICE);
  };
  hipStream_t s = ctx->stream;
  const int G = 256;
  if ((e = hipMemcpyAsync(dev + o_desc, desc.data(), n * sizeof(RegionDesc),
                          hipMemcpyHostToDevice, s)) != hipSuccess ||
      (e = hipMemcpyAsync(dev + o_xp, xprefix.data(), (n + 1) * 8, hipMemcpyHostToDevice, s)) !=
          hipSuccess ||
      (e = hipMemcpyAsync(dev + o_ap, aprefix.data(), (n + 1) * 8, hipMemcpyHostToDevice, s)) !=
          hipSuccess ||
      (e = hipMemcpyAsync(dev + o_in, alpha, 4 * A, hipMemcpyHostToDevice, s)) != hipSuccess ||
      (e = hipMemsetAsync(dev + o_cnt, 0, 4 * A, s)) != hipSuccess)
    return fail(e, "upload");
  const RegionDesc* dR = (const RegionDesc*)(dev + o_desc);
  const int64_t* dxp = (const int64_t*)(dev + o_xp);
  const int64_t* dap = (const int64_t*)(dev + o_ap);
  uint32_t* dcnt = (uint32_t*)(dev + o_cnt);
  const int64_t* doffs = (const int64_t*)(dev + o_offs);
  const int ga = (int)((A + G - 1) / G);
  const int64_t gx64 = (X + G - 1) / G;
  if (gx64 > INT32_MAX) {
    set_error("region tables: too many configurations in one call");
    return done(TBH_EINVAL);
  }
  const int gx = (int)gx64;
  if (dtype == TBH_F32) {
    tbh_compactify_kernel<float><<<ga, G, 0, s>>>(dR, dap, (int)n, A, (const float*)(dev + o_in),
                                                  (float*)(dev + o_out));
    if ((e = hipGetLastError()) != hipSuccess) return fail(e, "compactify");
    tbh_configs_kernel<float, 0><<<gx, G, 0, s>>>(dR, dxp, (int)n, X,
                                                  (const float*)(dev + o_out), dcnt, doffs,
                                                  nullptr);
  } else {
    tbh_compactify_kernel<int32_t><<<ga, G, 0, s>>>(dR, dap, (int)n, A,
                                                    (const int32_t*)(dev + o_in),
                                                    (int32_t*)(dev + o_out));
    if ((e = hipGetLastError()) != hipSuccess) return fail(e, "compactify");
    tbh_configs_kernel<int32_t, 0><<<gx, G, 0, s>>>(dR, dxp, (int)n, X,
                                                    (const int32_t*)(dev + o_out), dcnt, doffs,
                                                    nullptr);
  }
  if ((e = hipGetLastError()) != hipSuccess) return fail(e, "count");
  if ((e = hipMemcpyAsync(alpha, dev + o_out, 4 * A, hipMemcpyDeviceToHost, s)) != hipSuccess ||
      (e = hipMemcpyAsync(cnt.data(), dcnt, 4 * A, hipMemcpyDeviceToHost, s)) != hipSuccess ||
      (e = hipStreamSynchronize(s)) != hipSuccess)
    return fail(e, "count readback");
  for (int64_t k = 0; k < A; k++) offs[k + 1] = offs[k] + cnt[k];
  const int64_t T = offs[A];
  *total = (size_t)T;
  for (size_t i = 0; i <= n; i++) group_ptr[i] = (size_t)offs[aprefix[i]];
  if (!configs || (size_t)T > cap) {
    if (configs) {
      set_error("region tables: " + std::to_string(T) + " configurations exceed the capacity " +
                std::to_string(cap));
      return done(TBH_ENOMEM);
    }
    return done(0);
  }
  if (T > 0) {
    cfg_off = ctx->pool.alloc((size_t)T * 4);
    if (cfg_off < 0) {
      set_error("pool exhausted (region table configurations)");
      return done(TBH_ENOMEM);
    }
    uint32_t* dcfg = (uint32_t*)(ctx->base + cfg_off);
    if ((e = hipMemcpyAsync(dev + o_offs, offs.data(), 8 * (A + 1), hipMemcpyHostToDevice, s)) !=
            hipSuccess ||
        (e = hipMemsetAsync(dcnt, 0, 4 * A, s)) != hipSuccess)
      return fail(e, "offsets");
    if (dtype == TBH_F32)
      tbh_configs_kernel<float, 1><<<gx, G, 0, s>>>(dR, dxp, (int)n, X,
                                                    (const float*)(dev + o_out), dcnt, doffs,
                                                    dcfg);
    else
      tbh_configs_kernel<int32_t, 1><<<gx, G, 0, s>>>(dR, dxp, (int)n, X,
                                                      (const int32_t*)(dev + o_out), dcnt, doffs,
                                                      dcfg);
    if ((e = hipGetLastError()) != hipSuccess) return fail(e, "write");
    if ((e = hipMemcpyAsync(configs, dcfg, (size_t)T * 4, hipMemcpyDeviceToHost, s)) !=
            hipSuccess ||
        (e = hipStreamSynchronize(s)) != hipSuccess)
      return fail(e, "configurations readback");
    // each group's slots were taken in arrival order: ascending within groups
    for (int64_t k = 0; k < A; k++) std::sort(configs + offs[k], configs + offs[k + 1]);
  }
  return done(0);
}
