// tbh_runtime.cpp -- context, device pool, executors and the C ABI of
// libtbh.so (include/tbh.h).
//
// contract_slices (src/dynamic_ob.jl:36-48) becomes: convert and validate the
// slices, plan each one on host threads, then either
//  - PLAIN: per slice, upload its steps, launch them one by one, read the
//    scalar back (the per-slice GPU loop of usecuda=true), or
//  - BATCHED: pack as many slices as the pool holds into one arena, upload
//    every step of the batch once, and run one launch per wave and kernel
//    kind across all slices, then one epilogue launch and one copy back.
#include "tbh_runtime.hpp"

#include <algorithm>
#include <array>
#include <chrono>
#include <cmath>
#include <cstring>
#include <fstream>
#include <system_error>

#include <sched.h>

namespace tbh {

static double ms_since(std::chrono::steady_clock::time_point t0) {
  return std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
}

static thread_local std::string g_err;
void set_error(const std::string& m) { g_err = m; }

#define HIP_TRY(x)                                                          \
  do {                                                                      \
    hipError_t e_ = (x);                                                    \
    if (e_ != hipSuccess) {                                                 \
      set_error(std::string(#x) + ": " + hipGetErrorString(e_));            \
      return TBH_EDEVICE;                                                   \
    }                                                                       \
  } while (0)

// ------------------------------------------------------------ pool -------
void Pool::init(size_t bytes) {
  cap_ = bytes / 256 * 256;
  free_.clear();
  inuse_.clear();
  free_[0] = cap_;
  used_ = peak_ = 0;
}

int64_t Pool::alloc(size_t bytes) {
  bytes = std::max<size_t>(256, (bytes + 255) / 256 * 256);
  for (auto it = free_.begin(); it != free_.end(); ++it) {
    if (it->second >= bytes) {
      int64_t off = it->first;
      size_t rest = it->second - bytes;
      free_.erase(it);
      if (rest) free_[off + (int64_t)bytes] = rest;
      inuse_[off] = bytes;
      used_ += bytes;
      peak_ = std::max(peak_, used_);
      return off;
    }
  }
  return -1;
}

void Pool::free(int64_t off) {
  auto it = inuse_.find(off);
  if (it == inuse_.end()) return;
  size_t sz = it->second;
  inuse_.erase(it);
  used_ -= sz;
  auto nx = free_.lower_bound(off);
  if (nx != free_.end() && nx->first == off + (int64_t)sz) {
    sz += nx->second;
    nx = free_.erase(nx);
  }
  if (nx != free_.begin()) {
    auto pv = std::prev(nx);
    if (pv->first + (int64_t)pv->second == off) {
      pv->second += sz;
      return;
    }
  }
  free_[off] = sz;
}

size_t Pool::largest_free() const {
  size_t m = 0;
  for (auto& kv : free_) m = std::max(m, kv.second);
  return m;
}

int Staging::ensure(size_t bytes) {
  if (bytes <= cap) return 0;
  if (host) (void)hipHostFree(host);
  host = nullptr;
  cap = 0;
  size_t want = std::max<size_t>(bytes, (size_t)1 << 20);
  if (hipHostMalloc(&host, want, hipHostMallocDefault) != hipSuccess) {
    host = nullptr;
    set_error("hipHostMalloc of staging buffer failed");
    return TBH_ENOMEM;
  }
  cap = want;
  return 0;
}

Staging::~Staging() {
  if (host) (void)hipHostFree(host);
}

// ------------------------------------------------------- profiler --------
hipEvent_t Profiler::get() {
  if (used == pool.size()) {
    hipEvent_t e = nullptr;
    if (hipEventCreate(&e) != hipSuccess) return nullptr;
    pool.push_back(e);
  }
  return pool[used++];
}

hipEvent_t Profiler::begin(hipStream_t s) {
  hipEvent_t a = get();
  if (a) (void)hipEventRecord(a, s);
  return a;
}

void Profiler::end(hipStream_t s, hipEvent_t a, int kind, int level, int nops, double bytes,
                   double macs, const char* tag) {
  hipEvent_t b = get();
  if (!a || !b) return;
  (void)hipEventRecord(b, s);
  recs.push_back({a, b, kind, level, nops, bytes, macs, tag ? tag : ""});
}

void Profiler::collect(tbh_stats& st) {
  const char* path = std::getenv("TBH_PROFILE_LEVELS");
  FILE* f = path ? std::fopen(path, "a") : nullptr;
  for (const Rec& r : recs) {
    float ms = 0;
    if (hipEventElapsedTime(&ms, r.a, r.b) != hipSuccess) continue;
    st.kind_ms[r.kind] += ms;
    st.kind_bytes[r.kind] += r.bytes;
    st.kind_macs[r.kind] += r.macs;
    st.kind_launches[r.kind]++;
    if (f)
      std::fprintf(f, "%d %d %d %.6f %.6g %.6g %s\n", r.level, r.kind, r.nops, ms, r.bytes, r.macs,
                   r.tag.c_str());
  }
  if (f) std::fclose(f);
  recs.clear();
  used = 0;
}

Profiler::~Profiler() {
  for (hipEvent_t e : pool) (void)hipEventDestroy(e);
}

// ------------------------------------------------------- conversion ------
static uint32_t f2b(float f) {
  uint32_t u;
  std::memcpy(&u, &f, 4);
  return u;
}

int convert_slice(const tbh_slice& s, tbh_dtype dtype, bool copy, SliceOwned* out) {
  if (dtype != TBH_F32 && dtype != TBH_I32) {
    set_error("unknown dtype");
    return TBH_EINVAL;
  }
  if (s.nv < 0 || s.nleaves < 0) {
    set_error("negative nv or nleaves");
    return TBH_EINVAL;
  }
  SliceOwned& o = *out;
  o.empty_graph = s.nv == 0;
  if (dtype == TBH_F32) {
    o.rbits = f2b((float)s.r);
  } else {
    if (s.r != std::floor(s.r) || std::fabs(s.r) >= (double)(1 << 28)) {
      set_error("Int32 slices need an integer r with |r| < 2^28");
      return TBH_EINVAL;
    }
    o.rbits = (uint32_t)(int32_t)s.r;
  }
  o.in.nv = s.nv;
  o.in.nleaves = s.nv == 0 ? 0 : s.nleaves;
  o.in.r = s.r;
  o.in.unit_wbits = dtype == TBH_F32 ? f2b(1.0f) : 1u;
  o.wbits.clear();
  if (s.weights && s.nv > 0) {
    o.wbits.resize(s.nv);
    std::memcpy(o.wbits.data(), s.weights, 4 * (size_t)s.nv);
    if (dtype == TBH_I32)
      for (int32_t v = 0; v < s.nv; v++) {
        int32_t w = (int32_t)o.wbits[v];
        if (w <= -(1 << 28) || w >= (1 << 28)) {
          set_error("Int32 weights must satisfy |w| < 2^28");
          return TBH_EINVAL;
        }
      }
    o.in.wbits = o.wbits.data();
  } else {
    o.in.wbits = nullptr;
  }
  if (dtype == TBH_I32) {
    // every finite entry of the contraction is a sum over a subset of the
    // vertex weights (each vertex tensor is a leaf once), plus r at the end:
    // sum |w| + |r| bounds them all
    double tot = std::fabs(s.r);
    if (o.in.wbits)
      for (int32_t v = 0; v < s.nv; v++) tot += std::fabs((double)(int32_t)o.wbits[v]);
    else
      tot += s.nv;
    if (tot >= (double)(1 << 28)) {
      set_error("Int32 slices need sum |w| + |r| < 2^28 (a finite value of the contraction "
                "could otherwise leave the range above the tropical zero)");
      return TBH_EINVAL;
    }
  }
  if (o.in.nleaves > 0) {
    if (!s.ix_ptr || !s.ix || (o.in.nleaves > 1 && !s.tree)) {
      set_error("slice without ix_ptr / ix / tree");
      return TBH_EINVAL;
    }
    if (copy) {
      o.ix_ptr.assign(s.ix_ptr, s.ix_ptr + o.in.nleaves + 1);
      o.ix.assign(s.ix, s.ix + o.ix_ptr.back());
      if (o.in.nleaves > 1) o.tree.assign(s.tree, s.tree + 2 * (size_t)(o.in.nleaves - 1));
      o.in.ix_ptr = o.ix_ptr.data();
      o.in.ix = o.ix.data();
      o.in.tree = o.tree.data();
    } else {
      o.in.ix_ptr = s.ix_ptr;
      o.in.ix = s.ix;
      o.in.tree = s.tree;
    }
  }
  return 0;
}

tbh_tuning default_tuning() {
  static const tbh_tuning t = [] {
    tbh_tuning d{};
    const PlanOptions o;
    d.chain_big = o.chain_big;
    d.chain_side = o.chain_side;
    d.chain_reduce = o.chain_reduce;
    d.split_tiles_bits = o.split_tiles_bits;
    d.split_min_k_bits = o.split_min_k_bits;
    d.ew_chunk = 8;
    d.direct_chunk = 4;
    d.ew_per_cu = 4;
    d.direct_max_reduce = o.direct_max_reduce;
    d.concurrent = 1;
    d.first_batch = 0.15;
    d.ew_prefetch = 0;
    d.ew_grid_mult = 16;
    d.gemm_grid_mult = 16;
    d.gemm_tiles = 1;
    d.chain_cost = o.chain_cost;
    if (const char* e = std::getenv("TBH_CHAIN"))
      std::sscanf(e, "%d,%d,%d", &d.chain_big, &d.chain_side, &d.chain_reduce);
    if (const char* e = std::getenv("TBH_SPLITK"))
      std::sscanf(e, "%d,%d", &d.split_tiles_bits, &d.split_min_k_bits);
    if (const char* e = std::getenv("TBH_EW_CHUNK")) d.ew_chunk = std::max(1, std::atoi(e));
    if (const char* e = std::getenv("TBH_DIRECT_CHUNK")) d.direct_chunk = std::max(1, std::atoi(e));
    if (const char* e = std::getenv("TBH_DIRECT_MAX_R"))
      d.direct_max_reduce = std::min(kMaxReduce, std::max(-1, std::atoi(e)));
    if (const char* e = std::getenv("TBH_CONCURRENT")) d.concurrent = std::atoi(e) != 0;
    if (const char* e = std::getenv("TBH_EW_PER_CU"))
      d.ew_per_cu = std::min(8, std::max(1, std::atoi(e)));
    if (const char* e = std::getenv("TBH_FIRST_BATCH"))
      d.first_batch = std::min(1.0, std::max(0.01, std::atof(e)));
    if (const char* e = std::getenv("TBH_EW_PF")) d.ew_prefetch = std::atoi(e) != 0;
    if (const char* e = std::getenv("TBH_EW_GRID_MULT"))
      d.ew_grid_mult = std::min(64, std::max(1, std::atoi(e)));
    if (const char* e = std::getenv("TBH_GEMM_GRID_MULT"))
      d.gemm_grid_mult = std::min(64, std::max(1, std::atoi(e)));
    if (const char* e = std::getenv("TBH_GEMM_TILES"))
      d.gemm_tiles = std::min(2, std::max(0, std::atoi(e)));
    if (const char* e = std::getenv("TBH_CHAIN_COST"))
      d.chain_cost = std::min(4096, std::max(0, std::atoi(e)));
    return d;
  }();
  return t;
}

// CPUs this process may use: the affinity mask, capped by a cgroup CPU quota
// (cgroup v2 cpu.max or v1 cpu.cfs_quota_us / cpu.cfs_period_us).
int usable_cpus() {
  int n = (int)std::max(1u, std::thread::hardware_concurrency());
  cpu_set_t set;
  if (sched_getaffinity(0, sizeof set, &set) == 0) n = std::min(n, std::max(1, CPU_COUNT(&set)));
  auto cap = [&](double quota, double period) {
    if (quota > 0 && period > 0) n = std::min(n, std::max(1, (int)std::ceil(quota / period)));
  };
  {
    std::ifstream f("/sys/fs/cgroup/cpu.max");
    std::string q;
    double period = 0;
    if (f >> q >> period && q != "max") cap(std::atof(q.c_str()), period);
  }
  {
    std::ifstream fq("/sys/fs/cgroup/cpu/cpu.cfs_quota_us"), fp("/sys/fs/cgroup/cpu/cpu.cfs_period_us");
    double quota = 0, period = 0;
    if (fq >> quota && fp >> period) cap(quota, period);
  }
  return n;
}

PlanOptions plan_options(const tbh_tuning& t, bool batched) {
  PlanOptions o;
  o.batched = batched;
  o.chain_big = t.chain_big;
  o.chain_side = t.chain_side;
  o.chain_reduce = t.chain_reduce;
  o.split_tiles_bits = t.split_tiles_bits;
  o.split_min_k_bits = t.split_min_k_bits;
  static const int scatter_bits = [] {
    const char* e = std::getenv("TBH_GEMM_SCATTER");
    return e ? std::atoi(e) : PlanOptions().gemm_scatter_bits;
  }();
  o.gemm_scatter_bits = scatter_bits;
  static const std::array<int, 3> gemm_min = [] {
    const PlanOptions d;
    std::array<int, 3> v = {d.gemm_min_k_mn, d.gemm_min_fill, d.gemm_min_k};
    if (const char* e = std::getenv("TBH_GEMM_MIN"))
      std::sscanf(e, "%d,%d,%d", &v[0], &v[1], &v[2]);
    return v;
  }();
  o.gemm_min_k_mn = gemm_min[0];
  o.gemm_min_fill = gemm_min[1];
  o.gemm_min_k = gemm_min[2];
  o.direct_max_reduce = t.direct_max_reduce;
  o.gemm_tiles = t.gemm_tiles;
  o.chain_cost = t.chain_cost;
  static const bool reuse_order = [] {
    const char* e = std::getenv("TBH_REUSE_ORDER");
    return e ? std::atoi(e) != 0 : PlanOptions().reuse_order;
  }();
  o.reuse_order = reuse_order;
  return o;
}

PlanOptions plan_options(const tbh_ctx* ctx, bool batched) {
  return plan_options(ctx->tune, batched);
}

void WorkerPool::start(int threads) {
  stop();
  stopping_ = false;
  for (int t = 0; t < threads; t++)
    th_.emplace_back([this] {
      for (;;) {
        Ticket tk;
        {
          std::unique_lock<std::mutex> g(m_);
          cv_.wait(g, [&] { return stopping_ || !q_.empty(); });
          if (q_.empty()) return;  // stopping, and nothing left to run
          tk = std::move(q_.front());
          q_.pop_front();
        }
        {
          std::lock_guard<std::mutex> g(tk.b->m);
          if (tk.b->closed) {  // its poster finished the work itself
            tk.b->skipped++;
            tk.b->cv.notify_all();
            continue;
          }
          tk.b->started++;
        }
        tk.b->body(tk.k);
        {
          std::lock_guard<std::mutex> g(tk.b->m);
          tk.b->finished++;
        }
        tk.b->cv.notify_all();
      }
    });
}

void WorkerPool::stop() {
  {
    std::lock_guard<std::mutex> g(m_);
    stopping_ = true;
  }
  cv_.notify_all();
  for (auto& t : th_) t.join();
  th_.clear();
}

std::shared_ptr<WorkerPool::Batch> WorkerPool::post(int tickets, std::function<void(int)> body) {
  auto b = std::make_shared<Batch>();
  b->body = std::move(body);
  b->total = tickets;
  {
    std::lock_guard<std::mutex> g(m_);
    for (int k = 0; k < tickets; k++) q_.push_back({b, k});
  }
  cv_.notify_all();
  return b;
}

void WorkerPool::wait(Batch& b) {
  std::unique_lock<std::mutex> g(b.m);
  b.cv.wait(g, [&] { return b.finished + b.skipped >= b.total; });
}

void WorkerPool::close_and_wait(Batch& b) {
  std::unique_lock<std::mutex> g(b.m);
  b.closed = true;  // tickets that have not started skip the body from now on
  b.cv.wait(g, [&] { return b.finished >= b.started; });
}

int plan_all(tbh_ctx* ctx, const std::vector<const SliceOwned*>& slices, bool batched,
             std::vector<SlicePlan>* plans) {
  plans->assign(slices.size(), SlicePlan());
  const PlanOptions opt = plan_options(ctx, batched);
  std::atomic<size_t> next{0};
  std::atomic<int> bad{-1};
  std::string first_err;
  std::mutex em;
  auto worker = [&]() {
    for (;;) {
      size_t i = next.fetch_add(1);
      if (i >= slices.size() || bad.load() >= 0) return;
      std::string err;
      if (plan_slice(slices[i]->in, opt, &(*plans)[i], &err)) {
        std::lock_guard<std::mutex> g(em);
        if (bad.load() < 0) {
          bad = (int)i;
          first_err = "slice " + std::to_string(i) + ": " + err;
        }
      }
    }
  };
  const int nt = (int)std::min<size_t>(ctx->plan_threads, slices.size());
  if (nt <= 1) {
    worker();
  } else {
    // nt - 1 pool threads and this one
    auto b = ctx->workers.post(nt - 1, [&](int) { worker(); });
    worker();
    WorkerPool::close_and_wait(*b);
  }
  if (bad.load() >= 0) {
    set_error(first_err);
    return TBH_EINVAL;
  }
  return 0;
}

PlanStream::PlanStream(const std::vector<const SliceOwned*>& slices, const PlanOptions& opt,
                       WorkerPool& pool, int threads)
    : slices_(slices), plans_(slices.size()), done_(slices.size(), 0) {
  auto worker = [this, opt](int) {
    for (;;) {
      const size_t i = next_.fetch_add(1);
      if (i >= slices_.size()) return;
      {
        std::lock_guard<std::mutex> g(m_);
        if (i > bad_) return;
      }
      std::string err;
      const int rc = plan_slice(slices_[i]->in, opt, &plans_[i], &err);
      {
        std::lock_guard<std::mutex> g(m_);
        if (rc && i < bad_) {
          bad_ = i;
          bad_err_ = "slice " + std::to_string(i) + ": " + err;
        }
        done_[i] = 1;
      }
      cv_.notify_all();
    }
  };
  const int nt = (int)std::min<size_t>(std::max(1, threads), std::max<size_t>(1, slices.size()));
  if (pool.size() == 0) {
    worker(0);  // no pool threads: plan everything now
  } else {
    batch_ = pool.post(nt, worker);
  }
}

PlanStream::~PlanStream() {
  {
    std::lock_guard<std::mutex> g(m_);
    bad_ = 0;  // stop the workers after their current slice
  }
  if (batch_) WorkerPool::wait(*batch_);
}

const SlicePlan* PlanStream::get(size_t i, std::string* err) {
  std::unique_lock<std::mutex> g(m_);
  cv_.wait(g, [&] { return done_[i] || i >= bad_; });
  if (i >= bad_) {
    if (err) *err = bad_err_;
    return nullptr;
  }
  return &plans_[i];
}

// ---------------------------------------------------- descriptor packs ---
namespace {

struct Pack {
  std::vector<char> buf;
  size_t put(const void* p, size_t n) {
    size_t off = (buf.size() + 255) / 256 * 256;
    buf.resize(off + n);
    if (n) std::memcpy(buf.data() + off, p, n);
    return off;
  }
  size_t reserve(size_t n) {
    size_t off = (buf.size() + 255) / 256 * 256;
    buf.resize(off + n);
    return off;
  }
};

EwDesc make_ew(const POp& op, const std::vector<PTensor>& T, uint64_t base) {
  EwDesc d = op.ew;
  d.a = base + (uint64_t)T[op.a].off;
  d.b = op.b >= 0 ? base + (uint64_t)T[op.b].off : 0;
  d.c = base + (uint64_t)T[op.c].off;
  return d;
}

GemmDesc make_gemm(const POp& op, const std::vector<PTensor>& T, uint64_t base) {
  GemmDesc d;
  std::memset(&d, 0, sizeof d);
  d.a = base + (uint64_t)T[op.a].off;
  d.b = base + (uint64_t)T[op.b].off;
  d.c = base + (uint64_t)T[op.c].off;
  d.M = op.M;
  d.N = op.N;
  d.K = op.K;
  d.batch = op.batch;
  d.lda = op.M;
  d.ldb = op.K;
  d.ldc = op.M;
  d.sA = (int64_t)op.M * op.K;
  d.sB = (int64_t)op.K * op.N;
  d.sC = (int64_t)op.M * op.N;
  d.shape = op.shape;
  d.scatter = op.scatter;
  std::memcpy(d.pc, op.pc, sizeof d.pc);
  std::memcpy(d.ep, op.ep, sizeof d.ep);
  d.ep_n = op.ep_n;
  d.ep_split = op.ep_split;
  d.tiles_m = (op.M + gemm_bm(d.shape) - 1) / gemm_bm(d.shape);
  d.tiles_n = (op.N + gemm_bn(d.shape) - 1) / gemm_bn(d.shape);
  return d;
}

int64_t gemm_tiles(const GemmDesc& d) { return (int64_t)d.tiles_m * d.tiles_n * d.batch; }

// EW workgroups: claims of `chunk` tiles, ew_per_cu resident per CU (no more
// than fit: 5 by the 32 KB of LDS each; 4 measured faster than 5 on the
// 3-regular n=300 slices), times ew_grid_mult.  A grid of 16x the resident
// workgroups measured faster than a persistent one (tiled EW 157 -> 153 ms
// per n=300 solve): workgroups that start as others finish even out claims
// of uneven cost.
int ew_grid(tbh_ctx* ctx, int dtype, int64_t total, int chunk) {
  int64_t blocks = (total + chunk - 1) / chunk;
  const int occ = ctx->tune.ew_prefetch ? ctx->ew_pf_occ[dtype ? 1 : 0] : ctx->ew_occ[dtype ? 1 : 0];
  const int64_t per_cu = (int64_t)std::min(ctx->tune.ew_per_cu, occ) * ctx->tune.ew_grid_mult;
  return (int)std::max<int64_t>(1, std::min<int64_t>(blocks, (int64_t)ctx->num_cus * per_cu));
}

// GEMM workgroups: the resident workgroups per CU (occupancy query) times
// gemm_grid_mult.  16x measured faster than a persistent grid (GEMM 161 ->
// 138 ms per n=300 solve): the tile costs of a level's GEMMs differ (K from
// 2^5 up), and late-starting workgroups balance them.
int gemm_grid(tbh_ctx* ctx, int64_t total, int dtype, bool vec, int shape) {
  static const int forced =
      std::getenv("TBH_GEMM_PER_CU") ? std::atoi(std::getenv("TBH_GEMM_PER_CU")) : 0;
  int per_cu = forced > 0 ? forced : ctx->gemm_occ[dtype ? 1 : 0][vec ? 1 : 0][shape];
  if (per_cu <= 0) per_cu = 2;  // occupancy not queried (tbh_create fills it)
  const int64_t blocks = (int64_t)ctx->num_cus * per_cu * ctx->tune.gemm_grid_mult;
  return (int)std::max<int64_t>(1, std::min<int64_t>(total, blocks));
}

// direct steps: wave tiles, 4 waves per workgroup
int direct_grid(tbh_ctx* ctx, int64_t total, int chunk) {
  const int64_t waves = (total + chunk - 1) / chunk;
  return (int)std::max<int64_t>(1, std::min<int64_t>((waves + 3) / 4, (int64_t)ctx->num_cus * 8));
}

// Waves (direct) and workgroups (tiled) take small chunks of the tile space
// round-robin, which spreads expensive steps evenly; long contiguous ranges
// measured slower (costly steps cluster in a few ranges).  Tiled claims of 8
// tiles measured best with the 16x grid once a workgroup's consecutive tiles
// reuse repeating sub-tensors from LDS (n=300 solve: 4 -> 268 ms median, 8 ->
// 265, 16 -> 279; before that reuse 2 -> 291, 4 -> 280, 8 -> 279), but only
// on launches large enough to give every workgroup of the grid such a claim:
// smaller launches (serving rounds, small solves) were slower with 8 than
// with 4, so a claim above 4 is halved, down to 4, until the launch fills the
// grid with claims.
int direct_chunk(tbh_ctx* ctx, int64_t) { return std::max(1, ctx->tune.direct_chunk); }
int ew_chunk(tbh_ctx* ctx, int64_t total) {
  int c = std::max(1, ctx->tune.ew_chunk);
  const int64_t grid_cap = (int64_t)ctx->num_cus * ctx->tune.ew_per_cu * ctx->tune.ew_grid_mult;
  while (c > 4 && total < (int64_t)c * grid_cap) c >>= 1;
  return c;
}

}  // namespace

// ------------------------------------------------------ BATCHED ----------
namespace {

// launch kinds: 0 direct EW, 1 tiled EW, 2 + s GEMM with tile shape s
constexpr int kKinds = 5;
int op_kind(const POp& op) {
  return op.kind == kGemm ? 2 + op.shape : (op.direct ? 0 : 1);
}
int64_t op_tiles(const POp& op) {
  if (op.kind != kGemm) return op.tiles;
  const int sh = op.shape, bm = gemm_bm(sh), bn = gemm_bn(sh);
  return (int64_t)((op.M + bm - 1) / bm) * ((op.N + bn - 1) / bn) * op.batch;
}
size_t desc_size(int kind) { return kind >= 2 ? sizeof(GemmDesc) : sizeof(EwDesc); }
int stat_kind(int kind) { return kind < 2 ? kind : 2; }  // tbh_stats: direct, tiled, GEMM

}  // namespace

size_t desc_bytes(const SlicePlan& P) {
  if (P.empty) return sizeof(EpiDesc) + 4;
  // a step: its descriptor, its prefix entry and (per-step profiling) its
  // one-step prefix pair
  return P.ops.size() * (std::max(sizeof(EwDesc), sizeof(GemmDesc)) + 3 * sizeof(int64_t)) +
         P.leaves.size() * sizeof(LeafDesc) + sizeof(EpiDesc) + 4;
}

size_t desc_batch_overhead(int levels) {
  const size_t slots = (size_t)kKinds * (size_t)(levels + 1);
  return 256 * (3 + 3 * slots) + sizeof(int64_t) * slots;
}

namespace {

// Runs f(i, t) for i in [0, n) on the caller (slot 0) and up to threads - 1
// pool threads (slots 1..), or on the caller alone for small n; pool threads
// busy elsewhere never hold it up.
template <typename F>
void parallel_for(size_t n, WorkerPool& pool, int threads, F f) {
  const int nt = (int)std::min<size_t>(std::max(1, threads), std::max<size_t>(1, n / 16));
  if (nt <= 1 || pool.size() == 0) {
    for (size_t i = 0; i < n; i++) f(i, 0);
    return;
  }
  std::atomic<size_t> next{0};
  auto b = pool.post(nt - 1, [&](int t) {
    for (size_t i; (i = next.fetch_add(1)) < n;) f(i, t + 1);
  });
  for (size_t i; (i = next.fetch_add(1)) < n;) f(i, 0);
  WorkerPool::close_and_wait(*b);
}

}  // namespace

// The arena region of a batched run when `largest` bytes are the largest free
// block of the pool: that block less a reserve for the step descriptors.
size_t region_bytes(size_t largest) {
  const size_t reserve = std::min<size_t>(
      (size_t)8 << 30, std::max(largest / 8, std::min<size_t>(largest / 4, (size_t)16 << 20)));
  return largest - reserve;
}

// Batches run one after another on the stream through one arena region
// (stream order keeps a batch's kernels behind the previous batch's).  A
// batch's launches are issued by the context's launch thread (a launch call
// blocks while the device's queue is deep), and its descriptors go up on a
// copy stream, so the next batch is packed and uploaded while this one is
// still being issued and run.  A batch is packed in two passes on the host:
// pass 1 (serial, integer work only) counts the steps of every (kernel kind,
// level), gives each slice its first index in each and builds the tile
// prefixes; pass 2 (host threads, one slice at a time) writes every
// descriptor straight into its final place in a pinned staging buffer (two,
// used alternately), which goes to the device in one copy.  The previous
// batch is waited for only after the next one is enqueued.
BatchRunner::BatchRunner(tbh_ctx* ctx) : ctx_(ctx) {}

BatchRunner::~BatchRunner() { close(); }

int BatchRunner::open() {
  if (region_off_ >= 0) return 0;
  const size_t largest = ctx_->pool.largest_free();
  region_ = largest >= 256 ? region_bytes(largest) : 0;
  region_off_ = region_ >= 256 ? ctx_->pool.alloc(region_) : -1;
  if (region_off_ < 0) {
    region_ = 0;
    set_error("pool too small for a slice arena");
    return TBH_ENOMEM;
  }
  arena_base_ = (uint64_t)(ctx_->base + region_off_);
  desc_budget_ = ctx_->pool.largest_free() / 2;
  for (auto& ev : evs_)
    if (!ev && hipEventCreateWithFlags(&ev, hipEventDisableTiming) != hipSuccess) ev = nullptr;
  return 0;
}

void BatchRunner::close() {
  const int issued = ctx_->issuer.wait();
  ctx_->issuer.reset();  // a failed batch is released below
  if (issued && !prev_.on) sync_all();
  if (prev_.on) {
    // an unfinished batch: wait for it without reporting (error paths)
    sync_all();
    ctx_->pool.free(prev_.doff);
    prev_.on = false;
  }
  for (auto& ev : evs_)
    if (ev) (void)hipEventDestroy(ev);
  evs_[0] = evs_[1] = nullptr;
  if (region_off_ >= 0) ctx_->pool.free(region_off_);
  region_off_ = -1;
  region_ = 0;
  desc_budget_ = 0;
  if (ctx_->profiling) ctx_->prof.collect(ctx_->stats);
}

// Waits for both streams (error paths: no kernel may still use the arena
// or the descriptors when they are released).
void BatchRunner::sync_all() {
  (void)hipStreamSynchronize(ctx_->copy_stream);
  (void)hipStreamSynchronize(ctx_->stream);
  if (ctx_->stream2) (void)hipStreamSynchronize(ctx_->stream2);
}

// Waits until every launch of the newest batch is issued; reports a failed
// launch (the batch stays pending, for close() to wait for and release).
int BatchRunner::join_issue() { return ctx_->issuer.wait(); }

// Waits for the batch in flight, releases its descriptors and copies its
// results to their destination.
int BatchRunner::finish(void** done_tag) {
  if (done_tag) *done_tag = nullptr;
  if (!prev_.on) return 0;
  auto t_wait = std::chrono::steady_clock::now();
  hipError_t e = prev_.done ? hipEventSynchronize(prev_.done) : hipStreamSynchronize(ctx_->stream);
  ctx_->stats.host_ms[3] += ms_since(t_wait);
  ctx_->pool.free(prev_.doff);
  prev_.on = false;
  if (done_tag) *done_tag = prev_.tag;
  if (e != hipSuccess) {
    set_error(std::string("batched execution: ") + hipGetErrorString(e));
    return TBH_EDEVICE;
  }
  std::memcpy(prev_.dst, ctx_->results_pin[prev_.parity].host, 4 * prev_.nb);
  return 0;
}

int BatchRunner::flush(void** done_tag) {
  if (done_tag) *done_tag = nullptr;
  const int rc = join_issue();
  return rc ? rc : finish(done_tag);
}

// One batch's launch sequence, issued by the batch's own thread: the wait
// for its descriptor upload, the leaves, one launch per kernel kind and
// level, the epilogue and the copy of the results back, then its done event.
struct IssueJob {
  int dtype = 0;
  bool prof = false, per_step = false;
  int bidx = 0, H = 0;
  size_t nb = 0;
  int64_t nleaves = 0;
  size_t o_leaves = 0, o_epi = 0, o_res = 0;
  char* dbase = nullptr;     // the batch's descriptors on the device
  const char* hb = nullptr;  // ... and in the pinned staging buffer (step tags)
  hipEvent_t up = nullptr, done = nullptr;
  std::vector<int32_t> cnt;
  std::vector<std::vector<int64_t>> pfx;
  std::vector<size_t> o_desc, o_pfx, o_one;
  std::vector<double> lvl_bytes, lvl_macs;
  std::vector<std::vector<std::pair<double, double>>> step_work;
};

namespace {

hipError_t issue_batch(tbh_ctx* ctx, IssueJob& j) {
  hipError_t e = hipSetDevice(ctx->device);
  auto t_issue = std::chrono::steady_clock::now();
  const int H = j.H, dtype = j.dtype;
  const bool prof = j.prof;
  auto slot = [H](int k, int L) { return k * (H + 1) + L; };
  const auto& cnt = j.cnt;
  const auto& pfx = j.pfx;
  char* dbase = j.dbase;
  // a launch failure forced from the host (error-path tests), checked
  // before every launch of the batch; nothing is launched in its place
  auto injected = [ctx] {
    int64_t v = ctx->inject_launch_failure.load();
    while (v >= 0)
      if (ctx->inject_launch_failure.compare_exchange_weak(v, v - 1)) return v == 0;
    return false;
  };
  if (e == hipSuccess) e = hipStreamWaitEvent(ctx->stream, j.up, 0);
  if (e == hipSuccess && j.nleaves && injected()) e = hipErrorLaunchFailure;
  if (e == hipSuccess)
    e = launch_leaves(dtype, (const LeafDesc*)(dbase + j.o_leaves), j.nleaves, ctx->stream);
  ctx->stats.launches += j.nleaves ? 1 : 0;
  hipEvent_t ev = nullptr;
  // one launch per kernel kind and level (direct EW, tiled EW, GEMM); in
  // the per-step profiling mode every step alone, tagged with its shape
  // GEMM launches on a second stream beside a level's EW launches (not
  // while profiling): a level's steps are independent, so the VALU-bound
  // GEMM tiles and the memory-bound EW tiles can share the CUs.  Each
  // stream waits for the other's work of earlier levels (events recorded
  // lazily, only when the other stream launched since the last sync).
  const bool conc = ctx->tune.concurrent != 0 && !prof;
  hipStream_t sk[2] = {ctx->stream, conc ? ctx->stream2 : ctx->stream};
  auto& evpool = ctx->lvl_events[j.bidx & 1];
  size_t evn = 0;
  // dirty[s]: stream s launched work the other stream has not waited for
  // (the upload wait and the leaves are stream-0 work)
  bool dirty[2] = {true, false};
  auto sync_to = [&](int s) {  // stream s waits for the other stream's work so far
    const int o = 1 - s;
    if (!conc || !dirty[o]) return hipSuccess;
    if (evn == evpool.size()) {
      hipEvent_t x;
      hipError_t r = hipEventCreateWithFlags(&x, hipEventDisableTiming);
      if (r != hipSuccess) return r;
      evpool.push_back(x);
    }
    hipEvent_t x = evpool[evn++];
    hipError_t r = hipEventRecord(x, sk[o]);
    if (r == hipSuccess) r = hipStreamWaitEvent(sk[s], x, 0);
    dirty[o] = false;
    return r;
  };
  auto launch = [&](int k, const char* descs, const int64_t* prefix, int nops, int64_t total) {
    if (injected()) return hipErrorLaunchFailure;
    if (k == 0)
      return launch_ew_direct(dtype, reinterpret_cast<const EwDesc*>(descs), prefix, nops, total,
                              direct_grid(ctx, total, direct_chunk(ctx, total)),
                              direct_chunk(ctx, total), sk[0]);
    if (k == 1)
      return (ctx->tune.ew_prefetch ? launch_ew_pf : launch_ew)(
          dtype, reinterpret_cast<const EwDesc*>(descs), prefix, nops, total,
                       ew_grid(ctx, dtype, total, ew_chunk(ctx, total)), ew_chunk(ctx, total),
                       sk[0]);
    return launch_gemm(dtype, true, k - 2, reinterpret_cast<const GemmDesc*>(descs), prefix, nops,
                       total, gemm_grid(ctx, total, dtype, true, k - 2), sk[1]);
  };
  for (int L = 1; L <= H && e == hipSuccess; L++) {
    if (conc) {
      // a stream launching work of this level first waits for the other
      // stream's work of earlier levels (both syncs before any launch)
      bool has[2] = {false, false};
      for (int k = 0; k < kKinds; k++)
        if (cnt[slot(k, L)]) has[k >= 2 ? 1 : 0] = true;
      if (has[0] && e == hipSuccess) e = sync_to(0);
      if (has[1] && e == hipSuccess) e = sync_to(1);
      if (has[0]) dirty[0] = true;
      if (has[1]) dirty[1] = true;
    }
    for (int k = 0; k < kKinds && e == hipSuccess; k++) {
      const int w = slot(k, L);
      if (!cnt[w]) continue;
      const int64_t total = pfx[w].back();
      if (j.per_step) {
        for (int32_t q = 0; q < cnt[w] && e == hipSuccess; q++) {
          char tag_s[160];
          if (k >= 2) {
            const GemmDesc& d = reinterpret_cast<const GemmDesc*>(j.hb + j.o_desc[w])[q];
            std::snprintf(tag_s, sizeof tag_s, "M=%d N=%d K=%d batch=%d", d.M, d.N, d.K, d.batch);
          } else {
            const EwDesc& d = reinterpret_cast<const EwDesc*>(j.hb + j.o_desc[w])[q];
            std::snprintf(tag_s, sizeof tag_s, "nT=%d nO=%d nR=%d nRo=%d nSA=%d nSB=%d op=%d",
                          d.nT, d.nO, d.nR, d.nRo, d.nSA, d.nSB, d.mlay);
          }
          const int64_t tiles = pfx[w][q + 1] - pfx[w][q];
          ev = ctx->prof.begin(ctx->stream);
          e = launch(k, dbase + j.o_desc[w] + q * desc_size(k),
                     reinterpret_cast<const int64_t*>(dbase + j.o_one[w]) + 2 * q, 1, tiles);
          ctx->prof.end(ctx->stream, ev, stat_kind(k), L, 1, j.step_work[w][q].first,
                        j.step_work[w][q].second, tag_s);
          ctx->stats.launches++;
        }
        continue;
      }
      if (prof) ev = ctx->prof.begin(ctx->stream);
      e = launch(k, dbase + j.o_desc[w], reinterpret_cast<const int64_t*>(dbase + j.o_pfx[w]),
                 cnt[w], total);
      if (prof)
        ctx->prof.end(ctx->stream, ev, stat_kind(k), L, cnt[w], j.lvl_bytes[w], j.lvl_macs[w]);
      ctx->stats.launches++;
    }
    ctx->stats.waves++;
  }
  if (e == hipSuccess && conc) e = sync_to(0);  // the last GEMMs before the epilogue
  if (e == hipSuccess && injected()) e = hipErrorLaunchFailure;
  if (e == hipSuccess) {
    e = launch_epilogue(dtype, (const EpiDesc*)(dbase + j.o_epi), (int64_t)j.nb, dbase + j.o_res,
                        ctx->stream);
    ctx->stats.launches++;
  }
  if (e == hipSuccess)
    e = hipMemcpyAsync(ctx->results_pin[j.bidx & 1].host, dbase + j.o_res, 4 * j.nb,
                       hipMemcpyDeviceToHost, ctx->stream);
  if (e == hipSuccess && j.done) e = hipEventRecord(j.done, ctx->stream);
  ctx->stats.host_ms[2] += ms_since(t_issue);
  return e;
}

}  // namespace

Issuer::~Issuer() { stop(); }

void Issuer::start(tbh_ctx* ctx) {
  ctx_ = ctx;
  stop_ = false;
  th_ = std::thread([this] {
    (void)hipSetDevice(ctx_->device);
    for (;;) {
      std::unique_ptr<IssueJob> j;
      {
        std::unique_lock<std::mutex> g(m_);
        cv_.wait(g, [&] { return stop_ || job_; });
        if (!job_) return;  // stopping, nothing in hand
        j = std::move(job_);
      }
      const hipError_t r = issue_batch(ctx_, *j);
      {
        std::lock_guard<std::mutex> g(m_);
        if (r != hipSuccess && err_ == hipSuccess) {
          err_ = r;
          msg_ = std::string("batched launch: ") + hipGetErrorString(r);
        }
        pending_ = false;
      }
      cv_.notify_all();
    }
  });
}

void Issuer::stop() {
  if (!th_.joinable()) return;
  {
    std::lock_guard<std::mutex> g(m_);
    stop_ = true;
  }
  cv_.notify_all();
  th_.join();
}

void Issuer::submit(std::unique_ptr<IssueJob> job) {
  {
    std::unique_lock<std::mutex> g(m_);
    cv_.wait(g, [&] { return !pending_; });  // one sequence in hand at a time
    job_ = std::move(job);
    pending_ = true;
  }
  cv_.notify_all();
}

int Issuer::wait() {
  std::unique_lock<std::mutex> g(m_);
  cv_.wait(g, [&] { return !pending_; });
  if (err_ != hipSuccess) {
    set_error(msg_);
    return TBH_EDEVICE;
  }
  return 0;
}

void Issuer::reset() {
  std::lock_guard<std::mutex> g(m_);
  err_ = hipSuccess;
  msg_.clear();
}

int BatchRunner::enqueue(int dtype, const std::vector<const SliceOwned*>& slices,
                         const std::vector<const SlicePlan*>& plans, bool add_r, uint32_t* dst,
                         void* tag, void** done_tag) {
  if (done_tag) *done_tag = nullptr;
  tbh_ctx* ctx = ctx_;
  const size_t nb = plans.size();
  if (region_off_ < 0) {
    set_error("batched run not open");
    return TBH_EINVAL;
  }
  if (nb == 0) return 0;
  const uint32_t zero_bits = dtype == TBH_F32 ? f2b(0.0f) : 0u;
  const bool per_step = ctx->profiling && std::getenv("TBH_PROFILE_STEPS");
  const bool prof = ctx->profiling;
  const int bidx = bidx_++;
  auto t_pack = std::chrono::steady_clock::now();
  std::vector<uint64_t> sbase(nb);
  int H = 0;
  {
    uint64_t acc = 0;
    for (size_t i = 0; i < nb; i++) {
      sbase[i] = arena_base_ + acc;
      acc += plans[i]->empty ? 0 : (uint64_t)plans[i]->arena_bytes;
      H = std::max(H, plans[i]->nlevels);
    }
    if (acc > region_) {
      set_error("batch arenas (" + std::to_string(acc) + " bytes) exceed the region (" +
                std::to_string(region_) + ")");
      return TBH_ENOMEM;
    }
  }
  const int W = kKinds * (H + 1);  // (kind, level) slots
  auto slot = [H](int k, int L) { return k * (H + 1) + L; };
  // pass 1
  std::vector<int32_t> cnt(W, 0), start(nb * W);
  std::vector<std::vector<int64_t>> pfx(W, std::vector<int64_t>(1, 0));
  std::vector<double> lvl_bytes(W, 0.0), lvl_macs(W, 0.0);
  std::vector<std::vector<std::pair<double, double>>> step_work(per_step ? W : 0);
  std::vector<int64_t> leaf_start(nb + 1, 0);
  for (size_t i = 0; i < nb; i++) {
    const SlicePlan& P = *plans[i];
    std::memcpy(&start[i * W], cnt.data(), W * sizeof(int32_t));
    leaf_start[i + 1] = leaf_start[i] + (P.empty ? 0 : (int64_t)P.leaves.size());
    if (P.empty) continue;
    for (const POp& op : P.ops) {
      const int w = slot(op_kind(op), op.level);
      cnt[w]++;
      pfx[w].push_back(pfx[w].back() + op_tiles(op));
      lvl_bytes[w] += op.bytes;
      lvl_macs[w] += op.macs;
      if (per_step) step_work[w].push_back({op.bytes, op.macs});
      if (op.kind == kEw) {
        ctx->stats.ew_ops++;
        ctx->stats.ew_ops_algo_bytes += op.bytes;
      } else {
        ctx->stats.gemm_ops++;
        ctx->stats.gemm_ops_algo_ops += 2 * op.macs;
      }
    }
    ctx->stats.total_algo_ops += 2 * P.node_macs;
    ctx->stats.leaf_tensors += P.leaves.size();
  }
  const int64_t nleaves = leaf_start[nb];
  // layout of the upload (256-byte aligned arrays)
  size_t off = 0;
  auto take = [&](size_t bytes) {
    const size_t o = (off + 255) / 256 * 256;
    off = o + bytes;
    return o;
  };
  const size_t o_leaves = take((size_t)nleaves * sizeof(LeafDesc));
  const size_t o_epi = take(nb * sizeof(EpiDesc));
  const size_t o_res = take(4 * nb);
  std::vector<size_t> o_desc(W), o_pfx(W), o_one(W);
  for (int w = 0; w < W; w++) {
    if (!cnt[w]) continue;
    o_desc[w] = take((size_t)cnt[w] * desc_size(w / (H + 1)));
    o_pfx[w] = take((size_t)(cnt[w] + 1) * sizeof(int64_t));
    if (per_step) o_one[w] = take((size_t)cnt[w] * 2 * sizeof(int64_t));
  }
  const size_t bytes = off;
  Staging& stg = (bidx & 1) ? ctx->staging2 : ctx->staging;
  int rc = stg.ensure(bytes);
  if (!rc) rc = ctx->results_pin[bidx & 1].ensure(4 * nb);
  if (rc) return rc;
  char* hb = static_cast<char*>(stg.host);
  for (int w = 0; w < W; w++) {
    if (!cnt[w]) continue;
    std::memcpy(hb + o_pfx[w], pfx[w].data(), pfx[w].size() * sizeof(int64_t));
    if (per_step) {
      int64_t* one = reinterpret_cast<int64_t*>(hb + o_one[w]);
      for (int32_t k = 0; k < cnt[w]; k++) {
        one[2 * k] = 0;
        one[2 * k + 1] = pfx[w][k + 1] - pfx[w][k];
      }
    }
  }
  // pass 2
  std::vector<std::vector<int32_t>> local(std::max(1, ctx->plan_threads),
                                          std::vector<int32_t>(W));
  parallel_for(nb, ctx->workers, ctx->plan_threads, [&](size_t i, int t) {
    const SlicePlan& P = *plans[i];
    const uint64_t b = sbase[i];
    EpiDesc e;
    std::memset(&e, 0, sizeof e);
    e.rbits = add_r ? slices[i]->rbits : zero_bits;
    e.root = P.empty ? 0 : b + (uint64_t)P.tensors[P.root].off;
    std::memcpy(hb + o_epi + i * sizeof(EpiDesc), &e, sizeof e);
    if (P.empty) return;
    LeafDesc* ld = reinterpret_cast<LeafDesc*>(hb + o_leaves) + leaf_start[i];
    for (const PLeaf& L : P.leaves) {
      ld->p = b + (uint64_t)P.tensors[L.tensor].off;
      ld->kind = L.kind;
      ld->wbits = (int32_t)L.wbits;
      ld++;
    }
    std::vector<int32_t>& loc = local[t];
    std::fill(loc.begin(), loc.end(), 0);
    const int32_t* st = &start[i * W];
    for (const POp& op : P.ops) {
      const int k = op_kind(op), w = slot(k, op.level);
      const int32_t idx = st[w] + loc[w]++;
      if (k >= 2)
        reinterpret_cast<GemmDesc*>(hb + o_desc[w])[idx] = make_gemm(op, P.tensors, b);
      else
        reinterpret_cast<EwDesc*>(hb + o_desc[w])[idx] = make_ew(op, P.tensors, b);
    }
  });
  ctx->stats.host_ms[1] += ms_since(t_pack);
  // device descriptors: if the reserve cannot hold two batches' worth,
  // wait for the batch in flight first
  int64_t doff = ctx->pool.alloc(bytes);
  void* early_tag = nullptr;
  if (doff < 0) {
    if ((rc = join_issue())) return rc;
    if ((rc = finish(&early_tag))) return rc;
    doff = ctx->pool.alloc(bytes);
  }
  if (doff < 0) {
    set_error("pool exhausted (descriptors)");
    return TBH_ENOMEM;
  }
  // the upload goes on the copy stream, beside the batch in flight (its
  // descriptors are a pool block of their own); the issuing thread makes the
  // compute stream wait for it
  char* dbase = ctx->base + doff;
  const hipEvent_t up = ctx->up_events[bidx & 1];
  hipError_t e = hipMemcpyAsync(dbase, hb, bytes, hipMemcpyHostToDevice, ctx->copy_stream);
  if (e == hipSuccess) e = hipEventRecord(up, ctx->copy_stream);
  if (e != hipSuccess) {
    (void)hipStreamSynchronize(ctx->copy_stream);
    ctx->pool.free(doff);
    set_error(std::string("descriptor upload: ") + hipGetErrorString(e));
    return TBH_EDEVICE;
  }
  // every launch of the previous batch is issued before this batch's first
  if ((rc = join_issue())) {
    (void)hipStreamSynchronize(ctx->copy_stream);
    ctx->pool.free(doff);
    return rc;
  }
  auto job = std::make_unique<IssueJob>();
  job->dtype = dtype;
  job->prof = prof;
  job->per_step = per_step;
  job->bidx = bidx;
  job->H = H;
  job->nb = nb;
  job->nleaves = nleaves;
  job->o_leaves = o_leaves;
  job->o_epi = o_epi;
  job->o_res = o_res;
  job->dbase = dbase;
  job->hb = hb;
  job->up = up;
  job->done = evs_[bidx & 1];
  job->cnt = std::move(cnt);
  job->pfx = std::move(pfx);
  job->o_desc = std::move(o_desc);
  job->o_pfx = std::move(o_pfx);
  job->o_one = std::move(o_one);
  job->lvl_bytes = std::move(lvl_bytes);
  job->lvl_macs = std::move(lvl_macs);
  job->step_work = std::move(step_work);
  ctx->issuer.submit(std::move(job));
  // this batch runs while the previous one is waited for
  void* tag_prev = nullptr;
  if ((rc = finish(&tag_prev))) {
    (void)join_issue();
    sync_all();
    ctx->pool.free(doff);
    return rc;
  }
  if (done_tag) *done_tag = early_tag ? early_tag : tag_prev;
  prev_.on = true;
  prev_.done = evs_[bidx & 1];
  prev_.doff = doff;
  prev_.parity = bidx & 1;
  prev_.nb = nb;
  prev_.dst = dst;
  prev_.tag = tag;
  return 0;
}

int execute_batched(tbh_ctx* ctx, int dtype, const std::vector<const SliceOwned*>& slices,
                    const PlanOf& plan_of, bool add_r, uint32_t* results) {
  const size_t n = slices.size();
  if (n == 0) return 0;
  BatchRunner run(ctx);
  int status = run.open();
  if (status) return status;
  const size_t region = run.region(), dbudget = run.desc_budget();
  std::string err;
  // the first batch is kept to a fraction of the region so that the GPU
  // starts while the later slices are still being planned
  const double first_frac = std::min(1.0, std::max(0.01, ctx->tune.first_batch));
  size_t i0 = 0;
  for (int bidx = 0; i0 < n && status == 0; bidx++) {
    auto t_plan = std::chrono::steady_clock::now();
    // admit slices while their arenas fit the region (waiting for plans)
    const size_t cap = bidx == 0 ? (size_t)(first_frac * (double)region) : region;
    std::vector<const SlicePlan*> plans;
    std::vector<const SliceOwned*> sl;
    size_t i1 = i0, arena = 0, desc = 0;
    int levels = 0;
    while (i1 < n) {
      const SlicePlan* P = plan_of(i1, &err);
      if (!P) {
        set_error(err);
        status = TBH_EINVAL;
        break;
      }
      const size_t a = P->empty ? 0 : (size_t)P->arena_bytes, d = desc_bytes(*P);
      const int lv = std::max(levels, P->nlevels);
      // a batch is bounded by its arenas and by its descriptors
      if (i1 > i0 &&
          (arena + a > cap || desc + d + desc_batch_overhead(lv) > dbudget))
        break;
      if (a > region || d + desc_batch_overhead(lv) > dbudget) {
        set_error(a > region ? "slice arena (" + std::to_string(a) + " bytes) does not fit the pool"
                             : "slice descriptors (" + std::to_string(d) +
                                   " bytes) do not fit the pool");
        status = TBH_ENOMEM;
        break;
      }
      arena += a;
      desc += d;
      levels = lv;
      plans.push_back(P);
      sl.push_back(slices[i1]);
      i1++;
    }
    ctx->stats.host_ms[0] += ms_since(t_plan);
    if (status) break;
    status = run.enqueue(dtype, sl, plans, add_r, results + i0, nullptr, nullptr);
    i0 = i1;
  }
  if (status == 0) status = run.flush(nullptr);
  return status;
}

// -------------------------------------------------------- PLAIN ----------
namespace {

// A PLAIN slice's descriptors, packed for an arena at device address b.
struct PlainPack {
  Pack pk;
  size_t o_leaves = 0, o_epi = 0, o_res = 0;
  std::vector<size_t> o_op, o_pre;
  std::vector<int64_t> tiles;
  int64_t nleaves = 0;
};

void pack_plain(const SliceOwned& slice, const SlicePlan& P, bool add_r, uint32_t zero_bits,
                uint64_t b, PlainPack* out) {
  PlainPack& q = *out;
  std::vector<LeafDesc> leaves;
  leaves.reserve(P.leaves.size());
  for (const PLeaf& L : P.leaves) {
    LeafDesc d;
    d.p = b + (uint64_t)P.tensors[L.tensor].off;
    d.kind = L.kind;
    d.wbits = (int32_t)L.wbits;
    leaves.push_back(d);
  }
  q.nleaves = (int64_t)leaves.size();
  EpiDesc e;
  std::memset(&e, 0, sizeof e);
  e.root = b + (uint64_t)P.tensors[P.root].off;
  e.rbits = add_r ? slice.rbits : zero_bits;
  q.o_leaves = q.pk.put(leaves.data(), leaves.size() * sizeof(LeafDesc));
  q.o_epi = q.pk.put(&e, sizeof e);
  q.o_res = q.pk.reserve(4);
  q.o_op.resize(P.ops.size());
  q.o_pre.resize(P.ops.size());
  q.tiles.resize(P.ops.size());
  for (size_t k = 0; k < P.ops.size(); k++) {
    const POp& op = P.ops[k];
    int64_t pre[2] = {0, 0};
    if (op.kind == kEw) {
      EwDesc d = make_ew(op, P.tensors, b);
      q.o_op[k] = q.pk.put(&d, sizeof d);
      pre[1] = d.tiles;
    } else {
      GemmDesc d = make_gemm(op, P.tensors, b);
      q.o_op[k] = q.pk.put(&d, sizeof d);
      pre[1] = gemm_tiles(d);
    }
    q.tiles[k] = pre[1];
    q.o_pre[k] = q.pk.put(pre, sizeof pre);
  }
}

size_t round256(size_t x) { return (x + 255) / 256 * 256; }

}  // namespace

// Device bytes the PLAIN executor needs for one slice: its arena, then its
// descriptors.
size_t plain_device_bytes(const SliceOwned& slice, const SlicePlan& P) {
  if (P.empty) return 0;
  PlainPack q;
  pack_plain(slice, P, true, 0, 0, &q);
  return round256((size_t)std::max<int64_t>(P.arena_bytes, 256)) + q.pk.buf.size();
}

// scratch == nullptr: the arena and the descriptors come from the context's
// pool and the descriptors go up through its pinned staging buffer.
// Otherwise they live in the caller's device memory at scratch (at least
// plain_device_bytes), the descriptors go up from pageable host memory and
// the result comes back the same way: nothing is allocated (no profiling).
int execute_plain(tbh_ctx* ctx, int dtype, const SliceOwned& slice, const SlicePlan& P,
                  bool add_r, uint32_t* result, char* scratch) {
  const uint32_t zero_bits = dtype == TBH_F32 ? f2b(0.0f) : 0u;
  if (P.empty) {
    *result = add_r ? slice.rbits : zero_bits;
    return 0;
  }
  const bool ext = scratch != nullptr;
  const size_t arena_bytes = round256((size_t)std::max<int64_t>(P.arena_bytes, 256));
  int64_t arena_off = -1, doff = -1;
  char* arena = scratch;
  if (!ext) {
    arena_off = ctx->pool.alloc(arena_bytes);
    if (arena_off < 0) {
      set_error("slice arena does not fit the pool");
      return TBH_ENOMEM;
    }
    arena = ctx->base + arena_off;
  }
  PlainPack q;
  pack_plain(slice, P, add_r, zero_bits, (uint64_t)arena, &q);
  const size_t bytes = q.pk.buf.size();
  char* dbase = nullptr;
  const void* src = q.pk.buf.data();
  if (ext) {
    dbase = scratch + arena_bytes;
  } else {
    doff = ctx->pool.alloc(bytes);
    if (doff < 0) {
      ctx->pool.free(arena_off);
      set_error("pool exhausted (descriptors)");
      return TBH_ENOMEM;
    }
    dbase = ctx->base + doff;
    int rc = ctx->staging.ensure(bytes);
    if (rc) {
      ctx->pool.free(doff);
      ctx->pool.free(arena_off);
      return rc;
    }
    std::memcpy(ctx->staging.host, q.pk.buf.data(), bytes);
    src = ctx->staging.host;
  }
  const bool prof = ctx->profiling && !ext;
  hipError_t err = hipMemcpyAsync(dbase, src, bytes, hipMemcpyHostToDevice, ctx->stream);
  if (err == hipSuccess)
    err = launch_leaves(dtype, (const LeafDesc*)(dbase + q.o_leaves), q.nleaves, ctx->stream);
  ctx->stats.launches++;
  for (size_t k = 0; k < P.ops.size() && err == hipSuccess; k++) {
    const POp& op = P.ops[k];
    const int64_t tiles = q.tiles[k];
    const int64_t* pre = (const int64_t*)(dbase + q.o_pre[k]);
    hipEvent_t ev = prof ? ctx->prof.begin(ctx->stream) : nullptr;
    if (op.kind == kEw) {
      if (op.direct)
        err = launch_ew_direct(dtype, (const EwDesc*)(dbase + q.o_op[k]), pre, 1, tiles,
                               direct_grid(ctx, tiles, direct_chunk(ctx, tiles)),
                               direct_chunk(ctx, tiles), ctx->stream);
      else
        err = (ctx->tune.ew_prefetch ? launch_ew_pf : launch_ew)(
            dtype, (const EwDesc*)(dbase + q.o_op[k]), pre, 1, tiles,
                        ew_grid(ctx, dtype, tiles, ew_chunk(ctx, tiles)), ew_chunk(ctx, tiles),
                        ctx->stream);
      ctx->stats.ew_ops++;
      ctx->stats.ew_ops_algo_bytes += op.bytes;
    } else {
      err = launch_gemm(dtype, true, op.shape, (const GemmDesc*)(dbase + q.o_op[k]),
                        pre, 1, tiles, gemm_grid(ctx, tiles, dtype, true, op.shape),
                        ctx->stream);
      ctx->stats.gemm_ops++;
      ctx->stats.gemm_ops_algo_ops += 2 * op.macs;
    }
    if (ev)
      ctx->prof.end(ctx->stream, ev, op.kind == kGemm ? 2 : (op.direct ? 0 : 1), (int)k + 1, 1,
                    op.bytes, op.macs);
    ctx->stats.launches++;
  }
  ctx->stats.total_algo_ops += 2 * P.node_macs;
  ctx->stats.leaf_tensors += P.leaves.size();
  if (err == hipSuccess)
    err = launch_epilogue(dtype, (const EpiDesc*)(dbase + q.o_epi), 1, dbase + q.o_res,
                          ctx->stream);
  ctx->stats.launches++;
  uint32_t res_local = 0;
  void* res_dst = ext ? (void*)&res_local : ctx->staging.host;
  if (err == hipSuccess)
    err = hipMemcpyAsync(res_dst, dbase + q.o_res, 4, hipMemcpyDeviceToHost, ctx->stream);
  hipError_t es = hipStreamSynchronize(ctx->stream);
  if (prof) ctx->prof.collect(ctx->stats);
  if (!ext) {
    ctx->pool.free(doff);
    ctx->pool.free(arena_off);
  }
  if (err != hipSuccess || es != hipSuccess) {
    set_error(std::string("plain slice: ") + hipGetErrorString(err != hipSuccess ? err : es));
    return TBH_EDEVICE;
  }
  std::memcpy(result, res_dst, 4);
  return 0;
}

}  // namespace tbh

using namespace tbh;

// ------------------------------------------- the no-allocation scratch ---
// Device memory of tbh_contract_slices_noalloc: a file-scope array that the
// code object holds on every GPU the library runs on, from load to process
// end.  In host code `scratch` names a host-side stand-in, never the
// device's copy: its device address is looked up once per device
// (hipGetSymbolAddress, with that device current, in tbh_create) and only
// that address is passed to kernels and copies.
static __device__ __align__(256) char scratch[TBH_NOALLOC_SCRATCH_BYTES];

namespace {
constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;                        // guards the two tables
char* g_scratch_ptr[kMaxDevices] = {};          // device address of `scratch`
std::string g_scratch_err[kMaxDevices];         // why the lookup failed
std::mutex g_scratch_use[kMaxDevices];          // one call per GPU at a time

// Called from tbh_create with `device` current: resolves the device's copy of
// the array once.
void resolve_scratch(int device) {
  if (device < 0 || device >= kMaxDevices) return;
  std::lock_guard<std::mutex> g(g_scratch_mu);
  if (g_scratch_ptr[device]) return;
  void* p = nullptr;
  const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
  if (e != hipSuccess || !p) {
    g_scratch_err[device] = std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e);
    return;
  }
  g_scratch_ptr[device] = static_cast<char*>(p);
}
}  // namespace

// ---------------------------------------------------------------- C ABI --
extern "C" {

const char* tbh_last_error(void) { return g_err.c_str(); }

int tbh_create(int device, size_t pool_bytes, tbh_ctx** out) {
  if (!out) {
    set_error("null output pointer");
    return TBH_EINVAL;
  }
  *out = nullptr;
  int ndev = 0;
  HIP_TRY(hipGetDeviceCount(&ndev));
  if (device < 0 || device >= ndev) {
    set_error("device " + std::to_string(device) + " not present (" + std::to_string(ndev) +
              " devices)");
    return TBH_EINVAL;
  }
  HIP_TRY(hipSetDevice(device));
  auto ctx = std::make_unique<tbh_ctx>();
  ctx->device = device;
  hipDeviceProp_t prop;
  HIP_TRY(hipGetDeviceProperties(&prop, device));
  ctx->num_cus = prop.multiProcessorCount;
  for (int t = 0; t < 2; t++) {
    ctx->ew_occ[t] = ew_blocks_per_cu(t);
    ctx->ew_pf_occ[t] = ew_pf_blocks_per_cu(t);
    for (int v = 0; v < 2; v++)
      for (int s = 0; s < 3; s++) ctx->gemm_occ[t][v][s] = gemm_blocks_per_cu(t, v != 0, s);
  }
  size_t fr = 0, tot = 0;
  HIP_TRY(hipMemGetInfo(&fr, &tot));
  if (pool_bytes == 0) pool_bytes = (size_t)(fr * 0.8);
  if (pool_bytes > fr) {
    set_error("pool of " + std::to_string(pool_bytes) + " bytes exceeds free device memory (" +
              std::to_string(fr) + ")");
    return TBH_ENOMEM;
  }
  auto release = [&] {
    for (hipEvent_t& x : ctx->up_events)
      if (x) (void)hipEventDestroy(x);
    for (hipStream_t s : {ctx->stream, ctx->stream2, ctx->copy_stream})
      if (s) (void)hipStreamDestroy(s);
  };
  if (hipStreamCreateWithFlags(&ctx->stream, hipStreamNonBlocking) != hipSuccess ||
      hipStreamCreateWithFlags(&ctx->stream2, hipStreamNonBlocking) != hipSuccess ||
      hipStreamCreateWithFlags(&ctx->copy_stream, hipStreamNonBlocking) != hipSuccess ||
      hipEventCreateWithFlags(&ctx->up_events[0], hipEventDisableTiming) != hipSuccess ||
      hipEventCreateWithFlags(&ctx->up_events[1], hipEventDisableTiming) != hipSuccess) {
    release();
    set_error("stream or event creation failed");
    return TBH_EDEVICE;
  }
  void* p = nullptr;
  hipError_t e = hipMalloc(&p, pool_bytes);
  if (e != hipSuccess) {
    release();
    set_error(std::string("hipMalloc pool: ") + hipGetErrorString(e));
    return TBH_ENOMEM;
  }
  ctx->base = (char*)p;
  ctx->pool.init(pool_bytes);
  ctx->stats.pool_bytes = ctx->pool.capacity();
  // planning threads leave two of the process's CPUs to the caller and the
  // launch thread (on the MI355X host's 16-CPU share, 14 threads: 256 ms
  // per n=300 solve against 260 ms with 16)
  ctx->plan_threads = std::max(1, std::min(16, usable_cpus() - 2));
  if (const char* e = std::getenv("TBH_PLAN_THREADS"))
    ctx->plan_threads = std::min(64, std::max(1, std::atoi(e)));
  ctx->workers.start(2 * ctx->plan_threads);
  try {
    ctx->issuer.start(ctx.get());
  } catch (const std::system_error&) {
    ctx->workers.stop();
    (void)hipFree(ctx->base);
    release();
    set_error("could not start the launch thread");
    return TBH_ENOMEM;
  }
  if (const char* e = std::getenv("TBH_INJECT_LAUNCH_FAILURE"))
    ctx->inject_launch_failure = std::max<int64_t>(-1, std::atoll(e));
  ctx->tune = default_tuning();
  resolve_scratch(device);  // for tbh_contract_slices_noalloc (device is current)
  *out = ctx.release();
  return 0;
}

int tbh_serve_stop(tbh_ctx* ctx);

int tbh_destroy(tbh_ctx* ctx) {
  if (!ctx) return 0;
  tbh_serve_stop(ctx);
  ctx->issuer.stop();
  (void)hipSetDevice(ctx->device);
  if (ctx->stream) (void)hipStreamSynchronize(ctx->stream);
  if (ctx->stream2) (void)hipStreamSynchronize(ctx->stream2);
  if (ctx->copy_stream) (void)hipStreamSynchronize(ctx->copy_stream);
  if (ctx->base) (void)hipFree(ctx->base);
  for (auto& v : ctx->lvl_events)
    for (hipEvent_t ev : v) (void)hipEventDestroy(ev);
  for (hipEvent_t ev : ctx->up_events)
    if (ev) (void)hipEventDestroy(ev);
  if (ctx->stream) (void)hipStreamDestroy(ctx->stream);
  if (ctx->stream2) (void)hipStreamDestroy(ctx->stream2);
  if (ctx->copy_stream) (void)hipStreamDestroy(ctx->copy_stream);
  delete ctx;
  return 0;
}

static int contract_impl(tbh_ctx* ctx, tbh_dtype dtype, const tbh_slice* slices, size_t n,
                         void* results, tbh_mode mode, bool add_r) {
  if (!ctx || (n && (!slices || !results))) {
    set_error("null argument");
    return TBH_EINVAL;
  }
  if (mode != TBH_PLAIN && mode != TBH_BATCHED) {
    set_error("unknown mode");
    return TBH_EINVAL;
  }
  std::lock_guard<std::mutex> g(ctx->mu);
  HIP_TRY(hipSetDevice(ctx->device));
  auto t_plan = std::chrono::steady_clock::now();
  std::vector<SliceOwned> owned(n);
  std::vector<const SliceOwned*> ptrs(n);
  for (size_t i = 0; i < n; i++) {
    int rc = convert_slice(slices[i], dtype, false, &owned[i]);
    if (rc) {
      set_error("slice " + std::to_string(i) + ": " + tbh_last_error());
      return rc;
    }
    ptrs[i] = &owned[i];
  }
  uint32_t* out = (uint32_t*)results;
  if (mode == TBH_BATCHED) {
    // slices are planned on host threads while earlier batches run
    PlanStream ps(ptrs, plan_options(ctx, true), ctx->workers, ctx->plan_threads);
    ctx->stats.host_ms[0] += ms_since(t_plan);
    return execute_batched(ctx, dtype, ptrs,
                           [&](size_t i, std::string* err) { return ps.get(i, err); }, add_r, out);
  }
  std::vector<SlicePlan> plans;
  int rc = plan_all(ctx, ptrs, false, &plans);
  ctx->stats.host_ms[0] += ms_since(t_plan);
  if (rc) return rc;
  for (size_t i = 0; i < n; i++) {
    rc = execute_plain(ctx, dtype, owned[i], plans[i], add_r, out + i);
    if (rc) return rc;
  }
  return 0;
}

int tbh_contract_slices(tbh_ctx* ctx, tbh_dtype dtype, const tbh_slice* slices, size_t n,
                        void* results, tbh_mode mode) {
  return contract_impl(ctx, dtype, slices, n, results, mode, true);
}

// tbh_contract_slices(..., TBH_PLAIN) with every device byte in the GPU's
// copy of `scratch`: no hipMalloc / hipFree / pinned host allocation, no
// pool.  Every slice is planned and measured before anything runs, so a
// slice too large for the array fails the call before any work.  Nothing in
// the array needs to start at zero: each slice's leaves, step outputs and
// descriptors are written by this call before they are read.
int tbh_contract_slices_noalloc(tbh_ctx* ctx, tbh_dtype dtype, const tbh_slice* slices, size_t n,
                                void* results) {
  if (!ctx || (n && (!slices || !results))) {
    set_error("null argument");
    return TBH_EINVAL;
  }
  std::lock_guard<std::mutex> g(ctx->mu);
  HIP_TRY(hipSetDevice(ctx->device));
  char* base = nullptr;
  {
    std::lock_guard<std::mutex> gs(g_scratch_mu);
    if (ctx->device < kMaxDevices) {
      base = g_scratch_ptr[ctx->device];
      if (!base) set_error(g_scratch_err[ctx->device]);
    } else {
      set_error("device index beyond the no-allocation table");
    }
  }
  if (!base) return TBH_EDEVICE;
  std::vector<SliceOwned> owned(n);
  std::vector<const SliceOwned*> ptrs(n);
  for (size_t i = 0; i < n; i++) {
    int rc = convert_slice(slices[i], dtype, false, &owned[i]);
    if (rc) {
      set_error("slice " + std::to_string(i) + ": " + tbh_last_error());
      return rc;
    }
    ptrs[i] = &owned[i];
  }
  std::vector<SlicePlan> plans;
  int rc = plan_all(ctx, ptrs, false, &plans);
  if (rc) return rc;
  for (size_t i = 0; i < n; i++) {
    const size_t need = plain_device_bytes(owned[i], plans[i]);
    if (need > (size_t)TBH_NOALLOC_SCRATCH_BYTES) {
      set_error("slice " + std::to_string(i) + " needs " + std::to_string(need) +
                " bytes of device memory; tbh_contract_slices_noalloc holds at most " +
                std::to_string((size_t)TBH_NOALLOC_SCRATCH_BYTES) +
                " (64 MiB) per GPU: use tbh_contract_slices");
      return TBH_ENOMEM;
    }
  }
  std::lock_guard<std::mutex> gu(g_scratch_use[ctx->device]);
  uint32_t* out = (uint32_t*)results;
  for (size_t i = 0; i < n; i++) {
    rc = execute_plain(ctx, dtype, owned[i], plans[i], true, out + i, base);
    if (rc) return rc;
  }
  return 0;
}

int tbh_solve_slice(tbh_ctx* ctx, tbh_dtype dtype, const tbh_slice* slice, void* result,
                    tbh_mode mode) {
  return contract_impl(ctx, dtype, slice, 1, result, mode, false);
}

int tbh_contract_max(tbh_ctx* ctx, tbh_dtype dtype, const tbh_slice* slices, size_t n,
                     void* result, tbh_mode mode) {
  if (n == 0) {
    set_error("maximum over zero slices");
    return TBH_EINVAL;
  }
  std::vector<uint32_t> r(n);
  int rc = contract_impl(ctx, dtype, slices, n, r.data(), mode, true);
  if (rc) return rc;
  if (dtype == TBH_F32) {
    float m = -INFINITY;
    for (uint32_t b : r) {
      float f;
      std::memcpy(&f, &b, 4);
      m = std::max(m, f);
    }
    *(float*)result = m;
  } else {
    int32_t m = INT32_MIN;
    for (uint32_t b : r) m = std::max(m, (int32_t)b);
    *(int32_t*)result = m;
  }
  return 0;
}

int tbh_maxplus_gemm(tbh_ctx* ctx, tbh_dtype dtype, int64_t M, int64_t N, int64_t K, int64_t batch,
                     const void* A, int64_t lda, int64_t strideA, const void* B, int64_t ldb,
                     int64_t strideB, void* C, int64_t ldc, int64_t strideC,
                     tbh_ptr_kind ptr_kind) {
  if (!ctx || dtype < 0 || dtype > 1) {
    set_error("bad context or dtype");
    return TBH_EINVAL;
  }
  if (M < 0 || N < 0 || K < 0 || batch < 0 || M > INT32_MAX || N > INT32_MAX || K > INT32_MAX ||
      batch > INT32_MAX) {
    set_error("sizes must be in [0, 2^31)");
    return TBH_EINVAL;
  }
  if (lda < std::max<int64_t>(1, M) || ldb < std::max<int64_t>(1, K) ||
      ldc < std::max<int64_t>(1, M)) {
    set_error("leading dimension smaller than the matrix");
    return TBH_EINVAL;
  }
  if (M == 0 || N == 0 || batch == 0) return 0;
  if (!A || !B || !C) {
    set_error("null matrix pointer");
    return TBH_EINVAL;
  }
  std::lock_guard<std::mutex> g(ctx->mu);
  HIP_TRY(hipSetDevice(ctx->device));
  const int64_t ea = (batch - 1) * strideA + (K > 0 ? (K - 1) * lda + M : 0);
  const int64_t eb = (batch - 1) * strideB + (N - 1) * ldb + K;
  const int64_t ec = (batch - 1) * strideC + (N - 1) * ldc + M;
  if (strideA < 0 || strideB < 0 || strideC < 0) {
    set_error("negative batch stride");
    return TBH_EINVAL;
  }
  std::vector<int64_t> offs;
  char* dA = (char*)A;
  char* dB = (char*)B;
  char* dC = (char*)C;
  auto release = [&]() {
    for (int64_t o : offs) ctx->pool.free(o);
  };
  if (ptr_kind == TBH_HOST) {
    for (auto [pp, ne] : {std::pair<char**, int64_t>{&dA, ea}, {&dB, eb}, {&dC, ec}}) {
      int64_t o = ctx->pool.alloc((size_t)std::max<int64_t>(ne, 1) * 4);
      if (o < 0) {
        release();
        set_error("GEMM operands do not fit the pool");
        return TBH_ENOMEM;
      }
      offs.push_back(o);
      *pp = ctx->base + o;
    }
    if (K > 0) HIP_TRY(hipMemcpyAsync(dA, A, 4 * ea, hipMemcpyHostToDevice, ctx->stream));
    HIP_TRY(hipMemcpyAsync(dB, B, 4 * std::max<int64_t>(eb, 0), hipMemcpyHostToDevice, ctx->stream));
    // C goes back as its whole span; when the span holds entries outside
    // the M x N blocks (ldc > M, or gaps between batch entries) C is
    // uploaded first, so that those entries come back unchanged
    const bool dense_c = ldc == M && (batch == 1 || strideC == N * ldc);
    if (!dense_c) HIP_TRY(hipMemcpyAsync(dC, C, 4 * ec, hipMemcpyHostToDevice, ctx->stream));
  } else if (ptr_kind != TBH_DEVICE) {
    set_error("unknown ptr_kind");
    return TBH_EINVAL;
  }
  GemmDesc d;
  std::memset(&d, 0, sizeof d);
  d.a = (uint64_t)dA;
  d.b = (uint64_t)dB;
  d.c = (uint64_t)dC;
  d.M = (int32_t)M;
  d.N = (int32_t)N;
  d.K = (int32_t)K;
  d.batch = (int32_t)batch;
  d.lda = lda;
  d.ldb = ldb;
  d.ldc = ldc;
  d.sA = strideA;
  d.sB = strideB;
  d.sC = strideC;
  d.shape = gemm_shape(M, N);
  d.scatter = 0;
  d.tiles_m = (int32_t)((M + gemm_bm(d.shape) - 1) / gemm_bm(d.shape));
  d.tiles_n = (int32_t)((N + gemm_bn(d.shape) - 1) / gemm_bn(d.shape));
  const int64_t total = gemm_tiles(d);
  const bool vec = M % 4 == 0 && K % 4 == 0 && lda % 4 == 0 && ldb % 4 == 0 && ldc % 4 == 0 &&
                   strideA % 4 == 0 && strideB % 4 == 0 && strideC % 4 == 0 &&
                   ((uint64_t)dA % 16 == 0) && ((uint64_t)dB % 16 == 0) && ((uint64_t)dC % 16 == 0);
  int64_t doff = ctx->pool.alloc(sizeof d + 64);
  if (doff < 0) {
    release();
    set_error("pool exhausted");
    return TBH_ENOMEM;
  }
  offs.push_back(doff);
  char* dd = ctx->base + doff;
  int64_t pre[2] = {0, total};
  int rc = ctx->staging.ensure(256);
  if (rc) {
    release();
    return rc;
  }
  std::memcpy(ctx->staging.host, &d, sizeof d);
  std::memcpy((char*)ctx->staging.host + sizeof d, pre, sizeof pre);
  hipError_t e = hipMemcpyAsync(dd, ctx->staging.host, sizeof d + sizeof pre,
                                hipMemcpyHostToDevice, ctx->stream);
  hipEvent_t ev = ctx->profiling ? ctx->prof.begin(ctx->stream) : nullptr;
  if (e == hipSuccess)
    e = launch_gemm(dtype, vec, d.shape, (const GemmDesc*)dd, (const int64_t*)(dd + sizeof d), 1,
                    total, gemm_grid(ctx, total, dtype, vec, d.shape), ctx->stream);
  if (ev)
    ctx->prof.end(ctx->stream, ev, 2, 0, 1, 4.0 * (double)batch * (M * K + K * N + M * N),
                  (double)M * N * K * batch);
  ctx->stats.launches++;
  if (e == hipSuccess && ptr_kind == TBH_HOST)
    e = hipMemcpyAsync(C, dC, 4 * ec, hipMemcpyDeviceToHost, ctx->stream);
  hipError_t es = hipStreamSynchronize(ctx->stream);
  if (ctx->profiling) ctx->prof.collect(ctx->stats);
  release();
  if (e != hipSuccess || es != hipSuccess) {
    set_error(std::string("maxplus_gemm: ") + hipGetErrorString(e != hipSuccess ? e : es));
    return TBH_EDEVICE;
  }
  return 0;
}

static int view_slice(const tbh_slice* s, SliceIn* in) {
  if (!s) {
    set_error("null slice");
    return TBH_EINVAL;
  }
  in->nv = s->nv;
  in->nleaves = s->nv == 0 ? 0 : s->nleaves;
  in->ix_ptr = s->ix_ptr;
  in->ix = s->ix;
  in->tree = s->tree;
  in->r = s->r;
  return 0;
}

int tbh_complexity(const tbh_slice* slice, double* tc, double* sc, double* rwc) {
  SliceIn in;
  if (view_slice(slice, &in)) return TBH_EINVAL;
  std::string err;
  if (slice_complexity(in, tc, sc, rwc, &err)) {
    set_error(err);
    return TBH_EINVAL;
  }
  return 0;
}

int tbh_peak_memory(const tbh_slice* slice, double* log2_elems) {
  SliceIn in;
  if (view_slice(slice, &in)) return TBH_EINVAL;
  std::string err;
  double all;
  if (slice_peak_memory(in, log2_elems, &all, &err)) {
    set_error(err);
    return TBH_EINVAL;
  }
  return 0;
}

int tbh_all_memory(const tbh_slice* slice, double* log2_elems) {
  SliceIn in;
  if (view_slice(slice, &in)) return TBH_EINVAL;
  std::string err;
  double peak;
  if (slice_peak_memory(in, &peak, log2_elems, &err)) {
    set_error(err);
    return TBH_EINVAL;
  }
  return 0;
}

int tbh_slice_footprint(const tbh_slice* slice, tbh_dtype dtype, size_t* bytes) {
  SliceOwned o;
  if (!slice || !bytes) {
    set_error("null argument");
    return TBH_EINVAL;
  }
  int rc = convert_slice(*slice, dtype, false, &o);
  if (rc) return rc;
  SlicePlan P;
  PlanOptions opt;
  std::string err;
  if (plan_slice(o.in, opt, &P, &err)) {
    set_error(err);
    return TBH_EINVAL;
  }
  *bytes = P.empty ? 0 : (size_t)P.arena_bytes;
  return 0;
}

int tbh_plan_steps(const tbh_slice* slice, tbh_dtype dtype, tbh_mode mode, tbh_step_info* out,
                   int cap) {
  SliceOwned o;
  if (!slice || (cap > 0 && !out)) {
    set_error("null argument");
    return TBH_EINVAL;
  }
  int rc = convert_slice(*slice, dtype, false, &o);
  if (rc) return rc;
  SlicePlan P;
  const PlanOptions opt = plan_options(default_tuning(), mode == TBH_BATCHED);
  std::string err;
  if (plan_slice(o.in, opt, &P, &err)) {
    set_error(err);
    return TBH_EINVAL;
  }
  int n = (int)P.ops.size();
  for (int i = 0; i < n && i < cap; i++) {
    const POp& op = P.ops[i];
    tbh_step_info& s = out[i];
    std::memset(&s, 0, sizeof s);
    s.kind = op.kind;
    s.level = op.level;
    s.rank_a = P.tensors[op.a].rank;
    s.rank_b = op.b >= 0 ? P.tensors[op.b].rank : -1;
    s.rank_c = P.tensors[op.c].rank;
    s.n_reduced = op.nR;
    s.M = op.M;
    s.N = op.N;
    s.K = op.K;
    s.batch = op.batch;
    // how many of A's / B's lowest positions a tile reads contiguously
    if (op.kind == kEw) {
      for (int k = 0; k < op.ew.nSA && op.ew.sA[k] == k; k++) s.lanes_a++;
      for (int k = 0; k < op.ew.nSB && op.ew.sB[k] == k; k++) s.lanes_b++;
      s.tile_bits = op.ew.nT;
      s.outer_reduced = op.ew.nRo;
      s.direct = op.direct ? 1 : 0;
      s.outer_product = op.ew.mlay;
      for (int k = 0; k < op.ew.nT; k++) {
        s.tile_a_only += op.ew.tA[k] >= 0 && op.ew.tB[k] < 0;
        s.tile_b_only += op.ew.tB[k] >= 0 && op.ew.tA[k] < 0;
      }
    }
    if (op.kind == kGemm) s.gemm_store = op.scatter;
    s.macs = op.macs;
    s.bytes = op.bytes;
  }
  return n;
}

static bool serving_active(tbh_ctx* ctx) {
  std::lock_guard<std::mutex> g(ctx->serve_mu);
  return (bool)ctx->serve;
}

static int check_tuning(const tbh_tuning& t) {
  if (t.chain_big < 1 || t.chain_big > kMaxRank || t.chain_side < 1 || t.chain_reduce < 1 ||
      t.chain_reduce > kMaxReduce || t.split_tiles_bits < 0 || t.split_tiles_bits > 24 ||
      t.split_min_k_bits < 4 || t.split_min_k_bits > 24 || t.ew_chunk < 1 || t.ew_chunk > 64 ||
      t.direct_chunk < 1 || t.direct_chunk > 64 || t.ew_per_cu < 1 || t.ew_per_cu > 8 ||
      t.direct_max_reduce < -1 || t.direct_max_reduce > kMaxReduce || t.concurrent < 0 ||
      t.concurrent > 1 ||
      !(t.first_batch > 0.0 && t.first_batch <= 1.0) || t.ew_prefetch < 0 || t.ew_prefetch > 1 ||
      t.ew_grid_mult < 1 || t.ew_grid_mult > 64 || t.gemm_grid_mult < 1 || t.gemm_grid_mult > 64 ||
      t.gemm_tiles < 0 || t.gemm_tiles > 2 || t.chain_cost < 0 || t.chain_cost > 4096) {
    set_error("tuning parameter out of range");
    return TBH_EINVAL;
  }
  return 0;
}

int tbh_get_tuning(tbh_ctx* ctx, tbh_tuning* out) {
  if (!ctx || !out) {
    set_error("null argument");
    return TBH_EINVAL;
  }
  std::lock_guard<std::mutex> g(ctx->mu);
  *out = ctx->tune;
  return 0;
}

int tbh_set_tuning(tbh_ctx* ctx, const tbh_tuning* t) {
  if (!ctx || !t) {
    set_error("null argument");
    return TBH_EINVAL;
  }
  if (int rc = check_tuning(*t)) return rc;
  if (serving_active(ctx)) {
    set_error("tuning cannot change while serving");
    return TBH_EBUSY;
  }
  std::lock_guard<std::mutex> g(ctx->mu);
  ctx->tune = *t;
  return 0;
}

// True when every finite value of the slices' contractions is an integer
// below 2^24 in magnitude (Int32 always): then every sum is exact and any
// grouping of the tree gives identical bits.
static bool exact_sums(tbh_dtype dtype, const tbh_slice* slices, size_t n) {
  if (dtype != TBH_F32) return true;
  for (size_t i = 0; i < n; i++) {
    const tbh_slice& s = slices[i];
    if (s.r != std::floor(s.r)) return false;
    double tot = std::fabs(s.r);
    if (s.weights && s.nv > 0) {
      const float* w = static_cast<const float*>(s.weights);
      for (int32_t v = 0; v < s.nv; v++) {
        if (w[v] != std::floor(w[v])) return false;
        tot += std::fabs((double)w[v]);
      }
    } else {
      tot += s.nv;
    }
    if (tot >= (double)(1 << 24)) return false;
  }
  return true;
}

// Coordinate descent over the tuning parameters, each over a short ordered
// list of values: from the current value, its neighbours in the list are
// timed (one BATCHED contraction each) and the fastest kept if it beats the
// current by 2%; passes repeat until nothing improves or the trials run out.
int tbh_autotune(tbh_ctx* ctx, tbh_dtype dtype, const tbh_slice* slices, size_t n, int max_trials,
                 tbh_tuning* best, double* best_ms) {
  if (!ctx || (n && !slices) || !best || !best_ms) {
    set_error("null argument");
    return TBH_EINVAL;
  }
  if (n == 0) {
    set_error("autotune needs slices");
    return TBH_EINVAL;
  }
  if (serving_active(ctx)) {
    set_error("autotune cannot run while serving");
    return TBH_EBUSY;
  }
  std::vector<uint32_t> ref(n), res(n);
  tbh_tuning cur;
  {
    std::lock_guard<std::mutex> g(ctx->mu);
    cur = ctx->tune;
  }
  const tbh_tuning start = cur;
  auto run = [&](const tbh_tuning& t, double* ms) {
    {
      std::lock_guard<std::mutex> g(ctx->mu);
      ctx->tune = t;
    }
    auto t0 = std::chrono::steady_clock::now();
    const int rc = contract_impl(ctx, dtype, slices, n, res.data(), TBH_BATCHED, true);
    *ms = ms_since(t0);
    return rc;
  };
  auto restore = [&](const tbh_tuning& t) {
    std::lock_guard<std::mutex> g(ctx->mu);
    ctx->tune = t;
  };
  // a tuning's time is the best of two runs: single runs vary by a few
  // percent, enough to pass the 2% margin below by chance
  auto time2 = [&](const tbh_tuning& t, double* ms) {
    int r = run(t, ms);
    if (r || res != ref) return r;  // the caller reports differing results
    double m2 = 0;
    r = run(t, &m2);
    if (!r) *ms = std::min(*ms, m2);
    return r;
  };
  double cur_ms = 0;
  int rc = run(cur, &cur_ms);  // warm-up; its results are the reference
  if (!rc) {
    ref = res;
    rc = time2(cur, &cur_ms);
  }
  if (rc) {
    restore(start);
    return rc;
  }
  int trials = 2;
  struct Param {
    std::vector<double> values;
    double tbh_tuning::*d;
    int32_t tbh_tuning::*i;
  };
  const std::vector<Param> params = {
      {{15, 16, 17, 18, 19}, nullptr, &tbh_tuning::chain_big},
      {{13, 14, 15, 16, 17, 18}, nullptr, &tbh_tuning::chain_side},
      {{4, 5, 6, 7, 8, 9}, nullptr, &tbh_tuning::chain_reduce},
      {{7, 8, 9, 10, 11}, nullptr, &tbh_tuning::split_tiles_bits},
      {{9, 10, 11, 12}, nullptr, &tbh_tuning::split_min_k_bits},
      {{1, 2, 4, 8}, nullptr, &tbh_tuning::ew_chunk},
      {{1, 2, 4, 8}, nullptr, &tbh_tuning::direct_chunk},
      {{2, 3, 4, 5}, nullptr, &tbh_tuning::ew_per_cu},
      {{-1, 0, 1, 2, 3}, nullptr, &tbh_tuning::direct_max_reduce},
      {{0, 1}, nullptr, &tbh_tuning::concurrent},
      {{0.05, 0.1, 0.15, 0.25, 0.4}, &tbh_tuning::first_batch, nullptr},
      {{0, 1}, nullptr, &tbh_tuning::ew_prefetch},
      {{1, 4, 16, 64}, nullptr, &tbh_tuning::ew_grid_mult},
      {{1, 4, 16, 64}, nullptr, &tbh_tuning::gemm_grid_mult},
      {{0, 1, 2}, nullptr, &tbh_tuning::gemm_tiles},
      {{0, 8, 12, 16, 24}, nullptr, &tbh_tuning::chain_cost}};
  // chain re-association regroups a chain's sums: for Float32 slices whose
  // sums are not all exact (non-integer weights or r, or sums reaching
  // 2^24) a result may move by a rounding step, so the chain parameters are
  // left out of the search and every candidate keeps the caller's grouping
  const bool exact = exact_sums(dtype, slices, n);
  auto searched = [&](const Param& p) {
    return exact || (p.i != &tbh_tuning::chain_big && p.i != &tbh_tuning::chain_side &&
                     p.i != &tbh_tuning::chain_reduce && p.i != &tbh_tuning::chain_cost);
  };
  auto get = [](const tbh_tuning& t, const Param& p) { return p.d ? t.*p.d : (double)(t.*p.i); };
  auto set = [](tbh_tuning& t, const Param& p, double v) {
    if (p.d)
      t.*p.d = v;
    else
      t.*p.i = (int32_t)v;
  };
  for (bool improved = true; improved && trials < max_trials;) {
    improved = false;
    for (const Param& p : params) {
      if (!searched(p)) continue;
      // neighbours of the current value in the list
      const double v0 = get(cur, p);
      size_t at = 0;
      for (size_t k = 0; k < p.values.size(); k++)
        if (std::fabs(p.values[k] - v0) < std::fabs(p.values[at] - v0)) at = k;
      for (int dir : {-1, 1}) {
        const long k = (long)at + dir;
        if (k < 0 || k >= (long)p.values.size() || trials >= max_trials) continue;
        tbh_tuning t = cur;
        set(t, p, p.values[k]);
        if (t.chain_side >= t.chain_big || check_tuning(t)) continue;
        double ms = 0;
        rc = time2(t, &ms);
        trials++;
        if (rc) {
          restore(start);
          return rc;
        }
        if (res != ref) {
          restore(start);
          set_error("autotune: results differ between tunings");
          return TBH_EINVAL;
        }
        if (ms < 0.98 * cur_ms) {
          cur = t;
          cur_ms = ms;
          improved = true;
        }
      }
    }
  }
  restore(cur);
  *best = cur;
  *best_ms = cur_ms;
  return 0;
}

int tbh_valu_peak(tbh_ctx* ctx, tbh_dtype dtype, double* tropical_ops_per_s) {
  if (!ctx || !tropical_ops_per_s) {
    set_error("null argument");
    return TBH_EINVAL;
  }
  if (dtype != TBH_F32 && dtype != TBH_I32) {
    set_error("unknown dtype");
    return TBH_EINVAL;
  }
  std::lock_guard<std::mutex> g(ctx->mu);
  HIP_TRY(hipSetDevice(ctx->device));
  const int blocks = ctx->num_cus * 8, iters = 4096;
  const size_t bytes = (size_t)blocks * 256 * 4;
  const int64_t off = ctx->pool.alloc(bytes);
  if (off < 0) {
    set_error("pool exhausted");
    return TBH_ENOMEM;
  }
  hipEvent_t a = nullptr, b = nullptr;
  hipError_t e = hipEventCreate(&a);
  if (e == hipSuccess) e = hipEventCreate(&b);
  if (e == hipSuccess) e = launch_valu_probe(dtype, ctx->base + off, blocks, 64, ctx->stream);
  if (e == hipSuccess) e = hipEventRecord(a, ctx->stream);
  if (e == hipSuccess) e = launch_valu_probe(dtype, ctx->base + off, blocks, iters, ctx->stream);
  if (e == hipSuccess) e = hipEventRecord(b, ctx->stream);
  if (e == hipSuccess) e = hipEventSynchronize(b);
  float ms = 0;
  if (e == hipSuccess) e = hipEventElapsedTime(&ms, a, b);
  if (a) (void)hipEventDestroy(a);
  if (b) (void)hipEventDestroy(b);
  ctx->pool.free(off);
  if (e != hipSuccess) {
    set_error(std::string("valu probe: ") + hipGetErrorString(e));
    return TBH_EDEVICE;
  }
  // 64 maxima x 2 k values per iteration per thread, 2 ops per MAC
  const double ops = 2.0 * 128.0 * iters * (double)blocks * 256.0;
  *tropical_ops_per_s = ops / (ms * 1e-3);
  return 0;
}

int tbh_get_stats(tbh_ctx* ctx, tbh_stats* out) {
  if (!ctx || !out) {
    set_error("null argument");
    return TBH_EINVAL;
  }
  std::lock_guard<std::mutex> g(ctx->mu);
  *out = ctx->stats;
  out->pool_bytes = ctx->pool.capacity();
  out->pool_peak_bytes = ctx->pool.peak();
  return 0;
}

int tbh_set_profiling(tbh_ctx* ctx, int on) {
  if (!ctx) {
    set_error("null argument");
    return TBH_EINVAL;
  }
  std::lock_guard<std::mutex> g(ctx->mu);
  ctx->profiling = on != 0;
  return 0;
}

int tbh_reset_stats(tbh_ctx* ctx) {
  if (!ctx) {
    set_error("null argument");
    return TBH_EINVAL;
  }
  std::lock_guard<std::mutex> g(ctx->mu);
  ctx->stats = tbh_stats{};
  ctx->pool.reset_peak();
  return 0;
}

}  // extern "C"
