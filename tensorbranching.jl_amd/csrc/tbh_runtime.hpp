// tbh_runtime.hpp -- context, device pool and executors behind the C ABI.
#pragma once
#include <hip/hip_runtime.h>

#include <atomic>
#include <condition_variable>
#include <deque>
#include <functional>
#include <map>
#include <memory>
#include <mutex>
#include <string>
#include <thread>
#include <unordered_map>
#include <vector>

#include "tbh.h"
#include "tbh_internal.hpp"
#include "tbh_plan.hpp"

namespace tbh {

void set_error(const std::string& m);

// launchers (tbh_kernels.hip)
hipError_t launch_ew(int dtype, const EwDesc* ops, const int64_t* prefix, int nops, int64_t total,
                     int grid, int chunk, hipStream_t s);
hipError_t launch_ew_direct(int dtype, const EwDesc* ops, const int64_t* prefix, int nops,
                            int64_t total, int grid, int chunk, hipStream_t s);
hipError_t launch_gemm(int dtype, bool vec, int shape, const GemmDesc* ops, const int64_t* prefix,
                       int nops, int64_t total, int grid, hipStream_t s);
hipError_t launch_leaves(int dtype, const LeafDesc* leaves, int64_t n, hipStream_t s);
hipError_t launch_epilogue(int dtype, const EpiDesc* e, int64_t n, void* out, hipStream_t s);
hipError_t launch_valu_probe(int dtype, void* out, int blocks, int iters, hipStream_t s);
// resident tiled-EW workgroups per CU (LDS and registers), from the runtime
int ew_blocks_per_cu(int dtype);
// the same step with the next tile's sub-tensors prefetched into registers
hipError_t launch_ew_pf(int dtype, const EwDesc* ops, const int64_t* prefix, int nops,
                        int64_t total, int grid, int chunk, hipStream_t s);
int ew_pf_blocks_per_cu(int dtype);
// resident workgroups per CU of one GEMM kernel (dtype, VEC, tile shape)
int gemm_blocks_per_cu(int dtype, bool vec, int shape);

// First-fit allocator over one device allocation (offsets, 256-byte units).
class Pool {
 public:
  void init(size_t bytes);
  int64_t alloc(size_t bytes);  // -1 when it does not fit
  void free(int64_t off);
  size_t capacity() const { return cap_; }
  size_t used() const { return used_; }
  size_t peak() const { return peak_; }
  size_t largest_free() const;
  void reset_peak() { peak_ = used_; }

 private:
  size_t cap_ = 0, used_ = 0, peak_ = 0;
  std::map<int64_t, size_t> free_;   // offset -> size
  std::map<int64_t, size_t> inuse_;  // offset -> size
};

struct Staging {  // pinned host buffer reused for descriptor uploads
  void* host = nullptr;
  size_t cap = 0;
  int ensure(size_t bytes);
  ~Staging();
};

struct ServeState;
struct IssueJob;  // one batch's launch sequence (tbh_runtime.cpp)

// A context's launch thread.  The thread that packs a batch hands its launch
// sequence over and goes on to pack the next one (a launch call blocks while
// the device's queue is deep); one sequence is in hand at a time.
class Issuer {
 public:
  ~Issuer();
  void start(tbh_ctx* ctx);
  void stop();
  void submit(std::unique_ptr<IssueJob> job);
  // Waits until the sequence in hand, if any, is issued; 0, or TBH_EDEVICE
  // (error set) when a launch of it failed -- until reset().
  int wait();
  void reset();

 private:
  tbh_ctx* ctx_ = nullptr;
  std::thread th_;
  std::mutex m_;
  std::condition_variable cv_;
  std::unique_ptr<IssueJob> job_;
  bool pending_ = false, stop_ = false;
  hipError_t err_ = hipSuccess;
  std::string msg_;
};

// Host threads that live as long as the context: slice planning, descriptor
// packing and the serving planner post work to them instead of starting
// threads per call.  The pool holds two threads per planning thread, so a
// batch can be packed while later slices are still being planned.  post()
// queues `tickets` runs of body(k), k = 0..tickets-1, and returns at once;
// wait() blocks until all of them have returned.  Bodies
// never wait for other posted work, so posts from several threads cannot
// deadlock.
class WorkerPool {
 public:
  struct Batch {
    std::function<void(int)> body;
    int total = 0;
    int started = 0, finished = 0;  // tickets that ran the body
    int skipped = 0;                // tickets that found the batch closed
    bool closed = false;            // tickets not started yet skip the body
    std::mutex m;
    std::condition_variable cv;
  };
  void start(int threads);
  void stop();
  ~WorkerPool() { stop(); }
  int size() const { return (int)th_.size(); }
  std::shared_ptr<Batch> post(int tickets, std::function<void(int)> body);
  static void wait(Batch& b);
  // closes b (its tickets not started yet will not run) and waits for the
  // ones running: for posters that do the work themselves as well
  static void close_and_wait(Batch& b);

 private:
  struct Ticket {
    std::shared_ptr<Batch> b;
    int k;
  };
  std::mutex m_;
  std::condition_variable cv_;
  std::deque<Ticket> q_;
  bool stopping_ = false;
  std::vector<std::thread> th_;
};

// HIP-event timing of launches on the context's stream (profiling mode).
struct Profiler {
  struct Rec {
    hipEvent_t a, b;
    int kind, level, nops;
    double bytes, macs;
    std::string tag;
  };
  std::vector<hipEvent_t> pool;
  size_t used = 0;
  std::vector<Rec> recs;
  hipEvent_t get();
  // record around one launch: call begin() before and end() after it
  hipEvent_t begin(hipStream_t s);
  void end(hipStream_t s, hipEvent_t a, int kind, int level, int nops, double bytes, double macs,
           const char* tag = nullptr);
  // after the stream is synchronized: fold the records into stats
  void collect(tbh_stats& st);
  ~Profiler();
};

}  // namespace tbh

struct tbh_ctx {
  int device = 0;
  int num_cus = 256;
  int ew_occ[2] = {4, 4};     // resident tiled-EW workgroups per CU, per element type
  int ew_pf_occ[2] = {4, 4};  // the same for the prefetching kernel
  int gemm_occ[2][2][3] = {};  // resident GEMM workgroups per CU [dtype][vec][shape],
                               // from the occupancy query in tbh_create
  hipStream_t stream = nullptr;
  hipStream_t stream2 = nullptr;  // GEMM launches beside the EW launches of a level
  hipStream_t copy_stream = nullptr;   // batched descriptor uploads, beside the batch in flight
  hipEvent_t up_events[2] = {nullptr, nullptr};  // upload done, per batch parity
  std::vector<hipEvent_t> lvl_events[2];  // cross-stream level barriers, per batch parity
  char* base = nullptr;  // device pool
  tbh::Pool pool;
  tbh::Staging staging, staging2;  // descriptor uploads (ping-pong across batches)
  tbh::Staging results_pin[2];     // per-batch results read back (by batch parity)
  std::mutex mu;         // one executor call at a time
  tbh_stats stats{};
  int plan_threads = 8;
  tbh::WorkerPool workers;  // plan_threads host threads (planning, packing)
  tbh_tuning tune{};   // planner / launch parameters (tbh_set_tuning, tbh_autotune)
  bool profiling = false;
  tbh::Profiler prof;
  // error-path tests: after this many batched launches the next one reports
  // a launch failure instead of launching, once (-1: never;
  // TBH_INJECT_LAUNCH_FAILURE at tbh_create)
  std::atomic<int64_t> inject_launch_failure{-1};
  tbh::Issuer issuer;  // batched launch sequences
  std::mutex serve_mu;   // guards `serve` (never held across an execution)
  std::shared_ptr<tbh::ServeState> serve;  // created by tbh_serve_start
};

namespace tbh {

// Converted input slice with its own storage.
struct SliceOwned {
  SliceIn in;
  std::vector<uint32_t> wbits;
  std::vector<int32_t> ix_ptr, ix, tree;
  uint32_t rbits = 0;
  bool empty_graph = false;  // nv == 0: result is r
};

int convert_slice(const tbh_slice& s, tbh_dtype dtype, bool copy, SliceOwned* out);

// Plans slices on background host threads in index order; get(i) blocks
// until slice i is planned and returns nullptr (error in *err) if planning
// failed at or before i.
class PlanStream {
 public:
  PlanStream(const std::vector<const SliceOwned*>& slices, const PlanOptions& opt,
             WorkerPool& pool, int threads);
  ~PlanStream();
  const SlicePlan* get(size_t i, std::string* err);

 private:
  const std::vector<const SliceOwned*>& slices_;
  std::vector<SlicePlan> plans_;
  std::vector<char> done_;
  std::mutex m_;
  std::condition_variable cv_;
  std::atomic<size_t> next_{0};
  size_t bad_ = (size_t)-1;
  std::string bad_err_;
  std::shared_ptr<WorkerPool::Batch> batch_;
};

// The arena region a batched run takes when `largest` bytes are the
// largest free block of the pool (the rest stays for step descriptors).
size_t region_bytes(size_t largest);

// Runs batches of planned slices through one arena region of the pool, one
// batch in flight behind the one being enqueued.  enqueue() packs a batch
// (its arenas must fit region() together), launches it, then waits for the
// batch enqueued before it, whose results it copies to that batch's `dst`
// and whose `tag` it returns in *done_tag (nullptr: none).  flush() waits
// for the last batch the same way.  After an error the results of the
// batch being enqueued and of the one in flight are lost; close() (also the
// destructor) waits for the streams and releases the region.  The caller
// holds ctx->mu from open() to close().
class BatchRunner {
 public:
  explicit BatchRunner(tbh_ctx* ctx);
  ~BatchRunner();
  int open();
  void close();
  size_t region() const { return region_; }
  // device bytes a batch's descriptors may take: half of what the pool
  // keeps beside the region, so that two batches' descriptors fit at once
  size_t desc_budget() const { return desc_budget_; }
  bool busy() const { return prev_.on; }
  int enqueue(int dtype, const std::vector<const SliceOwned*>& slices,
              const std::vector<const SlicePlan*>& plans, bool add_r, uint32_t* dst, void* tag,
              void** done_tag);
  int flush(void** done_tag);
  // waits until every launch of the newest batch is issued (a failed launch
  // is reported by the next enqueue or flush)
  void wait_issued() { (void)join_issue(); }

 private:
  int finish(void** done_tag);
  int join_issue();
  void sync_all();
  struct Pending {
    bool on = false;
    hipEvent_t done = nullptr;
    int64_t doff = -1;
    int parity = 0;
    size_t nb = 0;
    uint32_t* dst = nullptr;
    void* tag = nullptr;
  };
  tbh_ctx* ctx_;
  int64_t region_off_ = -1;
  size_t region_ = 0;
  size_t desc_budget_ = 0;
  uint64_t arena_base_ = 0;
  int bidx_ = 0;
  hipEvent_t evs_[2] = {nullptr, nullptr};
  Pending prev_;
};

// Upper bounds of a batch's descriptor upload: per slice (its steps,
// leaves, epilogue and result) and per batch of at most `levels` levels
// (the 256-byte alignment of every (kernel kind, level) array).
size_t desc_bytes(const SlicePlan& P);
size_t desc_batch_overhead(int levels);

// Executes the plans of `slices` (plan_of(i) may block until slice i is
// planned, nullptr = planning failed); results[i] (dtype bits) = root + r
// (add_r) or root.  Used by the C ABI and by the serving scheduler.
using PlanOf = std::function<const SlicePlan*(size_t, std::string*)>;
int execute_batched(tbh_ctx* ctx, int dtype, const std::vector<const SliceOwned*>& slices,
                    const PlanOf& plan_of, bool add_r, uint32_t* results);
// scratch: device memory for the slice's arena and descriptors (at least
// plain_device_bytes) instead of the pool; nullptr = the pool
int execute_plain(tbh_ctx* ctx, int dtype, const SliceOwned& slice, const SlicePlan& plan,
                  bool add_r, uint32_t* result, char* scratch = nullptr);
size_t plain_device_bytes(const SliceOwned& slice, const SlicePlan& plan);
int plan_all(tbh_ctx* ctx, const std::vector<const SliceOwned*>& slices, bool batched,
             std::vector<SlicePlan>* plans);
// a tuning (the context's) as planner options
PlanOptions plan_options(const tbh_tuning& t, bool batched);
PlanOptions plan_options(const tbh_ctx* ctx, bool batched);
// default tuning (environment overrides TBH_CHAIN="big,side,reduce",
// TBH_SPLITK="tile_bits,min_k_bits", TBH_CHAIN_COST, TBH_EW_CHUNK,
// TBH_DIRECT_CHUNK, TBH_FIRST_BATCH and the others listed where it is
// defined, read once)
tbh_tuning default_tuning();
// CPUs the process may use (affinity mask and cgroup CPU quota)
int usable_cpus();

}  // namespace tbh
