// Host-side check of the planner, built by tests/test_host_sanitize.py with
// AddressSanitizer and UndefinedBehaviorSanitizer (g++, host code only):
// plans every slice of a file in the batched and plain lowerings and under
// tuning variants that reach the planner's other branches (every eligible
// EW step direct; no consumer-ordered GEMM output; no re-association), and
// evaluates the complexity and memory models.  Exits non-zero on a planner
// error; the sanitizers abort on any report.
//
// File format (little-endian int32): n, then per slice nv, nleaves,
// len(ix_ptr), len(ix), len(tree), followed by those three arrays.
#include <cstdio>
#include <string>
#include <vector>

#include "tbh_plan.hpp"

int main(int argc, char** argv) {
  if (argc < 2) return 2;
  FILE* f = std::fopen(argv[1], "rb");
  if (!f) return 2;
  int n = 0;
  if (std::fread(&n, 4, 1, f) != 1) return 2;
  long steps = 0, plans = 0;
  for (int i = 0; i < n; i++) {
    int h[5];
    if (std::fread(h, 4, 5, f) != 5) return 2;
    std::vector<int32_t> ixp(h[2]), ix(h[3]), tree(h[4]);
    if (std::fread(ixp.data(), 4, h[2], f) != (size_t)h[2] ||
        std::fread(ix.data(), 4, h[3], f) != (size_t)h[3] ||
        std::fread(tree.data(), 4, h[4], f) != (size_t)h[4])
      return 2;
    tbh::SliceIn in;
    in.nv = h[0];
    in.nleaves = h[1];
    in.ix_ptr = ixp.data();
    in.ix = ix.data();
    in.tree = tree.empty() ? nullptr : tree.data();
    in.unit_wbits = 0x3f800000u;  // Float32 1
    std::vector<tbh::PlanOptions> variants(9);
    variants[1].batched = false;
    variants[2].direct_max_reduce = tbh::kMaxReduce;
    variants[3].gemm_scatter_bits = -1;
    variants[4].reassociate = false;
    variants[4].split_k = false;
    variants[5].gemm_tiles = 0;  // 128 x 128 GEMM tiles only
    variants[6].gemm_tiles = 2;  // 256 x 64 tiles up to N = 128
    variants[7].chain_cost = 0;  // chain rotations under the rank caps alone
    variants[8].reuse_order = false;  // tiled EW outer labels in C's order
    for (const auto& opt : variants) {
      tbh::SlicePlan P;
      std::string err;
      if (tbh::plan_slice(in, opt, &P, &err)) {
        std::fprintf(stderr, "slice %d: %s\n", i, err.c_str());
        return 1;
      }
      steps += (long)P.ops.size();
      plans++;
    }
    double tc = 0, sc = 0, rwc = 0, peak = 0, all = 0;
    std::string err;
    if (tbh::slice_complexity(in, &tc, &sc, &rwc, &err) ||
        tbh::slice_peak_memory(in, &peak, &all, &err)) {
      std::fprintf(stderr, "slice %d: %s\n", i, err.c_str());
      return 1;
    }
  }
  std::fclose(f);
  std::printf("planned %ld slice variants, %ld steps\n", plans, steps);
  return 0;
}
