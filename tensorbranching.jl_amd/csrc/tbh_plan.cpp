// tbh_plan.cpp -- lowers a slice's binary contraction tree to EW / GEMM steps.
//
// Semantics follow the reference path: leaf tensors of IndependentSet under
// SizeMax (GenericTensorNetworks generate_tensors), intermediate labels by
// parse_eincode's rule (a label is kept iff it occurs outside the subtree;
// uncompress, src/types.jl:75-79), and each node a tropical contraction that
// max-reduces the labels it drops (OMEinsum's binary rule).  The two modes:
//
//  - plain (OMEinsum's lowering, the reference GPU path): per node, permute
//    and reduce the inputs to (a,k,t) / (k,b,t), one batched max-plus GEMM to
//    (a,b,t), permute to the node's label order; steps run one by one.
//  - batched: each node is one EW step reading its inputs in whatever layout
//    they have and writing its output in the layout its consumer wants, or a
//    tiled GEMM step when the node is GEMM-shaped (its inputs are written in
//    GEMM order by their producers); steps are levelled into waves (ALAP) so
//    a wave of many slices is one launch per kernel kind.
//
// Default layout of a tensor: labels ordered by when they are eliminated,
// latest first, so long-lived labels sit at the low (contiguous) positions of
// every tensor and consecutive lanes stay coalesced along the spine.
#include "tbh_plan.hpp"

#include <algorithm>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <map>
#include <numeric>

namespace tbh {

namespace {

int fail(std::string* err, const std::string& m) {
  if (err) *err = m;
  return -1;
}

int validate(const SliceIn& s, std::string* err) {
  if (s.nv < 0 || s.nleaves < 0) return fail(err, "negative nv or nleaves");
  if (s.nleaves == 0) return 0;
  if (!s.ix_ptr || !s.ix) return fail(err, "missing leaf labels");
  if (s.nleaves > 1 && !s.tree) return fail(err, "missing tree");
  if (s.ix_ptr[0] != 0) return fail(err, "ix_ptr[0] must be 0");
  for (int32_t i = 0; i < s.nleaves; i++) {
    int32_t n = s.ix_ptr[i + 1] - s.ix_ptr[i];
    if (n < 1 || n > 2)
      return fail(err, "leaf " + std::to_string(i) + " has rank " + std::to_string(n) +
                           " (MIS leaves have rank 1 or 2)");
    for (int32_t k = s.ix_ptr[i]; k < s.ix_ptr[i + 1]; k++)
      if (s.ix[k] < 0 || s.ix[k] >= s.nv)
        return fail(err, "leaf label " + std::to_string(s.ix[k]) + " out of range");
    if (n == 2 && s.ix[s.ix_ptr[i]] == s.ix[s.ix_ptr[i] + 1])
      return fail(err, "edge leaf with a repeated label");
  }
  if (s.nleaves > 1) {
    int32_t tot = 2 * s.nleaves - 1;
    std::vector<char> used(tot, 0);
    for (int32_t j = 0; j + 1 < s.nleaves; j++)
      for (int q = 0; q < 2; q++) {
        int32_t c = s.tree[2 * j + q];
        if (c < 0 || c >= tot || (c >= s.nleaves && c - s.nleaves >= j) || used[c])
          return fail(err, "malformed tree at node " + std::to_string(j));
        used[c] = 1;
      }
  }
  return 0;
}

// labels of leaves then internal nodes, each in first-appearance order
void compute_labels(const SliceIn& s, std::vector<std::vector<int32_t>>& lab) {
  int32_t nl = s.nleaves, nn = std::max(0, nl - 1);
  lab.assign(nl + nn, {});
  std::vector<std::vector<int32_t>> cnt(nl + nn);
  std::vector<int32_t> total(std::max(1, s.nv), 0), acc(std::max(1, s.nv), 0);
  for (int32_t i = 0; i < nl; i++)
    for (int32_t k = s.ix_ptr[i]; k < s.ix_ptr[i + 1]; k++) {
      lab[i].push_back(s.ix[k]);
      cnt[i].push_back(1);
      total[s.ix[k]]++;
    }
  for (int32_t j = 0; j < nn; j++) {
    int32_t a = s.tree[2 * j], b = s.tree[2 * j + 1], o = nl + j;
    for (int32_t c : {a, b})
      for (size_t k = 0; k < lab[c].size(); k++) acc[lab[c][k]] += cnt[c][k];
    for (int32_t c : {a, b})
      for (size_t k = 0; k < lab[c].size(); k++) {
        int32_t l = lab[c][k];
        if (acc[l] > 0) {
          if (acc[l] < total[l]) {
            lab[o].push_back(l);
            cnt[o].push_back(acc[l]);
          }
          acc[l] = 0;
        }
      }
  }
}

int pos_in(const std::vector<int32_t>& v, int32_t x) {
  for (size_t i = 0; i < v.size(); i++)
    if (v[i] == x) return (int)i;
  return -1;
}

// Sets the positions of L's labels in a label -> position table for the
// lifetime of the object, then clears them.
struct PosMap {
  std::vector<int8_t>& m;
  const std::vector<int32_t>& L;
  PosMap(std::vector<int8_t>& m_, const std::vector<int32_t>& L_) : m(m_), L(L_) {
    for (size_t i = 0; i < L.size(); i++) m[L[i]] = (int8_t)i;
  }
  ~PosMap() {
    for (int32_t l : L) m[l] = -1;
  }
};

struct Builder {
  const SliceIn& s;
  const PlanOptions& opt;
  SlicePlan& P;
  std::vector<std::vector<int32_t>> layout;  // per tensor
  std::string* err;
  // label -> position in the A, B, C of the EW step being built, else -1
  std::vector<int8_t> posA_, posB_, posC_;

  int add_tensor(const std::vector<int32_t>& lay) {
    PTensor t;
    t.rank = (int)lay.size();
    t.bytes = (int64_t)4 << t.rank;
    P.tensors.push_back(t);
    layout.push_back(lay);
    return (int)P.tensors.size() - 1;
  }

  // EW step C = max over dropped labels of A (+ B)
  int ew(int a, int b, int c) {
    const auto &LA = layout[a], &LC = layout[c];
    static const std::vector<int32_t> none;
    const auto& LB = b >= 0 ? layout[b] : none;
    POp op;
    op.kind = kEw;
    op.a = a;
    op.b = b;
    op.c = c;
    op.nP = (int)LC.size();
    if (op.nP > kMaxRank) return fail(err, "tensor rank above 30");
    for (int32_t l : LC)
      if (pos_in(LA, l) < 0 && pos_in(LB, l) < 0)
        return fail(err, "output label missing from inputs");
    // reduced labels, ordered by their position in A then B
    std::vector<int32_t> R;
    R.reserve(LA.size() + LB.size());
    for (int32_t l : LA)
      if (pos_in(LC, l) < 0) R.push_back(l);
    for (int32_t l : LB)
      if (pos_in(LC, l) < 0 && pos_in(LA, l) < 0) R.push_back(l);
    if ((R.size() > 8 && op.nP < 20) || R.size() > (size_t)kMaxReduce) {
      // split-K: keep the last s reduced labels in a temporary (more
      // outputs, shorter loops), then reduce them in a second step
      const int s_keep = R.size() > (size_t)kMaxReduce && op.nP >= 20
                             ? (int)R.size() - kMaxReduce
                             : std::min((int)R.size() - 4, 20 - op.nP);
      if (op.nP + s_keep > 28) return fail(err, "EW step too large to split");
      std::vector<int32_t> tl = LC;
      tl.insert(tl.end(), R.end() - s_keep, R.end());
      const int t = add_tensor(tl);
      if (ew(a, b, t)) return -1;
      return ew(t, -1, c);
    }
    op.nR = (int)R.size();
    // label -> position tables of A, B and C for this step (O(1) lookups;
    // cleared when the step is built)
    const PosMap mA(posA_, LA), mB(posB_, LB), mC(posC_, LC);
    auto inA = [&](int32_t l) { return posA_[l] >= 0; };
    auto inB = [&](int32_t l) { return posB_[l] >= 0; };
    // Ri: the reduced labels held in the sub-tensors (R's first |Ri|, i.e.
    // the lowest A positions); the rest are outer reduced labels
    std::vector<int32_t> Ri = R;
    auto inRi = [&](int32_t l) { return std::find(Ri.begin(), Ri.end(), l) != Ri.end(); };
    std::vector<int32_t> T;
    T.reserve(kTileBits + 1);
    // direct step: the tile is C's lowest min(nP, 10) positions and the
    // operands are read straight from memory; taken when every reduced label
    // fits the sub-tensor budget and each operand is small (cache resident)
    // or aligned: its lowest positions are lane labels (C positions 0..5) or
    // reduced labels, so the loads are coalesced
    {
      int RA = 0, RB = 0;
      for (int32_t l : R) RA += inA(l), RB += inB(l);
      const int nTd = std::min(op.nP, 10);
      int na = 0, nb = 0;
      for (int p = 0; p < nTd; p++) na += inA(LC[p]), nb += inB(LC[p]);
      auto aligned = [&](const std::vector<int32_t>& L) {
        for (int p = 0; p < std::min<int>(5, (int)L.size()); p++) {
          const int pc = posC_[L[p]];
          if (!inRi(L[p]) && !(pc >= 0 && pc < std::min(6, nTd))) return false;
        }
        return true;
      };
      const bool okA = LA.size() <= 13 || aligned(LA);
      const bool okB = b < 0 || LB.size() <= 13 || aligned(LB);
      if (na + RA <= kSubBits && nb + RB <= kSubBits && okA && okB &&
          (int)R.size() <= opt.direct_max_reduce) {
        T.assign(LC.begin(), LC.begin() + nTd);
        op.direct = true;
      }
    }
    if (!op.direct) {
      // tiled: for r_in inner reduced labels, the tile takes C's lowest
      // positions, then enough of A's and B's lowest output labels that
      // every operand is read contiguously, then grows toward 2^12 outputs
      // within the LDS budget (a tile's loads are its only latency hiding)
      auto select = [&](int r_in, std::vector<int32_t>& Tt) {
        Ri.assign(R.begin(), R.begin() + r_in);
        int RA = 0, RB = 0;
        for (int32_t l : Ri) RA += inA(l), RB += inB(l);
        Tt.clear();
        if (RA > kSubBits || RB > kSubBits) return -1;
        int nTA = 0, nTB = 0;
        auto try_add = [&](int32_t l) {
          if (std::find(Tt.begin(), Tt.end(), l) != Tt.end()) return true;
          if ((int)Tt.size() >= kTileBits) return false;
          const int ia = inA(l), ib = inB(l);
          if (nTA + ia + RA > kSubBits || nTB + ib + RB > kSubBits) return false;
          Tt.push_back(l);
          nTA += ia;
          nTB += ib;
          return true;
        };
        auto cover = [&](const std::vector<int32_t>& L, int want) {
          int got = 0;
          for (int32_t l : L) {
            if (got >= want) break;
            if (inRi(l) || (posC_[l] >= 0 && try_add(l)))
              got++;
            else
              break;
          }
        };
        for (int p = 0; p < std::min(5, op.nP); p++) try_add(LC[p]);
        cover(LA, 5);
        cover(LB, 5);
        for (int p = 5; p < op.nP; p++) try_add(LC[p]);
        return RA + nTA + (RB + nTB) * 16;  // sub-tensor bits of A, B
      };
      // modelled cycles per MAC of one workgroup: per chunk of the outer
      // reduced loop a fixed cost (two barriers, load latency) plus staging
      // (4 cycles per staged element per wave), plus 16 cycles per MAC
      // column of a wave; a tile below 256 outputs leaves threads idle
      auto cost = [&](int nt, int r_in, int sa, int sb) {
        const double cnt = std::ldexp(1.0, std::max(0, nt - 8));
        const double chunk = 400.0 + 16.0 * (std::ldexp(1.0, sa) + std::ldexp(1.0, sb)) / 256.0 +
                             16.0 * cnt * std::ldexp(1.0, r_in);
        return chunk / std::ldexp(1.0, nt + r_in);
      };
      int best_r = -1;
      double best_c = 0;
      std::vector<int32_t> Tt;
      Tt.reserve(kTileBits + 1);
      for (int r_in = (int)R.size(); r_in >= 0; r_in--) {
        const int bits = select(r_in, Tt);
        if (bits < 0) continue;
        const double c = cost((int)Tt.size(), r_in, bits % 16, b >= 0 ? bits / 16 : -100);
        if (best_r < 0 || c < best_c * 0.999) {
          best_r = r_in;
          best_c = c;
        }
      }
      if (best_r < 0 || (int)R.size() - best_r > kMaxReduce)
        return fail(err, "EW step: reduced labels exceed a tile");
      select(best_r, T);
    }
    std::sort(T.begin(), T.end(),
              [&](int32_t x, int32_t y) { return posC_[x] < posC_[y]; });
    // a full binary tile: tile bits 0..4 stay C's lowest tile positions (a
    // lane's store covers whole 128-byte lines); bits 8..11 take MA A-only
    // then 4 - MA B-only labels when the tile has them (MA = 2, else 3 or
    // 1), so each thread's 16 outputs are an outer product of 2^MA A values
    // and 2^(4-MA) B values per reduced configuration
    int mlay = 0;
    if (!op.direct && b >= 0 && (int)T.size() == kTileBits) {
      std::vector<int32_t> ao, bo, other;
      for (int k = 5; k < kTileBits; k++) {
        const int32_t l = T[k];
        if (inA(l) && !inB(l))
          ao.push_back(l);
        else if (inB(l) && !inA(l))
          bo.push_back(l);
        else
          other.push_back(l);
      }
      int ma = 0;
      if (ao.size() >= 2 && bo.size() >= 2)
        ma = 2;
      else if (ao.size() >= 3 && bo.size() >= 1)
        ma = 3;
      else if (ao.size() >= 1 && bo.size() >= 3)
        ma = 1;
      if (ma) {
        std::vector<int32_t> mid = other;  // bits 5..7
        mid.insert(mid.end(), ao.begin() + ma, ao.end());
        mid.insert(mid.end(), bo.begin() + (4 - ma), bo.end());
        std::sort(mid.begin(), mid.end(),
                  [&](int32_t x, int32_t y) { return posC_[x] < posC_[y]; });
        T.resize(5);
        T.insert(T.end(), mid.begin(), mid.end());
        T.insert(T.end(), ao.begin(), ao.begin() + ma);
        T.insert(T.end(), bo.begin(), bo.begin() + (4 - ma));
        mlay = ma;
      }
    }
    std::vector<int32_t> O;
    O.reserve(LC.size());
    for (int32_t l : LC)
      if (std::find(T.begin(), T.end(), l) == T.end()) O.push_back(l);
    std::vector<int32_t> SA, SB;  // sub-tensor labels in operand order
    SA.reserve(kSubBits + 1);
    SB.reserve(kSubBits + 1);
    for (int32_t l : LA)
      if (inRi(l) || std::find(T.begin(), T.end(), l) != T.end()) SA.push_back(l);
    for (int32_t l : LB)
      if (inRi(l) || std::find(T.begin(), T.end(), l) != T.end()) SB.push_back(l);
    // a workgroup's consecutive tiles take consecutive outer configurations,
    // and the tiled kernel does not stage a sub-tensor again whose address
    // repeats: the outer labels absent from the larger sub-tensor come first
    // (lowest outer bits), so that it repeats most often
    if (!op.direct && b >= 0 && opt.reuse_order) {
      const bool bigA = SA.size() >= SB.size();
      std::stable_partition(O.begin(), O.end(),
                            [&](int32_t l) { return bigA ? !inA(l) : !inB(l); });
    }
    std::vector<int32_t> Ro(R.begin() + Ri.size(), R.end());  // outer reduced
    EwDesc& d = op.ew;
    std::memset(&d, -1, sizeof d);
    d.a = d.b = d.c = 0;
    d.nT = (int)T.size();
    d.nO = (int)O.size();
    d.nR = (int)Ri.size();
    d.nRo = (int)Ro.size();
    d.mlay = mlay;
    d.nSA = (int)SA.size();
    d.nSB = (int)SB.size();
    d.tiles = 1 << d.nO;
    for (int i = 0; i < d.nRo; i++) {
      d.qA[i] = (int8_t)posA_[Ro[i]];
      d.qB[i] = (int8_t)posB_[Ro[i]];
    }
    for (int k = 0; k < d.nT; k++) {
      d.tC[k] = (int8_t)posC_[T[k]];
      d.tA[k] = (int8_t)pos_in(SA, T[k]);
      d.tB[k] = (int8_t)pos_in(SB, T[k]);
    }
    for (int i = 0; i < d.nR; i++) {
      d.rA[i] = (int8_t)pos_in(SA, Ri[i]);
      d.rB[i] = (int8_t)pos_in(SB, Ri[i]);
    }
    for (int k = 0; k < d.nSA; k++) d.sA[k] = (int8_t)posA_[SA[k]];
    for (int k = 0; k < d.nSB; k++) d.sB[k] = (int8_t)posB_[SB[k]];
    for (int k = 0; k < d.nO; k++) {
      d.oC[k] = (int8_t)posC_[O[k]];
      d.oA[k] = (int8_t)posA_[O[k]];
      d.oB[k] = (int8_t)posB_[O[k]];
    }
    op.tiles = d.tiles;
    op.macs = std::ldexp(1.0, op.nP + op.nR);
    op.bytes = 4.0 * (std::ldexp(1.0, (int)LA.size()) + (b >= 0 ? std::ldexp(1.0, (int)LB.size()) : 0) +
                      std::ldexp(1.0, op.nP));
    P.ops.push_back(op);
    P.n_ew++;
    P.ew_bytes += op.bytes;
    return 0;
  }

  // GEMM step: A in (a,k,t), B in (k,b,t), C in (a,b,t)
  int gemm(int a, int b, int c, int na, int nb, int nk, int nt) {
    POp op;
    op.kind = kGemm;
    op.a = a;
    op.b = b;
    op.c = c;
    op.M = 1 << na;
    op.N = 1 << nb;
    op.K = 1 << nk;
    op.batch = 1 << nt;
    op.shape = gemm_shape(op.M, op.N, opt.gemm_tiles);
    op.macs = std::ldexp(1.0, na + nb + nk + nt);
    op.bytes = 4.0 * (std::ldexp(1.0, na + nk + nt) + std::ldexp(1.0, nk + nb + nt) +
                      std::ldexp(1.0, na + nb + nt));
    P.ops.push_back(op);
    P.n_gemm++;
    P.gemm_macs += op.macs;
    return 0;
  }
};

// A node runs as a tiled GEMM when M, N >= 64 and K >= 16 (gemm_min_k_mn), or when its
// outputs fill at least a quarter of a 128 x 128 register-blocked tile and K
// is long enough (>= 64) to amortise the partly idle tile; smaller outputs
// run as EW steps (a partly filled tile still computes the whole tile).
bool gemm_shaped(const PlanOptions& opt, int na, int nb, int nk) {
  const int64_t M = (int64_t)1 << std::min(na, 30), N = (int64_t)1 << std::min(nb, 30),
                K = (int64_t)1 << std::min(nk, 30);
  if (M >= 64 && N >= 64 && K >= opt.gemm_min_k_mn) return true;
  return std::min<int64_t>(M, 128) * std::min<int64_t>(N, 128) >= opt.gemm_min_fill &&
         K >= opt.gemm_min_k;
}

// How a GEMM writing the labels G (m bits, n bits, batch bits) stores into a
// tensor of layout L (GemmDesc::scatter), or 0 when a permute step is
// needed: every one of L's lowest gemm_scatter_bits positions must be a tile
// bit of one epilogue pass, so each tile writes whole runs of that many
// elements.  1 when L's positions 0 and 1 are rows 0 and 1 (the tile's own
// row quads are 16-byte runs of C), else 2 (staged through LDS, ep the
// pass's tile bits in ascending position of L).
int gemm_scatter_mode(const PlanOptions& opt, const std::vector<int32_t>& L,
                      const std::vector<int32_t>& G, int na, int nb, POp* g) {
  if (opt.gemm_scatter_bits < 0 || L.size() != G.size() || (int)G.size() > kGemmScatter - 8 ||
      na < 2)
    return 0;
  const int shape = gemm_shape((int64_t)1 << na, (int64_t)1 << nb, opt.gemm_tiles);
  const int bm = __builtin_ctz(gemm_bm(shape)), bn = __builtin_ctz(gemm_bn(shape));
  const int vm = std::min(na, bm), vn = std::min(nb, bn);
  const int split = vm + vn > 13 ? 1 : 0;  // an LDS pass holds kEpiElems = 2^13 outputs
  // the pass's tile bits: rows below vm, columns below vn (less the top one)
  std::vector<std::pair<int, int>> bits;  // (position in L, id)
  for (int b = 0; b < vm; b++) bits.push_back({pos_in(L, G[b]), b});
  for (int b = 0; b < vn - split; b++) bits.push_back({pos_in(L, G[na + b]), 32 + b});
  std::sort(bits.begin(), bits.end());
  // (positions 0 and 1 at least: the LDS-staged epilogue stores quads)
  const int low = std::min<int>(std::max(opt.gemm_scatter_bits, 2), (int)L.size());
  for (int p = 0; p < low; p++)
    if (p >= (int)bits.size() || bits[p].first != p) return 0;
  if ((int)bits.size() < 10) return 0;
  g->ep_n = (int)bits.size();
  g->ep_split = split;
  for (size_t k = 0; k < bits.size(); k++) g->ep[k] = (int8_t)bits[k].second;
  // rows 0 and 1 at positions 0 and 1: each thread's row quads are 16-byte
  // runs of C, stored directly (L2 completes the lines of a tile)
  return bits[0].second == 0 && bits[1].second == 1 ? 1 : 2;
}

// Split-K for a GEMM node (batched plans): a node with few 128 x 128 tiles
// and a long K loop keeps its s highest K labels as batch labels of a
// temporary, reduced afterwards by an EW step, so that its tiles spread over
// the CUs.  Each part keeps K >= 2^opt.split_min_k_bits.
int gemm_split(const PlanOptions& opt, int na, int nb, int nk, int nt) {
  if (!opt.split_k) return 0;
  // log2 of the node's tile count
  const int tile_bits = (std::max(na, 7) - 7) + (std::max(nb, 7) - 7) + nt;
  int s = 0;
  while (tile_bits + s < opt.split_tiles_bits && nk - s - 1 >= opt.split_min_k_bits &&
         na + nb + nt + s + 1 <= opt.split_max_rank)
    s++;
  return s;
}

// Interval reuse of arena blocks: blocks freed by tensors that died at a level
// strictly below a tensor's birth may be reused by it (exact size classes).
int64_t assign_offsets(std::vector<PTensor>& T) {
  std::vector<int> idx(T.size());
  std::iota(idx.begin(), idx.end(), 0);
  std::stable_sort(idx.begin(), idx.end(), [&](int x, int y) { return T[x].birth < T[y].birth; });
  std::vector<int> by_death(idx);
  std::stable_sort(by_death.begin(), by_death.end(),
                   [&](int x, int y) { return T[x].death < T[y].death; });
  std::map<int64_t, std::vector<int64_t>> freel;  // class bytes -> offsets
  int64_t top = 0;
  size_t di = 0;
  for (int t : idx) {
    while (di < by_death.size() && T[by_death[di]].death < T[t].birth) {
      const PTensor& d = T[by_death[di]];
      if (d.off >= 0) freel[std::max<int64_t>(d.bytes, 16)].push_back(d.off);
      di++;
    }
    int64_t cls = std::max<int64_t>(T[t].bytes, 16);
    auto it = freel.find(cls);
    if (it != freel.end() && !it->second.empty()) {
      T[t].off = it->second.back();
      it->second.pop_back();
    } else {
      int64_t al = std::min<int64_t>(cls, 256);
      top = (top + al - 1) / al * al;
      T[t].off = top;
      top += cls;
    }
  }
  return (top + 255) / 256 * 256;
}

// Every address a step can form stays inside its tensors and the arena:
// positions are distinct and below the operand's rank, GEMM extents match the
// tensor sizes, tiles cover exactly the output, and tensors fit the arena.
int verify_plan(const SlicePlan& P, std::string* err) {
  auto bad = [&](const std::string& m) { return fail(err, "plan check: " + m); };
  for (const auto& t : P.tensors) {
    if (t.off < 0 || t.off % 16 || t.off + std::max<int64_t>(t.bytes, 16) > P.arena_bytes)
      return bad("tensor outside the arena");
    if (t.rank < 0 || t.rank > kMaxRank) return bad("rank");
  }
  for (const auto& op : P.ops) {
    const int ra = P.tensors[op.a].rank;
    const int rb = op.b >= 0 ? P.tensors[op.b].rank : 0;
    const int rc = P.tensors[op.c].rank;
    if (op.kind == kEw) {
      const EwDesc& d = op.ew;
      if (op.nP != rc || d.nT + d.nO != rc || d.nT < 0 || d.nT > kTileBits || d.nO < 0 ||
          d.nR + d.nRo != op.nR || d.nR < 0 || d.nR > kMaxReduce || d.nRo < 0 ||
          d.nRo > kMaxReduce || d.nSA < 0 || d.nSA > kSubBits || d.nSB < 0 ||
          d.nSB > kSubBits || (op.b < 0 && d.nSB != 0) || d.tiles != (1 << d.nO) ||
          op.tiles != d.tiles || (op.direct && (d.nT > 10 || d.nRo != 0)))
        return bad("EW shape");
      // every position of A, B, C is covered exactly once, every sub-tensor
      // bit by exactly one tile or reduced bit, and all indices in range
      auto take = [&](int8_t p, int r, uint64_t& used) {
        if (p < 0) return true;
        if (p >= r || ((used >> p) & 1)) return false;
        used |= 1ull << p;
        return true;
      };
      uint64_t uc = 0, ua = 0, ub = 0, usa = 0, usb = 0;
      for (int k = 0; k < d.nT; k++)
        if (d.tC[k] < 0 || !take(d.tC[k], rc, uc) || !take(d.tA[k], d.nSA, usa) ||
            !take(d.tB[k], d.nSB, usb))
          return bad("EW tile bits");
      for (int k = d.nT; k < kTileBits; k++)
        if (d.tC[k] != -1 || d.tA[k] != -1 || d.tB[k] != -1) return bad("EW tile padding");
      if (d.mlay != 0) {
        if (d.mlay < 1 || d.mlay > 3 || op.direct || op.b < 0 || d.nT != kTileBits)
          return bad("EW outer-product tile layout");
        for (int k = 8; k < kTileBits; k++) {
          const bool a_only = d.tA[k] >= 0 && d.tB[k] < 0, b_only = d.tB[k] >= 0 && d.tA[k] < 0;
          if (k < 8 + d.mlay ? !a_only : !b_only) return bad("EW outer-product tile layout");
        }
      }
      for (int i = 0; i < kMaxReduce; i++) {
        if (i >= d.nR) {
          if (d.rA[i] != -1 || d.rB[i] != -1) return bad("EW reduced padding");
          continue;
        }
        if (!take(d.rA[i], d.nSA, usa) || !take(d.rB[i], d.nSB, usb)) return bad("EW reduced bits");
      }
      for (int i = 0; i < kMaxReduce; i++) {
        if (i >= d.nRo) {
          if (d.qA[i] != -1 || d.qB[i] != -1) return bad("EW outer reduced padding");
          continue;
        }
        if ((d.qA[i] < 0 && d.qB[i] < 0) || !take(d.qA[i], ra, ua) || !take(d.qB[i], rb, ub))
          return bad("EW outer reduced bits");
      }
      for (int k = 0; k < kSubBits; k++) {
        if (k >= d.nSA && d.sA[k] != -1) return bad("EW sub-tensor padding");
        if (k >= d.nSB && d.sB[k] != -1) return bad("EW sub-tensor padding");
        if (k < d.nSA && !take(d.sA[k], ra, ua)) return bad("EW sub-tensor A");
        if (k < d.nSB && !take(d.sB[k], rb, ub)) return bad("EW sub-tensor B");
      }
      for (int k = 0; k < 32; k++) {
        if (k >= d.nO) {
          if (d.oC[k] != -1 || d.oA[k] != -1 || d.oB[k] != -1) return bad("EW outer padding");
          continue;
        }
        if (d.oC[k] < 0 || !take(d.oC[k], rc, uc) || !take(d.oA[k], ra, ua) ||
            !take(d.oB[k], rb, ub))
          return bad("EW outer bits");
      }
      auto full = [](int r) { return r >= 64 ? ~0ull : ((1ull << r) - 1); };
      if (uc != full(rc) || ua != full(ra) || (op.b >= 0 && ub != full(rb)) ||
          usa != full(d.nSA) || usb != full(d.nSB))
        return bad("EW step leaves a position uncovered");
    } else {
      const int64_t ea = (int64_t)op.M * op.K * op.batch, eb = (int64_t)op.K * op.N * op.batch,
                    ec = (int64_t)op.M * op.N * op.batch;
      if (op.b < 0 || ea != ((int64_t)1 << ra) || eb != ((int64_t)1 << rb) ||
          ec != ((int64_t)1 << rc))
        return bad("GEMM extents");
      if (op.M % 4 || op.K % 4) return bad("GEMM vector alignment");
      if (op.shape < 0 || op.shape > 2) return bad("GEMM tile shape");
      if (op.scatter) {
        // pc: a bijection of the output bits onto C's positions, -1 past
        // them; scatter 1: rows 0 and 1 at positions 0 and 1 (quad stores);
        // scatter 2: ep lists each valid tile bit of a pass once, at most
        // 2^13 outputs per pass
        if (op.scatter < 1 || op.scatter > 2) return bad("GEMM scatter mode");
        const int na = __builtin_ctz((uint32_t)op.M), nb = __builtin_ctz((uint32_t)op.N);
        if (op.M != (1 << na) || op.N != (1 << nb)) return bad("GEMM scatter extents");
        if (op.scatter == 1 && (op.pc[0] != 0 || op.pc[1] != 1))
          return bad("GEMM output map: rows 0, 1");
        if (op.scatter == 2) {
          const int shape = op.shape;
          const int vm = std::min(na, __builtin_ctz(gemm_bm(shape)));
          const int vn = std::min(nb, __builtin_ctz(gemm_bn(shape)));
          if (vm < 2 || op.ep_split != (vm + vn > 13 ? 1 : 0) || op.ep_n != vm + vn - op.ep_split ||
              op.ep_n < 10 || op.ep_n > 13)
            return bad("GEMM epilogue passes");
          uint64_t seen = 0;
          for (int k = 0; k < op.ep_n; k++) {
            const int b = op.ep[k];
            const bool ok = (b >= 0 && b < vm) || (b >= 32 && b - 32 < vn - op.ep_split);
            if (!ok || ((seen >> b) & 1)) return bad("GEMM epilogue bits");
            seen |= 1ull << b;
            // the first two at C positions 0 and 1 (quad stores)
            const int q = b < 32 ? b : na + b - 32;
            if (k < 2 && op.pc[q] != k) return bad("GEMM epilogue quads");
          }
        }
        uint64_t used = 0;
        for (int q = 0; q < kGemmScatter; q++) {
          const int p = op.pc[q];
          if (q >= rc) {
            if (p != -1) return bad("GEMM output map past the rank");
          } else if (p < 0 || p >= rc || ((used >> p) & 1)) {
            return bad("GEMM output map");
          } else {
            used |= 1ull << p;
          }
        }
      }
    }
  }
  return 0;
}

// Re-association of spine chains (batched mode only).  A node x = (Y, S)
// whose big child Y = (Y', S') also absorbs a small side tensor is rotated to
// x = (Y', F) with F = (S', S): the small side tensors are combined first and
// the big spine tensor Y' is read once instead of materialising Y.  F keeps
// the labels of S' U S that occur in Y' or outside x, so x's own labels are
// unchanged.  Accepted while F stays small and x reduces few labels.  For
// integer-valued tensors the result is the same bits (every sum is exact, so
// the contraction is order-invariant), which the tests check against the
// unmodified tree; non-integer Float32 sums regrouped this way may differ by
// a rounding step.
void reassociate(int32_t nl, int32_t nv, std::vector<int32_t>& tree,
                 const std::vector<std::vector<int32_t>>& lab0, const PlanOptions& opt) {
  const int32_t nn = nl - 1;
  if (nn < 3) return;
  const int W = (nv + 63) / 64;
  const int tot = nl + nn;
  std::vector<uint64_t> L((size_t)tot * W, 0);
  auto Lp = [&](int x) { return &L[(size_t)x * W]; };
  for (int x = 0; x < tot; x++)
    for (int32_t l : lab0[x]) Lp(x)[l >> 6] |= 1ull << (l & 63);
  auto rank = [&](int x) {
    int c = 0;
    for (int i = 0; i < W; i++) c += __builtin_popcountll(Lp(x)[i]);
    return c;
  };
  std::vector<int32_t> lc(tot, -1), rc(tot, -1), rk(tot, 0);
  for (int x = 0; x < tot; x++) rk[x] = rank(x);
  for (int32_t j = 0; j < nn; j++) {
    lc[nl + j] = tree[2 * j];
    rc[nl + j] = tree[2 * j + 1];
  }
  std::vector<uint64_t> F(W);
  // a combined side tensor must never count as big (F <= side < big), so each
  // rotation moves one node down the spine and the loop ends
  const int side = std::min(opt.chain_side, opt.chain_big - 1);
  for (int32_t j = 0; j < nn; j++) {
    const int x = nl + j;
    for (int guard = 0; guard < nn; guard++) {
      int Y = rk[lc[x]] >= rk[rc[x]] ? lc[x] : rc[x];
      int S = Y == lc[x] ? rc[x] : lc[x];
      if (Y < nl || rk[Y] < opt.chain_big || rk[S] > side) break;
      int Yp = rk[lc[Y]] >= rk[rc[Y]] ? lc[Y] : rc[Y];
      int Sp = Yp == lc[Y] ? rc[Y] : lc[Y];
      if (rk[Yp] < opt.chain_big || rk[Sp] > side) break;
      int rf = 0, nred = 0, u_ys = 0, u_x = 0, u_f = 0, u_new = 0;
      for (int i = 0; i < W; i++) {
        F[i] = (Lp(Sp)[i] | Lp(S)[i]) & (Lp(Yp)[i] | Lp(x)[i]);
        rf += __builtin_popcountll(F[i]);
        nred += __builtin_popcountll((Lp(Yp)[i] | F[i]) & ~Lp(x)[i]);
        u_ys += __builtin_popcountll(Lp(Yp)[i] | Lp(Sp)[i]);
        u_x += __builtin_popcountll(Lp(Y)[i] | Lp(S)[i]);
        u_f += __builtin_popcountll(Lp(Sp)[i] | Lp(S)[i]);
        u_new += __builtin_popcountll(Lp(Yp)[i] | F[i]);
      }
      if (rf > side) break;
      if (opt.chain_cost > 0) {
        // old: Y = Y' (x) S' then x = Y (x) S; new: F = S' (x) S then
        // x = Y' (x) F.  Reads of Y', S', S and the write of x are common;
        // the old pair writes and reads Y, the new one F.
        // (a weight per step kind -- EW MACs dearer than GEMM MACs -- chose
        // worse rotations on the 3-regular n=300 slices than one weight)
        const double w = opt.chain_cost / 64.0;
        const double c_old = 8 * std::ldexp(1.0, rk[Y]) +
                             w * (std::ldexp(1.0, u_ys) + std::ldexp(1.0, u_x));
        const double c_new = 8 * std::ldexp(1.0, rf) +
                             w * (std::ldexp(1.0, u_f) + std::ldexp(1.0, u_new));
        if (c_new >= c_old) break;
      } else if (nred > opt.chain_reduce) {
        break;
      }
      // node Y becomes F = (S', S); x = (Y', F)
      lc[Y] = Sp;
      rc[Y] = S;
      std::copy(F.begin(), F.end(), Lp(Y));
      rk[Y] = rf;
      lc[x] = Yp;
      rc[x] = Y;
    }
  }
  // serialise in post-order again
  std::vector<int32_t> id(tot, -1);
  for (int i = 0; i < nl; i++) id[i] = i;
  std::vector<int32_t> out;
  out.reserve(2 * nn);
  int32_t next = nl;
  std::vector<std::pair<int32_t, int>> st;
  st.emplace_back(nl + nn - 1, 0);
  while (!st.empty()) {
    auto [x, state] = st.back();
    st.pop_back();
    if (x < nl) continue;
    if (state == 0) {
      st.emplace_back(x, 1);
      st.emplace_back(rc[x], 0);
      st.emplace_back(lc[x], 0);
    } else {
      id[x] = next++;
      out.push_back(id[lc[x]]);
      out.push_back(id[rc[x]]);
    }
  }
  tree.swap(out);
}

}  // namespace

int slice_labels(const SliceIn& s, std::vector<std::vector<int32_t>>* labels, std::string* err) {
  if (validate(s, err)) return -1;
  compute_labels(s, *labels);
  return 0;
}

int slice_complexity(const SliceIn& s, double* tc, double* sc, double* rwc, std::string* err) {
  if (validate(s, err)) return -1;
  if (s.nleaves < 2) {
    *tc = -INFINITY;
    *sc = s.nleaves == 1 ? s.ix_ptr[1] - s.ix_ptr[0] : 0;
    *rwc = -INFINITY;
    return 0;
  }
  std::vector<std::vector<int32_t>> lab;
  compute_labels(s, lab);
  double t = 0, rw = 0;
  size_t scm = 0;
  std::vector<char> mark(std::max(1, s.nv), 0);
  for (int32_t j = 0; j + 1 < s.nleaves; j++) {
    const auto &A = lab[s.tree[2 * j]], &B = lab[s.tree[2 * j + 1]], &C = lab[s.nleaves + j];
    int u = 0;
    for (int32_t l : A) mark[l] = 1, u++;
    for (int32_t l : B)
      if (!mark[l]) u++;
    for (int32_t l : A) mark[l] = 0;
    t += std::ldexp(1.0, u);
    rw += std::ldexp(1.0, (int)A.size()) + std::ldexp(1.0, (int)B.size()) + std::ldexp(1.0, (int)C.size());
    scm = std::max(scm, C.size());
  }
  *tc = std::log2(t);
  *sc = (double)scm;
  *rwc = std::log2(rw);
  return 0;
}

// contraction_peak_memory and contraction_all_memory, src/utils.jl:197-229.
int slice_peak_memory(const SliceIn& s, double* peak, double* all, std::string* err) {
  if (validate(s, err)) return -1;
  if (s.nleaves == 0) {
    *peak = -INFINITY;
    *all = -INFINITY;
    return 0;
  }
  std::vector<std::vector<int32_t>> lab;
  compute_labels(s, lab);
  int32_t nl = s.nleaves, nn = nl - 1;
  double cur = 0;
  for (int32_t i = 0; i < nl; i++) cur += std::ldexp(1.0, (int)lab[i].size());
  double pk = cur, tot = 0;
  std::vector<double> ffs(std::max(1, nn), 0.0);
  for (int32_t j = 0; j < nn; j++) {
    double freed = 0;
    for (int q = 0; q < 2; q++) {
      int32_t c = s.tree[2 * j + q];
      if (c >= nl) freed += ffs[c - nl];
      ffs[j] += std::ldexp(1.0, (int)lab[c].size());
    }
    double alloc = std::ldexp(1.0, (int)lab[nl + j].size());
    tot += alloc;
    cur += alloc - freed;
    pk = std::max(pk, cur);
  }
  *peak = std::log2(pk);
  *all = nn > 0 ? std::log2(tot) : -INFINITY;
  return 0;
}

int plan_slice(const SliceIn& s_in, const PlanOptions& opt_in, SlicePlan* out, std::string* err) {
  const PlanOptions& opt = opt_in;
  SlicePlan& P = *out;
  P = SlicePlan();
  if (validate(s_in, err)) return -1;
  if (s_in.nv == 0 || s_in.nleaves == 0) {
    P.empty = true;
    return 0;
  }
  P.empty = false;
  SliceIn s = s_in;
  const int32_t nl = s.nleaves, nn = nl - 1;
  std::vector<std::vector<int32_t>> lab;
  compute_labels(s, lab);
  // tree-node contraction count of the given tree (the reference's tc)
  for (int32_t j = 0; j < nn; j++) {
    const auto &A = lab[s.tree[2 * j]], &Bv = lab[s.tree[2 * j + 1]];
    int u = (int)A.size();
    for (int32_t l : Bv)
      if (pos_in(A, l) < 0) u++;
    P.node_macs += std::ldexp(1.0, u);
  }
  std::vector<int32_t> tree2;
  if (opt.batched && opt.reassociate && nn >= 3 && !std::getenv("TBH_NO_REASSOCIATE")) {
    tree2.assign(s.tree, s.tree + 2 * (size_t)nn);
    reassociate(nl, s.nv, tree2, lab, opt);
    s.tree = tree2.data();
    compute_labels(s, lab);
  }

  // elimination step of each label (node index where it disappears)
  std::vector<int32_t> elim(s.nv, nn);
  for (int32_t j = 0; j < nn; j++) {
    const auto &A = lab[s.tree[2 * j]], &B = lab[s.tree[2 * j + 1]], &C = lab[nl + j];
    for (const auto* X : {&A, &B})
      for (int32_t l : *X)
        if (pos_in(C, l) < 0) elim[l] = j;
  }
  auto sorted_key = [&](std::vector<int32_t> v) {
    std::sort(v.begin(), v.end(), [&](int32_t x, int32_t y) {
      return elim[x] != elim[y] ? elim[x] > elim[y] : x < y;
    });
    return v;
  };

  Builder B{s, opt, P, {}, err, {}, {}, {}};
  B.posA_.assign(std::max(1, s.nv), -1);
  B.posB_.assign(std::max(1, s.nv), -1);
  B.posC_.assign(std::max(1, s.nv), -1);
  // a node becomes up to three steps (permutes, split-K reductions)
  P.ops.reserve(3 * (size_t)nn + 4);
  P.tensors.reserve(3 * (size_t)(nl + nn) + 4);
  B.layout.reserve(3 * (size_t)(nl + nn) + 4);
  // leaves are tensors 0..nl-1 (content is symmetric, so any layout works)
  for (int32_t i = 0; i < nl; i++) {
    B.add_tensor(opt.batched ? sorted_key(lab[i]) : lab[i]);
    PLeaf L;
    L.tensor = i;
    L.kind = (int)lab[i].size();
    L.wbits = 0;
    if (L.kind == 1) {
      int32_t v = lab[i][0];
      L.wbits = s.wbits ? s.wbits[v] : s.unit_wbits;
    }
    P.leaves.push_back(L);
  }
  // internal node output tensors nl..nl+nn-1 (layouts fixed below)
  for (int32_t j = 0; j < nn; j++) B.add_tensor(lab[nl + j]);

  // classify labels of node j
  struct Cls {
    std::vector<int32_t> a, b, k, t;  // a: A-only out, b: B-only out
    bool priv = false;
    int ca = -1, cb = -1;             // child tensors in A / B role
  };
  auto classify = [&](int32_t j) {
    Cls c;
    int32_t x = s.tree[2 * j], y = s.tree[2 * j + 1];
    const auto &LX = lab[x], &LY = lab[y], &LC = lab[nl + j];
    std::vector<int32_t> fx, fy;
    fx.reserve(LC.size());
    fy.reserve(LC.size());
    c.t.reserve(LC.size());
    for (int32_t l : LC) {
      bool inx = pos_in(LX, l) >= 0, iny = pos_in(LY, l) >= 0;
      if (inx && iny)
        c.t.push_back(l);
      else if (inx)
        fx.push_back(l);
      else
        fy.push_back(l);
    }
    for (int32_t l : LX) {
      if (pos_in(LC, l) >= 0) continue;
      if (pos_in(LY, l) >= 0)
        c.k.push_back(l);
      else
        c.priv = true;
    }
    for (int32_t l : LY)
      if (pos_in(LC, l) < 0 && pos_in(LX, l) < 0) c.priv = true;
    // A role: the side with more free labels (M >= N)
    if (fx.size() >= fy.size()) {
      c.ca = x;
      c.cb = y;
      c.a = fx;
      c.b = fy;
    } else {
      c.ca = y;
      c.cb = x;
      c.a = fy;
      c.b = fx;
    }
    return c;
  };
  auto concat = [](std::initializer_list<const std::vector<int32_t>*> parts) {
    std::vector<int32_t> r;
    for (auto* p : parts) r.insert(r.end(), p->begin(), p->end());
    return r;
  };

  if (opt.batched) {
    // top-down layout requests; GEMM-shaped nodes ask their children for
    // (a,k,t) / (k,b,t) and write (a,b,t)
    std::vector<char> is_gemm(nn, 0), has_req(nl + nn, 0);
    std::vector<std::vector<int32_t>> req(nl + nn);
    std::vector<Cls> cls(nn);
    std::vector<std::vector<int32_t>> gemm_out(nn);
    std::vector<int> split(nn, 0);
    for (int32_t j = nn - 1; j >= 0; j--) {
      cls[j] = classify(j);
      Cls& c = cls[j];
      const int o = nl + j;
      int na = (int)c.a.size(), nb = (int)c.b.size(), nk = (int)c.k.size();
      bool g = !c.priv && gemm_shaped(opt, na, nb, nk);
      is_gemm[j] = g;
      if (g) {
        // the order of the M, N and batch labels is free: with a layout
        // request from the consumer, follow it (the output then needs no
        // permute, or one the GEMM's own stores can do)
        auto by_req = [&](std::vector<int32_t> v) {
          if (!has_req[o] || opt.gemm_scatter_bits < 0) return sorted_key(v);
          std::sort(v.begin(), v.end(),
                    [&](int32_t x, int32_t y) { return pos_in(req[o], x) < pos_in(req[o], y); });
          return v;
        };
        auto sa = by_req(c.a), sb = by_req(c.b), sk = sorted_key(c.k), st = by_req(c.t);
        // split-K: the last ks K labels (the latest eliminated stay lowest)
        // become batch labels between the outputs and t
        const int ks = gemm_split(opt, na, nb, nk, (int)st.size());
        std::vector<int32_t> klo(sk.begin(), sk.end() - ks), khi(sk.end() - ks, sk.end());
        split[j] = ks;
        c.a = sa;
        c.b = sb;
        c.k = sk;
        c.t = st;
        gemm_out[j] = concat({&sa, &sb, &khi, &st});
        req[c.ca] = concat({&sa, &klo, &khi, &st});
        req[c.cb] = concat({&klo, &sb, &khi, &st});
        has_req[c.ca] = has_req[c.cb] = 1;
        const auto node_out = concat({&sa, &sb, &st});
        B.layout[o] = has_req[o] ? req[o] : node_out;
      } else {
        B.layout[o] = has_req[o] ? req[o] : sorted_key(lab[o]);
      }
    }
    for (int32_t i = 0; i < nl; i++)
      if (has_req[i]) B.layout[i] = req[i];
    // emit steps in post-order
    for (int32_t j = 0; j < nn; j++) {
      const int o = nl + j;
      Cls& c = cls[j];
      if (is_gemm[j]) {
        int dst = o;
        const int na = (int)c.a.size(), nb = (int)c.b.size();
        POp epi;
        const int scatter = B.layout[o] != gemm_out[j] && split[j] == 0
                                ? gemm_scatter_mode(opt, B.layout[o], gemm_out[j], na, nb, &epi)
                                : 0;
        if (B.layout[o] != gemm_out[j] && !scatter) dst = B.add_tensor(gemm_out[j]);
        B.gemm(c.ca, c.cb, dst, na, nb, (int)c.k.size() - split[j], (int)c.t.size() + split[j]);
        if (scatter) {
          POp& g = P.ops.back();
          g.scatter = scatter;
          std::fill(std::begin(g.pc), std::end(g.pc), (int8_t)-1);
          for (size_t q = 0; q < gemm_out[j].size(); q++)
            g.pc[q] = (int8_t)pos_in(B.layout[o], gemm_out[j][q]);
          std::memcpy(g.ep, epi.ep, sizeof g.ep);
          g.ep_n = epi.ep_n;
          g.ep_split = epi.ep_split;
        }
        if (dst != o && B.ew(dst, -1, o)) return -1;
      } else {
        if (B.ew(s.tree[2 * j], s.tree[2 * j + 1], o)) return -1;
      }
    }
  } else {
    // OMEinsum lowering, node outputs in parse_eincode label order
    for (int32_t j = 0; j < nn; j++) {
      const int o = nl + j;
      Cls c = classify(j);
      int x = s.tree[2 * j], y = s.tree[2 * j + 1];
      // groups as OMEinsum forms them: a, b, t follow the output, k input A
      c.ca = x;
      c.cb = y;
      c.a.clear();
      c.b.clear();
      c.k.clear();
      c.t.clear();
      const auto &LX = lab[x], &LY = lab[y], &LC = lab[o];
      for (int32_t l : LC) {
        bool inx = pos_in(LX, l) >= 0, iny = pos_in(LY, l) >= 0;
        if (inx && iny)
          c.t.push_back(l);
        else if (inx)
          c.a.push_back(l);
        else
          c.b.push_back(l);
      }
      for (int32_t l : LX)
        if (pos_in(LC, l) < 0 && pos_in(LY, l) >= 0) c.k.push_back(l);
      auto wantA = concat({&c.a, &c.k, &c.t}), wantB = concat({&c.k, &c.b, &c.t});
      auto gout = concat({&c.a, &c.b, &c.t});
      int ta = x, tb = y;
      if (B.layout[x] != wantA) {
        ta = B.add_tensor(wantA);
        if (B.ew(x, -1, ta)) return -1;
      }
      if (B.layout[y] != wantB) {
        tb = B.add_tensor(wantB);
        if (B.ew(y, -1, tb)) return -1;
      }
      int dst = o;
      if (gout != LC) dst = B.add_tensor(gout);
      int na = (int)c.a.size(), nb = (int)c.b.size(), nk = (int)c.k.size();
      bool g = gemm_shaped(opt, na, nb, nk);
      if (g)
        B.gemm(ta, tb, dst, na, nb, nk, (int)c.t.size());
      else if (B.ew(ta, tb, dst))
        return -1;
      if (dst != o && B.ew(dst, -1, o)) return -1;
    }
  }

  // root: the last node's output is a scalar; a single leaf is reduced
  if (nn == 0) {
    int sc = B.add_tensor({});
    if (B.ew(0, -1, sc)) return -1;
    P.root = sc;
  } else {
    P.root = nl + nn - 1;
  }
  if (P.tensors[P.root].rank != 0) return fail(err, "root tensor is not a scalar");

  // producers and levels
  const int nt = (int)P.tensors.size();
  std::vector<int> producer(nt, -1), consumer(nt, -1);
  for (size_t i = 0; i < P.ops.size(); i++) {
    producer[P.ops[i].c] = (int)i;
    for (int in : {P.ops[i].a, P.ops[i].b})
      if (in >= 0) consumer[in] = (int)i;
  }
  const int nops = (int)P.ops.size();
  if (opt.batched) {
    std::vector<int> asap(nops, 1);
    int H = 0;
    for (int i = 0; i < nops; i++) {
      int lv = 1;
      for (int in : {P.ops[i].a, P.ops[i].b})
        if (in >= 0 && producer[in] >= 0) lv = std::max(lv, asap[producer[in]] + 1);
      asap[i] = lv;
      H = std::max(H, lv);
    }
    // ALAP: as late as the consumer allows (side subtrees just in time)
    for (int i = nops - 1; i >= 0; i--) {
      int cons = consumer[P.ops[i].c];
      P.ops[i].level = cons < 0 ? H : P.ops[cons].level - 1;
    }
    P.nlevels = H;
    for (int t = 0; t < nt; t++) {
      P.tensors[t].birth = producer[t] >= 0 ? P.ops[producer[t]].level : 0;
      P.tensors[t].death = consumer[t] >= 0 ? P.ops[consumer[t]].level : H + 1;
    }
  } else {
    for (int i = 0; i < nops; i++) P.ops[i].level = i + 1;
    P.nlevels = nops;
    for (int t = 0; t < nt; t++) {
      P.tensors[t].birth = producer[t] >= 0 ? producer[t] + 1 : 0;
      P.tensors[t].death = consumer[t] >= 0 ? consumer[t] + 1 : nops + 1;
    }
  }
  for (auto& t : P.tensors) t.off = -1;
  P.arena_bytes = assign_offsets(P.tensors);
  if (verify_plan(P, err)) return -1;
  for (const auto& op : P.ops) {
    P.exec_macs += op.macs;
    P.exec_bytes += op.bytes;
  }
  return 0;
}

}  // namespace tbh
