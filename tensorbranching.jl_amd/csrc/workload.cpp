// workload.cpp -- host producers of sliced MIS tensor networks (tbh_workload.h).
//
// Not on the hot path: these stand in for the reference's graph generators,
// TreeSA order search and optimal-branching slicer (which need Julia), so that
// the contraction backend gets inputs of the shapes the reference produces.
// See tbh_workload.h for the reference file each function stands in for.
#include "tbh_workload.h"

#include <algorithm>
#include <atomic>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <memory>
#include <mutex>
#include <numeric>
#include <string>
#include <thread>
#include <vector>

namespace {

thread_local std::string g_err;

// ---------------------------------------------------------------- PRNG ----
struct Rng {  // xoshiro256** seeded by SplitMix64
  uint64_t s[4];
  explicit Rng(uint64_t seed) {
    uint64_t z = seed;
    for (auto& x : s) {
      z += 0x9E3779B97F4A7C15ULL;
      uint64_t t = z;
      t = (t ^ (t >> 30)) * 0xBF58476D1CE4E5B9ULL;
      t = (t ^ (t >> 27)) * 0x94D049BB133111EBULL;
      x = t ^ (t >> 31);
    }
  }
  static uint64_t rotl(uint64_t x, int k) { return (x << k) | (x >> (64 - k)); }
  uint64_t next() {
    uint64_t r = rotl(s[1] * 5, 7) * 9, t = s[1] << 17;
    s[2] ^= s[0];
    s[3] ^= s[1];
    s[1] ^= s[2];
    s[0] ^= s[3];
    s[2] ^= t;
    s[3] = rotl(s[3], 45);
    return r;
  }
  double uniform() { return (next() >> 11) * 0x1.0p-53; }
  uint64_t below(uint64_t n) {  // unbiased by rejection
    uint64_t lim = UINT64_MAX - UINT64_MAX % n;
    uint64_t x;
    do x = next();
    while (x >= lim);
    return x % n;
  }
};

}  // namespace

struct tbw_graph {
  int32_t nv = 0;
  std::vector<int32_t> edges;  // 2*ne
};

namespace {

// ------------------------------------------------------ label bookkeeping --
// Leaves of rawcode(IndependentSet(g)): vertex tensors [v] then edge tensors
// [u,v] (u<v).  A tree node keeps a label iff the label occurs outside it.
struct Tree {
  int32_t nleaves = 0;
  std::vector<int32_t> ix_ptr, ix, tree;
  double sc = 0, tc = -INFINITY;
};

// Node labels by the "occurs outside the subtree" rule; returns per node
// (leaves then internal nodes) sorted label vectors.
std::vector<std::vector<int32_t>> node_labels(const Tree& t, int32_t nv) {
  std::vector<std::vector<int32_t>> lab(t.nleaves + std::max(0, t.nleaves - 1));
  std::vector<std::vector<int32_t>> cnt(lab.size());
  std::vector<int32_t> total(nv, 0);
  for (int32_t i = 0; i < t.nleaves; i++)
    for (int32_t k = t.ix_ptr[i]; k < t.ix_ptr[i + 1]; k++) {
      lab[i].push_back(t.ix[k]);
      cnt[i].push_back(1);
      total[t.ix[k]]++;
    }
  std::vector<int32_t> acc(nv, 0);
  for (int32_t j = 0; j + 1 < t.nleaves; j++) {
    int32_t a = t.tree[2 * j], b = t.tree[2 * j + 1], o = t.nleaves + j;
    for (int side = 0; side < 2; side++) {
      int32_t c = side ? b : a;
      for (size_t k = 0; k < lab[c].size(); k++) acc[lab[c][k]] += cnt[c][k];
    }
    for (int side = 0; side < 2; side++) {
      int32_t c = side ? b : a;
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
  return lab;
}

void tree_complexity(Tree& t, int32_t nv) {
  if (t.nleaves < 2) {
    t.sc = t.nleaves == 1 ? t.ix_ptr[1] - t.ix_ptr[0] : 0;
    t.tc = -INFINITY;
    return;
  }
  auto lab = node_labels(t, nv);
  double tot = 0;
  int sc = 0;
  std::vector<char> mark(nv, 0);
  for (int32_t j = 0; j + 1 < t.nleaves; j++) {
    int32_t a = t.tree[2 * j], b = t.tree[2 * j + 1];
    int u = 0;
    for (int32_t l : lab[a]) mark[l] = 1, u++;
    for (int32_t l : lab[b])
      if (!mark[l]) u++;
    for (int32_t l : lab[a]) mark[l] = 0;
    tot += std::ldexp(1.0, u);
    sc = std::max<int>(sc, (int)lab[t.nleaves + j].size());
  }
  t.sc = sc;
  t.tc = std::log2(tot);
}

// ------------------------------------------------ elimination order -------
// Greedy min-fill with random tie-breaking on bitset adjacency.
struct BitGraph {
  int n = 0, words = 0;
  std::vector<uint64_t> bits;  // n * words
  uint64_t* row(int v) { return &bits[(size_t)v * words]; }
  const uint64_t* row(int v) const { return &bits[(size_t)v * words]; }
  void init(int nv) {
    n = nv;
    words = (nv + 63) / 64;
    bits.assign((size_t)n * words, 0);
  }
  void add(int u, int v) {
    row(u)[v >> 6] |= 1ULL << (v & 63);
    row(v)[u >> 6] |= 1ULL << (u & 63);
  }
};

struct OrderScore {
  int width = 0;      // max neighbourhood size at elimination
  double cost = 0;    // sum 2^deg
  bool better(const OrderScore& o) const {
    return width < o.width || (width == o.width && cost < o.cost);
  }
};

OrderScore min_fill_order(const BitGraph& g0, Rng& rng, std::vector<int32_t>& order,
                          double noise) {
  BitGraph g = g0;
  int n = g.n, W = g.words;
  std::vector<char> alive(n, 1);
  std::vector<int> deg(n), fill(n);
  auto popc_and = [&](const uint64_t* a, const uint64_t* b) {
    int c = 0;
    for (int i = 0; i < W; i++) c += __builtin_popcountll(a[i] & b[i]);
    return c;
  };
  auto compute_fill = [&](int v) {
    const uint64_t* rv = g.row(v);
    int d = 0, f2 = 0;
    for (int i = 0; i < W; i++) {
      uint64_t m = rv[i];
      while (m) {
        int u = i * 64 + __builtin_ctzll(m);
        m &= m - 1;
        d++;
        f2 += popc_and(g.row(u), rv);
      }
    }
    deg[v] = d;
    fill[v] = (d * (d - 1) - f2) / 2;
  };
  for (int v = 0; v < n; v++) compute_fill(v);
  order.clear();
  OrderScore sc;
  std::vector<int> cand;
  std::vector<char> touched(n, 0);
  std::vector<int> tl;
  for (int step = 0; step < n; step++) {
    int best = INT32_MAX;
    cand.clear();
    for (int v = 0; v < n; v++) {
      if (!alive[v]) continue;
      // key: fill first, degree as tie-break
      int key = fill[v] * 64 + std::min(deg[v], 63);
      if (key < best) {
        best = key;
        cand.clear();
        cand.push_back(v);
      } else if (key == best) {
        cand.push_back(v);
      }
    }
    int v = cand[rng.below(cand.size())];
    if (noise > 0 && rng.uniform() < noise) {  // occasional random exploration
      std::vector<int> al;
      for (int u = 0; u < n; u++)
        if (alive[u] && fill[u] <= fill[v] + 1) al.push_back(u);
      v = al[rng.below(al.size())];
    }
    order.push_back(v);
    sc.width = std::max(sc.width, deg[v]);
    sc.cost += std::ldexp(1.0, deg[v]);
    // make N(v) a clique, remove v
    std::vector<int> nb;
    const uint64_t* rv = g.row(v);
    for (int i = 0; i < W; i++) {
      uint64_t m = rv[i];
      while (m) {
        nb.push_back(i * 64 + __builtin_ctzll(m));
        m &= m - 1;
      }
    }
    for (size_t a = 0; a < nb.size(); a++)
      for (size_t b = a + 1; b < nb.size(); b++) g.add(nb[a], nb[b]);
    for (int u : nb) g.row(u)[v >> 6] &= ~(1ULL << (v & 63));
    std::fill(g.row(v), g.row(v) + W, 0);
    alive[v] = 0;
    // recompute fills within distance 2
    tl.clear();
    for (int u : nb) {
      if (!touched[u]) touched[u] = 1, tl.push_back(u);
      const uint64_t* ru = g.row(u);
      for (int i = 0; i < W; i++) {
        uint64_t m = ru[i];
        while (m) {
          int x = i * 64 + __builtin_ctzll(m);
          m &= m - 1;
          if (!touched[x]) touched[x] = 1, tl.push_back(x);
        }
      }
    }
    for (int u : tl) {
      touched[u] = 0;
      if (alive[u]) compute_fill(u);
    }
  }
  return sc;
}

BitGraph to_bitgraph(int32_t nv, const std::vector<int32_t>& edges) {
  BitGraph g;
  g.init(nv);
  for (size_t e = 0; e + 1 < edges.size(); e += 2) g.add(edges[e], edges[e + 1]);
  return g;
}

std::vector<int32_t> best_order(int32_t nv, const std::vector<int32_t>& edges,
                                int trials, Rng& rng, OrderScore* out_score) {
  BitGraph g = to_bitgraph(nv, edges);
  std::vector<int32_t> best, cur;
  OrderScore bs;
  bs.width = INT32_MAX;
  for (int t = 0; t < std::max(1, trials); t++) {
    OrderScore s = min_fill_order(g, rng, cur, t == 0 ? 0.0 : 0.02);
    if (s.better(bs)) bs = s, best = cur;
  }
  if (out_score) *out_score = bs;
  return best;
}

// ------------------------------------------- elimination order -> tree ----
// Tensors containing the next eliminated vertex are merged in ascending rank
// (the largest last); leftovers (one per component) are chained at the end.
Tree order_to_tree(int32_t nv, const std::vector<int32_t>& edges,
                   const std::vector<int32_t>& order) {
  Tree t;
  int32_t ne = (int32_t)edges.size() / 2;
  t.nleaves = nv + ne;
  t.ix_ptr.push_back(0);
  for (int32_t v = 0; v < nv; v++) {
    t.ix.push_back(v);
    t.ix_ptr.push_back((int32_t)t.ix.size());
  }
  for (int32_t e = 0; e < ne; e++) {
    int32_t u = edges[2 * e], v = edges[2 * e + 1];
    t.ix.push_back(std::min(u, v));
    t.ix.push_back(std::max(u, v));
    t.ix_ptr.push_back((int32_t)t.ix.size());
  }
  if (t.nleaves == 0) return t;
  struct Tn {
    int32_t id;
    std::vector<int32_t> lab;
  };
  std::vector<Tn> cur;
  std::vector<int32_t> cnt(nv, 0);
  for (int32_t i = 0; i < t.nleaves; i++) {
    Tn x;
    x.id = i;
    for (int32_t k = t.ix_ptr[i]; k < t.ix_ptr[i + 1]; k++) {
      x.lab.push_back(t.ix[k]);
      cnt[t.ix[k]]++;
    }
    cur.push_back(std::move(x));
  }
  int32_t next_id = t.nleaves;
  std::vector<int32_t> inb(nv, 0);
  auto merge = [&](Tn& a, Tn& b) {
    Tn o;
    o.id = next_id++;
    t.tree.push_back(a.id);
    t.tree.push_back(b.id);
    for (int32_t l : a.lab) inb[l]++;
    for (int32_t l : b.lab) inb[l]++;
    for (int side = 0; side < 2; side++)
      for (int32_t l : (side ? b.lab : a.lab)) {
        if (inb[l] == 0) continue;
        if (cnt[l] > inb[l]) {
          o.lab.push_back(l);
          cnt[l] -= inb[l] - 1;
        } else {
          cnt[l] = 0;
        }
        inb[l] = 0;
      }
    std::sort(o.lab.begin(), o.lab.end());
    return o;
  };
  for (int32_t v : order) {
    std::vector<Tn> bucket, rest;
    for (auto& x : cur) {
      if (std::binary_search(x.lab.begin(), x.lab.end(), v))
        bucket.push_back(std::move(x));
      else
        rest.push_back(std::move(x));
    }
    if (bucket.size() >= 2) {
      std::stable_sort(bucket.begin(), bucket.end(),
                       [](const Tn& a, const Tn& b) { return a.lab.size() < b.lab.size(); });
      Tn acc = std::move(bucket[0]);
      for (size_t i = 1; i < bucket.size(); i++) acc = merge(acc, bucket[i]);
      rest.push_back(std::move(acc));
    } else {
      for (auto& x : bucket) rest.push_back(std::move(x));
    }
    cur = std::move(rest);
  }
  while (cur.size() >= 2) {
    Tn a = std::move(cur[0]);
    Tn o = merge(a, cur[1]);
    cur.erase(cur.begin(), cur.begin() + 2);
    cur.insert(cur.begin(), std::move(o));
  }
  tree_complexity(t, nv);
  return t;
}

// ------------------------------------------------------------ TreeSA ------
// Simulated annealing over rotations of the binary contraction tree, in the
// manner of OMEinsumContractionOrders.TreeSA (initialize_code and refine,
// src/dynamic_ob.jl:50-54, src/refine.jl:1-25).  A rotation at node x swaps
// the sibling s of an internal child y with one of y's children; only y's
// tensor changes, and its labels are (L(p) | L(s)) & (L(q) | L(x)), so a move
// is scored in O(words).  Score: log2(sum_nodes 2^|union|) plus
// sc_weight * max(0, sc - sc_target).
struct SATree {
  int nl = 0, nn = 0, W = 0;
  std::vector<int32_t> lc, rc, par;
  std::vector<uint64_t> lab;
  int32_t root = -1;
  double tcsum = 0;
  std::vector<int32_t> hist;
  uint64_t* L(int x) { return &lab[(size_t)x * W]; }
  int rank(int x) {
    int c = 0;
    const uint64_t* a = L(x);
    for (int i = 0; i < W; i++) c += __builtin_popcountll(a[i]);
    return c;
  }
  int uni(int a, int b) {
    int c = 0;
    const uint64_t *x = L(a), *y = L(b);
    for (int i = 0; i < W; i++) c += __builtin_popcountll(x[i] | y[i]);
    return c;
  }
  int sc() {
    for (int r = (int)hist.size() - 1; r > 0; r--)
      if (hist[r]) return r;
    return 0;
  }
};

SATree to_satree(const Tree& t, int32_t nv) {
  SATree s;
  s.nl = t.nleaves;
  s.nn = std::max(0, t.nleaves - 1);
  s.W = std::max(1, (nv + 63) / 64);
  int tot = s.nl + s.nn;
  s.lc.assign(tot, -1);
  s.rc.assign(tot, -1);
  s.par.assign(tot, -1);
  s.lab.assign((size_t)tot * s.W, 0);
  auto lab = node_labels(t, nv);
  for (int x = 0; x < tot; x++)
    for (int32_t l : lab[x]) s.L(x)[l >> 6] |= 1ULL << (l & 63);
  for (int j = 0; j < s.nn; j++) {
    int x = s.nl + j;
    s.lc[x] = t.tree[2 * j];
    s.rc[x] = t.tree[2 * j + 1];
    s.par[s.lc[x]] = x;
    s.par[s.rc[x]] = x;
  }
  s.root = s.nn > 0 ? s.nl + s.nn - 1 : (s.nl ? 0 : -1);
  s.hist.assign(nv + 2, 0);
  for (int j = 0; j < s.nn; j++) {
    int x = s.nl + j;
    s.tcsum += std::ldexp(1.0, s.uni(s.lc[x], s.rc[x]));
    s.hist[s.rank(x)]++;
  }
  return s;
}

Tree from_satree(SATree& s, const Tree& t0, int32_t nv) {
  Tree t;
  t.nleaves = t0.nleaves;
  t.ix_ptr = t0.ix_ptr;
  t.ix = t0.ix;
  if (s.nn > 0) {
    std::vector<int32_t> id(s.nl + s.nn, -1);
    for (int i = 0; i < s.nl; i++) id[i] = i;
    int32_t next = s.nl;
    std::vector<std::pair<int32_t, int>> st;
    st.emplace_back(s.root, 0);
    while (!st.empty()) {
      auto [x, state] = st.back();
      st.pop_back();
      if (x < s.nl) continue;
      if (state == 0) {
        st.emplace_back(x, 1);
        st.emplace_back(s.rc[x], 0);
        st.emplace_back(s.lc[x], 0);
      } else {
        id[x] = next++;
        t.tree.push_back(id[s.lc[x]]);
        t.tree.push_back(id[s.rc[x]]);
      }
    }
  }
  tree_complexity(t, nv);
  return t;
}

void anneal(SATree& s, Rng& rng, double beta0, double beta1, int nbeta, int niters,
            int sc_target, double sc_weight) {
  if (s.nn < 2) return;
  int W = s.W;
  std::vector<uint64_t> ynew(W);
  int cur_sc = s.sc();
  for (int bi = 0; bi < nbeta; bi++) {
    double beta = nbeta == 1 ? beta1 : beta0 + (beta1 - beta0) * bi / (nbeta - 1);
    for (int it = 0; it < niters; it++) {
      for (int j = 0; j < s.nn; j++) {
        int x = s.nl + (int)rng.below(s.nn);
        int cand_y[2], nc = 0;
        if (s.lc[x] >= s.nl) cand_y[nc++] = s.lc[x];
        if (s.rc[x] >= s.nl) cand_y[nc++] = s.rc[x];
        if (!nc) continue;
        int y = cand_y[nc == 1 ? 0 : rng.below(2)];
        int sib = s.lc[x] == y ? s.rc[x] : s.lc[x];
        bool keep_left = rng.below(2);
        int p = keep_left ? s.lc[y] : s.rc[y];
        int q = keep_left ? s.rc[y] : s.lc[y];
        const uint64_t *Lp = s.L(p), *Lq = s.L(q), *Ls = s.L(sib), *Lx = s.L(x);
        int ry_new = 0, uy_new = 0, ux_new = 0;
        for (int i = 0; i < W; i++) {
          uint64_t ps = Lp[i] | Ls[i];
          ynew[i] = ps & (Lq[i] | Lx[i]);
          ry_new += __builtin_popcountll(ynew[i]);
          uy_new += __builtin_popcountll(ps);
          ux_new += __builtin_popcountll(ynew[i] | Lq[i]);
        }
        int uy_old = s.uni(p, q), ux_old = s.uni(y, sib), ry_old = s.rank(y);
        double tc_new = s.tcsum + std::ldexp(1.0, uy_new) + std::ldexp(1.0, ux_new) -
                        std::ldexp(1.0, uy_old) - std::ldexp(1.0, ux_old);
        // local energy of the two touched nodes, as TreeSA scores a rule
        int rx = s.rank(x);
        double lt_old = std::log2(std::ldexp(1.0, uy_old) + std::ldexp(1.0, ux_old));
        double lt_new = std::log2(std::ldexp(1.0, uy_new) + std::ldexp(1.0, ux_new));
        int ls_old = std::max(ry_old, rx), ls_new = std::max(ry_new, rx);
        double d = lt_new - lt_old + sc_weight * (std::max(ls_new, sc_target) -
                                                   std::max(ls_old, sc_target));
        s.hist[ry_old]--;
        s.hist[ry_new]++;
        int sc_new = ry_new > cur_sc ? ry_new : (s.hist[cur_sc] ? cur_sc : s.sc());
        if (d <= 0 || rng.uniform() < std::exp(-beta * d)) {
          // y = (p, sib); x = (y, q)
          s.lc[y] = p;
          s.rc[y] = sib;
          s.par[sib] = y;
          if (s.lc[x] == y)
            s.rc[x] = q;
          else
            s.lc[x] = q;
          s.par[q] = x;
          std::copy(ynew.begin(), ynew.end(), s.L(y));
          s.tcsum = tc_new;
          cur_sc = sc_new;
        } else {
          s.hist[ry_new]--;
          s.hist[ry_old]++;
        }
      }
    }
  }
}

Tree treesa(const Tree& t, int32_t nv, Rng& rng, int ntrials, double beta0, double beta1,
            int nbeta, int niters, int sc_target) {
  Tree best = t;
  for (int k = 0; k < ntrials; k++) {
    SATree s = to_satree(t, nv);
    anneal(s, rng, beta0, beta1, nbeta, niters, sc_target, 1.0);
    Tree c = from_satree(s, t, nv);
    if (c.sc < best.sc || (c.sc == best.sc && c.tc < best.tc)) best = std::move(c);
  }
  return best;
}

// ---------------------------------------------------------- branches ------
struct Branch {
  int32_t nv = 0;
  std::vector<int32_t> edges;  // current labels, order defines edge leaves
  std::vector<double> w;
  double r = 0;
  Tree t;                      // leaves: vertices then edges
};

struct Finished {
  int32_t nv = 0;
  std::vector<int32_t> edges;
  std::vector<double> w;
  double r = 0;
  Tree t;
};

Tree leaves_of(int32_t nv, const std::vector<int32_t>& edges) {
  Tree t;
  int32_t ne = (int32_t)edges.size() / 2;
  t.nleaves = nv + ne;
  t.ix_ptr.push_back(0);
  for (int32_t v = 0; v < nv; v++) {
    t.ix.push_back(v);
    t.ix_ptr.push_back((int32_t)t.ix.size());
  }
  for (int32_t e = 0; e < ne; e++) {
    t.ix.push_back(std::min(edges[2 * e], edges[2 * e + 1]));
    t.ix.push_back(std::max(edges[2 * e], edges[2 * e + 1]));
    t.ix_ptr.push_back((int32_t)t.ix.size());
  }
  return t;
}

// Remove the vertices in `drop`, relabel, and shrink the tree the way
// remove_tensors + reform_tree! do (src/utils.jl:46-58,102-131): leaves that
// touch a dropped vertex go, and nodes left with one child are bypassed.
Branch induced(const Branch& b, const std::vector<char>& drop) {
  Branch o;
  std::vector<int32_t> nid(b.nv, -1);
  for (int32_t v = 0; v < b.nv; v++)
    if (!drop[v]) {
      nid[v] = o.nv++;
      o.w.push_back(b.w[v]);
    }
  int32_t ne_old = (int32_t)b.edges.size() / 2;
  std::vector<int32_t> emap(ne_old, -1);
  int32_t ne_new = 0;
  for (int32_t e = 0; e < ne_old; e++) {
    int32_t u = nid[b.edges[2 * e]], v = nid[b.edges[2 * e + 1]];
    if (u >= 0 && v >= 0) {
      o.edges.push_back(u);
      o.edges.push_back(v);
      emap[e] = ne_new++;
    }
  }
  o.r = b.r;
  o.t = leaves_of(o.nv, o.edges);
  int32_t nl_new = o.t.nleaves;
  const Tree& t = b.t;
  std::vector<int32_t> m(t.nleaves + std::max(0, t.nleaves - 1), -1);
  for (int32_t v = 0; v < b.nv; v++) m[v] = nid[v];
  for (int32_t e = 0; e < ne_old; e++) m[b.nv + e] = emap[e] >= 0 ? o.nv + emap[e] : -1;
  int32_t next = nl_new;
  for (int32_t j = 0; j + 1 < t.nleaves; j++) {
    int32_t a = m[t.tree[2 * j]], c = m[t.tree[2 * j + 1]];
    int32_t& out = m[t.nleaves + j];
    if (a < 0)
      out = c;
    else if (c < 0)
      out = a;
    else {
      o.t.tree.push_back(a);
      o.t.tree.push_back(c);
      out = next++;
    }
  }
  tree_complexity(o.t, o.nv);
  return o;
}

// Simplicial-vertex reduction: if N(v) is a clique and w_v >= w_u for all
// u in N(v), some maximum independent set contains v: take it.
Branch kernelize(Branch b) {
  for (;;) {
    std::vector<std::vector<int32_t>> adj(b.nv);
    for (size_t e = 0; e + 1 < b.edges.size(); e += 2) {
      adj[b.edges[e]].push_back(b.edges[e + 1]);
      adj[b.edges[e + 1]].push_back(b.edges[e]);
    }
    for (auto& a : adj) std::sort(a.begin(), a.end());
    std::vector<char> drop(b.nv, 0);
    bool any = false;
    for (int32_t v = 0; v < b.nv; v++) {
      if (drop[v]) continue;
      bool ok = adj[v].size() <= 4;
      for (int32_t u : adj[v])
        if (drop[u] || b.w[u] > b.w[v]) ok = false;
      for (size_t i = 0; ok && i < adj[v].size(); i++)
        for (size_t j = i + 1; ok && j < adj[v].size(); j++)
          if (!std::binary_search(adj[adj[v][i]].begin(), adj[adj[v][i]].end(), adj[v][j]))
            ok = false;
      if (!ok) continue;
      b.r += b.w[v];
      drop[v] = 1;
      for (int32_t u : adj[v]) drop[u] = 1;
      any = true;
    }
    if (!any) return b;
    b = induced(b, drop);
  }
}

struct SliceResult {
  bool finished = false;
  Finished fin;
  std::vector<Branch> children;
};

SliceResult step(Branch b, int sc_target, bool refine, uint64_t seed, bool do_kernel) {
  SliceResult res;
  if (do_kernel) b = kernelize(std::move(b));
  Rng rng(seed);
  if (b.nv == 0) {
    res.finished = true;
    res.fin.r = b.r;
    return res;
  }
  // refine as TreeSARefiner does (src/refine.jl:1-19): short re-annealing
  if (refine && b.t.sc > sc_target - 3)
    b.t = treesa(b.t, b.nv, rng, 1, 1.0, 15.0, 8, 6, sc_target);
  if (b.t.sc <= sc_target) {
    res.finished = true;
    res.fin.nv = b.nv;
    res.fin.edges = std::move(b.edges);
    res.fin.w = std::move(b.w);
    res.fin.r = b.r;
    res.fin.t = std::move(b.t);
    return res;
  }
  // branch on the label occurring most often among the largest tensors
  auto lab = node_labels(b.t, b.nv);
  std::vector<int32_t> hits(b.nv, 0), deg(b.nv, 0);
  for (size_t e = 0; e < b.edges.size(); e++) deg[b.edges[e]]++;
  for (int32_t j = b.t.nleaves; j < (int32_t)lab.size(); j++)
    if ((double)lab[j].size() >= b.t.sc)
      for (int32_t l : lab[j]) hits[l]++;
  int32_t v = 0;
  for (int32_t u = 1; u < b.nv; u++)
    if (hits[u] > hits[v] || (hits[u] == hits[v] && deg[u] > deg[v])) v = u;
  std::vector<char> drop(b.nv, 0);
  drop[v] = 1;
  res.children.push_back(induced(b, drop));
  for (size_t e = 0; e + 1 < b.edges.size(); e += 2) {
    if (b.edges[e] == v) drop[b.edges[e + 1]] = 1;
    if (b.edges[e + 1] == v) drop[b.edges[e]] = 1;
  }
  Branch inc = induced(b, drop);
  inc.r += b.w[v];
  res.children.push_back(std::move(inc));
  return res;
}

// Initial code for a whole graph: best min-fill order -> tree, then TreeSA.
Tree initial_tree(int32_t nv, const std::vector<int32_t>& edges, int trials, Rng& rng,
                  int sc_target) {
  auto o = best_order(nv, edges, std::max(1, trials / 4), rng, nullptr);
  Tree t = order_to_tree(nv, edges, o);
  if (trials > 0) t = treesa(t, nv, rng, std::max(1, trials / 10), 0.01, 15.0, 150, std::max(10, trials / 4), sc_target);
  return t;
}

}  // namespace

struct tbw_sliceset {
  std::vector<Finished> items;
};

extern "C" {

const char* tbw_last_error(void) { return g_err.c_str(); }

tbw_graph* tbw_random_regular(int32_t n, int32_t d, uint64_t seed) {
  if (n <= 0 || d < 0 || d >= n || ((int64_t)n * d) % 2) {
    g_err = "random_regular: need 0 <= d < n and n*d even";
    return nullptr;
  }
  Rng rng(seed);
  std::vector<int32_t> stubs((size_t)n * d);
  for (int attempt = 0; attempt < 100000; attempt++) {
    for (int32_t v = 0; v < n; v++)
      for (int32_t k = 0; k < d; k++) stubs[(size_t)v * d + k] = v;
    for (size_t i = stubs.size(); i > 1; i--) std::swap(stubs[i - 1], stubs[rng.below(i)]);
    bool ok = true;
    std::vector<std::pair<int32_t, int32_t>> es;
    for (size_t i = 0; i < stubs.size() && ok; i += 2) {
      int32_t u = std::min(stubs[i], stubs[i + 1]), v = std::max(stubs[i], stubs[i + 1]);
      if (u == v) ok = false;
      es.emplace_back(u, v);
    }
    if (!ok) continue;
    std::sort(es.begin(), es.end());
    if (std::adjacent_find(es.begin(), es.end()) != es.end()) continue;
    auto* g = new tbw_graph;
    g->nv = n;
    for (auto& e : es) {
      g->edges.push_back(e.first);
      g->edges.push_back(e.second);
    }
    return g;
  }
  g_err = "random_regular: no simple graph found";
  return nullptr;
}

tbw_graph* tbw_erdos_renyi(int32_t n, double p, uint64_t seed) {
  if (n < 0 || p < 0 || p > 1) {
    g_err = "erdos_renyi: bad arguments";
    return nullptr;
  }
  Rng rng(seed);
  auto* g = new tbw_graph;
  g->nv = n;
  for (int32_t u = 0; u < n; u++)
    for (int32_t v = u + 1; v < n; v++)
      if (rng.uniform() < p) {
        g->edges.push_back(u);
        g->edges.push_back(v);
      }
  return g;
}

tbw_graph* tbw_ksg(int32_t m, int32_t n, double rho, uint64_t seed) {
  if (m <= 0 || n <= 0 || rho < 0 || rho > 1) {
    g_err = "ksg: bad arguments";
    return nullptr;
  }
  Rng rng(seed);
  int32_t N = m * n, k = (int32_t)std::lround(rho * N);
  std::vector<int32_t> sites(N);
  std::iota(sites.begin(), sites.end(), 0);
  for (int32_t i = 0; i < k; i++) std::swap(sites[i], sites[i + rng.below(N - i)]);
  sites.resize(k);
  std::sort(sites.begin(), sites.end());
  std::vector<int32_t> id(N, -1);
  for (int32_t i = 0; i < k; i++) id[sites[i]] = i;
  auto* g = new tbw_graph;
  g->nv = k;
  for (int32_t i = 0; i < k; i++) {
    int32_t x = sites[i] / n, y = sites[i] % n;
    for (int dx = 0; dx <= 1; dx++)
      for (int dy = -1; dy <= 1; dy++) {
        if (dx == 0 && dy <= 0) continue;
        int32_t xx = x + dx, yy = y + dy;
        if (xx >= m || yy < 0 || yy >= n) continue;
        int32_t j = id[xx * n + yy];
        if (j >= 0) {
          g->edges.push_back(std::min(i, j));
          g->edges.push_back(std::max(i, j));
        }
      }
  }
  return g;
}

tbw_graph* tbw_graph_from_edges(int32_t nv, int32_t ne, const int32_t* edges) {
  if (nv < 0 || ne < 0) {
    g_err = "graph: negative size";
    return nullptr;
  }
  std::vector<std::pair<int32_t, int32_t>> es;
  for (int32_t e = 0; e < ne; e++) {
    int32_t u = edges[2 * e], v = edges[2 * e + 1];
    if (u < 0 || v < 0 || u >= nv || v >= nv || u == v) {
      g_err = "graph: bad edge";
      return nullptr;
    }
    es.emplace_back(std::min(u, v), std::max(u, v));
  }
  std::sort(es.begin(), es.end());
  es.erase(std::unique(es.begin(), es.end()), es.end());
  auto* g = new tbw_graph;
  g->nv = nv;
  for (auto& e : es) {
    g->edges.push_back(e.first);
    g->edges.push_back(e.second);
  }
  return g;
}

int32_t tbw_graph_nv(const tbw_graph* g) { return g->nv; }
int32_t tbw_graph_ne(const tbw_graph* g) { return (int32_t)g->edges.size() / 2; }
const int32_t* tbw_graph_edges(const tbw_graph* g) { return g->edges.data(); }
void tbw_graph_free(tbw_graph* g) { delete g; }

int tbw_elimination_order(const tbw_graph* g, int32_t trials, uint64_t seed,
                          int32_t* order_out) {
  Rng rng(seed);
  auto o = best_order(g->nv, g->edges, trials, rng, nullptr);
  std::copy(o.begin(), o.end(), order_out);
  return 0;
}

tbw_sliceset* tbw_whole(const tbw_graph* g, const double* weights,
                        int32_t order_trials, uint64_t seed) {
  Rng rng(seed);
  auto* ss = new tbw_sliceset;
  Finished f;
  f.nv = g->nv;
  f.edges = g->edges;
  f.w.assign(g->nv, 1.0);
  if (weights) f.w.assign(weights, weights + g->nv);
  f.t = initial_tree(g->nv, g->edges, order_trials, rng, 20);
  ss->items.push_back(std::move(f));
  return ss;
}

tbw_sliceset* tbw_slice(const tbw_graph* g, const double* weights,
                        int32_t sc_target, int32_t order_trials, uint64_t seed,
                        int32_t nthreads, int32_t do_kernel) {
  return tbw_slice_limited(g, weights, sc_target, order_trials, seed, nthreads, do_kernel,
                           (int64_t)1 << 18);
}

tbw_sliceset* tbw_slice_limited(const tbw_graph* g, const double* weights,
                                int32_t sc_target, int32_t order_trials, uint64_t seed,
                                int32_t nthreads, int32_t do_kernel, int64_t max_slices) {
  if (max_slices < 1) {
    g_err = "slice: max_slices must be positive";
    return nullptr;
  }
  if (nthreads <= 0) nthreads = (int32_t)std::max(1u, std::thread::hardware_concurrency());
  Branch root;
  root.nv = g->nv;
  root.edges = g->edges;
  root.w.assign(g->nv, 1.0);
  if (weights) root.w.assign(weights, weights + g->nv);
  Rng rng(seed);
  root.t = initial_tree(g->nv, g->edges, order_trials, rng, sc_target);
  auto* ss = new tbw_sliceset;
  std::vector<Branch> frontier;
  frontier.push_back(std::move(root));
  const size_t kMaxSlices = (size_t)max_slices;
  uint64_t round_seed = seed * 0x9E3779B97F4A7C15ULL + 1;
  // the root uses the tree searched above; later rounds refine briefly
  bool first = true;
  while (!frontier.empty()) {
    std::vector<SliceResult> res(frontier.size());
    std::atomic<size_t> next{0};
    auto worker = [&]() {
      for (;;) {
        size_t i = next.fetch_add(1);
        if (i >= frontier.size()) return;
        res[i] = step(std::move(frontier[i]), sc_target, !first && order_trials > 0,
                      round_seed + i * 7919, do_kernel != 0);
      }
    };
    std::vector<std::thread> th;
    int nt = (int)std::min<size_t>(nthreads, frontier.size());
    for (int k = 0; k < nt; k++) th.emplace_back(worker);
    for (auto& x : th) x.join();
    first = false;
    round_seed = round_seed * 6364136223846793005ULL + 1442695040888963407ULL;
    std::vector<Branch> nf;
    for (auto& r : res) {
      if (r.finished)
        ss->items.push_back(std::move(r.fin));
      else
        for (auto& c : r.children) nf.push_back(std::move(c));
    }
    if (ss->items.size() + nf.size() > kMaxSlices) {
      g_err = "slice: more than " + std::to_string(kMaxSlices) + " slices; raise sc_target";
      delete ss;
      return nullptr;
    }
    frontier = std::move(nf);
  }
  return ss;
}

int64_t tbw_sliceset_count(const tbw_sliceset* s) { return (int64_t)s->items.size(); }

int tbw_sliceset_get(const tbw_sliceset* s, int64_t i, tbw_slice_view* out) {
  if (i < 0 || i >= (int64_t)s->items.size()) {
    g_err = "sliceset index out of range";
    return -1;
  }
  const Finished& f = s->items[i];
  out->nv = f.nv;
  out->ne = (int32_t)f.edges.size() / 2;
  out->edges = f.edges.data();
  out->weights = f.w.data();
  out->nleaves = f.t.nleaves;
  out->ix_ptr = f.t.ix_ptr.data();
  out->ix = f.t.ix.data();
  out->tree = f.t.tree.data();
  out->r = f.r;
  out->sc = f.t.sc;
  out->tc = f.t.tc;
  return 0;
}

void tbw_sliceset_free(tbw_sliceset* s) { delete s; }

}  // extern "C"
