// Multilevel vertex separators for the nested-dissection ordering.
//
// The subgraph is coarsened by heavy-edge matching until a few hundred
// vertices remain; the coarsest graph is bisected by breadth-first growth
// from several random seeds, the cut's lighter boundary becomes the
// separator, and a node-based Fiduccia-Mattheyses refinement (a separator
// vertex moves to one side and pulls its neighbours on the other side into
// the separator; the move sequence is rolled back to its best prefix) improves
// it at every level on the way back to the full graph.  This is the scheme of
// multilevel nested dissection (Karypis and Kumar, "A fast and high quality
// multilevel scheme for partitioning irregular graphs", SIAM J. Sci. Comput.
// 20 (1998); node refinement as in their METIS library's documentation),
// written from that description.  It plays the role of the METIS ordering the
// reference's GPU solver uses (cuDSS, vajax/analysis/solver_factories.py:895-922).
#include "separator.h"

#include <algorithm>
#include <numeric>
#include <queue>
#include <tuple>

namespace vjx {
namespace {

struct Graph {
  int n = 0;
  std::vector<int32_t> xadj{0}, adjncy, adjw, vw;
  int64_t total = 0;
};

// splitmix64: deterministic choices for a given seed
struct Rng {
  uint64_t s;
  uint64_t next() {
    uint64_t z = (s += 0x9e3779b97f4a7c15ULL);
    z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ULL;
    z = (z ^ (z >> 27)) * 0x94d049bb133111ebULL;
    return z ^ (z >> 31);
  }
  int below(int k) { return (int)(next() % (uint64_t)k); }
};

// Heavy-edge matching (random visiting order, coarse vertices no heavier than
// max_cw); false when it no longer shrinks the graph by 8 %.
bool coarsen(const Graph& g, Rng& rng, int64_t max_cw, Graph& c, std::vector<int32_t>& cmap) {
  const int n = g.n;
  std::vector<int32_t> match(n, -1), perm(n);
  std::iota(perm.begin(), perm.end(), 0);
  for (int i = n - 1; i > 0; --i) std::swap(perm[i], perm[rng.below(i + 1)]);
  for (int v : perm) {
    if (match[v] >= 0) continue;
    int best = -1, bw = -1;
    for (int e = g.xadj[v]; e < g.xadj[v + 1]; ++e) {
      const int u = g.adjncy[e];
      if (match[u] >= 0 || (int64_t)g.vw[u] + g.vw[v] > max_cw) continue;
      if (g.adjw[e] > bw) {
        bw = g.adjw[e];
        best = u;
      }
    }
    if (best >= 0) {
      match[v] = best;
      match[best] = v;
    } else {
      match[v] = v;
    }
  }
  cmap.assign(n, -1);
  int cn = 0;
  for (int v = 0; v < n; ++v)
    if (cmap[v] < 0) {
      cmap[v] = cn;
      cmap[match[v]] = cn;
      ++cn;
    }
  if ((double)cn > 0.92 * (double)n) return false;
  c = Graph();
  c.n = cn;
  c.total = g.total;
  c.vw.assign(cn, 0);
  c.xadj.assign(cn + 1, 0);
  std::vector<int32_t> pos(cn, -1);
  // coarse vertices were numbered in the order of their first fine vertex
  for (int v = 0; v < n; ++v) {
    const int u = match[v];
    if (u < v) continue;
    const int cv = cmap[v];
    const int32_t seg = (int32_t)c.adjncy.size();
    c.vw[cv] = g.vw[v] + (u != v ? g.vw[u] : 0);
    for (int w : {v, u}) {
      for (int e = g.xadj[w]; e < g.xadj[w + 1]; ++e) {
        const int cu = cmap[g.adjncy[e]];
        if (cu == cv) continue;
        if (pos[cu] >= seg) {
          c.adjw[pos[cu]] += g.adjw[e];
        } else {
          pos[cu] = (int32_t)c.adjncy.size();
          c.adjncy.push_back(cu);
          c.adjw.push_back(g.adjw[e]);
        }
      }
      if (u == v) break;
    }
    c.xadj[cv + 1] = (int32_t)c.adjncy.size();
  }
  return true;
}

// Part weights: [0], [1] the halves, [2] the separator.
struct Weights {
  int64_t w[3] = {0, 0, 0};
};

Weights weigh(const Graph& g, const std::vector<int8_t>& where) {
  Weights p;
  for (int v = 0; v < g.n; ++v) p.w[where[v]] += g.vw[v];
  return p;
}

// Breadth-first growth of side 0 from `seed` to half the weight; the
// lighter boundary of the cut becomes the separator.
void grow_separator(const Graph& g, int seed, std::vector<int8_t>& where) {
  where.assign(g.n, 1);
  const int64_t target = g.total / 2;
  int64_t w0 = 0;
  std::vector<int32_t> q;
  q.reserve(g.n);
  where[seed] = 0;
  w0 += g.vw[seed];
  q.push_back(seed);
  size_t h = 0;
  int scan = 0;
  while (w0 < target) {
    if (h == q.size()) {  // a disconnected remainder: restart from any vertex
      while (scan < g.n && where[scan] != 1) ++scan;
      if (scan == g.n) break;
      where[scan] = 0;
      w0 += g.vw[scan];
      q.push_back(scan);
      continue;
    }
    const int v = q[h++];
    for (int e = g.xadj[v]; e < g.xadj[v + 1] && w0 < target; ++e) {
      const int u = g.adjncy[e];
      if (where[u] == 1) {
        where[u] = 0;
        w0 += g.vw[u];
        q.push_back(u);
      }
    }
  }
  int64_t bw[2] = {0, 0};
  std::vector<char> bnd(g.n, 0);
  for (int v = 0; v < g.n; ++v)
    for (int e = g.xadj[v]; e < g.xadj[v + 1]; ++e)
      if (where[g.adjncy[e]] != where[v]) {
        bnd[v] = 1;
        bw[where[v]] += g.vw[v];
        break;
      }
  const int side = bw[0] <= bw[1] ? 0 : 1;
  for (int v = 0; v < g.n; ++v)
    if (bnd[v] && where[v] == side) where[v] = 2;
}

// One node-based FM pass; returns true if the separator got lighter.
bool refine_pass(const Graph& g, std::vector<int8_t>& where, Weights& pw, int64_t max_side,
                 std::vector<int32_t>& version, std::vector<char>& locked) {
  using Entry = std::tuple<int64_t, int32_t, int32_t>;  // gain, vertex, version
  std::priority_queue<Entry> q[2];
  auto gain = [&](int v, int to) {
    const int other = 1 - to;
    int64_t pulled = 0;
    for (int e = g.xadj[v]; e < g.xadj[v + 1]; ++e)
      if (where[g.adjncy[e]] == other) pulled += g.vw[g.adjncy[e]];
    return (int64_t)g.vw[v] - pulled;
  };
  auto push = [&](int v) {
    ++version[v];
    q[0].push({gain(v, 0), v, version[v]});
    q[1].push({gain(v, 1), v, version[v]});
  };
  std::fill(locked.begin(), locked.end(), 0);
  for (int v = 0; v < g.n; ++v)
    if (where[v] == 2) push(v);
  const int64_t start_sep = pw.w[2];
  int64_t best_sep = pw.w[2];
  int64_t best_imb = std::llabs(pw.w[0] - pw.w[1]);
  struct Move {
    int32_t v;
    int8_t to;
    int32_t pulled_end;
  };
  std::vector<Move> log;
  std::vector<int32_t> pulled;
  size_t best_len = 0;
  const int limit = std::min(400, std::max(50, g.n / 100));
  int since_best = 0;
  std::vector<int32_t> touched;
  while (true) {
    int cand[2] = {-1, -1};
    int64_t cg[2] = {0, 0};
    for (int to = 0; to < 2; ++to) {
      while (!q[to].empty()) {
        const auto [gn, v, ver] = q[to].top();
        if (where[v] != 2 || locked[v] || ver != version[v]) {
          q[to].pop();
          continue;
        }
        if (pw.w[to] + g.vw[v] <= max_side) {
          cand[to] = v;
          cg[to] = gn;
        }
        break;
      }
    }
    int to;
    if (cand[0] < 0 && cand[1] < 0) break;
    if (cand[0] < 0)
      to = 1;
    else if (cand[1] < 0)
      to = 0;
    else if (cg[0] != cg[1])
      to = cg[0] > cg[1] ? 0 : 1;
    else
      to = pw.w[0] <= pw.w[1] ? 0 : 1;
    const int v = cand[to];
    q[to].pop();
    const int other = 1 - to;
    locked[v] = 1;
    where[v] = (int8_t)to;
    pw.w[to] += g.vw[v];
    pw.w[2] -= g.vw[v];
    touched.clear();
    for (int e = g.xadj[v]; e < g.xadj[v + 1]; ++e) {
      const int u = g.adjncy[e];
      if (where[u] == other) {
        where[u] = 2;
        pw.w[other] -= g.vw[u];
        pw.w[2] += g.vw[u];
        pulled.push_back(u);
        touched.push_back(u);
        for (int f = g.xadj[u]; f < g.xadj[u + 1]; ++f)
          if (where[g.adjncy[f]] == 2) touched.push_back(g.adjncy[f]);
      } else if (where[u] == 2) {
        touched.push_back(u);
      }
    }
    log.push_back({(int32_t)v, (int8_t)to, (int32_t)pulled.size()});
    for (int u : touched)
      if (where[u] == 2 && !locked[u]) push(u);
    const int64_t imb = std::llabs(pw.w[0] - pw.w[1]);
    if (pw.w[2] < best_sep || (pw.w[2] == best_sep && imb < best_imb)) {
      best_sep = pw.w[2];
      best_imb = imb;
      best_len = log.size();
      since_best = 0;
    } else if (++since_best > limit) {
      break;
    }
  }
  // roll back to the best prefix
  while (log.size() > best_len) {
    const Move m = log.back();
    log.pop_back();
    const int32_t begin = log.empty() ? 0 : log.back().pulled_end;
    const int other = 1 - m.to;
    for (int32_t i = m.pulled_end - 1; i >= begin; --i) {
      const int u = pulled[i];
      where[u] = (int8_t)other;
      pw.w[2] -= g.vw[u];
      pw.w[other] += g.vw[u];
    }
    pulled.resize(begin);
    where[m.v] = 2;
    pw.w[m.to] -= g.vw[m.v];
    pw.w[2] += g.vw[m.v];
  }
  return best_sep < start_sep;
}

void refine(const Graph& g, std::vector<int8_t>& where, int64_t max_side, int passes) {
  Weights pw = weigh(g, where);
  std::vector<int32_t> version(g.n, 0);
  std::vector<char> locked(g.n, 0);
  for (int p = 0; p < passes; ++p)
    if (!refine_pass(g, where, pw, max_side, version, locked)) break;
}

}  // namespace

void vertex_separator(const std::vector<std::vector<int32_t>>& adj,
                      const std::vector<int32_t>& nodes, std::vector<int32_t>& loc, uint64_t seed,
                      std::vector<int8_t>& part) {
  const int k = (int)nodes.size();
  part.assign(k, 0);
  if (k < 3) return;
  for (int i = 0; i < k; ++i) loc[nodes[i]] = i;
  std::vector<Graph> lv(1);
  Graph& g0 = lv[0];
  g0.n = k;
  g0.total = k;
  g0.vw.assign(k, 1);
  g0.xadj.assign(k + 1, 0);
  for (int i = 0; i < k; ++i) {
    for (int u : adj[nodes[i]]) {
      const int j = loc[u];
      if (j >= 0 && j < k && nodes[j] == u) {
        g0.adjncy.push_back(j);
        g0.adjw.push_back(1);
      }
    }
    g0.xadj[i + 1] = (int32_t)g0.adjncy.size();
  }
  for (int i = 0; i < k; ++i) loc[nodes[i]] = -1;

  Rng rng{seed ^ ((uint64_t)k << 32)};
  constexpr int kCoarsest = 160;
  const int64_t max_cw = std::max<int64_t>(1, (int64_t)(1.5 * (double)k / kCoarsest));
  std::vector<std::vector<int32_t>> maps;
  while (lv.back().n > kCoarsest && lv.size() < 48) {
    Graph c;
    std::vector<int32_t> cmap;
    if (!coarsen(lv.back(), rng, max_cw, c, cmap)) break;
    lv.push_back(std::move(c));
    maps.push_back(std::move(cmap));
  }
  // halves may differ by up to 20 % of the weight
  const int64_t max_side = (int64_t)(0.6 * (double)k) + 1;

  // coarsest level: the best of several grown bisections
  const Graph& gc = lv.back();
  std::vector<int8_t> where, best;
  int64_t best_sep = -1, best_imb = 0;
  for (int trial = 0; trial < 8; ++trial) {
    grow_separator(gc, rng.below(gc.n), where);
    refine(gc, where, max_side, 8);
    const Weights w = weigh(gc, where);
    const int64_t imb = std::llabs(w.w[0] - w.w[1]);
    const bool ok = w.w[0] <= max_side && w.w[1] <= max_side;
    const bool best_ok = !best.empty() && weigh(gc, best).w[0] <= max_side &&
                         weigh(gc, best).w[1] <= max_side;
    if (best.empty() || (ok && !best_ok) ||
        (ok == best_ok && (w.w[2] < best_sep || (w.w[2] == best_sep && imb < best_imb)))) {
      best = where;
      best_sep = w.w[2];
      best_imb = imb;
    }
  }
  where.swap(best);
  // project and refine level by level
  for (int l = (int)lv.size() - 1; l > 0; --l) {
    const Graph& f = lv[l - 1];
    const std::vector<int32_t>& cmap = maps[l - 1];
    std::vector<int8_t> fw(f.n);
    for (int v = 0; v < f.n; ++v) fw[v] = where[cmap[v]];
    where.swap(fw);
    refine(f, where, max_side, 4);
  }
  for (int i = 0; i < k; ++i) part[i] = where[i];
}

}  // namespace vjx
