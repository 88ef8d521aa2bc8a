// Symbolic LU analysis on the host (see symbolic.h for the contract).
#include "symbolic.h"

#include "separator.h"

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <numeric>
#include <set>
#include <array>
#include <chrono>
#include <tuple>

namespace vjx {

bool csr_validate(int n, int nnz, const int32_t* indptr, const int32_t* indices, std::string& err) {
  if (n < 0 || nnz < 0) {
    err = "negative matrix size";
    return false;
  }
  if (indptr[0] != 0) {
    err = "CSR indptr[0] != 0";
    return false;
  }
  if (indptr[n] != nnz) {
    err = "CSR indptr[n] != nnz";
    return false;
  }
  for (int r = 0; r < n; ++r) {
    if (indptr[r + 1] < indptr[r]) {
      err = "CSR indptr not monotonic at row " + std::to_string(r);
      return false;
    }
    for (int e = indptr[r]; e < indptr[r + 1]; ++e) {
      int c = indices[e];
      if (c < 0 || c >= n) {
        err = "CSR column index out of range in row " + std::to_string(r);
        return false;
      }
      if (e > indptr[r] && c <= indices[e - 1]) {
        err = "CSR column indices not strictly sorted in row " + std::to_string(r);
        return false;
      }
    }
  }
  return true;
}

static void build_fact_lds(LuSymbolic& s);

// Stable counting sort of v by key(x) in [0, nkeys).
template <typename Rec, typename Key>
static void counting_sort(std::vector<Rec>& v, std::vector<Rec>& tmp, int64_t nkeys, Key key) {
  if (v.size() < 2 || nkeys <= 1) return;
  std::vector<int64_t> cnt((size_t)nkeys + 1, 0);
  for (const auto& x : v) cnt[(size_t)key(x) + 1]++;
  for (int64_t k = 0; k < nkeys; ++k) cnt[(size_t)k + 1] += cnt[(size_t)k];
  tmp.resize(v.size());
  for (const auto& x : v) tmp[(size_t)cnt[(size_t)key(x)]++] = x;
  v.swap(tmp);
}

static double ms_since(std::chrono::steady_clock::time_point t0) {
  return std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
}

// ---- orderings ----------------------------------------------------------

using Adj = std::vector<std::vector<int32_t>>;

// Symmetric adjacency of A + A^T without the diagonal, sorted and unique.
static void build_graph(int n, const int32_t* indptr, const int32_t* indices, Adj& adj) {
  adj.assign(n, {});
  for (int r = 0; r < n; ++r)
    for (int e = indptr[r]; e < indptr[r + 1]; ++e) {
      int c = indices[e];
      if (c == r) continue;
      adj[r].push_back(c);
      adj[c].push_back(r);
    }
  for (auto& a : adj) {
    std::sort(a.begin(), a.end());
    a.erase(std::unique(a.begin(), a.end()), a.end());
  }
}

// Rows with degree above max(16, 10 sqrt(n)) (supply rails) are taken out of
// the graph; returned in increasing degree order for the end of the ordering.
static std::vector<int32_t> split_dense(int n, Adj& adj) {
  const size_t thresh = std::max<size_t>(16, (size_t)(10.0 * std::sqrt((double)n)));
  std::vector<char> dense(n, 0);
  std::vector<std::pair<size_t, int>> dl;
  for (int v = 0; v < n; ++v)
    if (adj[v].size() > thresh) {
      dense[v] = 1;
      dl.push_back({adj[v].size(), v});
    }
  if (!dl.empty())
    for (int v = 0; v < n; ++v) {
      if (dense[v]) {
        adj[v].clear();
        continue;
      }
      auto& a = adj[v];
      a.erase(std::remove_if(a.begin(), a.end(), [&](int u) { return dense[u] != 0; }), a.end());
    }
  std::sort(dl.begin(), dl.end());
  std::vector<int32_t> out;
  for (auto& d : dl) out.push_back(d.second);
  return out;
}

// Minimum degree with explicit elimination on the subgraph induced by
// `nodes` (local copy of the adjacency); ties broken by vertex id.
static void md_subset(const Adj& adj, const std::vector<int32_t>& nodes, std::vector<int32_t>& loc,
                      std::vector<int32_t>& order) {
  const int k = (int)nodes.size();
  if (k == 0) return;
  for (int i = 0; i < k; ++i) loc[nodes[i]] = i;
  Adj g(k);
  for (int i = 0; i < k; ++i) {
    for (int u : adj[nodes[i]]) {
      int j = loc[u];
      if (j >= 0 && j < k && nodes[j] == u) g[i].push_back(j);
    }
    std::sort(g[i].begin(), g[i].end());
  }
  std::set<std::pair<size_t, int>> heap;
  for (int i = 0; i < k; ++i) heap.insert({g[i].size(), i});
  std::vector<int32_t> nb, merged;
  while (!heap.empty()) {
    int v = heap.begin()->second;
    heap.erase(heap.begin());
    order.push_back(nodes[v]);
    nb.swap(g[v]);
    g[v].clear();
    for (int u : nb) {
      heap.erase({g[u].size(), u});
      merged.clear();
      const auto& au = g[u];
      size_t i = 0, j = 0;
      while (i < au.size() || j < nb.size()) {
        int x;
        if (j >= nb.size() || (i < au.size() && au[i] < nb[j])) {
          x = au[i++];
        } else if (i >= au.size() || nb[j] < au[i]) {
          x = nb[j++];
        } else {
          x = au[i];
          ++i;
          ++j;
        }
        if (x != u && x != v) merged.push_back(x);
      }
      g[u].swap(merged);
      heap.insert({g[u].size(), u});
    }
    nb.clear();
  }
  for (int i = 0; i < k; ++i) loc[nodes[i]] = -1;
}

namespace {

// Nested dissection: vertex separators from BFS level structures, or
// multilevel separators (separator.cpp).
struct Dissector {
  const Adj& adj;
  int leaf;
  bool multilevel;
  std::vector<int32_t> stamp, dist, loc;
  std::vector<int32_t>& order;
  int next_stamp = 1;

  Dissector(const Adj& a, int n, int leaf_, bool ml, std::vector<int32_t>& ord)
      : adj(a), leaf(leaf_), multilevel(ml), stamp(n, 0), dist(n, -1), loc(n, -1), order(ord) {}

  void split(const std::vector<int32_t>& nodes, const std::vector<int8_t>& part) {
    std::vector<int32_t> A, B, S;
    for (size_t i = 0; i < nodes.size(); ++i)
      (part[i] == 0 ? A : part[i] == 1 ? B : S).push_back(nodes[i]);
    if (A.empty() || B.empty()) {
      md_subset(adj, nodes, loc, order);
      return;
    }
    run(A);
    run(B);
    md_subset(adj, S, loc, order);
  }

  // BFS inside the current subset (stamp == s); returns visit order and
  // fills dist.
  void bfs(int s, int start, std::vector<int32_t>& visit) {
    visit.clear();
    visit.push_back(start);
    dist[start] = 0;
    for (size_t h = 0; h < visit.size(); ++h) {
      int v = visit[h];
      for (int u : adj[v])
        if (stamp[u] == s && dist[u] < 0) {
          dist[u] = dist[v] + 1;
          visit.push_back(u);
        }
    }
  }
  void clear_dist(const std::vector<int32_t>& visit) {
    for (int v : visit) dist[v] = -1;
  }

  void run(std::vector<int32_t> nodes) {
    if ((int)nodes.size() <= leaf) {
      md_subset(adj, nodes, loc, order);
      return;
    }
    const int s = next_stamp++;
    for (int v : nodes) stamp[v] = s;
    // connected components
    std::vector<int32_t> visit;
    std::vector<std::vector<int32_t>> comps;
    for (int v : nodes) {
      if (dist[v] >= 0) continue;
      bfs(s, v, visit);
      comps.push_back(visit);
    }
    for (auto& c : comps) clear_dist(c);
    if (comps.size() > 1) {
      for (auto& c : comps) run(c);
      return;
    }
    if (multilevel) {
      std::vector<int8_t> part;
      vertex_separator(adj, nodes, loc, (uint64_t)s, part);
      split(nodes, part);
      return;
    }
    // pseudo-peripheral start: two sweeps to the far end
    int start = nodes[0];
    for (int sweep = 0; sweep < 2; ++sweep) {
      bfs(s, start, visit);
      int far = visit.back();
      const int dmax = dist[far];
      for (int v : visit)
        if (dist[v] == dmax && adj[v].size() < adj[far].size()) far = v;
      clear_dist(visit);
      start = far;
    }
    bfs(s, start, visit);
    const int nlev = dist[visit.back()] + 1;
    std::vector<int32_t> lsize(nlev, 0);
    for (int v : visit) lsize[dist[v]]++;
    if (nlev < 3) {
      clear_dist(visit);
      md_subset(adj, nodes, loc, order);
      return;
    }
    // level where the cumulative count crosses half; then the smallest level
    // within a window around it
    const int total = (int)visit.size();
    int m = 0, cum = 0;
    while (m < nlev - 1 && cum + lsize[m] < total / 2) cum += lsize[m++];
    const int w = std::max(1, nlev / 8);
    int best = m;
    for (int l = std::max(1, m - w); l <= std::min(nlev - 2, m + w); ++l)
      if (lsize[l] < lsize[best] || (lsize[l] == lsize[best] && std::abs(l - m) < std::abs(best - m)))
        best = l;
    if (best < 1) best = 1;
    if (best > nlev - 2) best = nlev - 2;
    std::vector<int32_t> A, B, S;
    for (int v : visit) {
      int d = dist[v];
      if (d < best)
        A.push_back(v);
      else if (d > best)
        B.push_back(v);
      else
        S.push_back(v);
    }
    // thin the separator: a vertex with no neighbour beyond it joins A
    std::vector<int32_t> S2;
    for (int v : S) {
      bool touches_b = false;
      for (int u : adj[v])
        if (stamp[u] == s && dist[u] == best + 1) {
          touches_b = true;
          break;
        }
      if (touches_b)
        S2.push_back(v);
      else
        A.push_back(v);
    }
    clear_dist(visit);
    if (A.empty() || B.empty()) {
      md_subset(adj, nodes, loc, order);
      return;
    }
    run(A);
    run(B);
    // separator last; minimum degree inside it on the separator graph
    md_subset(adj, S2, loc, order);
  }
};

}  // namespace

static void order_columns(int n, const int32_t* indptr, const int32_t* indices, int method,
                          int leaf, std::vector<int32_t>& order, int& n_dense) {
  Adj adj;
  build_graph(n, indptr, indices, adj);
  std::vector<int32_t> dense = split_dense(n, adj);
  n_dense = (int)dense.size();
  std::vector<char> is_dense(n, 0);
  for (int v : dense) is_dense[v] = 1;
  std::vector<int32_t> rest;
  rest.reserve(n);
  for (int v = 0; v < n; ++v)
    if (!is_dense[v]) rest.push_back(v);
  order.clear();
  order.reserve(n);
  if (method == 2 || method == 3) {
    // BFS separators: larger leaves trade a deeper elimination for less fill
    // (on the 65k-gate DAG, n = 130k, leaves of ~1000 rows cut the LU updates
    // by 2.7x against 48-row leaves); small blocks keep 48.  Multilevel
    // separators are narrow enough that 256-row leaves give the fewest
    // updates and levels together (65k-gate DAG: 25.9 M updates in 161
    // levels against 26.9 M in 229 with 1016-row leaves; 64x64 array
    // multiplier: 50.9 M in 345 against 58.5 M in 451)
    const int eff = leaf > 0 ? leaf : method == 3 ? 256 : std::max(48, n / 128);
    Dissector nd(adj, n, std::max(4, eff), method == 3, order);
    nd.run(rest);
  } else {
    std::vector<int32_t> loc(n, -1);
    md_subset(adj, rest, loc, order);
  }
  for (int v : dense) order.push_back(v);
}

namespace {

struct ColStore {
  std::vector<int64_t> ptr{0};
  std::vector<int32_t> idx;
  std::vector<double> val;
};

}  // namespace

// Left-looking LU with threshold partial pivoting (Gilbert-Peierls).
static bool gp_factor(int n, const std::vector<int32_t>& cp, const std::vector<int32_t>& ri,
                      const std::vector<double>& vx, const std::vector<int32_t>& q, double tol,
                      int64_t max_nnz, std::vector<int32_t>& pinv, ColStore& L, ColStore& U,
                      int& n_offdiag, double& min_ratio, std::string& err) {
  pinv.assign(n, -1);
  std::vector<double> x(n, 0.0);
  std::vector<char> mark(n, 0);
  std::vector<int32_t> xi(n), stack(n), pstack(n);
  n_offdiag = 0;
  min_ratio = 1.0;
  for (int k = 0; k < n; ++k) {
    const int col = q[k];
    int top = n;
    // depth-first reach of A(:,col) through the columns of L
    for (int e = cp[col]; e < cp[col + 1]; ++e) {
      int r0 = ri[e];
      if (mark[r0]) continue;
      int head = 0;
      stack[0] = r0;
      while (head >= 0) {
        int r = stack[head];
        int j = pinv[r];
        if (!mark[r]) {
          mark[r] = 1;
          pstack[head] = (j < 0) ? 0 : (int)L.ptr[j];
        }
        bool done = true;
        if (j >= 0) {
          int64_t end = L.ptr[j + 1];
          for (int64_t pp = pstack[head]; pp < end; ++pp) {
            int i = L.idx[pp];
            if (mark[i]) continue;
            pstack[head] = (int)(pp + 1);
            stack[++head] = i;
            done = false;
            break;
          }
        }
        if (done) {
          --head;
          xi[--top] = r;
        }
      }
    }
    for (int e = cp[col]; e < cp[col + 1]; ++e) x[ri[e]] = vx[e];
    for (int t = top; t < n; ++t) {
      int r = xi[t];
      int j = pinv[r];
      if (j < 0) continue;
      double ujk = x[r];
      for (int64_t pp = L.ptr[j]; pp < L.ptr[j + 1]; ++pp) x[L.idx[pp]] -= L.val[pp] * ujk;
    }
    int ipiv = -1;
    double amax = -1.0;
    for (int t = top; t < n; ++t) {
      int r = xi[t];
      if (pinv[r] >= 0) continue;
      double a = std::fabs(x[r]);
      if (a > amax || (a == amax && (ipiv < 0 || r < ipiv))) {
        amax = a;
        ipiv = r;
      }
    }
    if (ipiv < 0) {
      err = "structurally singular matrix at column " + std::to_string(col);
      return false;
    }
    if (pinv[col] < 0 && mark[col] && std::fabs(x[col]) >= tol * amax) ipiv = col;
    if (ipiv != col) ++n_offdiag;
    double piv = x[ipiv];
    if (amax > 0.0) min_ratio = std::min(min_ratio, std::fabs(piv) / amax);
    if (piv == 0.0) min_ratio = 0.0;
    for (int t = top; t < n; ++t) {
      int r = xi[t];
      if (pinv[r] >= 0) {
        U.idx.push_back(pinv[r]);
        U.val.push_back(x[r]);
      }
    }
    U.idx.push_back(k);
    U.val.push_back(piv);
    U.ptr.push_back((int64_t)U.idx.size());
    for (int t = top; t < n; ++t) {
      int r = xi[t];
      if (pinv[r] < 0 && r != ipiv) {
        L.idx.push_back(r);
        L.val.push_back(piv != 0.0 ? x[r] / piv : 0.0);
      }
    }
    L.ptr.push_back((int64_t)L.idx.size());
    if (max_nnz > 0 && (int64_t)(L.idx.size() + U.idx.size()) > max_nnz) {
      err = "LU fill beyond the fill limit of " + std::to_string(max_nnz) + " entries at pivot " +
            std::to_string(k) + " of " + std::to_string(n);
      return false;
    }
    pinv[ipiv] = k;
    for (int t = top; t < n; ++t) {
      x[xi[t]] = 0.0;
      mark[xi[t]] = 0;
    }
  }
  for (auto& r : L.idx) r = pinv[r];
  return true;
}

static void make_groups(LevelSchedule& s, int wide_threshold) {
  s.groups.clear();
  s.max_width = 0;
  s.max_group_levels = 0;
  int nl = (int)s.lvl_ptr.size() - 1;
  int l = 0;
  // a level is launched as a grid when it has many items or more than one
  // long segment (those then run one workgroup each, in parallel)
  // (one workgroup can overlap ~2 waves of short items or 2 rounds of
  // lane groups before a separate grid launch pays off)
  auto is_wide = [&](int lv) {
    const int w = s.lvl_ptr[lv + 1] - s.lvl_ptr[lv];
    const int nh = s.lvl_heavy.empty() ? 0 : s.lvl_heavy[lv];
    const int nm = s.lvl_mid.empty() ? 0 : s.lvl_mid[lv];
    return w >= wide_threshold || nh > 1 || nm > 64;
  };
  while (l < nl) {
    int w = s.lvl_ptr[l + 1] - s.lvl_ptr[l];
    s.max_width = std::max(s.max_width, w);
    if (is_wide(l)) {
      s.groups.push_back({l, l + 1, 1, w});
      ++l;
      continue;
    }
    int l0 = l, items = 0;
    while (l < nl && l - l0 < kMaxNarrowLevels) {
      int ww = s.lvl_ptr[l + 1] - s.lvl_ptr[l];
      if (is_wide(l)) break;
      s.max_width = std::max(s.max_width, ww);
      items += ww;
      ++l;
    }
    s.groups.push_back({l0, l, 0, items});
    s.max_group_levels = std::max(s.max_group_levels, l - l0);
  }
}

// Rows grouped by level; inside a level heavy rows (long segments) first.
template <typename SegLen>
static void rows_by_level(const std::vector<int32_t>& level, SegLen seglen, LevelSchedule& s) {
  int n = (int)level.size();
  int nl = 0;
  for (int v : level) nl = std::max(nl, v + 1);
  s.lvl_ptr.assign(nl + 1, 0);
  for (int v : level) s.lvl_ptr[v + 1]++;
  for (int l = 0; l < nl; ++l) s.lvl_ptr[l + 1] += s.lvl_ptr[l];
  s.items.assign(n, 0);
  s.lvl_heavy.assign(nl, 0);
  s.lvl_mid.assign(nl, 0);
  for (int i = 0; i < n; ++i) {
    const int c = reduction_class(seglen(i));
    if (c == 2) s.lvl_heavy[level[i]]++;
    if (c == 1) s.lvl_mid[level[i]]++;
  }
  std::vector<int32_t> fill[3];
  for (auto& f : fill) f.resize(nl);
  for (int l = 0; l < nl; ++l) {
    fill[2][l] = s.lvl_ptr[l];
    fill[1][l] = s.lvl_ptr[l] + s.lvl_heavy[l];
    fill[0][l] = s.lvl_ptr[l] + s.lvl_heavy[l] + s.lvl_mid[l];
  }
  for (int i = 0; i < n; ++i) {
    const int c = reduction_class(seglen(i));
    s.items[fill[c][level[i]]++] = i;
  }
}

static bool analyze_single(int n, const int32_t* indptr, const int32_t* indices,
                           const double* vals, const AnalyzeOptions& opt, LuSymbolic& s,
                           std::string& err) {
  const auto t0 = std::chrono::steady_clock::now();
  const int nnz = indptr[n];
  s = LuSymbolic();
  s.n = n;
  s.nnz_a = nnz;
  // CSC copy of A (original numbering) with the CSR entry id of every value
  std::vector<int32_t> cp(n + 1, 0), ri(nnz), eid(nnz);
  std::vector<double> vx(nnz);
  for (int e = 0; e < nnz; ++e) cp[indices[e] + 1]++;
  for (int c = 0; c < n; ++c) cp[c + 1] += cp[c];
  {
    std::vector<int32_t> fill(cp.begin(), cp.end() - 1);
    for (int r = 0; r < n; ++r)
      for (int e = indptr[r]; e < indptr[r + 1]; ++e) {
        int c = indices[e];
        int d = fill[c]++;
        ri[d] = r;
        vx[d] = vals[e];
        eid[d] = e;
      }
  }
  if (opt.ordering == 1) {
    s.q.resize(n);
    std::iota(s.q.begin(), s.q.end(), 0);
  } else {
    order_columns(n, indptr, indices, opt.ordering, opt.nd_leaf, s.q, s.n_dense);
  }
  s.analyze_ms[0] = ms_since(t0);
  s.qinv.assign(n, 0);
  for (int k = 0; k < n; ++k) s.qinv[s.q[k]] = k;

  ColStore L, U;
  if (!gp_factor(n, cp, ri, vx, s.q, opt.pivot_tol, opt.max_lu_nnz, s.pinv, L, U,
                 s.n_offdiag_pivots, s.min_pivot_ratio, err))
    return false;
  s.p.assign(n, 0);
  for (int r = 0; r < n; ++r) s.p[s.pinv[r]] = r;
  s.analyze_ms[1] = ms_since(t0) - s.analyze_ms[0];

  // LU pattern (permuted space) in CSR
  std::vector<int32_t> cnt(n + 1, 0);
  for (int k = 0; k < n; ++k) {
    for (int64_t pp = U.ptr[k]; pp < U.ptr[k + 1]; ++pp) cnt[U.idx[pp] + 1]++;
    for (int64_t pp = L.ptr[k]; pp < L.ptr[k + 1]; ++pp) cnt[L.idx[pp] + 1]++;
  }
  s.rowptr.assign(n + 1, 0);
  for (int i = 0; i < n; ++i) s.rowptr[i + 1] = s.rowptr[i] + cnt[i + 1];
  const int64_t nnz_lu = s.rowptr[n];
  if (nnz_lu > (int64_t)INT32_MAX) {
    err = "LU fill exceeds 2^31 entries";
    return false;
  }
  s.col.assign(nnz_lu, 0);
  {
    std::vector<int32_t> fill(s.rowptr.begin(), s.rowptr.end() - 1);
    for (int k = 0; k < n; ++k) {
      for (int64_t pp = U.ptr[k]; pp < U.ptr[k + 1]; ++pp) s.col[fill[U.idx[pp]]++] = k;
      for (int64_t pp = L.ptr[k]; pp < L.ptr[k + 1]; ++pp) s.col[fill[L.idx[pp]]++] = k;
    }
  }
  s.diag.assign(n, -1);
  for (int i = 0; i < n; ++i) {
    std::sort(s.col.begin() + s.rowptr[i], s.col.begin() + s.rowptr[i + 1]);
    for (int e = s.rowptr[i]; e < s.rowptr[i + 1]; ++e)
      if (s.col[e] == i) s.diag[i] = e;
    if (s.diag[i] < 0) {
      err = "internal: LU pattern misses a diagonal";
      return false;
    }
  }
  s.nnz_l = 0;
  for (int i = 0; i < n; ++i) s.nnz_l += s.diag[i] - s.rowptr[i];
  s.nnz_u = nnz_lu - s.nnz_l;
  auto find = [&](int i, int j) -> int32_t {
    auto b = s.col.begin() + s.rowptr[i], e = s.col.begin() + s.rowptr[i + 1];
    auto it = std::lower_bound(b, e, j);
    if (it == e || *it != j) return -1;
    return (int32_t)(it - s.col.begin());
  };
  s.src.assign(nnz_lu, -1);
  for (int c = 0; c < n; ++c)
    for (int d = cp[c]; d < cp[c + 1]; ++d) {
      int32_t pos = find(s.pinv[ri[d]], s.qinv[c]);
      if (pos < 0) {
        err = "internal: A entry outside LU pattern";
        return false;
      }
      s.src[pos] = eid[d];
    }

  s.analyze_ms[2] = ms_since(t0) - s.analyze_ms[0] - s.analyze_ms[1];
  // ---- dense trailing block (band programs only) ----
  int s0 = n;  // first dense pivot; n: none
  if (opt.band_width > 0 && opt.dense_min > 0 && n >= 2 * opt.dense_min) {
    auto trailing = [&](int K) {  // LU entries with row and column >= n - K
      const int lo = n - K;
      int64_t c = 0;
      for (int i = lo; i < n; ++i) {
        auto b = s.col.begin() + s.rowptr[i], e = s.col.begin() + s.rowptr[i + 1];
        c += e - std::lower_bound(b, e, lo);
      }
      return c;
    };
    int best = 0;
    for (int K = opt.dense_min; K <= n / 2; K += std::max(16, K / 8)) {
      if ((double)trailing(K) >= opt.dense_density * (double)K * (double)K)
        best = K;
      else if (best)
        break;
    }
    if (best) s0 = n - best;
  }
  if (s0 < n) {
    const int K = n - s0;
    s.dense_start = s0;
    s.dense_k = K;
    for (int i = s0; i < n; ++i)
      for (int e = s.rowptr[i]; e < s.rowptr[i + 1]; ++e)
        if (s.col[e] >= s0) {
          s.dense_slot.push_back(e);
          s.dense_pos.push_back((i - s0) * K + (s.col[e] - s0));
        }
  }
  // CSC view of the LU pattern: column k -> slots, sorted by row
  std::vector<int32_t> ccnt(n + 1, 0);
  for (int64_t e = 0; e < nnz_lu; ++e) ccnt[s.col[e] + 1]++;
  for (int k = 0; k < n; ++k) ccnt[k + 1] += ccnt[k];
  std::vector<int32_t> cslot(nnz_lu), crow(nnz_lu);
  {
    std::vector<int32_t> fill(ccnt.begin(), ccnt.end() - 1);
    for (int i = 0; i < n; ++i)
      for (int e = s.rowptr[i]; e < s.rowptr[i + 1]; ++e) {
        int d = fill[s.col[e]]++;
        cslot[d] = e;
        crow[d] = i;
      }
  }
  std::vector<char> in_gemm(n, 0);  // pivots whose dense-block updates are a matrix product
  if (s0 < n) {
    // earlier pivots with L rows and U columns inside the dense block
    const int K = n - s0;
    // a pivot whose contribution covers at least 1/32 of the block joins the
    // matrix product; the others update the block's slots by scalar updates
    // in the scalar program (the block is assembled after it)
    std::vector<int32_t> P;
    for (int k = 0; k < s0; ++k) {
      int64_t nl = 0, nu = 0;
      for (int d = ccnt[k]; d < ccnt[k + 1]; ++d) nl += crow[d] >= s0;
      for (int e = s.diag[k] + 1; e < s.rowptr[k + 1]; ++e) nu += s.col[e] >= s0;
      if (nl > 0 && nu > 0 && 32 * nl * nu >= (int64_t)K * K) {
        P.push_back(k);
        in_gemm[k] = 1;
      }
    }
    const int m = (int)P.size();
    s.dense_m = m;
    for (int p = 0; p < m; ++p) {
      const int k = P[p];
      for (int d = ccnt[k]; d < ccnt[k + 1]; ++d)
        if (crow[d] >= s0) {
          s.dl_slot.push_back(cslot[d]);
          s.dl_pos.push_back((crow[d] - s0) * m + p);
        }
      for (int e = s.diag[k] + 1; e < s.rowptr[k + 1]; ++e)
        if (s.col[e] >= s0) {
          s.du_slot.push_back(e);
          s.du_pos.push_back(p * K + (s.col[e] - s0));
        }
    }
    s.dense_flops = 2.0 * K * (double)K * m + 2.0 * K * (double)K * K / 3.0;
  }
  // pivot levels
  std::vector<int32_t> level(n, 0);
  int nlev = 0;
  for (int k = 0; k < n; ++k) {
    int lv = 0;
    for (int e = s.rowptr[k]; e < s.diag[k]; ++e) lv = std::max(lv, level[s.col[e]] + 1);
    for (int d = ccnt[k]; d < ccnt[k + 1]; ++d) {
      int r = crow[d];
      if (r >= k) break;
      lv = std::max(lv, level[r] + 1);
    }
    level[k] = lv;
    nlev = std::max(nlev, lv + 1);
  }
  if (s0 < n) {
    // the scalar program covers the pivots before the dense block only
    nlev = 0;
    for (int k = 0; k < s0; ++k) nlev = std::max(nlev, level[k] + 1);
  }
  s.n_fact_levels = nlev;
  // update contributions grouped by level, sorted by (target slot, pivot)
  std::vector<std::vector<int32_t>> by_level(nlev);
  for (int k = 0; k < s0; ++k) by_level[level[k]].push_back(k);
  // rows and columns of the dense block belong to no band's pivots
  auto lev = [&](int x) { return x >= s0 ? INT32_MAX : level[x]; };
  s.tgt_pos.clear();
  s.tgt_seg.assign(1, 0);
  s.upd_l.clear();
  s.upd_u.clear();
  s.upd_d.clear();
  s.fact.lvl_ptr.assign(1, 0);
  s.fact.items.clear();
  s.flops = 0.0;
  if (opt.band_width > 0) {
    // ---- band program ----
    const int W = std::min(opt.band_width, kMaxBandWidth);
    BandSchedule& B = s.fact_band;
    B = BandSchedule();
    B.band_task.assign(1, 0);
    B.band_hub.assign(1, 0);
    B.hub_cptr.assign(1, 0);
    s.fact_schur_lo.clear();
    s.fact_schur_hi.clear();
    s.fact_schur_heavy.clear();
    s.fact_schur_mid.clear();
    std::vector<int32_t> parent(n), comp_id(n, -1);
    auto root = [&](int v) {
      while (parent[v] != v) {
        parent[v] = parent[parent[v]];
        v = parent[v];
      }
      return v;
    };
    auto first_l = [&](int k) {  // first CSC entry of column k strictly below the diagonal
      int d0 = ccnt[k];
      while (d0 < ccnt[k + 1] && crow[d0] <= k) ++d0;
      return d0;
    };
    // owned: (task, level in band, target, pivot, L slot, U slot); schur: (target, pivot, L, U)
    std::vector<std::array<int32_t, 6>> own, own_tmp;
    std::vector<std::array<int32_t, 4>> sch_tmp;
    std::vector<std::array<int32_t, 4>> sch;
    std::vector<int32_t> piv, comps, task_of;
    std::vector<int64_t> cwork;
    for (int b0 = 0; b0 < nlev; b0 += W) {
      const int b1 = std::min(nlev, b0 + W), w = b1 - b0;
      piv.clear();
      for (int l = b0; l < b1; ++l) piv.insert(piv.end(), by_level[l].begin(), by_level[l].end());
      std::sort(piv.begin(), piv.end());
      for (int k : piv) parent[k] = k;
      for (int k : piv) {
        auto unite = [&](int a) {
          if (level[a] < b0) return;
          const int x = root(k), y = root(a);
          if (x != y) parent[std::max(x, y)] = std::min(x, y);
        };
        for (int e = s.rowptr[k]; e < s.diag[k]; ++e) unite(s.col[e]);  // L(k, k')
        for (int d = ccnt[k]; d < ccnt[k + 1]; ++d) {                      // U(k', k), k' < k
          if (crow[d] >= k) break;
          unite(crow[d]);
        }
      }
      comps.clear();
      cwork.clear();
      for (int k : piv) {
        const int r = root(k);
        if (comp_id[r] < 0) {
          comp_id[r] = (int32_t)comps.size();
          comps.push_back(r);
          cwork.push_back(0);
        }
        const int64_t nlk = ccnt[k + 1] - first_l(k), nuk = s.rowptr[k + 1] - s.diag[k] - 1;
        cwork[comp_id[r]] += nlk * nuk + 8;
      }
      const int nc = (int)comps.size();
      const int ntask = std::min(nc, std::max(1, opt.max_band_tasks));
      {
        std::vector<int32_t> order(nc);
        std::iota(order.begin(), order.end(), 0);
        std::stable_sort(order.begin(), order.end(),
                         [&](int a, int c) { return cwork[a] > cwork[c]; });
        task_of.assign(nc, 0);
        std::vector<std::pair<int64_t, int32_t>> heap;
        for (int t = 0; t < ntask; ++t) heap.push_back({0, t});
        auto cmp = [](const std::pair<int64_t, int32_t>& a, const std::pair<int64_t, int32_t>& c) {
          return a > c;
        };
        std::make_heap(heap.begin(), heap.end(), cmp);
        for (int k : order) {
          std::pop_heap(heap.begin(), heap.end(), cmp);
          task_of[k] = heap.back().second;
          heap.back().first += cwork[k];
          std::push_heap(heap.begin(), heap.end(), cmp);
        }
      }
      own.clear();
      sch.clear();
      for (int k : piv) {
        const int task = task_of[comp_id[root(k)]];
        const int d0 = first_l(k);
        const int nlk = ccnt[k + 1] - d0, nuk = s.rowptr[k + 1] - s.diag[k] - 1;
        s.flops += (double)nlk + 2.0 * (double)nlk * (double)nuk;
        for (int d = d0; d < ccnt[k + 1]; ++d) {
          const int i = crow[d], lpos = cslot[d];
          for (int e = s.diag[k] + 1; e < s.rowptr[k + 1]; ++e) {
            const int j = s.col[e];
            if (i >= s0 && j >= s0 && in_gemm[k]) continue;  // the dense block's product
            const int32_t t = find(i, j);
            if (t < 0) {
              err = "internal: fill closure violated";
              return false;
            }
            // a row or column of a pivot in this band belongs to the pivot's task
            if (lev(i) < b1 || lev(j) < b1)
              own.push_back({task, level[k] - b0, t, k, lpos, e});
            else
              sch.push_back({t, k, lpos, e});
          }
        }
      }
      // each task's pivot columns of L, scaled at the task's end
      {
        std::vector<std::vector<int32_t>> per(ntask);
        for (int k : piv) {
          auto& v = per[task_of[comp_id[root(k)]]];
          for (int d = first_l(k); d < ccnt[k + 1]; ++d) {
            v.push_back(cslot[d]);
            v.push_back(s.diag[k]);
          }
        }
        if (s.task_scale_ptr.empty()) s.task_scale_ptr.push_back(0);
        for (const auto& v : per) {
          s.task_scale.insert(s.task_scale.end(), v.begin(), v.end());
          s.task_scale_ptr.push_back((int32_t)(s.task_scale.size() / 2));
        }
      }
      for (int k : piv) comp_id[root(k)] = -1;
      // Pivots were visited in ascending order, so each target's updates are
      // already ordered by pivot; stable counting sorts group them by target
      // (owned: then by task and level) without disturbing that order.
      counting_sort(own, own_tmp, (int64_t)s.col.size(), [](const std::array<int32_t, 6>& a) {
        return (int64_t)a[2];
      });
      counting_sort(own, own_tmp, (int64_t)ntask * w,
                    [w](const std::array<int32_t, 6>& a) { return (int64_t)a[0] * w + a[1]; });
      counting_sort(sch, sch_tmp, (int64_t)s.col.size(),
                    [](const std::array<int32_t, 4>& a) { return (int64_t)a[0]; });
      // emit one target per (task, level, slot) with its contributions by pivot
      auto push_target = [&](int32_t slot) {
        B.items.push_back((int32_t)s.tgt_pos.size());
        s.tgt_pos.push_back(slot);
      };
      B.band_w.push_back(w);
      B.band_ptr_off.push_back((int32_t)B.tl_ptr.size());
      B.band_hm_off.push_back((int32_t)B.tl_heavy.size());
      size_t a0 = 0;
      std::vector<std::pair<size_t, size_t>> segs;
      for (int task = 0; task < ntask; ++task)
        for (int j = 0; j < w; ++j) {
          B.tl_ptr.push_back((int32_t)B.items.size());
          segs.clear();
          while (a0 < own.size() && own[a0][0] == task && own[a0][1] == j) {
            size_t a1 = a0 + 1;
            while (a1 < own.size() && own[a1][0] == task && own[a1][1] == j &&
                   own[a1][2] == own[a0][2])
              ++a1;
            segs.push_back({a0, a1});
            a0 = a1;
          }
          int nh = 0, nm = 0;
          for (int pass = 2; pass >= 0; --pass)
            for (const auto& sg : segs) {
              const int cls = reduction_class((int)(sg.second - sg.first));
              if (cls != pass) continue;
              nh += cls == 2;
              nm += cls == 1;
              push_target(own[sg.first][2]);
              for (size_t u = sg.first; u < sg.second; ++u) {
                s.upd_l.push_back(own[u][4]);
                s.upd_u.push_back(own[u][5]);
                s.upd_d.push_back(s.diag[own[u][3]]);
              }
              s.tgt_seg.push_back((int32_t)s.upd_l.size());
            }
          B.tl_heavy.push_back(nh);
          B.tl_mid.push_back(nm);
          if (j == w - 1) B.tl_ptr.push_back((int32_t)B.items.size());
        }
      B.band_task.push_back(B.band_task.back() + ntask);
      B.band_hub.push_back(0);
      B.max_tasks = std::max(B.max_tasks, ntask);
      // the band's Schur step: one target per slot, by class
      segs.clear();
      for (size_t c0 = 0; c0 < sch.size();) {
        size_t c1 = c0 + 1;
        while (c1 < sch.size() && sch[c1][0] == sch[c0][0]) ++c1;
        segs.push_back({c0, c1});
        c0 = c1;
      }
      s.fact_schur_lo.push_back((int32_t)B.items.size());
      int nh = 0, nm = 0;
      for (int pass = 2; pass >= 0; --pass)
        for (const auto& sg : segs) {
          const int cls = reduction_class((int)(sg.second - sg.first));
          if (cls != pass) continue;
          nh += cls == 2;
          nm += cls == 1;
          push_target(sch[sg.first][0]);
          for (size_t u = sg.first; u < sg.second; ++u) {
            s.upd_l.push_back(sch[u][2]);
            s.upd_u.push_back(sch[u][3]);
            s.upd_d.push_back(s.diag[sch[u][1]]);
          }
          s.tgt_seg.push_back((int32_t)s.upd_l.size());
        }
      s.fact_schur_hi.push_back((int32_t)B.items.size());
      s.fact_schur_heavy.push_back(nh);
      s.fact_schur_mid.push_back(nm);
    }
  }
  std::vector<std::tuple<int32_t, int32_t, int32_t, int32_t>> ups;
  for (int l = 0; l < nlev && opt.band_width <= 0; ++l) {
    ups.clear();
    for (int k : by_level[l]) {
      // L(:,k): rows > k in column k; U(k,:): cols > k in row k
      int d0 = ccnt[k];
      while (d0 < ccnt[k + 1] && crow[d0] <= k) ++d0;
      int nl = ccnt[k + 1] - d0;
      int nu = s.rowptr[k + 1] - s.diag[k] - 1;
      s.flops += (double)nl + 2.0 * (double)nl * (double)nu;
      for (int d = d0; d < ccnt[k + 1]; ++d) {
        int i = crow[d];
        int lpos = cslot[d];
        for (int e = s.diag[k] + 1; e < s.rowptr[k + 1]; ++e) {
          int j = s.col[e];
          int32_t t = find(i, j);
          if (t < 0) {
            err = "internal: fill closure violated";
            return false;
          }
          ups.emplace_back(t, k, lpos, e);
        }
      }
    }
    std::sort(ups.begin(), ups.end());
    // segments of this level: [begin, end) ranges of ups with equal target
    std::vector<std::pair<size_t, size_t>> segs;
    for (size_t a0 = 0; a0 < ups.size();) {
      size_t a1 = a0 + 1;
      while (a1 < ups.size() && std::get<0>(ups[a1]) == std::get<0>(ups[a0])) ++a1;
      segs.push_back({a0, a1});
      a0 = a1;
    }
    if (segs.empty()) continue;
    int heavy = 0, mid = 0;
    for (int pass = 2; pass >= 0; --pass)
      for (const auto& sg : segs) {
        const int cls = reduction_class((int)(sg.second - sg.first));
        if (cls != pass) continue;
        if (cls == 2) ++heavy;
        if (cls == 1) ++mid;
        s.fact.items.push_back((int32_t)s.tgt_pos.size());
        s.tgt_pos.push_back(std::get<0>(ups[sg.first]));
        for (size_t a1 = sg.first; a1 < sg.second; ++a1) {
          const auto& u = ups[a1];
          s.upd_l.push_back(std::get<2>(u));
          s.upd_u.push_back(std::get<3>(u));
          s.upd_d.push_back(s.diag[std::get<1>(u)]);
        }
        s.tgt_seg.push_back((int32_t)s.upd_l.size());
      }
    s.fact.lvl_ptr.push_back((int32_t)s.fact.items.size());
    s.fact.lvl_heavy.push_back(heavy);
    s.fact.lvl_mid.push_back(mid);
  }
  s.n_updates = (int64_t)s.upd_l.size();
  // L scaling list
  s.l_slot.clear();
  s.l_diag.clear();
  for (int i = 0; i < n; ++i)
    for (int e = s.rowptr[i]; e < s.diag[i]; ++e) {
      s.l_slot.push_back(e);
      s.l_diag.push_back(s.diag[s.col[e]]);
    }
  make_groups(s.fact, opt.wide_threshold);
  build_fact_lds(s);
  s.analyze_ms[3] = ms_since(t0) - s.analyze_ms[0] - s.analyze_ms[1] - s.analyze_ms[2];
  return true;
}

// ---- band programs of the triangular solves ----------------------------------

// level: per row; seg(i) -> (begin, end) of row i's off-diagonal segment in the
// LU arrays; deps(i, f) calls f(j) for every row j that row i reads.
template <typename Seg, typename Deps>
static void build_bands(const std::vector<int32_t>& level, Seg seg, Deps deps, int width,
                        int max_tasks, BandSchedule& B, const std::vector<char>* cut = nullptr) {
  B = BandSchedule();
  const int n = (int)level.size();
  if (n == 0 || width <= 0) return;
  width = std::min(width, kMaxBandWidth);
  int nl = 0;
  for (int v : level) nl = std::max(nl, v + 1);
  std::vector<int32_t> lptr(nl + 1, 0), lrows(n);
  for (int v : level) lptr[v + 1]++;
  for (int l = 0; l < nl; ++l) lptr[l + 1] += lptr[l];
  {
    std::vector<int32_t> fill(lptr.begin(), lptr.end() - 1);
    for (int i = 0; i < n; ++i) lrows[fill[level[i]]++] = i;
  }
  auto seglen = [&](int i) {
    auto r = seg(i);
    return r.second - r.first;
  };
  std::vector<char> has_hub(nl, 0);
  for (int i = 0; i < n; ++i)
    if (seglen(i) > kHubChunk) has_hub[level[i]] = 1;
  std::vector<int32_t> parent(n, -1), comp_id(n, -1);
  auto root = [&](int v) {
    while (parent[v] != v) {
      parent[v] = parent[parent[v]];
      v = parent[v];
    }
    return v;
  };
  B.band_task.assign(1, 0);
  B.band_hub.assign(1, 0);
  B.hub_cptr.assign(1, 0);
  std::vector<int32_t> rows, comps, task_of_comp;
  std::vector<int64_t> comp_work;
  for (int b0 = 0; b0 < nl;) {
    int b1 = std::min(nl, b0 + width);
    for (int l = b0; l < b1; ++l)
      if (has_hub[l] || (cut && l < (int)cut->size() && (*cut)[l])) {
        b1 = l + 1;
        break;
      }
    const int w = b1 - b0;
    // components of the band's non-hub rows
    rows.assign(lrows.begin() + lptr[b0], lrows.begin() + lptr[b1]);
    for (int i : rows) parent[i] = i;
    for (int i : rows) {
      if (seglen(i) > kHubChunk) continue;
      deps(i, [&](int j) {
        if (level[j] >= b0 && seglen(j) <= kHubChunk) {
          const int a = root(i), c = root(j);
          if (a != c) parent[std::max(a, c)] = std::min(a, c);
        }
      });
    }
    comps.clear();
    comp_work.clear();
    for (int i : rows) {
      if (seglen(i) > kHubChunk) continue;
      const int r = root(i);
      if (comp_id[r] < 0 || comp_id[r] >= (int)comps.size() || comps[comp_id[r]] != r) {
        comp_id[r] = (int32_t)comps.size();
        comps.push_back(r);
        comp_work.push_back(0);
      }
      comp_work[comp_id[r]] += seglen(i) + 8;
    }
    // pack components into tasks: longest first onto the least loaded task
    const int nc = (int)comps.size();
    const int ntask = std::min(nc, std::max(1, max_tasks));
    std::vector<int32_t> order(nc);
    std::iota(order.begin(), order.end(), 0);
    std::stable_sort(order.begin(), order.end(),
                     [&](int a, int c) { return comp_work[a] > comp_work[c]; });
    task_of_comp.assign(nc, 0);
    {
      std::vector<std::pair<int64_t, int32_t>> heap;  // (load, task) min-heap
      for (int t = 0; t < ntask; ++t) heap.push_back({0, t});
      auto cmp = [](const std::pair<int64_t, int32_t>& a, const std::pair<int64_t, int32_t>& c) {
        return a > c;
      };
      std::make_heap(heap.begin(), heap.end(), cmp);
      for (int k : order) {
        std::pop_heap(heap.begin(), heap.end(), cmp);
        auto& top = heap.back();
        task_of_comp[k] = top.second;
        top.first += comp_work[k];
        std::push_heap(heap.begin(), heap.end(), cmp);
      }
    }
    // per task, per level, per class: rows in ascending order
    std::vector<std::vector<int32_t>> bucket((size_t)ntask * w * 3);
    std::vector<int32_t> sorted_rows = rows;
    std::sort(sorted_rows.begin(), sorted_rows.end());
    for (int i : sorted_rows) {
      if (seglen(i) > kHubChunk) continue;
      const int t = task_of_comp[comp_id[root(i)]];
      const int cls = reduction_class(seglen(i));
      bucket[((size_t)t * w + (level[i] - b0)) * 3 + cls].push_back(i);
    }
    B.band_w.push_back(w);
    B.band_ptr_off.push_back((int32_t)B.tl_ptr.size());
    B.band_hm_off.push_back((int32_t)B.tl_heavy.size());
    for (int t = 0; t < ntask; ++t) {
      for (int j = 0; j < w; ++j) {
        B.tl_ptr.push_back((int32_t)B.items.size());
        const size_t k = ((size_t)t * w + j) * 3;
        for (int cls = 2; cls >= 0; --cls)
          B.items.insert(B.items.end(), bucket[k + cls].begin(), bucket[k + cls].end());
        B.tl_heavy.push_back((int32_t)bucket[k + 2].size());
        B.tl_mid.push_back((int32_t)bucket[k + 1].size());
      }
      B.tl_ptr.push_back((int32_t)B.items.size());
    }
    B.band_task.push_back(B.band_task.back() + ntask);
    B.max_tasks = std::max(B.max_tasks, ntask);
    // hub rows of this band, finished after its tasks
    for (int i : sorted_rows) {
      if (seglen(i) <= kHubChunk) continue;
      B.hub_rows.push_back(i);
      auto r = seg(i);
      for (int c = r.first; c < r.second; c += kHubChunk) {
        B.hub_cs.push_back(c);
        B.hub_ce.push_back(std::min(r.second, c + kHubChunk));
      }
      B.hub_cptr.push_back((int32_t)B.hub_cs.size());
    }
    B.band_hub.push_back((int32_t)B.hub_rows.size());
    for (int i : rows) comp_id[root(i)] = -1;
    b0 = b1;
  }
}

// Checks that a band program runs every row once and only after the rows it
// reads: earlier bands, or the same task at an earlier level; a hub row after
// every task row of its band.
template <typename Deps>
static bool check_bands(int n, const BandSchedule& B, Deps deps, std::string& err) {
  std::vector<int32_t> band(n, -1), task(n, -1), lvl(n, -1);
  std::vector<char> hub(n, 0);
  const int nb = (int)B.band_w.size();
  for (int b = 0; b < nb; ++b) {
    const int w = B.band_w[b];
    for (int t = B.band_task[b]; t < B.band_task[b + 1]; ++t) {
      const int tl = t - B.band_task[b];
      const int32_t* tp = &B.tl_ptr[B.band_ptr_off[b] + (size_t)tl * (w + 1)];
      for (int j = 0; j < w; ++j)
        for (int g = tp[j]; g < tp[j + 1]; ++g) {
          const int i = B.items[g];
          if (band[i] >= 0) {
            err = "row scheduled twice";
            return false;
          }
          band[i] = b;
          task[i] = t;
          lvl[i] = j;
        }
    }
    for (int h = B.band_hub[b]; h < B.band_hub[b + 1]; ++h) {
      const int i = B.hub_rows[h];
      if (band[i] >= 0) {
        err = "hub row scheduled twice";
        return false;
      }
      band[i] = b;
      hub[i] = 1;
    }
  }
  for (int i = 0; i < n; ++i) {
    if (band[i] < 0) {
      err = "row " + std::to_string(i) + " not scheduled";
      return false;
    }
    bool ok = true;
    deps(i, [&](int j) {
      if (band[j] < band[i]) return;
      if (band[j] > band[i] || hub[j]) {
        ok = false;
        return;
      }
      if (hub[i]) return;  // task rows of the same band run before its hubs
      if (task[j] != task[i] || lvl[j] >= lvl[i]) ok = false;
    });
    if (!ok) {
      err = "row " + std::to_string(i) + " runs before a row it reads";
      return false;
    }
  }
  return true;
}

// LDS form of a solve band: every task fits its band's LDS size, own rows come
// first, and every local index maps back to the row (column) it stands for.
template <typename Seg>
static bool check_tri_lds(const LuSymbolic& s, const BandSchedule& B, Seg seg, std::string& err) {
  for (size_t b = 0; b + 1 < B.band_task.size(); ++b) {
    if (B.band_lds.empty() || B.band_lds[b] <= 0) continue;
    const int w = B.band_w[b];
    for (int t = B.band_task[b]; t < B.band_task[b + 1]; ++t) {
      const int32_t* tk = &B.lds_task[(size_t)t * 4];
      const int32_t* sl = &B.lds_slots[tk[0]];
      if (tk[1] > B.band_lds[b] || tk[2] > tk[1]) {
        err = "solve LDS program: task exceeds its band's LDS size";
        return false;
      }
      const int32_t* tp = &B.tl_ptr[B.band_ptr_off[b] + (size_t)(t - B.band_task[b]) * (w + 1)];
      // entries-in-LDS form: the task's rows are its first tk[2] local rows
      // in item order, their entries one contiguous run of lds_col / lds_src
      if (tp[w] - tp[0] != tk[2] || tk[2] > B.band_rows[b]) {
        err = "solve LDS program: task rows do not match its record";
        return false;
      }
      const int64_t cb = tp[w] > tp[0] ? B.lds_rec[(size_t)tp[0] * 4 + 3] : 0;
      int64_t next = cb;
      for (int g = tp[0]; g < tp[w]; ++g) {
        const int i = B.items[g];
        const int32_t* rec = &B.lds_rec[(size_t)g * 4];
        auto r = seg(i);
        bool ok = rec[0] == g - tp[0] && sl[rec[0]] == i && rec[1] == r.first &&
                  rec[2] == r.second && rec[3] == next;
        for (int e = r.first; ok && e < r.second; ++e) {
          const size_t k = (size_t)rec[3] + (e - r.first);
          const int lc = B.lds_col[k];
          ok = lc < tk[1] && sl[lc] == s.col[e] && B.lds_src[k] == e;
        }
        next += r.second - r.first;
        if (ok && next - cb > B.band_ent[b]) ok = false;
        if (!ok) {
          err = "solve LDS program: row " + std::to_string(i) + " encoded wrongly";
          return false;
        }
      }
    }
  }
  return true;
}

bool lu_check_schedules(const LuSymbolic& s, std::string& err) {
  const int n = s.n;
  if (!check_tri_lds(s, s.fwd_band,
                     [&](int i) { return std::make_pair((int)s.rowptr[i], (int)s.diag[i]); }, err) ||
      !check_tri_lds(s, s.bwd_band,
                     [&](int i) { return std::make_pair((int)s.diag[i] + 1, (int)s.rowptr[i + 1]); },
                     err))
    return false;
  auto fdeps = [&](int i, auto f) {
    for (int e = s.rowptr[i]; e < s.diag[i]; ++e) f(s.col[e]);
  };
  auto bdeps = [&](int i, auto f) {
    for (int e = s.diag[i] + 1; e < s.rowptr[i + 1]; ++e) f(s.col[e]);
  };
  if (!s.fwd_band.band_w.empty() && !check_bands(n, s.fwd_band, fdeps, err)) {
    err = "forward band program: " + err;
    return false;
  }
  if (!s.bwd_band.band_w.empty() && !check_bands(n, s.bwd_band, bdeps, err)) {
    err = "backward band program: " + err;
    return false;
  }
  // refactorisation band program: stage = (band, phase, level) with phase 0 the
  // tasks and 1 the Schur step.  Every update reads slots whose last write is in
  // an earlier stage, and within a band's task phase each slot is written by
  // one task only (tasks of one launch run concurrently).
  const BandSchedule& B = s.fact_band;
  if (B.band_w.empty()) return true;
  // LDS encoding: every local index maps back to the slot the update reads
  for (size_t b = 0; b + 1 < B.band_task.size(); ++b) {
    if (s.band_lds[b] <= 0) continue;
    const int w = B.band_w[b];
    for (int t = B.band_task[b]; t < B.band_task[b + 1]; ++t) {
      const int32_t* tk = &s.lds_task[(size_t)t * 4];
      const int32_t* sl = &s.lds_slots[tk[0]];
      if (tk[1] > s.band_lds[b] || tk[2] > tk[1]) {
        err = "refactorisation LDS program: task exceeds its band's LDS size";
        return false;
      }
      const int32_t* tp = &B.tl_ptr[B.band_ptr_off[b] + (size_t)(t - B.band_task[b]) * (w + 1)];
      for (int g = tp[0]; g < tp[w]; ++g) {
        const int32_t* rec = &s.lds_item[(size_t)g * 4];
        const int tg = B.items[g];
        const int u0 = s.tgt_seg[tg];
        const bool inl = rec[2] == 1;  // single update inline: {packed, pivot}
        bool ok = (inl || rec[0] == tg) && rec[2] == s.tgt_seg[tg + 1] - u0 && rec[3] < tk[2] &&
                  sl[rec[3]] == s.tgt_pos[tg];
        for (int k = 0; ok && k < rec[2]; ++k) {
          const uint32_t a = inl ? (uint32_t)rec[0] : s.lds_upd[(size_t)(rec[1] + k) * 2];
          const uint32_t d = inl ? (uint32_t)rec[1] : s.lds_upd[(size_t)(rec[1] + k) * 2 + 1];
          const int l = (int)(a & 0xffffu), v = (int)(a >> 16);
          ok = l < tk[1] && v < tk[1] && (int)d < tk[1] && sl[l] == s.upd_l[u0 + k] &&
               sl[v] == s.upd_u[u0 + k] && sl[d] == s.upd_d[u0 + k];
        }
        if (!ok) {
          err = "refactorisation LDS program: item " + std::to_string(g) + " encoded wrongly";
          return false;
        }
      }
    }
  }
  const int64_t nnz = (int64_t)s.col.size();
  std::vector<int64_t> last_write(nnz, -1), task_band_write(nnz, -1);
  std::vector<int32_t> writer(nnz, -1), task_writer(nnz, -1);
  auto stage = [](int64_t b, int64_t phase, int64_t j) { return (b * 2 + phase) * 128 + j; };
  std::vector<std::pair<int32_t, int64_t>> items;  // (item, stage) in execution order
  std::vector<int32_t> item_task;
  for (size_t b = 0; b + 1 < B.band_task.size(); ++b) {
    const int w = B.band_w[b];
    for (int j = 0; j < w; ++j)
      for (int t = B.band_task[b]; t < B.band_task[b + 1]; ++t) {
        const int32_t* tp = &B.tl_ptr[B.band_ptr_off[b] + (size_t)(t - B.band_task[b]) * (w + 1)];
        for (int g = tp[j]; g < tp[j + 1]; ++g) {
          items.push_back({B.items[g], stage((int64_t)b, 0, j)});
          item_task.push_back(t);
        }
      }
    for (int g = s.fact_schur_lo[b]; g < s.fact_schur_hi[b]; ++g) {
      items.push_back({B.items[g], stage((int64_t)b, 1, 0)});
      item_task.push_back(-1);
    }
  }
  for (size_t k = 0; k < items.size(); ++k) {  // writes
    const int pos = s.tgt_pos[items[k].first];
    const int64_t st = items[k].second;
    if (item_task[k] >= 0 && last_write[pos] >= 0 && last_write[pos] / 128 == st / 128 &&
        writer[pos] != item_task[k]) {
      err = "refactorisation: two tasks of one launch write the same slot";
      return false;
    }
    last_write[pos] = std::max(last_write[pos], st);
    writer[pos] = item_task[k];
    if (item_task[k] >= 0) {
      task_band_write[pos] = st / 128;
      task_writer[pos] = item_task[k];
    }
  }
  for (size_t k = 0; k < items.size(); ++k) {  // reads see final values
    const int t = items[k].first;
    const int64_t st = items[k].second;
    for (int u = s.tgt_seg[t]; u < s.tgt_seg[t + 1]; ++u)
      for (int32_t slot : {s.upd_l[u], s.upd_u[u], s.upd_d[u]}) {
        if (last_write[slot] >= st) {
          err = "refactorisation: an update reads a slot not final at its stage";
          return false;
        }
        // a slot another task of the same launch writes is not readable
        // (tasks run concurrently; each holds its own copy in LDS)
        if (item_task[k] >= 0 && task_band_write[slot] == st / 128 &&
            task_writer[slot] != item_task[k]) {
          err = "refactorisation: an update reads a slot another task of its launch writes";
          return false;
        }
      }
  }
  // L scaled at the ends of the tasks: every L slot exactly once, by the task
  // that last wrote it in that band (or after all its writes), with its pivot
  // final; task updates read raw L of their own band's pivots, Schur updates
  // read L scaled in their band; the LDS encoding maps back to the slots.
  std::vector<int32_t> scaled_band(nnz, -1);
  for (size_t b = 0; b + 1 < B.band_task.size(); ++b)
    for (int t = B.band_task[b]; t < B.band_task[b + 1]; ++t)
      for (int e = s.task_scale_ptr[t]; e < s.task_scale_ptr[t + 1]; ++e) {
        const int l = s.task_scale[2 * (size_t)e], d = s.task_scale[2 * (size_t)e + 1];
        const int64_t end = stage((int64_t)b, 1, 0);
        bool ok = scaled_band[l] < 0 && last_write[l] < end && last_write[d] < end &&
                  !(task_band_write[l] == end / 128 - 1 && task_writer[l] != t) &&
                  !(task_band_write[d] == end / 128 - 1 && task_writer[d] != t);
        if (ok && s.band_lds[b] > 0) {
          const int32_t* tk = &s.lds_task[(size_t)t * 4];
          const int32_t* sl = &s.lds_slots[tk[0]];
          const uint32_t q = s.lds_scale[e];
          ok = (int)(q & 0xffffu) < tk[2] && (int)(q >> 16) < tk[1] && sl[q & 0xffffu] == l &&
               sl[q >> 16] == d;
        }
        if (!ok) {
          err = "refactorisation: L scaling entry " + std::to_string(e) + " misplaced";
          return false;
        }
        scaled_band[l] = (int32_t)b;
      }
  const int ds = s.dense_start >= 0 ? s.dense_start : s.n;
  for (int i = 0; i < s.n; ++i)
    for (int e = s.rowptr[i]; e < s.diag[i]; ++e)
      if (scaled_band[e] < 0 && !(i >= ds && s.col[e] >= ds)) {
        err = "refactorisation: L slot " + std::to_string(e) + " never scaled";
        return false;
      }
  if (ds < s.n) {
    // dense block: every slot inside it once, none written or scaled by the
    // scalar program; its inputs Lp / Up are final, scaled L and U slots
    const int K = s.dense_k, m = s.dense_m;
    std::vector<char> seen((size_t)K * K, 0);
    size_t expect = 0;
    for (int i = ds; i < s.n; ++i)
      for (int e = s.rowptr[i]; e < s.rowptr[i + 1]; ++e) expect += s.col[e] >= ds;
    bool ok = s.dense_slot.size() == expect && s.dense_pos.size() == expect;
    for (size_t q = 0; ok && q < s.dense_slot.size(); ++q) {
      const int e = s.dense_slot[q], pos = s.dense_pos[q];
      const int i = (int)(std::upper_bound(s.rowptr.begin(), s.rowptr.end(), e) -
                          s.rowptr.begin()) - 1;
      ok = pos >= 0 && pos < K * K && !seen[pos] && i >= ds && s.col[e] >= ds &&
           pos == (i - ds) * K + (s.col[e] - ds) && scaled_band[e] < 0;
      if (ok) seen[pos] = 1;
    }
    for (size_t q = 0; ok && q < s.dl_slot.size(); ++q)
      ok = s.dl_pos[q] >= 0 && s.dl_pos[q] < K * m && scaled_band[s.dl_slot[q]] >= 0;
    for (size_t q = 0; ok && q < s.du_slot.size(); ++q)
      ok = s.du_pos[q] >= 0 && s.du_pos[q] < m * K && s.col[s.du_slot[q]] >= ds;
    if (!ok) {
      err = "refactorisation: dense block maps inconsistent";
      return false;
    }
  }
  for (size_t k = 0; k < items.size(); ++k) {
    const int t = items[k].first;
    const int64_t b = items[k].second / 256;
    for (int u = s.tgt_seg[t]; u < s.tgt_seg[t + 1]; ++u)
      if (scaled_band[s.upd_l[u]] != b) {
        err = "refactorisation: an update reads L scaled in another band";
        return false;
      }
  }
  return true;
}

// LDS form of a solve band program: seg(i) gives row i's off-diagonal entry
// range; a band runs from LDS only if every one of its tasks fits.
template <typename Seg>
static void build_tri_lds(const LuSymbolic& s, Seg seg, BandSchedule& B) {
  B.lds_rec.assign(B.items.size() * 4, 0);
  B.lds_col.clear();
  B.lds_task.clear();
  B.lds_slots.clear();
  B.band_lds.clear();
  B.lds_src.clear();
  B.band_ent.clear();
  B.band_rows.clear();
  std::vector<int32_t> local(s.n, -1), rows;
  for (size_t b = 0; b + 1 < B.band_task.size(); ++b) {
    const int w = B.band_w[b];
    int need = 0, ent_max = 0, rows_max = 0;
    for (int t = B.band_task[b]; t < B.band_task[b + 1]; ++t) {
      const int32_t* tp = &B.tl_ptr[B.band_ptr_off[b] + (size_t)(t - B.band_task[b]) * (w + 1)];
      const size_t col0 = B.lds_col.size();
      rows.clear();
      for (int g = tp[0]; g < tp[w]; ++g) {
        local[B.items[g]] = (int32_t)rows.size();
        rows.push_back(B.items[g]);
      }
      const int nown = (int)rows.size();
      for (int g = tp[0]; g < tp[w]; ++g) {
        auto r = seg(B.items[g]);
        for (int e = r.first; e < r.second; ++e) {
          const int j = s.col[e];
          if (local[j] < 0) {
            local[j] = (int32_t)rows.size();
            rows.push_back(j);
          }
        }
      }
      const int cnt = (int)rows.size();
      const bool fits = cnt <= kLdsSlots;
      need = (need < 0 || !fits) ? -1 : std::max(need, cnt);
      B.lds_task.insert(B.lds_task.end(), {(int32_t)B.lds_slots.size(), cnt, nown, 0});
      B.lds_slots.insert(B.lds_slots.end(), rows.begin(), rows.end());
      for (int g = tp[0]; g < tp[w]; ++g) {
        const int i = B.items[g];
        auto r = seg(i);
        int32_t* rec = &B.lds_rec[(size_t)g * 4];
        rec[0] = fits ? local[i] : 0;
        rec[1] = r.first;
        rec[2] = r.second;
        rec[3] = (int32_t)B.lds_col.size();
        for (int e = r.first; e < r.second; ++e) {
          B.lds_col.push_back(fits ? (uint16_t)local[s.col[e]] : (uint16_t)0);
          B.lds_src.push_back(e);
        }
      }
      ent_max = std::max(ent_max, (int)(B.lds_col.size() - col0));
      rows_max = std::max(rows_max, nown);
      for (int j : rows) local[j] = -1;
    }
    B.band_lds.push_back(need > 0 ? need : 0);
    B.band_ent.push_back(need > 0 ? ent_max : 0);
    B.band_rows.push_back(need > 0 ? rows_max : 0);
  }
}

// Level sets and band programs of both solves from the LU pattern.
static void solve_schedules(LuSymbolic& s, const AnalyzeOptions& opt) {
  const int n = s.n;
  // the solves follow the factors' band width unless given their own
  const int sw = opt.band_width > 0 && opt.solve_band_width > 0 ? opt.solve_band_width
                                                                 : opt.band_width;
  std::vector<int32_t> lf(n, 0), lb(n, 0);
  for (int i = 0; i < n; ++i) {
    int lv = 0;
    for (int e = s.rowptr[i]; e < s.diag[i]; ++e) lv = std::max(lv, lf[s.col[e]] + 1);
    lf[i] = lv;
  }
  for (int i = n - 1; i >= 0; --i) {
    int lv = 0;
    for (int e = s.diag[i] + 1; e < s.rowptr[i + 1]; ++e) lv = std::max(lv, lb[s.col[e]] + 1);
    lb[i] = lv;
  }
  rows_by_level(lf, [&](int i) { return s.diag[i] - s.rowptr[i]; }, s.fwd);
  rows_by_level(lb, [&](int i) { return s.rowptr[i + 1] - s.diag[i] - 1; }, s.bwd);
  auto seg_f = [&](int i) { return std::make_pair((int)s.rowptr[i], (int)s.diag[i]); };
  auto seg_b = [&](int i) { return std::make_pair((int)s.diag[i] + 1, (int)s.rowptr[i + 1]); };
  auto deps_f = [&](int i, auto f) {
    for (int e = s.rowptr[i]; e < s.diag[i]; ++e) f(s.col[e]);
  };
  auto deps_b = [&](int i, auto f) {
    for (int e = s.diag[i] + 1; e < s.rowptr[i + 1]; ++e) f(s.col[e]);
  };
  // A band whose tasks do not fit a workgroup's LDS (FP64 values, entries,
  // records; kSolveLdsBytes) runs from HBM, several times slower: such a
  // band is split at its middle level and the program rebuilt, a few rounds.
  auto lde_fits = [&](const BandSchedule& B, size_t b, bool fwd) {
    if (B.band_lds.empty() || B.band_lds[b] <= 0) return false;
    const size_t nv = (size_t)B.band_lds[b], ne = (size_t)B.band_ent[b],
                 nr = (size_t)B.band_rows[b];
    const size_t rec = ((nv + ne) * 8 + 15) & ~(size_t)15;
    return rec + nr * 16 + (fwd ? 0 : 2 * nr * 8) + 2 * ne <= (size_t)kSolveLdsBytes;
  };
  auto build_split = [&](const std::vector<int32_t>& lv, auto seg, auto deps, bool fwd,
                         BandSchedule& B) {
    int nl = 0;
    for (int v : lv) nl = std::max(nl, v + 1);
    std::vector<char> cut(nl, 0);
    for (int round = 0; round < 6; ++round) {
      build_bands(lv, seg, deps, sw, opt.max_band_tasks, B, &cut);
      build_tri_lds(s, seg, B);
      bool changed = false;
      int l0 = 0;
      for (size_t b = 0; b < B.band_w.size(); ++b) {
        const int w = B.band_w[b];
        if (w > 1 && B.band_task[b + 1] > B.band_task[b] && !lde_fits(B, b, fwd)) {
          cut[l0 + w / 2 - 1] = 1;
          changed = true;
        }
        l0 += w;
      }
      if (!changed) break;
    }
  };
  build_split(lf, seg_f, deps_f, true, s.fwd_band);
  build_split(lb, seg_b, deps_b, false, s.bwd_band);
}

void lu_rebuild_solve_programs(LuSymbolic& s, const AnalyzeOptions& opt) {
  solve_schedules(s, opt);
  make_groups(s.fwd, opt.wide_threshold);
  make_groups(s.bwd, opt.wide_threshold);
}

// ---- block-diagonal systems --------------------------------------------------
//
// A batch of circuits packed into one system (Monte Carlo instances, several
// topologies) has a block-diagonal matrix: one block per connected component
// of the pattern of A + A^T.  Each distinct block pattern is analysed once
// (from the values of its first instance) and the results are laid out block
// after block; the level programs are merged level by level, so one launch
// per level serves every instance.

// Connected components of the pattern of A + A^T, labelled in order of their
// smallest vertex.
static int pattern_components(int n, const int32_t* indptr, const int32_t* indices,
                              std::vector<int32_t>& comp) {
  std::vector<int32_t> parent(n);
  std::iota(parent.begin(), parent.end(), 0);
  auto root = [&](int v) {
    while (parent[v] != v) {
      parent[v] = parent[parent[v]];
      v = parent[v];
    }
    return v;
  };
  for (int r = 0; r < n; ++r)
    for (int e = indptr[r]; e < indptr[r + 1]; ++e) {
      int a = root(r), b = root(indices[e]);
      if (a != b) parent[std::max(a, b)] = std::min(a, b);
    }
  comp.assign(n, -1);
  std::vector<int32_t> label(n, -1);
  int nc = 0;
  for (int v = 0; v < n; ++v) {
    const int r = root(v);
    if (label[r] < 0) label[r] = nc++;
    comp[v] = label[r];
  }
  return nc;
}

struct BlockPattern {
  std::vector<int32_t> verts;            // global vertices, ascending
  std::vector<int32_t> indptr, indices;  // local CSR (same entry order as global)
  uint64_t hash = 0;
};

static void local_pattern(const int32_t* indptr, const int32_t* indices,
                          const std::vector<int32_t>& g2l, BlockPattern& b) {
  b.indptr.assign(1, 0);
  b.indices.clear();
  uint64_t h = 1469598103934665603ull;
  auto mix = [&](uint64_t x) {
    h ^= x;
    h *= 1099511628211ull;
  };
  mix((uint64_t)b.verts.size());
  for (int v : b.verts) {
    for (int e = indptr[v]; e < indptr[v + 1]; ++e) {
      b.indices.push_back(g2l[indices[e]]);
      mix((uint64_t)g2l[indices[e]] + 1);
    }
    b.indptr.push_back((int32_t)b.indices.size());
    mix(0);
  }
  b.hash = h;
}

// Append the level program `src` (item ids offset by `off`) into per-level,
// per-class buckets.
static void bucket_schedule(const LevelSchedule& src, int32_t off,
                            std::vector<std::vector<int32_t>> (&bk)[3]) {
  const int nl = (int)src.lvl_ptr.size() - 1;
  for (auto& b : bk)
    if ((int)b.size() < nl) b.resize(nl);
  for (int l = 0; l < nl; ++l) {
    const int a = src.lvl_ptr[l], h = a + src.lvl_heavy[l], m = h + src.lvl_mid[l];
    for (int i = a; i < h; ++i) bk[2][l].push_back(src.items[i] + off);
    for (int i = h; i < m; ++i) bk[1][l].push_back(src.items[i] + off);
    for (int i = m; i < src.lvl_ptr[l + 1]; ++i) bk[0][l].push_back(src.items[i] + off);
  }
}

static void merge_schedule(std::vector<std::vector<int32_t>> (&bk)[3], LevelSchedule& s) {
  size_t nl = std::max({bk[0].size(), bk[1].size(), bk[2].size()});
  s.items.clear();
  s.lvl_ptr.assign(1, 0);
  s.lvl_heavy.clear();
  s.lvl_mid.clear();
  for (size_t l = 0; l < nl; ++l) {
    size_t before = s.items.size();
    for (int c = 2; c >= 0; --c)
      if (l < bk[c].size()) s.items.insert(s.items.end(), bk[c][l].begin(), bk[c][l].end());
    if (s.items.size() == before) continue;
    s.lvl_heavy.push_back(l < bk[2].size() ? (int32_t)bk[2][l].size() : 0);
    s.lvl_mid.push_back(l < bk[1].size() ? (int32_t)bk[1][l].size() : 0);
    s.lvl_ptr.push_back((int32_t)s.items.size());
  }
}

// LDS form of the refactorisation band tasks (see LuSymbolic::lds_*).  A band
// runs from LDS only if every one of its tasks fits in kLdsSlots values.
static void build_fact_lds(LuSymbolic& s) {
  const BandSchedule& B = s.fact_band;
  s.lds_item.assign(B.items.size() * 4, 0);
  s.lds_upd.clear();
  s.lds_task.clear();
  s.lds_slots.clear();
  s.band_lds.clear();
  s.lds_scale.assign(s.task_scale.size() / 2, 0u);
  std::vector<int32_t> local(s.col.size(), -1), slots;
  for (size_t b = 0; b + 1 < B.band_task.size(); ++b) {
    const int w = B.band_w[b];
    int need = 0;
    for (int t = B.band_task[b]; t < B.band_task[b + 1]; ++t) {
      const int32_t* tp = &B.tl_ptr[B.band_ptr_off[b] + (size_t)(t - B.band_task[b]) * (w + 1)];
      // written slots first, then the operands only read
      slots.clear();
      auto add = [&](int32_t sl) {
        if (local[sl] < 0) {
          local[sl] = (int32_t)slots.size();
          slots.push_back(sl);
        }
      };
      for (int g = tp[0]; g < tp[w]; ++g) add(s.tgt_pos[B.items[g]]);
      // the task's pivot columns of L are scaled in LDS and written back
      const int e0 = s.task_scale_ptr[t], e1 = s.task_scale_ptr[t + 1];
      for (int e = e0; e < e1; ++e) add(s.task_scale[2 * (size_t)e]);
      const int nwrite = (int)slots.size();
      for (int e = e0; e < e1; ++e) add(s.task_scale[2 * (size_t)e + 1]);
      for (int g = tp[0]; g < tp[w]; ++g) {
        const int tg = B.items[g];
        for (int u = s.tgt_seg[tg]; u < s.tgt_seg[tg + 1]; ++u)
          for (int32_t sl : {s.upd_l[u], s.upd_u[u], s.upd_d[u]})
            if (local[sl] < 0) {
              local[sl] = (int32_t)slots.size();
              slots.push_back(sl);
            }
      }
      const int cnt = (int)slots.size();
      const bool fits = cnt <= kLdsSlots;
      if (fits)
        for (int e = e0; e < e1; ++e)
          s.lds_scale[e] = (uint32_t)local[s.task_scale[2 * (size_t)e]] |
                           ((uint32_t)local[s.task_scale[2 * (size_t)e + 1]] << 16);
      need = (need < 0 || !fits) ? -1 : std::max(need, cnt);
      s.lds_task.insert(s.lds_task.end(), {(int32_t)s.lds_slots.size(), cnt, nwrite, 0});
      s.lds_slots.insert(s.lds_slots.end(), slots.begin(), slots.end());
      for (int g = tp[0]; g < tp[w]; ++g) {
        const int tg = B.items[g];
        const int u0 = s.tgt_seg[tg], u1 = s.tgt_seg[tg + 1];
        int32_t* rec = &s.lds_item[(size_t)g * 4];
        rec[0] = tg;
        rec[1] = (int32_t)(s.lds_upd.size() / 2);
        rec[2] = u1 - u0;
        rec[3] = fits ? local[s.tgt_pos[tg]] : 0;
        if (fits && u1 - u0 == 1) {
          // a single update travels in the record itself: {packed L | U << 16, pivot}
          const uint32_t l = (uint32_t)local[s.upd_l[u0]], v = (uint32_t)local[s.upd_u[u0]];
          rec[0] = (int32_t)(l | (v << 16));
          rec[1] = local[s.upd_d[u0]];
          continue;
        }
        for (int u = u0; u < u1; ++u) {
          const uint32_t l = fits ? (uint32_t)local[s.upd_l[u]] : 0u;
          const uint32_t v = fits ? (uint32_t)local[s.upd_u[u]] : 0u;
          const uint32_t d = fits ? (uint32_t)local[s.upd_d[u]] : 0u;
          s.lds_upd.push_back(l | (v << 16));
          s.lds_upd.push_back(d);
        }
      }
      for (int sl : slots) local[sl] = -1;
    }
    s.band_lds.push_back(need > 0 ? need : 0);
  }
}

// Merge the refactorisation band programs of the diagonal blocks (target ids
// offset by each block's first target): band b runs every block's band-b
// tasks in one launch (a narrower block band is padded with empty levels),
// then every block's band-b Schur items, class by class.
static void merge_fact_bands(const std::vector<std::pair<const LuSymbolic*, int32_t>>& parts,
                             const std::vector<int32_t>& slot_off, LuSymbolic& s) {
  BandSchedule& M = s.fact_band;
  M = BandSchedule();
  s.task_scale.clear();
  s.task_scale_ptr.assign(1, 0);
  s.fact_schur_lo.clear();
  s.fact_schur_hi.clear();
  s.fact_schur_heavy.clear();
  s.fact_schur_mid.clear();
  size_t nb = 0;
  for (const auto& p : parts) nb = std::max(nb, p.first->fact_band.band_w.size());
  if (nb == 0) return;
  M.band_task.assign(1, 0);
  M.band_hub.assign(1, 0);
  M.hub_cptr.assign(1, 0);
  for (size_t b = 0; b < nb; ++b) {
    int w = 1;
    for (const auto& p : parts) {
      const BandSchedule& B = p.first->fact_band;
      if (b < B.band_w.size()) w = std::max(w, B.band_w[b]);
    }
    M.band_w.push_back(w);
    M.band_ptr_off.push_back((int32_t)M.tl_ptr.size());
    M.band_hm_off.push_back((int32_t)M.tl_heavy.size());
    int ntask = 0;
    for (size_t pi = 0; pi < parts.size(); ++pi) {
      const auto& p = parts[pi];
      const BandSchedule& B = p.first->fact_band;
      if (b >= B.band_w.size()) continue;
      const int bw = B.band_w[b];
      for (int t = B.band_task[b]; t < B.band_task[b + 1]; ++t, ++ntask) {
        const LuSymbolic& P = *p.first;
        for (int e = P.task_scale_ptr[t]; e < P.task_scale_ptr[t + 1]; ++e) {
          s.task_scale.push_back(P.task_scale[2 * (size_t)e] + slot_off[pi]);
          s.task_scale.push_back(P.task_scale[2 * (size_t)e + 1] + slot_off[pi]);
        }
        s.task_scale_ptr.push_back((int32_t)(s.task_scale.size() / 2));
        const size_t tl = (size_t)(t - B.band_task[b]);
        const int32_t* tp = &B.tl_ptr[B.band_ptr_off[b] + tl * (bw + 1)];
        for (int j = 0; j < w; ++j) {
          M.tl_ptr.push_back((int32_t)M.items.size());
          if (j < bw) {
            for (int g = tp[j]; g < tp[j + 1]; ++g) M.items.push_back(B.items[g] + p.second);
            M.tl_heavy.push_back(B.tl_heavy[B.band_hm_off[b] + tl * bw + j]);
            M.tl_mid.push_back(B.tl_mid[B.band_hm_off[b] + tl * bw + j]);
          } else {
            M.tl_heavy.push_back(0);
            M.tl_mid.push_back(0);
          }
        }
        M.tl_ptr.push_back((int32_t)M.items.size());
      }
    }
    M.band_task.push_back(M.band_task.back() + ntask);
    M.band_hub.push_back(0);
    M.max_tasks = std::max(M.max_tasks, ntask);
    s.fact_schur_lo.push_back((int32_t)M.items.size());
    int nh = 0, nm = 0;
    for (int cls = 2; cls >= 0; --cls)
      for (const auto& p : parts) {
        const LuSymbolic& P = *p.first;
        if (b >= P.fact_band.band_w.size()) continue;
        const int lo = P.fact_schur_lo[b], h = lo + P.fact_schur_heavy[b],
                  m = h + P.fact_schur_mid[b], hi = P.fact_schur_hi[b];
        const int a = cls == 2 ? lo : (cls == 1 ? h : m), e = cls == 2 ? h : (cls == 1 ? m : hi);
        for (int g = a; g < e; ++g) M.items.push_back(P.fact_band.items[g] + p.second);
        if (cls == 2) nh += e - a;
        if (cls == 1) nm += e - a;
      }
    s.fact_schur_hi.push_back((int32_t)M.items.size());
    s.fact_schur_heavy.push_back(nh);
    s.fact_schur_mid.push_back(nm);
  }
}

bool lu_analyze(int n, const int32_t* indptr, const int32_t* indices, const double* vals,
                const AnalyzeOptions& opt, LuSymbolic& s, std::string& err) {
  const int nnz = indptr[n];
  if (!csr_validate(n, nnz, indptr, indices, err)) return false;
  std::vector<int32_t> comp;
  const int nc = n > 0 ? pattern_components(n, indptr, indices, comp) : 0;
  if (nc <= 1) {
    if (!analyze_single(n, indptr, indices, vals, opt, s, err)) return false;
    s.n_blocks = 1;
    s.n_block_patterns = 1;
    const auto t1 = std::chrono::steady_clock::now();
    solve_schedules(s, opt);
    s.analyze_ms[4] = ms_since(t1);
    make_groups(s.fwd, opt.wide_threshold);
    make_groups(s.bwd, opt.wide_threshold);
    return true;
  }
  // blocks and their local patterns
  std::vector<BlockPattern> blk(nc);
  for (int v = 0; v < n; ++v) blk[comp[v]].verts.push_back(v);
  std::vector<int32_t> g2l(n);
  for (auto& b : blk)
    for (size_t i = 0; i < b.verts.size(); ++i) g2l[b.verts[i]] = (int32_t)i;
  for (auto& b : blk) local_pattern(indptr, indices, g2l, b);
  // distinct patterns: analyse the first instance of each
  std::vector<int32_t> kind(nc, -1), rep;
  for (int c = 0; c < nc; ++c) {
    for (size_t k = 0; k < rep.size(); ++k) {
      const BlockPattern& r = blk[rep[k]];
      if (r.hash == blk[c].hash && r.indptr == blk[c].indptr && r.indices == blk[c].indices) {
        kind[c] = (int32_t)k;
        break;
      }
    }
    if (kind[c] < 0) {
      kind[c] = (int32_t)rep.size();
      rep.push_back(c);
    }
  }
  std::vector<LuSymbolic> sym(rep.size());
  for (size_t k = 0; k < rep.size(); ++k) {
    const BlockPattern& b = blk[rep[k]];
    std::vector<double> lv(b.indices.size());
    size_t u = 0;
    for (int v : b.verts)
      for (int e = indptr[v]; e < indptr[v + 1]; ++e) lv[u++] = vals[e];
    // packed batches keep every block on the scalar program (no dense block)
    AnalyzeOptions bopt = opt;
    bopt.dense_min = 0;
    if (!analyze_single((int)b.verts.size(), b.indptr.data(), b.indices.data(), lv.data(), bopt,
                        sym[k], err))
      return false;
  }
  // lay the blocks out one after another
  s = LuSymbolic();
  s.n = n;
  s.nnz_a = nnz;
  s.n_blocks = nc;
  s.n_block_patterns = (int)rep.size();
  int64_t nnz_lu = 0, n_upd = 0, n_tgt = 0, n_l = 0;
  for (int c = 0; c < nc; ++c) {
    const LuSymbolic& b = sym[kind[c]];
    nnz_lu += (int64_t)b.col.size();
    n_upd += (int64_t)b.upd_l.size();
    n_tgt += (int64_t)b.tgt_pos.size();
    n_l += (int64_t)b.l_slot.size();
  }
  if (nnz_lu > (int64_t)INT32_MAX || n_upd > (int64_t)INT32_MAX) {
    err = "LU of the packed system exceeds 2^31 entries";
    return false;
  }
  if (opt.max_lu_nnz > 0 && nnz_lu > opt.max_lu_nnz) {
    err = "LU fill of the packed system beyond the fill limit of " +
          std::to_string(opt.max_lu_nnz) + " entries";
    return false;
  }
  s.q.resize(n);
  s.p.resize(n);
  s.rowptr.assign(1, 0);
  s.col.reserve(nnz_lu);
  s.diag.reserve(n);
  s.src.reserve(nnz_lu);
  s.tgt_pos.reserve(n_tgt);
  s.tgt_seg.assign(1, 0);
  s.upd_l.reserve(n_upd);
  s.upd_u.reserve(n_upd);
  s.upd_d.reserve(n_upd);
  s.l_slot.reserve(n_l);
  s.l_diag.reserve(n_l);
  s.min_pivot_ratio = 1.0;
  std::vector<std::vector<int32_t>> bf[3];
  std::vector<std::pair<const LuSymbolic*, int32_t>> fact_parts;
  std::vector<int32_t> fact_soff;
  int32_t roff = 0, soff = 0, toff = 0, uoff = 0;
  std::vector<int32_t> gmap;
  for (int c = 0; c < nc; ++c) {
    const LuSymbolic& b = sym[kind[c]];
    const BlockPattern& bp = blk[c];
    const int nb = b.n;
    // local A entry -> global A entry
    gmap.resize(bp.indices.size());
    for (int r = 0; r < nb; ++r)
      for (int e = bp.indptr[r]; e < bp.indptr[r + 1]; ++e)
        gmap[e] = indptr[bp.verts[r]] + (e - bp.indptr[r]);
    for (int k = 0; k < nb; ++k) {
      s.q[roff + k] = bp.verts[b.q[k]];
      s.p[roff + k] = bp.verts[b.p[k]];
    }
    for (int r = 0; r < nb; ++r) {
      s.rowptr.push_back(b.rowptr[r + 1] + soff);
      s.diag.push_back(b.diag[r] + soff);
    }
    for (size_t e = 0; e < b.col.size(); ++e) {
      s.col.push_back(b.col[e] + roff);
      s.src.push_back(b.src[e] < 0 ? -1 : gmap[b.src[e]]);
    }
    for (size_t t = 0; t < b.tgt_pos.size(); ++t) {
      s.tgt_pos.push_back(b.tgt_pos[t] + soff);
      s.tgt_seg.push_back(b.tgt_seg[t + 1] + uoff);
    }
    for (size_t u = 0; u < b.upd_l.size(); ++u) {
      s.upd_l.push_back(b.upd_l[u] + soff);
      s.upd_u.push_back(b.upd_u[u] + soff);
      s.upd_d.push_back(b.upd_d[u] + soff);
    }
    for (size_t e = 0; e < b.l_slot.size(); ++e) {
      s.l_slot.push_back(b.l_slot[e] + soff);
      s.l_diag.push_back(b.l_diag[e] + soff);
    }
    bucket_schedule(b.fact, toff, bf);
    fact_parts.push_back({&b, toff});
    fact_soff.push_back(soff);
    s.nnz_l += b.nnz_l;
    s.nnz_u += b.nnz_u;
    s.n_updates += b.n_updates;
    s.flops += b.flops;
    s.n_dense += b.n_dense;
    s.n_fact_levels = std::max(s.n_fact_levels, b.n_fact_levels);
    s.n_offdiag_pivots += b.n_offdiag_pivots;
    s.min_pivot_ratio = std::min(s.min_pivot_ratio, b.min_pivot_ratio);
    roff += nb;
    soff += (int32_t)b.col.size();
    toff += (int32_t)b.tgt_pos.size();
    uoff += (int32_t)b.upd_l.size();
  }
  s.qinv.assign(n, 0);
  s.pinv.assign(n, 0);
  for (int k = 0; k < n; ++k) {
    s.qinv[s.q[k]] = k;
    s.pinv[s.p[k]] = k;
  }
  merge_schedule(bf, s.fact);
  merge_fact_bands(fact_parts, fact_soff, s);
  build_fact_lds(s);
  for (const auto& b : sym)
    for (int k = 0; k < 4; ++k) s.analyze_ms[k] += b.analyze_ms[k];
  const auto t1 = std::chrono::steady_clock::now();
  solve_schedules(s, opt);
  s.analyze_ms[4] = ms_since(t1);
  make_groups(s.fact, opt.wide_threshold);
  make_groups(s.fwd, opt.wide_threshold);
  make_groups(s.bwd, opt.wide_threshold);
  return true;
}

// Fixed-order sums of the GPU kernels: G lanes take strided partial sums,
// then a halving tree (lane t adds lane t + w).  G = 1 is the plain
// sequential sum.  Emulated here so the CPU program matches the device bit
// for bit.
template <int G, typename T, typename F>
static T lane_sum(int start, int end, F term) {
  T part[G];
  for (int t = 0; t < G; ++t) {
    T acc = (T)0;
    for (int u = start + t; u < end; u += G) acc += term(u);
    part[t] = acc;
  }
  for (int w = G / 2; w > 0; w >>= 1)
    for (int t = 0; t < w; ++t) part[t] += part[t + w];
  return part[0];
}

template <typename T, typename F>
static T class_sum(int cls, int start, int end, F term) {
  if (cls == 2) return lane_sum<kReduceLanes, T>(start, end, term);
  if (cls == 1) return lane_sum<kGroupLanes, T>(start, end, term);
  T acc = (T)0;
  for (int u = start; u < end; ++u) acc += term(u);
  return acc;
}

// Sum of one triangular-solve row: by length class; under band programs the
// long rows up to kHubChunk take the one-wavefront sum and longer rows chunk
// sums (workgroup class) added in order.
template <typename T, typename F>
static T row_sum(int start, int end, bool chunk_hubs, F term) {
  const int len = end - start;
  if (chunk_hubs && len > kHeavySegment && len <= kHubChunk)
    return lane_sum<kWaveLanes, T>(start, end, term);
  if (!chunk_hubs || len <= kHubChunk) return class_sum<T>(reduction_class(len), start, end, term);
  T acc = (T)0;
  for (int c = start; c < end; c += kHubChunk)
    acc += class_sum<T>(2, c, std::min(end, c + kHubChunk), term);
  return acc;
}

static int item_class(const LevelSchedule& S, int l, int it) {
  const int lo = S.lvl_ptr[l];
  if (it < lo + S.lvl_heavy[l]) return 2;
  if (it < lo + S.lvl_heavy[l] + S.lvl_mid[l]) return 1;
  return 0;
}

// The dense block on the CPU (FP64, whatever T): assemble, subtract the
// earlier pivots' contributions, blocked LU without pivoting in panels of
// kDensePanel columns, write back.  The same algorithm as the device kernels;
// the sums run in another order, so the two agree to rounding.
template <typename T>
static void dense_block_host(const LuSymbolic& s, T* lu) {
  const int K = s.dense_k, m = s.dense_m;
  std::vector<double> D((size_t)K * K, 0.0), Lp((size_t)K * m, 0.0), Up((size_t)m * K, 0.0);
  for (size_t q = 0; q < s.dense_slot.size(); ++q) D[s.dense_pos[q]] = (double)lu[s.dense_slot[q]];
  for (size_t q = 0; q < s.dl_slot.size(); ++q) Lp[s.dl_pos[q]] = (double)lu[s.dl_slot[q]];
  for (size_t q = 0; q < s.du_slot.size(); ++q) Up[s.du_pos[q]] = (double)lu[s.du_slot[q]];
  for (int i = 0; i < K; ++i)
    for (int p = 0; p < m; ++p) {
      const double l = Lp[(size_t)i * m + p];
      if (l == 0.0) continue;
      for (int j = 0; j < K; ++j) D[(size_t)i * K + j] -= l * Up[(size_t)p * K + j];
    }
  dense_lu_host(K, D.data());
  for (size_t q = 0; q < s.dense_slot.size(); ++q) lu[s.dense_slot[q]] = (T)D[s.dense_pos[q]];
}

void dense_lu_host(int K, double* D) {
  for (int o = 0; o < K; o += kDensePanel) {
    const int w = std::min(kDensePanel, K - o);
    for (int c = o; c < o + w; ++c) {  // the panel's diagonal block
      for (int r = c + 1; r < o + w; ++r) {
        D[(size_t)r * K + c] /= D[(size_t)c * K + c];
        for (int c2 = c + 1; c2 < o + w; ++c2)
          D[(size_t)r * K + c2] -= D[(size_t)r * K + c] * D[(size_t)c * K + c2];
      }
    }
    for (int r = o + w; r < K; ++r)  // L21 = A21 U11^-1
      for (int c = o; c < o + w; ++c) {
        double x = D[(size_t)r * K + c];
        for (int c2 = o; c2 < c; ++c2) x -= D[(size_t)r * K + c2] * D[(size_t)c2 * K + c];
        D[(size_t)r * K + c] = x / D[(size_t)c * K + c];
      }
    for (int j = o + w; j < K; ++j)  // U12 = L11^-1 A12
      for (int r = o; r < o + w; ++r) {
        double y = D[(size_t)r * K + j];
        for (int r2 = o; r2 < r; ++r2) y -= D[(size_t)r * K + r2] * D[(size_t)r2 * K + j];
        D[(size_t)r * K + j] = y;
      }
    for (int r = o + w; r < K; ++r)  // A22 -= L21 U12
      for (int c = o; c < o + w; ++c) {
        const double l = D[(size_t)r * K + c];
        for (int j = o + w; j < K; ++j) D[(size_t)r * K + j] -= l * D[(size_t)c * K + j];
      }
  }
}

template <typename T>
void lu_refactor_host(const LuSymbolic& s, const double* a, T* lu) {
  const int64_t nnz = s.rowptr[s.n];
  for (int64_t e = 0; e < nnz; ++e) lu[e] = s.src[e] >= 0 ? (T)a[s.src[e]] : (T)0;
  const int nl = (int)s.fact.lvl_ptr.size() - 1;
  auto term = [&](int u) { return lu[s.upd_l[u]] * lu[s.upd_u[u]] / lu[s.upd_d[u]]; };
  // Schur updates read L already scaled by its pivot
  auto term_scaled = [&](int u) { return lu[s.upd_l[u]] * lu[s.upd_u[u]]; };
  auto apply = [&](int t) {
    const int a = s.tgt_seg[t], e = s.tgt_seg[t + 1];
    lu[s.tgt_pos[t]] -= class_sum<T>(reduction_class(e - a), a, e, term);
  };
  auto apply_scaled = [&](int t) {
    const int a = s.tgt_seg[t], e = s.tgt_seg[t + 1];
    lu[s.tgt_pos[t]] -= class_sum<T>(reduction_class(e - a), a, e, term_scaled);
  };
  const BandSchedule& B = s.fact_band;
  const bool bands = B.band_task.size() > 1;
  for (size_t b = 0; b + 1 < B.band_task.size(); ++b) {
    // tasks are independent within a band: task by task, level by level,
    // each ending with the scaling of its pivots' L columns
    const int w = B.band_w[b];
    for (int t = B.band_task[b]; t < B.band_task[b + 1]; ++t) {
      const int32_t* tp = &B.tl_ptr[B.band_ptr_off[b] + (size_t)(t - B.band_task[b]) * (w + 1)];
      for (int g = tp[0]; g < tp[w]; ++g) apply(B.items[g]);
      for (int e = s.task_scale_ptr[t]; e < s.task_scale_ptr[t + 1]; ++e)
        lu[s.task_scale[2 * (size_t)e]] /= lu[s.task_scale[2 * (size_t)e + 1]];
    }
    for (int g = s.fact_schur_lo[b]; g < s.fact_schur_hi[b]; ++g) apply_scaled(B.items[g]);
  }
  for (int l = 0; l < nl; ++l)
    for (int it = s.fact.lvl_ptr[l]; it < s.fact.lvl_ptr[l + 1]; ++it) {
      const int t = s.fact.items[it];
      const T acc = class_sum<T>(item_class(s.fact, l, it), s.tgt_seg[t], s.tgt_seg[t + 1], term);
      lu[s.tgt_pos[t]] -= acc;
    }
  if (!bands)
    for (size_t e = 0; e < s.l_slot.size(); ++e) lu[s.l_slot[e]] /= lu[s.l_diag[e]];
  if (s.dense_start >= 0) dense_block_host<T>(s, lu);
}

template <typename T>
void lu_solve_host(const LuSymbolic& s, const T* lu, const double* b, double* x) {
  const int n = s.n;
  std::vector<T> y(n), z(n);
  for (int i = 0; i < n; ++i) y[i] = (T)b[s.p[i]];
  // Each row's sum depends only on its length class (and, under band
  // programs, on the chunking of rows longer than kHubChunk), not on the
  // order rows are visited, so rows are taken in dependency order here.
  const bool chunk_f = !s.fwd_band.band_w.empty(), chunk_b = !s.bwd_band.band_w.empty();
  for (int i = 0; i < n; ++i) {
    auto term = [&](int e) { return lu[e] * y[s.col[e]]; };
    const T acc = row_sum<T>(s.rowptr[i], s.diag[i], chunk_f, term);
    y[i] = y[i] - acc;
  }
  for (int i = n - 1; i >= 0; --i) {
    auto term = [&](int e) { return lu[e] * z[s.col[e]]; };
    const T acc = row_sum<T>(s.diag[i] + 1, s.rowptr[i + 1], chunk_b, term);
    z[i] = (y[i] - acc) / lu[s.diag[i]];
  }
  for (int i = 0; i < n; ++i) x[s.q[i]] = (double)z[i];
}

template void lu_refactor_host<double>(const LuSymbolic&, const double*, double*);
template void lu_refactor_host<float>(const LuSymbolic&, const double*, float*);
template void lu_solve_host<double>(const LuSymbolic&, const double*, const double*, double*);
template void lu_solve_host<float>(const LuSymbolic&, const float*, const double*, double*);

}  // namespace vjx
