// Symbolic LU analysis on the host (see symbolic.h for the contract).
#include "symbolic.h"

#include <algorithm>
#include <cmath>
#include <cstring>
#include <numeric>
#include <set>
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

// Minimum degree ordering on the graph of A + A^T with explicit elimination
// (cliques of eliminated neighbourhoods are materialised).  High-degree rows
// are removed first and appended at the end.
static void min_degree_order(int n, const int32_t* indptr, const int32_t* indices,
                             std::vector<int32_t>& order, int& n_dense) {
  std::vector<std::vector<int32_t>> adj(n);
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
  const size_t thresh = std::max<size_t>(16, (size_t)(10.0 * std::sqrt((double)n)));
  std::vector<char> dense(n, 0);
  std::vector<std::pair<size_t, int>> dense_list;
  n_dense = 0;
  for (int v = 0; v < n; ++v)
    if (adj[v].size() > thresh) {
      dense[v] = 1;
      dense_list.push_back({adj[v].size(), v});
      ++n_dense;
    }
  if (n_dense)
    for (int v = 0; v < n; ++v) {
      if (dense[v]) continue;
      auto& a = adj[v];
      a.erase(std::remove_if(a.begin(), a.end(), [&](int u) { return dense[u] != 0; }), a.end());
    }
  std::set<std::pair<size_t, int>> heap;
  for (int v = 0; v < n; ++v)
    if (!dense[v]) heap.insert({adj[v].size(), v});
  order.clear();
  order.reserve(n);
  std::vector<int32_t> nb, merged;
  while (!heap.empty()) {
    int v = heap.begin()->second;
    heap.erase(heap.begin());
    order.push_back(v);
    nb.swap(adj[v]);
    adj[v].clear();
    for (int u : nb) {
      heap.erase({adj[u].size(), u});
      merged.clear();
      const auto& au = adj[u];
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
      adj[u].swap(merged);
      heap.insert({adj[u].size(), u});
    }
    nb.clear();
  }
  std::sort(dense_list.begin(), dense_list.end());
  for (auto& d : dense_list) order.push_back(d.second);
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
                      std::vector<int32_t>& pinv, ColStore& L, ColStore& U, int& n_offdiag,
                      double& min_ratio, std::string& err) {
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
  int nl = (int)s.lvl_ptr.size() - 1;
  int l = 0;
  while (l < nl) {
    int w = s.lvl_ptr[l + 1] - s.lvl_ptr[l];
    s.max_width = std::max(s.max_width, w);
    if (w >= wide_threshold) {
      s.groups.push_back({l, l + 1, 1, w});
      ++l;
      continue;
    }
    int l0 = l, items = 0;
    while (l < nl) {
      int ww = s.lvl_ptr[l + 1] - s.lvl_ptr[l];
      if (ww >= wide_threshold) break;
      s.max_width = std::max(s.max_width, ww);
      items += ww;
      ++l;
    }
    s.groups.push_back({l0, l, 0, items});
  }
}

static void rows_by_level(const std::vector<int32_t>& level, LevelSchedule& s) {
  int n = (int)level.size();
  int nl = 0;
  for (int v : level) nl = std::max(nl, v + 1);
  s.lvl_ptr.assign(nl + 1, 0);
  for (int v : level) s.lvl_ptr[v + 1]++;
  for (int l = 0; l < nl; ++l) s.lvl_ptr[l + 1] += s.lvl_ptr[l];
  s.items.assign(n, 0);
  std::vector<int32_t> fill(s.lvl_ptr.begin(), s.lvl_ptr.end() - 1);
  for (int i = 0; i < n; ++i) s.items[fill[level[i]]++] = i;
}

bool lu_analyze(int n, const int32_t* indptr, const int32_t* indices, const double* vals,
                const AnalyzeOptions& opt, LuSymbolic& s, std::string& err) {
  const int nnz = indptr[n];
  if (!csr_validate(n, nnz, indptr, indices, err)) return false;
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
    min_degree_order(n, indptr, indices, s.q, s.n_dense);
  }
  s.qinv.assign(n, 0);
  for (int k = 0; k < n; ++k) s.qinv[s.q[k]] = k;

  ColStore L, U;
  if (!gp_factor(n, cp, ri, vx, s.q, opt.pivot_tol, s.pinv, L, U, s.n_offdiag_pivots,
                 s.min_pivot_ratio, err))
    return false;
  s.p.assign(n, 0);
  for (int r = 0; r < n; ++r) s.p[s.pinv[r]] = r;

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
  // update contributions grouped by level, sorted by (target slot, pivot)
  std::vector<std::vector<int32_t>> by_level(nlev);
  for (int k = 0; k < n; ++k) by_level[level[k]].push_back(k);
  s.tgt_pos.clear();
  s.tgt_seg.assign(1, 0);
  s.upd_l.clear();
  s.upd_u.clear();
  s.upd_d.clear();
  s.fact.lvl_ptr.assign(1, 0);
  s.fact.items.clear();
  s.flops = 0.0;
  std::vector<std::tuple<int32_t, int32_t, int32_t, int32_t>> ups;
  for (int l = 0; l < nlev; ++l) {
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
    for (size_t a = 0; a < ups.size(); ++a) {
      const auto& u = ups[a];
      if (a == 0 || std::get<0>(ups[a - 1]) != std::get<0>(u)) {
        if (a != 0) s.tgt_seg.push_back((int32_t)s.upd_l.size());
        s.fact.items.push_back((int32_t)s.tgt_pos.size());
        s.tgt_pos.push_back(std::get<0>(u));
      }
      s.upd_l.push_back(std::get<2>(u));
      s.upd_u.push_back(std::get<3>(u));
      s.upd_d.push_back(s.diag[std::get<1>(u)]);
    }
    if (!ups.empty()) s.tgt_seg.push_back((int32_t)s.upd_l.size());
    if (!ups.empty()) s.fact.lvl_ptr.push_back((int32_t)s.fact.items.size());
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
  // solve level sets
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
  rows_by_level(lf, s.fwd);
  rows_by_level(lb, s.bwd);
  make_groups(s.fact, opt.wide_threshold);
  make_groups(s.fwd, opt.wide_threshold);
  make_groups(s.bwd, opt.wide_threshold);
  return true;
}

template <typename T>
void lu_refactor_host(const LuSymbolic& s, const double* a, T* lu) {
  const int64_t nnz = s.rowptr[s.n];
  for (int64_t e = 0; e < nnz; ++e) lu[e] = s.src[e] >= 0 ? (T)a[s.src[e]] : (T)0;
  const int nl = (int)s.fact.lvl_ptr.size() - 1;
  for (int l = 0; l < nl; ++l)
    for (int it = s.fact.lvl_ptr[l]; it < s.fact.lvl_ptr[l + 1]; ++it) {
      int t = s.fact.items[it];
      T acc = 0;
      for (int u = s.tgt_seg[t]; u < s.tgt_seg[t + 1]; ++u)
        acc += lu[s.upd_l[u]] * lu[s.upd_u[u]] / lu[s.upd_d[u]];
      lu[s.tgt_pos[t]] -= acc;
    }
  for (size_t e = 0; e < s.l_slot.size(); ++e) lu[s.l_slot[e]] /= lu[s.l_diag[e]];
}

template <typename T>
void lu_solve_host(const LuSymbolic& s, const T* lu, const double* b, double* x) {
  const int n = s.n;
  std::vector<T> y(n), z(n);
  for (int i = 0; i < n; ++i) y[i] = (T)b[s.p[i]];
  for (int l = 0; l + 1 < (int)s.fwd.lvl_ptr.size(); ++l)
    for (int it = s.fwd.lvl_ptr[l]; it < s.fwd.lvl_ptr[l + 1]; ++it) {
      int i = s.fwd.items[it];
      T acc = 0;
      for (int e = s.rowptr[i]; e < s.diag[i]; ++e) acc += lu[e] * y[s.col[e]];
      y[i] = y[i] - acc;
    }
  for (int l = 0; l + 1 < (int)s.bwd.lvl_ptr.size(); ++l)
    for (int it = s.bwd.lvl_ptr[l]; it < s.bwd.lvl_ptr[l + 1]; ++it) {
      int i = s.bwd.items[it];
      T acc = 0;
      for (int e = s.diag[i] + 1; e < s.rowptr[i + 1]; ++e) acc += lu[e] * z[s.col[e]];
      z[i] = (y[i] - acc) / lu[s.diag[i]];
    }
  for (int i = 0; i < n; ++i) x[s.q[i]] = (double)z[i];
}

template void lu_refactor_host<double>(const LuSymbolic&, const double*, double*);
template void lu_refactor_host<float>(const LuSymbolic&, const double*, float*);
template void lu_solve_host<double>(const LuSymbolic&, const double*, const double*, double*);
template void lu_solve_host<float>(const LuSymbolic&, const float*, const double*, double*);

}  // namespace vjx
