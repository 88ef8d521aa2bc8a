// C-ABI layer: sparse LU handles, circuit handles, system build.
#include <hip/hip_runtime.h>

#include <chrono>
#include <cstring>
#include <map>
#include <memory>
#include <mutex>
#include <string>
#include <tuple>
#include <vector>

#include "circuit.h"
#include "transient.h"
#include "internal.h"

namespace vjx {

static thread_local std::string g_err;
void set_error(const std::string& msg) { g_err = msg; }

int pattern_upload(vjx_pattern* pat);

}  // namespace vjx

using namespace vjx;

extern "C" {

int vjx_version(void) { return 100; }

const char* vjx_last_error(void) { return g_err.c_str(); }

int vjx_device_count(void) {
  int n = 0;
  if (hipGetDeviceCount(&n) != hipSuccess) return 0;
  return n;
}

int vjx_pattern_create(int32_t n, int32_t nnz, const int32_t* indptr, const int32_t* indices,
                       vjx_pattern** out) {
  if (!out || !indptr || (nnz > 0 && !indices)) {
    set_error("vjx_pattern_create: null argument");
    return VJX_ERR_ARG;
  }
  std::string err;
  if (!csr_validate(n, nnz, indptr, indices, err)) {
    set_error("vjx_pattern_create: " + err);
    return VJX_ERR_ARG;
  }
  auto* p = new vjx_pattern();
  p->n = n;
  p->nnz = nnz;
  p->indptr.assign(indptr, indptr + n + 1);
  p->indices.assign(indices, indices + nnz);
  *out = p;
  return VJX_OK;
}

void vjx_pattern_destroy(vjx_pattern* pat) {
  if (!pat) return;
  if (pat->trans) vjx_pattern_destroy(pat->trans);
  if (pat->stream) (void)hipStreamDestroy(pat->stream);
  delete pat;
}

int vjx_lu_analyze(vjx_pattern* pat, const double* data_host) {
  if (!pat || (!data_host && pat->nnz)) {
    set_error("vjx_lu_analyze: null argument");
    return VJX_ERR_ARG;
  }
  std::string err;
  auto t0 = std::chrono::steady_clock::now();
  // analysed into a fresh record: a failed re-analysis leaves the previous
  // analysis (and the device factors built from it) in effect
  LuSymbolic fresh;
  AnalyzeOptions o = pat->opts;
  const bool nd = o.ordering == 2 || o.ordering == 3;
  // Automatic leaf (multilevel ordering, leaf 0): leaves of 256 rows keep the
  // fill and the update count low on large circuits, but inside a leaf the
  // minimum-degree order of a chain (an RC ladder, a ring) is one pivot level
  // per row, and pivot levels are what the band programs pay for.  Systems
  // of up to kAutoLeafRows rows, whose analysis takes milliseconds, are
  // analysed with both 256- and 4-row leaves, and the program with the lower
  // estimated refactorisation + solve time is kept: 2.8 us per pivot level
  // and 5e-5 us per update (both measured on MI355X: a 202-row RC ladder has
  // 202 levels with 256-row leaves and 9 with 4-row ones, 0.58 against 0.10
  // ms per Newton iteration).
  constexpr int kAutoLeafRows = 4096;
  const bool auto_leaf = o.ordering == 3 && o.nd_leaf == 0 && pat->n <= kAutoLeafRows;
  if (o.ordering == 3 && o.nd_leaf == 0) o.nd_leaf = 256;
  if (!lu_analyze(pat->n, pat->indptr.data(), pat->indices.data(), data_host, o, fresh, err)) {
    set_error("vjx_lu_analyze: " + err);
    if (err.find("fill limit") != std::string::npos) return VJX_ERR_FILL;
    return err.find("singular") != std::string::npos ? VJX_ERR_SINGULAR : VJX_ERR_ARG;
  }
  fresh.nd_leaf = nd ? o.nd_leaf : 0;
  if (auto_leaf) {
    auto cost = [](const LuSymbolic& s) { return 2.8 * s.n_fact_levels + 5e-5 * s.n_updates; };
    AnalyzeOptions o4 = o;
    o4.nd_leaf = 4;
    LuSymbolic alt;
    std::string err4;
    if (lu_analyze(pat->n, pat->indptr.data(), pat->indices.data(), data_host, o4, alt, err4) &&
        cost(alt) < cost(fresh)) {
      alt.nd_leaf = 4;
      fresh = std::move(alt);
    }
  }
  pat->lu.sym = std::move(fresh);
  pat->analyze_seconds =
      std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
  pat->lu.analyzed = true;
  pat->lu.factored_prec = -1;
  pat->uploaded = false;
  pat->n_analyses++;
  return VJX_OK;
}

int vjx_lu_get_info(const vjx_pattern* pat, vjx_lu_info* o) {
  if (!pat || !o) {
    set_error("vjx_lu_get_info: null argument");
    return VJX_ERR_ARG;
  }
  if (!pat->lu.analyzed) {
    set_error("vjx_lu_get_info: no analysis yet");
    return VJX_ERR_STATE;
  }
  const LuSymbolic& s = pat->lu.sym;
  o->n = s.n;
  o->nnz_a = s.nnz_a;
  o->nnz_l = s.nnz_l;
  o->nnz_u = s.nnz_u;
  o->n_updates = s.n_updates;
  o->nd_leaf = s.nd_leaf;
  o->flops = s.flops;
  o->fact_levels = s.n_fact_levels;
  o->fwd_levels = (int32_t)s.fwd.lvl_ptr.size() - 1;
  o->bwd_levels = (int32_t)s.bwd.lvl_ptr.size() - 1;
  if (!s.fact_band.band_w.empty()) {
    int nlaunch = 0;
    for (size_t b = 0; b + 1 < s.fact_band.band_task.size(); ++b)
      nlaunch += (s.fact_band.band_task[b + 1] > s.fact_band.band_task[b]) +
                 (s.fact_schur_hi[b] > s.fact_schur_lo[b]);
    o->fact_launches = nlaunch;
  } else {
    o->fact_launches = (int32_t)s.fact.groups.size();
  }
  o->fwd_launches = (int32_t)s.fwd.groups.size();
  o->bwd_launches = (int32_t)s.bwd.groups.size();
  o->n_dense_rows = s.n_dense;
  o->n_offdiag_pivots = s.n_offdiag_pivots;
  o->min_pivot_ratio = s.min_pivot_ratio;
  o->n_blocks = s.n_blocks;
  o->n_block_patterns = s.n_block_patterns;
  o->fwd_bands = (int32_t)s.fwd_band.band_w.size();
  o->bwd_bands = (int32_t)s.bwd_band.band_w.size();
  o->fwd_hub_rows = (int32_t)s.fwd_band.hub_rows.size();
  o->bwd_hub_rows = (int32_t)s.bwd_band.hub_rows.size();
  for (int k = 0; k < 5; ++k) o->analyze_ms[k] = s.analyze_ms[k];
  o->band_width = pat->opts.band_width;
  o->solve_band_width = pat->opts.solve_band_width;
  o->tuned_solve_ms = pat->tuned_solve_ms;
  o->tuned_factor_ms = pat->tuned_factor_ms;
  o->lu_pipeline = pat->lu.lu_deep;
  o->dense_k = s.dense_k;
  o->dense_m = s.dense_m;
  o->dense_flops = s.dense_flops;
  o->fact_task_updates = o->fact_schur_updates = o->fact_task_targets = o->fact_schur_targets = 0;
  if (!s.fact_band.band_w.empty()) {
    std::vector<char> schur(s.fact_band.items.size(), 0);
    for (size_t b = 0; b < s.fact_schur_lo.size(); ++b)
      for (int g = s.fact_schur_lo[b]; g < s.fact_schur_hi[b]; ++g) schur[g] = 1;
    for (size_t g = 0; g < s.fact_band.items.size(); ++g) {
      const int t = s.fact_band.items[g];
      const int64_t u = s.tgt_seg[t + 1] - s.tgt_seg[t];
      if (schur[g]) {
        o->fact_schur_updates += u;
        o->fact_schur_targets += 1;
      } else {
        o->fact_task_updates += u;
        o->fact_task_targets += 1;
      }
    }
  } else {
    o->fact_task_updates = s.n_updates;
    o->fact_task_targets = (int64_t)s.tgt_pos.size();
  }
  return VJX_OK;
}

int vjx_lu_get_perm(const vjx_pattern* pat, int32_t* p, int32_t* q) {
  if (!pat || !pat->lu.analyzed) {
    set_error("vjx_lu_get_perm: no analysis yet");
    return VJX_ERR_STATE;
  }
  if (p) std::memcpy(p, pat->lu.sym.p.data(), sizeof(int32_t) * pat->n);
  if (q) std::memcpy(q, pat->lu.sym.q.data(), sizeof(int32_t) * pat->n);
  return VJX_OK;
}

int vjx_lu_check_schedules(const vjx_pattern* pat) {
  if (!pat || !pat->lu.analyzed) {
    set_error("vjx_lu_check_schedules: no analysis yet");
    return VJX_ERR_STATE;
  }
  std::string err;
  if (!lu_check_schedules(pat->lu.sym, err)) {
    set_error("vjx_lu_check_schedules: " + err);
    return VJX_ERR_STATE;
  }
  return VJX_OK;
}

int vjx_lu_autotune(vjx_pattern* pat, const double* data_dev) {
  if (!pat || (!data_dev && pat->nnz)) {
    set_error("vjx_lu_autotune: null argument");
    return VJX_ERR_ARG;
  }
  int rc = pattern_ensure_analyzed(pat, data_dev);
  if (rc < 0) return rc;
  rc = pattern_autotune_factor(pat, data_dev);
  if (rc < 0) return rc;
  rc = pattern_autotune_solves(pat, data_dev);
  if (rc < 0) return rc;
  VJX_TRY_HIP(hipStreamSynchronize(pat->stream));
  return VJX_OK;
}

int vjx_pattern_set_band_width(vjx_pattern* pat, int32_t width) {
  if (!pat || width < 0 || width > kMaxBandWidth) {
    set_error("vjx_pattern_set_band_width: width must be in [0, " +
              std::to_string(kMaxBandWidth) + "]");
    return VJX_ERR_ARG;
  }
  if (pat->opts.band_width != width) {
    pat->opts.band_width = width;
    // the schedules change: analyse again on next use
    pat->lu.analyzed = false;
    pat->lu.factored_prec = -1;
    pat->uploaded = false;
  }
  return VJX_OK;
}

int vjx_pattern_set_ordering(vjx_pattern* pat, int32_t ordering, int32_t nd_leaf) {
  const bool nd = ordering == 2 || ordering == 3;
  if (!pat || ordering < 0 || ordering > 3 || (nd && nd_leaf != 0 && nd_leaf < 4)) {
    set_error("vjx_pattern_set_ordering: ordering must be 0 (minimum degree), 1 (natural), "
              "2 or 3 (nested dissection, BFS or multilevel separators; leaf 0 = automatic "
              "or >= 4)");
    return VJX_ERR_ARG;
  }
  if (pat->opts.ordering != ordering || (nd && pat->opts.nd_leaf != nd_leaf)) {
    pat->opts.ordering = ordering;
    if (nd) pat->opts.nd_leaf = nd_leaf;
    pat->lu.analyzed = false;
    pat->lu.factored_prec = -1;
    pat->uploaded = false;
  }
  return VJX_OK;
}

int vjx_pattern_set_fill_limit(vjx_pattern* pat, int64_t max_lu_nnz) {
  if (!pat || max_lu_nnz < 0) {
    set_error("vjx_pattern_set_fill_limit: null pattern or negative limit");
    return VJX_ERR_ARG;
  }
  pat->opts.max_lu_nnz = max_lu_nnz;
  return VJX_OK;
}

int vjx_pattern_set_dense(vjx_pattern* pat, int32_t min_rows, double density) {
  if (!pat || min_rows < 0 || !(density > 0.0 && density <= 1.0)) {
    set_error("vjx_pattern_set_dense: min_rows >= 0 and 0 < density <= 1");
    return VJX_ERR_ARG;
  }
  if (pat->opts.dense_min != min_rows || pat->opts.dense_density != density) {
    pat->opts.dense_min = min_rows;
    pat->opts.dense_density = density;
    pat->lu.analyzed = false;
    pat->lu.factored_prec = -1;
    pat->uploaded = false;
  }
  return VJX_OK;
}

int vjx_lu_get_pattern(const vjx_pattern* pat, int32_t* rowptr, int32_t* col) {
  if (!pat || !pat->lu.analyzed) {
    set_error("vjx_lu_get_pattern: no analysis yet");
    return VJX_ERR_STATE;
  }
  const auto& s = pat->lu.sym;
  if (rowptr) std::memcpy(rowptr, s.rowptr.data(), sizeof(int32_t) * (s.n + 1));
  if (col) std::memcpy(col, s.col.data(), sizeof(int32_t) * s.col.size());
  return VJX_OK;
}

int vjx_lu_refactor_cpu(vjx_pattern* pat, const double* data, int32_t prec, double* lu_out) {
  if (!pat || !data) {
    set_error("vjx_lu_refactor_cpu: null argument");
    return VJX_ERR_ARG;
  }
  if (!pat->lu.analyzed) {
    int rc = vjx_lu_analyze(pat, data);
    if (rc < 0) return rc;
  }
  const auto& s = pat->lu.sym;
  const size_t nnz = s.col.size();
  if (prec == VJX_PREC_F32) {
    std::vector<float> lu(nnz);
    lu_refactor_host<float>(s, data, lu.data());
    if (lu_out)
      for (size_t e = 0; e < nnz; ++e) lu_out[e] = lu[e];
  } else {
    std::vector<double> lu(nnz);
    lu_refactor_host<double>(s, data, lu.data());
    if (lu_out) std::memcpy(lu_out, lu.data(), nnz * sizeof(double));
  }
  return VJX_OK;
}

int vjx_lu_solve_cpu(vjx_pattern* pat, const double* data, int32_t prec, const double* b,
                     double* x) {
  if (!pat || !data || !b || !x) {
    set_error("vjx_lu_solve_cpu: null argument");
    return VJX_ERR_ARG;
  }
  if (!pat->lu.analyzed) {
    int rc = vjx_lu_analyze(pat, data);
    if (rc < 0) return rc;
  }
  const auto& s = pat->lu.sym;
  const size_t nnz = s.col.size();
  if (prec == VJX_PREC_F32) {
    std::vector<float> lu(nnz);
    lu_refactor_host<float>(s, data, lu.data());
    lu_solve_host<float>(s, lu.data(), b, x);
  } else {
    std::vector<double> lu(nnz);
    lu_refactor_host<double>(s, data, lu.data());
    lu_solve_host<double>(s, lu.data(), b, x);
  }
  return VJX_OK;
}

int vjx_lu_refactor(vjx_pattern* pat, const double* data_dev, int32_t prec) {
  if (!pat || (!data_dev && pat->nnz)) {
    set_error("vjx_lu_refactor: null argument");
    return VJX_ERR_ARG;
  }
  if (prec != VJX_PREC_F64 && prec != VJX_PREC_F32) {
    set_error("vjx_lu_refactor: precision must be VJX_PREC_F64 or VJX_PREC_F32");
    return VJX_ERR_ARG;
  }
  int rc = pattern_ensure_analyzed(pat, data_dev);
  if (rc < 0) return rc;
  rc = pattern_refactor_async(pat, data_dev, prec, true);
  if (rc < 0) return rc;
  VJX_TRY_HIP(hipStreamSynchronize(pat->stream));
  return pattern_check_pivots(pat);
}

int vjx_lu_solve(vjx_pattern* pat, const double* data_dev, const double* b_dev, double* x_dev,
                 int32_t refine_steps) {
  if (!pat || !b_dev || !x_dev) {
    set_error("vjx_lu_solve: null argument");
    return VJX_ERR_ARG;
  }
  if (pat->lu.factored_prec < 0) {
    set_error("vjx_lu_solve: call vjx_lu_refactor first");
    return VJX_ERR_STATE;
  }
  if (pat->lu.factored_prec == VJX_PREC_F32 && refine_steps > 0 && !data_dev) {
    set_error("vjx_lu_solve: refinement needs the FP64 matrix");
    return VJX_ERR_ARG;
  }
  int rc = pattern_solve_async(pat, data_dev, b_dev, +1.0, x_dev, refine_steps);
  if (rc < 0) return rc;
  VJX_TRY_HIP(hipStreamSynchronize(pat->stream));
  return VJX_OK;
}

int vjx_spmv(vjx_pattern* pat, const double* data_dev, const double* x_dev, double* y_dev) {
  if (!pat || !data_dev || !x_dev || !y_dev) {
    set_error("vjx_spmv: null argument");
    return VJX_ERR_ARG;
  }
  int rc = pattern_upload_csr(pat);
  if (rc < 0) return rc;
  launch_spmv(pat->n, pat->d_indptr.p, pat->d_indices.p, data_dev, x_dev, nullptr, 1.0, y_dev,
              pat->hub_rows.view(), pat->stream);
  VJX_TRY_HIP(hipGetLastError());
  VJX_TRY_HIP(hipStreamSynchronize(pat->stream));
  return VJX_OK;
}

// A^T through the handle's second pattern: gather the values on the device,
// factorise (analysis on the first call, re-pivot on a pivot warning), solve.
int vjx_lu_solve_transpose(vjx_pattern* pat, const double* data_dev, const double* b_dev,
                           double* x_dev, int32_t prec, int32_t refine_steps) {
  if (!pat || (!data_dev && pat->nnz) || !b_dev || !x_dev) {
    set_error("vjx_lu_solve_transpose: null argument");
    return VJX_ERR_ARG;
  }
  if (prec != VJX_PREC_F64 && prec != VJX_PREC_F32) {
    set_error("vjx_lu_solve_transpose: precision must be VJX_PREC_F64 or VJX_PREC_F32");
    return VJX_ERR_ARG;
  }
  if (!pat->trans) {
    const int n = pat->n, nnz = pat->nnz;
    std::vector<int32_t> tp(n + 1, 0), ti(nnz), src(nnz);
    for (int e = 0; e < nnz; ++e) ++tp[pat->indices[e] + 1];
    for (int j = 0; j < n; ++j) tp[j + 1] += tp[j];
    std::vector<int32_t> next(tp.begin(), tp.end() - 1);
    // rows of A in increasing order: columns of each row of A^T come out sorted
    for (int i = 0; i < n; ++i)
      for (int e = pat->indptr[i]; e < pat->indptr[i + 1]; ++e) {
        const int pos = next[pat->indices[e]]++;
        ti[pos] = i;
        src[pos] = e;
      }
    vjx_pattern* t = nullptr;
    int rc = vjx_pattern_create(n, nnz, tp.data(), ti.data(), &t);
    if (rc < 0) return rc;
    t->opts = pat->opts;
    hipError_t e1 = pat->d_trans_src.upload(src.data(), nnz);
    hipError_t e2 = pat->d_trans_data.alloc(std::max(nnz, 1));
    if (e1 != hipSuccess || e2 != hipSuccess) {
      vjx_pattern_destroy(t);
      VJX_TRY_HIP(e1 != hipSuccess ? e1 : e2);
    }
    pat->trans = t;
  }
  vjx_pattern* t = pat->trans;
  int rc = pattern_stream(t);
  if (rc < 0) return rc;
  launch_gather(pat->nnz, pat->d_trans_src.p, data_dev, pat->d_trans_data.p, t->stream);
  VJX_TRY_HIP(hipGetLastError());
  VJX_TRY_HIP(hipStreamSynchronize(t->stream));
  rc = vjx_lu_refactor(t, pat->d_trans_data.p, prec);
  if (rc < 0) return rc;
  if (rc == VJX_WARN_PIVOT) {
    std::vector<double> h(pat->nnz);
    VJX_TRY_HIP(hipMemcpy(h.data(), pat->d_trans_data.p, sizeof(double) * pat->nnz,
                          hipMemcpyDeviceToHost));
    rc = vjx_lu_analyze(t, h.data());
    if (rc < 0) return rc;
    rc = vjx_lu_refactor(t, pat->d_trans_data.p, prec);
    if (rc < 0) return rc;
  }
  int rc2 = vjx_lu_solve(t, pat->d_trans_data.p, b_dev, x_dev, refine_steps);
  return rc2 < 0 ? rc2 : rc;
}

int vjx_lu_download(vjx_pattern* pat, double* lu_host) {
  if (!pat || !lu_host || pat->lu.factored_prec < 0) {
    set_error("vjx_lu_download: nothing factorised");
    return VJX_ERR_STATE;
  }
  const size_t nnz = pat->lu.sym.col.size();
  if (pat->lu.factored_prec == VJX_PREC_F32) {
    std::vector<float> tmp(nnz);
    VJX_TRY_HIP(hipMemcpy(tmp.data(), pat->lu.lu32.p, nnz * sizeof(float), hipMemcpyDeviceToHost));
    for (size_t e = 0; e < nnz; ++e) lu_host[e] = tmp[e];
  } else {
    VJX_TRY_HIP(hipMemcpy(lu_host, pat->lu.lu64.p, nnz * sizeof(double), hipMemcpyDeviceToHost));
  }
  return VJX_OK;
}

// ---- host-buffer solve with a pattern-keyed cache ----
static std::mutex g_cache_mu;
static std::map<std::tuple<int, int, std::vector<int32_t>, std::vector<int32_t>>,
                std::unique_ptr<vjx_pattern, void (*)(vjx_pattern*)>>
    g_cache;

// op 0: A x = b, op 1: A^T x = b, op 2: y = A x
static int host_op(const char* name, int op, int32_t n, int32_t nnz, const int32_t* indptr,
                   const int32_t* indices, const double* data, const double* b, double* x) {
  if (!indptr || !indices || !data || !b || !x) {
    set_error(std::string(name) + ": null argument");
    return VJX_ERR_ARG;
  }
  std::lock_guard<std::mutex> lock(g_cache_mu);
  auto key = std::make_tuple((int)n, (int)nnz, std::vector<int32_t>(indptr, indptr + n + 1),
                             std::vector<int32_t>(indices, indices + nnz));
  auto it = g_cache.find(key);
  vjx_pattern* pat = nullptr;
  if (it == g_cache.end()) {
    int rc = vjx_pattern_create(n, nnz, indptr, indices, &pat);
    if (rc < 0) return rc;
    // the transposed solve and the product analyse nothing on A itself
    rc = op == 0 ? vjx_lu_analyze(pat, data) : VJX_OK;
    if (rc < 0) {
      vjx_pattern_destroy(pat);
      return rc;
    }
    g_cache.emplace(key, std::unique_ptr<vjx_pattern, void (*)(vjx_pattern*)>(pat, vjx_pattern_destroy));
  } else {
    pat = it->second.get();
  }
  DBuf<double> d_data, d_b, d_x;
  VJX_TRY_HIP(d_data.upload(data, std::max(nnz, 1)));
  VJX_TRY_HIP(d_b.upload(b, std::max(n, 1)));
  VJX_TRY_HIP(d_x.alloc(std::max(n, 1)));
  if (op != 0) {
    int rc = op == 1 ? vjx_lu_solve_transpose(pat, d_data.p, d_b.p, d_x.p, VJX_PREC_F64, 0)
                     : vjx_spmv(pat, d_data.p, d_b.p, d_x.p);
    if (rc < 0) return rc;
    VJX_TRY_HIP(hipMemcpy(x, d_x.p, n * sizeof(double), hipMemcpyDeviceToHost));
    return rc;
  }
  int rc = vjx_lu_refactor(pat, d_data.p, VJX_PREC_F64);
  if (rc < 0) return rc;
  if (rc == VJX_WARN_PIVOT) {
    // pivot order from the first matrix no longer fits: re-pivot on this one
    rc = vjx_lu_analyze(pat, data);
    if (rc < 0) return rc;
    rc = vjx_lu_refactor(pat, d_data.p, VJX_PREC_F64);
    if (rc < 0) return rc;
  }
  int rc2 = vjx_lu_solve(pat, d_data.p, d_b.p, d_x.p, 0);
  if (rc2 < 0) return rc2;
  VJX_TRY_HIP(hipMemcpy(x, d_x.p, n * sizeof(double), hipMemcpyDeviceToHost));
  return rc;
}

int vjx_solve_host(int32_t n, int32_t nnz, const int32_t* indptr, const int32_t* indices,
                   const double* data, const double* b, double* x) {
  return host_op("vjx_solve_host", 0, n, nnz, indptr, indices, data, b, x);
}

int vjx_solve_transpose_host(int32_t n, int32_t nnz, const int32_t* indptr,
                             const int32_t* indices, const double* data, const double* b,
                             double* x) {
  return host_op("vjx_solve_transpose_host", 1, n, nnz, indptr, indices, data, b, x);
}

int vjx_dot_host(int32_t n, int32_t nnz, const int32_t* indptr, const int32_t* indices,
                 const double* data, const double* x, double* y) {
  return host_op("vjx_dot_host", 2, n, nnz, indptr, indices, data, x, y);
}

void vjx_clear_cache(void) {
  std::lock_guard<std::mutex> lock(g_cache_mu);
  g_cache.clear();
}

}  // extern "C"
