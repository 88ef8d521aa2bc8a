// Sparse LU handle helpers: device upload of the symbolic analysis and the
// asynchronous refactor / solve sequences used by the C ABI and the
// transient driver.
#include <cmath>
#include <cstdlib>
#include <cstring>

#include "circuit.h"

namespace vjx {

// Event pairs around individual launches, folded in after a synchronisation.
namespace {
struct KTimer {
  bool on = false;
  std::vector<hipEvent_t> pool;
  size_t used = 0;
  std::vector<std::pair<size_t, int>> pending;  // (start event index, class)
  double ms[KC_COUNT] = {};
  int64_t launches[KC_COUNT] = {};
};
KTimer g_kt;
hipEvent_t kt_event() {
  if (g_kt.used == g_kt.pool.size()) {
    hipEvent_t e = nullptr;
    if (hipEventCreate(&e) != hipSuccess) return nullptr;
    g_kt.pool.push_back(e);
  }
  return g_kt.pool[g_kt.used++];
}
}  // namespace

void ktimer_enable(bool on) { g_kt.on = on; }
bool ktimer_on() { return g_kt.on; }
void ktimer_begin(hipStream_t s) {
  if (!g_kt.on) return;
  hipEvent_t e = kt_event();
  if (e) (void)hipEventRecord(e, s);
}
void ktimer_end(int cls, hipStream_t s) {
  if (!g_kt.on || g_kt.used == 0) return;
  const size_t start = g_kt.used - 1;
  hipEvent_t e = kt_event();
  if (!e) return;
  (void)hipEventRecord(e, s);
  g_kt.pending.push_back({start, cls});
}
void ktimer_collect() {
  for (const auto& p : g_kt.pending) {
    float t = 0.f;
    if (hipEventElapsedTime(&t, g_kt.pool[p.first], g_kt.pool[p.first + 1]) == hipSuccess) {
      g_kt.ms[p.second] += t;
      g_kt.launches[p.second] += 1;
    }
  }
  g_kt.pending.clear();
  g_kt.used = 0;
}
void ktimer_reset() {
  g_kt.pending.clear();
  g_kt.used = 0;
  for (int c = 0; c < KC_COUNT; ++c) {
    g_kt.ms[c] = 0.0;
    g_kt.launches[c] = 0;
  }
}
void ktimer_read(double* ms, int64_t* launches) {
  for (int c = 0; c < KC_COUNT; ++c) {
    ms[c] = g_kt.ms[c];
    launches[c] = g_kt.launches[c];
  }
}

int pattern_stream(vjx_pattern* pat) {
  if (pat->stream) return VJX_OK;
  VJX_TRY_HIP(hipStreamCreateWithFlags(&pat->stream, hipStreamNonBlocking));
  VJX_TRY_HIP(pat->d_status.alloc(1));
  VJX_TRY_HIP(hipHostMalloc((void**)&pat->h_status, sizeof(DevStatus), hipHostMallocDefault));
  std::memset(pat->h_status, 0, sizeof(DevStatus));
  return VJX_OK;
}

int pattern_upload_csr(vjx_pattern* pat) {
  int rc = pattern_stream(pat);
  if (rc < 0) return rc;
  if (pat->d_indptr.p == nullptr) {
    VJX_TRY_HIP(pat->d_indptr.upload(pat->indptr));
    VJX_TRY_HIP(pat->d_indices.upload(pat->indices));
    VJX_TRY_HIP(pat->hub_rows.build(pat->indptr.data(), pat->n, 1));
  }
  return VJX_OK;
}

hipError_t HubPlan::build(const int32_t* seg, int n, int nparts) {
  std::vector<int32_t> hi, hp(1, 0), hcs, hce;
  for (int i = 0; i < n; ++i) {
    const int u0 = seg[i], u1 = seg[i + 1];
    if (u1 - u0 <= kHeavySegment) continue;
    hi.push_back(i);
    for (int c = u0; c < u1; c += kHubChunk) {
      hcs.push_back(c);
      hce.push_back(u1 - c > kHubChunk ? c + kHubChunk : u1);
    }
    hp.push_back((int32_t)hcs.size());
  }
  n_items = (int)hi.size();
  n_chunks = (int)hcs.size();
  hipError_t e = item.upload(hi);
  if (e == hipSuccess) e = ptr.upload(hp);
  if (e == hipSuccess) e = cs.upload(hcs);
  if (e == hipSuccess) e = ce.upload(hce);
  if (e == hipSuccess) e = part.alloc(n_chunks);
  if (e == hipSuccess && nparts > 1) e = part2.alloc(n_chunks);
  if (e == hipSuccess && nparts > 2) e = part3.alloc(n_chunks);
  if (e == hipSuccess && nparts > 4) e = part4.alloc(n_chunks);
  if (e == hipSuccess && nparts > 4) e = part5.alloc(n_chunks);
  return e;
}

static hipError_t upload_sched(DevSchedule& d, const LevelSchedule& h) {
  hipError_t e = d.items.upload(h.items);
  if (e != hipSuccess) return e;
  e = d.lvl_dev.upload(h.lvl_ptr);
  if (e != hipSuccess) return e;
  e = d.heavy_dev.upload(h.lvl_heavy);
  if (e != hipSuccess) return e;
  e = d.mid_dev.upload(h.lvl_mid);
  d.lvl_ptr = h.lvl_ptr;
  d.lvl_heavy = h.lvl_heavy;
  d.lvl_mid = h.lvl_mid;
  d.groups = h.groups;
  return e;
}

static hipError_t upload_band(DevBand& d, const BandSchedule& h) {
  d.on = !h.band_w.empty();
  if (!d.on) return hipSuccess;
  hipError_t e = d.items.upload(h.items);
  if (e == hipSuccess) e = d.tl_ptr.upload(h.tl_ptr);
  if (e == hipSuccess) e = d.tl_heavy.upload(h.tl_heavy);
  if (e == hipSuccess) e = d.tl_mid.upload(h.tl_mid);
  if (e == hipSuccess) e = d.hub_rows.upload(h.hub_rows);
  if (e == hipSuccess) e = d.hub_cptr.upload(h.hub_cptr);
  if (e == hipSuccess) e = d.hub_cs.upload(h.hub_cs);
  if (e == hipSuccess) e = d.hub_ce.upload(h.hub_ce);
  if (e == hipSuccess) e = d.part.alloc(h.hub_cs.size());
  if (e == hipSuccess) e = d.lds_rec.upload(h.lds_rec);
  if (e == hipSuccess) e = d.lds_col.upload(h.lds_col);
  if (e == hipSuccess) e = d.lds_task.upload(h.lds_task);
  if (e == hipSuccess) e = d.lds_slots.upload(h.lds_slots);
  if (e == hipSuccess) e = d.lds_src.upload(h.lds_src);
  return e;
}

int pattern_upload(vjx_pattern* pat) {
  if (pat->uploaded) return VJX_OK;
  int rc = pattern_upload_csr(pat);
  if (rc < 0) return rc;
  DevLU& L = pat->lu;
  const LuSymbolic& s = L.sym;
  VJX_TRY_HIP(L.rowptr.upload(s.rowptr));
  VJX_TRY_HIP(L.col.upload(s.col));
  VJX_TRY_HIP(L.diag.upload(s.diag));
  VJX_TRY_HIP(L.src.upload(s.src));
  VJX_TRY_HIP(L.tgt_pos.upload(s.tgt_pos));
  VJX_TRY_HIP(L.tgt_seg.upload(s.tgt_seg));
  VJX_TRY_HIP(L.upd_l.upload(s.upd_l));
  VJX_TRY_HIP(L.upd_u.upload(s.upd_u));
  VJX_TRY_HIP(L.upd_d.upload(s.upd_d));
  VJX_TRY_HIP(L.l_slot.upload(s.l_slot));
  VJX_TRY_HIP(L.l_diag.upload(s.l_diag));
  VJX_TRY_HIP(L.task_scale.upload(s.task_scale));
  VJX_TRY_HIP(L.task_scale_ptr.upload(s.task_scale_ptr));
  VJX_TRY_HIP(L.lds_scale.upload(s.lds_scale));
  VJX_TRY_HIP(L.dense_slot.upload(s.dense_slot));
  VJX_TRY_HIP(L.dense_pos.upload(s.dense_pos));
  VJX_TRY_HIP(L.dl_slot.upload(s.dl_slot));
  VJX_TRY_HIP(L.dl_pos.upload(s.dl_pos));
  VJX_TRY_HIP(L.du_slot.upload(s.du_slot));
  VJX_TRY_HIP(L.du_pos.upload(s.du_pos));
  VJX_TRY_HIP(L.dense_D.alloc((size_t)s.dense_k * s.dense_k));
  VJX_TRY_HIP(L.dense_Lp.alloc((size_t)s.dense_k * s.dense_m));
  VJX_TRY_HIP(L.dense_Up.alloc((size_t)s.dense_k * s.dense_m));
  if (s.fact_band.band_w.empty()) {
    L.upd_lu.release();
  } else {
    // the Schur steps read both slots of an update with one 8-byte load
    std::vector<int32_t> lu2(2 * s.upd_l.size());
    for (size_t u = 0; u < s.upd_l.size(); ++u) {
      lu2[2 * u] = s.upd_l[u];
      lu2[2 * u + 1] = s.upd_u[u];
    }
    VJX_TRY_HIP(L.upd_lu.upload(lu2));
  }
  VJX_TRY_HIP(L.p.upload(s.p));
  VJX_TRY_HIP(L.q.upload(s.q));
  VJX_TRY_HIP(upload_sched(L.fact, s.fact));
  VJX_TRY_HIP(upload_sched(L.fwd, s.fwd));
  VJX_TRY_HIP(upload_sched(L.bwd, s.bwd));
  VJX_TRY_HIP(upload_band(L.fwd_band, s.fwd_band));
  VJX_TRY_HIP(upload_band(L.bwd_band, s.bwd_band));
  VJX_TRY_HIP(upload_band(L.fact_band, s.fact_band));
  VJX_TRY_HIP(L.lds_item.upload(s.lds_item));
  VJX_TRY_HIP(L.lds_upd.upload(s.lds_upd));
  VJX_TRY_HIP(L.lds_task.upload(s.lds_task));
  VJX_TRY_HIP(L.lds_slots.upload(s.lds_slots));
  // VJX_LU_LDS=0 keeps every LU band task in HBM (A/B measurements)
  const char* lds_env = std::getenv("VJX_LU_LDS");
  L.use_lds = !(lds_env && lds_env[0] == '0');
  // VJX_LU_DEEP=1/2 forces the shallow / deep LU band pipeline (A/B
  // measurements); by default bands with more than two tasks per CU take the
  // shallow one
  // (when unset the pipeline stays what the autotuner or the default chose)
  const char* deep_env = std::getenv("VJX_LU_DEEP");
  if (deep_env && deep_env[0] >= '0' && deep_env[0] <= '2') L.lu_deep = deep_env[0] - '0';
  // Solve band tasks run from LDS when everything they read fits the
  // device's LDS per workgroup (less the kernel's static arrays), else from
  // HBM; VJX_SOLVE_LDS=0 keeps every task in HBM (A/B measurements)
  const char* tri_env = std::getenv("VJX_SOLVE_LDS");
  L.fwd_band.use_lds = L.bwd_band.use_lds = !(tri_env && tri_env[0] == '0');
  int dev = 0, lds_max = 0;
  VJX_TRY_HIP(hipGetDevice(&dev));
  VJX_TRY_HIP(hipDeviceGetAttribute(&lds_max, hipDeviceAttributeMaxSharedMemoryPerBlock, dev));
  VJX_TRY_HIP(hipDeviceGetAttribute(&L.n_cu, hipDeviceAttributeMultiprocessorCount, dev));
  L.fwd_band.lde_bytes = L.bwd_band.lde_bytes = std::max(0, lds_max - kTriStaticLds);
  const size_t nnz = s.col.size();
  const size_t n = (size_t)s.n;
  VJX_TRY_HIP(L.lu64.alloc(nnz));
  VJX_TRY_HIP(L.lu32.alloc(nnz));
  VJX_TRY_HIP(L.y64.alloc(n));
  VJX_TRY_HIP(L.z64.alloc(n));
  VJX_TRY_HIP(L.r64.alloc(n));
  VJX_TRY_HIP(L.x64.alloc(n));
  VJX_TRY_HIP(L.y32.alloc(n));
  VJX_TRY_HIP(L.z32.alloc(n));
  pat->uploaded = true;
  return VJX_OK;
}

int pattern_ensure_analyzed(vjx_pattern* pat, const double* data_dev) {
  int rc = pattern_stream(pat);
  if (rc < 0) return rc;
  if (!pat->lu.analyzed) {
    std::vector<double> h(pat->nnz);
    if (pat->nnz)
      VJX_TRY_HIP(hipMemcpy(h.data(), data_dev, sizeof(double) * pat->nnz, hipMemcpyDeviceToHost));
    rc = vjx_lu_analyze(pat, h.data());
    if (rc < 0) return rc;
  }
  return pattern_upload(pat);
}

int pattern_refactor_async(vjx_pattern* pat, const double* data_dev, int prec, bool reset_status) {
  DevLU& L = pat->lu;
  if (reset_status) launch_status_reset(pat->d_status.p, pat->stream);
  if (prec == VJX_PREC_F32)
    launch_lu_refactor<float>(L, data_dev, L.lu32.p, pat->d_status.p, pat->stream);
  else
    launch_lu_refactor<double>(L, data_dev, L.lu64.p, pat->d_status.p, pat->stream);
  VJX_TRY_HIP(hipGetLastError());
  L.factored_prec = prec;
  return VJX_OK;
}

// Autotuner of the solve launch plan: rebuilds the two solves' band programs
// for each candidate band width, times the solves on the device with HIP
// events and keeps the fastest.  Results do not depend on the width (each
// row's sum is fixed), only the number of launches and the work per task.
int pattern_autotune_solves(vjx_pattern* pat, const double* data_dev) {
  DevLU& L = pat->lu;
  if (!L.analyzed || pat->opts.band_width <= 0) return VJX_OK;
  int rc = pattern_ensure_analyzed(pat, data_dev);
  if (rc < 0) return rc;
  hipStream_t s = pat->stream;
  const int prec = L.factored_prec < 0 ? VJX_PREC_F64 : L.factored_prec;
  rc = pattern_refactor_async(pat, data_dev, prec, true);
  if (rc < 0) return rc;
  const int n = pat->n;
  VJX_TRY_HIP(hipMemsetAsync(L.r64.p, 0, sizeof(double) * n, s));
  hipEvent_t ev[2];
  VJX_TRY_HIP(hipEventCreate(&ev[0]));
  VJX_TRY_HIP(hipEventCreate(&ev[1]));
  auto solve = [&]() {
    if (prec == VJX_PREC_F32)
      launch_lu_solve<float>(L, L.lu32.p, L.r64.p, 1.0, L.x64.p, L.y32.p, L.z32.p, s);
    else
      launch_lu_solve<double>(L, L.lu64.p, L.r64.p, 1.0, L.x64.p, L.y64.p, L.z64.p, s);
  };
  const int candidates[] = {8, 12, 16, 24, 32, 48, 64};
  int best = 0;
  float best_ms = 0.f;
  rc = VJX_OK;
  for (int w : candidates) {
    pat->opts.solve_band_width = w;
    lu_rebuild_solve_programs(L.sym, pat->opts);
    hipError_t e = upload_band(L.fwd_band, L.sym.fwd_band);
    if (e == hipSuccess) e = upload_band(L.bwd_band, L.sym.bwd_band);
    if (e != hipSuccess) {
      set_error(std::string("vjx_lu_autotune: upload failed: ") + hipGetErrorString(e));
      rc = VJX_ERR_HIP;
      break;
    }
    solve();  // warm-up
    (void)hipEventRecord(ev[0], s);
    for (int r = 0; r < 5; ++r) solve();
    (void)hipEventRecord(ev[1], s);
    e = hipEventSynchronize(ev[1]);
    float ms = 0.f;
    if (e == hipSuccess) e = hipEventElapsedTime(&ms, ev[0], ev[1]);
    if (e != hipSuccess) {
      set_error(std::string("vjx_lu_autotune: timing failed: ") + hipGetErrorString(e));
      rc = VJX_ERR_HIP;
      break;
    }
    if (best == 0 || ms < best_ms) {
      best = w;
      best_ms = ms;
    }
  }
  (void)hipEventDestroy(ev[0]);
  (void)hipEventDestroy(ev[1]);
  if (rc < 0) return rc;
  pat->opts.solve_band_width = best;
  lu_rebuild_solve_programs(L.sym, pat->opts);
  VJX_TRY_HIP(upload_band(L.fwd_band, L.sym.fwd_band));
  VJX_TRY_HIP(upload_band(L.bwd_band, L.sym.bwd_band));
  pat->tuned_solve_ms = best_ms / 5.0;
  return VJX_OK;
}

// Autotuner of the refactorisation's launch plan: times the numeric
// refactorisation on the device for each LU band-task pipeline (chosen per
// band by its task count, shallow and low-register, or deep prefetch: the same
// sums in the same order) and for refactorisation bands of 12, 16 and 24
// pivot levels (each a fresh analysis of the same matrix: the same ordering
// and pivots, a different split of the updates between band tasks and Schur
// steps, which changes the factors at rounding level), and keeps the fastest
// pair.
int pattern_autotune_factor(vjx_pattern* pat, const double* data_dev) {
  DevLU& L = pat->lu;
  int rc = pattern_ensure_analyzed(pat, data_dev);
  if (rc < 0) return rc;
  if (pat->opts.band_width <= 0) return VJX_OK;
  hipStream_t s = pat->stream;
  const int prec = L.factored_prec < 0 ? VJX_PREC_F64 : L.factored_prec;
  hipEvent_t ev[2];
  VJX_TRY_HIP(hipEventCreate(&ev[0]));
  VJX_TRY_HIP(hipEventCreate(&ev[1]));
  // the fastest of three rounds of five refactorisations (the minimum is the
  // figure least disturbed by other work on the device)
  auto time_refactor = [&](float& ms) -> int {
    int r = pattern_refactor_async(pat, data_dev, prec, true);  // warm-up
    if (r < 0) return r;
    ms = -1.f;
    for (int round = 0; round < 3; ++round) {
      VJX_TRY_HIP(hipEventRecord(ev[0], s));
      for (int k = 0; k < 5; ++k) {
        r = pattern_refactor_async(pat, data_dev, prec, true);
        if (r < 0) return r;
      }
      VJX_TRY_HIP(hipEventRecord(ev[1], s));
      VJX_TRY_HIP(hipEventSynchronize(ev[1]));
      float t = 0.f;
      VJX_TRY_HIP(hipEventElapsedTime(&t, ev[0], ev[1]));
      if (ms < 0.f || t < ms) ms = t;
    }
    ms /= 5.0f;
    return VJX_OK;
  };
  std::vector<double> host(pat->nnz);
  if (pat->nnz)
    VJX_TRY_HIP(hipMemcpy(host.data(), data_dev, sizeof(double) * pat->nnz, hipMemcpyDeviceToHost));
  const int width0 = pat->opts.band_width;
  const int deep0 = L.lu_deep;
  int best_w = width0, best_deep = deep0;
  float best_ms = -1.f;
  int widths[] = {width0, 12, 16, 24};
  int n_widths = 4;
  // VJX_LU_BANDS=w pins the band width (A/B measurements and profiling
  // passes that must see the plan of another run); the pipeline is still timed
  if (const char* env = std::getenv("VJX_LU_BANDS")) {
    const int w = std::atoi(env);
    if (w > 0 && w <= kMaxBandWidth) {
      widths[0] = w;
      n_widths = 1;
    }
  }
  for (int wi = 0; wi < n_widths && rc >= 0; ++wi) {
    const int w = widths[wi];
    if (wi > 0 && w == width0) continue;
    if (w != pat->opts.band_width) {
      pat->opts.band_width = w;
      rc = vjx_lu_analyze(pat, host.data());
      if (rc < 0) break;
      rc = pattern_upload(pat);
      if (rc < 0) break;
    }
    // the current plan (timed first) is kept unless another is at least 10 %
    // faster: plans closer than that trade places from run to run with the
    // timing noise, and another band width also changes the order of the
    // Schur sums (the rounding of the factors)
    for (int i = 0; i < 3 && rc >= 0; ++i) {
      const int deep = (deep0 + i) % 3;
      L.lu_deep = deep;
      float ms = 0.f;
      rc = time_refactor(ms);
      if (rc >= 0 && (best_ms < 0.f || ms < 0.9f * best_ms)) {
        best_ms = ms;
        best_w = w;
        best_deep = deep;
      }
    }
  }
  (void)hipEventDestroy(ev[0]);
  (void)hipEventDestroy(ev[1]);
  if (rc < 0) return rc;
  if (pat->opts.band_width != best_w) {
    pat->opts.band_width = best_w;
    rc = vjx_lu_analyze(pat, host.data());
    if (rc < 0) return rc;
    rc = pattern_upload(pat);
    if (rc < 0) return rc;
  }
  L.lu_deep = best_deep;
  pat->tuned_factor_ms = best_ms;
  return VJX_OK;
}

int pattern_check_pivots(vjx_pattern* pat) {
  VJX_TRY_HIP(hipMemcpy(pat->h_status, pat->d_status.p, sizeof(DevStatus), hipMemcpyDeviceToHost));
  double maxl;
  std::memcpy(&maxl, &pat->h_status->maxl_bits, sizeof(double));
  if (pat->h_status->nonfinite_in > 0) {
    set_error("vjx_lu_refactor: " + std::to_string(pat->h_status->nonfinite_in) +
              " non-finite matrix entries");
    return VJX_WARN_PIVOT;
  }
  if (pat->h_status->nonfinite > 0 || maxl > 1e12) {
    set_error("vjx_lu_refactor: pivot growth (max |L| = " + std::to_string(maxl) +
              ", non-finite = " + std::to_string(pat->h_status->nonfinite) +
              "); call vjx_lu_analyze to re-pivot");
    return VJX_WARN_PIVOT;
  }
  return VJX_OK;
}

int pattern_solve_async(vjx_pattern* pat, const double* data_dev, const double* rhs, double scale,
                        double* x, int refine_steps) {
  DevLU& L = pat->lu;
  hipStream_t s = pat->stream;
  const int n = pat->n;
  if (L.factored_prec == VJX_PREC_F64) {
    launch_lu_solve<double>(L, L.lu64.p, rhs, scale, x, L.y64.p, L.z64.p, s);
  } else {
    launch_lu_solve<float>(L, L.lu32.p, rhs, scale, x, L.y32.p, L.z32.p, s);
    if (refine_steps > 0) {
      int rc = pattern_upload_csr(pat);
      if (rc < 0) return rc;
    }
    for (int r = 0; r < refine_steps; ++r) {
      // residual of A x = scale*rhs, in the reference's form
      // (solver_factories.py:911-921): r = rhs + A x for scale -1.
      launch_spmv(n, pat->d_indptr.p, pat->d_indices.p, data_dev, x, rhs, -scale, L.r64.p,
                  pat->hub_rows.view(), s);
      launch_lu_solve<float>(L, L.lu32.p, L.r64.p, scale, L.x64.p, L.y32.p, L.z32.p, s);
      launch_axpy(n, x, L.x64.p, s);
    }
  }
  VJX_TRY_HIP(hipGetLastError());
  return VJX_OK;
}

}  // namespace vjx
