// Transient driver: Newton iterations and the adaptive / fixed-step time loop.
//
// All vectors (solution, charges, histories, Jacobian, LU factors) stay in
// HBM.  The host runs only the scalar logic of the reference's step control
// -- integration coefficients, predictor weights, accept/reject and the next
// dt (vajax/analysis/transient/full_mna.py:1840-2194,
// vajax/analysis/transient/adaptive.py:94-339) -- in the same FP64
// arithmetic, reading back one small status record per Newton iteration.
#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstring>
#include <limits>
#include <vector>

#include "circuit.h"
#include "transient.h"

namespace vjx {

namespace {

bool g_profile = false;
vjx_phase_times g_phase = {};

double bits_to_double(unsigned long long b) {
  double d;
  std::memcpy(&d, &b, sizeof(d));
  return d;
}

struct Work {
  DBuf<double> Xa, Xw, Xsave, Vp, Qp, dQ, Q2, Qn, dQn, mrc, hmrc, floor, Vmax, f, csr, delta, vsrc, isrc,
      Ivs, hist, vout, iout, zeros;
  // limit states (limited devices): at the step's start, of the current
  // Newton iterate, written by the last build, saved for a re-pivot redo
  DBuf<double> ls_step, ls_iter, ls_out, ls_save;
  DBuf<DevStatus> st, st_lte;
  // [0] Newton status, [1] LU status, [2] LTE status: mapped pinned host
  // memory the iteration's last kernels write (d_hst: its device address)
  DevStatus* h_st = nullptr;
  DevStatus* d_hst = nullptr;
  hipError_t alloc_status() {
    hipError_t e = hipHostMalloc((void**)&h_st, 3 * sizeof(DevStatus), hipHostMallocMapped);
    if (e != hipSuccess) return e;
    std::memset(h_st, 0, 3 * sizeof(DevStatus));
    return hipHostGetDevicePointer((void**)&d_hst, h_st, 0);
  }
  ~Work() {
    if (h_st) (void)hipHostFree(h_st);
  }
};

struct Driver {
  vjx_circuit* c;
  const vjx_tran_config* cfg;
  hipStream_t s;
  Work w;
  int n_total, n_unk, m, k, n_aug;
  int n_lim = 0;
  int64_t total_evals = 0;
  int repivots = 0;
  // L + U entries of the first analysis the monitor replaces; re-analyses run
  // under a fill limit of four times this
  int64_t fill_base = 0;
  double gmin_dev = -1.0;  // device simparam $gmin; < 0: the options' gmin

  // Pinned staging of the scalar uploads, two records used in turn: the
  // step loop does not wait for a step's last upload (its final build's)
  // before the next step writes the next one; every Newton iteration
  // synchronises, so a record is free again two uploads later.
  DevScalars* h_sc2 = nullptr;
  int sc_flip = 0;
  // the record on the device: a Newton solve's iterations and its final
  // build use the same scalars (the device reads no iteration count), so
  // only the first of them is uploaded
  DevScalars sc_dev{};
  bool sc_dev_valid = false;
  int upload_scalars(double c0, double c1, double d1, double c2, double gshunt, int it) {
    if (!h_sc2) {
      VJX_TRY_HIP(hipHostMalloc((void**)&h_sc2, 2 * sizeof(DevScalars), hipHostMallocDefault));
      std::memset(h_sc2, 0, 2 * sizeof(DevScalars));
    }
    DevScalars v{};
    std::memset(&v, 0, sizeof(v));
    v.c0 = c0;
    v.c1 = c1;
    v.c2 = c2;
    v.d1 = d1;
    v.gmin = cfg->gmin;
    v.gmin_dev = gmin_dev < 0.0 ? cfg->gmin : gmin_dev;
    v.gshunt = gshunt;
    v.reltol = cfg->reltol;
    v.vntol = cfg->vntol;
    v.abstol = cfg->abstol;
    v.nr_damping = cfg->nr_damping;
    v.max_step = 1e30;
    v.iteration = 0;  // it: the Newton iteration, informational only
    (void)it;
    if (sc_dev_valid && std::memcmp(&v, &sc_dev, sizeof(v)) == 0) return VJX_OK;
    sc_dev = v;
    sc_dev_valid = true;
    DevScalars* h = h_sc2 + sc_flip;
    sc_flip ^= 1;
    *h = v;
    VJX_TRY_HIP(hipMemcpyAsync(c->d_sc.p, h, sizeof(DevScalars), hipMemcpyHostToDevice, s));
    return VJX_OK;
  }
  // source values at t, evaluated on the device (no host round trip)
  int set_sources(double t) {
    launch_sources2(c->vsrc_table(), w.vsrc.p, c->isrc_table(), w.isrc.p, t, s);
    VJX_TRY_HIP(hipGetLastError());
    return VJX_OK;
  }

  // LTE of the step (adaptive.py:221-339), queued behind every Newton
  // iteration of a transient step and exported with its status: the
  // iteration that converges (or the last one) leaves the step's LTE norm in
  // h_st[2] at the same synchronisation, so the step needs no round trip of
  // its own.  It reads only the iterate, the predictor and V_max, none of
  // which the step's final build changes.
  bool lte_hook = false;
  double lte_scale = 0.0;
  void enqueue_lte() {
    if (!lte_hook) return;
    // the iteration's first kernel cleared st_lte
    launch_lte(n_total, w.Xw.p, w.Vp.p, w.Vmax.p, lte_scale, cfg->reltol, cfg->abstol,
               w.st_lte.p, s);
    launch_status_export(w.st_lte.p, nullptr, w.d_hst + 2, s);
  }

  hipGraph_t graph = nullptr;
  hipGraphExec_t graph_exec = nullptr;
  const double* graph_args[4] = {nullptr, nullptr, nullptr, nullptr};
  int graph_launches = 0;

  void drop_graph() {
    if (graph_exec) (void)hipGraphExecDestroy(graph_exec);
    if (graph) (void)hipGraphDestroy(graph);
    graph_exec = nullptr;
    graph = nullptr;
  }
  ~Driver() {
    drop_graph();
    if (h_sc2) (void)hipHostFree(h_sc2);
    for (auto& e : ev)
      if (e) (void)hipEventDestroy(e);
  }

  // after a stream sync: fold the event pairs of the last iteration in
  void collect_phase_times() {
    ktimer_collect();
    if (!ev_pending) return;
    float a = 0.f, b = 0.f, d = 0.f;
    if (hipEventElapsedTime(&a, ev[0], ev[1]) == hipSuccess &&
        hipEventElapsedTime(&b, ev[1], ev[2]) == hipSuccess &&
        hipEventElapsedTime(&d, ev[2], ev[3]) == hipSuccess) {
      g_phase.build_ms += a;
      g_phase.lu_ms += b;
      g_phase.solve_ms += d;
      g_phase.n_iterations += 1;
    }
    ev_pending = false;
  }

  // Symbolic analysis (ordering + pivots) from the Jacobian at Xw with the
  // current scalars; run before the timed loop.
  int ensure_analysis(const double* Qp, const double* dQ, const double* Q2) {
    vjx_pattern* pat = c->pat;
    if (pat->lu.analyzed && pat->uploaded) return VJX_OK;
    circuit_build_async(c, w.Xw.p, w.vsrc.p, w.isrc.p, Qp, dQ, Q2, nullptr, w.csr.p, w.f.p,
                        w.Qn.p, w.mrc.p, w.Ivs.p, nullptr, nullptr, nullptr, s);
    // pivots are chosen on the matrix the solves will see
    launch_noi_apply(c->n_noi_zero, c->noi_zero.p, c->n_noi, c->noi_diag.p, c->noi_rows.p,
                     w.csr.p, w.f.p, s);
    VJX_TRY_HIP(hipGetLastError());
    VJX_TRY_HIP(hipStreamSynchronize(s));
    int rc = pattern_ensure_analyzed(pat, w.csr.p);
    if (rc < 0 || !cfg->autotune) return rc;
    rc = pattern_autotune_factor(pat, w.csr.p);
    if (rc < 0) return rc;
    rc = pattern_autotune_solves(pat, w.csr.p);
    if (rc < 0) return rc;
    return circuit_autotune_eval(c, w.Xw.p, s);
  }

  // One Newton iteration (solver_factories.py:230-325) on Xw, kernels only
  // (capturable): save X, reset status, build, refactor, solve, delta test,
  // damped update.
  hipEvent_t ev[4] = {nullptr, nullptr, nullptr, nullptr};
  bool ev_pending = false;

  int enqueue_iteration(const double* Qp, const double* dQ, const double* Q2,
                        const double* floor) {
    vjx_pattern* pat = c->pat;
    const bool prof = g_profile && !cfg->use_graph;
    if (prof && !ev[0])
      for (auto& e : ev) VJX_TRY_HIP(hipEventCreate(&e));
    // per-launch events only in eager profiling passes (never under capture)
    ktimer_enable(prof);
    // save the iterate (and limit states) for a re-pivot redo; clear the LU,
    // Newton and LTE status records
    launch_iter_begin(n_total + m, w.Xw.p, w.Xsave.p, n_lim, n_lim ? w.ls_iter.p : nullptr,
                      n_lim ? w.ls_save.p : nullptr, pat->d_status.p, w.st.p, w.st_lte.p, s);
    if (prof) VJX_TRY_HIP(hipEventRecord(ev[0], s));
    circuit_build_async(c, w.Xw.p, w.vsrc.p, w.isrc.p, Qp, dQ, Q2, floor, w.csr.p, w.f.p, w.Qn.p,
                        w.mrc.p, w.Ivs.p, w.st.p, n_lim ? w.ls_iter.p : nullptr,
                        n_lim ? w.ls_out.p : nullptr, s);
    if (n_lim) {
      // limit-state convergence terms, then the state feeds the next iteration
      launch_lim_stats(n_lim, w.ls_iter.p, w.ls_out.p, w.st.p, s);
      launch_copy(n_lim, w.ls_out.p, w.ls_iter.p, s);
    }
    // NOI rows / columns of the system the solve sees (the residual test above
    // used the unmodified f); their delta is then exactly 0, so the delta test
    // needs no mask
    launch_noi_apply(c->n_noi_zero, c->noi_zero.p, c->n_noi, c->noi_diag.p, c->noi_rows.p,
                     w.csr.p, w.f.p, s);
    if (prof) VJX_TRY_HIP(hipEventRecord(ev[1], s));
    int rc = pattern_refactor_async(pat, w.csr.p, cfg->precision, false);
    if (rc < 0) return rc;
    if (prof) VJX_TRY_HIP(hipEventRecord(ev[2], s));
    rc = pattern_solve_async(pat, w.csr.p, w.f.p, -1.0, w.delta.p, cfg->refine_steps);
    if (rc < 0) return rc;
    if (prof) {
      VJX_TRY_HIP(hipEventRecord(ev[3], s));
      ev_pending = true;
    }
    launch_delta(n_unk, m, n_total, w.delta.p, w.Xw.p, c->d_sc.p, w.st.p, s);
    launch_update(n_unk, m, n_total, w.delta.p, w.Xw.p, c->d_sc.p, w.st.p, s);
    launch_noi_clamp(c->n_noi, c->noi_rows.p, w.Xw.p, s);
    launch_status_export(w.st.p, pat->d_status.p, w.d_hst, s);
    return VJX_OK;
  }

  // Capture and instantiate the Newton-iteration graph for these history
  // buffers without launching it, so that its one-time cost falls before a
  // timed loop rather than inside its first step.
  int prepare_graph(const double* Qp, const double* dQ, const double* Q2, const double* floor) {
    vjx_pattern* pat = c->pat;
    if (!cfg->use_graph || !pat->lu.analyzed || !pat->uploaded) return VJX_OK;
    const double* args[4] = {Qp, dQ, Q2, floor};
    if (graph_exec && std::memcmp(args, graph_args, sizeof(args)) == 0) return VJX_OK;
    drop_graph();
    pat->stream = s;
    VJX_TRY_HIP(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
    int rc = enqueue_iteration(Qp, dQ, Q2, floor);
    hipGraph_t g = nullptr;
    hipError_t e = hipStreamEndCapture(s, &g);
    if (rc < 0) return rc;
    VJX_TRY_HIP(e);
    graph = g;
    VJX_TRY_HIP(hipGraphInstantiate(&graph_exec, graph, nullptr, nullptr, 0));
    std::memcpy(graph_args, args, sizeof(args));
    return VJX_OK;
  }

  int nr_iteration(const double* Qp, const double* dQ, const double* Q2, const double* floor) {
    vjx_pattern* pat = c->pat;
    if (!pat->lu.analyzed || !pat->uploaded) {
      int rc = ensure_analysis(Qp, dQ, Q2);
      if (rc < 0) return rc;
      drop_graph();
    }
    pat->stream = s;
    ++total_evals;
    if (!cfg->use_graph) {
      int rc = enqueue_iteration(Qp, dQ, Q2, floor);
      if (rc < 0) return rc;
      enqueue_lte();
      VJX_TRY_HIP(hipGetLastError());
      return VJX_OK;
    }
    int rc = prepare_graph(Qp, dQ, Q2, floor);
    if (rc < 0) return rc;
    VJX_TRY_HIP(hipGraphLaunch(graph_exec, s));
    ++graph_launches;
    enqueue_lte();
    VJX_TRY_HIP(hipGetLastError());
    return VJX_OK;
  }

  // nr_solve while-loop semantics; returns iterations, converged, max_f.
  int nr_solve(double c0, double c1, double d1, double c2, double gshunt, const double* Qp,
               const double* dQ, const double* Q2, const double* floor, int& iters, bool& conv,
               double& max_f) {
    iters = 0;
    conv = false;
    max_f = std::numeric_limits<double>::infinity();
    vjx_pattern* pat = c->pat;
    bool redone = false;  // this iteration already re-pivoted once
    // at most one re-pivot per solve that does not cure the factor: one that
    // leaves it non-finite or above the growth threshold (a diverging
    // iterate, whose Jacobian holds entries of 1e30 and beyond, is no
    // pivoting problem), or an analysis that fails, disarms the monitor for
    // the rest of this solve
    bool armed = true;
    while (!conv && iters < cfg->tran_itl) {
      int rc = upload_scalars(c0, c1, d1, c2, gshunt, iters);
      if (rc < 0) return rc;
      rc = nr_iteration(Qp, dQ, Q2, floor);
      if (rc < 0) return rc;
      // the iteration exported both status records as its last kernel
      VJX_TRY_HIP(hipStreamSynchronize(s));
      collect_phase_times();
      // pivot monitor: a finite matrix whose factor is non-finite or grows
      // beyond the threshold is re-pivoted (at most once per iteration) and
      // the iteration redone from the saved iterate.  A non-finite matrix
      // (device eval overflow) is not a pivoting problem: the iteration
      // stands, its non-finite update fails the convergence tests, as the
      // reference's solve would (solver_factories.py:1225-1297).  A matrix
      // the analysis finds singular keeps the previous pivots (UMFPACK's
      // singular-matrix warning is tolerated, umfpack_ffi.cpp:297, 320).
      const DevStatus& lu_st = w.h_st[1];
      const double maxl = bits_to_double(lu_st.maxl_bits);
      const double growth = cfg->repivot_growth > 0.0 ? cfg->repivot_growth : 1e12;
      const bool bad_factor = lu_st.nonfinite > 0 || maxl > growth;
      if (redone && bad_factor) armed = false;
      if (lu_st.nonfinite_in == 0 && !redone && bad_factor && armed && repivots < 64) {
        redone = true;
        std::vector<double> h(c->nse);
        VJX_TRY_HIP(hipMemcpy(h.data(), w.csr.p, sizeof(double) * c->nse, hipMemcpyDeviceToHost));
        if (!fill_base) fill_base = (int64_t)pat->lu.sym.col.size();
        const int64_t user_limit = pat->opts.max_lu_nnz;
        pat->opts.max_lu_nnz =
            user_limit > 0 ? std::min(user_limit, 4 * fill_base) : 4 * fill_base;
        rc = vjx_lu_analyze(pat, h.data());
        pat->opts.max_lu_nnz = user_limit;
        if (rc == VJX_ERR_SINGULAR || rc == VJX_ERR_FILL) {
          armed = false;  // the previous pivots stay; the iteration stands
        } else {
          if (rc < 0) return rc;
          rc = pattern_upload(pat);
          if (rc < 0) return rc;
          drop_graph();  // factor buffers were reallocated
          ++repivots;
          launch_copy(n_total + m, w.Xsave.p, w.Xw.p, s);
          if (n_lim) launch_copy(n_lim, w.ls_save.p, w.ls_iter.p, s);
          --total_evals;
          continue;
        }
      }
      redone = false;
      const bool res_ok = w.h_st->res_fail == 0;
      const bool delta_ok = (iters == 0) || w.h_st->delta_fail == 0;
      max_f = bits_to_double(w.h_st->maxf_bits);
      conv = res_ok && delta_ok;
      if (n_lim) {
        // VACASK preventedConvergence (solver_factories.py:317-323)
        const double lim_delta = bits_to_double(w.h_st->limd_bits);
        const double lim_ref =
            std::max(bits_to_double(w.h_st->limr_bits) * cfg->reltol, cfg->vntol);
        conv = conv && lim_delta < lim_ref && iters >= 1;
      }
      ++iters;
    }
    // final build at X_final for Q, I_vsource, max_res_contrib (:689-713)
    int rc = upload_scalars(c0, c1, d1, c2, gshunt, iters);
    if (rc < 0) return rc;
    circuit_build_async(c, w.Xw.p, w.vsrc.p, w.isrc.p, Qp, dQ, Q2, floor, nullptr, w.f.p, w.Qn.p,
                        w.mrc.p, w.Ivs.p, nullptr, n_lim ? w.ls_iter.p : nullptr, nullptr, s);
    ++total_evals;
    launch_dqdt(n_unk, c0, c1, d1, c2, w.Qn.p, Qp, dQ, Q2, w.dQn.p, s);
    VJX_TRY_HIP(hipGetLastError());
    return VJX_OK;
  }
};

}  // namespace

}  // namespace vjx

using namespace vjx;

extern "C" int vjx_transient_run(vjx_circuit* c, const vjx_tran_config* cfg, vjx_tran_result* res) {
  if (!c || !cfg || !res || !res->times || !res->V_out || !res->I_out) {
    set_error("vjx_transient_run: null argument");
    return VJX_ERR_ARG;
  }
  ktimer_enable(false);
  if (cfg->n_replay > 1) {
    bool ok = cfg->replay_times_host != nullptr && !cfg->fixed_step;
    for (int i = 1; ok && i < cfg->n_replay; ++i)
      ok = cfg->replay_times_host[i] >= cfg->replay_times_host[i - 1];
    if (!ok) {
      set_error("vjx_transient_run: replay needs a non-decreasing time grid and the adaptive loop");
      return VJX_ERR_ARG;
    }
  }
  if (cfg->max_steps < 2 || !(cfg->dt > 0.0) || cfg->tran_itl < 1 ||
      (cfg->precision != VJX_PREC_F64 && cfg->precision != VJX_PREC_F32)) {
    set_error("vjx_transient_run: invalid configuration");
    return VJX_ERR_ARG;
  }
  Driver D;
  D.c = c;
  D.cfg = cfg;
  int rc = pattern_stream(c->pat);
  if (rc < 0) return rc;
  D.s = c->pat->stream;
  hipStream_t s = D.s;
  const int n_total = c->n_total, n_unk = c->n_unk, m = c->m, k = c->k;
  D.n_total = n_total;
  D.n_unk = n_unk;
  D.m = m;
  D.k = k;
  D.n_aug = c->n_aug;
  Work& w = D.w;
  const int nx = n_total + m;
  const int max_hist = cfg->max_order + 2;
  VJX_TRY_HIP(w.Xa.alloc(nx));
  VJX_TRY_HIP(w.Xw.alloc(nx));
  VJX_TRY_HIP(w.Xsave.alloc(nx));
  VJX_TRY_HIP(w.Vp.alloc(n_total));
  for (auto* b : {&w.Qp, &w.dQ, &w.Q2, &w.Qn, &w.dQn, &w.mrc, &w.hmrc, &w.floor, &w.zeros})
    VJX_TRY_HIP(b->alloc(std::max(n_unk, 1)));
  VJX_TRY_HIP(w.Vmax.alloc(n_total));
  VJX_TRY_HIP(w.f.alloc(std::max(c->n_aug, 1)));
  VJX_TRY_HIP(w.delta.alloc(std::max(c->n_aug, 1)));
  VJX_TRY_HIP(w.csr.alloc(std::max(c->nse, 1)));
  VJX_TRY_HIP(w.vsrc.alloc(std::max(m, 1)));
  VJX_TRY_HIP(w.isrc.alloc(std::max(k, 1)));
  VJX_TRY_HIP(w.Ivs.alloc(std::max(m, 1)));
  VJX_TRY_HIP(w.hist.alloc((size_t)max_hist * n_total));
  const int mo = std::max(m, 1);
  // output rows on the device: all of them, or a chunk flushed to the host
  // result each time it fills
  const int out_rows = cfg->output_chunk > 0 ? std::min(cfg->output_chunk, cfg->max_steps)
                                             : cfg->max_steps;
  int out_base = 0;  // step index of the chunk's first row
  VJX_TRY_HIP(w.vout.alloc((size_t)out_rows * n_total));
  VJX_TRY_HIP(w.iout.alloc((size_t)out_rows * mo));
  VJX_TRY_HIP(w.st.alloc(1));
  VJX_TRY_HIP(w.st_lte.alloc(1));
  VJX_TRY_HIP(w.alloc_status());
  const int nl = c->n_lim;
  D.n_lim = nl;
  if (nl) {
    for (auto* b : {&w.ls_step, &w.ls_iter, &w.ls_out, &w.ls_save}) VJX_TRY_HIP(b->alloc(nl));
    // the DC solve and the time loop both start from zero limit states
    // (full_mna.py:913, 1028-1033)
    launch_fill(nl, 0.0, w.ls_iter.p, s);
    launch_fill(nl, 0.0, w.ls_step.p, s);
  }
  launch_fill(n_unk, 0.0, w.zeros.p, s);
  launch_fill((int)((size_t)out_rows * mo), 0.0, w.iout.p, s);
  launch_fill(m, 0.0, w.Ivs.p, s);

  // ---------------- initial state (full_mna.py:871-957) ----------------
  auto t_dc0 = std::chrono::steady_clock::now();
  std::vector<double> x0(nx, 0.0);
  if (!cfg->uic && cfg->v_init_host)
    for (int i = 0; i < n_total; ++i) x0[i] = cfg->v_init_host[i];
  x0[0] = 0.0;
  VJX_TRY_HIP(hipMemcpy(w.Xw.p, x0.data(), sizeof(double) * nx, hipMemcpyHostToDevice));
  rc = D.set_sources(0.0);
  if (rc < 0) return rc;
  launch_fill(n_unk, cfg->abstol, w.floor.p, s);
  int dc_iters = 0;
  bool dc_conv = true;
  double dc_maxf = 0.0;
  if (cfg->uic) {
    rc = D.upload_scalars(0.0, 0.0, 0.0, 0.0, 0.0, 1);
    if (rc < 0) return rc;
    circuit_build_async(c, w.Xw.p, w.vsrc.p, w.isrc.p, w.zeros.p, w.zeros.p, w.zeros.p, nullptr,
                        nullptr, w.f.p, w.Qn.p, w.mrc.p, w.Ivs.p, nullptr, nullptr, nullptr, s);
    ++D.total_evals;
    launch_fill(n_unk, 0.0, w.mrc.p, s);
  } else {
    rc = D.nr_solve(0.0, 0.0, 0.0, 0.0, 0.0, w.zeros.p, w.zeros.p, w.zeros.p, w.floor.p, dc_iters,
                    dc_conv, dc_maxf);
    if (rc < 0) return rc;
  }
  // state from the initial solution
  launch_copy(nx, w.Xw.p, w.Xa.p, s);
  launch_copy(n_unk, w.Qn.p, w.Qp.p, s);
  launch_fill(n_unk, 0.0, w.dQ.p, s);
  launch_fill(n_unk, 0.0, w.Q2.p, s);
  launch_copy(n_unk, w.mrc.p, w.hmrc.p, s);
  launch_fill((int)((size_t)max_hist * n_total), 0.0, w.hist.p, s);
  launch_copy(n_total, w.Xa.p, w.hist.p, s);
  // V_max_historic = |X0|: accept kernel on a zero vector computes max(0,|x|)
  launch_fill(n_total, 0.0, w.Vmax.p, s);
  // copy rows [out_base, end) of the chunk to the host result
  auto flush_rows = [&](int end) -> int {
    const size_t rows = (size_t)(end - out_base);
    if (!rows) return VJX_OK;
    VJX_TRY_HIP(hipStreamSynchronize(s));
    VJX_TRY_HIP(hipMemcpy(res->V_out + (size_t)out_base * n_total, w.vout.p,
                          sizeof(double) * rows * n_total, hipMemcpyDeviceToHost));
    VJX_TRY_HIP(hipMemcpy(res->I_out + (size_t)out_base * mo, w.iout.p,
                          sizeof(double) * rows * mo, hipMemcpyDeviceToHost));
    return VJX_OK;
  };
  // enqueue the output of step idx (rows are written in increasing order)
  auto emit = [&](int idx, int zero_current) -> int {
    if (idx - out_base == out_rows) {
      int r2 = flush_rows(idx);
      if (r2 < 0) return r2;
      out_base = idx;
    }
    const size_t row = (size_t)(idx - out_base);
    launch_output(n_total, m, w.Xa.p, w.Ivs.p, zero_current, w.vout.p + row * n_total,
                  w.iout.p + row * mo, s);
    return VJX_OK;
  };
  // row 0: the initial state; the adaptive loop records the DC source
  // currents there (full_mna.py:1021-1023), the fixed-step loop leaves them
  // at zero (full_mna.py:1259-1263)
  rc = emit(0, cfg->fixed_step ? 1 : 0);
  if (rc < 0) return rc;
  {
    // Vmax = |X0|
    DBuf<double> tmp;
    VJX_TRY_HIP(tmp.alloc(nx));
    launch_accept(n_total, 0, 0, w.Xa.p, tmp.p, tmp.p, w.Vmax.p, nullptr, nullptr, nullptr,
                  nullptr, nullptr, nullptr, nullptr, s);
    VJX_TRY_HIP(hipStreamSynchronize(s));
  }
  // symbolic analysis before the timed loop, from a transient-like Jacobian
  // (backward-Euler c0 at the first step size) when the DC solve did not
  // already provide one
  {
    const double dt0 = cfg->tran_fs != 1.0 ? cfg->dt * cfg->tran_fs : cfg->dt;
    rc = D.upload_scalars(1.0 / dt0, -1.0 / dt0, 0.0, 0.0, 0.0, 0);
    if (rc < 0) return rc;
    rc = D.ensure_analysis(w.Qp.p, w.dQ.p, w.Q2.p);
    if (rc < 0) return rc;
  }
  // the time loop's Newton graph (fixed history buffers) is built here, as
  // setup, not inside the first step
  rc = D.prepare_graph(w.Qp.p, w.dQ.p, w.Q2.p, w.floor.p);
  if (rc < 0) return rc;
  VJX_TRY_HIP(hipGetLastError());
  VJX_TRY_HIP(hipStreamSynchronize(s));
  res->dc_time = std::chrono::duration<double>(std::chrono::steady_clock::now() - t_dc0).count();
  res->dc_converged = dc_conv ? 1 : 0;
  res->dc_iterations = dc_iters;

  std::vector<double> times(cfg->max_steps, 0.0);
  int step_idx = 1;
  int64_t total_nr = 0;
  int rejected = 0;
  int replay_nonconv = 0;
  // measurement window: opened (after a stream synchronisation) at the first
  // step attempt whose accepted-step count reaches window_from_step
  struct Window {
    bool open = false;
    int step0 = 0, rej0 = 0;
    int64_t nr0 = 0;
    std::chrono::steady_clock::time_point t0;
    int start(hipStream_t st, int step, int rej, int64_t nr, bool mark) {
      if (mark) launch_window_mark(0, st);
      VJX_TRY_HIP(hipStreamSynchronize(st));
      open = true;
      step0 = step;
      rej0 = rej;
      nr0 = nr;
      t0 = std::chrono::steady_clock::now();
      return VJX_OK;
    }
  } window;
  double t = 0.0;
  double dt = cfg->dt;
  auto t0 = std::chrono::steady_clock::now();

  if (cfg->fixed_step) {
    // ---------------- fixed-step loop (full_mna.py:1198-1441) ----------------
    if (cfg->tran_fs != 1.0) dt = dt * cfg->tran_fs;
    int n_steps = (int)std::min<double>(std::floor(cfg->t_stop / dt), (double)(cfg->max_steps - 1));
    if (n_steps < 1) n_steps = 1;
    const double inv = 1.0 / dt;
    double c0 = inv, c1 = -inv, d1 = 0.0, c2 = 0.0;
    if (cfg->method == VJX_METHOD_TRAP) {
      c0 = 2.0 * inv;
      c1 = -2.0 * inv;
      d1 = -1.0;
    } else if (cfg->method == VJX_METHOD_GEAR2) {
      c0 = 1.5 * inv;
      c1 = -2.0 * inv;
      c2 = 0.5 * inv;
    }
    launch_fill(n_unk, cfg->abstol, w.floor.p, s);
    for (int i = 0; i < n_steps; ++i) {
      if (i + 1 >= cfg->window_from_step && !window.open) {
        rc = window.start(s, i + 1, 0, total_nr, cfg->window_from_step > 0);
        if (rc < 0) return rc;
      }
      const double t_next = ((double)i + 1.0) * dt;
      rc = D.set_sources(t_next);
      if (rc < 0) return rc;
      int iters;
      bool conv;
      double maxf;
      if (nl) launch_copy(nl, w.ls_step.p, w.ls_iter.p, s);
      rc = D.nr_solve(c0, c1, d1, c2, 0.0, w.Qp.p, w.dQ.p, w.Q2.p, w.floor.p, iters, conv, maxf);
      if (rc < 0) return rc;
      total_nr += iters;
      // the limit state advances with a converged step (full_mna.py:1349-1353)
      if (nl && conv) launch_copy(nl, w.ls_iter.p, w.ls_step.p, s);
      times[i + 1] = t_next;
      launch_fixed_accept(nx, n_unk, w.Xw.p, w.Xa.p, w.Qn.p, w.dQn.p, w.Qp.p, w.dQ.p, w.Q2.p, s);
      rc = emit(i + 1, 0);
      if (rc < 0) return rc;
    }
    step_idx = n_steps + 1;
    res->min_dt_used = dt;
    res->max_dt_used = dt;
  } else {
    // ---------------- adaptive loop (full_mna.py:780-1196, 1840-2194) ----------------
    double max_dt = cfg->max_dt;
    if (cfg->tran_minpts > 0) {
      const double hm = cfg->t_stop / cfg->tran_minpts;
      const double eff = std::min(max_dt, hm);
      if (eff < max_dt) max_dt = eff;
    }
    if (cfg->tran_fs != 1.0) dt = dt * cfg->tran_fs;
    std::vector<double> dt_hist(max_hist, dt);
    std::vector<int> slot(max_hist);
    for (int j = 0; j < max_hist; ++j) slot[j] = j;
    int hist_count = 1, warmup = 0, consec = 0;
    double min_dt_used = dt, max_dt_used = dt;
    auto hist_ptr = [&](int j) { return w.hist.p + (size_t)slot[j] * n_total; };
    const bool replay = cfg->n_replay > 1 && cfg->replay_times_host;
    if (replay) dt = cfg->replay_times_host[1] - cfg->replay_times_host[0];
    while (replay ? step_idx < std::min(cfg->n_replay, cfg->max_steps)
                  : (t < cfg->t_stop && step_idx < cfg->max_steps)) {
      if (step_idx >= cfg->window_from_step && !window.open) {
        rc = window.start(s, step_idx, rejected, total_nr, cfg->window_from_step > 0);
        if (rc < 0) return rc;
      }
      const double dt_cur = dt;
      if (!(dt_cur > 0.0)) {
        // full_mna.py:2076 clamps the next dt to t_stop - t_next also after a
        // rejected step, so rejecting a step that ends on t_stop leaves dt = 0.
        // The reference then builds with 1/dt = inf: every Newton iteration is
        // non-finite, the step is a Newton failure at min_dt, and it is
        // force-accepted with the previous solution and zero source currents
        // (full_mna.py:2059-2085).  Restated without the non-finite arithmetic,
        // which would otherwise trip the re-pivot monitor.
        total_nr += cfg->tran_itl;
        consec = 0;
        const double last_good = hist_count > 0 ? dt_hist[0] : cfg->min_dt * 16.0;
        double new_dt = std::max(last_good / 8.0, cfg->min_dt * 2.0);
        new_dt = std::min(std::max(new_dt, cfg->min_dt), max_dt);
        new_dt = std::min(new_dt, cfg->t_stop - t);
        times[step_idx] = t;
        rc = emit(step_idx, 1);
        if (rc < 0) return rc;
        ++step_idx;
        min_dt_used = std::min(min_dt_used, dt_cur);
        max_dt_used = std::max(max_dt_used, dt_cur);
        dt = new_dt;
        if (replay && step_idx < cfg->n_replay)
          dt = cfg->replay_times_host[step_idx] - cfg->replay_times_host[step_idx - 1];
        continue;
      }
      const double inv = 1.0 / dt_cur;
      double c0, c1, d1, c2, e_int;
      if (cfg->method == VJX_METHOD_BE) {
        c0 = inv; c1 = -inv; d1 = 0.0; c2 = 0.0; e_int = -0.5;
      } else if (cfg->method == VJX_METHOD_TRAP) {
        c0 = 2.0 * inv; c1 = -2.0 * inv; d1 = -1.0; c2 = 0.0; e_int = 1.0 / 24.0;
      } else {
        const double dt_prev = dt_hist[0];
        const double om = dt_prev > 0 ? dt_cur / dt_prev : 1.0;
        const double op1 = 1.0 + om;
        c0 = (1.0 + 2.0 * om) / (op1 * dt_cur);
        c1 = -op1 / dt_cur;
        d1 = 0.0;
        c2 = (om * om) / (op1 * dt_cur);
        e_int = -2.0 / 9.0;
      }
      const double t_next = replay ? cfg->replay_times_host[step_idx] : t + dt_cur;
      rc = D.set_sources(t_next);
      if (rc < 0) return rc;
      const bool can_predict = (warmup >= cfg->warmup_steps) && (hist_count >= 2);
      // predictor (adaptive.py:94-218)
      int order = std::max(std::min(hist_count - 1, cfg->max_order), 0);
      double cs0 = dt_hist[0], cs1 = dt_hist[0] + dt_hist[1];
      double tau1 = -(dt_cur + cs0) / dt_cur;
      double tau2 = -(dt_cur + cs1) / dt_cur;
      double ratio1 = dt_cur / std::max(dt_hist[0], 1e-30);
      double err1 = -ratio1 * (1 + ratio1) / 2;
      const double tau0 = -1.0;
      double den0 = (tau0 - tau1) * (tau0 - tau2);
      double den1 = (tau1 - tau0) * (tau1 - tau2);
      double den2 = (tau2 - tau0) * (tau2 - tau1);
      if (std::fabs(den0) < 1e-30) den0 = 1e-30;
      if (std::fabs(den1) < 1e-30) den1 = 1e-30;
      if (std::fabs(den2) < 1e-30) den2 = 1e-30;
      double l0 = (0 - tau1) * (0 - tau2) / den0;
      double l1 = (0 - tau0) * (0 - tau2) / den1;
      double l2 = (0 - tau0) * (0 - tau1) / den2;
      double err2 = -tau0 * tau1 * tau2 / 6;
      double a[3], pred_err;
      if (order == 0) {
        a[0] = 1.0; a[1] = 0.0; a[2] = 0.0; pred_err = 1.0;
      } else if (order == 1) {
        a[0] = 1.0 + ratio1; a[1] = -ratio1; a[2] = 0.0; pred_err = err1;
      } else {
        a[0] = l0; a[1] = l1; a[2] = l2; pred_err = err2;
      }
      for (int j = order + 1; j < 3; ++j) a[j] = 0.0;
      // predictor, and the step's residual floor from the history maxima
      launch_predict(n_total, m, w.Xa.p, hist_ptr(0), hist_ptr(1), hist_ptr(2), a[0], a[1], a[2],
                     can_predict ? 1 : 0, w.Vp.p, w.Xw.p, n_unk, w.hmrc.p, cfg->reltol,
                     cfg->abstol, w.floor.p, s);
      double ramp = cfg->gshunt_steps > 0
                        ? std::min(std::max((double)step_idx / cfg->gshunt_steps, 0.0), 1.0)
                        : 1.0;
      const double gshunt = cfg->gshunt_init + ramp * (cfg->gshunt_target - cfg->gshunt_init);
      int iters;
      bool conv;
      double maxf;
      if (nl) launch_copy(nl, w.ls_step.p, w.ls_iter.p, s);
      // LTE (adaptive.py:221-339): computed behind each Newton iteration on
      // its iterate, read with the last one's status
      D.lte_hook = true;
      D.lte_scale = e_int / (e_int - pred_err);
      rc = D.nr_solve(c0, c1, d1, c2, gshunt, w.Qp.p, w.dQ.p, w.Q2.p, w.floor.p, iters, conv, maxf);
      D.lte_hook = false;
      if (rc < 0) return rc;
      total_nr += iters;
      const bool nr_failed = !conv;
      const bool at_min = dt_cur <= cfg->min_dt;
      bool fail_min = nr_failed && at_min;
      const double lte_norm = bits_to_double(w.h_st[2].maxf_bits);
      const int lorder = std::min(hist_count - 1, cfg->max_order);
      double factor = std::pow(cfg->lte_ratio / std::max(lte_norm, 1e-10), 1.0 / (lorder + 1));
      factor = std::min(std::max(factor, 0.1), cfg->grow_factor);
      const double dt_lte = dt_cur * factor;
      const bool lte_reject = (dt_cur / dt_lte > cfg->redo_factor) && can_predict && conv;
      const bool nr_reject = nr_failed && !at_min;
      const bool force = consec >= cfg->max_consecutive_rejects;
      const bool reject = (lte_reject || nr_reject) && !force && !replay;
      bool accept = !reject;
      if (accept)
        consec = 0;
      else if (lte_reject)
        ++consec;
      const double dt_from_lte = can_predict ? dt_lte : dt_cur;
      const double dt_halved = std::max(dt_cur * cfg->tran_ft, cfg->min_dt);
      const double last_good = hist_count > 0 ? dt_hist[0] : cfg->min_dt * 16.0;
      const double dt_rec = std::max(last_good / 8.0, cfg->min_dt * 2.0);
      double new_dt = nr_failed ? (at_min ? dt_rec : dt_halved) : dt_from_lte;
      new_dt = std::min(std::max(new_dt, cfg->min_dt), max_dt);
      new_dt = std::min(new_dt, cfg->t_stop - t_next);
      bool use_new = accept && !fail_min;
      if (replay) {
        // the step is the replayed grid's: taken whatever the LTE says
        accept = true;
        use_new = true;
        fail_min = false;
        if (nr_failed) ++replay_nonconv;
        if (step_idx + 1 < cfg->n_replay)
          new_dt = cfg->replay_times_host[step_idx + 1] - cfg->replay_times_host[step_idx];
      }
      if (accept) t = t_next;
      if (use_new) {
        // the limit state advances with the new solution (full_mna.py:2162-2167)
        if (nl) launch_copy(nl, w.ls_iter.p, w.ls_step.p, s);
        // roll history: the oldest slot receives V_new and becomes slot 0
        int last = slot[max_hist - 1];
        for (int j = max_hist - 1; j > 0; --j) slot[j] = slot[j - 1];
        slot[0] = last;
        for (int j = max_hist - 1; j > 0; --j) dt_hist[j] = dt_hist[j - 1];
        dt_hist[0] = dt_cur;
        launch_accept(n_total, n_unk, m, w.Xw.p, w.Xa.p, hist_ptr(0), w.Vmax.p, w.Qn.p, w.dQn.p,
                      w.Qp.p, w.dQ.p, w.Q2.p, w.mrc.p, w.hmrc.p, s);
        hist_count = std::min(hist_count + 1, max_hist);
        ++warmup;
      }
      if (accept) {
        times[step_idx] = t_next;
        rc = emit(step_idx, fail_min ? 1 : 0);
        if (rc < 0) return rc;
        ++step_idx;
        min_dt_used = std::min(min_dt_used, dt_cur);
        max_dt_used = std::max(max_dt_used, dt_cur);
      } else {
        ++rejected;
      }
      dt = new_dt;
    }
    res->min_dt_used = min_dt_used;
    res->max_dt_used = max_dt_used;
  }
  if (window.open && cfg->window_from_step > 0) launch_window_mark(1, s);
  VJX_TRY_HIP(hipGetLastError());
  VJX_TRY_HIP(hipStreamSynchronize(s));
  const auto t_end = std::chrono::steady_clock::now();
  res->wall_time = std::chrono::duration<double>(t_end - t0).count();
  res->n_steps = step_idx;
  res->rejected_steps = (cfg->n_replay > 1 && cfg->replay_times_host) ? replay_nonconv : rejected;
  if (window.open) {
    res->window_steps = step_idx - window.step0;
    res->window_rejected = rejected - window.rej0;
    res->window_nr_iterations = total_nr - window.nr0;
    res->window_time = std::chrono::duration<double>(t_end - window.t0).count();
  } else {
    res->window_steps = 0;
    res->window_rejected = 0;
    res->window_nr_iterations = 0;
    res->window_time = 0.0;
  }
  res->total_nr_iterations = total_nr;
  res->total_evals = D.total_evals;
  res->repivots = D.repivots;
  std::memcpy(res->times, times.data(), sizeof(double) * step_idx);
  rc = flush_rows(step_idx);
  if (rc < 0) return rc;
  return dc_conv ? VJX_OK : VJX_WARN_NOCONV;
}

extern "C" int vjx_newton_solve(vjx_circuit* c, const vjx_tran_config* cfg,
                                const vjx_newton_args* a, vjx_newton_result* r) {
  if (!c || !cfg || !a || !r || !a->X_init || !r->X || !r->Q || !r->dQdt || !r->I_vsource ||
      !r->max_res_contrib || (c->m && !a->vsource_vals) || (c->k && !a->isource_vals)) {
    set_error("vjx_newton_solve: null argument");
    return VJX_ERR_ARG;
  }
  if (cfg->tran_itl < 1 || (cfg->precision != VJX_PREC_F64 && cfg->precision != VJX_PREC_F32)) {
    set_error("vjx_newton_solve: invalid configuration");
    return VJX_ERR_ARG;
  }
  ktimer_enable(false);
  Driver D;
  D.c = c;
  D.cfg = cfg;
  D.gmin_dev = a->gmin;
  int rc = pattern_stream(c->pat);
  if (rc < 0) return rc;
  D.s = c->pat->stream;
  hipStream_t s = D.s;
  const int n_total = c->n_total, n_unk = c->n_unk, m = c->m, k = c->k;
  D.n_total = n_total;
  D.n_unk = n_unk;
  D.m = m;
  D.k = k;
  D.n_aug = c->n_aug;
  Work& w = D.w;
  const int nx = n_total + m;
  VJX_TRY_HIP(w.Xw.alloc(nx));
  VJX_TRY_HIP(w.Xsave.alloc(nx));
  for (auto* b : {&w.Qn, &w.dQn, &w.mrc, &w.floor, &w.zeros})
    VJX_TRY_HIP(b->alloc(std::max(n_unk, 1)));
  VJX_TRY_HIP(w.f.alloc(std::max(c->n_aug, 1)));
  VJX_TRY_HIP(w.delta.alloc(std::max(c->n_aug, 1)));
  VJX_TRY_HIP(w.csr.alloc(std::max(c->nse, 1)));
  VJX_TRY_HIP(w.vsrc.alloc(std::max(m, 1)));
  VJX_TRY_HIP(w.isrc.alloc(std::max(k, 1)));
  VJX_TRY_HIP(w.Ivs.alloc(std::max(m, 1)));
  VJX_TRY_HIP(w.st.alloc(1));
  VJX_TRY_HIP(w.st_lte.alloc(1));
  VJX_TRY_HIP(w.alloc_status());
  const int nl = c->n_lim;
  D.n_lim = nl;
  if (nl) {
    for (auto* b : {&w.ls_iter, &w.ls_out, &w.ls_save}) VJX_TRY_HIP(b->alloc(nl));
    if (a->limit_state_in)
      launch_copy(nl, a->limit_state_in, w.ls_iter.p, s);
    else
      launch_fill(nl, 0.0, w.ls_iter.p, s);
  }
  launch_fill(n_unk, 0.0, w.zeros.p, s);
  launch_fill(m, 0.0, w.Ivs.p, s);
  launch_copy(nx, a->X_init, w.Xw.p, s);
  launch_copy(m, a->vsource_vals, w.vsrc.p, s);
  launch_copy(k, a->isource_vals, w.isrc.p, s);
  if (a->res_tol_floor)
    launch_copy(n_unk, a->res_tol_floor, w.floor.p, s);
  else
    launch_fill(n_unk, cfg->abstol, w.floor.p, s);
  VJX_TRY_HIP(hipGetLastError());
  const double* Qp = a->Q_prev ? a->Q_prev : w.zeros.p;
  const double* dQ = a->dQdt_prev ? a->dQdt_prev : w.zeros.p;
  const double* Q2 = a->Q_prev2 ? a->Q_prev2 : w.zeros.p;
  if (!c->pat->lu.analyzed || !c->pat->uploaded) {
    // symbolic analysis on the Jacobian at X_init with these coefficients
    rc = D.upload_scalars(a->c0, a->c1, a->d1, a->c2, a->gshunt, 0);
    if (rc < 0) return rc;
    rc = D.ensure_analysis(Qp, dQ, Q2);
    if (rc < 0) return rc;
  }
  int iters = 0;
  bool conv = false;
  double maxf = 0.0;
  rc = D.nr_solve(a->c0, a->c1, a->d1, a->c2, a->gshunt, Qp, dQ, Q2, w.floor.p, iters, conv, maxf);
  if (rc < 0) return rc;
  launch_copy(nx, w.Xw.p, r->X, s);
  launch_copy(n_unk, w.Qn.p, r->Q, s);
  launch_copy(n_unk, w.dQn.p, r->dQdt, s);
  launch_copy(std::max(m, 1), w.Ivs.p, r->I_vsource, s);
  launch_copy(n_unk, w.mrc.p, r->max_res_contrib, s);
  if (nl && r->limit_state) launch_copy(nl, w.ls_iter.p, r->limit_state, s);
  VJX_TRY_HIP(hipGetLastError());
  VJX_TRY_HIP(hipStreamSynchronize(s));
  r->iterations = iters;
  r->converged = conv ? 1 : 0;
  r->max_f = maxf;
  r->repivots = D.repivots;
  return conv ? VJX_OK : VJX_WARN_NOCONV;
}

extern "C" int vjx_profile_enable(int32_t enable) {
  vjx::g_profile = enable != 0;
  return VJX_OK;
}

extern "C" int vjx_profile_reset(void) {
  vjx::g_phase = vjx_phase_times{};
  vjx::ktimer_reset();
  return VJX_OK;
}

extern "C" int vjx_profile_read(vjx_phase_times* out) {
  if (!out) {
    vjx::set_error("vjx_profile_read: null argument");
    return VJX_ERR_ARG;
  }
  *out = vjx::g_phase;
  vjx::ktimer_read(out->kernel_ms, out->kernel_launches);
  return VJX_OK;
}
