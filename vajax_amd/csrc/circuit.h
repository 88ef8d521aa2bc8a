// Handle types behind the C ABI and the helpers shared by api.cpp,
// circuit.cpp and transient.cpp.
#pragma once

#include <memory>
#include <vector>

#include "internal.h"

struct vjx_pattern {
  int n = 0;
  int nnz = 0;
  std::vector<int32_t> indptr, indices;
  vjx::AnalyzeOptions opts;
  vjx::DevLU lu;
  bool uploaded = false;
  vjx::DBuf<int32_t> d_indptr, d_indices;
  vjx::HubPlan hub_rows;  // rows of A with more than kHeavySegment entries
  hipStream_t stream = nullptr;
  vjx::DBuf<vjx::DevStatus> d_status;
  vjx::DevStatus* h_status = nullptr;  // pinned
  double analyze_seconds = 0.0;
  int n_analyses = 0;
  double tuned_solve_ms = 0.0;   // solve time at the width the autotuner kept
  double tuned_factor_ms = 0.0;  // refactorisation time of the plan it kept
  // transposed solves: a second handle for the pattern of A^T and, for each
  // of its entries, the position of the same value in A's data
  vjx_pattern* trans = nullptr;
  vjx::DBuf<int32_t> d_trans_src;
  vjx::DBuf<double> d_trans_data;
  ~vjx_pattern() {
    if (h_status) (void)hipHostFree(h_status);
  }
};

struct vjx_circuit {
  int n_total = 0, n_unk = 0, m = 0, k = 0, n_aug = 0, nse = 0;
  std::vector<std::unique_ptr<vjx::DevModel>> models;
  vjx::DBuf<int32_t> slot_diag, jac_seg, res_seg, br_ptr, br_ent, vs_p, vs_n, vs_wave, is_wave,
      is_dst, conv_mask;
  vjx::HubPlan hub_slots, hub_rows;  // long stamping / residual segments
  vjx::DBuf<double> slot_b, vs_par, is_par;
  // pwl points (empty when no source of the kind is a pwl)
  vjx::DBuf<int32_t> vs_pwl_ptr, is_pwl_ptr;
  vjx::DBuf<double> vs_pwl_t, vs_pwl_v, is_pwl_t, is_pwl_v;
  vjx::SourceTable vsrc_table() const {
    return {m, vs_wave.p, vs_par.p, vs_pwl_ptr.p, vs_pwl_t.p, vs_pwl_v.p};
  }
  vjx::SourceTable isrc_table() const {
    return {k, is_wave.p, is_par.p, is_pwl_ptr.p, is_pwl_t.p, is_pwl_v.p};
  }
  int n_jac_contrib = 0, n_res_contrib = 0;
  vjx::DBuf<double> jac_contrib, rf, rq;
  // limited devices: flat limit-state count, lim_rhs contributions (same
  // slots as rf / rq, zero where no limited device writes) and a scratch
  // limit-state output for builds whose caller discards it
  int n_lim = 0;
  vjx::DBuf<double> lr, lq, ls_scratch;
  // noise-only internal nodes: their rows, the CSR slots zeroed (their rows,
  // their columns in node rows, diagonals excluded) and their diagonals
  int n_noi = 0, n_noi_zero = 0;
  vjx::DBuf<int32_t> noi_rows, noi_zero, noi_diag;
  vjx_pattern* pat = nullptr;
  vjx::DBuf<vjx::DevScalars> d_sc;
  vjx::DevScalars* h_sc = nullptr;  // pinned
  bool has_conv_mask = false;
  ~vjx_circuit();
};

namespace vjx {

int pattern_stream(vjx_pattern* pat);
int pattern_upload_csr(vjx_pattern* pat);
int pattern_ensure_analyzed(vjx_pattern* pat, const double* data_dev);
int pattern_upload(vjx_pattern* pat);
int pattern_refactor_async(vjx_pattern* pat, const double* data_dev, int prec, bool reset_status);
int pattern_check_pivots(vjx_pattern* pat);
// Pick the solves' band width by timing candidates on the device.
int pattern_autotune_solves(vjx_pattern* pat, const double* data_dev);
// Pick the refactorisation's band width and band-task pipeline by timing them.
int pattern_autotune_factor(vjx_pattern* pat, const double* data_dev);
// Solve A x = scale*rhs (scale = +1 or -1) with the current factors; for FP32
// factors run refine_steps FP64 residual corrections.
int pattern_solve_async(vjx_pattern* pat, const double* data_dev, const double* rhs, double scale,
                        double* x, int refine_steps);

void launch_spmv(int n, const int32_t* indptr, const int32_t* indices, const double* data,
                 const double* x, const double* base, double sgn, double* out,
                 const HubChunks& hub, hipStream_t s);
void launch_status_reset(vjx::DevStatus* st, hipStream_t s);
// Xsave[0, n) = X, ls_save[0, nl) = ls, and the status records a, b, c
// (each may be null) cleared: one launch.
void launch_iter_begin(int n, const double* X, double* Xsave, int nl, const double* ls,
                       double* ls_save, vjx::DevStatus* a, vjx::DevStatus* b, vjx::DevStatus* c,
                       hipStream_t s);

// Enqueue one system build: eval, stamping reductions, residual.
// ls_in (NULL: zeros) / ls_out: flat limit states of the limited devices.
// only_model >= 0 evaluates that model group alone (the others and the
// current sources contribute nothing; the structural stamps remain).
void circuit_build_async(vjx_circuit* c, const double* X, const double* vsrc, const double* isrc,
                         const double* Q_prev, const double* dQdt_prev, const double* Q_prev2,
                         const double* res_floor, double* csr, double* f, double* Q, double* mrc,
                         double* I_vs, vjx::DevStatus* st, const double* ls_in, double* ls_out,
                         hipStream_t s, int only_model = -1);

// Autotuner of the eval launch plan: for every model group of at least
// 8192 devices without limit states, times its eval kernel at workgroups of
// 256, 128 and 64 lanes (HIP events, best of three rounds of five launches)
// on the iterate X and keeps another size only when it is 10 % faster.
// Results do not depend on the size (one lane per device); the contribution
// buffers it writes are scratch that the next build overwrites.
int circuit_autotune_eval(vjx_circuit* c, const double* X, hipStream_t s);

}  // namespace vjx
