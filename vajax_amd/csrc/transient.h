// Launchers of the Newton / time-step element-wise kernels.
#pragma once

#include "internal.h"

namespace vjx {

void launch_delta(int n_unk, int m, int n_total, const double* delta, const double* X,
                  const DevScalars* sc, DevStatus* st, hipStream_t s);
void launch_update(int n_unk, int m, int n_total, const double* delta, double* X,
                   const DevScalars* sc, const DevStatus* st, hipStream_t s);
// predictor over the n_total + m unknowns; with floor given, also
// floor[i] = max(hmrc[i] * reltol, abstol) for i < n_unk
void launch_predict(int n_total, int m, const double* Xa, const double* h0, const double* h1,
                    const double* h2, double a0, double a1, double a2, int can_predict, double* Vp,
                    double* Xw, int n_unk, const double* hmrc, double reltol, double abstol,
                    double* floor, hipStream_t s);
void launch_fill(int n, double v, double* x, hipStream_t s);
void launch_noi_apply(int nz, const int32_t* zero, int nd, const int32_t* diag,
                      const int32_t* rows, double* csr, double* f, hipStream_t s);
void launch_noi_clamp(int nd, const int32_t* rows, double* X, hipStream_t s);
void launch_copy(int n, const double* a, double* b, hipStream_t s);
// Xa = Xw over nx; Q2 = Qp, Qp = Qn, dQ = dQn over n_unk (one launch)
void launch_fixed_accept(int nx, int n_unk, const double* Xw, double* Xa, const double* Qn,
                         const double* dQn, double* Qp, double* dQ, double* Q2, hipStream_t s);
// out[0] = *a and, when b is given, out[1] = *b (out: mapped pinned host memory)
void launch_status_export(const DevStatus* a, const DevStatus* b, DevStatus* out, hipStream_t s);
void launch_window_mark(int which, hipStream_t s);
// b[i] = a[idx[i]], i < n
void launch_gather(int n, const int32_t* idx, const double* a, double* b, hipStream_t s);
void launch_dqdt(int n, double c0, double c1, double d1, double c2, const double* Q,
                 const double* Qp, const double* dQ, const double* Q2, double* out,
                 hipStream_t s);
void launch_lte(int n, const double* Vn, const double* Vp, const double* Vmax, double scale,
                double reltol, double abstol, DevStatus* st, hipStream_t s);
void launch_accept(int n_total, int n_unk, int m, const double* Xw, double* Xa, double* hist_new,
                   double* Vmax, const double* Qn, const double* dQn, double* Qp, double* dQ,
                   double* Q2, const double* mrc, double* hmrc, hipStream_t s);
void launch_output(int n_total, int m, const double* Xa, const double* Ivs, int zero_current,
                   double* vrow, double* irow, hipStream_t s);

}  // namespace vjx
