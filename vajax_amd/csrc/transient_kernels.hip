// Element-wise kernels of the Newton / time-step loop (FP64 throughout).
#include "block_flags.h"
#include "circuit.h"
#include "transient.h"

namespace vjx {

namespace {

__device__ __forceinline__ unsigned long long dbits(double x) {
  return (unsigned long long)__double_as_longlong(x);
}
__device__ __forceinline__ double bitsd(unsigned long long b) {
  return __longlong_as_double((long long)b);
}
// jnp.maximum semantics: NaN propagates.
__device__ __forceinline__ double jmax(double a, double b) {
  if (a != a || b != b) return __longlong_as_double(0x7ff8000000000000LL);
  return a > b ? a : b;
}
__device__ __forceinline__ double jmin(double a, double b) {
  if (a != a || b != b) return __longlong_as_double(0x7ff8000000000000LL);
  return a < b ? a : b;
}

// delta test of solver_factories.py:283-299
__global__ void __launch_bounds__(256) k_delta(int n_unk, int m, int n_total,
                                               const double* __restrict__ delta,
                                               const double* __restrict__ X,
                                               const DevScalars* __restrict__ sc, DevStatus* st) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  bool fail = false;
  double adv = 0.0, ad = 0.0;
  if (i < n_unk + m) {
    const double d = delta[i];
    double cd, xr, atol;
    if (i < n_unk) {
      cd = d * sc->nr_damping;
      xr = X[1 + i];
      atol = sc->vntol;
      adv = fabs(d);
    } else {
      cd = d;
      xr = X[n_total + (i - n_unk)];
      atol = sc->abstol;
    }
    const double tol = jmax(fabs(xr) * sc->reltol, atol);
    fail = !(fabs(cd) < tol);
    ad = fabs(d);
  }
  block_flags(dbits(adv), dbits(ad), fail ? 1u : 0u, &st->maxdv_bits, &st->maxdelta_bits,
              &st->delta_fail);
}

// damped update of solver_factories.py:301-309
__global__ void __launch_bounds__(256) k_update(int n_unk, int m, int n_total,
                                                const double* __restrict__ delta, double* X,
                                                const DevScalars* __restrict__ sc,
                                                const DevStatus* __restrict__ st) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n_unk + m) return;
  if (i < n_unk) {
    const double maxdv = bitsd(st->maxdv_bits);
    const double vs = jmin(1.0, sc->max_step / jmax(maxdv, 1e-30));
    X[1 + i] = X[1 + i] + delta[i] * vs * sc->nr_damping;
  } else {
    const int j = n_total + (i - n_unk);
    X[j] = X[j] + delta[i];
  }
}

// Predictor (adaptive.py:94-218) and, for the node rows, the residual
// tolerance floor max(reltol * history max |contribution|, abstol)
// (full_mna.py:1922-1928) of the step, in one launch.
__global__ void k_predict(int n_total, int m, const double* __restrict__ Xa,
                          const double* __restrict__ h0, const double* __restrict__ h1,
                          const double* __restrict__ h2, double a0, double a1, double a2,
                          int can_predict, double* Vp, double* Xw, int n_unk,
                          const double* __restrict__ hmrc, double reltol, double abstol,
                          double* floor) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (floor && i < n_unk) floor[i] = jmax(hmrc[i] * reltol, abstol);
  if (i < n_total) {
    const double v = a0 * h0[i] + a1 * h1[i] + a2 * h2[i];
    Vp[i] = v;
    Xw[i] = can_predict ? v : Xa[i];
  } else if (i < n_total + m) {
    Xw[i] = Xa[i];
  }
}

__global__ void k_fill(int n, double v, double* x) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) x[i] = v;
}

__global__ void k_copy(int n, const double* __restrict__ a, double* b) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) b[i] = a[i];
}

// The fixed-step loop's accepted step in one launch (full_mna.py:1355-1365):
// X_accepted = X_new, and the charge history shifts, Q_prev2 = Q_prev,
// Q_prev = Q_new, dQ/dt_prev = dQ/dt_new (each element read before written).
__global__ void k_fixed_accept(int nx, int n_unk, const double* __restrict__ Xw,
                               double* __restrict__ Xa, const double* __restrict__ Qn,
                               const double* __restrict__ dQn, double* __restrict__ Qp,
                               double* __restrict__ dQ, double* __restrict__ Q2) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < nx) Xa[i] = Xw[i];
  if (i < n_unk) {
    Q2[i] = Qp[i];
    Qp[i] = Qn[i];
    dQ[i] = dQn[i];
  }
}

// The Newton and LU status records written by one thread straight into
// mapped pinned host memory at the end of an iteration (no copy launches on
// the critical path).
__global__ void k_status_export(const DevStatus* __restrict__ a, const DevStatus* __restrict__ b,
                                DevStatus* out) {
  if (threadIdx.x != 0) return;
  out[0] = *a;
  if (b) out[1] = *b;
}

// Marks the boundaries of a measurement window in kernel traces (a profiler
// keeps the dispatches between two of these); it does no work.
__global__ void k_window_mark(int which) { (void)which; }

__global__ void k_gather(int n, const int32_t* __restrict__ idx, const double* __restrict__ a,
                         double* b) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) b[i] = a[idx[i]];
}

__global__ void k_dqdt(int n, double c0, double c1, double d1, double c2,
                       const double* __restrict__ Q, const double* __restrict__ Qp,
                       const double* __restrict__ dQ, const double* __restrict__ Q2, double* out) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) out[i] = c0 * Q[i] + c1 * Qp[i] + d1 * dQ[i] + c2 * Q2[i];
}

// LTE ratio of adaptive.py:254-271
__global__ void __launch_bounds__(256) k_lte(int n, const double* __restrict__ Vn,
                                             const double* __restrict__ Vp,
                                             const double* __restrict__ Vmax, double scale,
                                             double reltol, double abstol, DevStatus* st) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  double r = 0.0;
  if (i < n) {
    const double lte = (Vn[i] - Vp[i]) * scale;
    const double vc = jmax(fabs(Vn[i]), fabs(Vp[i]));
    const double vs = jmax(Vmax[i], vc);
    const double tol = reltol * vs + abstol;
    r = fabs(lte) / tol;
    if (r != r) r = __longlong_as_double(0x7ff8000000000000LL);
  }
  block_flags(dbits(r), 0, 0u, &st->maxf_bits, nullptr, nullptr);
}

// state update on use_new_solution (full_mna.py:2078-2113)
__global__ void k_accept(int n_total, int n_unk, int m, const double* __restrict__ Xw, double* Xa,
                         double* hist_new, double* Vmax, const double* __restrict__ Qn,
                         const double* __restrict__ dQn, double* Qp, double* dQ, double* Q2,
                         const double* __restrict__ mrc, double* hmrc) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n_total) {
    const double v = Xw[i];
    Xa[i] = v;
    hist_new[i] = v;
    Vmax[i] = jmax(Vmax[i], fabs(v));
  }
  if (i < m) Xa[n_total + i] = Xw[n_total + i];
  if (i < n_unk) {
    Q2[i] = Qp[i];
    Qp[i] = Qn[i];
    dQ[i] = dQn[i];
    hmrc[i] = jmax(hmrc[i], mrc[i]);
  }
}

__global__ void k_output(int n_total, int m, const double* __restrict__ Xa,
                         const double* __restrict__ Ivs, int zero_current, double* vrow,
                         double* irow) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n_total) vrow[i] = Xa[i];
  if (i < m) irow[i] = zero_current ? 0.0 : Ivs[i];
}

inline int nb(long long n, int b = 256) { return (int)((n + b - 1) / b); }

}  // namespace

void launch_delta(int n_unk, int m, int n_total, const double* delta, const double* X,
                  const DevScalars* sc, DevStatus* st, hipStream_t s) {
  k_delta<<<nb(n_unk + m), 256, 0, s>>>(n_unk, m, n_total, delta, X, sc, st);
}
void launch_update(int n_unk, int m, int n_total, const double* delta, double* X,
                   const DevScalars* sc, const DevStatus* st, hipStream_t s) {
  k_update<<<nb(n_unk + m), 256, 0, s>>>(n_unk, m, n_total, delta, X, sc, st);
}
void launch_predict(int n_total, int m, const double* Xa, const double* h0, const double* h1,
                    const double* h2, double a0, double a1, double a2, int can_predict, double* Vp,
                    double* Xw, int n_unk, const double* hmrc, double reltol, double abstol,
                    double* floor, hipStream_t s) {
  // n_unk < n_total: the grid covers the floor rows too
  k_predict<<<nb(n_total + m), 256, 0, s>>>(n_total, m, Xa, h0, h1, h2, a0, a1, a2, can_predict,
                                            Vp, Xw, n_unk, hmrc, reltol, abstol, floor);
}
// Noise-only internal nodes in the Newton solve (solver_factories.py:1268-1276):
// zero their rows and their columns in node rows, put 1 on their diagonals,
// 0 in their right-hand side (three disjoint index sets).
__global__ void k_noi_apply(int nz, const int32_t* __restrict__ zero, int nd,
                            const int32_t* __restrict__ diag, const int32_t* __restrict__ rows,
                            double* csr, double* f) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < nz) {
    csr[zero[i]] = 0.0;
  } else if (i < nz + nd) {
    csr[diag[i - nz]] = 1.0;
  } else if (i < nz + 2 * nd) {
    f[rows[i - nz - nd]] = 0.0;
  }
}

// NOI nodes held at 0 V after the update (solver_factories.py:311-313).
__global__ void k_noi_clamp(int nd, const int32_t* __restrict__ rows, double* X) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < nd) X[1 + rows[i]] = 0.0;
}

void launch_noi_apply(int nz, const int32_t* zero, int nd, const int32_t* diag,
                      const int32_t* rows, double* csr, double* f, hipStream_t s) {
  const int n = nz + 2 * nd;
  if (n) k_noi_apply<<<nb(n), 256, 0, s>>>(nz, zero, nd, diag, rows, csr, f);
}

void launch_noi_clamp(int nd, const int32_t* rows, double* X, hipStream_t s) {
  if (nd) k_noi_clamp<<<nb(nd), 256, 0, s>>>(nd, rows, X);
}

void launch_fill(int n, double v, double* x, hipStream_t s) {
  if (n) k_fill<<<nb(n), 256, 0, s>>>(n, v, x);
}
void launch_copy(int n, const double* a, double* b, hipStream_t s) {
  if (n) k_copy<<<nb(n), 256, 0, s>>>(n, a, b);
}
void launch_fixed_accept(int nx, int n_unk, const double* Xw, double* Xa, const double* Qn,
                         const double* dQn, double* Qp, double* dQ, double* Q2, hipStream_t s) {
  const int n = nx > n_unk ? nx : n_unk;
  if (n) k_fixed_accept<<<nb(n), 256, 0, s>>>(nx, n_unk, Xw, Xa, Qn, dQn, Qp, dQ, Q2);
}
void launch_window_mark(int which, hipStream_t s) { k_window_mark<<<1, 64, 0, s>>>(which); }
void launch_status_export(const DevStatus* a, const DevStatus* b, DevStatus* out, hipStream_t s) {
  k_status_export<<<1, 64, 0, s>>>(a, b, out);
}
void launch_gather(int n, const int32_t* idx, const double* a, double* b, hipStream_t s) {
  if (n) k_gather<<<nb(n), 256, 0, s>>>(n, idx, a, b);
}
void launch_dqdt(int n, double c0, double c1, double d1, double c2, const double* Q,
                 const double* Qp, const double* dQ, const double* Q2, double* out,
                 hipStream_t s) {
  if (n) k_dqdt<<<nb(n), 256, 0, s>>>(n, c0, c1, d1, c2, Q, Qp, dQ, Q2, out);
}
void launch_lte(int n, const double* Vn, const double* Vp, const double* Vmax, double scale,
                double reltol, double abstol, DevStatus* st, hipStream_t s) {
  k_lte<<<nb(n), 256, 0, s>>>(n, Vn, Vp, Vmax, scale, reltol, abstol, st);
}
void launch_accept(int n_total, int n_unk, int m, const double* Xw, double* Xa, double* hist_new,
                   double* Vmax, const double* Qn, const double* dQn, double* Qp, double* dQ,
                   double* Q2, const double* mrc, double* hmrc, hipStream_t s) {
  k_accept<<<nb(n_total > m ? n_total : m), 256, 0, s>>>(n_total, n_unk, m, Xw, Xa, hist_new, Vmax, Qn, dQn, Qp,
                                           dQ, Q2, mrc, hmrc);
}
void launch_output(int n_total, int m, const double* Xa, const double* Ivs, int zero_current,
                   double* vrow, double* irow, hipStream_t s) {
  k_output<<<nb(n_total + m), 256, 0, s>>>(n_total, m, Xa, Ivs, zero_current, vrow, irow);
}

}  // namespace vjx
