// C++ CPU restatement of the transient step loop -- test infrastructure and
// the CPU baseline, never part of the product library.
//
// The same algorithm as the numpy oracle (oracle/transient.py, oracle/mna.py,
// oracle/models.py), restated for speed on host cores:
//
// * system build (vajax/analysis/mna_builder.py:759-1076, mna.py:489-719):
//   devices are evaluated in parallel (OpenMP over instances); each device
//   writes its stamps to its own slots, and every CSR entry / residual row is
//   then summed over its contributions in stamping order (model, instance,
//   key) -- the order of the sequential scatter-add -- so the build is
//   deterministic and independent of the thread count;
// * sparse LU (the role of UMFPACK, vajax/sparse/umfpack_ffi.cpp:189-326, in a
//   KLU-style form): a fill-reducing column order computed once at setup, a
//   left-looking Gilbert-Peierls factorisation with threshold partial
//   pivoting (0.1, preferring the diagonal, umfpack_ffi.cpp:153) on the first
//   matrix, then numeric refactorisation on the same pattern and pivot order
//   for every later Newton iteration; a pivot that collapses (zero,
//   non-finite or below 1e-8 of its column) triggers a fresh pivoting
//   factorisation.  Serial, as UMFPACK / KLU are;
// * Newton (solver_factories.py:230-325, 574-727), adaptive BDF2 / trap / BE
//   step control (full_mna.py:780-1196, 1840-2194; adaptive.py:94-339) and the
//   fixed-step loop (full_mna.py:1198-1441), line by line as in
//   oracle/transient.py, including the zero-dt step after a rejected step
//   that ends on t_stop (full_mna.py:2076).
//
// Built with -ffp-contract=off so expressions round as written.

#include <omp.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <limits>
#include <string>
#include <vector>

namespace {

constexpr double KB = 1.380649e-23;
constexpr double QE = 1.602176634e-19;
constexpr double TEMP_K = 300.15;
const double VT = KB * TEMP_K / QE;

enum { K_RES = 0, K_CAP = 1, K_MOS = 2 };
enum { M_BE = 0, M_TRAP = 1, M_GEAR2 = 2 };

thread_local std::string g_err;

using clk = std::chrono::steady_clock;
inline double secs(clk::time_point a) {
  return std::chrono::duration<double>(clk::now() - a).count();
}

inline double nan0(double x) { return std::isnan(x) ? 0.0 : x; }

}  // namespace

extern "C" {

struct ct_model {
  int32_t kind;   // 0 resistor, 1 capacitor, 2 mosfet
  int32_t count;  // instances
  const int64_t* nodes;    // [count * T] global node index (T = 2 or 4)
  const int64_t* csr_pos;  // [count * T * T] CSR slot or -1
  const int64_t* res_idx;  // [count * T] residual row or -1
  const double* par;       // resistor r / capacitor c: [count]; mosfet: [count * 11]
};

struct ct_sources {
  int32_t count;
  const int64_t* p;
  const int64_t* n;
  const int32_t* wave;  // 0 dc, 1 pulse
  const double* par;    // [count * 8]: dc, val0, val1, rise, fall, width, period, delay
};

struct ct_circuit_desc {
  int32_t n_total, n_unk, m, n_models;
  int64_t nse;
  const int64_t* indptr;
  const int64_t* indices;
  const int64_t* diag_pos;
  int32_t nb;
  const int64_t* b_pos;
  const double* b_val;
  ct_sources vs, is;
  ct_model models[4];
  const int64_t* col_order;  // fill-reducing column order q[k] of the n_aug matrix
  double gmin;
};

struct ct_options {
  double reltol, abstol, vntol;
  int32_t tran_itl;
  double nr_damping;
  int32_t uic;
};

struct ct_adaptive {
  double lte_ratio, redo_factor, reltol, abstol, min_dt, max_dt;
  int32_t tran_minpts, warmup_steps, max_order;
  double grow_factor, gshunt_init;
  int32_t gshunt_steps;
  double gshunt_target, tran_fs, tran_ft;
  int32_t method, max_consecutive_rejects;
};

struct ct_stats {
  int64_t n_steps, rejected, total_nr, lu_factors, lu_refactors, body_iterations;
  double wall, min_dt_used, max_dt_used, t_build, t_lu, t_solve;
  int32_t dc_converged, dc_iterations, threads;
};

const char* ct_last_error(void) { return g_err.c_str(); }

}  // extern "C"

namespace {

// ---------------------------------------------------------------------------
// Sparse LU: left-looking Gilbert-Peierls with threshold pivoting, then
// refactorisation on the fixed pattern / pivot order.
// ---------------------------------------------------------------------------
struct SparseLU {
  int n = 0;
  std::vector<int64_t> q;                       // column order
  std::vector<int64_t> cp;                      // CSC of A: column pointers
  std::vector<int32_t> ci;                      // row indices
  std::vector<int64_t> csrc;                    // index into the CSR data
  std::vector<int64_t> Lp, Up;                  // column pointers
  std::vector<int32_t> Li, Ui;                  // pivot-space row indices
  std::vector<double> Lx, Ux, Ud;               // values, U diagonal
  std::vector<int32_t> pinv;                    // original row -> pivot index
  std::vector<double> x;                        // work (original rows / pivot space)
  std::vector<int32_t> xi, stack, mark;         // DFS work
  std::vector<int64_t> pstack;
  bool factored = false;
  int64_t n_factor = 0, n_refactor = 0;
  double tol = 0.1;

  void init(int n_, const int64_t* indptr, const int64_t* indices, const int64_t* order) {
    n = n_;
    q.assign(order, order + n);
    const int64_t nnz = indptr[n];
    cp.assign(n + 1, 0);
    for (int64_t k = 0; k < nnz; ++k) cp[indices[k] + 1]++;
    for (int j = 0; j < n; ++j) cp[j + 1] += cp[j];
    ci.resize(nnz);
    csrc.resize(nnz);
    std::vector<int64_t> next(cp.begin(), cp.end() - 1);
    for (int r = 0; r < n; ++r)
      for (int64_t k = indptr[r]; k < indptr[r + 1]; ++k) {
        const int64_t d = next[indices[k]]++;
        ci[d] = r;
        csrc[d] = k;
      }
    x.assign(n, 0.0);
    xi.resize(n);
    stack.resize(n);
    pstack.resize(n);
    mark.assign(n, -1);
    pinv.assign(n, -1);
  }

  // Full factorisation with pivot search.  Returns false if a column has no
  // usable pivot.
  bool factor(const double* a) {
    Lp.assign(1, 0);
    Up.assign(1, 0);
    Li.clear(); Lx.clear(); Ui.clear(); Ux.clear();
    Ud.assign(n, 0.0);
    std::fill(pinv.begin(), pinv.end(), -1);
    std::fill(mark.begin(), mark.end(), -1);
    // L rows are kept in original numbering during the factorisation
    for (int k = 0; k < n; ++k) {
      const int64_t j = q[k];
      // reach of A(:,j) through the columns of L computed so far, in
      // topological order: xi[top..n-1]
      int top = n;
      for (int64_t p = cp[j]; p < cp[j + 1]; ++p) {
        const int r0 = ci[p];
        if (mark[r0] == k) continue;
        int head = 0;
        stack[0] = r0;
        while (head >= 0) {
          const int r = stack[head];
          const int kk = pinv[r];
          if (mark[r] != k) {
            mark[r] = k;
            pstack[head] = kk < 0 ? 0 : Lp[kk];
          }
          bool done = true;
          if (kk >= 0) {
            const int64_t end = Lp[kk + 1];
            for (int64_t t = pstack[head]; t < end; ++t) {
              const int rr = Li[t];
              if (mark[rr] == k) continue;
              pstack[head] = t + 1;
              stack[++head] = rr;
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
      for (int t = top; t < n; ++t) x[xi[t]] = 0.0;
      for (int64_t p = cp[j]; p < cp[j + 1]; ++p) x[ci[p]] = a[csrc[p]];
      for (int t = top; t < n; ++t) {
        const int r = xi[t];
        const int kk = pinv[r];
        if (kk < 0) continue;
        const double v = x[r];
        for (int64_t s = Lp[kk]; s < Lp[kk + 1]; ++s) x[Li[s]] -= Lx[s] * v;
      }
      int ipiv = -1;
      double amax = -1.0;
      for (int t = top; t < n; ++t) {
        const int r = xi[t];
        if (pinv[r] < 0) {
          const double av = std::fabs(x[r]);
          if (av > amax) { amax = av; ipiv = r; }
        } else {
          Ui.push_back(pinv[r]);
          Ux.push_back(x[r]);
        }
      }
      if (ipiv < 0 || !(amax > 0.0) || !std::isfinite(amax)) {
        g_err = "singular matrix at pivot step " + std::to_string(k);
        return false;
      }
      if (pinv[j] < 0 && mark[j] == k && std::fabs(x[j]) >= tol * amax) ipiv = (int)j;
      const double piv = x[ipiv];
      Ud[k] = piv;
      pinv[ipiv] = k;
      for (int t = top; t < n; ++t) {
        const int r = xi[t];
        if (pinv[r] < 0) {
          Li.push_back(r);
          Lx.push_back(x[r] / piv);
        }
      }
      Lp.push_back((int64_t)Li.size());
      Up.push_back((int64_t)Ui.size());
    }
    for (auto& r : Li) r = pinv[r];
    factored = true;
    ++n_factor;
    return true;
  }

  // Numeric refactorisation on the stored pattern and pivot order.  Returns
  // false when a pivot collapses (the caller then factors afresh).
  bool refactor(const double* a) {
    for (int k = 0; k < n; ++k) {
      const int64_t j = q[k];
      for (int64_t s = Up[k]; s < Up[k + 1]; ++s) x[Ui[s]] = 0.0;
      for (int64_t s = Lp[k]; s < Lp[k + 1]; ++s) x[Li[s]] = 0.0;
      x[k] = 0.0;
      for (int64_t p = cp[j]; p < cp[j + 1]; ++p) x[pinv[ci[p]]] = a[csrc[p]];
      for (int64_t s = Up[k]; s < Up[k + 1]; ++s) {
        const int pp = Ui[s];
        const double v = x[pp];
        Ux[s] = v;
        for (int64_t t = Lp[pp]; t < Lp[pp + 1]; ++t) x[Li[t]] -= Lx[t] * v;
      }
      const double piv = x[k];
      double amax = std::fabs(piv);
      for (int64_t s = Lp[k]; s < Lp[k + 1]; ++s) amax = std::max(amax, std::fabs(x[Li[s]]));
      if (!(std::fabs(piv) > 1e-8 * amax) || !std::isfinite(amax)) return false;
      Ud[k] = piv;
      for (int64_t s = Lp[k]; s < Lp[k + 1]; ++s) Lx[s] = x[Li[s]] / piv;
    }
    ++n_refactor;
    return true;
  }

  bool numeric(const double* a) {
    if (factored && refactor(a)) return true;
    return factor(a);
  }

  // x_out = A^{-1} b
  void solve(const double* b, double* out) {
    for (int i = 0; i < n; ++i) x[pinv[i]] = b[i];
    for (int k = 0; k < n; ++k) {
      const double v = x[k];
      if (v != 0.0)
        for (int64_t s = Lp[k]; s < Lp[k + 1]; ++s) x[Li[s]] -= Lx[s] * v;
    }
    for (int k = n - 1; k >= 0; --k) {
      const double v = x[k] / Ud[k];
      x[k] = v;
      if (v != 0.0)
        for (int64_t s = Up[k]; s < Up[k + 1]; ++s) x[Ui[s]] -= Ux[s] * v;
    }
    for (int k = 0; k < n; ++k) out[q[k]] = x[k];
  }
};

// ---------------------------------------------------------------------------
// Device models (oracle/models.py)
// ---------------------------------------------------------------------------
inline double softplus(double u) { return std::max(u, 0.0) + std::log1p(std::exp(-std::fabs(u))); }
inline double sigmoid(double u) {
  const double e = std::exp(-std::fabs(u));
  return u >= 0.0 ? 1.0 / (1.0 + e) : e / (1.0 + e);
}

// EKV long-channel drain current in the NMOS frame, analytic derivatives,
// linear terminal charges (ekv_longchannel.va:58-96 with the oracle's
// charge model).  k = sgn, vgoff, phi, gamma, beta0, theta, cgs, cgd, cgb, cdb, csb.
inline void eval_mosfet(double vgb, double vsb, double vdb, const double* k, double rr[4],
                        double rc[4], double jr[16], double jc[16]) {
  const double sgn = k[0], vgoff = k[1], phi = k[2], gamma = k[3], beta0 = k[4], theta = k[5];
  const double VG = sgn * vgb, VS = sgn * vsb, VD = sgn * vdb;
  const double vgp = VG + vgoff;
  const double g2 = 0.5 * gamma;
  const double t = vgp + g2 * g2;
  const bool big = t > 1e-12;
  const double s = std::sqrt(big ? t : 1.0);
  const double vp = big ? vgp - phi - gamma * (s - g2) : vgp - phi - gamma * gamma / 2.0;
  const double dvp = big ? 1.0 - gamma / (2.0 * s) : 1.0;
  const double arg = phi + vp + 4.0 * VT;
  const bool pos = arg > 1e-24;
  const double sa = std::sqrt(pos ? arg : 1.0);
  const double n = pos ? 1.0 + gamma / (2.0 * sa) : 1.0 + gamma / 2.0 * 1e48;
  const double dn = pos ? -gamma / (4.0 * arg * sa) : 0.0;
  const double den = 1.0 + theta * vp;
  const double beta = beta0 / den;
  const double dbeta = -beta0 * theta / (den * den);
  const double uf = 0.5 * ((vp - VS) / VT);
  const double ur = 0.5 * ((vp - VD) / VT);
  const double lf = softplus(uf), lr = softplus(ur);
  const double fpf = lf * sigmoid(uf), fpr = lr * sigmoid(ur);
  const double vt2 = VT * VT;
  const double ispec = 2.0 * n * beta * vt2;
  const double idiff = lf * lf - lr * lr;
  const double i_d = ispec * idiff;
  const double dispec = 2.0 * vt2 * (dn * beta + n * dbeta);
  const double gvp = dispec * idiff + ispec * (fpf - fpr) / VT;
  const double gg = gvp * dvp;
  const double gs = -ispec * fpf / VT;
  const double gd = ispec * fpr / VT;
  const double gb = -(gg + gs + gd);
  const double iext = sgn * i_d;
  rr[0] = iext; rr[1] = 0.0; rr[2] = -iext; rr[3] = 0.0;
  const double jrv[16] = {gd, gg, gs, gb, 0, 0, 0, 0, -gd, -gg, -gs, -gb, 0, 0, 0, 0};
  std::memcpy(jr, jrv, sizeof(jrv));
  const double cgs = k[6], cgd = k[7], cgb = k[8], cdb = k[9], csb = k[10];
  const double vgs = vgb - vsb, vgd = vgb - vdb;
  rc[1] = cgs * vgs + cgd * vgd + cgb * vgb;
  rc[0] = -cgd * vgd + cdb * vdb;
  rc[2] = -cgs * vgs + csb * vsb;
  rc[3] = -(cgb * vgb + cdb * vdb + csb * vsb);
  const double jcv[16] = {cgd + cdb, -cgd, 0.0, -cdb,
                          -cgd, cgs + cgd + cgb, -cgs, -cgb,
                          0.0, -cgs, cgs + csb, -csb,
                          -cdb, -cgb, -csb, cdb + csb + cgb};
  std::memcpy(jc, jcv, sizeof(jcv));
}

// oracle/sources.py (vajax/analysis/sources.py:31-63)
double pulse(double t, const double* p) {
  const double v0 = p[1], v1 = p[2], rise = p[3], fall = p[4], width = p[5], period = p[6],
               delay = p[7];
  double t_in = t >= delay ? std::fmod(t - delay, period) : 0.0;
  if (t_in < 0.0) t_in += period;
  if (t < delay) return v0;
  if (t_in < rise) return v0 + (v1 - v0) * t_in / rise;
  if (t_in < rise + width) return v1;
  if (t_in < rise + width + fall) return v1 - (v1 - v0) * (t_in - rise - width) / fall;
  return v0;
}

void source_values(const ct_sources& s, double t, double* out) {
  for (int i = 0; i < s.count; ++i) {
    const double* p = s.par + (size_t)i * 8;
    out[i] = s.wave[i] == 0 ? p[0] : pulse(t, p);
  }
}

// ---------------------------------------------------------------------------
// Circuit and system build
// ---------------------------------------------------------------------------
struct Model {
  int kind, count, T;
  std::vector<int64_t> nodes, csr_pos, res_idx;
  std::vector<double> par;
  int64_t jac_off, res_off;  // offsets into the flat contribution arrays
};

struct Built {
  std::vector<double> csr, f, Q, Ivs, mrc;
};

struct Circuit {
  int n_total, n_unk, m, n_aug, k_is;
  int64_t nse;
  std::vector<int64_t> indptr, indices, diag_pos, b_pos;
  std::vector<double> b_val;
  std::vector<int64_t> vs_p, vs_n, is_p, is_n;
  std::vector<int32_t> vs_wave, is_wave;
  std::vector<double> vs_par, is_par;
  ct_sources vs{}, is{};
  std::vector<Model> models;
  double gmin;
  // gather lists: CSR slot -> jac contributions; residual row -> res
  // contributions (isource stamps first, then devices)
  std::vector<int64_t> jptr, jsrc, rptr, rsrc;
  std::vector<double> jval, rval, qval;  // contributions (qval indexed like rval)
  int64_t n_jc = 0, n_rc = 0;
  SparseLU lu;
  ct_stats st{};

  void build(const double* X, const double* vsrc, const double* isrc, const double* Q_prev,
             double c0, double gshunt, double c1, double d1, const double* dQ, double c2,
             const double* Q2, Built& out) {
    const auto t0 = clk::now();
    const double* V = X;
    const double* Ib = X + n_total;
    // isource stamps: +I at p, -I at n (mna_builder.py:830-837)
    for (int i = 0; i < k_is; ++i) {
      rval[2 * i] = nan0(isrc[i]);
      rval[2 * i + 1] = nan0(-isrc[i]);
      qval[2 * i] = 0.0;
      qval[2 * i + 1] = 0.0;
    }
    for (auto& md : models) {
      double* jv = jval.data() + md.jac_off;
      double* rv = rval.data() + md.res_off;
      double* qv = qval.data() + md.res_off;
      const int64_t* nd = md.nodes.data();
      const double* pr = md.par.data();
      const int cnt = md.count;
      if (md.kind == K_RES || md.kind == K_CAP) {
#pragma omp parallel for schedule(static)
        for (int d = 0; d < cnt; ++d) {
          const double vab = V[nd[2 * d]] - V[nd[2 * d + 1]];
          if (md.kind == K_RES) {
            const double i = vab / pr[d];
            const double g = 1.0 / pr[d];
            rv[2 * d] = nan0(i); rv[2 * d + 1] = nan0(-i);
            qv[2 * d] = 0.0; qv[2 * d + 1] = 0.0;
            const double gj[4] = {g, -g, -g, g};
            for (int e = 0; e < 4; ++e) jv[4 * d + e] = nan0(gj[e] + c0 * 0.0);
          } else {
            const double qq = pr[d] * vab;
            rv[2 * d] = 0.0; rv[2 * d + 1] = 0.0;
            qv[2 * d] = nan0(qq); qv[2 * d + 1] = nan0(-qq);
            const double cj[4] = {pr[d], -pr[d], -pr[d], pr[d]};
            for (int e = 0; e < 4; ++e) jv[4 * d + e] = nan0(0.0 + c0 * cj[e]);
          }
        }
      } else {
#pragma omp parallel for schedule(static)
        for (int d = 0; d < cnt; ++d) {
          const int64_t* nn = nd + 4 * d;
          const double vgb = V[nn[1]] - V[nn[3]];
          const double vsb = V[nn[2]] - V[nn[3]];
          const double vdb = V[nn[0]] - V[nn[3]];
          double rr[4], rc[4], jr[16], jc[16];
          eval_mosfet(vgb, vsb, vdb, pr + 11 * (size_t)d, rr, rc, jr, jc);
          for (int e = 0; e < 4; ++e) {
            rv[4 * d + e] = nan0(rr[e]);
            qv[4 * d + e] = nan0(rc[e]);
          }
          for (int e = 0; e < 16; ++e) jv[16 * d + e] = nan0(jr[e] + c0 * jc[e]);
        }
      }
    }
    out.csr.resize(nse);
    double* csr = out.csr.data();
    const double gd = gmin + gshunt;
#pragma omp parallel for schedule(static, 4096)
    for (int64_t s = 0; s < nse; ++s) {
      double acc = 0.0;
      for (int64_t t = jptr[s]; t < jptr[s + 1]; ++t) acc += jval[jsrc[t]];
      csr[s] = acc;
    }
    for (int i = 0; i < n_unk; ++i) csr[diag_pos[i]] += gd;
    for (size_t t = 0; t < b_pos.size(); ++t) csr[b_pos[t]] += b_val[t];

    out.f.resize(n_aug);
    out.Q.resize(n_unk);
    out.mrc.resize(n_unk);
    out.Ivs.resize(m);
    std::vector<double>& fres = fres_;
    fres.resize(n_unk);
#pragma omp parallel for schedule(static, 4096)
    for (int r = 0; r < n_unk; ++r) {
      double a = 0.0, q = 0.0, mx = 0.0;
      for (int64_t t = rptr[r]; t < rptr[r + 1]; ++t) {
        const int64_t c = rsrc[t];
        a += rval[c];
        mx = std::max(mx, std::fabs(rval[c]));
        if (c >= 2 * (int64_t)k_is) q += qval[c];
      }
      fres[r] = a - 0.0;
      out.Q[r] = q;
      out.mrc[r] = mx;
    }
    for (int i = 0; i < m; ++i) out.Ivs[i] = vs_p[i] > 0 ? -fres[vs_p[i] - 1] : 0.0;
    if (m) {
      // contrib via the sequential scatter of mna.py:489-536 (p entries, then n)
      std::vector<double>& contrib = contrib_;
      contrib.assign(n_unk, 0.0);
      for (int i = 0; i < m; ++i)
        contrib[vs_p[i] > 0 ? vs_p[i] - 1 : 0] += vs_p[i] > 0 ? Ib[i] : 0.0;
      for (int i = 0; i < m; ++i)
        contrib[vs_n[i] > 0 ? vs_n[i] - 1 : 0] += vs_n[i] > 0 ? -Ib[i] : 0.0;
      for (int r = 0; r < n_unk; ++r) fres[r] = fres[r] + contrib[r];
    }
    double* f = out.f.data();
#pragma omp parallel for schedule(static, 4096)
    for (int r = 0; r < n_unk; ++r) {
      double v = fres[r] - 0.0 + c0 * (out.Q[r] - 0.0) + c1 * Q_prev[r] + d1 * dQ[r] + c2 * Q2[r];
      f[r] = v + gd * V[r + 1];
    }
    for (int i = 0; i < m; ++i) f[n_unk + i] = V[vs_p[i]] - V[vs_n[i]] - vsrc[i];
    st.t_build += secs(t0);
  }

  std::vector<double> fres_, contrib_;
};

// ---------------------------------------------------------------------------
// Newton (oracle/transient.py nr_solve)
// ---------------------------------------------------------------------------
struct NR {
  std::vector<double> X, Q, dQdt, Ivs, mrc;
  int iterations = 0;
  bool converged = false;
  double max_f = 0.0;
  bool singular = false;
};

bool nr_solve(Circuit& c, const ct_options& o, const double* X_init, const double* vsrc,
              const double* isrc, const double* Q_prev, double c0, double gshunt, double c1,
              double d1, const double* dQ, double c2, const double* Q2, const double* floor,
              NR& r, Built& b) {
  const int n_total = c.n_total, n_unk = c.n_unk, m = c.m, n_aug = c.n_aug;
  r.X.assign(X_init, X_init + n_total + m);
  r.iterations = 0;
  r.converged = false;
  r.max_f = std::numeric_limits<double>::infinity();
  std::vector<double> rhs(n_aug), delta(n_aug);
  while (!r.converged && r.iterations < o.tran_itl) {
    c.build(r.X.data(), vsrc, isrc, Q_prev, c0, gshunt, c1, d1, dQ, c2, Q2, b);
    const double* f = b.f.data();
    double maxf = 0.0;
    int res_bad = 0;
#pragma omp parallel for reduction(max : maxf) reduction(| : res_bad) schedule(static, 4096)
    for (int i = 0; i < n_aug; ++i) {
      const double af = std::fabs(f[i]);
      maxf = std::max(maxf, af);
      const double tol = i < n_unk ? std::max(b.mrc[i] * o.reltol, floor[i]) : o.vntol;
      if (!(af < tol)) res_bad |= 1;
    }
    r.max_f = n_aug ? maxf : 0.0;
    for (int i = 0; i < n_aug; ++i) rhs[i] = -f[i];
    auto tl = clk::now();
    if (!c.lu.numeric(b.csr.data())) {
      r.singular = true;
      return false;
    }
    c.st.t_lu += secs(tl);
    tl = clk::now();
    c.lu.solve(rhs.data(), delta.data());
    c.st.t_solve += secs(tl);
    double max_vd = 0.0;
    int del_bad = 0;
    const double* X = r.X.data();
    const int it = r.iterations;
#pragma omp parallel for reduction(max : max_vd) reduction(| : del_bad) schedule(static, 4096)
    for (int i = 0; i < n_aug; ++i) {
      const bool node = i < n_unk;
      const double cd = node ? delta[i] * o.nr_damping : delta[i];
      const double xr = node ? X[i + 1] : X[n_total + (i - n_unk)];
      const double tol = std::max(std::fabs(xr) * o.reltol, node ? o.vntol : o.abstol);
      if (it != 0 && !(std::fabs(cd) < tol)) del_bad |= 1;
      if (node) max_vd = std::max(max_vd, std::fabs(delta[i]));
    }
    const bool delta_ok = (it == 0) || !del_bad;
    const double v_scale = std::min(1.0, 1e30 / std::max(n_unk ? max_vd : 0.0, 1e-30));
    double* Xw = r.X.data();
    for (int i = 0; i < n_unk; ++i) Xw[i + 1] += delta[i] * v_scale * o.nr_damping;
    for (int i = 0; i < m; ++i) Xw[n_total + i] += delta[n_unk + i];
    r.converged = !res_bad && delta_ok;
    r.iterations += 1;
  }
  c.build(r.X.data(), vsrc, isrc, Q_prev, c0, gshunt, c1, d1, dQ, c2, Q2, b);
  r.Q = b.Q;
  r.Ivs = b.Ivs;
  r.mrc = b.mrc;
  r.dQdt.resize(n_unk);
  for (int i = 0; i < n_unk; ++i)
    r.dQdt[i] = c0 * b.Q[i] + c1 * Q_prev[i] + d1 * dQ[i] + c2 * Q2[i];
  return true;
}

// DC operating point or UIC start (oracle/transient.py initial_state).
bool initial_state(Circuit& c, const ct_options& o, const double* x_init, std::vector<double>& X,
                   std::vector<double>& Q, std::vector<double>& Ivs, std::vector<double>& mrc) {
  const int n_total = c.n_total, n_unk = c.n_unk, m = c.m;
  std::vector<double> vs0(m), is0(c.k_is), zeros(n_unk, 0.0);
  source_values(c.vs, 0.0, vs0.data());
  source_values(c.is, 0.0, is0.data());
  X.assign(x_init, x_init + n_total + m);
  Built b;
  if (o.uic) {
    std::fill(X.begin(), X.begin() + n_total, 0.0);
    c.build(X.data(), vs0.data(), is0.data(), zeros.data(), 0.0, 0.0, 0.0, 0.0, zeros.data(), 0.0,
            zeros.data(), b);
    Q = b.Q;
    Ivs = b.Ivs;
    mrc.assign(n_unk, 0.0);
    c.st.dc_converged = 1;
    c.st.dc_iterations = 0;
    return true;
  }
  std::vector<double> floor(n_unk, o.abstol);
  NR r;
  if (!nr_solve(c, o, X.data(), vs0.data(), is0.data(), zeros.data(), 0.0, 0.0, 0.0, 0.0,
                zeros.data(), 0.0, zeros.data(), floor.data(), r, b))
    return false;
  X = r.X;
  Q = r.Q;
  Ivs = r.Ivs;
  mrc = r.mrc;
  c.st.dc_converged = r.converged;
  c.st.dc_iterations = r.iterations;
  return true;
}

void coeffs(int method, double dt_cur, double dt_prev, double& c0, double& c1, double& d1,
            double& c2, double& e) {
  const double inv = 1.0 / dt_cur;
  if (method == M_BE) {
    c0 = inv; c1 = -inv; d1 = 0.0; c2 = 0.0; e = -0.5;
  } else if (method == M_TRAP) {
    c0 = 2.0 * inv; c1 = -2.0 * inv; d1 = -1.0; c2 = 0.0; e = 1.0 / 24.0;
  } else {
    const double om = dt_prev > 0 ? dt_cur / dt_prev : 1.0;
    const double op1 = 1.0 + om;
    c0 = (1.0 + 2.0 * om) / (op1 * dt_cur);
    c1 = -op1 / dt_cur;
    d1 = 0.0;
    c2 = (om * om) / (op1 * dt_cur);
    e = -2.0 / 9.0;
  }
}

}  // namespace

extern "C" {

int ct_create(const ct_circuit_desc* d, void** out) {
  if (!d || !out) {
    g_err = "ct_create: null argument";
    return -1;
  }
  auto* c = new Circuit();
  c->n_total = d->n_total;
  c->n_unk = d->n_unk;
  c->m = d->m;
  c->n_aug = d->n_unk + d->m;
  c->nse = d->nse;
  c->gmin = d->gmin;
  c->indptr.assign(d->indptr, d->indptr + c->n_aug + 1);
  c->indices.assign(d->indices, d->indices + c->nse);
  c->diag_pos.assign(d->diag_pos, d->diag_pos + c->n_unk);
  c->b_pos.assign(d->b_pos, d->b_pos + d->nb);
  c->b_val.assign(d->b_val, d->b_val + d->nb);
  auto copy_src = [](const ct_sources& s, std::vector<int64_t>& p, std::vector<int64_t>& n,
                     std::vector<int32_t>& w, std::vector<double>& par, ct_sources& o) {
    p.assign(s.p, s.p + s.count);
    n.assign(s.n, s.n + s.count);
    w.assign(s.wave, s.wave + s.count);
    par.assign(s.par, s.par + (size_t)s.count * 8);
    o.count = s.count;
    o.p = p.data();
    o.n = n.data();
    o.wave = w.data();
    o.par = par.data();
  };
  copy_src(d->vs, c->vs_p, c->vs_n, c->vs_wave, c->vs_par, c->vs);
  copy_src(d->is, c->is_p, c->is_n, c->is_wave, c->is_par, c->is);
  c->k_is = d->is.count;
  int64_t joff = 0, roff = 2 * (int64_t)c->k_is;
  for (int i = 0; i < d->n_models; ++i) {
    const ct_model& s = d->models[i];
    Model md;
    md.kind = s.kind;
    md.count = s.count;
    md.T = s.kind == K_MOS ? 4 : 2;
    const size_t N = (size_t)s.count;
    md.nodes.assign(s.nodes, s.nodes + N * md.T);
    md.csr_pos.assign(s.csr_pos, s.csr_pos + N * md.T * md.T);
    md.res_idx.assign(s.res_idx, s.res_idx + N * md.T);
    md.par.assign(s.par, s.par + N * (s.kind == K_MOS ? 11 : 1));
    md.jac_off = joff;
    md.res_off = roff;
    joff += (int64_t)N * md.T * md.T;
    roff += (int64_t)N * md.T;
    c->models.push_back(std::move(md));
  }
  c->n_jc = joff;
  c->n_rc = roff;
  c->jval.assign(joff, 0.0);
  c->rval.assign(roff, 0.0);
  c->qval.assign(roff, 0.0);
  // gather lists by counting sort; stable in (model, instance, key) order
  c->jptr.assign(c->nse + 1, 0);
  c->rptr.assign(c->n_unk + 1, 0);
  for (int i = 0; i < c->k_is; ++i) {
    if (c->is_p[i] > 0) c->rptr[c->is_p[i]]++;
    if (c->is_n[i] > 0) c->rptr[c->is_n[i]]++;
  }
  for (auto& md : c->models) {
    for (auto p : md.csr_pos)
      if (p >= 0) c->jptr[p + 1]++;
    for (auto r : md.res_idx)
      if (r >= 0) c->rptr[r + 1]++;
  }
  for (int64_t s = 0; s < c->nse; ++s) c->jptr[s + 1] += c->jptr[s];
  for (int r = 0; r < c->n_unk; ++r) c->rptr[r + 1] += c->rptr[r];
  c->jsrc.resize(c->jptr[c->nse]);
  c->rsrc.resize(c->rptr[c->n_unk]);
  std::vector<int64_t> jn(c->jptr.begin(), c->jptr.end() - 1), rn(c->rptr.begin(), c->rptr.end() - 1);
  for (int i = 0; i < c->k_is; ++i) {
    if (c->is_p[i] > 0) c->rsrc[rn[c->is_p[i] - 1]++] = 2 * i;
    if (c->is_n[i] > 0) c->rsrc[rn[c->is_n[i] - 1]++] = 2 * i + 1;
  }
  for (auto& md : c->models) {
    for (size_t e = 0; e < md.csr_pos.size(); ++e)
      if (md.csr_pos[e] >= 0) c->jsrc[jn[md.csr_pos[e]]++] = md.jac_off + (int64_t)e;
    for (size_t e = 0; e < md.res_idx.size(); ++e)
      if (md.res_idx[e] >= 0) c->rsrc[rn[md.res_idx[e]]++] = md.res_off + (int64_t)e;
  }
  c->lu.init(c->n_aug, c->indptr.data(), c->indices.data(), d->col_order);
  *out = c;
  return 0;
}

void ct_destroy(void* h) { delete static_cast<Circuit*>(h); }

// One system build at X (the oracle's mna.build_system); for tests.
int ct_build_system(void* h, const double* X, const double* vsrc, const double* isrc,
                    const double* Q_prev, double c0, double gshunt, double c1, double d1,
                    const double* dQ, double c2, const double* Q2, double* csr, double* f,
                    double* Q, double* Ivs, double* mrc) {
  auto* c = static_cast<Circuit*>(h);
  Built b;
  c->build(X, vsrc, isrc, Q_prev, c0, gshunt, c1, d1, dQ, c2, Q2, b);
  std::memcpy(csr, b.csr.data(), sizeof(double) * c->nse);
  std::memcpy(f, b.f.data(), sizeof(double) * c->n_aug);
  std::memcpy(Q, b.Q.data(), sizeof(double) * c->n_unk);
  if (c->m) std::memcpy(Ivs, b.Ivs.data(), sizeof(double) * c->m);
  std::memcpy(mrc, b.mrc.data(), sizeof(double) * c->n_unk);
  return 0;
}

// A x = b with the circuit's pattern (numeric factorisation or refactorisation
// on the stored pivot order, as inside Newton); for tests.
int ct_lu_solve(void* h, const double* csr, const double* b, double* x, int32_t fresh,
                int64_t* counts) {
  auto* c = static_cast<Circuit*>(h);
  if (fresh) c->lu.factored = false;
  if (!c->lu.numeric(csr)) return -2;
  c->lu.solve(b, x);
  if (counts) {
    counts[0] = c->lu.n_factor;
    counts[1] = c->lu.n_refactor;
    counts[2] = c->lu.Lp.back();
    counts[3] = c->lu.Up.back();
  }
  return 0;
}

// Adaptive transient (oracle/transient.py run_adaptive).  times [max_steps],
// V [max_steps * n_total], I [max_steps * max(m, 1)].
int ct_run_adaptive(void* h, double t_stop, double dt, const ct_options* o, const ct_adaptive* a,
                    int64_t max_steps, const double* x_init, int32_t threads, double* times,
                    double* Vout, double* Iout, ct_stats* stats) {
  auto* c = static_cast<Circuit*>(h);
  if (threads > 0) omp_set_num_threads(threads);
  c->st = ct_stats{};
  c->st.threads = omp_get_max_threads();
  c->lu.factored = false;
  c->lu.n_factor = c->lu.n_refactor = 0;
  const int n_total = c->n_total, n_unk = c->n_unk, m = c->m, mo = std::max(m, 1);
  double max_dt = a->max_dt;
  if (a->tran_minpts > 0) {
    const double eff = std::min(max_dt, t_stop / a->tran_minpts);
    if (eff < max_dt) max_dt = eff;
  }
  if (a->tran_fs != 1.0) dt = dt * a->tran_fs;
  std::vector<double> X, Q_prev, Ivs, hist_mrc;
  if (!initial_state(*c, *o, x_init, X, Q_prev, Ivs, hist_mrc)) return -2;
  const int max_hist = a->max_order + 2;
  std::vector<std::vector<double>> V_hist(max_hist, std::vector<double>(n_total, 0.0));
  std::copy(X.begin(), X.begin() + n_total, V_hist[0].begin());
  std::vector<double> dt_hist(max_hist, dt);
  int hist_count = 1;
  times[0] = 0.0;
  std::copy(X.begin(), X.begin() + n_total, Vout);
  for (int i = 0; i < mo; ++i) Iout[i] = m ? Ivs[i] : 0.0;
  double t = 0.0;
  int64_t step_idx = 1;
  int warmup = 0;
  std::vector<double> dQdt_prev(n_unk, 0.0), Q_prev2(n_unk, 0.0);
  int64_t total_nr = 0, rejected = 0;
  int consec = 0;
  double min_dt_used = dt, max_dt_used = dt;
  std::vector<double> V_max_hist(n_total);
  for (int i = 0; i < n_total; ++i) V_max_hist[i] = std::fabs(X[i]);
  std::vector<double> vs(m), isv(c->k_is), V_pred(n_total), X_init(n_total + m), floor(n_unk);
  NR r;
  Built b;
  const auto t0 = clk::now();
  int64_t iters = 0;
  while (t < t_stop && step_idx < max_steps) {
    ++iters;
    const double dt_cur = dt;
    if (dt_cur <= 0.0) {
      total_nr += o->tran_itl;
      consec = 0;
      const double last_good = hist_count > 0 ? dt_hist[0] : a->min_dt * 16.0;
      double new_dt = std::max(last_good / 8.0, a->min_dt * 2.0);
      new_dt = std::min(std::max(new_dt, a->min_dt), max_dt);
      new_dt = std::min(new_dt, t_stop - t);
      times[step_idx] = t;
      std::copy(X.begin(), X.begin() + n_total, Vout + step_idx * n_total);
      for (int i = 0; i < mo; ++i) Iout[step_idx * mo + i] = 0.0;
      ++step_idx;
      min_dt_used = std::min(min_dt_used, dt_cur);
      max_dt_used = std::max(max_dt_used, dt_cur);
      dt = new_dt;
      continue;
    }
    double c0, c1, d1, c2, e_int;
    coeffs(a->method, dt_cur, dt_hist[0], c0, c1, d1, c2, e_int);
    const double t_next = t + dt_cur;
    source_values(c->vs, t_next, vs.data());
    source_values(c->is, t_next, isv.data());
    const bool can_predict = (warmup >= a->warmup_steps) && (hist_count >= 2);
    // predictor (oracle/transient.py predict)
    const int order = std::max(std::min(hist_count - 1, a->max_order), 0);
    const double cs0 = dt_hist[0], cs1 = dt_hist[0] + dt_hist[1];
    const double tau1 = -(dt_cur + cs0) / dt_cur, tau2 = -(dt_cur + cs1) / dt_cur;
    const double ratio1 = dt_cur / std::max(dt_hist[0], 1e-30);
    const double err1 = -ratio1 * (1 + ratio1) / 2;
    const double tau0 = -1.0;
    double den0 = (tau0 - tau1) * (tau0 - tau2);
    double den1 = (tau1 - tau0) * (tau1 - tau2);
    double den2 = (tau2 - tau0) * (tau2 - tau1);
    if (std::fabs(den0) < 1e-30) den0 = 1e-30;
    if (std::fabs(den1) < 1e-30) den1 = 1e-30;
    if (std::fabs(den2) < 1e-30) den2 = 1e-30;
    const double l0 = (0 - tau1) * (0 - tau2) / den0;
    const double l1 = (0 - tau0) * (0 - tau2) / den1;
    const double l2 = (0 - tau0) * (0 - tau1) / den2;
    const double err2 = -tau0 * tau1 * tau2 / 6;
    double cf[3], pred_err;
    if (order == 0) {
      cf[0] = 1.0; cf[1] = 0.0; cf[2] = 0.0; pred_err = 1.0;
    } else if (order == 1) {
      cf[0] = 1.0 + ratio1; cf[1] = -ratio1; cf[2] = 0.0; pred_err = err1;
    } else {
      cf[0] = l0; cf[1] = l1; cf[2] = l2; pred_err = err2;
    }
    for (int j = order + 1; j < 3; ++j) cf[j] = 0.0;
    for (int i = 0; i < n_total; ++i)
      V_pred[i] = cf[0] * V_hist[0][i] + cf[1] * V_hist[1][i] + cf[2] * V_hist[2][i];
    std::copy(X.begin(), X.end(), X_init.begin());
    if (can_predict) std::copy(V_pred.begin(), V_pred.end(), X_init.begin());
    const double ramp = a->gshunt_steps > 0
                            ? std::min(std::max((double)step_idx / a->gshunt_steps, 0.0), 1.0)
                            : 1.0;
    const double gshunt = a->gshunt_init + ramp * (a->gshunt_target - a->gshunt_init);
    for (int i = 0; i < n_unk; ++i) floor[i] = std::max(hist_mrc[i] * a->reltol, a->abstol);
    if (!nr_solve(*c, *o, X_init.data(), vs.data(), isv.data(), Q_prev.data(), c0, gshunt, c1, d1,
                  dQdt_prev.data(), c2, Q_prev2.data(), floor.data(), r, b))
      return -2;
    total_nr += r.iterations;
    const bool nr_failed = !r.converged;
    const bool at_min = dt_cur <= a->min_dt;
    const bool fail_min = nr_failed && at_min;
    // LTE (oracle/transient.py lte_timestep)
    const double ratio = e_int / (e_int - pred_err);
    double lte_norm = 0.0;
    for (int i = 0; i < n_total; ++i) {
      const double vn = r.X[i];
      const double lte = (vn - V_pred[i]) * ratio;
      const double vcur = std::max(std::fabs(vn), std::fabs(V_pred[i]));
      const double vscale = std::max(V_max_hist[i], vcur);
      const double tol = a->reltol * vscale + a->abstol;
      lte_norm = std::max(lte_norm, std::fabs(lte) / tol);
    }
    const int lorder = std::min(hist_count - 1, a->max_order);
    double factor = std::pow(a->lte_ratio / std::max(lte_norm, 1e-10), 1.0 / (lorder + 1));
    factor = std::min(std::max(factor, 0.1), a->grow_factor);
    const double dt_lte = dt_cur * factor;
    const bool lte_reject = (dt_cur / dt_lte > a->redo_factor) && can_predict && r.converged;
    const bool nr_reject = nr_failed && !at_min;
    const bool force = consec >= a->max_consecutive_rejects;
    const bool reject = (lte_reject || nr_reject) && !force;
    const bool accept = !reject;
    if (accept)
      consec = 0;
    else if (lte_reject)
      ++consec;
    const double dt_from_lte = can_predict ? dt_lte : dt_cur;
    const double dt_halved = std::max(dt_cur * a->tran_ft, a->min_dt);
    const double last_good = hist_count > 0 ? dt_hist[0] : a->min_dt * 16.0;
    const double dt_rec = std::max(last_good / 8.0, a->min_dt * 2.0);
    double new_dt = nr_failed ? (at_min ? dt_rec : dt_halved) : dt_from_lte;
    new_dt = std::min(std::max(new_dt, a->min_dt), max_dt);
    new_dt = std::min(new_dt, t_stop - t_next);
    const bool use_new = accept && !fail_min;
    if (accept) t = t_next;
    if (use_new) {
      X = r.X;
      Q_prev2 = Q_prev;
      Q_prev = r.Q;
      dQdt_prev = r.dQdt;
      std::rotate(V_hist.begin(), V_hist.end() - 1, V_hist.end());
      std::copy(r.X.begin(), r.X.begin() + n_total, V_hist[0].begin());
      for (int j = max_hist - 1; j > 0; --j) dt_hist[j] = dt_hist[j - 1];
      dt_hist[0] = dt_cur;
      hist_count = std::min(hist_count + 1, max_hist);
      ++warmup;
      for (int i = 0; i < n_total; ++i) V_max_hist[i] = std::max(V_max_hist[i], std::fabs(r.X[i]));
      for (int i = 0; i < n_unk; ++i) hist_mrc[i] = std::max(hist_mrc[i], r.mrc[i]);
    }
    if (accept) {
      times[step_idx] = t_next;
      std::copy(X.begin(), X.begin() + n_total, Vout + step_idx * n_total);
      for (int i = 0; i < mo; ++i)
        Iout[step_idx * mo + i] = (fail_min || !m) ? 0.0 : r.Ivs[i];
      ++step_idx;
      min_dt_used = std::min(min_dt_used, dt_cur);
      max_dt_used = std::max(max_dt_used, dt_cur);
    } else {
      ++rejected;
    }
    dt = new_dt;
  }
  c->st.wall = secs(t0);
  c->st.n_steps = step_idx;
  c->st.rejected = rejected;
  c->st.total_nr = total_nr;
  c->st.body_iterations = iters;
  c->st.min_dt_used = min_dt_used;
  c->st.max_dt_used = max_dt_used;
  c->st.lu_factors = c->lu.n_factor;
  c->st.lu_refactors = c->lu.n_refactor;
  *stats = c->st;
  return 0;
}

// Fixed-step transient (oracle/transient.py run_fixed): n_steps steps of dt
// (already scaled by tran_fs by the caller).
int ct_run_fixed(void* h, double dt, int32_t method, const ct_options* o, int64_t n_steps,
                 double abstol, const double* x_init, int32_t threads, double* times, double* Vout,
                 double* Iout, ct_stats* stats) {
  auto* c = static_cast<Circuit*>(h);
  if (threads > 0) omp_set_num_threads(threads);
  c->st = ct_stats{};
  c->st.threads = omp_get_max_threads();
  c->lu.factored = false;
  c->lu.n_factor = c->lu.n_refactor = 0;
  const int n_total = c->n_total, n_unk = c->n_unk, m = c->m, mo = std::max(m, 1);
  std::vector<double> X, Q_prev, Ivs, mrc;
  if (!initial_state(*c, *o, x_init, X, Q_prev, Ivs, mrc)) return -2;
  const double inv = 1.0 / dt;
  double c0 = inv, c1 = -inv, d1 = 0.0, c2 = 0.0;
  if (method == M_TRAP) {
    c0 = 2.0 * inv; c1 = -2.0 * inv; d1 = -1.0;
  } else if (method == M_GEAR2) {
    c0 = 1.5 * inv; c1 = -2.0 * inv; c2 = 0.5 * inv;
  }
  times[0] = 0.0;
  std::copy(X.begin(), X.begin() + n_total, Vout);
  for (int i = 0; i < mo; ++i) Iout[i] = 0.0;
  std::vector<double> dQdt_prev(n_unk, 0.0), Q_prev2(n_unk, 0.0), floor(n_unk, abstol);
  std::vector<double> vs(m), isv(c->k_is);
  int64_t total_nr = 0;
  NR r;
  Built b;
  const auto t0 = clk::now();
  for (int64_t i = 0; i < n_steps; ++i) {
    const double t_next = ((double)i + 1.0) * dt;
    source_values(c->vs, t_next, vs.data());
    source_values(c->is, t_next, isv.data());
    if (!nr_solve(*c, *o, X.data(), vs.data(), isv.data(), Q_prev.data(), c0, 0.0, c1, d1,
                  dQdt_prev.data(), c2, Q_prev2.data(), floor.data(), r, b))
      return -2;
    total_nr += r.iterations;
    times[i + 1] = t_next;
    std::copy(r.X.begin(), r.X.begin() + n_total, Vout + (i + 1) * n_total);
    for (int k = 0; k < mo; ++k) Iout[(i + 1) * mo + k] = m ? r.Ivs[k] : 0.0;
    X = r.X;
    Q_prev2 = Q_prev;
    Q_prev = r.Q;
    dQdt_prev = r.dQdt;
  }
  c->st.wall = secs(t0);
  c->st.n_steps = n_steps + 1;
  c->st.total_nr = total_nr;
  c->st.lu_factors = c->lu.n_factor;
  c->st.lu_refactors = c->lu.n_refactor;
  *stats = c->st;
  return 0;
}

}  // extern "C"
