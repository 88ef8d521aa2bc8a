// HIP kernels of the transient core for gfx950 (MI355X).
//
// Stamping: one lane per device instance reads structure-of-arrays node
// indices and parameters (coalesced), evaluates the model and writes every
// Jacobian / residual contribution to its slot in a contribution list that is
// pre-sorted by CSR slot (resp. residual row).  A second kernel reduces each
// slot's contiguous segment in a fixed order, so csr_data and f are
// bit-identical run to run without atomics.
//
// LU: right-looking numeric refactorisation over a level schedule computed
// on the host.  Every item of a level is one update target whose segment of
// contributions L(i,k)*U(k,j)/U(k,k) is summed in fixed order; wide levels run
// as a grid, runs of narrow levels as one block with barriers between levels.
// Triangular solves use row level sets in the same way.
//
// Compiled with -ffp-contract=off so every FP64 expression rounds exactly as
// written (and as the numpy oracle evaluates it).
#include "block_flags.h"
#include "circuit.h"
#include "internal.h"

#include <algorithm>
#include <cmath>

namespace vjx {

namespace {

constexpr double KB = 1.380649e-23;
constexpr double QE = 1.602176634e-19;
constexpr double TEMP_K = 300.15;
constexpr double VT = KB * TEMP_K / QE;

__device__ __forceinline__ double nan0(double x) { return x != x ? 0.0 : x; }

__device__ __forceinline__ unsigned long long dbits(double x) {
  return (unsigned long long)__double_as_longlong(x);
}

// ---- models -------------------------------------------------------------

template <int MID>
struct ModelTraits;
template <>
struct ModelTraits<VJX_MODEL_RESISTOR> {
  static constexpr int NV = 1, NR = 2, NJ = 4, NP = 1, NL = 0;
};
template <>
struct ModelTraits<VJX_MODEL_CAPACITOR> {
  static constexpr int NV = 1, NR = 2, NJ = 4, NP = 1, NL = 0;
};
template <>
struct ModelTraits<VJX_MODEL_MOSFET> {
  static constexpr int NV = 3, NR = 4, NJ = 16, NP = 11, NL = 0;
};
template <>
struct ModelTraits<VJX_MODEL_DIODE> {
  static constexpr int NV = 1, NR = 2, NJ = 4, NP = 4, NL = 1;
};
template <>
struct ModelTraits<VJX_MODEL_EKV26> {
  static constexpr int NV = 3, NR = 4, NJ = 16, NP = 52, NL = 0;
};

// resistor.va:23-27 -- I(A,B) <+ V(A,B)/r
__device__ __forceinline__ void eval_resistor(const double* vb, const double* par, int i, int N,
                                              double* rr, double* rc, double* jr, double* jc) {
  const double r = par[i];
  const double cur = vb[0] / r;
  const double g = 1.0 / r;
  rr[0] = cur;
  rr[1] = -cur;
  rc[0] = 0.0;
  rc[1] = 0.0;
  jr[0] = g;
  jr[1] = -g;
  jr[2] = -g;
  jr[3] = g;
  jc[0] = jc[1] = jc[2] = jc[3] = 0.0;
}

// capacitor.va:20-23 -- I(A,B) <+ ddt(c*V(A,B))
__device__ __forceinline__ void eval_capacitor(const double* vb, const double* par, int i, int N,
                                               double* rr, double* rc, double* jr, double* jc) {
  const double c = par[i];
  const double q = c * vb[0];
  rr[0] = 0.0;
  rr[1] = 0.0;
  rc[0] = q;
  rc[1] = -q;
  jr[0] = jr[1] = jr[2] = jr[3] = 0.0;
  jc[0] = c;
  jc[1] = -c;
  jc[2] = -c;
  jc[3] = c;
}

__device__ __forceinline__ double softplus(double u) {
  return fmax(u, 0.0) + log1p(exp(-fabs(u)));
}
__device__ __forceinline__ double sigmoid(double u) {
  const double e = exp(-fabs(u));
  return u >= 0.0 ? 1.0 / (1.0 + e) : e / (1.0 + e);
}

// EKV 2.6 long-channel drain current (ekv_longchannel.va:58-96) in the NMOS
// frame with polarity sgn, analytic derivatives, linear terminal charges.
// Parameter columns: sgn, vgoff, phi, gamma, beta0, theta, cgs, cgd, cgb, cdb, csb.
// Branch voltages: vb[0] = V(g,b), vb[1] = V(s,b), vb[2] = V(d,b).
// Residual order d, g, s, b; Jacobian keys row-major over (d, g, s, b).
__device__ __forceinline__ void eval_mosfet(const double* vb, const double* par, int i, int N,
                                            double* rr, double* rc, double* jr, double* jc) {
  const double sgn = par[0 * (size_t)N + i];
  const double vgoff = par[1 * (size_t)N + i];
  const double phi = par[2 * (size_t)N + i];
  const double gamma = par[3 * (size_t)N + i];
  const double beta0 = par[4 * (size_t)N + i];
  const double theta = par[5 * (size_t)N + i];
  const double vgb = vb[0], vsb = vb[1], vdb = vb[2];
  const double VG = sgn * vgb;
  const double VS = sgn * vsb;
  const double VD = sgn * vdb;
  const double vgp = VG + vgoff;
  const double g2 = 0.5 * gamma;
  const double t = vgp + g2 * g2;
  double vp, dvp;
  if (t > 1e-12) {
    const double s = sqrt(t);
    vp = vgp - phi - gamma * (s - g2);
    dvp = 1.0 - gamma / (2.0 * s);
  } else {
    vp = vgp - phi - gamma * gamma / 2.0;
    dvp = 1.0;
  }
  const double arg = phi + vp + 4.0 * VT;
  double n, dn;
  if (arg > 1e-24) {
    const double sa = sqrt(arg);
    n = 1.0 + gamma / (2.0 * sa);
    dn = -gamma / (4.0 * arg * sa);
  } else {
    n = 1.0 + gamma / 2.0 * 1e48;
    dn = 0.0;
  }
  const double den = 1.0 + theta * vp;
  const double beta = beta0 / den;
  const double dbeta = -beta0 * theta / (den * den);
  const double uf = 0.5 * ((vp - VS) / VT);
  const double ur = 0.5 * ((vp - VD) / VT);
  const double lf = softplus(uf);
  const double lr = softplus(ur);
  const double fpf = lf * sigmoid(uf);
  const double fpr = lr * sigmoid(ur);
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
  rr[0] = iext;
  rr[1] = 0.0;
  rr[2] = -iext;
  rr[3] = 0.0;
  jr[0] = gd;
  jr[1] = gg;
  jr[2] = gs;
  jr[3] = gb;
  jr[4] = jr[5] = jr[6] = jr[7] = 0.0;
  jr[8] = -gd;
  jr[9] = -gg;
  jr[10] = -gs;
  jr[11] = -gb;
  jr[12] = jr[13] = jr[14] = jr[15] = 0.0;
  const double cgs = par[6 * (size_t)N + i];
  const double cgd = par[7 * (size_t)N + i];
  const double cgb = par[8 * (size_t)N + i];
  const double cdb = par[9 * (size_t)N + i];
  const double csb = par[10 * (size_t)N + i];
  const double vgs = vgb - vsb;
  const double vgd = vgb - vdb;
  rc[1] = cgs * vgs + cgd * vgd + cgb * vgb;
  rc[0] = -cgd * vgd + cdb * vdb;
  rc[2] = -cgs * vgs + csb * vsb;
  rc[3] = -(cgb * vgb + cdb * vdb + csb * vsb);
  jc[0] = cgd + cdb;
  jc[1] = -cgd;
  jc[2] = 0.0;
  jc[3] = -cdb;
  jc[4] = -cgd;
  jc[5] = cgs + cgd + cgb;
  jc[6] = -cgs;
  jc[7] = -cgb;
  jc[8] = 0.0;
  jc[9] = -cgs;
  jc[10] = cgs + csb;
  jc[11] = -csb;
  jc[12] = -cdb;
  jc[13] = -cgb;
  jc[14] = -csb;
  jc[15] = cdb + csb + cgb;
}

// SPICE junction-voltage limiting, vajax/analysis/limiting.py:22-82 (VACASK
// DEVpnjlim): each branch as written there, logs of max(., 1e-30).
__device__ __forceinline__ double pnjlim(double vnew, double vold, double vt, double vcrit) {
  const double delta = vnew - vold;
  const double arg = delta / vt;
  if (vnew > vcrit && fabs(delta) > 2 * vt) {
    if (vold > 0) {
      return arg >= 0 ? vold + vt * log(fmax(1.0 + arg, 1e-30))
                      : vold - vt * log(fmax(1.0 - arg, 1e-30));
    }
    return vnew > 0 ? vt * log(fmax(vnew / vt, 1e-30)) : vcrit;
  }
  if (vnew < 0) return fmax(vnew, vold > 0 ? -vold - 1.0 : 2.0 * vold - 1.0);
  return vnew;
}

// FET gate-voltage limiting, limiting.py:85-107.
__device__ __forceinline__ double fetlim(double vnew, double vold, double vto) {
  const double step = vold >= vto ? 2 * fabs(vold - vto) + 2 : 0.5;
  return vold + fmin(fmax(vnew - vold, -step), step);
}

// Junction diode with $limit: vdl = pnjlim(V(A,K), limit_state_in, nvt,
// vcrit); I(A,K) <+ is*(exp(vdl/nvt) - 1) + ddt(cj*vdl); limit_state_out =
// vdl; lim_rhs[r] = J[r, A] * (vdl - V(A,K)) for the resistive and reactive
// Jacobians (openvaf_jax/codegen/function_builder.py:1126-1212).
// Parameter columns: is, nvt, vcrit, cj.
__device__ __forceinline__ void eval_diode(const double* vb, const double* par, int i, int N,
                                           const double* li, double* rr, double* rc, double* jr,
                                           double* jc, double* lr, double* lc, double* lo) {
  const double is = par[0 * (size_t)N + i];
  const double nvt = par[1 * (size_t)N + i];
  const double vcrit = par[2 * (size_t)N + i];
  const double cj = par[3 * (size_t)N + i];
  const double vd = vb[0];
  const double vdl = pnjlim(vd, li[0], nvt, vcrit);
  const double e = exp(vdl / nvt);
  const double cur = is * (e - 1.0);
  const double g = is * e / nvt;
  const double q = cj * vdl;
  const double dl = vdl - vd;
  rr[0] = cur;
  rr[1] = -cur;
  rc[0] = q;
  rc[1] = -q;
  jr[0] = g;
  jr[1] = -g;
  jr[2] = -g;
  jr[3] = g;
  jc[0] = cj;
  jc[1] = -cj;
  jc[2] = -cj;
  jc[3] = cj;
  lr[0] = g * dl;
  lr[1] = -(g * dl);
  lc[0] = cj * dl;
  lc[1] = -(cj * dl);
  lo[0] = vdl;
}

// One instance of model MID; li / lr / lc / lo (limit state in, lim_rhs
// resistive / reactive, limit state out) are used by limited models only.
// ---- EKV 2.6 (vajax/devices/models/ekv/ekv.va) --------------------------------
//
// The EKV 2.6 charge-based MOSFET of ekv.va: short- and narrow-channel effects
// on the pinch-off voltage, velocity saturation and channel-length modulation,
// mobility reduction by the vertical field (E0 form, or THETA when E0 = 0),
// impact-ionisation substrate current, the node charges QD, QS, QG, and the
// drain/source junction diodes with breakdown, trap-assisted tunnelling and
// depletion charge.  Derivatives by forward-mode dual numbers over the three
// branch voltages V(g,b), V(s,b), V(d,b), in place of the file's hand-written
// derivatives.  Left out: noise, and the series-resistance correction of the
// drain current (ekv.va:697-700), which is the identity at the default RSH = 0
// (the host rejects RSH > 0).  Temperature and geometry terms are the instance
// cache below; the thermal voltage is k T / q at the device temperature.
struct D3 {
  double v, d0, d1, d2;
};
__device__ __forceinline__ D3 dc(double v) { return D3{v, 0.0, 0.0, 0.0}; }
__device__ __forceinline__ D3 operator+(D3 a, D3 b) {
  return D3{a.v + b.v, a.d0 + b.d0, a.d1 + b.d1, a.d2 + b.d2};
}
__device__ __forceinline__ D3 operator-(D3 a, D3 b) {
  return D3{a.v - b.v, a.d0 - b.d0, a.d1 - b.d1, a.d2 - b.d2};
}
__device__ __forceinline__ D3 operator-(D3 a) { return D3{-a.v, -a.d0, -a.d1, -a.d2}; }
__device__ __forceinline__ D3 operator*(D3 a, D3 b) {
  return D3{a.v * b.v, a.d0 * b.v + a.v * b.d0, a.d1 * b.v + a.v * b.d1, a.d2 * b.v + a.v * b.d2};
}
__device__ __forceinline__ D3 operator*(double k, D3 a) {
  return D3{k * a.v, k * a.d0, k * a.d1, k * a.d2};
}
__device__ __forceinline__ D3 operator*(D3 a, double k) { return k * a; }
__device__ __forceinline__ D3 operator+(D3 a, double k) { return D3{a.v + k, a.d0, a.d1, a.d2}; }
__device__ __forceinline__ D3 operator+(double k, D3 a) { return a + k; }
__device__ __forceinline__ D3 operator-(D3 a, double k) { return D3{a.v - k, a.d0, a.d1, a.d2}; }
__device__ __forceinline__ D3 operator-(double k, D3 a) { return D3{k - a.v, -a.d0, -a.d1, -a.d2}; }
__device__ __forceinline__ D3 operator/(D3 a, D3 b) {
  const double q = a.v / b.v;
  return D3{q, (a.d0 - q * b.d0) / b.v, (a.d1 - q * b.d1) / b.v, (a.d2 - q * b.d2) / b.v};
}
__device__ __forceinline__ D3 operator/(D3 a, double k) { return D3{a.v / k, a.d0 / k, a.d1 / k, a.d2 / k}; }
__device__ __forceinline__ D3 operator/(double k, D3 b) { return dc(k) / b; }
__device__ __forceinline__ D3 dsqrt(D3 a) {
  const double r = sqrt(a.v), g = 0.5 / r;
  return D3{r, g * a.d0, g * a.d1, g * a.d2};
}
__device__ __forceinline__ D3 dlog(D3 a) {
  const double g = 1.0 / a.v;
  return D3{log(a.v), g * a.d0, g * a.d1, g * a.d2};
}
__device__ __forceinline__ D3 dexp(D3 a) {
  const double e = exp(a.v);
  return D3{e, e * a.d0, e * a.d1, e * a.d2};
}

// Normalized inversion charge q from x = (VP - V)/Vt, the file's approximate
// inverse of x = 2 q + ln q (ekv.va:428-444): returns yk = q.
__device__ __forceinline__ D3 ekv_q(D3 x) {
  if (x.v > -0.35) {
    const D3 z0 = 2.0 / (1.3 + x - dlog(x + 1.6));
    const D3 zk = (2.0 + z0) / (1.0 + x + dlog(z0));
    return (1.0 + x + dlog(zk)) / (2.0 + zk);
  }
  if (x.v > -15.0) {
    const D3 z0 = 1.55 + dexp(-x);
    const D3 zk = (2.0 + z0) / (1.0 + x + dlog(z0));
    return (1.0 + x + dlog(zk)) / (2.0 + zk);
  }
  if (x.v > -23.0) return 1.0 / (2.0 + dexp(-x));
  return dexp(x) + 1e-64;
}

// Instance cache columns, in order (the voltage-independent terms of
// ekv.va:298-388 and :767-788, computed once on the host by models.py).
enum Ekv26Col {
  EK_TYPE, EK_VTO_S, EK_PHI_T, EK_GAMMA_S, EK_KP_WEFF, EK_LEFF, EK_UCRIT_T, EK_LAMBDA, EK_LC,
  EK_WETA_W, EK_LETA_L, EK_DVFB, EK_E0, EK_THETA, EK_T0, EK_ETA_QI, EK_IBN, EK_IBA_IBB,
  EK_LC_IBB, EK_WLCOX, EK_VC, EK_LOG_VC_VT, EK_VT,
  // junction diodes (ekv.va:763-848)
  EK_M, EK_JK, EK_IS_D, EK_IS_S, EK_BV, EK_XJBV, EK_GMIN, EK_TUN_W, EK_TUN_PD, EK_TUN_AD,
  EK_TUN_PS, EK_TUN_AS, EK_KG, EK_KSW, EK_KA, EK_VTSSWG, EK_VTSSW, EK_VTS, EK_CJA_D, EK_CJP_D,
  EK_CJG, EK_CJA_S, EK_CJP_S, EK_MJ, EK_MJSW, EK_MJSWG, EK_PB, EK_PBSW, EK_PBSWG, EK_NCOL
};

// Column k of instance i's cache, read where it is used.
struct Ekv26Cols {
  const double* p;
  int N, i;
  __device__ __forceinline__ double operator[](int k) const { return p[(size_t)k * N + i]; }
};

// Trap-assisted tunnelling term  exp(v k vt / max(vt + v, 1e-3)) - 1  (ekv.va:800-802).
__device__ __forceinline__ D3 ekv_tunnel(D3 v, double k, double vt) {
  const D3 den = vt + v.v > 1.0e-3 ? vt + v : dc(1.0e-3);
  return dexp(v * k * vt / den) - 1.0;
}

// One junction diode between the drain (or source) and the bulk, v = TYPE V(x,b):
// current (ekv.va:792-803) and charge (ekv.va:818-831), both without TYPE M.
__device__ __forceinline__ void ekv_junction(D3 v, double is, double tun_p, double tun_a,
                                             double cja, double cjp, const Ekv26Cols& c, D3& cur,
                                             D3& q) {
  D3 arg = -v * c[EK_JK];
  if (arg.v < -40.0) arg = dc(-40.0);
  const D3 t0 = (c[EK_BV] - v) * c[EK_JK];
  const D3 fbd = t0.v > 70.0 ? dc(1.0) : 1.0 + c[EK_XJBV] * dexp(-t0);
  D3 tun = -c[EK_TUN_W] * ekv_tunnel(v, c[EK_KG], c[EK_VTSSWG]);
  tun = tun - tun_p * ekv_tunnel(v, c[EK_KSW], c[EK_VTSSW]);
  tun = tun - tun_a * ekv_tunnel(v, c[EK_KA], c[EK_VTS]);
  cur = is * (1.0 - dexp(arg)) * fbd + v * c[EK_GMIN] + tun;
  D3 cap;
  if (v.v > 0.0) {
    cap = cja * dexp(-c[EK_MJ] * dlog(1.0 + v / c[EK_PB])) +
          cjp * dexp(-c[EK_MJSW] * dlog(1.0 + v / c[EK_PBSW])) +
          c[EK_CJG] * dexp(-c[EK_MJSWG] * dlog(1.0 + v / c[EK_PBSWG]));
  } else {
    cap = cja * (1.0 - c[EK_MJ] * v / c[EK_PB]) + cjp * (1.0 - c[EK_MJSW] * v / c[EK_PBSW]) +
          c[EK_CJG] * (1.0 - c[EK_MJSWG] * v / c[EK_PBSWG]);
  }
  q = cap * v;
}

// Branch voltages vb[0] = V(g,b), vb[1] = V(s,b), vb[2] = V(d,b); residual
// order d, g, s, b; Jacobian keys row-major over (d, g, s, b).
__device__ __forceinline__ void eval_ekv26(const double* vb, const double* par, int i, int N,
                                           double* rr, double* rc, double* jr, double* jc) {
  const Ekv26Cols c{par, N, i};
  const double TYPE = c[EK_TYPE], VTO_S = c[EK_VTO_S], PHI_T = c[EK_PHI_T];
  const double GAMMA_S = c[EK_GAMMA_S], KP_Weff = c[EK_KP_WEFF], Leff = c[EK_LEFF];
  const double UCRIT_T = c[EK_UCRIT_T], LAMBDA = c[EK_LAMBDA], Lc = c[EK_LC];
  const double WETA_W = c[EK_WETA_W], LETA_L = c[EK_LETA_L], deltaVFB = c[EK_DVFB];
  const double E0 = c[EK_E0], THETA = c[EK_THETA], T0 = c[EK_T0], eta_qi = c[EK_ETA_QI];
  const double IBN = c[EK_IBN], IBA_IBB = c[EK_IBA_IBB], Lc_IBB = c[EK_LC_IBB];
  const double WLCox = c[EK_WLCOX], Vc = c[EK_VC], log_Vc_Vt = c[EK_LOG_VC_VT], Vt = c[EK_VT];
  const double inv_Vt = 1.0 / Vt, Vt_01 = 0.1 * Vt, Vt_4 = 4.0 * Vt;
  const double Vt_Vt_2 = 2.0 * Vt * Vt, Vt_Vt_16 = 16.0 * Vt * Vt;
  const double GAMMA_sqrt_PHI = GAMMA_S * sqrt(PHI_T);
  const double Lc_LAMBDA = Lc * LAMBDA, Lc_UCRIT = Lc * UCRIT_T, inv_UCRIT = 1.0 / UCRIT_T;
  const D3 vg = D3{TYPE * vb[0], TYPE, 0.0, 0.0};
  const D3 vs = D3{TYPE * vb[1], 0.0, TYPE, 0.0};
  const D3 vd = D3{TYPE * vb[2], 0.0, 0.0, TYPE};
  // source/drain interchange (ekv.va:390-402)
  const bool fwd = !(vd.v - vs.v < 0);
  const D3 VG = vg, VS = fwd ? vs : vd, VD = fwd ? vd : vs;
  const double Mode = fwd ? 1.0 : -1.0;
  // effective gate voltage and pinch-off voltage (ekv.va:403-425)
  const D3 VGstar = VG - VTO_S - deltaVFB + PHI_T + GAMMA_sqrt_PHI;
  const D3 VGprime = 0.5 * (VGstar + dsqrt(VGstar * VGstar + 2.0 * Vt_Vt_16));
  const D3 PHI_VS = PHI_T + VS;
  const D3 sqrt_PHI_VS = dsqrt(0.5 * (PHI_VS + dsqrt(PHI_VS * PHI_VS + Vt_Vt_16)));
  const D3 PHI_VD = PHI_T + VD;
  const D3 sqrt_PHI_VD = dsqrt(0.5 * (PHI_VD + dsqrt(PHI_VD * PHI_VD + Vt_Vt_16)));
  const D3 big_sqrt_VP0 = dsqrt(VGprime + 0.25 * GAMMA_S * GAMMA_S);
  const D3 VP0 = VGprime - PHI_T - GAMMA_S * (big_sqrt_VP0 - 0.5 * GAMMA_S);
  const D3 sqrt_PHI_VP0 = dsqrt(VP0 + PHI_T + Vt_01);
  const D3 GAMMAstar = GAMMA_S - LETA_L * (sqrt_PHI_VS + sqrt_PHI_VD) + WETA_W * sqrt_PHI_VP0;
  const D3 GAMMAprime = 0.5 * (GAMMAstar + dsqrt(GAMMAstar * GAMMAstar + Vt_01));
  const D3 big_sqrt_VP = dsqrt(VGprime + 0.25 * GAMMAprime * GAMMAprime);
  const D3 VP = VGprime - PHI_T - GAMMAprime * (big_sqrt_VP - 0.5 * GAMMAprime);
  // forward normalized current (ekv.va:427-447)
  const D3 qf = ekv_q((VP - VS) * inv_Vt);
  const D3 if_ = qf * (1.0 + qf);
  const D3 sqrt_if = dsqrt(if_);
  // saturation voltage (ekv.va:449-459)
  const double Vt_Vc = Vt / Vc;
  const D3 VDSS = Vc * (dsqrt(0.25 + sqrt_if * Vt_Vc) - 0.5);
  const D3 Vds = 0.5 * (VD - VS);
  const D3 deltaV_2 = Vt_Vt_16 * (LAMBDA * (sqrt_if - VDSS * inv_Vt) + 15.625e-3);
  const D3 Vip = dsqrt(VDSS * VDSS + deltaV_2) - dsqrt((Vds - VDSS) * (Vds - VDSS) + deltaV_2);
  const D3 VDSSprime =
      Vc * (dsqrt(0.25 + (sqrt_if - 0.75 * dlog(if_)) * Vt_Vc) - 0.5) + log_Vc_Vt;
  // reverse normalized current with the saturation voltage (ekv.va:461-486)
  const D3 Vdsprime = Vds - VDSSprime;
  const D3 qrp = ekv_q((VP - Vds - VS - dsqrt(VDSSprime * VDSSprime + deltaV_2) +
                        dsqrt(Vdsprime * Vdsprime + deltaV_2)) *
                       inv_Vt);
  const D3 irprime = qrp * (1.0 + qrp);
  // channel length modulation and velocity saturation (ekv.va:489-493)
  const D3 deltaL = Lc_LAMBDA * dlog(1.0 + (Vds - Vip) / Lc_UCRIT);
  const D3 Lprime = Leff - deltaL + (Vds + Vip) * inv_UCRIT;
  const double Lmin = 0.1 * Leff;
  const D3 Leq = 0.5 * (Lprime + dsqrt(Lprime * Lprime + Lmin * Lmin));
  // reverse normalized current (ekv.va:498-518)
  const D3 qr = ekv_q((VP - VD) * inv_Vt);
  const D3 ir = qr * (1.0 + qr);
  // normalized charges and the transconductance factor (ekv.va:519-557)
  const D3 sif2 = 0.25 + if_, sir2 = 0.25 + ir;
  const D3 sif = dsqrt(sif2), sir = dsqrt(sir2);
  const D3 sif_sir_2 = (sif + sir) * (sif + sir);
  const D3 sqrt_PHI_VP_2 = 2.0 * dsqrt(VP + PHI_T + 1.0e-6);
  const D3 n_1 = GAMMA_S / sqrt_PHI_VP_2;
  const D3 n_1_n = GAMMA_S / (sqrt_PHI_VP_2 + GAMMA_S);
  const D3 qi = -((1.0 + n_1) * Vt *
                  ((0.66666666 + 0.66666666) * (sir2 + sir * sif + sif2) / (sif + sir) - 1.0));
  const D3 qb = -0.5 * GAMMA_S * sqrt_PHI_VP_2 - n_1_n * qi;
  D3 beta;
  if (E0 == 0.0) {
    const D3 VPprime = 0.5 * (VP + dsqrt(VP * VP + Vt_Vt_2));
    beta = KP_Weff / (Leq * (1.0 + THETA * VPprime));
  } else {
    const D3 e = qb + eta_qi * qi;
    const D3 E0_Q_1 = e.v > 0.0 ? 1.0 + T0 * e : 1.0 - T0 * e;
    beta = KP_Weff * (1.0 + T0 * GAMMA_sqrt_PHI) / (Leq * E0_Q_1);
  }
  // slope factor and drain current (ekv.va:560-565)
  const D3 n = 1.0 + GAMMA_S / (2.0 * dsqrt(PHI_T + VP + Vt_4));
  const D3 Id = Vt_Vt_2 * n * beta * (if_ - irprime);
  // impact ionisation (ekv.va:709-725)
  D3 Isub = dc(0.0);
  const D3 Vib = VD - VS - (IBN + IBN) * VDSS;
  if (Vib.v > 0.0 && IBA_IBB > 0.0) {
    D3 x = -Lc_IBB / Vib;
    if (x.v < -35.0) x = dc(-35.0);
    Isub = IBA_IBB * Vib * dexp(x) * Id;
  }
  // node charges (ekv.va:727-740)
  const D3 sif3 = sif * sif2, sir3 = sir * sir2;
  const D3 sqrt_PHI_VP2_2 = 2.0 * dsqrt(PHI_T + 0.5 * VP);
  const D3 n_Vt_COX = (1.0 + GAMMAprime / sqrt_PHI_VP2_2) * (Vt * WLCox);
  const D3 QD = -(n_Vt_COX *
                  (0.266666666 * (3.0 * sir3 + 6.0 * sir2 * sif + 4.0 * sir * sif2 + 2.0 * sif3) /
                       sif_sir_2 -
                   0.5));
  const D3 QS = -(n_Vt_COX *
                  (0.266666666 * (3.0 * sif3 + 6.0 * sif2 * sir + 4.0 * sif * sir2 + 2.0 * sir3) /
                       sif_sir_2 -
                   0.5));
  const D3 QI = QS + QD;
  const D3 QB = WLCox * (-0.5 * GAMMAprime * sqrt_PHI_VP_2 + VGprime - VGstar) -
                QI * GAMMAprime / (GAMMAprime + sqrt_PHI_VP2_2);
  const D3 QG = -QI - QB;
  // junction diodes on the unswapped branches (ekv.va:790-847)
  D3 ijd, qjd, ijs, qjs;
  ekv_junction(vd, c[EK_IS_D], c[EK_TUN_PD], c[EK_TUN_AD], c[EK_CJA_D], c[EK_CJP_D], c, ijd, qjd);
  ekv_junction(vs, c[EK_IS_S], c[EK_TUN_PS], c[EK_TUN_AS], c[EK_CJA_S], c[EK_CJP_S], c, ijs, qjs);
  const double TM = TYPE * c[EK_M];
  // terminal currents and charges (ekv.va:741-754, 803, 815, 832, 847)
  const D3 ids = (TYPE * Mode) * Id;
  const D3 isub = TYPE * Isub;
  const D3 rd = (fwd ? ids + isub : ids) + TM * ijd;
  const D3 rs = (fwd ? -ids : isub - ids) + TM * ijs;
  const D3 rb = -(isub + TM * (ijd + ijs));
  const D3 qd = TYPE * (fwd ? QD : QS) + TM * qjd;
  const D3 qs = TYPE * (fwd ? QS : QD) + TM * qjs;
  const D3 qg = TYPE * QG;
  const D3 qbn = -(qd + qs + qg);
  // Jacobian columns (d, g, s, b) from the branch derivatives (g, s, d)
  auto put = [&](double* r, double* j, int row, const D3& x) {
    r[row] = x.v;
    j[4 * row + 0] = x.d2;
    j[4 * row + 1] = x.d0;
    j[4 * row + 2] = x.d1;
    j[4 * row + 3] = -(x.d0 + x.d1 + x.d2);
  };
  put(rr, jr, 0, rd);
  put(rr, jr, 1, dc(0.0));
  put(rr, jr, 2, rs);
  put(rr, jr, 3, rb);
  put(rc, jc, 0, qd);
  put(rc, jc, 1, qg);
  put(rc, jc, 2, qs);
  put(rc, jc, 3, qbn);
}

template <int MID>
__device__ __forceinline__ void eval_model(const double* vb, const double* par, int i, int N,
                                           const double* li, double* rr, double* rc, double* jr,
                                           double* jc, double* lr, double* lc, double* lo) {
  if constexpr (MID == VJX_MODEL_RESISTOR) eval_resistor(vb, par, i, N, rr, rc, jr, jc);
  if constexpr (MID == VJX_MODEL_CAPACITOR) eval_capacitor(vb, par, i, N, rr, rc, jr, jc);
  if constexpr (MID == VJX_MODEL_MOSFET) eval_mosfet(vb, par, i, N, rr, rc, jr, jc);
  if constexpr (MID == VJX_MODEL_EKV26) eval_ekv26(vb, par, i, N, rr, rc, jr, jc);
  if constexpr (MID == VJX_MODEL_DIODE) eval_diode(vb, par, i, N, li, rr, rc, jr, jc, lr, lc, lo);
}

// Eval + stamp: lane = device instance.
template <int MID>
__global__ void __launch_bounds__(256) k_eval(int N, const int32_t* __restrict__ vn1,
                                              const int32_t* __restrict__ vn2,
                                              const int32_t* __restrict__ jdst,
                                              const int32_t* __restrict__ rdst,
                                              const double* __restrict__ par,
                                              const double* __restrict__ X,
                                              const DevScalars* __restrict__ sc, double* jac,
                                              double* rf, double* rq, double* lrf, double* lrq,
                                              const double* __restrict__ ls_in, double* ls_out) {
  using Tr = ModelTraits<MID>;
  constexpr int NL = Tr::NL > 0 ? Tr::NL : 1;
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= N) return;
  double vb[Tr::NV];
#pragma unroll
  for (int k = 0; k < Tr::NV; ++k)
    vb[k] = X[vn1[(size_t)k * N + i]] - X[vn2[(size_t)k * N + i]];
  double rr[Tr::NR], rc[Tr::NR], jr[Tr::NJ], jc[Tr::NJ];
  double lr[Tr::NR], lc[Tr::NR], li[NL], lo[NL];
  if constexpr (Tr::NL > 0) {
#pragma unroll
    for (int s = 0; s < Tr::NL; ++s) li[s] = ls_in ? ls_in[(size_t)i * Tr::NL + s] : 0.0;
  }
  eval_model<MID>(vb, par, i, N, li, rr, rc, jr, jc, lr, lc, lo);
  if constexpr (Tr::NL > 0) {
#pragma unroll
    for (int s = 0; s < Tr::NL; ++s) ls_out[(size_t)i * Tr::NL + s] = lo[s];
#pragma unroll
    for (int r = 0; r < Tr::NR; ++r) {
      const int d = rdst[(size_t)r * N + i];
      if (d >= 0) {
        lrf[d] = nan0(lr[r]);
        lrq[d] = nan0(lc[r]);
      }
    }
  }
  if (jac) {
    const double c0 = sc->c0;
#pragma unroll
    for (int e = 0; e < Tr::NJ; ++e) {
      const int d = jdst[(size_t)e * N + i];
      if (d >= 0) jac[d] = nan0(jr[e] + c0 * jc[e]);
    }
  }
#pragma unroll
  for (int r = 0; r < Tr::NR; ++r) {
    const int d = rdst[(size_t)r * N + i];
    if (d >= 0) {
      rf[d] = nan0(rr[r]);
      rq[d] = nan0(rc[r]);
    }
  }
}

// Inputs of the split eval contract gathered into the raw eval's layout:
// branch voltages [n_v][N] from device_params columns, the eval input block
// [n_par][N] from shared_cache or device_cache columns.
__global__ void k_split_gather(int N, int n_v, int n_par, int n_dp, int n_dc,
                               const int32_t* __restrict__ vpos, const int32_t* __restrict__ csrc,
                               const double* __restrict__ sc, const double* __restrict__ dp,
                               const double* __restrict__ dc, double* vb, double* par) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= N) return;
  for (int k = 0; k < n_v; ++k) vb[(size_t)k * N + i] = dp[(size_t)i * n_dp + vpos[k]];
  for (int j = 0; j < n_par; ++j) {
    const int c = csrc[j];
    par[(size_t)j * N + i] = c >= 0 ? sc[c] : dc[(size_t)i * n_dc + ~c];
  }
}

// $mfactor on every residual, Jacobian and lim_rhs output; lim_rhs zero for
// models without $limit (the raw eval does not write them); small-signal
// residuals zero.
__global__ void k_split_finish(int N, int n_res, int n_jac, int limited, const double* mf,
                               double* rr, double* rc, double* jr, double* jc, double* lr,
                               double* lc, double* sr, double* sq) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= N) return;
  const double m = mf ? *mf : 1.0;
  for (int r = 0; r < n_res; ++r) {
    const size_t e = (size_t)i * n_res + r;
    rr[e] *= m;
    rc[e] *= m;
    lr[e] = limited ? lr[e] * m : 0.0;
    lc[e] = limited ? lc[e] * m : 0.0;
    sr[e] = 0.0;
    sq[e] = 0.0;
  }
  for (int j = 0; j < n_jac; ++j) {
    const size_t e = (size_t)i * n_jac + j;
    jr[e] *= m;
    jc[e] *= m;
  }
}

// Raw eval (openvaf_jax contract), row-major outputs per instance.
template <int MID>
__global__ void __launch_bounds__(256)
    k_model_eval(int N, const double* __restrict__ vbr, const double* __restrict__ par,
                 const double* __restrict__ lio, double* rro, double* rco, double* jro,
                 double* jco, double* lro, double* lco, double* loo) {
  using Tr = ModelTraits<MID>;
  constexpr int NL = Tr::NL > 0 ? Tr::NL : 1;
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= N) return;
  double vb[Tr::NV];
#pragma unroll
  for (int k = 0; k < Tr::NV; ++k) vb[k] = vbr[(size_t)k * N + i];
  double rr[Tr::NR], rc[Tr::NR], jr[Tr::NJ], jc[Tr::NJ];
  double lr[Tr::NR], lc[Tr::NR], li[NL], lo[NL];
  if constexpr (Tr::NL > 0) {
#pragma unroll
    for (int s = 0; s < Tr::NL; ++s) li[s] = lio ? lio[(size_t)i * Tr::NL + s] : 0.0;
  }
  eval_model<MID>(vb, par, i, N, li, rr, rc, jr, jc, lr, lc, lo);
  if constexpr (Tr::NL > 0) {
#pragma unroll
    for (int r = 0; r < Tr::NR; ++r) {
      if (lro) lro[(size_t)i * Tr::NR + r] = lr[r];
      if (lco) lco[(size_t)i * Tr::NR + r] = lc[r];
    }
#pragma unroll
    for (int s = 0; s < Tr::NL; ++s)
      if (loo) loo[(size_t)i * Tr::NL + s] = lo[s];
  }
#pragma unroll
  for (int r = 0; r < Tr::NR; ++r) {
    rro[(size_t)i * Tr::NR + r] = rr[r];
    rco[(size_t)i * Tr::NR + r] = rc[r];
  }
#pragma unroll
  for (int e = 0; e < Tr::NJ; ++e) {
    jro[(size_t)i * Tr::NJ + e] = jr[e];
    jco[(size_t)i * Tr::NJ + e] = jc[e];
  }
}

__global__ void k_isrc(int k, const double* __restrict__ isrc, const int32_t* __restrict__ dst,
                       double* rf, double* rq) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= k) return;
  const double v = isrc[i];
  const int dp = dst[2 * i], dn = dst[2 * i + 1];
  if (dp >= 0) {
    rf[dp] = nan0(v * 1.0);
    rq[dp] = 0.0;
  }
  if (dn >= 0) {
    rf[dn] = nan0(v * -1.0);
    rq[dn] = 0.0;
  }
}

// ---- fixed-order workgroup reductions ---------------------------------------
//
// Long segments (a supply rail's CSR slot or KCL row, a dense LU row) are
// summed by one workgroup: kReduceLanes strided partial sums, then a halving
// tree in LDS.  The order is fixed, so results are deterministic; the CPU
// reference program emulates the same order.

constexpr int RB = kReduceLanes;

// All threads of the block call; threads >= RB contribute nothing.  The tree
// is the halving tree r[t] += r[t + w] for w = 128 .. 1: its first two steps
// run in LDS in one pass ((r[t] + r[t+128]) + (r[t+64] + r[t+192]) for
// t < 64), the last six as wavefront shuffles, so the block waits on two
// barriers instead of nine; the additions and their order are unchanged.
// Workgroup barrier for data exchanged through LDS only: waits for this
// wave's LDS operations (lgkmcnt(0); s_waitcnt immediate 0xC07F leaves
// vmcnt and expcnt at their maxima in the gfx9 encoding), not for its
// outstanding global loads, so prefetches stay in flight across it.  The
// empty asm statements keep the compiler from moving memory accesses
// across the barrier.  Not for data exchanged through global memory.
__device__ __forceinline__ void lds_barrier() {
  asm volatile("" ::: "memory");
  __builtin_amdgcn_s_waitcnt(0xC07F);
  __builtin_amdgcn_s_barrier();
  asm volatile("" ::: "memory");
}

// The tree part of block_sum: lane t (< RB) holds its strided partial.
template <typename T>
__device__ __forceinline__ T block_tree(T acc, T* red) {
  static_assert(RB == 256, "tree layout assumes 256 lanes");
  const int t = threadIdx.x;
  if (t < RB) red[t] = acc;
  lds_barrier();
  if (t < 64) {
    T a = (red[t] + red[t + 128]) + (red[t + 64] + red[t + 192]);
#pragma unroll
    for (int w = 32; w > 0; w >>= 1) a += __shfl_down(a, w, 64);
    if (t == 0) red[RB] = a;
  }
  lds_barrier();
  const T total = red[RB];
  lds_barrier();
  return total;
}

template <typename T, typename F>
__device__ __forceinline__ T block_sum(int start, int end, F term, T* red) {
  T acc = (T)0;
  if ((int)threadIdx.x < RB)
    for (int u = start + (int)threadIdx.x; u < end; u += RB) acc += term(u);
  return block_tree<T>(acc, red);
}

// G consecutive lanes (G divides the wavefront): strided partials, then the
// halving tree through shuffles; lane 0 of the group returns the sum.  All G
// lanes must call.
template <int G, typename T, typename F>
__device__ __forceinline__ T group_sum(int start, int end, int lane, F term) {
  T acc = (T)0;
  for (int u = start + lane; u < end; u += G) acc += term(u);
#pragma unroll
  for (int w = G / 2; w > 0; w >>= 1) acc += __shfl_down(acc, w, G);
  return acc;
}

template <typename F>
__device__ __forceinline__ double block_max(int start, int end, F term, double* red) {
  const int t = threadIdx.x;
  if (t < RB) {
    double m = 0.0;
    for (int u = start + t; u < end; u += RB) m = fmax(m, term(u));
    red[t] = m;
  }
  __syncthreads();
  for (int w = RB / 2; w > 0; w >>= 1) {
    if (t < w) red[t] = fmax(red[t], red[t + w]);
    __syncthreads();
  }
  const double total = red[0];
  __syncthreads();
  return total;
}

// ---- stamping reductions ----------------------------------------------------

__device__ __forceinline__ double finish_slot(double acc, int s, const int32_t* slot_diag,
                                              const double* slot_b, const DevScalars* sc) {
  if (slot_diag[s]) acc += sc->gmin + sc->gshunt;
  const double b = slot_b[s];
  if (b != 0.0) acc += b;
  return acc;
}

// Chunk pre-pass for hub segments: chunk c sums [start[c], end[c]) with one
// workgroup; the reducing kernels then add a segment's chunk partials in order.
__global__ void __launch_bounds__(256) k_chunk_sum(const int32_t* __restrict__ cs,
                                                   const int32_t* __restrict__ ce,
                                                   const double* __restrict__ v, double* part) {
  __shared__ double red[RB + 1];
  const int c = blockIdx.x;
  const double s = block_sum<double>(cs[c], ce[c], [&](int u) { return v[u]; }, red);
  if (threadIdx.x == 0) part[c] = s;
}

// residual rows: sums of rf and rq and max |rf| per chunk
__global__ void __launch_bounds__(256) k_chunk_res(const int32_t* __restrict__ cs,
                                                   const int32_t* __restrict__ ce,
                                                   const double* __restrict__ rf,
                                                   const double* __restrict__ rq,
                                                   const double* __restrict__ lr,
                                                   const double* __restrict__ lq, double* pf,
                                                   double* pq, double* pm, double* pl,
                                                   double* plq) {
  __shared__ double red[RB + 1];
  const int c = blockIdx.x;
  const double f = block_sum<double>(cs[c], ce[c], [&](int u) { return rf[u]; }, red);
  const double q = block_sum<double>(cs[c], ce[c], [&](int u) { return rq[u]; }, red);
  const double m = block_max(cs[c], ce[c], [&](int u) { return fabs(rf[u]); }, red);
  double l = 0.0, lc = 0.0;
  if (lr) {  // uniform across the block
    l = block_sum<double>(cs[c], ce[c], [&](int u) { return lr[u]; }, red);
    lc = block_sum<double>(cs[c], ce[c], [&](int u) { return lq[u]; }, red);
  }
  if (threadIdx.x == 0) {
    pf[c] = f;
    pq[c] = q;
    pm[c] = m;
    if (lr) {
      pl[c] = l;
      plq[c] = lc;
    }
  }
}

// SpMV rows: sums of a[e] * x[ci[e]] per chunk
__global__ void __launch_bounds__(256) k_chunk_spmv(const int32_t* __restrict__ cs,
                                                    const int32_t* __restrict__ ce,
                                                    const double* __restrict__ a,
                                                    const int32_t* __restrict__ ci,
                                                    const double* __restrict__ x, double* part) {
  __shared__ double red[RB + 1];
  const int c = blockIdx.x;
  const double s = block_sum<double>(cs[c], ce[c], [&](int e) { return a[e] * x[ci[e]]; }, red);
  if (threadIdx.x == 0) part[c] = s;
}

// Blocks [0, n_heavy) finish one hub slot each from its chunk partials; the
// other blocks reduce one slot per lane.
__global__ void __launch_bounds__(256) k_reduce_jac(int nse, const int32_t* __restrict__ seg,
                                                    const double* __restrict__ contrib,
                                                    const int32_t* __restrict__ slot_diag,
                                                    const double* __restrict__ slot_b,
                                                    const DevScalars* __restrict__ sc,
                                                    double* __restrict__ csr, HubChunks hub) {
  if ((int)blockIdx.x < hub.n_items) {
    if (threadIdx.x == 0) {
      const int s = hub.item[blockIdx.x];
      double tot = 0.0;
      for (int c = hub.ptr[blockIdx.x]; c < hub.ptr[blockIdx.x + 1]; ++c) tot += hub.part[c];
      csr[s] = finish_slot(tot, s, slot_diag, slot_b, sc);
    }
    return;
  }
  const int n_heavy = hub.n_items;
  const int s = (blockIdx.x - n_heavy) * blockDim.x + threadIdx.x;
  if (s >= nse) return;
  const int u0 = seg[s], u1 = seg[s + 1];
  if (u1 - u0 > kHeavySegment) return;
  double acc = 0.0;
  for (int u = u0; u < u1; ++u) acc += contrib[u];
  csr[s] = finish_slot(acc, s, slot_diag, slot_b, sc);
}

// Residual of one node row from its reduced sums (mna_builder.py:1012-1076);
// returns the value entering the convergence test and whether it fails.
// f, q: the row's sums of resistive and reactive residual contributions; lf,
// lq: its lim_rhs sums (zero without limited models).
__device__ __forceinline__ bool finish_row(const ResidualArgs& a, int i, double f, double q,
                                           double mx, double lf, double lq, double& fchk) {
  const DevScalars* sc = a.sc;
  double fr = f - lf;
  if (a.m > 0) {
    double br = 0.0;
    for (int e = a.br_ptr[i]; e < a.br_ptr[i + 1]; ++e) {
      const int k = a.br_ent[e];
      if (k >= 0) {
        a.I_vs[k] = -fr;
        br += a.X[a.n_total + k];
      } else {
        br += -a.X[a.n_total + ~k];
      }
    }
    fr = fr + br;
  }
  const double qp = a.Q_prev ? a.Q_prev[i] : 0.0;
  const double dq = a.dQdt_prev ? a.dQdt_prev[i] : 0.0;
  const double q2 = a.Q_prev2 ? a.Q_prev2[i] : 0.0;
  double t = fr - 0.0;
  t = t + sc->c0 * (q - lq);
  t = t + sc->c1 * qp;
  t = t + sc->d1 * dq;
  t = t + sc->c2 * q2;
  const double eff = sc->gmin + sc->gshunt;
  t = t + eff * a.X[i + 1];
  a.f[i] = t;
  if (a.Q) a.Q[i] = q;
  if (a.mrc) a.mrc[i] = mx;
  fchk = 0.0;
  const bool checked = a.conv_mask == nullptr || a.conv_mask[i] != 0;
  if (!checked) return false;
  const double floor = a.res_floor ? a.res_floor[i] : sc->abstol;
  const double tol = fmax(mx * sc->reltol, floor);
  fchk = fabs(t);
  return !(fchk < tol);
}

__global__ void __launch_bounds__(256) k_residual(ResidualArgs a) {
  const HubChunks& hub = a.hub;
  if ((int)blockIdx.x < hub.n_items) {
    if (threadIdx.x == 0) {
      const int i = hub.item[blockIdx.x];
      double f = 0.0, q = 0.0, mx = 0.0, lf = 0.0, lq = 0.0;
      for (int c = hub.ptr[blockIdx.x]; c < hub.ptr[blockIdx.x + 1]; ++c) {
        f += hub.part[c];
        q += hub.part2[c];
        mx = fmax(mx, hub.part3[c]);
        if (a.lr) {
          lf += hub.part4[c];
          lq += hub.part5[c];
        }
      }
      double fchk;
      const bool fail = finish_row(a, i, f, q, mx, lf, lq, fchk);
      if (a.st) {
        if (fchk > 0.0) atomicMax(&a.st->maxf_bits, dbits(fchk));
        if (fail) atomicAdd(&a.st->res_fail, 1u);
      }
    }
    return;
  }
  const int i = (blockIdx.x - hub.n_items) * blockDim.x + threadIdx.x;
  bool fail = false;
  double fchk = 0.0;
  if (i < a.n_unk) {
    const int u0 = a.res_seg[i], u1 = a.res_seg[i + 1];
    if (u1 - u0 <= kHeavySegment) {
      double f = 0.0, q = 0.0, mx = 0.0, lf = 0.0, lq = 0.0;
      for (int u = u0; u < u1; ++u) {
        const double v = a.rf[u];
        f += v;
        q += a.rq[u];
        mx = fmax(mx, fabs(v));
      }
      if (a.lr)
        for (int u = u0; u < u1; ++u) {
          lf += a.lr[u];
          lq += a.lq[u];
        }
      fail = finish_row(a, i, f, q, mx, lf, lq, fchk);
    }
  } else if (i < a.n_unk + a.m) {
    const int k = i - a.n_unk;
    const int p = a.vs_p[k], nn = a.vs_n[k];
    const double t = a.X[p] - a.X[nn] - a.vsrc[k];
    a.f[i] = t;
    if (p == 0) a.I_vs[k] = 0.0;
    fchk = fabs(t);
    fail = !(fchk < a.sc->vntol);
  }
  if (a.st) block_flags(dbits(fchk), 0, fail ? 1u : 0u, &a.st->maxf_bits, nullptr, &a.st->res_fail);
}

// ---- LU refactor ------------------------------------------------------------

// LU values <- matrix values; counts non-finite matrix entries into
// st->nonfinite_in (a non-finite Jacobian is not a pivoting problem).
template <typename T>
__global__ void __launch_bounds__(256) k_lu_gather(int nnz, const int32_t* __restrict__ src,
                                                   const double* __restrict__ csr, T* lu,
                                                   DevStatus* st) {
  const int e = blockIdx.x * blockDim.x + threadIdx.x;
  unsigned int bad = 0;
  if (e < nnz) {
    const int s = src[e];
    const double v = s >= 0 ? csr[s] : 0.0;
    lu[e] = (T)v;
    bad = fabs(v) <= 1.7976931348623157e308 ? 0u : 1u;
  }
  if (st) block_flags(0ull, 0ull, bad, nullptr, nullptr, &st->nonfinite_in);
}

// v[i] = src[slots[i]] for i < n by the whole block, 8 per thread and round:
// the round's slot loads, then the value loads they address, then the LDS
// stores, so a round costs two memory latencies whatever n is.
template <typename T>
__device__ __forceinline__ void gather_slots(int n, const int32_t* __restrict__ slots,
                                             const T* src, T* v) {
  const int bd = (int)blockDim.x;
  for (int i0 = (int)threadIdx.x; i0 < n; i0 += 8 * bd) {
    int sl[8];
#pragma unroll
    for (int u = 0; u < 8; ++u) sl[u] = i0 + u * bd < n ? slots[i0 + u * bd] : 0;
    T vv[8];
#pragma unroll
    for (int u = 0; u < 8; ++u)
      if (i0 + u * bd < n) vv[u] = src[sl[u]];
#pragma unroll
    for (int u = 0; u < 8; ++u)
      if (i0 + u * bd < n) v[i0 + u * bd] = vv[u];
  }
}

// dst[slots[i]] = v[i] for i < n by the whole block, 8 per thread and round.
template <typename T>
__device__ __forceinline__ void scatter_slots(int n, const int32_t* __restrict__ slots,
                                              const T* v, T* dst) {
  const int bd = (int)blockDim.x;
  for (int i0 = (int)threadIdx.x; i0 < n; i0 += 8 * bd) {
    int sl[8];
#pragma unroll
    for (int u = 0; u < 8; ++u) sl[u] = i0 + u * bd < n ? slots[i0 + u * bd] : 0;
#pragma unroll
    for (int u = 0; u < 8; ++u)
      if (i0 + u * bd < n) dst[sl[u]] = v[i0 + u * bd];
  }
}

// Task tables of a band program: task t of a band with w levels has item
// offsets tl_ptr[t*(w+1) ...] and per-level workgroup / lane-group counts.
struct BandProg {
  const int32_t* items;
  const int32_t* tl_ptr;
  const int32_t* tl_heavy;
  const int32_t* tl_mid;
};

struct LuProg {
  const int32_t* items;  // item -> target id
  const int32_t* tpos;
  const int32_t* tseg;
  const int32_t* ul;
  const int32_t* uu;
  const int32_t* ud;
  const int2* lu2;  // {ul, uu} interleaved: one 8-byte load per scaled update
};

template <typename T, bool SCALED = false>
struct LuTerm {
  const LuProg& P;
  const T* lu;
  __device__ __forceinline__ T operator()(int u) const {
    // SCALED: L(i,k) already divided by its pivot (the Schur steps of band
    // programs), otherwise the raw L(i,k) is divided here
    if constexpr (SCALED) {
      const int2 q = P.lu2[u];
      return lu[q.x] * lu[q.y];
    }
    return lu[P.ul[u]] * lu[P.uu[u]] / lu[P.ud[u]];
  }
};

// The end of a band task: divide the L entries of the task's own pivots'
// columns (all final now) by their pivots, in HBM, and fold max |L| and the
// count of non-finite entries into the status (the pivot-growth monitor).
// Every thread of the block calls it; the caller has synchronised the block
// after its last writes.
template <typename T>
__device__ __forceinline__ void task_scale_l(int e0, int e1, const int2* __restrict__ sc, T* lu,
                                             DevStatus* st) {
  double mag = 0.0;
  unsigned int bad = 0;
  for (int e = e0 + (int)threadIdx.x; e < e1; e += (int)blockDim.x) {
    const int2 q = sc[e];
    const T v = lu[q.x] / lu[q.y];
    lu[q.x] = v;
    const double m = fabs((double)v);
    if (m <= 1.7976931348623157e308)
      mag = fmax(mag, m);
    else
      ++bad;
  }
  if (st) block_flags(dbits(mag), 0, bad, &st->maxl_bits, nullptr, &st->nonfinite);
}

template <typename T, bool SCALED = false>
__device__ __forceinline__ void lu_light(int t, const LuProg& P, T* lu) {
  T acc = (T)0;
  const LuTerm<T, SCALED> term{P, lu};
  for (int u = P.tseg[t], e = P.tseg[t + 1]; u < e; ++u) acc += term(u);
  lu[P.tpos[t]] -= acc;
}

// all kGroupLanes lanes of a group call
template <typename T, bool SCALED = false>
__device__ __forceinline__ void lu_mid(int t, int lane, const LuProg& P, T* lu) {
  const T acc =
      group_sum<kGroupLanes, T>(P.tseg[t], P.tseg[t + 1], lane, LuTerm<T, SCALED>{P, lu});
  if (lane == 0) lu[P.tpos[t]] -= acc;
}

template <typename T, bool SCALED = false>
__device__ __forceinline__ void lu_heavy(int t, const LuProg& P, T* lu, T* red) {
  const T acc = block_sum<T>(P.tseg[t], P.tseg[t + 1], LuTerm<T, SCALED>{P, lu}, red);
  if (threadIdx.x == 0) lu[P.tpos[t]] -= acc;
}

constexpr int kGroupsPerBlock = 256 / kGroupLanes;

// One wide level: items are target ids in level order (identity): nh heavy
// (a workgroup each), nm medium (kGroupLanes lanes each), then short ones.
template <typename T, bool SCALED = false>
__global__ void __launch_bounds__(256) k_lu_wide(int lo, int hi, int nh, int nm, LuProg P, T* lu) {
  __shared__ T red[RB + 1];
  const int b = blockIdx.x;
  const int nmb = (nm + kGroupsPerBlock - 1) / kGroupsPerBlock;
  if (b < nh) {
    lu_heavy<T, SCALED>(P.items[lo + b], P, lu, red);
    return;
  }
  if (b < nh + nmb) {
    const int grp = (b - nh) * kGroupsPerBlock + (int)threadIdx.x / kGroupLanes;
    if (grp < nm)
      lu_mid<T, SCALED>(P.items[lo + nh + grp], (int)threadIdx.x % kGroupLanes, P, lu);
    return;
  }
  const int g = lo + nh + nm + (b - nh - nmb) * blockDim.x + threadIdx.x;
  if (g < hi) lu_light<T, SCALED>(P.items[g], P, lu);
}

// A run of narrow levels in one workgroup; level bounds cached in LDS.
template <typename T>
__global__ void __launch_bounds__(512) k_lu_narrow(int l0, int l1, const int32_t* __restrict__ lvl,
                                                   const int32_t* __restrict__ lheavy,
                                                   const int32_t* __restrict__ lmid, LuProg P,
                                                   T* lu) {
  __shared__ T red[RB + 1];
  extern __shared__ int32_t s_lv[];
  const int nl = l1 - l0;
  int32_t* s_hv = s_lv + nl + 1;
  int32_t* s_md = s_hv + nl;
  for (int j = threadIdx.x; j <= nl; j += blockDim.x) s_lv[j] = lvl[l0 + j];
  for (int j = threadIdx.x; j < nl; j += blockDim.x) {
    s_hv[j] = lheavy[l0 + j];
    s_md[j] = lmid[l0 + j];
  }
  __syncthreads();
  const int lane = (int)threadIdx.x % kGroupLanes;
  const int ngrp = (int)blockDim.x / kGroupLanes;
  for (int l = 0; l < nl; ++l) {
    const int lo = s_lv[l], hi = s_lv[l + 1], nh = s_hv[l], nm = s_md[l];
    for (int g = lo + nh + nm + (int)threadIdx.x; g < hi; g += blockDim.x)
      lu_light<T>(P.items[g], P, lu);
    for (int grp = (int)threadIdx.x / kGroupLanes; grp < nm; grp += ngrp)
      lu_mid<T>(P.items[lo + nh + grp], lane, P, lu);
    for (int h = 0; h < nh; ++h) lu_heavy<T>(P.items[lo + h], P, lu, red);
    __syncthreads();
  }
}

// One band of the refactorisation: workgroup = task (a group of connected
// pivots of the band), walking the band's levels; its items are the targets
// owned by the task (rows or columns of its own pivots).  The band's updates
// of later pivots' rows and columns (the Schur part) follow as one wide launch.
template <typename T>
__global__ void __launch_bounds__(256) k_lu_band(int w, int ptr_off, int hm_off, int task0,
                                                 BandProg B, LuProg P, const int32_t* sc_ptr,
                                                 const int2* sc, T* lu, DevStatus* st) {
  __shared__ T red[RB + 1];
  __shared__ int32_t s_ptr[kMaxBandWidth + 1], s_h[kMaxBandWidth], s_m[kMaxBandWidth];
  const int t = blockIdx.x;
  const int32_t* tp = B.tl_ptr + ptr_off + (size_t)t * (w + 1);
  const int32_t* th = B.tl_heavy + hm_off + (size_t)t * w;
  const int32_t* tm = B.tl_mid + hm_off + (size_t)t * w;
  for (int j = threadIdx.x; j <= w; j += blockDim.x) s_ptr[j] = tp[j];
  for (int j = threadIdx.x; j < w; j += blockDim.x) {
    s_h[j] = th[j];
    s_m[j] = tm[j];
  }
  __syncthreads();
  const int lane = (int)threadIdx.x % kGroupLanes;
  const int ngrp = (int)blockDim.x / kGroupLanes;
  for (int j = 0; j < w; ++j) {
    const int lo = s_ptr[j], hi = s_ptr[j + 1], nh = s_h[j], nm = s_m[j];
    if (lo == hi) continue;  // uniform across the block
    for (int g = lo + nh + nm + (int)threadIdx.x; g < hi; g += blockDim.x)
      lu_light<T>(B.items[g], P, lu);
    for (int grp = (int)threadIdx.x / kGroupLanes; grp < nm; grp += ngrp)
      lu_mid<T>(B.items[lo + nh + grp], lane, P, lu);
    for (int h = 0; h < nh; ++h) lu_heavy<T>(B.items[lo + h], P, lu, red);
    __syncthreads();
  }
  task_scale_l<T>(sc_ptr[task0 + t], sc_ptr[task0 + t + 1], sc, lu, st);
}

// The same band task with its slots in LDS: the task's targets and every
// operand its updates read are loaded once, the levels run on LDS (one
// dependent global load per item for its record, the rest on-chip), and the
// targets are written back.  Updates reference slots by 16-bit local index:
// packed word a = (L index | U index << 16), word b = pivot index.  Same
// fixed-order sums as the HBM form, so the factors are bit-identical.
template <typename T>
struct LdsTerm {
  const T* v;
  const uint2* up;
  __device__ __forceinline__ T operator()(int k) const {
    const uint2 a = up[k];
    return v[a.x & 0xffffu] * v[a.x >> 16] / v[a.y];
  }
};

// KP, RD and KB set the short-item pipeline (items per lane, levels ahead,
// batch of further items): the deep form hides more latency per task, the
// shallow one needs fewer registers and so keeps more tasks per CU when a
// band has many of them.
template <typename T, int KP, int RD, int KB>
__global__ void __launch_bounds__(256) k_lu_band_lds(int w, int ptr_off, int hm_off, int task0,
                                                     BandProg B, const int4* __restrict__ recs,
                                                     const uint2* __restrict__ upd,
                                                     const int4* __restrict__ tasks,
                                                     const int32_t* __restrict__ slots,
                                                     const int32_t* __restrict__ sc_ptr,
                                                     const uint32_t* __restrict__ lsc, T* lu,
                                                     DevStatus* st) {
  extern __shared__ __align__(16) unsigned char lds_raw[];
  T* vals = reinterpret_cast<T*>(lds_raw);
  __shared__ T red[RB + 1];
  __shared__ int32_t s_ptr[kMaxBandWidth + 1], s_h[kMaxBandWidth], s_m[kMaxBandWidth];
  const int t = blockIdx.x;
  const int4 tk = tasks[task0 + t];  // slot offset, slot count, written count
  const int32_t* tp = B.tl_ptr + ptr_off + (size_t)t * (w + 1);
  const int32_t* th = B.tl_heavy + hm_off + (size_t)t * w;
  const int32_t* tm = B.tl_mid + hm_off + (size_t)t * w;
  for (int j = threadIdx.x; j <= w; j += blockDim.x) s_ptr[j] = tp[j];
  for (int j = threadIdx.x; j < w; j += blockDim.x) {
    s_h[j] = th[j];
    s_m[j] = tm[j];
  }
  gather_slots<T>(tk.y, slots + tk.x, lu, vals);
  __syncthreads();
  const LdsTerm<T> term{vals, upd};
  const int lane = (int)threadIdx.x % kGroupLanes;
  const int ngrp = (int)blockDim.x / kGroupLanes;
  // Short items (at most kShortSegment = 3 updates).  An item with one
  // update carries it in its record ({packed word, pivot}, z = 1), so it
  // needs no dependent load; others read their updates from lds_upd.  Each
  // lane's first KP items of a level (tid, tid + 256, ...) are software
  // pipelined: records RD levels ahead, multi-update items' updates UD levels
  // ahead.  Items beyond KP per lane go in batches of KB whose loads are all
  // in flight together.  An absent item has w = -1 and z = 0.
  constexpr int UD = 1;
  const int bd = (int)blockDim.x;
  const int tid = (int)threadIdx.x;
  auto fetch_recs = [&](int4 (&R)[KP], int jj) {
    const bool in = jj < w;
    const int a = in ? s_ptr[jj] + s_h[jj] + s_m[jj] + tid : 0;
    const int e = in ? s_ptr[jj + 1] : 0;
#pragma unroll
    for (int k = 0; k < KP; ++k) {
      const int g = a + k * bd;
      R[k] = (in && g < e) ? recs[g] : make_int4(0, 0, 0, -1);
    }
  };
  auto fetch_upds = [&](uint2 (&U)[KP][3], const int4 (&R)[KP]) {
#pragma unroll
    for (int k = 0; k < KP; ++k)
#pragma unroll
      for (int q = 0; q < 3; ++q)
        if (R[k].z > 1 && q < R[k].z) U[k][q] = upd[R[k].y + q];
  };
  // the sum of one short item and its write; u holds the updates of a
  // multi-update item (unused for z <= 1)
  auto short_item = [&](const int4& r, const uint2 (&u)[3]) {
    if (r.w < 0) return;
    T acc = (T)0;
    if (r.z == 1) {
      acc += vals[(uint32_t)r.x & 0xffffu] * vals[(uint32_t)r.x >> 16] / vals[r.y];
    } else {
#pragma unroll
      for (int q = 0; q < 3; ++q)
        if (q < r.z) acc += vals[u[q].x & 0xffffu] * vals[u[q].x >> 16] / vals[u[q].y];
    }
    vals[r.w] -= acc;
  };
  int4 R[RD + 1][KP];      // R[s]: records of level j + s
  uint2 U[UD + 1][KP][3];  // U[s]: updates of level j + s
#pragma unroll
  for (int s = 0; s < RD; ++s) fetch_recs(R[s], s);
#pragma unroll
  for (int s = 0; s < UD; ++s) fetch_upds(U[s], R[s]);
  for (int j = 0; j < w; ++j) {
    const int lo = s_ptr[j], hi = s_ptr[j + 1], nh = s_h[j], nm = s_m[j];
    fetch_recs(R[RD], j + RD);
    fetch_upds(U[UD], R[UD]);
#pragma unroll
    for (int k = 0; k < KP; ++k) short_item(R[0][k], U[0][k]);
    // further short items of this lane (levels wider than KP workgroups)
    for (int g0 = lo + nh + nm + tid + KP * bd; g0 < hi; g0 += KB * bd) {
      int4 rb[KB];
#pragma unroll
      for (int u = 0; u < KB; ++u) {
        const int g = g0 + u * bd;
        rb[u] = g < hi ? recs[g] : make_int4(0, 0, 0, -1);
      }
      uint2 ub[KB][3];
#pragma unroll
      for (int u = 0; u < KB; ++u)
#pragma unroll
        for (int q = 0; q < 3; ++q)
          if (rb[u].z > 1 && q < rb[u].z) ub[u][q] = upd[rb[u].y + q];
#pragma unroll
      for (int u = 0; u < KB; ++u) short_item(rb[u], ub[u]);
    }
    for (int grp = (int)threadIdx.x / kGroupLanes; grp < nm; grp += ngrp) {
      const int4 r = recs[lo + nh + grp];
      const T acc = group_sum<kGroupLanes, T>(r.y, r.y + r.z, lane, term);
      if (lane == 0) vals[r.w] -= acc;
    }
    for (int h = 0; h < nh; ++h) {
      const int4 r = recs[lo + h];
      const T acc = block_sum<T>(r.y, r.y + r.z, term, red);
      if (threadIdx.x == 0) vals[r.w] -= acc;
    }
    lds_barrier();  // the levels exchange values through LDS only
#pragma unroll
    for (int k = 0; k < KP; ++k) {
#pragma unroll
      for (int s = 0; s < RD; ++s) R[s][k] = R[s + 1][k];
#pragma unroll
      for (int s = 0; s < UD; ++s)
#pragma unroll
        for (int q = 0; q < 3; ++q) U[s][k][q] = U[s + 1][k][q];
    }
  }
  __syncthreads();
  {
    // the task's pivot columns of L, final now: scaled in LDS before write-back
    double mag = 0.0;
    unsigned int bad = 0;
    for (int e = sc_ptr[task0 + t] + (int)threadIdx.x; e < sc_ptr[task0 + t + 1];
         e += (int)blockDim.x) {
      const uint32_t q = lsc[e];
      const T v = vals[q & 0xffffu] / vals[q >> 16];
      vals[q & 0xffffu] = v;
      const double m = fabs((double)v);
      if (m <= 1.7976931348623157e308)
        mag = fmax(mag, m);
      else
        ++bad;
    }
    if (st) block_flags(dbits(mag), 0, bad, &st->maxl_bits, nullptr, &st->nonfinite);
  }
  __syncthreads();
  scatter_slots<T>(tk.z, slots + tk.x, vals, lu);
}

template <typename T>
__global__ void __launch_bounds__(256) k_lu_scale(int n, const int32_t* __restrict__ ls,
                                                  const int32_t* __restrict__ ld, T* lu,
                                                  DevStatus* st) {
  // grid-stride over a capped grid: few workgroups, hence few status atomics
  double mag = 0.0;
  unsigned int bad = 0;
  for (int e = blockIdx.x * blockDim.x + threadIdx.x; e < n; e += gridDim.x * blockDim.x) {
    const T v = lu[ls[e]] / lu[ld[e]];
    lu[ls[e]] = v;
    const double m = fabs((double)v);
    if (m <= 1.7976931348623157e308)
      mag = fmax(mag, m);
    else
      ++bad;
  }
  if (st) block_flags(dbits(mag), 0, bad, &st->maxl_bits, nullptr, &st->nonfinite);
}

// ---- triangular solves ------------------------------------------------------

template <typename T>
__global__ void __launch_bounds__(256) k_solve_init(int n, const int32_t* __restrict__ p,
                                                    const double* __restrict__ rhs, double scale,
                                                    T* y) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  if (scale == -1.0)
    y[i] = -(T)rhs[p[i]];
  else
    y[i] = (T)rhs[p[i]];
}

struct TriProg {
  const int32_t* items;
  const int32_t* rp;
  const int32_t* col;
  const int32_t* dg;
};

template <typename T, bool FWD>
__device__ __forceinline__ void tri_light(int i, const TriProg& P, const T* __restrict__ lu, T* y,
                                          T* z) {
  T acc = (T)0;
  if (FWD) {
    for (int e = P.rp[i], d = P.dg[i]; e < d; ++e) acc += lu[e] * y[P.col[e]];
    y[i] = y[i] - acc;
  } else {
    const int d = P.dg[i];
    for (int e = d + 1, end = P.rp[i + 1]; e < end; ++e) acc += lu[e] * z[P.col[e]];
    z[i] = (y[i] - acc) / lu[d];
  }
}

template <typename T>
struct TriTerm {
  const T* lu;
  const T* v;
  const int32_t* col;
  __device__ __forceinline__ T operator()(int e) const { return lu[e] * v[col[e]]; }
};

// A long row of a solve band: one wavefront, lane = lane in the wavefront.
template <typename T, bool FWD>
__device__ __forceinline__ void tri_wave(int i, int lane, const TriProg& P,
                                         const T* __restrict__ lu, T* y, T* z) {
  if (FWD) {
    const T acc = group_sum<kWaveLanes, T>(P.rp[i], P.dg[i], lane, TriTerm<T>{lu, y, P.col});
    if (lane == 0) y[i] = y[i] - acc;
  } else {
    const int d = P.dg[i];
    const T acc = group_sum<kWaveLanes, T>(d + 1, P.rp[i + 1], lane, TriTerm<T>{lu, z, P.col});
    if (lane == 0) z[i] = (y[i] - acc) / lu[d];
  }
}

template <typename T, bool FWD>
__device__ __forceinline__ void tri_mid(int i, int lane, const TriProg& P,
                                        const T* __restrict__ lu, T* y, T* z) {
  if (FWD) {
    const T acc = group_sum<kGroupLanes, T>(P.rp[i], P.dg[i], lane, TriTerm<T>{lu, y, P.col});
    if (lane == 0) y[i] = y[i] - acc;
  } else {
    const int d = P.dg[i];
    const T acc = group_sum<kGroupLanes, T>(d + 1, P.rp[i + 1], lane, TriTerm<T>{lu, z, P.col});
    if (lane == 0) z[i] = (y[i] - acc) / lu[d];
  }
}

template <typename T, bool FWD>
__device__ __forceinline__ void tri_heavy(int i, const TriProg& P, const T* __restrict__ lu, T* y,
                                          T* z, T* red) {
  if (FWD) {
    const T acc = block_sum<T>(P.rp[i], P.dg[i], TriTerm<T>{lu, y, P.col}, red);
    if (threadIdx.x == 0) y[i] = y[i] - acc;
  } else {
    const int d = P.dg[i];
    const T acc = block_sum<T>(d + 1, P.rp[i + 1], TriTerm<T>{lu, z, P.col}, red);
    if (threadIdx.x == 0) z[i] = (y[i] - acc) / lu[d];
  }
}

template <typename T, bool FWD>
__global__ void __launch_bounds__(256) k_tri_wide(int lo, int hi, int nh, int nm, TriProg P,
                                                  const T* __restrict__ lu, T* y, T* z) {
  __shared__ T red[RB + 1];
  const int b = blockIdx.x;
  const int nmb = (nm + kGroupsPerBlock - 1) / kGroupsPerBlock;
  if (b < nh) {
    tri_heavy<T, FWD>(P.items[lo + b], P, lu, y, z, red);
    return;
  }
  if (b < nh + nmb) {
    const int grp = (b - nh) * kGroupsPerBlock + (int)threadIdx.x / kGroupLanes;
    if (grp < nm)
      tri_mid<T, FWD>(P.items[lo + nh + grp], (int)threadIdx.x % kGroupLanes, P, lu, y, z);
    return;
  }
  const int g = lo + nh + nm + (b - nh - nmb) * blockDim.x + threadIdx.x;
  if (g < hi) tri_light<T, FWD>(P.items[g], P, lu, y, z);
}

template <typename T, bool FWD>
__global__ void __launch_bounds__(512) k_tri_narrow(int l0, int l1, const int32_t* __restrict__ lvl,
                                                    const int32_t* __restrict__ lheavy,
                                                    const int32_t* __restrict__ lmid, TriProg P,
                                                    const T* __restrict__ lu, T* y, T* z) {
  __shared__ T red[RB + 1];
  extern __shared__ int32_t s_lv[];
  const int nl = l1 - l0;
  int32_t* s_hv = s_lv + nl + 1;
  int32_t* s_md = s_hv + nl;
  for (int j = threadIdx.x; j <= nl; j += blockDim.x) s_lv[j] = lvl[l0 + j];
  for (int j = threadIdx.x; j < nl; j += blockDim.x) {
    s_hv[j] = lheavy[l0 + j];
    s_md[j] = lmid[l0 + j];
  }
  __syncthreads();
  const int lane = (int)threadIdx.x % kGroupLanes;
  const int ngrp = (int)blockDim.x / kGroupLanes;
  for (int l = 0; l < nl; ++l) {
    const int lo = s_lv[l], hi = s_lv[l + 1], nh = s_hv[l], nm = s_md[l];
    for (int g = lo + nh + nm + (int)threadIdx.x; g < hi; g += blockDim.x)
      tri_light<T, FWD>(P.items[g], P, lu, y, z);
    for (int grp = (int)threadIdx.x / kGroupLanes; grp < nm; grp += ngrp)
      tri_mid<T, FWD>(P.items[lo + nh + grp], lane, P, lu, y, z);
    for (int h = 0; h < nh; ++h) tri_heavy<T, FWD>(P.items[lo + h], P, lu, y, z, red);
    __syncthreads();
  }
}

// One band of a triangular solve: workgroup = task, walking the band's w
// levels of its own rows (connected components packed by the host) with a
// barrier between levels.  Rows of earlier bands are final at launch.
template <typename T, bool FWD>
__global__ void __launch_bounds__(256) k_tri_band(int w, int ptr_off, int hm_off, BandProg B,
                                                  TriProg P, const T* __restrict__ lu, T* y, T* z) {
  __shared__ int32_t s_ptr[kMaxBandWidth + 1], s_h[kMaxBandWidth], s_m[kMaxBandWidth];
  const int t = blockIdx.x;
  const int32_t* tp = B.tl_ptr + ptr_off + (size_t)t * (w + 1);
  const int32_t* th = B.tl_heavy + hm_off + (size_t)t * w;
  const int32_t* tm = B.tl_mid + hm_off + (size_t)t * w;
  for (int j = threadIdx.x; j <= w; j += blockDim.x) s_ptr[j] = tp[j];
  for (int j = threadIdx.x; j < w; j += blockDim.x) {
    s_h[j] = th[j];
    s_m[j] = tm[j];
  }
  __syncthreads();
  const int lane = (int)threadIdx.x % kGroupLanes;
  const int ngrp = (int)blockDim.x / kGroupLanes;
  for (int j = 0; j < w; ++j) {
    const int lo = s_ptr[j], hi = s_ptr[j + 1], nh = s_h[j], nm = s_m[j];
    if (lo == hi) continue;  // uniform across the block
    for (int g = lo + nh + nm + (int)threadIdx.x; g < hi; g += blockDim.x)
      tri_light<T, FWD>(B.items[g], P, lu, y, z);
    for (int grp = (int)threadIdx.x / kGroupLanes; grp < nm; grp += ngrp)
      tri_mid<T, FWD>(B.items[lo + nh + grp], lane, P, lu, y, z);
    for (int h = (int)threadIdx.x / kWaveLanes; h < nh; h += (int)blockDim.x / kWaveLanes)
      tri_wave<T, FWD>(B.items[lo + h], (int)threadIdx.x % kWaveLanes, P, lu, y, z);
    __syncthreads();
  }
}

// The band task with everything its levels read in LDS.  The prologue issues
// all of the task's global loads at once -- vector values of its rows and of
// the rows they read, its row records, the values and local columns of its
// entries (a contiguous run of lds_col / lds_src), and for the backward solve
// each row's right-hand side and pivot -- so the task waits on memory latency
// once instead of once per level; the levels then run on LDS and barriers
// only.  Rows, classes and sums are those of the HBM band form (light rows
// sequential, medium rows 16-lane sums, long rows one-wavefront sums), so
// results are bit-identical.
//
// Dynamic LDS (the host sizes it from the band's largest task):
//   v[nv] T | ent[ne] T | (16-byte aligned) rec[nr] int4 |
//   backward only: rhs[nr] T, piv[nr] T | ecol[ne] uint16
template <typename T>
struct LdeTerm {
  const T* ent;
  const T* v;
  const uint16_t* ecol;
  int shift;  // LDS entry index of LU slot e is e + shift
  __device__ __forceinline__ T operator()(int e) const {
    return ent[e + shift] * v[ecol[e + shift]];
  }
};

__host__ __device__ inline size_t lde_rec_offset(int nv, int ne, size_t tsize) {
  return (((size_t)nv + (size_t)ne) * tsize + 15) & ~(size_t)15;
}

__host__ __device__ inline size_t lde_col_offset(int nv, int ne, int nr, size_t tsize, bool fwd) {
  return lde_rec_offset(nv, ne, tsize) + (size_t)nr * 16 + (fwd ? 0 : 2 * (size_t)nr * tsize);
}

template <typename T, bool FWD>
__global__ void __launch_bounds__(256) k_tri_band_lde(int w, int ptr_off, int hm_off, int task0,
                                                      int nv, int ne_max, int nr_max, BandProg B,
                                                      const int4* __restrict__ recs,
                                                      const uint16_t* __restrict__ lcol,
                                                      const int32_t* __restrict__ esrc,
                                                      const int4* __restrict__ tasks,
                                                      const int32_t* __restrict__ slots,
                                                      const T* __restrict__ lu, T* y, T* z) {
  extern __shared__ __align__(16) unsigned char lds_raw[];
  T* v = reinterpret_cast<T*>(lds_raw);
  T* ent = v + nv;
  int4* rec = reinterpret_cast<int4*>(lds_raw + lde_rec_offset(nv, ne_max, sizeof(T)));
  T* rhs = reinterpret_cast<T*>(rec + nr_max);
  T* piv = rhs + nr_max;
  uint16_t* ecol =
      reinterpret_cast<uint16_t*>(lds_raw + lde_col_offset(nv, ne_max, nr_max, sizeof(T), FWD));
  __shared__ int32_t s_ptr[kMaxBandWidth + 1], s_h[kMaxBandWidth], s_m[kMaxBandWidth];
  const int t = blockIdx.x;
  const int tid = (int)threadIdx.x;
  const int4 tk = tasks[task0 + t];  // slot offset, slot count, own rows
  const int32_t* tp = B.tl_ptr + ptr_off + (size_t)t * (w + 1);
  const int32_t* th = B.tl_heavy + hm_off + (size_t)t * w;
  const int32_t* tm = B.tl_mid + hm_off + (size_t)t * w;
  for (int j = tid; j <= w; j += blockDim.x) s_ptr[j] = tp[j];
  for (int j = tid; j < w; j += blockDim.x) {
    s_h[j] = th[j];
    s_m[j] = tm[j];
  }
  const int g0 = tp[0];
  const int nr = tk.z;  // the task's rows: items [g0, g0 + nr)
  if (nr == 0) return;  // uniform across the block
  const int4 rf = recs[g0];
  const int4 rl = recs[g0 + nr - 1];
  const int cb = rf.w;                          // first entry in lds_col / lds_src
  const int ne = rl.w + (rl.z - rl.y) - cb;     // the task's entries (<= ne_max)
  const T* src = FWD ? y : z;
  const int bd = (int)blockDim.x;
  // Every load below is independent of the others in its round: each thread
  // first issues a round's index loads, then the value loads they address,
  // then stores to LDS, so a round costs two memory latencies.
  constexpr int EC = 16;  // entries per thread and round
  int si[EC] = {};
  uint16_t cc[EC] = {};
  auto entry_indices = [&](int k0) {
#pragma unroll
    for (int u = 0; u < EC; ++u) {
      const int k = k0 + u * bd;
      if (k < ne) {
        si[u] = esrc[cb + k];
        cc[u] = lcol[cb + k];
      }
    }
  };
  entry_indices(tid);  // in flight while the vector values load
  gather_slots<T>(tk.y, slots + tk.x, src, v);
  for (int k = tid; k < nr; k += bd) {
    const int4 r = recs[g0 + k];
    rec[k] = r;
    if (!FWD) {
      rhs[k] = y[B.items[g0 + k]];
      piv[k] = lu[r.y - 1];  // the diagonal precedes the U part
    }
  }
  // entries: this round's values, then the next round's indices, then the
  // stores (the next indices stay in flight across the stores)
  for (int k0 = tid; k0 < ne; k0 += EC * bd) {
    T val[EC];
    uint16_t cur[EC];
#pragma unroll
    for (int u = 0; u < EC; ++u) {
      if (k0 + u * bd < ne) val[u] = lu[si[u]];
      cur[u] = cc[u];
    }
    entry_indices(k0 + EC * bd);
#pragma unroll
    for (int u = 0; u < EC; ++u) {
      const int k = k0 + u * bd;
      if (k < ne) {
        ent[k] = val[u];
        ecol[k] = cur[u];
      }
    }
  }
  __syncthreads();
  auto finish = [&](int k, const int4& r, T acc) {
    if (FWD)
      v[r.x] = v[r.x] - acc;
    else
      v[r.x] = (rhs[k] - acc) / piv[k];
  };
  const int lane = tid % kGroupLanes;
  const int ngrp = (int)blockDim.x / kGroupLanes;
  for (int j = 0; j < w; ++j) {
    const int lo = s_ptr[j], hi = s_ptr[j + 1], nh = s_h[j], nm = s_m[j];
    if (lo == hi) continue;  // uniform across the block
    for (int g = lo + nh + nm + tid; g < hi; g += blockDim.x) {
      const int k = g - g0;
      const int4 r = rec[k];
      const LdeTerm<T> term{ent, v, ecol, r.w - cb - r.y};
      T acc = (T)0;
      for (int e = r.y; e < r.z; ++e) acc += term(e);
      finish(k, r, acc);
    }
    for (int grp = tid / kGroupLanes; grp < nm; grp += ngrp) {
      const int k = lo + nh + grp - g0;
      const int4 r = rec[k];
      const T acc =
          group_sum<kGroupLanes, T>(r.y, r.z, lane, LdeTerm<T>{ent, v, ecol, r.w - cb - r.y});
      if (lane == 0) finish(k, r, acc);
    }
    for (int h = tid / kWaveLanes; h < nh; h += (int)blockDim.x / kWaveLanes) {
      const int k = lo + h - g0;
      const int4 r = rec[k];
      const T acc = group_sum<kWaveLanes, T>(r.y, r.z, tid % kWaveLanes,
                                             LdeTerm<T>{ent, v, ecol, r.w - cb - r.y});
      if (tid % kWaveLanes == 0) finish(k, r, acc);
    }
    lds_barrier();  // the levels exchange values through LDS only
  }
  T* dst = FWD ? y : z;
  scatter_slots<T>(nr, slots + tk.x, v, dst);
}

// Rows longer than kHubChunk: chunk partials over the grid, then one lane per
// row adds them in chunk order.
template <typename T, bool FWD>
__global__ void __launch_bounds__(256) k_tri_hub_part(const int32_t* __restrict__ cs,
                                                      const int32_t* __restrict__ ce, TriProg P,
                                                      const T* __restrict__ lu, const T* y,
                                                      const T* z, T* part) {
  __shared__ T red[RB + 1];
  const int c = blockIdx.x;
  const T s = block_sum<T>(cs[c], ce[c], TriTerm<T>{lu, FWD ? y : z, P.col}, red);
  if (threadIdx.x == 0) part[c] = s;
}

template <typename T, bool FWD>
__global__ void k_tri_hub_finish(int h0, int h1, const int32_t* __restrict__ rows,
                                 const int32_t* __restrict__ cptr, const T* __restrict__ part,
                                 TriProg P, const T* __restrict__ lu, T* y, T* z) {
  const int h = h0 + blockIdx.x * blockDim.x + threadIdx.x;
  if (h >= h1) return;
  const int i = rows[h];
  T acc = (T)0;
  for (int c = cptr[h]; c < cptr[h + 1]; ++c) acc += part[c];
  if (FWD)
    y[i] = y[i] - acc;
  else
    z[i] = (y[i] - acc) / lu[P.dg[i]];
}

template <typename T>
__global__ void __launch_bounds__(256) k_solve_finish(int n, const int32_t* __restrict__ q,
                                                      const T* __restrict__ z, double* x) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  x[q[i]] = (double)z[i];
}

// out = base + sgn * (A x)   (base may be null); long rows by a workgroup
__global__ void __launch_bounds__(256) k_spmv(int n, const int32_t* __restrict__ rp,
                                              const int32_t* __restrict__ ci,
                                              const double* __restrict__ a,
                                              const double* __restrict__ x,
                                              const double* __restrict__ base, double sgn,
                                              double* out, HubChunks hub) {
  if ((int)blockIdx.x < hub.n_items) {
    if (threadIdx.x == 0) {
      const int i = hub.item[blockIdx.x];
      double acc = 0.0;
      for (int c = hub.ptr[blockIdx.x]; c < hub.ptr[blockIdx.x + 1]; ++c) acc += hub.part[c];
      if (sgn != 1.0) acc = sgn * acc;
      out[i] = base ? base[i] + acc : acc;
    }
    return;
  }
  const int i = (blockIdx.x - hub.n_items) * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const int e0 = rp[i], e1 = rp[i + 1];
  if (e1 - e0 > kHeavySegment) return;
  double acc = 0.0;
  for (int e = e0; e < e1; ++e) acc += a[e] * x[ci[e]];
  if (sgn != 1.0) acc = sgn * acc;
  out[i] = base ? base[i] + acc : acc;
}

__global__ void k_axpy(int n, double* x, const double* d) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) x[i] = x[i] + d[i];
}

// Independent-source waveforms, vajax/analysis/sources.py: dc :25-28,
// pulse :31-63, sine :66-78, pwl :81-106 (jnp.interp: clamped ends, the
// segment found by a right-sided search), exp :109-125, am :128-156,
// fm :159-186.  Parameter columns as documented in vjx.h.
__device__ double wave_value(const SourceTable& S, int i, double t) {
  const size_t m = (size_t)S.n;
  const double* p = S.par;
  auto P = [&](int c) { return p[c * m + i]; };
  const double kPi = 3.141592653589793;
  switch (S.wave[i]) {
    case VJX_WAVE_PULSE: {
      const double v0 = P(1), v1 = P(2), r = P(3), f = P(4), wd = P(5), per = P(6), d = P(7);
      if (t < d) return v0;
      double tin = fmod(t - d, per);
      if (tin < 0.0) tin += per;
      if (tin < r) return v0 + (v1 - v0) * tin / r;
      if (tin < r + wd) return v1;
      if (tin < r + wd + f) return v1 - (v1 - v0) * (tin - r - wd) / f;
      return v0;
    }
    case VJX_WAVE_SINE:
      return P(1) + P(2) * sin(2.0 * kPi * P(3) * t + P(4));
    case VJX_WAVE_PWL: {
      const double per = P(1);
      if (per > 0.0) {
        t = fmod(t, per);
        if (t < 0.0) t += per;
      }
      const int a = S.pwl_ptr[i], b = S.pwl_ptr[i + 1];
      const double* xp = S.pwl_t;
      const double* fp = S.pwl_v;
      if (t < xp[a]) return fp[a];
      if (t > xp[b - 1]) return fp[b - 1];
      if (b - a == 1) return fp[a];
      // first index in (a, b) with xp > t, clamped to [a + 1, b - 1]
      int lo = a, hi = b;
      while (lo < hi) {
        const int mid = (lo + hi) >> 1;
        if (xp[mid] <= t)
          lo = mid + 1;
        else
          hi = mid;
      }
      const int j = lo < a + 1 ? a + 1 : (lo > b - 1 ? b - 1 : lo);
      const double dx = xp[j] - xp[j - 1];
      if (fabs(dx) <= 4.930380657631324e-32) return fp[j - 1];  // spacing(DBL_EPSILON)
      return fp[j - 1] + ((t - xp[j - 1]) / dx) * (fp[j] - fp[j - 1]);
    }
    case VJX_WAVE_EXP: {
      const double v0 = P(1), v1 = P(2), d = P(3), td = P(4), t1 = P(5), t2 = P(6);
      const double tf = d + td;
      if (t < d) return v0;
      if (t < tf) return v0 + (v1 - v0) * (1.0 - exp(-(t - d) / t1));
      const double risen = v0 + (v1 - v0) * (1.0 - exp(-td / t1));
      return risen + (v0 - risen) * (1.0 - exp(-(t - tf) / t2));
    }
    case VJX_WAVE_AM:
    case VJX_WAVE_FM: {
      const double dc = P(1), a = P(2), f = P(3), ph = P(4), d = P(5), mf = P(6), mph = P(7),
                   mi = P(8);
      if (S.wave[i] == VJX_WAVE_AM) {
        if (t < d) return dc + a * sin(ph * kPi / 180) * (1 + mi * sin(mph * kPi / 180));
        const double carrier = sin(2 * kPi * f * (t - d) + ph * kPi / 180);
        const double modulation = 1 + mi * sin(2 * kPi * mf * (t - d) + mph * kPi / 180);
        return dc + a * carrier * modulation;
      }
      if (t < d) return dc + a * sin(ph * kPi / 180 + mi * sin(mph * kPi / 180));
      const double cph = 2 * kPi * f * (t - d) + ph * kPi / 180;
      const double mod = mi * sin(2 * kPi * mf * (t - d) + mph * kPi / 180);
      return dc + a * sin(cph + mod);
    }
    default:
      return P(0);
  }
}

__global__ void k_sources(SourceTable src, double t, double* out) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < src.n) out[i] = wave_value(src, i, t);
}

// Both source tables at t in one launch: workgroups [0, nba) the first,
// the rest the second.
__global__ void k_sources2(SourceTable a, double* out_a, int nba, SourceTable b, double* out_b,
                           double t) {
  const bool first = (int)blockIdx.x < nba;
  const int blk = first ? (int)blockIdx.x : (int)blockIdx.x - nba;
  const int i = blk * blockDim.x + threadIdx.x;
  if (first) {
    if (i < a.n) out_a[i] = wave_value(a, i, t);
  } else if (i < b.n) {
    out_b[i] = wave_value(b, i, t);
  }
}

__device__ __forceinline__ void status_clear(DevStatus* st) {
  st->maxf_bits = 0;
  st->maxdv_bits = 0;
  st->maxdelta_bits = 0;
  st->maxl_bits = 0;
  st->limd_bits = 0;
  st->limr_bits = 0;
  st->res_fail = 0;
  st->delta_fail = 0;
  st->nonfinite = 0;
  st->nonfinite_in = 0;
}

__global__ void k_status_reset(DevStatus* st) {
  if (threadIdx.x == 0 && blockIdx.x == 0) status_clear(st);
}

// The start of a Newton iteration in one launch: the iterate and the limit
// states saved for a re-pivot redo, and up to three status records cleared.
__global__ void __launch_bounds__(256)
    k_iter_begin(int n, const double* __restrict__ X, double* __restrict__ Xsave, int nl,
                 const double* __restrict__ ls, double* __restrict__ ls_save, DevStatus* a,
                 DevStatus* b, DevStatus* c) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) Xsave[i] = X[i];
  if (i < nl) ls_save[i] = ls[i];
  if (i == 0) {
    if (a) status_clear(a);
    if (b) status_clear(b);
    if (c) status_clear(c);
  }
}

inline int nblk(long long n, int b) { return (int)((n + b - 1) / b); }

}  // namespace

static bool four_terminal(int id) { return id == VJX_MODEL_MOSFET || id == VJX_MODEL_EKV26; }
int model_n_v(int id) { return four_terminal(id) ? 3 : 1; }
int model_n_res(int id) { return four_terminal(id) ? 4 : 2; }
int model_n_jac(int id) { return four_terminal(id) ? 16 : 4; }
int model_n_par(int id) {
  return id == VJX_MODEL_MOSFET  ? ModelTraits<VJX_MODEL_MOSFET>::NP
         : id == VJX_MODEL_EKV26 ? ModelTraits<VJX_MODEL_EKV26>::NP
         : id == VJX_MODEL_DIODE ? ModelTraits<VJX_MODEL_DIODE>::NP
                                 : 1;
}
int model_n_lim(int id) { return id == VJX_MODEL_DIODE ? 1 : 0; }

void launch_eval(const DevModel& m, const double* X, const DevScalars* sc, EvalOut out,
                 hipStream_t s) {
  if (m.count == 0) return;
  const int B = m.block;  // 64, 128 or 256 (the kernels' launch bound)
  const int G = nblk(m.count, B);
  const double* li = out.ls_in ? out.ls_in + m.lim_off : nullptr;
  double* lo = out.ls_out ? out.ls_out + m.lim_off : nullptr;
  ktimer_begin(s);
  switch (m.model_id) {
    case VJX_MODEL_RESISTOR:
      k_eval<VJX_MODEL_RESISTOR><<<G, B, 0, s>>>(m.count, m.vn1.p, m.vn2.p, m.jac_dst.p,
                                                 m.res_dst.p, m.par.p, X, sc, out.jac, out.rf,
                                                 out.rq, nullptr, nullptr, nullptr, nullptr);
      break;
    case VJX_MODEL_CAPACITOR:
      k_eval<VJX_MODEL_CAPACITOR><<<G, B, 0, s>>>(m.count, m.vn1.p, m.vn2.p, m.jac_dst.p,
                                                  m.res_dst.p, m.par.p, X, sc, out.jac, out.rf,
                                                  out.rq, nullptr, nullptr, nullptr, nullptr);
      break;
    case VJX_MODEL_MOSFET:
      k_eval<VJX_MODEL_MOSFET><<<G, B, 0, s>>>(m.count, m.vn1.p, m.vn2.p, m.jac_dst.p,
                                               m.res_dst.p, m.par.p, X, sc, out.jac, out.rf,
                                               out.rq, nullptr, nullptr, nullptr, nullptr);
      break;
    case VJX_MODEL_EKV26:
      k_eval<VJX_MODEL_EKV26><<<G, B, 0, s>>>(m.count, m.vn1.p, m.vn2.p, m.jac_dst.p,
                                              m.res_dst.p, m.par.p, X, sc, out.jac, out.rf,
                                              out.rq, nullptr, nullptr, nullptr, nullptr);
      break;
    case VJX_MODEL_DIODE:
      k_eval<VJX_MODEL_DIODE><<<G, B, 0, s>>>(m.count, m.vn1.p, m.vn2.p, m.jac_dst.p,
                                              m.res_dst.p, m.par.p, X, sc, out.jac, out.rf, out.rq,
                                              out.lr, out.lq, li, lo);
      break;
  }
  ktimer_end(KC_EVAL, s);
}

void launch_model_eval_raw(int id, int count, const double* vb, const double* par,
                           const double* li, double* rr, double* rc, double* jr, double* jc,
                           double* lr, double* lc, double* lo, hipStream_t s) {
  if (count == 0) return;
  const int B = 256, G = nblk(count, B);
  switch (id) {
    case VJX_MODEL_RESISTOR:
      k_model_eval<VJX_MODEL_RESISTOR><<<G, B, 0, s>>>(count, vb, par, li, rr, rc, jr, jc, lr, lc,
                                                       lo);
      break;
    case VJX_MODEL_CAPACITOR:
      k_model_eval<VJX_MODEL_CAPACITOR><<<G, B, 0, s>>>(count, vb, par, li, rr, rc, jr, jc, lr,
                                                        lc, lo);
      break;
    case VJX_MODEL_MOSFET:
      k_model_eval<VJX_MODEL_MOSFET><<<G, B, 0, s>>>(count, vb, par, li, rr, rc, jr, jc, lr, lc,
                                                     lo);
      break;
    case VJX_MODEL_EKV26:
      k_model_eval<VJX_MODEL_EKV26><<<G, B, 0, s>>>(count, vb, par, li, rr, rc, jr, jc, lr, lc,
                                                    lo);
      break;
    case VJX_MODEL_DIODE:
      k_model_eval<VJX_MODEL_DIODE><<<G, B, 0, s>>>(count, vb, par, li, rr, rc, jr, jc, lr, lc,
                                                    lo);
      break;
  }
}

// Elementwise pnjlim (kind 0: p1 = vt, p2 = vcrit) or fetlim (kind 1: p1 = vto).
__global__ void k_limit(int kind, int n, const double* __restrict__ vnew,
                        const double* __restrict__ vold, double p1, double p2, double* out) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  out[i] = kind == 0 ? pnjlim(vnew[i], vold[i], p1, p2) : fetlim(vnew[i], vold[i], p1);
}

void launch_split_gather(int N, int n_v, int n_par, int n_dp, int n_dc, const int32_t* vpos,
                         const int32_t* csrc, const double* sc, const double* dp, const double* dc,
                         double* vb, double* par, hipStream_t s) {
  if (N > 0) k_split_gather<<<nblk(N, 256), 256, 0, s>>>(N, n_v, n_par, n_dp, n_dc, vpos, csrc, sc,
                                                         dp, dc, vb, par);
}
void launch_split_finish(int N, int n_res, int n_jac, int limited, const double* mf, double* rr,
                         double* rc, double* jr, double* jc, double* lr, double* lc, double* sr,
                         double* sq, hipStream_t s) {
  if (N > 0) k_split_finish<<<nblk(N, 256), 256, 0, s>>>(N, n_res, n_jac, limited, mf, rr, rc, jr,
                                                         jc, lr, lc, sr, sq);
}

void launch_limit(int kind, int n, const double* vnew, const double* vold, double p1, double p2,
                  double* out, hipStream_t s) {
  if (n == 0) return;
  k_limit<<<nblk(n, 256), 256, 0, s>>>(kind, n, vnew, vold, p1, p2, out);
}

// Limit-state convergence terms (solver_factories.py:319-321): max
// |ls_out - ls_in| and max |ls_in| into the status (order-independent maxima).
__global__ void __launch_bounds__(256) k_lim_stats(int n, const double* __restrict__ ls_in,
                                                   const double* __restrict__ ls_out,
                                                   DevStatus* st) {
  double dmax = 0.0, rmax = 0.0;
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
    dmax = fmax(dmax, fabs(ls_out[i] - ls_in[i]));
    rmax = fmax(rmax, fabs(ls_in[i]));
  }
  block_flags(dbits(dmax), dbits(rmax), 0u, &st->limd_bits, &st->limr_bits, nullptr);
}

void launch_lim_stats(int n, const double* ls_in, const double* ls_out, DevStatus* st,
                      hipStream_t s) {
  if (n == 0) return;
  k_lim_stats<<<std::min(nblk(n, 256), 64), 256, 0, s>>>(n, ls_in, ls_out, st);
}

void launch_isrc_contrib(int k, const double* isrc, const int32_t* is_dst, double* rf, double* rq,
                         hipStream_t s) {
  if (k == 0) return;
  k_isrc<<<nblk(k, 64), 64, 0, s>>>(k, isrc, is_dst, rf, rq);
}

void launch_reduce_jac(int nse, const int32_t* seg, const double* contrib, const int32_t* slot_diag,
                       const double* slot_b, const DevScalars* sc, double* csr,
                       const HubChunks& hub, hipStream_t s) {
  if (nse == 0) return;
  ktimer_begin(s);
  if (hub.n_chunks > 0) k_chunk_sum<<<hub.n_chunks, 256, 0, s>>>(hub.cs, hub.ce, contrib, hub.part);
  k_reduce_jac<<<hub.n_items + nblk(nse, 256), 256, 0, s>>>(nse, seg, contrib, slot_diag, slot_b,
                                                            sc, csr, hub);
  ktimer_end(KC_STAMP, s);
}

void launch_residual(const ResidualArgs& a, hipStream_t s) {
  const int n = a.n_unk + a.m;
  if (n == 0) return;
  const HubChunks& h = a.hub;
  ktimer_begin(s);
  if (h.n_chunks > 0)
    k_chunk_res<<<h.n_chunks, 256, 0, s>>>(h.cs, h.ce, a.rf, a.rq, a.lr, a.lq, h.part, h.part2,
                                           h.part3, h.part4, h.part5);
  k_residual<<<h.n_items + nblk(n, 256), 256, 0, s>>>(a);
  ktimer_end(KC_STAMP, s);
}

template <typename T>
static void run_fact_bands(DevLU& L, T* lu, DevStatus* st, hipStream_t s) {
  const LuSymbolic& S = L.sym;
  const int2* sc = reinterpret_cast<const int2*>(L.task_scale.p);
  const int32_t* scp = L.task_scale_ptr.p;
  const uint32_t* lsc = L.lds_scale.p;
  const BandSchedule& H = S.fact_band;
  const LuProg P{L.fact_band.items.p, L.tgt_pos.p, L.tgt_seg.p, L.upd_l.p, L.upd_u.p, L.upd_d.p,
                 reinterpret_cast<const int2*>(L.upd_lu.p)};
  const BandProg B{L.fact_band.items.p, L.fact_band.tl_ptr.p, L.fact_band.tl_heavy.p,
                   L.fact_band.tl_mid.p};
  for (size_t b = 0; b + 1 < H.band_task.size(); ++b) {
    const int ntask = H.band_task[b + 1] - H.band_task[b];
    if (ntask > 0) {
      ktimer_begin(s);
      const int nv = S.band_lds.empty() ? 0 : S.band_lds[b];
      const bool deep = L.lu_deep > 0 ? L.lu_deep == 2 : ntask <= 2 * L.n_cu;
      if (nv > 0 && L.use_lds && deep)
        k_lu_band_lds<T, 4, 3, 8><<<ntask, 256, sizeof(T) * (size_t)nv, s>>>(
            H.band_w[b], H.band_ptr_off[b], H.band_hm_off[b], H.band_task[b], B,
            reinterpret_cast<const int4*>(L.lds_item.p), reinterpret_cast<const uint2*>(L.lds_upd.p),
            reinterpret_cast<const int4*>(L.lds_task.p), L.lds_slots.p, scp, lsc, lu, st);
      else if (nv > 0 && L.use_lds)
        k_lu_band_lds<T, 1, 2, 2><<<ntask, 256, sizeof(T) * (size_t)nv, s>>>(
            H.band_w[b], H.band_ptr_off[b], H.band_hm_off[b], H.band_task[b], B,
            reinterpret_cast<const int4*>(L.lds_item.p), reinterpret_cast<const uint2*>(L.lds_upd.p),
            reinterpret_cast<const int4*>(L.lds_task.p), L.lds_slots.p, scp, lsc, lu, st);
      else
        k_lu_band<T><<<ntask, 256, 0, s>>>(H.band_w[b], H.band_ptr_off[b], H.band_hm_off[b],
                                           H.band_task[b], B, P, scp, sc, lu, st);
      ktimer_end(KC_LU_TASKS, s);
    }
    const int lo = S.fact_schur_lo[b], hi = S.fact_schur_hi[b];
    if (hi > lo) {
      const int nh = S.fact_schur_heavy[b], nm = S.fact_schur_mid[b];
      const int grid = nh + nblk(nm, kGroupsPerBlock) + nblk(hi - lo - nh - nm, 256);
      ktimer_begin(s);
      k_lu_wide<T, true><<<grid, 256, 0, s>>>(lo, hi, nh, nm, P, lu);
      ktimer_end(KC_LU_SCHUR, s);
    }
  }
}

template <typename T>
static void run_fact_groups(DevLU& L, T* lu, DevStatus* st, hipStream_t s) {
  if (L.fact_band.on) {
    run_fact_bands<T>(L, lu, st, s);
    return;
  }
  const LuProg P{L.fact.items.p, L.tgt_pos.p, L.tgt_seg.p, L.upd_l.p, L.upd_u.p, L.upd_d.p,
                 nullptr};
  for (const auto& g : L.fact.groups) {
    ktimer_begin(s);
    if (g.wide) {
      const int lo = L.fact.lvl_ptr[g.lvl_begin], hi = L.fact.lvl_ptr[g.lvl_end];
      const int nh = L.fact.lvl_heavy[g.lvl_begin], nm = L.fact.lvl_mid[g.lvl_begin];
      const int grid = nh + nblk(nm, kGroupsPerBlock) + nblk(hi - lo - nh - nm, 256);
      k_lu_wide<T><<<grid, 256, 0, s>>>(lo, hi, nh, nm, P, lu);
    } else {
      const int nl = g.lvl_end - g.lvl_begin;
      const size_t shm = sizeof(int32_t) * (3 * nl + 1);
      k_lu_narrow<T><<<1, L.narrow_block, shm, s>>>(g.lvl_begin, g.lvl_end, L.fact.lvl_dev.p,
                                                    L.fact.heavy_dev.p, L.fact.mid_dev.p, P, lu);
    }
    ktimer_end(KC_LU_TASKS, s);
  }
}

template <typename T>
void launch_lu_refactor(DevLU& L, const double* csr, T* lu, DevStatus* st, hipStream_t s) {
  const int nnz = (int)L.col.n;
  ktimer_begin(s);
  k_lu_gather<T><<<nblk(nnz, 256), 256, 0, s>>>(nnz, L.src.p, csr, lu, st);
  ktimer_end(KC_LU_GATHER, s);
  run_fact_groups<T>(L, lu, st, s);
  launch_dense_block<T>(L, lu, st, s);
  // band programs scaled L inside their tasks
  const int nl = L.fact_band.on ? 0 : (int)L.l_slot.n;
  if (nl > 0) {
    ktimer_begin(s);
    k_lu_scale<T><<<std::min(nblk(nl, 256), 2048), 256, 0, s>>>(nl, L.l_slot.p, L.l_diag.p, lu, st);
    ktimer_end(KC_LU_SCALE, s);
  }
}

template <typename T, bool FWD>
static void run_tri_groups(DevLU& L, const DevSchedule& S, const T* lu, T* y, T* z, hipStream_t s) {
  const TriProg P{S.items.p, L.rowptr.p, L.col.p, L.diag.p};
  for (const auto& g : S.groups) {
    ktimer_begin(s);
    if (g.wide) {
      const int lo = S.lvl_ptr[g.lvl_begin], hi = S.lvl_ptr[g.lvl_end];
      const int nh = S.lvl_heavy[g.lvl_begin], nm = S.lvl_mid[g.lvl_begin];
      const int grid = nh + nblk(nm, kGroupsPerBlock) + nblk(hi - lo - nh - nm, 256);
      k_tri_wide<T, FWD><<<grid, 256, 0, s>>>(lo, hi, nh, nm, P, lu, y, z);
    } else {
      const int nl = g.lvl_end - g.lvl_begin;
      const size_t shm = sizeof(int32_t) * (3 * nl + 1);
      k_tri_narrow<T, FWD><<<1, L.narrow_block, shm, s>>>(g.lvl_begin, g.lvl_end, S.lvl_dev.p,
                                                          S.heavy_dev.p, S.mid_dev.p, P, lu, y, z);
    }
    ktimer_end(KC_TRI_TASKS, s);
  }
}

template <typename T, bool FWD>
static void run_tri_bands(DevLU& L, const BandSchedule& H, DevBand& D, const T* lu, T* y, T* z,
                          void* part_override, hipStream_t s) {
  const TriProg P{nullptr, L.rowptr.p, L.col.p, L.diag.p};
  const BandProg B{D.items.p, D.tl_ptr.p, D.tl_heavy.p, D.tl_mid.p};
  T* part = reinterpret_cast<T*>(part_override ? part_override : (void*)D.part.p);
  for (size_t b = 0; b + 1 < H.band_task.size(); ++b) {
    const int ntask = H.band_task[b + 1] - H.band_task[b];
    if (ntask > 0) {
      ktimer_begin(s);
      const int nv = H.band_lds.empty() ? 0 : H.band_lds[b];
      const int ne = nv > 0 ? H.band_ent[b] : 0, nr = nv > 0 ? H.band_rows[b] : 0;
      const size_t lde =
          nv > 0 ? lde_col_offset(nv, ne, nr, sizeof(T), FWD) + 2 * (size_t)ne : 0;
      if (nv > 0 && D.use_lds && lde <= (size_t)D.lde_bytes)
        k_tri_band_lde<T, FWD><<<ntask, 256, lde, s>>>(
            H.band_w[b], H.band_ptr_off[b], H.band_hm_off[b], H.band_task[b], nv, ne, nr, B,
            reinterpret_cast<const int4*>(D.lds_rec.p), D.lds_col.p, D.lds_src.p,
            reinterpret_cast<const int4*>(D.lds_task.p), D.lds_slots.p, lu, y, z);
      else
        k_tri_band<T, FWD><<<ntask, 256, 0, s>>>(H.band_w[b], H.band_ptr_off[b],
                                                 H.band_hm_off[b], B, P, lu, y, z);
      ktimer_end(KC_TRI_TASKS, s);
    }
    const int h0 = H.band_hub[b], h1 = H.band_hub[b + 1];
    if (h1 > h0) {
      const int c0 = H.hub_cptr[h0], c1 = H.hub_cptr[h1];
      ktimer_begin(s);
      k_tri_hub_part<T, FWD><<<c1 - c0, 256, 0, s>>>(D.hub_cs.p + c0, D.hub_ce.p + c0, P, lu, y, z,
                                                      part + c0);
      k_tri_hub_finish<T, FWD><<<nblk(h1 - h0, 64), 64, 0, s>>>(h0, h1, D.hub_rows.p,
                                                                 D.hub_cptr.p, part, P, lu, y, z);
      ktimer_end(KC_TRI_HUB, s);
    }
  }
}

template <typename T>
void launch_lu_solve(DevLU& L, const T* lu, const double* rhs, double scale, double* x, T* y, T* z,
                     hipStream_t s, void* const* hub_parts) {
  const int n = L.sym.n;
  if (n == 0) return;
  k_solve_init<T><<<nblk(n, 256), 256, 0, s>>>(n, L.p.p, rhs, scale, y);
  if (L.fwd_band.on)
    run_tri_bands<T, true>(L, L.sym.fwd_band, L.fwd_band, lu, y, z,
                           hub_parts ? hub_parts[0] : nullptr, s);
  else
    run_tri_groups<T, true>(L, L.fwd, lu, y, z, s);
  if (L.bwd_band.on)
    run_tri_bands<T, false>(L, L.sym.bwd_band, L.bwd_band, lu, y, z,
                            hub_parts ? hub_parts[1] : nullptr, s);
  else
    run_tri_groups<T, false>(L, L.bwd, lu, y, z, s);
  k_solve_finish<T><<<nblk(n, 256), 256, 0, s>>>(n, L.q.p, z, x);
}

template void launch_lu_refactor<double>(DevLU&, const double*, double*, DevStatus*, hipStream_t);
template void launch_lu_refactor<float>(DevLU&, const double*, float*, DevStatus*, hipStream_t);
template void launch_lu_solve<double>(DevLU&, const double*, const double*, double, double*,
                                      double*, double*, hipStream_t, void* const*);
template void launch_lu_solve<float>(DevLU&, const float*, const double*, double, double*, float*,
                                     float*, hipStream_t, void* const*);

void launch_spmv(int n, const int32_t* indptr, const int32_t* indices, const double* data,
                 const double* x, const double* base, double sgn, double* out,
                 const HubChunks& hub, hipStream_t s) {
  if (n == 0) return;
  if (hub.n_chunks > 0)
    k_chunk_spmv<<<hub.n_chunks, 256, 0, s>>>(hub.cs, hub.ce, data, indices, x, hub.part);
  k_spmv<<<hub.n_items + nblk(n, 256), 256, 0, s>>>(n, indptr, indices, data, x, base, sgn, out,
                                                    hub);
}

void launch_status_reset(DevStatus* st, hipStream_t s) { k_status_reset<<<1, 64, 0, s>>>(st); }
void launch_iter_begin(int n, const double* X, double* Xsave, int nl, const double* ls,
                       double* ls_save, DevStatus* a, DevStatus* b, DevStatus* c, hipStream_t s) {
  k_iter_begin<<<nblk(std::max(std::max(n, nl), 1), 256), 256, 0, s>>>(n, X, Xsave, nl, ls,
                                                                       ls_save, a, b, c);
}

void launch_axpy(int n, double* x, const double* d, hipStream_t s) {
  if (n == 0) return;
  k_axpy<<<nblk(n, 256), 256, 0, s>>>(n, x, d);
}

void launch_sources(const SourceTable& src, double t, double* out, hipStream_t s) {
  if (src.n == 0) return;
  k_sources<<<nblk(src.n, 64), 64, 0, s>>>(src, t, out);
}

void launch_sources2(const SourceTable& a, double* out_a, const SourceTable& b, double* out_b,
                     double t, hipStream_t s) {
  const int nba = nblk(a.n, 64), nbb = nblk(b.n, 64);
  if (nba + nbb == 0) return;
  k_sources2<<<nba + nbb, 64, 0, s>>>(a, out_a, nba, b, out_b, t);
}

}  // namespace vjx
