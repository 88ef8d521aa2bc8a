"""EKV 2.6 MOSFET in numpy, one row per instance (oracle; test infrastructure).

A restatement of vajax/devices/models/ekv/ekv.va, written top to bottom
in the file's own variable names:

* ``cache(params)`` -- every voltage-independent quantity: the temperature
  update (ekv.va:298-375), the geometry and reverse-short-channel terms
  (:376-388, :414-415) and the junction areas and temperature update
  (:767-788).  $vt(T) is k T / q with the exact SI constants at the device
  temperature T = 300.15 K + TRISE (or TEMP when given), Tnom = 25 C unless
  TNOM is given.
* ``eval_ekv26(vgb, vsb, vdb, k)`` -- the intrinsic model (:390-740), its
  contributions (:741-754) and the junction diodes (:790-847), as the
  openvaf_jax eval contract ``(res_resist [N,4], res_react [N,4],
  jac_resist [N,16], jac_react [N,16])`` with terminals (d, g, s, b).
  Derivatives by forward-mode dual numbers over (V(g,b), V(s,b), V(d,b)):
  the derivative of every branch of the file's code is the derivative that
  OpenVAF takes of it.  Branches are ``np.where`` selections.

Left out, as in the HIP eval: noise, and the series-resistance correction
of Id (:697-700), the identity at RSH = 0 (a nonzero RSH is rejected).

Pinning (tests/test_ekv26_cpu.py): the Jacobian against central finite
differences of the residuals; KCL and charge conservation; source/drain
symmetry; the normalized-charge interpolation against the exact root of
2 q + ln q = x; the long-channel limit against the textbook EKV drain
current and its 40/60 saturation charge partition and Cgg = W L Cox in
strong inversion; the junction diode against its closed form; and the
short-channel terms (LETA roll-off and DIBL, WETA, the Q0 flat-band shift,
E0 mobility reduction, LAMBDA channel-length modulation, UCRIT velocity
saturation, impact ionisation) by the direction each must move the device
in, against the same device with the term off.  Those terms are pinned in
direction, not in value: no closed form fixes their magnitudes.  No golden
vectors for this model exist in the reference tree (its comparison test,
tests/test_osdi_jax_comparison.py:304-320, needs compiled OSDI).
"""

from __future__ import annotations

import numpy as np

P_EPS0 = 8.85418792394420013968e-12
P_CELSIUS0 = 273.15
NOT_GIVEN = 1e21
SIM_TEMP_K = 300.15


def vt_of(T):
    return 1.380649e-23 * T / 1.602176634e-19


def cache(p: dict) -> dict:
    """The file's initial-step quantities per instance, by the file's names."""
    g = {k: np.asarray(v, dtype=np.float64) for k, v in p.items()}
    if np.any(g["rsh"] != 0.0):
        raise ValueError("RSH > 0 is not supported")
    TYPE = g["type"]
    epssil = 11.7 * P_EPS0
    eps_COX = epssil / g["cox"]
    Lc = np.sqrt(eps_COX * g["xj"])
    E0 = g["e0"]
    with np.errstate(divide="ignore"):
        T0 = np.where(E0 == 0.0, 0.0, g["cox"] / (epssil * E0))
    V0 = (g["q0"] + g["q0"]) / g["cox"]
    eta_qi = np.where(TYPE > 0, 0.5, 0.3333333333333)
    T = np.where(g["temp"] == NOT_GIVEN, SIM_TEMP_K + g["trise"], g["temp"] + P_CELSIUS0)
    Tnom = np.where(g["tnom"] == NOT_GIVEN, 25.0 + P_CELSIUS0, g["tnom"] + P_CELSIUS0)
    Vt = vt_of(T)
    Eg = 1.16 - 7.02e-4 * T * T / (T + 1108.0)
    refEg = 1.16 - (7.02e-4 * Tnom * Tnom) / (Tnom + 1108.0)
    deltaT = T - Tnom
    ratioT = T / Tnom
    VTO_T = g["vto"] - g["tcv"] * deltaT
    KP_T = g["kp"] * np.power(ratioT, g["bex"])
    UCRIT_T = g["ucrit"] * np.power(ratioT, g["ucex"])
    IBB_T = g["ibb"] * (1.0 + g["ibbt"] * deltaT)
    PHI_T = g["phi"] * ratioT - 3.0 * Vt * np.log(ratioT) - refEg * ratioT + Eg
    tmp2 = PHI_T - 0.2
    PHI_T = 0.5 * (tmp2 + np.sqrt(tmp2 * tmp2 + Vt * Vt)) + 0.2
    Leff = g["l"] + g["dl"]
    Weff = g["w"] + g["dw"]
    Vc = UCRIT_T * Leff
    AWL = 1.0 / np.sqrt(Weff * Leff)
    AVTO, AKP, AGAMMA = g["avto"], g["akp"], g["agamma"]
    VTO_S = np.where(TYPE > 0, np.where(AVTO != 1e-6, AWL * (AVTO - 1e-6) + VTO_T, VTO_T),
                     np.where(AVTO != 1e-6, AWL * (1e-6 - AVTO) - VTO_T, -VTO_T))
    KP_Weff = Weff * np.where(AKP != 1e-6, KP_T * (1 + (AKP - 1e-6) * AWL), KP_T)
    GAMMA_S = np.where(AGAMMA != 1e-6, g["gamma"] + (AGAMMA - 1e-6) * AWL, g["gamma"])
    vL = 0.28 * (Leff / (g["lk"] * g["ns"]) - 0.1)
    sqv = 1.0 / (1.0 + 0.5 * (vL + np.sqrt(vL * vL + 1.936e-3)))
    deltaVFB = np.where(V0 == 0.0, 0.0, V0 * sqv * sqv)
    HDIF = g["hdif"]
    as_i = np.where((g["as"] == 0.0) & (HDIF > 0.0), 2.0 * HDIF * Weff, g["as"])
    ps_i = np.where((g["ps"] == 0.0) & (HDIF > 0.0), 4.0 * HDIF + 1.0 * Weff, g["ps"])
    ad_i = np.where((g["ad"] == 0.0) & (HDIF > 0.0), 2.0 * HDIF * Weff, g["ad"])
    pd_i = np.where((g["pd"] == 0.0) & (HDIF > 0.0), 4.0 * HDIF + 1.0 * Weff, g["pd"])
    temp_arg = np.exp((refEg / vt_of(Tnom) - Eg / Vt + g["tp_xti"] * np.log(ratioT)) / g["xd_n"])
    return dict(
        TYPE=TYPE, M=g["m"], VTO_S=VTO_S, PHI_T=PHI_T, GAMMA_S=GAMMA_S, KP_Weff=KP_Weff,
        Leff=Leff, Weff=Weff, UCRIT_T=UCRIT_T, LAMBDA=g["lambda"], Lc=Lc,
        WETA_W=3.0 * eps_COX * g["weta"] * g["m"] / Weff,
        LETA_L=eps_COX * g["leta"] * g["ns"] / Leff, deltaVFB=deltaVFB, E0=E0,
        THETA=g["theta"], T0=T0, eta_qi=eta_qi, IBN=g["ibn"], IBA_IBB=g["iba"] / IBB_T,
        Lc_IBB=Lc * IBB_T, WLCox=Weff * Leff * g["cox"], Vc=Vc,
        log_Vc_Vt=Vt * (np.log(0.5 * Vc / Vt) - 0.6), Vt=Vt, ratioT=ratioT,
        deltaT=deltaT, as_i=as_i, ps_i=ps_i, ad_i=ad_i, pd_i=pd_i,
        js_t=g["xd_js"] * temp_arg, jsw_t=g["xd_jsw"] * temp_arg,
        jswg_t=g["xd_jswg"] * temp_arg,
        pb_t=g["xd_pb"] - g["tp_pb"] * deltaT, pbsw_t=g["xd_pbsw"] - g["tp_pbsw"] * deltaT,
        pbswg_t=g["xd_pbswg"] - g["tp_pbswg"] * deltaT,
        cj_t=g["xd_cj"] * (1.0 + g["tp_cj"] * deltaT),
        cjsw_t=g["xd_cjsw"] * (1.0 + g["tp_cjsw"] * deltaT),
        cjswg_t=g["xd_cjswg"] * (1.0 + g["tp_cjswg"] * deltaT),
        njts_t=g["xd_njts"] * (1.0 + (ratioT - 1.0) * g["tp_njts"]),
        njtssw_t=g["xd_njtssw"] * (1.0 + (ratioT - 1.0) * g["tp_njtssw"]),
        njtsswg_t=g["xd_njtsswg"] * (1.0 + (ratioT - 1.0) * g["tp_njtsswg"]),
        xd_n=g["xd_n"], xd_bv=g["xd_bv"], xd_xjbv=g["xd_xjbv"], xd_gmin=g["xd_gmin"],
        xd_vts=g["xd_vts"], xd_vtssw=g["xd_vtssw"], xd_vtsswg=g["xd_vtsswg"],
        xd_mj=g["xd_mj"], xd_mjsw=g["xd_mjsw"], xd_mjswg=g["xd_mjswg"],
    )


class D:
    """Dual number array: value ``v`` [N] and partials ``g`` [3, N] over
    (V(g,b), V(s,b), V(d,b))."""

    __slots__ = ("v", "g")
    __array_ufunc__ = None  # ndarray (op) D defers to D's reflected operators

    def __init__(self, v, g=None):
        self.v = np.asarray(v, dtype=np.float64)
        self.g = np.zeros((3,) + self.v.shape) if g is None else g

    @staticmethod
    def _c(x):
        return x if isinstance(x, D) else D(np.broadcast_to(np.asarray(x, np.float64), ()))

    def __add__(self, o):
        if isinstance(o, D):
            return D(self.v + o.v, self.g + o.g)
        return D(self.v + o, self.g)

    __radd__ = __add__

    def __sub__(self, o):
        if isinstance(o, D):
            return D(self.v - o.v, self.g - o.g)
        return D(self.v - o, self.g)

    def __rsub__(self, o):
        return D(o - self.v, -self.g)

    def __neg__(self):
        return D(-self.v, -self.g)

    def __mul__(self, o):
        if isinstance(o, D):
            return D(self.v * o.v, self.g * o.v + self.v * o.g)
        return D(o * self.v, o * self.g)

    __rmul__ = __mul__

    def __truediv__(self, o):
        if isinstance(o, D):
            q = self.v / o.v
            return D(q, (self.g - q * o.g) / o.v)
        return D(self.v / o, self.g / o)

    def __rtruediv__(self, o):
        return D(np.broadcast_to(np.asarray(o, np.float64), self.v.shape).copy()) / self


def dsqrt(a: D) -> D:
    r = np.sqrt(a.v)
    return D(r, (0.5 / r) * a.g)


def dlog(a: D) -> D:
    return D(np.log(a.v), (1.0 / a.v) * a.g)


def dexp(a: D) -> D:
    e = np.exp(a.v)
    return D(e, e * a.g)


def where(m, a: D, b: D) -> D:
    a, b = D._c(a), D._c(b)
    return D(np.where(m, a.v, b.v), np.where(m, a.g, b.g))


def const(x, like: D) -> D:
    return D(np.broadcast_to(np.asarray(x, np.float64), like.v.shape).copy())


def interp_q(x: D) -> D:
    """The file's normalized charge yk for x = (VP - V)/Vt (ekv.va:428-444)."""
    with np.errstate(all="ignore"):
        z0a = 2.0 / (1.3 + x - dlog(x + 1.6))
        zka = (2.0 + z0a) / (1.0 + x + dlog(z0a))
        ya = (1.0 + x + dlog(zka)) / (2.0 + zka)
        z0b = 1.55 + dexp(-x)
        zkb = (2.0 + z0b) / (1.0 + x + dlog(z0b))
        yb = (1.0 + x + dlog(zkb)) / (2.0 + zkb)
        yc = 1.0 / (2.0 + dexp(-x))
        yd = dexp(x) + 1e-64
    return where(x.v > -0.35, ya, where(x.v > -15.0, yb, where(x.v > -23.0, yc, yd)))


def _junction(v: D, is_x, ad, pd, k):
    """Drain- or source-bulk diode current and charge (ekv.va:790-847)
    without the TYPE*M factor; ``ad``/``pd`` are that side's area/perimeter."""
    ratioT, Vt, xd_n = k["ratioT"], k["Vt"], k["xd_n"]
    with np.errstate(all="ignore"):
        arg = -v * ratioT / (Vt * xd_n)
        arg = where(arg.v < -40.0, const(-40.0, v), arg)
        tmp0 = (-v + k["xd_bv"]) * ratioT / (Vt * xd_n)
        fbd = where(tmp0.v > 70, const(1.0, v), 1.0 + k["xd_xjbv"] * dexp(-tmp0))

        def tun(coef, nj, vts):
            den = where(vts + v.v > 1.0e-3, vts + v, const(1.0e-3, v))
            return coef * (dexp(v * ratioT / (Vt * nj) * vts / den) - 1.0)

        t = -tun(k["Weff"] * k["jswg_t"], k["njtsswg_t"], k["xd_vtsswg"])
        t = t - tun(pd * k["jsw_t"], k["njtssw_t"], k["xd_vtssw"])
        t = t - tun(ad * k["js_t"], k["njts_t"], k["xd_vts"])
        cur = is_x * (1.0 - dexp(arg)) * fbd + v * k["xd_gmin"] + t
        pos = v.v > 0.0
        cp = (k["cj_t"] * ad * dexp(-k["xd_mj"] * dlog(1.0 + v / k["pb_t"]))
              + k["cjsw_t"] * pd * dexp(-k["xd_mjsw"] * dlog(1.0 + v / k["pbsw_t"]))
              + k["cjswg_t"] * k["Weff"] * dexp(-k["xd_mjswg"] * dlog(1.0 + v / k["pbswg_t"])))
        cn = (k["cj_t"] * ad * (1.0 - k["xd_mj"] * v / k["pb_t"])
              + k["cjsw_t"] * pd * (1.0 - k["xd_mjsw"] * v / k["pbsw_t"])
              + k["cjswg_t"] * k["Weff"] * (1.0 - k["xd_mjswg"] * v / k["pbswg_t"]))
    q = where(pos, cp, cn) * v
    return cur, q


def intrinsic(VG: D, VS: D, VD: D, k: dict) -> dict:
    """ekv.va:403-740 in the source/drain-ordered frame (VD >= VS)."""
    Vt = k["Vt"]
    inv_Vt = 1.0 / Vt
    Vt_01, Vt_4 = 0.1 * Vt, 4.0 * Vt
    Vt_Vt_2, Vt_Vt_16 = 2.0 * Vt * Vt, 16.0 * Vt * Vt
    PHI_T, GAMMA_S = k["PHI_T"], k["GAMMA_S"]
    GAMMA_sqrt_PHI = GAMMA_S * np.sqrt(PHI_T)
    LAMBDA, Lc, UCRIT_T, Vc = k["LAMBDA"], k["Lc"], k["UCRIT_T"], k["Vc"]
    VGstar = VG - k["VTO_S"] - k["deltaVFB"] + PHI_T + GAMMA_sqrt_PHI
    VGprime = 0.5 * (VGstar + dsqrt(VGstar * VGstar + 2.0 * Vt_Vt_16))
    PHI_VS = PHI_T + VS
    sqrt_PHI_VS = dsqrt(0.5 * (PHI_VS + dsqrt(PHI_VS * PHI_VS + Vt_Vt_16)))
    PHI_VD = PHI_T + VD
    sqrt_PHI_VD = dsqrt(0.5 * (PHI_VD + dsqrt(PHI_VD * PHI_VD + Vt_Vt_16)))
    big_sqrt_VP0 = dsqrt(VGprime + 0.25 * GAMMA_S * GAMMA_S)
    VP0 = VGprime - PHI_T - GAMMA_S * (big_sqrt_VP0 - 0.5 * GAMMA_S)
    sqrt_PHI_VP0 = dsqrt(VP0 + PHI_T + Vt_01)
    GAMMAstar = GAMMA_S - k["LETA_L"] * (sqrt_PHI_VS + sqrt_PHI_VD) + k["WETA_W"] * sqrt_PHI_VP0
    GAMMAprime = 0.5 * (GAMMAstar + dsqrt(GAMMAstar * GAMMAstar + Vt_01))
    big_sqrt_VP = dsqrt(VGprime + 0.25 * GAMMAprime * GAMMAprime)
    VP = VGprime - PHI_T - GAMMAprime * (big_sqrt_VP - 0.5 * GAMMAprime)
    yk = interp_q((VP - VS) * inv_Vt)
    if_ = yk * (1.0 + yk)
    sqrt_if = dsqrt(if_)
    Vt_Vc = Vt / Vc
    VDSS = Vc * (dsqrt(0.25 + sqrt_if * Vt_Vc) - 0.5)
    Vds = 0.5 * (VD - VS)
    deltaV_2 = Vt_Vt_16 * (LAMBDA * (sqrt_if - VDSS * inv_Vt) + 15.625e-3)
    sqrt_VDSS_deltaV = dsqrt(VDSS * VDSS + deltaV_2)
    sqrt_Vds_VDSS_deltaV = dsqrt((Vds - VDSS) * (Vds - VDSS) + deltaV_2)
    Vip = sqrt_VDSS_deltaV - sqrt_Vds_VDSS_deltaV
    VDSSprime = Vc * (dsqrt(0.25 + (sqrt_if - 0.75 * dlog(if_)) * Vt_Vc) - 0.5) + k["log_Vc_Vt"]
    Vdsprime = Vds - VDSSprime
    sqrt_VDSSprime_deltaV = dsqrt(VDSSprime * VDSSprime + deltaV_2)
    sqrt_Vds_VDSSprime_deltaV = dsqrt(Vdsprime * Vdsprime + deltaV_2)
    yk = interp_q((VP - Vds - VS - sqrt_VDSSprime_deltaV + sqrt_Vds_VDSSprime_deltaV) * inv_Vt)
    irprime = yk * (1.0 + yk)
    deltaL = (Lc * LAMBDA) * dlog(1.0 + (Vds - Vip) / (Lc * UCRIT_T))
    Lprime = k["Leff"] - deltaL + (Vds + Vip) * (1.0 / UCRIT_T)
    Lmin = 0.1 * k["Leff"]
    Leq = 0.5 * (Lprime + dsqrt(Lprime * Lprime + Lmin * Lmin))
    yk = interp_q((VP - VD) * inv_Vt)
    ir = yk * (1.0 + yk)
    sif2 = 0.25 + if_
    sir2 = 0.25 + ir
    sif = dsqrt(sif2)
    sir = dsqrt(sir2)
    sif_sir_2 = (sif + sir) * (sif + sir)
    VP_PHI_eps = VP + PHI_T + 1.0e-6
    sqrt_PHI_VP_2 = 2.0 * dsqrt(VP_PHI_eps)
    n_1 = GAMMA_S / sqrt_PHI_VP_2
    n_1_n = GAMMA_S / (sqrt_PHI_VP_2 + GAMMA_S)
    qi = -((1.0 + n_1) * Vt * ((0.66666666 + 0.66666666) * (sir2 + sir * sif + sif2) / (sif + sir)
                              - 1.0))
    qb = -0.5 * GAMMA_S * sqrt_PHI_VP_2 - n_1_n * qi
    E0, T0 = k["E0"], k["T0"]
    VPprime = 0.5 * (VP + dsqrt(VP * VP + Vt_Vt_2))
    beta_theta = k["KP_Weff"] / (Leq * (1.0 + k["THETA"] * VPprime))
    e = qb + k["eta_qi"] * qi
    E0_Q_1 = where(e.v > 0.0, 1.0 + T0 * e, 1.0 - T0 * e)
    beta_e0 = k["KP_Weff"] * (1.0 + T0 * GAMMA_sqrt_PHI) / (Leq * E0_Q_1)
    beta = where(E0 == 0.0, beta_theta, beta_e0)
    n = 1.0 + GAMMA_S / (2.0 * dsqrt(PHI_T + VP + Vt_4))
    Ispec = Vt_Vt_2 * n * beta
    Id = Ispec * (if_ - irprime)
    Vib = VD - VS - (k["IBN"] + k["IBN"]) * VDSS
    on = (Vib.v > 0.0) & (k["IBA_IBB"] > 0.0)
    with np.errstate(all="ignore"):
        Vib_s = where(on, Vib, const(1.0, Vib))
        x = -k["Lc_IBB"] / Vib_s
        x = where(x.v < -35.0, const(-35.0, x), x)
        Isub = where(on, k["IBA_IBB"] * Vib_s * dexp(x) * Id, const(0.0, Vib))
    WLCox = k["WLCox"]
    sif3 = sif * sif2
    sir3 = sir * sir2
    tmp1 = dsqrt(PHI_T + 0.5 * VP)
    sqrt_PHI_VP2_2 = tmp1 + tmp1
    n_Vt_COX = (1.0 + GAMMAprime / sqrt_PHI_VP2_2) * (Vt * WLCox)
    QD = -(n_Vt_COX * (0.266666666 * (3.0 * sir3 + 6.0 * sir2 * sif + 4.0 * sir * sif2
                                      + 2.0 * sif3) / sif_sir_2 - 0.5))
    QS = -(n_Vt_COX * (0.266666666 * (3.0 * sif3 + 6.0 * sif2 * sir + 4.0 * sif * sir2
                                      + 2.0 * sir3) / sif_sir_2 - 0.5))
    QI = QS + QD
    QB = WLCox * (-0.5 * GAMMAprime * sqrt_PHI_VP_2 + VGprime - VGstar) \
        - QI * GAMMAprime / (GAMMAprime + sqrt_PHI_VP2_2)
    QG = -QI - QB
    return dict(Id=Id, Isub=Isub, QD=QD, QS=QS, QG=QG, QB=QB, VP=VP, n=n, if_=if_, ir=ir,
                irprime=irprime, beta=beta, VDSS=VDSS)


def _cols(x: D):
    """(d, g, s, b) Jacobian columns of one output from its (g, s, d) partials."""
    gg, gs, gd = x.g[0], x.g[1], x.g[2]
    return [gd, gg, gs, -(gg + gs + gd)]


def eval_ekv26(vgb, vsb, vdb, k: dict):
    TYPE = k["TYPE"]
    N = np.shape(vgb)[0]
    z = np.zeros(N)
    e0, e1, e2 = np.array([1.0, 0, 0])[:, None], np.array([0, 1.0, 0])[:, None], \
        np.array([0, 0, 1.0])[:, None]
    vg = D(TYPE * vgb, e0 * TYPE)
    vs = D(TYPE * vsb, e1 * TYPE)
    vd = D(TYPE * vdb, e2 * TYPE)
    fwd = ~(vd.v - vs.v < 0)
    VS = where(fwd, vs, vd)
    VD = where(fwd, vd, vs)
    Mode = np.where(fwd, 1.0, -1.0)
    with np.errstate(all="ignore"):
        r = intrinsic(vg, VS, VD, k)
    ijd, qjd = _junction(vd, k["js_t"] * k["ad_i"] + k["jsw_t"] * k["pd_i"]
                         + k["jswg_t"] * k["Weff"], k["ad_i"], k["pd_i"], k)
    ijs, qjs = _junction(vs, k["js_t"] * k["as_i"] + k["jsw_t"] * k["ps_i"]
                         + k["jswg_t"] * k["Weff"], k["as_i"], k["ps_i"], k)
    TM = TYPE * k["M"]
    ids = (TYPE * Mode) * r["Id"]
    isub = TYPE * r["Isub"]
    zero = const(0.0, ids)
    rd = where(fwd, ids + isub, ids) + TM * ijd
    rs = where(fwd, -ids, isub - ids) + TM * ijs
    rb = -(isub + TM * (ijd + ijs))
    qd = TYPE * where(fwd, r["QD"], r["QS"]) + TM * qjd
    qs = TYPE * where(fwd, r["QS"], r["QD"]) + TM * qjs
    qg = TYPE * r["QG"]
    qbn = -(qd + qs + qg)
    res_r = np.stack([rd.v, z, rs.v, rb.v], axis=1)
    res_c = np.stack([qd.v, qg.v, qs.v, qbn.v], axis=1)
    jac_r = np.stack(sum((_cols(x) for x in (rd, zero, rs, rb)), []), axis=1)
    jac_c = np.stack(sum((_cols(x) for x in (qd, qg, qs, qbn)), []), axis=1)
    return res_r, res_c, jac_r, jac_c
