"""Per-model parameter columns handed to the HIP eval kernels.

Each model's kernel reads a structure-of-arrays parameter block
``[n_par][count]`` (float64).  For the resistor and capacitor this is the
instance parameter itself (vajax/devices/models/resistor.va:19,
capacitor.va:17); for the MOSFET it is the per-instance init cache computed
once at setup, the analogue of the cache that the reference's
``prepare_static_inputs`` builds from the model's init function
(vajax/analysis/openvaf_models.py:889-921):

    sgn   = type                         (+1 NMOS, -1 PMOS)
    vgoff = (phi - vto) + gamma*sqrt(phi)   so that VG' = VG + vgoff
    phi, gamma, beta0 = kp*(w/l), theta
    cgs = cgd = cox*w*l/2 + cov*w,  cgb = 0,  cdb = csb = cj*w

and for the diode its init constants is, nvt = n*vt, the pnjlim critical
voltage vcrit = nvt*ln(nvt/(sqrt(2)*is)), and cj.  The EKV 2.6 model
(``ekv26``) reads the 52-column cache of :func:`ekv26_cache`.
"""

from __future__ import annotations

from typing import Dict

import numpy as np

MOSFET_CACHE_COLUMNS = ("sgn", "vgoff", "phi", "gamma", "beta0", "theta",
                        "cgs", "cgd", "cgb", "cdb", "csb")


def mosfet_cache(p: Dict[str, np.ndarray]) -> np.ndarray:
    w, l = p["w"], p["l"]
    cgs = 0.5 * p["cox"] * w * l + p["cov"] * w
    cols = {
        "sgn": p["type"],
        "vgoff": (p["phi"] - p["vto"]) + p["gamma"] * np.sqrt(p["phi"]),
        "phi": p["phi"],
        "gamma": p["gamma"],
        "beta0": p["kp"] * (w / l),
        "theta": p["theta"],
        "cgs": cgs,
        "cgd": cgs,
        "cgb": np.zeros_like(w),
        "cdb": p["cj"] * w,
        "csb": p["cj"] * w,
    }
    return np.ascontiguousarray(np.stack([np.asarray(cols[c], dtype=np.float64)
                                          for c in MOSFET_CACHE_COLUMNS]))


# thermal voltage at the reference's default 27 C (k T / q, exact SI constants)
VT = 1.380649e-23 * 300.15 / 1.602176634e-19
DIODE_COLUMNS = ("is", "nvt", "vcrit", "cj")


def diode_cache(p: Dict[str, np.ndarray]) -> np.ndarray:
    nvt = p["n"] * VT
    cols = {"is": p["is"], "nvt": nvt, "vcrit": nvt * np.log(nvt / (np.sqrt(2.0) * p["is"])),
            "cj": p["cj"]}
    return np.ascontiguousarray(np.stack([np.asarray(cols[c], dtype=np.float64)
                                          for c in DIODE_COLUMNS]))


# EKV 2.6 instance cache: every voltage-independent term of
# vajax/devices/models/ekv/ekv.va (temperature update :298-375, geometry and
# reverse-short-channel terms :376-388, junction areas and temperature
# update :767-788), in the column order the eval kernel reads.
EKV26_CACHE_COLUMNS = (
    "type", "vto_s", "phi_t", "gamma_s", "kp_weff", "leff", "ucrit_t", "lambda", "lc",
    "weta_w", "leta_l", "dvfb", "e0", "theta", "t0", "eta_qi", "ibn", "iba_ibb", "lc_ibb",
    "wlcox", "vc", "log_vc_vt", "vt",
    "m", "jk", "is_d", "is_s", "bv", "xjbv", "gmin", "tun_w", "tun_pd", "tun_ad", "tun_ps",
    "tun_as", "kg", "ksw", "ka", "vtsswg", "vtssw", "vts", "cja_d", "cjp_d", "cjg", "cja_s",
    "cjp_s", "mj", "mjsw", "mjswg", "pb", "pbsw", "pbswg",
)
# physical constants of ekv.va:93-98 (P_EPS0, P_CELSIUS0); $vt(T) = k T / q
# with the exact SI constants
EPS0 = 8.85418792394420013968e-12
CELSIUS0 = 273.15
SIM_TEMP_K = 300.15
NOT_GIVEN = 1e21


def _vt(t):
    return 1.380649e-23 * t / 1.602176634e-19


def ekv26_cache(p: Dict[str, np.ndarray]) -> np.ndarray:
    """The EKV 2.6 init cache, one column per EKV26_CACHE_COLUMNS entry."""
    f = {k: np.asarray(v, dtype=np.float64) for k, v in p.items()}
    if np.any(f["rsh"] != 0.0):
        raise ValueError("ekv26: RSH > 0 (series resistance correction) is not supported")
    typ = f["type"]
    epssil = 11.7 * EPS0
    eps_cox = epssil / f["cox"]
    lc = np.sqrt(eps_cox * f["xj"])
    e0 = f["e0"]
    t0 = np.where(e0 == 0.0, 0.0, f["cox"] / (epssil * np.where(e0 == 0.0, 1.0, e0)))
    v0 = (f["q0"] + f["q0"]) / f["cox"]
    eta_qi = np.where(typ > 0, 0.5, 0.3333333333333)
    T = np.where(f["temp"] == NOT_GIVEN, SIM_TEMP_K + f["trise"], f["temp"] + CELSIUS0)
    tnom = np.where(f["tnom"] == NOT_GIVEN, 25.0 + CELSIUS0, f["tnom"] + CELSIUS0)
    vt = _vt(T)
    eg = 1.16 - 7.02e-4 * T * T / (T + 1108.0)
    ref_eg = 1.16 - (7.02e-4 * tnom * tnom) / (tnom + 1108.0)
    dT = T - tnom
    ratio = T / tnom
    vto_t = f["vto"] - f["tcv"] * dT
    kp_t = f["kp"] * ratio ** f["bex"]
    ucrit_t = f["ucrit"] * ratio ** f["ucex"]
    ibb_t = f["ibb"] * (1.0 + f["ibbt"] * dT)
    phi_t = f["phi"] * ratio - 3.0 * vt * np.log(ratio) - ref_eg * ratio + eg
    t2 = phi_t - 0.2
    phi_t = 0.5 * (t2 + np.sqrt(t2 * t2 + vt * vt)) + 0.2
    leff = f["l"] + f["dl"]
    weff = f["w"] + f["dw"]
    vc = ucrit_t * leff
    awl = 1.0 / np.sqrt(weff * leff)
    avto, akp, agam = f["avto"], f["akp"], f["agamma"]
    vto_s = np.where(typ > 0,
                     np.where(avto != 1e-6, awl * (avto - 1e-6) + vto_t, vto_t),
                     np.where(avto != 1e-6, awl * (1e-6 - avto) - vto_t, -vto_t))
    kp_weff = weff * np.where(akp != 1e-6, kp_t * (1 + (akp - 1e-6) * awl), kp_t)
    gamma_s = np.where(agam != 1e-6, f["gamma"] + (agam - 1e-6) * awl, f["gamma"])
    vl = 0.28 * (leff / (f["lk"] * f["ns"]) - 0.1)
    sqv = 1.0 / (1.0 + 0.5 * (vl + np.sqrt(vl * vl + 1.936e-3)))
    dvfb = np.where(v0 == 0.0, 0.0, v0 * sqv * sqv)
    # junctions
    hd = f["hdif"]
    as_i = np.where((f["as"] == 0.0) & (hd > 0.0), 2.0 * hd * weff, f["as"])
    ps_i = np.where((f["ps"] == 0.0) & (hd > 0.0), 4.0 * hd + 1.0 * weff, f["ps"])
    ad_i = np.where((f["ad"] == 0.0) & (hd > 0.0), 2.0 * hd * weff, f["ad"])
    pd_i = np.where((f["pd"] == 0.0) & (hd > 0.0), 4.0 * hd + 1.0 * weff, f["pd"])
    targ = np.exp((ref_eg / _vt(tnom) - eg / vt + f["tp_xti"] * np.log(ratio)) / f["xd_n"])
    js, jsw, jswg = f["xd_js"] * targ, f["xd_jsw"] * targ, f["xd_jswg"] * targ
    cj = f["xd_cj"] * (1.0 + f["tp_cj"] * dT)
    cjsw = f["xd_cjsw"] * (1.0 + f["tp_cjsw"] * dT)
    cjswg = f["xd_cjswg"] * (1.0 + f["tp_cjswg"] * dT)
    njts = f["xd_njts"] * (1.0 + (ratio - 1.0) * f["tp_njts"])
    njtssw = f["xd_njtssw"] * (1.0 + (ratio - 1.0) * f["tp_njtssw"])
    njtsswg = f["xd_njtsswg"] * (1.0 + (ratio - 1.0) * f["tp_njtsswg"])
    cols = {
        "type": typ, "vto_s": vto_s, "phi_t": phi_t, "gamma_s": gamma_s, "kp_weff": kp_weff,
        "leff": leff, "ucrit_t": ucrit_t, "lambda": f["lambda"], "lc": lc,
        "weta_w": 3.0 * eps_cox * f["weta"] * f["m"] / weff,
        "leta_l": eps_cox * f["leta"] * f["ns"] / leff, "dvfb": dvfb, "e0": e0,
        "theta": f["theta"], "t0": t0, "eta_qi": eta_qi, "ibn": f["ibn"],
        "iba_ibb": f["iba"] / ibb_t, "lc_ibb": lc * ibb_t, "wlcox": weff * leff * f["cox"],
        "vc": vc, "log_vc_vt": vt * (np.log(0.5 * vc / vt) - 0.6), "vt": vt,
        "m": f["m"], "jk": ratio / (vt * f["xd_n"]),
        "is_d": js * ad_i + jsw * pd_i + jswg * weff,
        "is_s": js * as_i + jsw * ps_i + jswg * weff,
        "bv": f["xd_bv"], "xjbv": f["xd_xjbv"], "gmin": f["xd_gmin"],
        "tun_w": weff * jswg, "tun_pd": pd_i * jsw, "tun_ad": ad_i * js,
        "tun_ps": ps_i * jsw, "tun_as": as_i * js,
        "kg": ratio / (vt * njtsswg), "ksw": ratio / (vt * njtssw), "ka": ratio / (vt * njts),
        "vtsswg": f["xd_vtsswg"], "vtssw": f["xd_vtssw"], "vts": f["xd_vts"],
        "cja_d": cj * ad_i, "cjp_d": cjsw * pd_i, "cjg": cjswg * weff,
        "cja_s": cj * as_i, "cjp_s": cjsw * ps_i,
        "mj": f["xd_mj"], "mjsw": f["xd_mjsw"], "mjswg": f["xd_mjswg"],
        "pb": f["xd_pb"] - f["tp_pb"] * dT, "pbsw": f["xd_pbsw"] - f["tp_pbsw"] * dT,
        "pbswg": f["xd_pbswg"] - f["tp_pbswg"] * dT,
    }
    n = len(typ)
    return np.ascontiguousarray(np.stack([np.broadcast_to(np.asarray(cols[c], np.float64), (n,))
                                          for c in EKV26_CACHE_COLUMNS]))


def param_block(model: str, params: Dict[str, np.ndarray]) -> np.ndarray:
    if model == "resistor":
        return np.ascontiguousarray(params["r"][None, :], dtype=np.float64)
    if model == "capacitor":
        return np.ascontiguousarray(params["c"][None, :], dtype=np.float64)
    if model == "mosfet":
        return mosfet_cache(params)
    if model == "diode":
        return diode_cache(params)
    if model == "ekv26":
        return ekv26_cache(params)
    raise ValueError(f"no eval kernel for model {model!r}")
