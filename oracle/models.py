"""Device evaluation in numpy, one row per instance (oracle).

Each ``eval_*`` returns the openvaf_jax eval contract restricted to what these
models produce: ``(res_resist [N,R], res_react [N,R], jac_resist [N,J],
jac_react [N,J])``; the diode, the one model with a ``$limit``, also returns
``lim_rhs_resist [N,R], lim_rhs_react [N,R], limit_state_out [N,1]``
(openvaf_jax/codegen/function_builder.py:967-996); for the others lim_rhs,
small-signal and limit-state outputs are zero.  Residual node order is the
terminal order; Jacobian keys are all (row, col) terminal pairs, row-major.

* resistor: ``I(A,B) <+ V(A,B)/r`` (vajax/devices/models/resistor.va:23-27).
* capacitor: ``I(A,B) <+ ddt(c*V(A,B))`` (vajax/devices/models/capacitor.va:20-23).
* mosfet: EKV 2.6 long-channel drain current
  (vajax/devices/models/ekv_longchannel/ekv_longchannel.va:58-96) evaluated in
  the NMOS frame (voltages and current multiplied by the polarity ``type``),
  with ln(1+exp(u)) in the overflow-safe form max(u,0) + log1p(exp(-|u|)),
  analytic derivatives, and linear terminal charges
  Q_g = cgs*Vgs + cgd*Vgd + cgb*Vgb, Q_d = -cgd*Vgd + cdb*Vdb,
  Q_s = -cgs*Vgs + csb*Vsb, Q_b = -(cgb*Vgb + cdb*Vdb + csb*Vsb)
  with cgs = cgd = cox*w*l/2 + cov*w, cdb = csb = cj*w, cgb = 0.
  PSP103 is not available, so MOSFET arithmetic against the reference's model
  is parity unpinned; the formula itself is pinned by closed-form and
  finite-difference tests.
* diode: the junction diode as a Verilog-A model with ``$limit`` compiles
  under openvaf_jax (instruction.py:628-660, 705-870;
  function_builder.py:1126-1212):
  vd = V(A,K); vdl = $limit(vd, "pnjlim", n*vt, vcrit) with
  vcrit = n*vt*ln(n*vt/(sqrt(2)*is)) and vold = limit_state_in;
  I(A,K) <+ is*(exp(vdl/(n*vt)) - 1) + ddt(cj*vdl); limit_state_out = vdl;
  lim_rhs[i] = J[i, A] * (vdl - vd) for the resistive and reactive parts.
"""

from __future__ import annotations

import numpy as np

# Boltzmann constant / elementary charge (exact SI values) at 300.15 K, the
# reference's default simulation temperature (27 C).
KB = 1.380649e-23
QE = 1.602176634e-19
TEMP_K = 300.15
VT = KB * TEMP_K / QE


def eval_resistor(v_ab: np.ndarray, r: np.ndarray):
    i = v_ab / r
    g = 1.0 / r
    res_r = np.stack([i, -i], axis=1)
    res_c = np.zeros_like(res_r)
    jac_r = np.stack([g, -g, -g, g], axis=1)
    jac_c = np.zeros_like(jac_r)
    return res_r, res_c, jac_r, jac_c


def eval_capacitor(v_ab: np.ndarray, c: np.ndarray):
    q = c * v_ab
    res_c = np.stack([q, -q], axis=1)
    res_r = np.zeros_like(res_c)
    jac_c = np.stack([c, -c, -c, c], axis=1)
    jac_r = np.zeros_like(jac_c)
    return res_r, res_c, jac_r, jac_c


def diode_cache(p: dict) -> dict:
    """Per-instance constants of the diode (its init cache)."""
    nvt = p["n"] * VT
    return {"is": np.asarray(p["is"], dtype=np.float64), "nvt": nvt,
            "vcrit": nvt * np.log(nvt / (np.sqrt(2.0) * p["is"])),
            "cj": np.asarray(p["cj"], dtype=np.float64)}


def eval_diode(vd: np.ndarray, vold: np.ndarray, k: dict):
    from .limiting import pnjlim

    vdl = pnjlim(vd, vold, k["nvt"], k["vcrit"])
    e = np.exp(vdl / k["nvt"])
    i = k["is"] * (e - 1.0)
    g = k["is"] * e / k["nvt"]
    q = k["cj"] * vdl
    cj = k["cj"]
    dl = vdl - vd
    res_r = np.stack([i, -i], axis=1)
    res_c = np.stack([q, -q], axis=1)
    jac_r = np.stack([g, -g, -g, g], axis=1)
    jac_c = np.stack([cj, -cj, -cj, cj], axis=1)
    lim_r = np.stack([g * dl, -(g * dl)], axis=1)
    lim_c = np.stack([cj * dl, -(cj * dl)], axis=1)
    return res_r, res_c, jac_r, jac_c, lim_r, lim_c, vdl[:, None]


def mosfet_cache(p: dict) -> dict:
    """Per-instance constants computed once at setup (the model's init cache)."""
    w, l = p["w"], p["l"]
    cgs = 0.5 * p["cox"] * w * l + p["cov"] * w
    return {
        "sgn": p["type"].astype(np.float64),
        "vgoff": (p["phi"] - p["vto"]) + p["gamma"] * np.sqrt(p["phi"]),
        "phi": p["phi"].astype(np.float64),
        "gamma": p["gamma"].astype(np.float64),
        "beta0": p["kp"] * (w / l),
        "theta": p["theta"].astype(np.float64),
        "cgs": cgs,
        "cgd": cgs.copy(),
        "cgb": np.zeros_like(w),
        "cdb": p["cj"] * w,
        "csb": p["cj"] * w,
    }


def _softplus(u):
    return np.maximum(u, 0.0) + np.log1p(np.exp(-np.abs(u)))


def _sigmoid(u):
    e = np.exp(-np.abs(u))
    return np.where(u >= 0.0, 1.0 / (1.0 + e), e / (1.0 + e))


def eval_mosfet(vgb: np.ndarray, vsb: np.ndarray, vdb: np.ndarray, k: dict):
    sgn = k["sgn"]
    gamma = k["gamma"]
    phi = k["phi"]
    VG = sgn * vgb
    VS = sgn * vsb
    VD = sgn * vdb
    vgp = VG + k["vgoff"]
    g2 = 0.5 * gamma
    t = vgp + g2 * g2
    big = t > 1e-12
    s = np.sqrt(np.where(big, t, 1.0))
    vp = np.where(big, vgp - phi - gamma * (s - g2), vgp - phi - gamma * gamma / 2.0)
    dvp = np.where(big, 1.0 - gamma / (2.0 * s), 1.0)
    arg = phi + vp + 4.0 * VT
    pos = arg > 1e-24
    sa = np.sqrt(np.where(pos, arg, 1.0))
    n = np.where(pos, 1.0 + gamma / (2.0 * sa), 1.0 + gamma / 2.0 * 1e48)
    dn = np.where(pos, -gamma / (4.0 * arg * sa), 0.0)
    den = 1.0 + k["theta"] * vp
    beta = k["beta0"] / den
    dbeta = -k["beta0"] * k["theta"] / (den * den)
    uf = 0.5 * ((vp - VS) / VT)
    ur = 0.5 * ((vp - VD) / VT)
    lf = _softplus(uf)
    lr = _softplus(ur)
    fpf = lf * _sigmoid(uf)
    fpr = lr * _sigmoid(ur)
    vt2 = VT * VT
    ispec = 2.0 * n * beta * vt2
    idiff = lf * lf - lr * lr
    i_d = ispec * idiff
    dispec = 2.0 * vt2 * (dn * beta + n * dbeta)
    gvp = dispec * idiff + ispec * (fpf - fpr) / VT
    gg = gvp * dvp
    gs = -ispec * fpf / VT
    gd = ispec * fpr / VT
    gb = -(gg + gs + gd)
    iext = sgn * i_d
    zero = np.zeros_like(iext)
    res_r = np.stack([iext, zero, -iext, zero], axis=1)
    jac_r = np.stack([gd, gg, gs, gb, zero, zero, zero, zero,
                      -gd, -gg, -gs, -gb, zero, zero, zero, zero], axis=1)
    cgs, cgd, cgb, cdb, csb = k["cgs"], k["cgd"], k["cgb"], k["cdb"], k["csb"]
    vgs = vgb - vsb
    vgd = vgb - vdb
    qg = cgs * vgs + cgd * vgd + cgb * vgb
    qd = -cgd * vgd + cdb * vdb
    qs = -cgs * vgs + csb * vsb
    qb = -(cgb * vgb + cdb * vdb + csb * vsb)
    res_c = np.stack([qd, qg, qs, qb], axis=1)
    jac_c = np.stack([
        cgd + cdb, -cgd, zero, -cdb,
        -cgd, cgs + cgd + cgb, -cgs, -cgb,
        zero, -cgs, cgs + csb, -csb,
        -cdb, -cgb, -csb, cdb + csb + cgb,
    ], axis=1)
    return res_r, res_c, jac_r, jac_c
