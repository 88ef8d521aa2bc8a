"""Oracle setup and system build (CSR-direct), restated from the reference.

Setup (independent of ``vajax_amd.topology``; dictionaries and sets instead of
vectorised numpy, so the two implementations check each other):

* node numbering: ground 0, other names sorted from 1 (engine.py:459-468);
* stamp indices with -1 for ground (mna_builder.py:44-118);
* CSR pattern = sorted unique (row, col) of the trial COO build, where masked
  entries fold onto (0, 0) (full_mna.py:423-479, mna.py:300-335, 400-469,
  650-667).

``build_system`` follows ``_build_system_csr_direct`` (mna_builder.py:759-1010)
and ``_build_residual`` (mna_builder.py:1012-1076):

1. current sources stamp +I at p and -I at n into f_resist and max|.| (:830-837);
2. per model (in model order): branch voltages V[n1]-V[n2], eval, stamp
   ``jac_resist + c0*jac_react`` into csr_data, res_resist into f_resist and
   max|.|, res_react into Q, and for limited models lim_rhs_resist /
   lim_rhs_react into their own sums (:965-966) -- NaN -> 0 on every stamped
   value; sums are sequential in (device, key) order, the order of XLA's
   scatter-add on CPU;
3. csr_data[diag] += gmin_opt + gshunt; B/B^T +-1 (:969-989);
4. f_resist -= lim_rhs_resist (:1039); I_vsource = -f_resist[p-1]
   (mna.py:564-596); f_resist += B*I_branch (mna.py:489-536);
   f_node = f_resist - 0 + c0*(Q - lim_rhs_react) + c1*Q_prev
   + d1*dQdt_prev + c2*Q_prev2, then + (gmin_opt+gshunt)*V[1:]
   (mna.py:673-719); f_branch = V[p] - V[n] - E (mna.py:539-561).

``gmin_opt`` is the netlist option gmin used for diagonal regularisation
(mna_builder.py:384); the device simparam gmin is not read by these models.
"""

from __future__ import annotations

from dataclasses import dataclass
from typing import Dict, List

import numpy as np

from vajax_amd.netlist import Netlist

from . import ekv26 as EKV
from . import models as M

TERMS = {"resistor": 2, "capacitor": 2, "mosfet": 4, "diode": 2, "ekv26": 4}
VPAIRS = {"resistor": [(0, 1)], "capacitor": [(0, 1)], "mosfet": [(1, 3), (2, 3), (0, 3)],
          "ekv26": [(1, 3), (2, 3), (0, 3)],
          "diode": [(0, 1)]}
# limit states per instance (one per $limit call)
N_LIM = {"resistor": 0, "capacitor": 0, "mosfet": 0, "diode": 1, "ekv26": 0}


@dataclass
class OracleModel:
    name: str
    nodes: np.ndarray  # [N, T] global node index
    res_idx: np.ndarray  # [N, T]
    csr_pos: np.ndarray  # [N, T*T]
    params: Dict[str, np.ndarray]
    cache: Dict[str, np.ndarray]
    lim_off: int = 0  # first limit state of this group in the flat array


@dataclass
class OracleCircuit:
    node_names: Dict[str, int]
    n_total: int
    n_unk: int
    m: int
    n_aug: int
    models: List[OracleModel]
    vs_p: np.ndarray
    vs_n: np.ndarray
    is_p: np.ndarray
    is_n: np.ndarray
    indptr: np.ndarray
    indices: np.ndarray
    diag_pos: np.ndarray
    b_pos: List[tuple]  # (csr position, value) in stamping order
    net: Netlist
    gmin: float = 1e-12
    n_lim: int = 0  # limit states: model order, instance, state (mna_builder.py:387-401)
    # residual rows of noise-only internal nodes, sorted, and the residual
    # convergence mask (False for internal device nodes and NOI nodes),
    # solver_factories.py:37-137
    noi_res: np.ndarray = None
    conv_mask: np.ndarray = None

    @property
    def nse(self):
        return len(self.indices)


def setup(net: Netlist, gmin: float = 1e-12) -> OracleCircuit:
    ground = net.ground
    names = {ground: 0}
    for i, n in enumerate(sorted(set(net.nodes) - {ground}), start=1):
        names[n] = i
    gidx = np.array([names[n] for n in net.nodes], dtype=np.int64)
    n_total = len(names)
    n_unk = n_total - 1
    m = net.vsources.count
    n_aug = n_unk + m

    entries = set()
    raw_models = []
    for mname, grp in net.groups.items():
        T = TERMS[mname]
        nodes = gidx[grp.terminals]
        rows = []
        for r in range(T):
            for c in range(T):
                rr = nodes[:, r] - 1
                cc = nodes[:, c] - 1
                ok = (rr >= 0) & (cc >= 0)
                rows.append((np.where(ok, rr, -1), np.where(ok, cc, -1)))
        for rr, cc in rows:
            ok = rr >= 0
            entries.update(zip(rr[ok].tolist(), cc[ok].tolist()))
            if not np.all(ok):
                entries.add((0, 0))
        raw_models.append((mname, grp, nodes, rows))
    vs_p = gidx[net.vsources.terminals[:, 0]] if m else np.zeros(0, np.int64)
    vs_n = gidx[net.vsources.terminals[:, 1]] if m else np.zeros(0, np.int64)
    for i in range(m):
        br = n_unk + i
        for node in (vs_p[i], vs_n[i]):
            if node > 0:
                entries.add((node - 1, br))
                entries.add((br, node - 1))
            else:
                entries.add((0, 0))
    for i in range(n_unk):
        entries.add((i, i))

    ordered = sorted(entries)
    pos_of = {e: k for k, e in enumerate(ordered)}
    indptr = np.zeros(n_aug + 1, dtype=np.int64)
    for r, _ in ordered:
        indptr[r + 1] += 1
    indptr = np.cumsum(indptr)
    indices = np.array([c for _, c in ordered], dtype=np.int64)

    models = []
    for mname, grp, nodes, rows in raw_models:
        T = TERMS[mname]
        csr = np.full((grp.count, T * T), -1, dtype=np.int64)
        for k, (rr, cc) in enumerate(rows):
            csr[:, k] = [pos_of[(a, b)] if a >= 0 else -1 for a, b in zip(rr.tolist(), cc.tolist())]
        res_idx = np.where(nodes > 0, nodes - 1, -1)
        cache = (M.mosfet_cache(grp.params) if mname == "mosfet" else
                 M.diode_cache(grp.params) if mname == "diode" else
                 EKV.cache(grp.params) if mname == "ekv26" else {})
        models.append(OracleModel(mname, nodes, res_idx, csr, grp.params, cache))
    n_lim = 0
    for om in models:
        om.lim_off = n_lim
        n_lim += N_LIM[om.name] * om.nodes.shape[0]
    diag_pos = np.array([pos_of[(i, i)] for i in range(n_unk)], dtype=np.int64)
    # B/B^T in the reference's stamping order: all b_p, all b_n, all bt_p, all bt_n
    b_pos = []
    for i in range(m):
        if vs_p[i] > 0:
            b_pos.append((pos_of[(vs_p[i] - 1, n_unk + i)], 1.0))
    for i in range(m):
        if vs_n[i] > 0:
            b_pos.append((pos_of[(vs_n[i] - 1, n_unk + i)], -1.0))
    for i in range(m):
        if vs_p[i] > 0:
            b_pos.append((pos_of[(n_unk + i, vs_p[i] - 1)], 1.0))
    for i in range(m):
        if vs_n[i] > 0:
            b_pos.append((pos_of[(n_unk + i, vs_n[i] - 1)], -1.0))
    k = net.isources.count
    is_p = gidx[net.isources.terminals[:, 0]] if k else np.zeros(0, np.int64)
    is_n = gidx[net.isources.terminals[:, 1]] if k else np.zeros(0, np.int64)
    noi_res = np.array(sorted(names[nm] - 1 for nm in set(net.noi_nodes)), dtype=np.int64)
    conv_mask = np.ones(n_unk, dtype=bool)
    for nm in list(net.internal_nodes) + list(net.noi_nodes):
        conv_mask[names[nm] - 1] = False
    return OracleCircuit(names, n_total, n_unk, m, n_aug, models, vs_p, vs_n, is_p, is_n,
                         indptr, indices, diag_pos, b_pos, net, gmin, n_lim, noi_res, conv_mask)


def eval_model(om: OracleModel, V: np.ndarray, lim_in=None):
    """(res_resist, res_react, jac_resist, jac_react, lim_rhs_resist,
    lim_rhs_react, limit_state_out) of one model group at node voltages V;
    lim_in is the group's slice of the limit state."""
    n = om.nodes
    if om.name == "diode":
        return M.eval_diode(V[n[:, 0]] - V[n[:, 1]], lim_in[:, 0], om.cache)
    rr, rc, jr, jc = _eval_plain(om, V)
    return rr, rc, jr, jc, None, None, None


def _eval_plain(om: OracleModel, V: np.ndarray):
    n = om.nodes
    if om.name == "resistor":
        return M.eval_resistor(V[n[:, 0]] - V[n[:, 1]], om.params["r"])
    if om.name == "capacitor":
        return M.eval_capacitor(V[n[:, 0]] - V[n[:, 1]], om.params["c"])
    if om.name == "mosfet":
        vgb = V[n[:, 1]] - V[n[:, 3]]
        vsb = V[n[:, 2]] - V[n[:, 3]]
        vdb = V[n[:, 0]] - V[n[:, 3]]
        return M.eval_mosfet(vgb, vsb, vdb, om.cache)
    if om.name == "ekv26":
        vgb = V[n[:, 1]] - V[n[:, 3]]
        vsb = V[n[:, 2]] - V[n[:, 3]]
        vdb = V[n[:, 0]] - V[n[:, 3]]
        return EKV.eval_ekv26(vgb, vsb, vdb, om.cache)
    raise ValueError(om.name)


def _nan0(x):
    return np.where(np.isnan(x), 0.0, x)


def build_system(cir: OracleCircuit, X, vsrc, isrc, Q_prev, c0, gshunt=0.0, c1=0.0, d1=0.0,
                 dQdt_prev=None, c2=0.0, Q_prev2=None):
    """Returns (csr_data, f_aug, Q, I_vsource, max_res_contrib) with the limit
    state at zero."""
    return build_system_lim(cir, X, vsrc, isrc, Q_prev, c0, gshunt, c1, d1, dQdt_prev, c2,
                            Q_prev2)[:5]


def build_system_lim(cir: OracleCircuit, X, vsrc, isrc, Q_prev, c0, gshunt=0.0, c1=0.0, d1=0.0,
                     dQdt_prev=None, c2=0.0, Q_prev2=None, lim_in=None):
    """Returns (csr_data, f_aug, Q, I_vsource, max_res_contrib, limit_state_out).

    lim_in (default zeros) is the flat limit state; the lim_rhs sums of the
    limited models are scattered like their residuals and subtracted from
    f_resist and, inside c0*(Q - .), from the charges (mna_builder.py:1039,
    mna.py:673-719)."""
    lim_in = np.zeros(cir.n_lim) if lim_in is None else np.asarray(lim_in, dtype=np.float64)
    lim_out = np.zeros(cir.n_lim)
    lim_r = np.zeros(cir.n_unk)
    lim_q = np.zeros(cir.n_unk)
    n_total, n_unk = cir.n_total, cir.n_unk
    V = X[:n_total]
    I_branch = X[n_total:]
    csr = np.zeros(cir.nse)
    f_res = np.zeros(n_unk)
    Q = np.zeros(n_unk)
    mrc = np.zeros(n_unk)

    k = len(isrc)
    if k:
        idx = np.stack([np.where(cir.is_p > 0, cir.is_p - 1, -1),
                        np.where(cir.is_n > 0, cir.is_n - 1, -1)], axis=1).ravel()
        val = (np.asarray(isrc)[:, None] * np.array([1.0, -1.0])[None, :]).ravel()
        ok = idx >= 0
        np.add.at(f_res, idx[ok], _nan0(val[ok]))
        np.maximum.at(mrc, idx[ok], _nan0(np.abs(val[ok])))

    for om in cir.models:
        N = om.nodes.shape[0]
        nl = N_LIM[om.name]
        li = lim_in[om.lim_off:om.lim_off + N * nl].reshape(N, nl) if nl else None
        rr, rc, jr, jc, lr, lc, lo = eval_model(om, V, li)
        if nl:
            lim_out[om.lim_off:om.lim_off + N * nl] = lo.ravel()
        comb = jr + c0 * jc
        pos = om.csr_pos.ravel()
        ok = pos >= 0
        np.add.at(csr, pos[ok], _nan0(comb.ravel()[ok]))
        ridx = om.res_idx.ravel()
        ok = ridx >= 0
        np.add.at(f_res, ridx[ok], _nan0(rr.ravel()[ok]))
        np.maximum.at(mrc, ridx[ok], _nan0(np.abs(rr.ravel()[ok])))
        np.add.at(Q, ridx[ok], _nan0(rc.ravel()[ok]))
        if lr is not None:
            np.add.at(lim_r, ridx[ok], _nan0(lr.ravel()[ok]))
            np.add.at(lim_q, ridx[ok], _nan0(lc.ravel()[ok]))

    csr[cir.diag_pos] += cir.gmin + gshunt
    for p, v in cir.b_pos:
        csr[p] += v

    f_res = f_res - lim_r
    I_vs = np.where(cir.vs_p > 0, -f_res[np.maximum(cir.vs_p - 1, 0)], 0.0) if cir.m else np.zeros(0)
    if cir.m:
        contrib = np.zeros(n_unk)
        pi = np.where(cir.vs_p > 0, cir.vs_p - 1, 0)
        ni = np.where(cir.vs_n > 0, cir.vs_n - 1, 0)
        np.add.at(contrib, np.concatenate([pi, ni]),
                  np.concatenate([np.where(cir.vs_p > 0, I_branch, 0.0),
                                  np.where(cir.vs_n > 0, -I_branch, 0.0)]))
        f_res = f_res + contrib
    dQ = dQdt_prev if dQdt_prev is not None else np.zeros(n_unk)
    Q2 = Q_prev2 if Q_prev2 is not None else np.zeros(n_unk)
    f_node = (f_res - np.zeros(n_unk) + c0 * (Q - lim_q) + c1 * Q_prev
              + d1 * dQ + c2 * Q2)
    f_node = f_node + (cir.gmin + gshunt) * V[1:]
    f_branch = V[cir.vs_p] - V[cir.vs_n] - np.asarray(vsrc) if cir.m else np.zeros(0)
    return csr, np.concatenate([f_node, f_branch]), Q, I_vs, mrc, lim_out


def csr_matrix(cir: OracleCircuit, data):
    import scipy.sparse as sp

    return sp.csr_matrix((data, cir.indices, cir.indptr), shape=(cir.n_aug, cir.n_aug))
