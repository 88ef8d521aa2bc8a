"""Setup for the transient core: node numbering, MNA indices, CSR pattern.

Everything here runs once per circuit on the host (numpy) and produces the
integer maps that the HIP kernels consume:

* node numbering -- ground 0, other names in sorted order from 1
  (vajax/analysis/engine.py:459-468); MNA row = node index - 1; branch rows of
  voltage sources follow in device order (vajax/analysis/mna.py:196-227);
* per-model stamp indices ``res_indices`` / ``jac_row`` / ``jac_col`` with -1
  for ground (vajax/analysis/mna_builder.py:44-118);
* the CSR pattern: sorted unique ``row * n_aug + col`` over every device
  Jacobian key, the node-row diagonal and the voltage-source B/B^T entries,
  masked entries folding onto (0, 0) exactly as the reference's trial COO
  build does (vajax/analysis/transient/full_mna.py:423-513,
  vajax/analysis/mna.py:300-335, 400-469, 650-667);
* CSR slot positions for device entries, diagonal and B/B^T
  (vajax/analysis/mna_builder.py:126-255);
* the deterministic reduction maps used by the stamping kernels: every
  (device, Jacobian key) contribution is assigned a slot in a list sorted by
  (CSR slot, model order, device, key) so that the per-slot sum has a fixed
  order and needs no atomics.
"""

from __future__ import annotations

from dataclasses import dataclass, field
from typing import Dict, List, Tuple

import numpy as np

from .netlist import MODEL_TERMINALS, Netlist

# Residual node names and Jacobian keys per model (the openvaf_jax
# ``dae_metadata`` contract, openvaf_jax/__init__.py:253-280).
MODEL_RES_NODES: Dict[str, List[str]] = {m: list(t) for m, t in MODEL_TERMINALS.items()}
MODEL_JAC_KEYS: Dict[str, List[Tuple[str, str]]] = {
    m: [(r, c) for r in t for c in t] for m, t in MODEL_TERMINALS.items()
}
# Branch voltages each model reads, as (terminal, terminal) pairs -- the
# ``voltage_node1/2`` gather of vajax/analysis/mna_builder.py:915.
MODEL_VOLTAGES: Dict[str, List[Tuple[str, str]]] = {
    "resistor": [("A", "B")],
    "capacitor": [("A", "B")],
    "mosfet": [("g", "b"), ("s", "b"), ("d", "b")],
    "ekv26": [("g", "b"), ("s", "b"), ("d", "b")],
    "diode": [("A", "K")],
}
MODEL_IDS = {"resistor": 0, "capacitor": 1, "mosfet": 2, "diode": 3, "ekv26": 4}
# limit states per instance (one per $limit call), flat over models in order,
# instances, states (vajax/analysis/mna_builder.py:387-401)
MODEL_LIMIT_STATES = {"resistor": 0, "capacitor": 0, "mosfet": 0, "diode": 1, "ekv26": 0}


@dataclass
class ModelStamp:
    model: str
    count: int
    node_terms: np.ndarray  # int32 [N, n_term] global node index (0 = ground)
    v_node1: np.ndarray  # int32 [N, n_v]
    v_node2: np.ndarray  # int32 [N, n_v]
    res_idx: np.ndarray  # int32 [N, n_res]  (-1 = ground)
    jac_row: np.ndarray  # int32 [N, n_jac]
    jac_col: np.ndarray  # int32 [N, n_jac]
    csr_pos: np.ndarray = None  # int32 [N, n_jac] (-1 = dropped)
    params: Dict[str, np.ndarray] = field(default_factory=dict)


@dataclass
class Topology:
    """Integer structure of one circuit's augmented MNA system."""

    node_names: Dict[str, int]
    n_total: int
    n_unknowns: int
    n_vsources: int
    n_augmented: int
    n_external: int
    models: Dict[str, ModelStamp]
    vs_node_p: np.ndarray  # int32 [m]
    vs_node_n: np.ndarray
    is_node_p: np.ndarray  # int32 [k]
    is_node_n: np.ndarray
    is_f_idx: np.ndarray  # int32 [k, 2] residual index (-1 = ground)
    indptr: np.ndarray  # int32 [n_aug + 1]
    indices: np.ndarray  # int32 [nse]
    unique_linear: np.ndarray  # int64 [nse]
    diag_pos: np.ndarray  # int32 [n_unknowns]
    vs_pos: Dict[str, np.ndarray]  # b_p, b_n, bt_p, bt_n (int32 [m], -1 invalid)

    @property
    def nse(self) -> int:
        return int(self.indices.shape[0])


def number_nodes(net: Netlist) -> Tuple[Dict[str, int], np.ndarray]:
    """Ground 0, remaining names sorted -> 1..; returns (name map, local->global)."""
    others = sorted(n for n in net.nodes if n != net.ground)
    names = {net.ground: 0}
    for i, n in enumerate(others, start=1):
        names[n] = i
    local_to_global = np.array([names[n] for n in net.nodes], dtype=np.int64)
    return names, local_to_global


def build_topology(net: Netlist) -> Topology:
    node_names, l2g = number_nodes(net)
    n_total = len(node_names)
    n_unk = n_total - 1
    m = net.vsources.count
    n_aug = n_unk + m

    models: Dict[str, ModelStamp] = {}
    for model, grp in net.groups.items():
        terms = MODEL_TERMINALS[model]
        tpos = {t: i for i, t in enumerate(terms)}
        gterm = l2g[grp.terminals].astype(np.int64)  # [N, n_term]
        vpairs = MODEL_VOLTAGES[model]
        v1 = np.stack([gterm[:, tpos[a]] for a, _ in vpairs], axis=1)
        v2 = np.stack([gterm[:, tpos[b]] for _, b in vpairs], axis=1)
        res = np.stack([gterm[:, tpos[t]] for t in MODEL_RES_NODES[model]], axis=1) - 1
        keys = MODEL_JAC_KEYS[model]
        jr = np.stack([gterm[:, tpos[r]] for r, _ in keys], axis=1) - 1
        jc = np.stack([gterm[:, tpos[c]] for _, c in keys], axis=1) - 1
        valid = (jr >= 0) & (jc >= 0)
        jr = np.where(valid, jr, -1)
        jc = np.where(valid, jc, -1)
        models[model] = ModelStamp(
            model=model,
            count=grp.count,
            node_terms=gterm.astype(np.int32),
            v_node1=v1.astype(np.int32),
            v_node2=v2.astype(np.int32),
            res_idx=res.astype(np.int32),
            jac_row=jr.astype(np.int32),
            jac_col=jc.astype(np.int32),
            params=grp.params,
        )

    vs_p = l2g[net.vsources.terminals[:, 0]].astype(np.int32) if m else np.zeros(0, np.int32)
    vs_n = l2g[net.vsources.terminals[:, 1]].astype(np.int32) if m else np.zeros(0, np.int32)
    k = net.isources.count
    is_p = l2g[net.isources.terminals[:, 0]].astype(np.int32) if k else np.zeros(0, np.int32)
    is_n = l2g[net.isources.terminals[:, 1]].astype(np.int32) if k else np.zeros(0, np.int32)
    is_f = np.stack([np.where(is_p != 0, is_p - 1, -1), np.where(is_n != 0, is_n - 1, -1)],
                    axis=1).astype(np.int32) if k else np.zeros((0, 2), np.int32)

    # --- CSR pattern from the trial COO set (full_mna.py:461-479) ---
    lin_parts = []
    for ms in models.values():
        valid = ms.jac_row >= 0
        r = np.where(valid, ms.jac_row, 0).astype(np.int64).ravel()
        c = np.where(valid, ms.jac_col, 0).astype(np.int64).ravel()
        lin_parts.append(r * n_aug + c)
    if m:
        br = np.arange(m, dtype=np.int64) + n_unk
        vp = vs_p > 0
        vn = vs_n > 0
        pm = np.where(vp, vs_p.astype(np.int64) - 1, 0)
        nm = np.where(vn, vs_n.astype(np.int64) - 1, 0)
        lin_parts.append(pm * n_aug + np.where(vp, br, 0))
        lin_parts.append(nm * n_aug + np.where(vn, br, 0))
        lin_parts.append(np.where(vp, br, 0) * n_aug + pm)
        lin_parts.append(np.where(vn, br, 0) * n_aug + nm)
    d = np.arange(n_unk, dtype=np.int64)
    lin_parts.append(d * n_aug + d)
    linear = np.concatenate(lin_parts) if lin_parts else np.zeros(0, np.int64)
    unique_linear = np.unique(linear)
    nse = unique_linear.shape[0]
    rows = (unique_linear // n_aug).astype(np.int64)
    cols = (unique_linear % n_aug).astype(np.int32)
    indptr = np.zeros(n_aug + 1, dtype=np.int32)
    indptr[1:] = np.cumsum(np.bincount(rows, minlength=n_aug))
    indices = cols
    _validate_csr(indptr, indices, n_aug)

    def find(r, c, valid):
        lin = r.astype(np.int64) * n_aug + c.astype(np.int64)
        pos = np.clip(np.searchsorted(unique_linear, lin), 0, max(nse - 1, 0))
        ok = (unique_linear[pos] == lin) & valid
        return np.where(ok, pos, -1).astype(np.int32)

    for ms in models.values():
        valid = (ms.jac_row >= 0) & (ms.jac_col >= 0)
        ms.csr_pos = find(ms.jac_row, ms.jac_col, valid)
    diag_pos = find(d, d, np.ones(n_unk, dtype=bool))
    if np.any(diag_pos < 0):
        raise ValueError("CSR pattern is missing node-row diagonal entries")
    vs_pos = {}
    if m:
        br = np.arange(m, dtype=np.int64) + n_unk
        vp = vs_p > 0
        vn = vs_n > 0
        vs_pos = {
            "b_p": find(np.where(vp, vs_p - 1, 0), np.where(vp, br, 0), vp),
            "b_n": find(np.where(vn, vs_n - 1, 0), np.where(vn, br, 0), vn),
            "bt_p": find(np.where(vp, br, 0), np.where(vp, vs_p - 1, 0), vp),
            "bt_n": find(np.where(vn, br, 0), np.where(vn, vs_n - 1, 0), vn),
        }

    return Topology(
        node_names=node_names,
        n_total=n_total,
        n_unknowns=n_unk,
        n_vsources=m,
        n_augmented=n_aug,
        n_external=n_total,
        models=models,
        vs_node_p=vs_p,
        vs_node_n=vs_n,
        is_node_p=is_p,
        is_node_n=is_n,
        is_f_idx=is_f,
        indptr=indptr,
        indices=indices,
        unique_linear=unique_linear,
        diag_pos=diag_pos,
        vs_pos=vs_pos,
    )


def _validate_csr(indptr: np.ndarray, indices: np.ndarray, n: int) -> None:
    """Structural checks of vajax/analysis/transient/full_mna.py:481-513."""
    nse = indices.shape[0]
    if indptr[0] != 0:
        raise ValueError(f"CSR indptr[0]={indptr[0]} != 0")
    if indptr[-1] != nse:
        raise ValueError(f"CSR indptr[-1]={indptr[-1]} != nse={nse}")
    if np.any(np.diff(indptr) < 0):
        raise ValueError("CSR indptr not monotonic")
    if nse and (indices.min() < 0 or indices.max() >= n):
        raise ValueError("CSR column index out of range")
    if nse > 1:
        row_start = np.zeros(nse, dtype=bool)
        starts = indptr[:-1][indptr[:-1] < nse]
        row_start[starts] = True
        interior = ~row_start[1:]
        if np.any(np.diff(indices.astype(np.int64))[interior] <= 0):
            raise ValueError("CSR column indices not strictly sorted within a row")


@dataclass
class ReductionMap:
    """Fixed-order segmented reductions for atomic-free stamping.

    ``jac_seg`` partitions the contribution list by CSR slot; contribution
    ``q`` of the sorted list comes from model ``jac_src_model[q]`` and flat
    index ``jac_src_flat[q]`` (= device * n_jac + key).  ``jac_dst`` maps every
    (model, device, key) to its sorted position (or -1 when dropped), which is
    where the eval kernel writes it.  ``res_*`` is the same for residual rows,
    with current-source contributions ahead of device contributions.
    """

    jac_seg: np.ndarray  # int32 [nse + 1]
    jac_dst: Dict[str, np.ndarray]  # model -> int32 [N * n_jac]
    res_seg: np.ndarray  # int32 [n_unk + 1]
    res_dst: Dict[str, np.ndarray]  # model -> int32 [N * n_res]
    is_dst: np.ndarray  # int32 [k * 2]
    n_jac_contrib: int
    n_res_contrib: int


def build_reduction_map(topo: Topology) -> ReductionMap:
    model_list = list(topo.models.keys())
    # Jacobian contributions, key (slot, model order, device, key) == stable sort
    # by slot of the concatenation in (model, device, key) order.
    slots, owners = [], []
    for mi, name in enumerate(model_list):
        ms = topo.models[name]
        pos = ms.csr_pos.ravel().astype(np.int64)
        slots.append(pos)
        owners.append(np.full(pos.shape[0], mi, dtype=np.int64))
    all_slots = np.concatenate(slots) if slots else np.zeros(0, np.int64)
    keep = all_slots >= 0
    order = np.argsort(np.where(keep, all_slots, np.iinfo(np.int64).max), kind="stable")
    n_keep = int(keep.sum())
    order = order[:n_keep]
    dst_flat = np.full(all_slots.shape[0], -1, dtype=np.int64)
    dst_flat[order] = np.arange(n_keep)
    jac_seg = np.zeros(topo.nse + 1, dtype=np.int32)
    jac_seg[1:] = np.cumsum(np.bincount(all_slots[keep], minlength=topo.nse))
    jac_dst = {}
    off = 0
    for name in model_list:
        n = topo.models[name].csr_pos.size
        jac_dst[name] = dst_flat[off:off + n].astype(np.int32)
        off += n

    # Residual contributions: isources first (mna_builder.py:830-837), then models.
    n_unk = topo.n_unknowns
    rows = [topo.is_f_idx.ravel().astype(np.int64)]
    for name in model_list:
        rows.append(topo.models[name].res_idx.ravel().astype(np.int64))
    all_rows = np.concatenate(rows)
    keep = all_rows >= 0
    order = np.argsort(np.where(keep, all_rows, np.iinfo(np.int64).max), kind="stable")
    n_keep = int(keep.sum())
    order = order[:n_keep]
    dst_flat = np.full(all_rows.shape[0], -1, dtype=np.int64)
    dst_flat[order] = np.arange(n_keep)
    res_seg = np.zeros(n_unk + 1, dtype=np.int32)
    res_seg[1:] = np.cumsum(np.bincount(all_rows[keep], minlength=n_unk))
    off = rows[0].shape[0]
    is_dst = dst_flat[:off].astype(np.int32)
    res_dst = {}
    for name, r in zip(model_list, rows[1:]):
        res_dst[name] = dst_flat[off:off + r.shape[0]].astype(np.int32)
        off += r.shape[0]
    return ReductionMap(
        jac_seg=jac_seg,
        jac_dst=jac_dst,
        res_seg=res_seg,
        res_dst=res_dst,
        is_dst=is_dst,
        n_jac_contrib=int(jac_seg[-1]),
        n_res_contrib=int(res_seg[-1]),
    )
