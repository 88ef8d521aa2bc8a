"""Seeded synthetic circuits used for parity tests and benchmarks.

All generators draw from ``numpy.random.default_rng(seed)`` so a (generator,
size, seed) triple names one circuit exactly.  Node names are zero-padded so
that Python's sorted order (the reference's node numbering,
vajax/analysis/engine.py:459-468) equals construction order.

* :func:`rc_ladder` -- series R / shunt C low-pass ladder driven by a pulse.
* :func:`ring_oscillator` -- odd-stage CMOS inverter ring with load caps and a
  one-shot current kick on the first stage output.
* :func:`gate_dag` -- random DAG of 2-input NAND/NOR gates with bounded fanout
  and 64 pulse-train primary inputs.  Gate inputs are drawn from a window of
  recently created signals (geometric distance), which gives the netlist the
  locality of a placed design instead of the expander structure of a uniform
  random graph.
* :func:`monte_carlo` -- per-device threshold/mobility variation of a netlist.
* :func:`with_ekv26` -- the same netlist with its MOSFETs as EKV 2.6 devices.

MOSFETs use the EKV long-channel substitute model (``mosfet`` model type) at
the technology values in :data:`NMOS_DEFAULTS` / :data:`PMOS_DEFAULTS`.
"""

from __future__ import annotations

import copy
from typing import Dict

import numpy as np

from .netlist import (
    MODEL_PARAMS,
    SOURCE_DEFAULTS,
    WAVE_DC,
    WAVE_PULSE,
    DeviceGroup,
    Netlist,
    NetlistBuilder,
    SourceGroup,
    empty_sources,
)

NMOS_DEFAULTS: Dict[str, float] = {
    "type": 1.0,
    "vto": 0.35,
    "gamma": 0.4,
    "phi": 0.8,
    "kp": 200e-6,
    "theta": 0.2,
    "cox": 0.02,
    "cov": 2e-10,
    "cj": 5e-10,
}
# EKV 2.6 (``ekv26``) technology values for the same circuits: a short-channel
# process at a 1 V supply (oxide 2.3 nm, junction depth 30 nm), with mobility
# reduction, velocity saturation, channel-length modulation, short/narrow
# channel effects, impact ionisation and the drain/source junctions.
EKV26_NMOS: Dict[str, float] = {
    "type": 1.0, "vto": 0.35, "gamma": 0.4, "phi": 0.8, "kp": 300e-6, "cox": 0.015,
    "xj": 30e-9, "e0": 1.0e8, "ucrit": 5.0e6, "lambda": 0.5, "leta": 0.1, "weta": 0.2,
    "q0": 0.0, "hdif": 0.1e-6, "xd_cj": 1.0e-3, "xd_cjsw": 1.0e-10, "xd_cjswg": 1.0e-10,
}
EKV26_PMOS: Dict[str, float] = dict(EKV26_NMOS, type=-1.0, vto=-0.35, kp=120e-6)

PMOS_DEFAULTS: Dict[str, float] = {
    "type": -1.0,
    "vto": 0.35,
    "gamma": 0.4,
    "phi": 0.8,
    "kp": 80e-6,
    "theta": 0.2,
    "cox": 0.02,
    "cov": 2e-10,
    "cj": 5e-10,
}

VDD = 1.0


def rc_ladder(sections: int = 200, seed: int = 0, t_stop: float = 10e-9, dt: float = 1e-12) -> Netlist:
    """RC low-pass ladder: in -> R1 -> n001 -> R2 -> ... -> n{sections}.

    R_i ~ U[1k, 10k] ohm in series, C_i ~ U[0.1, 1] pF from n_i to ground,
    driven by a 0->1 V pulse with 10 ps rise/fall and 1 ns width.
    """
    rng = np.random.default_rng(seed)
    r = rng.uniform(1e3, 1e4, sections)
    c = rng.uniform(1e-13, 1e-12, sections)
    width = len(str(sections))
    b = NetlistBuilder(f"rc_ladder_{sections}_s{seed}")
    b.source("vsource", "v1", "in", "0", type="pulse", val0=0.0, val1=1.0, rise=10e-12,
             fall=10e-12, width=1e-9, delay=0.0, period=1.0)
    prev = "in"
    for i in range(sections):
        node = f"n{i + 1:0{max(3, width)}d}"
        b.device("resistor", f"r{i + 1}", [prev, node], r=r[i])
        b.device("capacitor", f"c{i + 1}", [node, "0"], c=c[i])
        prev = node
    return b.build(analysis={"stop": t_stop, "step": dt},
                   options={"tran_method": "gear2", "icmode": "op"})


def diode_ladder(sections: int = 100, seed: int = 0, t_stop: float = 5e-9,
                 dt: float = 1e-12) -> Netlist:
    """RC ladder with clamp diodes, the carrier of device limiting (pnjlim).

    in -> R_i -> n_i, C_i from n_i to ground, and at every node a diode to
    ground (anode n_i) plus, at every other node, one from ground (anode 0):
    R ~ U[100, 1k] ohm, C ~ U[10, 100] fF, is ~ U[1e-15, 1e-13] A,
    n ~ U[1, 1.5], cj ~ U[1, 10] fF.  Driven by a -1.5 -> 2 V pulse, so both
    diode polarities are driven hard into forward bias (already at the DC
    point) and the junction voltages go through pnjlim.
    """
    rng = np.random.default_rng(seed)
    r = rng.uniform(100.0, 1e3, sections)
    c = rng.uniform(10e-15, 100e-15, sections)
    width = max(3, len(str(sections)))
    b = NetlistBuilder(f"diode_ladder_{sections}_s{seed}")
    b.source("vsource", "v1", "in", "0", type="pulse", val0=-1.5, val1=2.0, rise=20e-12,
             fall=20e-12, width=1e-9, delay=0.2e-9, period=1.0)
    prev = "in"
    for i in range(sections):
        node = f"n{i + 1:0{width}d}"
        b.device("resistor", f"r{i + 1}", [prev, node], r=r[i])
        b.device("capacitor", f"c{i + 1}", [node, "0"], c=c[i])
        b.device("diode", f"d{i + 1}", [node, "0"], **{"is": rng.uniform(1e-15, 1e-13),
                                                     "n": rng.uniform(1.0, 1.5),
                                                     "cj": rng.uniform(1e-15, 10e-15)})
        if i % 2 == 0:
            b.device("diode", f"dr{i + 1}", ["0", node], **{"is": rng.uniform(1e-15, 1e-13),
                                                           "n": rng.uniform(1.0, 1.5),
                                                           "cj": rng.uniform(1e-15, 10e-15)})
        prev = node
    return b.build(analysis={"stop": t_stop, "step": dt},
                   options={"tran_method": "gear2", "icmode": "op"})


def diode_clamp_chain(stages: int = 6, vdc: float = 5.0, r: float = 50.0) -> Netlist:
    """A hard DC operating point: a 5 V source into ``stages`` sections of
    series R (50 ohm) with two stacked junction diodes (is 1e-14 A, n 1) and a
    10 k resistor to ground at each section.  From the mid-rail start, plain
    Newton needs 16 iterations; with fewer allowed the DC solve goes through
    the homotopy chain."""
    b = NetlistBuilder(f"diode_clamp_{stages}")
    b.source("vsource", "v1", "in", "0", type="dc", dc=vdc)
    prev = "in"
    p = {"is": 1e-14, "n": 1.0, "cj": 1e-15}
    for i in range(stages):
        node, k = f"d{i + 1}", f"k{i + 1}"
        b.device("resistor", f"r{i + 1}", [prev, node], r=r)
        b.device("diode", f"da{i + 1}", [node, k], **p)
        b.device("diode", f"db{i + 1}", [k, "0"], **p)
        b.device("resistor", f"rk{i + 1}", [k, "0"], r=1e4)
        prev = node
    return b.build(analysis={"stop": 1e-9, "step": 1e-12}, options={"icmode": "op"})


def ring_oscillator(stages: int = 101, seed: int = 0, t_stop: float = 50e-9, dt: float = 1e-12,
                    model: str = "mosfet") -> Netlist:
    """CMOS inverter ring with per-stage random W/L and load caps.

    Each stage: NMOS(out_i, in_i, 0, 0) and PMOS(out_i, in_i, vdd, vdd) with
    in_i = out_{i-1}; W ~ U[0.2, 2] um, L ~ U[45, 90] nm, C_load ~ U[1, 5] fF.
    Started from all-zero initial conditions (icmode uic) with a 100 uA,
    50 ps current pulse drawn out of stage 0.  ``model`` as for
    :func:`gate_dag`.
    """
    nmos, pmos = _technology(model)
    if stages % 2 == 0:
        raise ValueError("ring needs an odd number of stages")
    rng = np.random.default_rng(seed)
    wn = rng.uniform(0.2e-6, 2e-6, stages)
    ln = rng.uniform(45e-9, 90e-9, stages)
    wp = rng.uniform(0.2e-6, 2e-6, stages)
    lp = rng.uniform(45e-9, 90e-9, stages)
    cl = rng.uniform(1e-15, 5e-15, stages)
    b = NetlistBuilder(f"ring_{stages}_s{seed}")
    b.source("vsource", "vdd", "vdd", "0", type="dc", dc=VDD)
    out = [f"out{i:03d}" for i in range(stages)]
    for i in range(stages):
        src = out[i - 1]
        b.device(model, f"mn{i:03d}", [out[i], src, "0", "0"], w=wn[i], l=ln[i], **nmos)
        b.device(model, f"mp{i:03d}", [out[i], src, "vdd", "vdd"], w=wp[i], l=lp[i], **pmos)
        b.device("capacitor", f"cl{i:03d}", [out[i], "0"], c=cl[i])
    b.source("isource", "ikick", out[0], "0", type="pulse", val0=0.0, val1=100e-6, rise=10e-12,
             fall=10e-12, width=50e-12, delay=0.0, period=1.0)
    return b.build(analysis={"stop": t_stop, "step": dt},
                   options={"tran_method": "gear2", "icmode": "uic"})


def _technology(model: str):
    """(NMOS, PMOS) instance parameters of the synthetic circuits for a model."""
    if model == "mosfet":
        return NMOS_DEFAULTS, PMOS_DEFAULTS
    if model == "ekv26":
        return EKV26_NMOS, EKV26_PMOS
    raise ValueError(f"model must be 'mosfet' or 'ekv26', not {model!r}")


def _dag_inputs(rng: np.random.Generator, n_gates: int, n_pi: int, fanout_max: int,
                window: int, mean_dist: float) -> np.ndarray:
    """Pick two distinct driver signals per gate from the recent-signal window."""
    n_sig = n_pi + n_gates
    fanout = np.zeros(n_sig, dtype=np.int64)
    inputs = np.empty((n_gates, 2), dtype=np.int64)
    p = 1.0 / mean_dist
    # pre-draw candidate distances in bulk; fall back to a scan when exhausted
    draws = rng.geometric(p, size=(n_gates, 24))
    for g in range(n_gates):
        avail = n_pi + g  # signals 0..avail-1 exist
        lo = max(0, avail - window)
        chosen = []
        for d in draws[g]:
            if len(chosen) == 2:
                break
            if d > avail - lo:
                continue
            s = avail - int(d)
            if fanout[s] >= fanout_max or s in chosen:
                continue
            chosen.append(s)
        if len(chosen) < 2:
            # deterministic fallback: nearest signals with spare fanout
            s = avail - 1
            while len(chosen) < 2 and s >= 0:
                if fanout[s] < fanout_max and s not in chosen:
                    chosen.append(s)
                s -= 1
            if len(chosen) < 2:
                raise RuntimeError("gate DAG generator ran out of fanout capacity")
        for s in chosen:
            fanout[s] += 1
        inputs[g] = chosen
    return inputs


def gate_dag(n_gates: int = 20000, seed: int = 0, n_pi: int = 64, fanout_max: int = 4,
             window: int = 96, mean_dist: float = 12.0, t_stop: float = 1e-9,
             dt: float = 1e-12, vdd_rise: float = 100e-12, model: str = "mosfet") -> Netlist:
    """Random DAG of 2-input NAND/NOR gates (4 MOSFETs each).

    Gate g picks NAND or NOR with probability 1/2 and two distinct inputs
    among the last ``window`` signals (primary inputs and earlier gate
    outputs) at geometric distance with mean ``mean_dist``; a signal drives
    at most ``fanout_max`` gate inputs.  Each gate adds an output node
    ``g<k>y`` and a series-stack node ``g<k>m``.  Primary inputs are pulse
    trains: delay ~ U[0, 1 ns], width ~ U[0.2, 2] ns, rise = fall = 20 ps,
    period = 2 width + rise + fall.  W ~ U[0.2, 2] um, L ~ U[45, 90] nm.
    The transient starts from zero initial conditions (icmode uic) with the
    supply ramped 0 -> 1 V over ``vdd_rise`` (a power-up): the DC point of a CMOS netlist
    whose cut-off stack nodes are held only by gmin is numerically singular
    for plain Newton from the mid-rail guess, while every transient Newton
    matrix is regularised by c0*C.  ``model`` is "mosfet" (the long-channel
    substitute) or "ekv26" (EKV 2.6 at :data:`EKV26_NMOS` / :data:`EKV26_PMOS`);
    both draw the same random numbers, so they give the same netlist topology.
    """
    _technology(model)
    rng = np.random.default_rng(seed)
    is_nand = rng.random(n_gates) < 0.5
    inputs = _dag_inputs(rng, n_gates, n_pi, fanout_max, window, mean_dist)
    n_mos = 4 * n_gates
    w = rng.uniform(0.2e-6, 2e-6, n_mos)
    l = rng.uniform(45e-9, 90e-9, n_mos)
    pi_delay = rng.uniform(0.0, 1e-9, n_pi)
    pi_width = rng.uniform(0.2e-9, 2e-9, n_pi)

    gw = len(str(n_gates))
    gw = max(6, gw)
    pw = max(2, len(str(n_pi)))
    # node table: ground, vdd, PIs, then per gate (m, y)
    nodes = ["0", "vdd"] + [f"pi{i:0{pw}d}" for i in range(n_pi)]
    base = len(nodes)
    gate_names_m = [f"g{g:0{gw}d}m" for g in range(n_gates)]
    gate_names_y = [f"g{g:0{gw}d}y" for g in range(n_gates)]
    inter = [None] * (2 * n_gates)
    inter[0::2] = gate_names_m
    inter[1::2] = gate_names_y
    nodes.extend(inter)
    GND, VDDN = 0, 1
    sig_node = np.empty(n_pi + n_gates, dtype=np.int64)
    sig_node[:n_pi] = 2 + np.arange(n_pi)
    m_node = base + 2 * np.arange(n_gates)
    y_node = m_node + 1
    sig_node[n_pi:] = y_node
    a = sig_node[inputs[:, 0]]
    bb = sig_node[inputs[:, 1]]

    # four devices per gate in order: P1, P2, N1, N2
    term = np.empty((n_gates, 4, 4), dtype=np.int64)
    typ = np.empty((n_gates, 4), dtype=np.float64)
    typ[:, 0:2] = -1.0
    typ[:, 2:4] = 1.0
    vdd_col = np.full(n_gates, VDDN)
    gnd_col = np.full(n_gates, GND)
    # NAND: P(y,a,vdd,vdd) P(y,b,vdd,vdd) N(y,a,m,0) N(m,b,0,0)
    # NOR:  P(m,a,vdd,vdd) P(y,b,m,vdd)   N(y,a,0,0) N(y,b,0,0)
    nd = is_nand
    term[:, 0] = np.stack([np.where(nd, y_node, m_node), a, vdd_col, vdd_col], axis=1)
    term[:, 1] = np.stack([y_node, bb, np.where(nd, vdd_col, m_node), vdd_col], axis=1)
    term[:, 2] = np.stack([y_node, a, np.where(nd, m_node, gnd_col), gnd_col], axis=1)
    term[:, 3] = np.stack([np.where(nd, m_node, y_node), bb, gnd_col, gnd_col], axis=1)
    terminals = term.reshape(n_mos, 4)
    typ = typ.reshape(n_mos)

    params = {"type": typ, "w": w, "l": l}
    if model == "mosfet":
        for k in ("vto", "gamma", "phi", "kp", "theta", "cox", "cov", "cj"):
            params[k] = np.where(typ > 0, NMOS_DEFAULTS[k], PMOS_DEFAULTS[k]).astype(np.float64)
    else:
        for k, v in MODEL_PARAMS["ekv26"].items():
            if k not in params:
                params[k] = np.where(typ > 0, EKV26_NMOS.get(k, v),
                                     EKV26_PMOS.get(k, v)).astype(np.float64)
    mos = DeviceGroup(model=model, terminals=terminals, params=params, names=None)

    m_src = 1 + n_pi
    vs_terms = np.zeros((m_src, 2), dtype=np.int64)
    vs_terms[0] = [VDDN, GND]
    vs_terms[1:, 0] = 2 + np.arange(n_pi)
    vs_wave = np.full(m_src, WAVE_PULSE, dtype=np.int32)
    vsp = {k: np.full(m_src, v, dtype=np.float64) for k, v in SOURCE_DEFAULTS.items()}
    # supply: power-up ramp 0 -> VDD over vdd_rise, then held (one pulse with
    # a period far beyond any run)
    vsp["dc"][0] = VDD
    vsp["val0"][0] = 0.0
    vsp["val1"][0] = VDD
    vsp["rise"][0] = vdd_rise
    vsp["width"][0] = 1.0
    vsp["period"][0] = 10.0
    vsp["delay"][0] = 0.0
    vsp["val0"][1:] = 0.0
    vsp["val1"][1:] = VDD
    vsp["rise"][1:] = 20e-12
    vsp["fall"][1:] = 20e-12
    vsp["width"][1:] = pi_width
    vsp["period"][1:] = 2 * pi_width + 40e-12
    vsp["delay"][1:] = pi_delay
    vs = SourceGroup(kind="vsource", names=["vdd"] + [f"vpi{i:0{pw}d}" for i in range(n_pi)],
                     terminals=vs_terms, wave=vs_wave, params=vsp)
    return Netlist(
        title=f"gate_dag_{n_gates}_s{seed}" + ("" if model == "mosfet" else f"_{model}"),
        nodes=nodes,
        ground="0",
        groups={model: mos},
        vsources=vs,
        isources=empty_sources("isource"),
        analysis={"stop": t_stop, "step": dt},
        options={"tran_method": "gear2", "icmode": "uic"},
    )


def monte_carlo(net: Netlist, seed: int, sigma: float = 0.03) -> Netlist:
    """Copy of ``net`` with per-MOSFET Vth and mobility variation.

    dVth ~ N(0, sigma * vto) added to vto, kp scaled by N(1, sigma), for the
    MOSFETs of either model.
    """
    rng = np.random.default_rng(seed)
    out = copy.deepcopy(net)
    out.title = f"{net.title}_mc{seed}"
    g = out.groups.get("mosfet", out.groups.get("ekv26"))
    if g is not None:
        n = g.count
        g.params["vto"] = g.params["vto"] + rng.normal(0.0, 1.0, n) * sigma * g.params["vto"]
        g.params["kp"] = g.params["kp"] * rng.normal(1.0, sigma, n)
    return out


def with_ekv26(net: Netlist) -> Netlist:
    """Copy of ``net`` whose MOSFETs (the long-channel ``mosfet`` group) are
    EKV 2.6 devices: same terminals, polarity and W/L, the technology values
    of :data:`EKV26_NMOS` / :data:`EKV26_PMOS`.  The group keeps its place in
    the netlist's group order."""
    out = copy.deepcopy(net)
    g = out.groups.get("mosfet")
    if g is None:
        return out
    typ = np.asarray(g.params["type"], dtype=np.float64)
    params = {"type": typ, "w": np.asarray(g.params["w"], np.float64),
              "l": np.asarray(g.params["l"], np.float64)}
    for k, v in MODEL_PARAMS["ekv26"].items():
        if k not in params:
            params[k] = np.where(typ > 0, EKV26_NMOS.get(k, v), EKV26_PMOS.get(k, v)).astype(
                np.float64)
    groups = {}
    for name, grp in out.groups.items():
        if name == "mosfet":
            groups["ekv26"] = DeviceGroup(model="ekv26", terminals=grp.terminals, params=params,
                                          names=grp.names)
        else:
            groups[name] = grp
    out.groups = groups
    out.title = f"{net.title}_ekv26"
    return out


def pack(nets, title: str = "batch") -> Netlist:
    """Pack several circuits into one block-diagonal netlist.

    Instance ``b`` keeps its devices and sources; its node ``x`` becomes
    ``i<b>.x`` (ground stays shared).  The prefixes keep each instance's
    nodes contiguous in the sorted node numbering, so the packed MNA matrix
    has one diagonal block per instance and the sparse LU analyses each
    distinct block pattern once.  All instances advance on the packed
    system's common time grid.
    """
    bw = max(3, len(str(max(len(nets) - 1, 0))))
    nodes = ["0"]
    groups: Dict[str, list] = {}
    vs: list = []
    is_: list = []
    internal: list = []
    noi: list = []
    for b, net in enumerate(nets):
        pre = f"i{b:0{bw}d}."
        internal += [pre + n for n in net.internal_nodes]
        noi += [pre + n for n in net.noi_nodes]
        gidx = net.nodes.index(net.ground)
        remap = np.empty(len(net.nodes), dtype=np.int64)
        for k, name in enumerate(net.nodes):
            if k == gidx:
                remap[k] = 0
            else:
                remap[k] = len(nodes)
                nodes.append(pre + name)
        for model, g in net.groups.items():
            groups.setdefault(model, []).append((remap[g.terminals], g.params))
        for src, acc in ((net.vsources, vs), (net.isources, is_)):
            if src.count:
                acc.append((pre, remap[src.terminals], src))

    def cat_sources(kind, acc):
        if not acc:
            return empty_sources(kind)
        return SourceGroup(
            kind=kind,
            names=[pre + n for pre, _, s in acc for n in s.names],
            terminals=np.concatenate([t for _, t, _ in acc]),
            wave=np.concatenate([s.wave for _, _, s in acc]).astype(np.int32),
            params={k: np.concatenate([s.params[k] for _, _, s in acc]) for k in SOURCE_DEFAULTS},
            pwl=([s.pwl_points(i) for _, _, s in acc for i in range(s.count)]
                 if any(s.pwl is not None for _, _, s in acc) else None),
        )

    out_groups = {}
    for model, parts in groups.items():
        keys = parts[0][1].keys()
        out_groups[model] = DeviceGroup(
            model=model,
            terminals=np.concatenate([t for t, _ in parts]),
            params={k: np.concatenate([p[k] for _, p in parts]) for k in keys},
            names=None)
    first = nets[0]
    return Netlist(
        title=title,
        nodes=nodes,
        ground="0",
        groups=out_groups,
        vsources=cat_sources("vsource", vs),
        isources=cat_sources("isource", is_),
        analysis=dict(first.analysis),
        options=dict(first.options),
        internal_nodes=internal,
        noi_nodes=noi,
    )


def batch_mc(n_instances: int = 256, gates: int = 2000, ragged=(500, 1000, 2000, 4000),
             seed: int = 0, sigma: float = 0.03, model: str = "mosfet") -> Netlist:
    """Monte Carlo batch plus ragged topologies, packed into one netlist.

    ``n_instances`` copies of the ``gates``-gate DAG (seed ``seed``) with
    per-device Vth / mobility variation (instance ``b`` drawn with seed
    ``seed + 1 + b``), followed by one DAG of each size in ``ragged`` (seeds
    ``seed + 1000 + j``).  ``model`` as for :func:`gate_dag`.
    """
    base = gate_dag(gates, seed=seed, model=model)
    nets = [monte_carlo(base, seed + 1 + b, sigma) for b in range(n_instances)]
    nets += [gate_dag(g, seed=seed + 1000 + j, model=model) for j, g in enumerate(ragged)]
    suffix = "" if model == "mosfet" else f"_{model}"
    return pack(nets, title=f"batch_mc{n_instances}x{gates}_ragged{len(ragged)}_s{seed}{suffix}")


def array_multiplier_netlist(n: int = 64) -> str:
    """VACASK netlist text of an n x n array multiplier built from CMOS gate
    subcircuits, the topology the reference's mul64 benchmark describes
    (vajax/benchmarks/data/mul64/runme.sim:8-13 and the comments of its
    generator): gates not / nor / and (NAND + inverter) on the nmos / pmos
    wrappers of the including netlist, xor = not(nor(and(a, not b),
    and(not a, b))), half adder = xor + and, full adder = two xor, two and, a
    nor and an inverter for the carry; partial products pp[i][j] = a[i] and
    b[j]; adder rows: row 0 half adders, rows 1..n-2 a half adder at column 0
    and full adders at columns 1..n-2; outputs p[0] = pp[0][0] through two
    inverters, p[1..n-1] the column-0 sums, p[n..2n-3] the last row's sums at
    columns 1..n-2 and p[2n-2] its carry at column n-2.  Written here, so a
    netlist's generate block of that name is never executed."""
    L = ["global vdd vss", ""]

    def sub(name, ports, insts):
        L.append(f"subckt {name}({' '.join(ports)})")
        for iname, master, nodes, params in insts:
            p = "".join(f" {k}={v}" for k, v in params.items())
            L.append(f"  {iname} ({' '.join(nodes)}) {master}{p}")
        L.append("ends")
        L.append("")

    P = {"w": "1u", "l": "0.2u"}
    N = {"w": "0.5u", "l": "0.2u"}
    sub("not", ["out", "in"], [("mp", "pmos", ["out", "in", "vdd", "vdd"], P),
                               ("mn", "nmos", ["out", "in", "vss", "vss"], N)])
    sub("nor", ["out", "in1", "in2"], [
        ("mp2", "pmos", ["int", "in2", "vdd", "vdd"], P),
        ("mp1", "pmos", ["out", "in1", "int", "vdd"], P),
        ("mn1", "nmos", ["out", "in1", "vss", "vss"], N),
        ("mn2", "nmos", ["out", "in2", "vss", "vss"], N)])
    sub("and", ["out", "in1", "in2"], [
        ("mp2", "pmos", ["outx", "in2", "vdd", "vdd"], P),
        ("mp1", "pmos", ["outx", "in1", "vdd", "vdd"], P),
        ("mn1", "nmos", ["outx", "in1", "int", "vss"], N),
        ("mn2", "nmos", ["int", "in2", "vss", "vss"], N),
        ("mp3", "pmos", ["out", "outx", "vdd", "vdd"], P),
        ("mn3", "nmos", ["out", "outx", "vss", "vss"], N)])
    sub("xor", ["out", "a", "b"], [
        ("inv_a", "not", ["na", "a"], {}), ("inv_b", "not", ["nb", "b"], {}),
        ("a1", "and", ["t1", "a", "nb"], {}), ("a2", "and", ["t2", "na", "b"], {}),
        ("n1", "nor", ["t3", "t1", "t2"], {}), ("inv_out", "not", ["out", "t3"], {})])
    sub("ha", ["sum", "cout", "a", "b"], [("x1", "xor", ["sum", "a", "b"], {}),
                                         ("a1", "and", ["cout", "a", "b"], {})])
    sub("fa", ["sum", "cout", "a", "b", "cin"], [
        ("x1", "xor", ["ab_xor", "a", "b"], {}), ("x2", "xor", ["sum", "ab_xor", "cin"], {}),
        ("a1", "and", ["ab_and", "a", "b"], {}), ("a2", "and", ["xor_cin", "ab_xor", "cin"], {}),
        ("n1", "nor", ["cout_n", "ab_and", "xor_cin"], {}),
        ("inv_cout", "not", ["cout", "cout_n"], {})])
    last = n - 2

    def sum_name(row, col):
        if col == 0:
            return f"p{row + 1}"
        if row == last:
            return f"p{n + col - 1}" if col < n - 1 else f"s{row}_{col}"
        return f"s{row}_{col}"

    def carry_name(row, col):
        if row == last and col == n - 2:
            return f"p{2 * n - 2}"
        return f"c{row}_{col}"

    insts = []
    for i in range(n):
        for j in range(n):
            insts.append((f"pp_{i}_{j}", "and", [f"pp{i}_{j}", f"a{i}", f"b{j}"], {}))
    insts.append(("buf_p0_inv", "not", ["p0_n", "pp0_0"], {}))
    insts.append(("buf_p0", "not", ["p0", "p0_n"], {}))
    for j in range(n - 1):
        insts.append((f"ha_0_{j}", "ha", [sum_name(0, j), carry_name(0, j), f"pp0_{j + 1}",
                                          f"pp1_{j}"], {}))
    for row in range(1, n - 1):
        for j in range(n - 1):
            b_in = f"pp{row + 1}_{j}"
            a_in = sum_name(row - 1, j + 1) if j < n - 2 else carry_name(row - 1, n - 2)
            if j == 0:
                insts.append((f"ha_{row}_{j}", "ha", [sum_name(row, j), carry_name(row, j), a_in,
                                                      b_in], {}))
            else:
                insts.append((f"fa_{row}_{j}", "fa", [sum_name(row, j), carry_name(row, j), a_in,
                                                      b_in, carry_name(row, j - 1)], {}))
    ports = ([f"a{i}" for i in range(n)] + [f"b{i}" for i in range(n)]
             + [f"p{i}" for i in range(2 * n - 1)])
    sub(f"multiplier_{n}x{n}", ports, insts)
    return "\n".join(L)


def array_multiplier_testbench(n: int = 64, a_bits=None, b_bits=None, t_stop: float = 2e-9,
                               dt: float = 2e-12, vdd: float = 1.2) -> str:
    """VACASK netlist text of an n x n array multiplier testbench in the
    arrangement of the reference's mul64 benchmark
    (vajax/benchmarks/data/mul64/runme.sim): nmos / pmos wrapper subcircuits
    (w, l; PSP103 cards, which the front end maps onto the EKV substitute),
    the multiplier from :func:`array_multiplier_netlist`, every input driven
    through a 1 ohm resistor by a 0.1 ns pulse from 0 to its bit (all bits
    high by default), supplies vdd / vss, a uic transient with the
    benchmark's tran_lteratio = 1.5."""
    a_bits = [1] * n if a_bits is None else list(a_bits)
    b_bits = [1] * n if b_bits is None else list(b_bits)
    L = [f"// {n}x{n} array multiplier testbench", "",
         "subckt nmos (d g s b)", "parameters w=1u l=0.2u",
         "  m (d g s b) psp103n w=w l=l", "ends", "",
         "subckt pmos (d g s b)", "parameters w=1u l=0.2u",
         "  m (d g s b) psp103p w=w l=l", "ends", "",
         "model psp103n psp103va (type=1)", "model psp103p psp103va (type=-1)", "",
         array_multiplier_netlist(n), "",
         "model v vsource", "model r sp_resistor", "",
         f"parameters vdd={vdd!r}", "",
         "subckt drv(out)", "  parameters v0=0 v1=1", "  parameters rs=1 delay=0.1n rise=0.1n",
         '  vdrv (int 0) v type="pulse" val0=v0*vdd val1=v1*vdd delay=delay rise=rise',
         "  rdrv (int out) r r=rs", "ends", "",
         "vdd (vdd 0) v dc=vdd", "vss (vss 0) v dc=0", "",
         "subckt mul_test()"]
    ports = ([f"a{i}" for i in range(n)] + [f"b{i}" for i in range(n)]
             + [f"p{i}" for i in range(2 * n - 1)])
    L.append(f"  u1 ({' '.join(ports)}) multiplier_{n}x{n}")
    for i in range(n):
        L.append(f"  da{i} (a{i}) drv v0=0 v1={int(a_bits[i])}")
    for i in range(n):
        L.append(f"  db{i} (b{i}) drv v0=0 v1={int(b_bits[i])}")
    L += ["ends", "", "control", "  options tran_lteratio=1.5", '  elaborate circuit("mul_test")',
          f'  analysis tr tran stop={t_stop!r} step={dt!r} icmode="uic"', "endc", ""]
    return "\n".join(L)


def array_multiplier_mosfets(n: int) -> int:
    """MOSFETs of :func:`array_multiplier_netlist`: and 6, not 2, nor 4,
    xor 22, half adder 28, full adder 62; n^2 partial products, two
    inverters, n - 1 half adders in row 0 and, in each of rows 1..n-2, one
    half adder and n - 2 full adders."""
    return 6 * n * n + 4 + 28 * (n - 1) + (n - 2) * (28 + 62 * (n - 2))
