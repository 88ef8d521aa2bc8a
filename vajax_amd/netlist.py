"""Flat circuit description consumed by the transient core.

A :class:`Netlist` is the flattened leaf-device view that the reference builds in
``CircuitEngine.parse()`` (vajax/analysis/engine.py:421-485): a set of named
nodes, one ground node, and device instances grouped by model type.  Instances
of one model are stored as structure-of-arrays (terminal index columns and one
float64 column per parameter), which is the layout the HIP eval kernels read.

Node names are kept as given; numbering (ground = 0, the other names in Python
sorted order from 1) is applied by :mod:`vajax_amd.topology`, following
vajax/analysis/engine.py:459-468.

Voltage and current sources carry a waveform record.  Supported waveform
types are the reference's ``dc``, ``pulse``, ``sine``, ``pwl``, ``exp``,
``am`` and ``fm`` with its parameter names and defaults
(vajax/analysis/sources.py:204-294); a pwl source's ``wave`` list of
time/value pairs is kept per source in :attr:`SourceGroup.pwl`.
"""

from __future__ import annotations

from dataclasses import dataclass, field
from typing import Dict, List, Optional

import numpy as np

# Model types known to the transient core, with their terminal names.  The
# MOSFET entry is the EKV 2.6 long-channel core
# (vajax/devices/models/ekv_longchannel/ekv_longchannel.va) extended with a
# polarity flag and linear terminal capacitances; it stands in for PSP103,
# whose Verilog-A source is not available.
MODEL_TERMINALS: Dict[str, List[str]] = {
    "resistor": ["A", "B"],
    "capacitor": ["A", "B"],
    "mosfet": ["d", "g", "s", "b"],
    # junction diode, anode A, cathode K; its junction voltage goes through
    # $limit(..., "pnjlim", ...) (vajax/analysis/limiting.py:22-82)
    "diode": ["A", "K"],
    # EKV 2.6 with its charges and junction diodes (vajax/devices/models/ekv/ekv.va)
    "ekv26": ["d", "g", "s", "b"],
}

# EKV 2.6 instance parameters and their defaults (ekv.va:237-296).  "temp" and
# "tnom" keep the file's not-given marker 1e21: the device then runs at the
# simulation temperature plus "trise", with Tnom = 25 C.  "noise" is accepted
# and ignored (no noise analysis here).
EKV26_PARAMS: Dict[str, float] = {
    "type": 1.0, "noise": 1.0, "trise": 0.0, "temp": 1e21, "tnom": 1e21,
    "l": 10e-6, "w": 10e-6, "m": 1.0, "ns": 1.0,
    "as": 0.0, "ad": 0.0, "ps": 0.0, "pd": 0.0,
    "cox": 2.0e-3, "xj": 300e-9, "vto": 0.5, "tcv": 1.0e-3, "gamma": 0.7, "phi": 0.5,
    "kp": 150e-6, "bex": -1.5, "theta": 0.0, "e0": 1.0e8, "ucrit": 2.0e6, "ucex": 0.8,
    "lambda": 0.8, "dl": -0.01e-6, "dw": -0.01e-6, "weta": 0.2, "leta": 0.3,
    "q0": 230e-6, "lk": 0.4e-6, "iba": 5.0e8, "ibb": 4.0e8, "ibbt": 9.0e-4, "ibn": 1.0,
    "rsh": 0.0, "hdif": 0.5e-6, "avto": 1e-6, "akp": 1e-6, "agamma": 1e-6,
    "af": 1.0, "kf": 0.0,
    "xd_n": 1.0, "xd_js": 1.0e-09, "xd_jsw": 1.0e-12, "xd_jswg": 1.0e-12,
    "xd_mj": 0.900, "xd_mjsw": 0.700, "xd_mjswg": 0.700,
    "xd_pb": 0.800, "xd_pbsw": 0.600, "xd_pbswg": 0.600,
    "xd_cj": 1.0e-09, "xd_cjsw": 1.0e-12, "xd_cjswg": 1.0e-12,
    "xd_gmin": 0.0, "xd_xjbv": 0.0, "xd_bv": 10.0,
    "xd_njts": 1.0, "xd_njtssw": 1.0, "xd_njtsswg": 1.0,
    "xd_vts": 0.0, "xd_vtssw": 0.0, "xd_vtsswg": 0.0,
    "tp_xti": 3.0, "tp_cj": 0.0, "tp_cjsw": 0.0, "tp_cjswg": 0.0,
    "tp_pb": 0.0, "tp_pbsw": 0.0, "tp_pbswg": 0.0,
    "tp_njts": 0.0, "tp_njtssw": 0.0, "tp_njtsswg": 0.0,
}

# Per-model instance parameters and defaults.
MODEL_PARAMS: Dict[str, Dict[str, float]] = {
    # vajax/devices/models/resistor.va:19
    "resistor": {"r": 1.0},
    # vajax/devices/models/capacitor.va:17
    "capacitor": {"c": 1e-12},
    # EKV long-channel defaults (ekv_longchannel.va:44-56) plus polarity and
    # capacitance coefficients of this build's substitute model.
    "mosfet": {
        "type": 1.0,  # +1 NMOS, -1 PMOS
        "w": 10e-6,
        "l": 10e-6,
        "vto": 0.5,
        "gamma": 0.7,
        "phi": 0.5,
        "kp": 20e-6,
        "theta": 50e-3,
        "cox": 0.0,  # F/m^2, intrinsic gate capacitance per area
        "cov": 0.0,  # F/m, gate overlap capacitance per width
        "cj": 0.0,  # F/m, drain/source junction capacitance per width
    },
    # SPICE diode defaults: saturation current, emission coefficient, a
    # linear junction capacitance
    "diode": {"is": 1e-14, "n": 1.0, "cj": 0.0},
    "ekv26": EKV26_PARAMS,
}

SOURCE_DEFAULTS: Dict[str, float] = {
    # vajax/analysis/sources.py:219-227 (pulse defaults)
    "dc": 0.0,
    "val0": 0.0,
    "val1": 1.0,
    "rise": 1e-9,
    "fall": 1e-9,
    "width": 1e-6,
    "period": 2e-6,
    "delay": 0.0,
    # sine / am / fm (sources.py:231-236, 264-285)
    "sinedc": 0.0,
    "ampl": 1.0,
    "freq": 1e3,
    "phase": 0.0,
    "modfreq": 1e3,
    "modphase": 0.0,
    "modindex": 0.5,
    # exp (sources.py:254-262; val0, val1 and delay shared with pulse)
    "td2": 1e-9,
    "tau1": 1e-10,
    "tau2": 1e-10,
    # pwl (sources.py:238-252)
    "offset": 0.0,
    "scale": 1.0,
    "stretch": 1.0,
    "pwlperiod": 0.0,
}

WAVE_DC = 0
WAVE_PULSE = 1
WAVE_SINE = 2
WAVE_PWL = 3
WAVE_EXP = 4
WAVE_AM = 5
WAVE_FM = 6
WAVE_CODES = {"dc": WAVE_DC, "pulse": WAVE_PULSE, "sine": WAVE_SINE, "pwl": WAVE_PWL,
              "exp": WAVE_EXP, "am": WAVE_AM, "fm": WAVE_FM}


@dataclass
class DeviceGroup:
    """All instances of one model type, as structure-of-arrays."""

    model: str
    terminals: np.ndarray  # int64 [N, n_terminals], indices into Netlist.nodes
    params: Dict[str, np.ndarray]  # name -> float64 [N]
    names: Optional[List[str]] = None

    @property
    def count(self) -> int:
        return int(self.terminals.shape[0])


@dataclass
class SourceGroup:
    """Independent sources (vsource or isource) in device order."""

    kind: str  # "vsource" | "isource"
    names: List[str]
    terminals: np.ndarray  # int64 [m, 2] (p, n)
    wave: np.ndarray  # int32 [m] waveform code
    params: Dict[str, np.ndarray]  # SOURCE_DEFAULTS keys -> float64 [m]
    # per source, the pwl ``wave`` list [t0, v0, t1, v1, ...] (empty for the
    # other waveforms); None when no source is a pwl
    pwl: Optional[List[np.ndarray]] = None

    @property
    def count(self) -> int:
        return len(self.names)

    def pwl_points(self, i: int) -> np.ndarray:
        if self.pwl is None:
            return np.zeros(0)
        return np.asarray(self.pwl[i], dtype=np.float64)


def empty_sources(kind: str) -> SourceGroup:
    return SourceGroup(
        kind=kind,
        names=[],
        terminals=np.zeros((0, 2), dtype=np.int64),
        wave=np.zeros(0, dtype=np.int32),
        params={k: np.zeros(0) for k in SOURCE_DEFAULTS},
    )


@dataclass
class Netlist:
    """Flattened circuit: nodes, device groups (model order = first appearance)."""

    title: str
    nodes: List[str]
    ground: str
    groups: Dict[str, DeviceGroup]
    vsources: SourceGroup
    isources: SourceGroup
    analysis: Dict[str, float] = field(default_factory=dict)
    options: Dict[str, object] = field(default_factory=dict)
    # Internal device nodes are left out of the Newton residual test, and
    # noise-only internal (NOI) nodes are also held at 0 V with their rows and
    # columns of the linear system replaced by the identity
    # (vajax/analysis/solver_factories.py:37-137, 1268-1276).
    internal_nodes: List[str] = field(default_factory=list)
    noi_nodes: List[str] = field(default_factory=list)

    def device_count(self) -> int:
        return sum(g.count for g in self.groups.values())


class NetlistBuilder:
    """Incremental construction of a :class:`Netlist` from named instances."""

    def __init__(self, title: str, ground: str = "0"):
        self.title = title
        self.ground = ground
        self._node_index: Dict[str, int] = {}
        self._nodes: List[str] = []
        self._dev_terms: Dict[str, List[List[int]]] = {}
        self._dev_params: Dict[str, Dict[str, List[float]]] = {}
        self._dev_names: Dict[str, List[str]] = {}
        self._src: Dict[str, Dict[str, list]] = {
            k: {"names": [], "terms": [], "wave": [], "pwl": [], **{p: [] for p in SOURCE_DEFAULTS}}
            for k in ("vsource", "isource")
        }
        self.node(ground)

    def node(self, name: str) -> int:
        idx = self._node_index.get(name)
        if idx is None:
            idx = len(self._nodes)
            self._node_index[name] = idx
            self._nodes.append(name)
        return idx

    def device(self, model: str, name: str, terminals, **params) -> None:
        if model not in MODEL_TERMINALS:
            raise ValueError(f"unknown model type {model!r}")
        if len(terminals) != len(MODEL_TERMINALS[model]):
            raise ValueError(f"{model} needs {len(MODEL_TERMINALS[model])} terminals")
        if model not in self._dev_terms:
            self._dev_terms[model] = []
            self._dev_params[model] = {p: [] for p in MODEL_PARAMS[model]}
            self._dev_names[model] = []
        self._dev_terms[model].append([self.node(t) for t in terminals])
        defaults = MODEL_PARAMS[model]
        unknown = set(params) - set(defaults)
        if unknown:
            raise ValueError(f"unknown {model} parameters {sorted(unknown)}")
        for p, d in defaults.items():
            self._dev_params[model][p].append(float(params.get(p, d)))
        self._dev_names[model].append(name)

    def source(self, kind: str, name: str, p: str, n: str, type: str = "dc", wave=None,
               **params) -> None:
        """A vsource / isource; ``wave`` is a pwl source's [t0, v0, t1, v1, ...]
        (sources.py:238-252, default [0, 0])."""
        if kind not in self._src:
            raise ValueError(f"unknown source kind {kind!r}")
        if type not in WAVE_CODES:
            raise ValueError(f"unsupported waveform {type!r}")
        if "sinephase" in params and "phase" not in params:
            # am / fm accept sinephase for phase (sources.py:268, 280)
            params["phase"] = params.pop("sinephase")
        unknown = set(params) - set(SOURCE_DEFAULTS)
        if unknown:
            raise ValueError(f"unknown source parameters {sorted(unknown)}")
        pts = np.zeros(0)
        if type == "pwl":
            pts = np.asarray([0.0, 0.0] if wave is None else wave, dtype=np.float64).ravel()
            if pts.size < 2 or pts.size % 2:
                raise ValueError(f"pwl source {name}: wave needs time/value pairs")
            if np.any(np.diff(pts[0::2]) < 0):
                raise ValueError(f"pwl source {name}: times must not decrease")
        elif wave is not None:
            raise ValueError(f"source {name}: wave is a pwl parameter")
        s = self._src[kind]
        s["names"].append(name)
        s["terms"].append([self.node(p), self.node(n)])
        s["wave"].append(WAVE_CODES[type])
        s["pwl"].append(pts)
        for k, d in SOURCE_DEFAULTS.items():
            s[k].append(float(params.get(k, d)))

    def build(self, analysis=None, options=None, internal_nodes=(), noi_nodes=()) -> Netlist:
        groups = {}
        for model, terms in self._dev_terms.items():
            groups[model] = DeviceGroup(
                model=model,
                terminals=np.asarray(terms, dtype=np.int64).reshape(-1, len(MODEL_TERMINALS[model])),
                params={k: np.asarray(v, dtype=np.float64) for k, v in self._dev_params[model].items()},
                names=self._dev_names[model],
            )
        srcs = {}
        for kind, s in self._src.items():
            srcs[kind] = SourceGroup(
                kind=kind,
                names=list(s["names"]),
                terminals=np.asarray(s["terms"], dtype=np.int64).reshape(-1, 2),
                wave=np.asarray(s["wave"], dtype=np.int32),
                params={k: np.asarray(s[k], dtype=np.float64) for k in SOURCE_DEFAULTS},
                pwl=list(s["pwl"]) if any(w == WAVE_PWL for w in s["wave"]) else None,
            )
        return Netlist(
            title=self.title,
            nodes=list(self._nodes),
            ground=self.ground,
            groups=groups,
            vsources=srcs["vsource"],
            isources=srcs["isource"],
            analysis=dict(analysis or {}),
            options=dict(options or {}),
            internal_nodes=[n for n in internal_nodes],
            noi_nodes=[n for n in noi_nodes],
        )


# Parameter columns of each waveform in the C ABI's [VJX_SRC_NPAR][m] block
# (include/vjx.h); column 0 is always dc.
SRC_NPAR = 10
_WAVE_COLUMNS = {
    WAVE_DC: (),
    WAVE_PULSE: ("val0", "val1", "rise", "fall", "width", "period", "delay"),
    WAVE_SINE: ("sinedc", "ampl", "freq", "phase"),
    WAVE_EXP: ("val0", "val1", "delay", "td2", "tau1", "tau2"),
    WAVE_AM: ("sinedc", "ampl", "freq", "phase", "delay", "modfreq", "modphase", "modindex"),
    WAVE_FM: ("sinedc", "ampl", "freq", "phase", "delay", "modfreq", "modphase", "modindex"),
    WAVE_PWL: ("pwlperiod",),
}


def pack_sources(src: SourceGroup):
    """(wave int32[m], params float64[SRC_NPAR][m], pwl_ptr, pwl_t, pwl_v) for
    the C ABI.  pwl points are stored as the reference evaluates them:
    times * stretch, values * scale + offset (vajax/analysis/sources.py:100-101);
    the three pwl arrays are None when no source is a pwl."""
    m = src.count
    wave = np.ascontiguousarray(src.wave, dtype=np.int32)
    par = np.zeros((SRC_NPAR, m), dtype=np.float64)
    if m:
        par[0] = src.params["dc"]
    for i in range(m):
        for c, name in enumerate(_WAVE_COLUMNS[int(wave[i])], start=1):
            par[c, i] = src.params[name][i]
    if not np.any(wave == WAVE_PWL):
        return wave, par, None, None, None
    ptr = np.zeros(m + 1, dtype=np.int32)
    ts, vs = [], []
    for i in range(m):
        if int(wave[i]) == WAVE_PWL:
            w = src.pwl_points(i)
            ts.append(w[0::2] * src.params["stretch"][i])
            vs.append(w[1::2] * src.params["scale"][i] + src.params["offset"][i])
            ptr[i + 1] = ptr[i] + len(ts[-1])
        else:
            ptr[i + 1] = ptr[i]
    tt = np.ascontiguousarray(np.concatenate(ts), dtype=np.float64)
    vv = np.ascontiguousarray(np.concatenate(vs), dtype=np.float64)
    return wave, par, ptr, tt, vv
