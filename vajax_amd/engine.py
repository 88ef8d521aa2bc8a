"""Transient analysis API mirroring vajax's ``CircuitEngine``.

    engine = CircuitEngine(netlist)        # a vajax_amd.netlist.Netlist
    engine.parse()                         # numbering, pattern, reduction maps
    engine.prepare(t_stop=..., dt=..., precision="f64")
    result = engine.run_transient()        # TransientResult(times, voltages, currents, stats)

Contract followed: vajax/analysis/engine.py:115-189 (TransientResult),
:421-485 (parse), :824-982 (prepare) and :985-1023 (run_transient);
AdaptiveConfig / SimulationOptions defaults from
vajax/analysis/transient/adaptive.py:37-91 and vajax/analysis/options.py:35-192.
The time loop itself runs on the GPU through ``vjx_transient_run``.
"""

from __future__ import annotations

import ctypes as C
import math
from dataclasses import dataclass, field, fields, replace
from typing import Any, Dict, Optional, Tuple

import numpy as np

from . import _lib as L
from .models import param_block
from .netlist import Netlist, pack_sources
from .sparse import SparseLU
from .topology import Topology, build_reduction_map, build_topology

_METHOD_ALIASES = {"be": "be", "euler": "be", "backward_euler": "be", "trap": "trap",
                   "trapezoidal": "trap", "am2": "trap", "gear2": "gear2", "gear": "gear2",
                   "bdf2": "gear2", "bdf": "gear2"}


@dataclass
class AdaptiveConfig:
    """Step-control parameters (vajax/analysis/transient/adaptive.py:37-91)."""

    lte_ratio: float = 3.5
    redo_factor: float = 2.5
    reltol: float = 1e-3
    abstol: float = 1e-12
    min_dt: float = 1e-15
    max_dt: float = float("inf")
    tran_minpts: int = 50
    warmup_steps: int = 2
    max_order: int = 2
    grow_factor: float = 2.0
    nr_convtol: float = 1.0
    gshunt_init: float = 0.0
    gshunt_steps: int = 5
    gshunt_target: float = 0.0
    tran_fs: float = 0.25
    tran_ft: float = 0.25
    integration_method: str = "trap"
    max_consecutive_rejects: int = 5


# output rows kept on the device by default before chunked copies to the host
OUTPUT_RESIDENT_BYTES = 8 << 30


@dataclass
class SimulationOptions:
    """Options used on the transient path (vajax/analysis/options.py:35-192)."""

    nr_damping: float = 1.0
    tran_itl: int = 20
    reltol: float = 1e-3
    abstol: float = 1e-12
    vntol: float = 1e-6
    gmin: float = 1e-12
    icmode: str = "op"
    fixed_step_transient: bool = False
    # re-run the symbolic analysis when the factors' max |L| exceeds this
    # (pivot growth; 0 keeps the library default 1e12)
    repivot_growth: float = 0.0
    # DC operating point: Newton caps and the homotopy chain (options.py:63-67,
    # 148-190)
    op_itl: int = 100
    op_itlcont: int = 50
    op_homotopy: Tuple[str, ...] = ("gdev", "gshunt", "src")
    homotopy_gminsteps: int = 100
    homotopy_gminfactor: float = 10.0
    homotopy_maxgminfactor: float = 10.0
    homotopy_mingminfactor: float = 1.00005
    homotopy_startgmin: float = 1e-3
    homotopy_maxgmin: float = 1e2
    homotopy_mingmin: float = 1e-15
    homotopy_srcsteps: int = 100
    homotopy_srcstep: float = 0.01
    homotopy_srcscale: float = 3.0
    homotopy_minsrcstep: float = 1e-7


@dataclass
class TransientResult:
    """vajax/analysis/engine.py:115-189."""

    times: np.ndarray
    voltages: Dict[str, np.ndarray]
    currents: Dict[str, np.ndarray]
    stats: Dict[str, Any]

    @property
    def num_steps(self) -> int:
        return len(self.times)

    def voltage(self, node: str) -> np.ndarray:
        if node in self.voltages:
            return self.voltages[node]
        low = node.lower()
        for k, v in self.voltages.items():
            if k.lower() == low:
                return v
        raise KeyError(f"Node '{node}' not found. Available: {self.node_names}")

    @property
    def node_names(self):
        return list(self.voltages.keys())

    def current(self, source: str) -> np.ndarray:
        if source in self.currents:
            return self.currents[source]
        low = source.lower()
        for k, v in self.currents.items():
            if k.lower() == low:
                return v
        raise KeyError(f"Source '{source}' not found. Available: {self.source_names}")

    @property
    def source_names(self):
        return list(self.currents.keys())


def method_name(m: str) -> str:
    try:
        return _METHOD_ALIASES[str(m).lower()]
    except KeyError:
        raise ValueError(f"Unknown integration method: {m}. Supported: be, trap, gear2") from None


def _soa(a: np.ndarray) -> np.ndarray:
    """[N, k] -> contiguous int32 [k, N]."""
    return np.ascontiguousarray(np.asarray(a, dtype=np.int64).T, dtype=np.int32)


def circuit_desc(topo: Topology, net: Netlist, red=None):
    """The C ABI's vjx_circuit_desc for a topology, and the numpy buffers it
    points into (keep them alive until vjx_circuit_create returns)."""
    red = build_reduction_map(topo) if red is None else red
    n_unk, m, k = topo.n_unknowns, topo.n_vsources, net.isources.count
    nse = topo.nse
    keep = []  # keep numpy buffers alive while the C side copies them

    def hold(a):
        keep.append(a)
        return a

    slot_diag = np.zeros(nse, np.int32)
    slot_diag[topo.diag_pos] = 1
    slot_b = np.zeros(nse, np.float64)
    for key, val in (("b_p", 1.0), ("b_n", -1.0), ("bt_p", 1.0), ("bt_n", -1.0)):
        pos = topo.vs_pos.get(key, np.zeros(0, np.int32))
        pos = pos[pos >= 0]
        if np.any(slot_b[pos] != 0.0):
            raise ValueError("voltage source with both terminals on one node")
        slot_b[pos] += val
    br_ptr = np.zeros(n_unk + 1, np.int32)
    ents = [[] for _ in range(n_unk)]
    for kk in range(m):
        p = int(topo.vs_node_p[kk])
        if p > 0:
            ents[p - 1].append(kk)
    for kk in range(m):
        nn = int(topo.vs_node_n[kk])
        if nn > 0:
            ents[nn - 1].append(~kk)
    br_ptr[1:] = np.cumsum([len(e) for e in ents])
    br_ent = np.array([x for e in ents for x in e], dtype=np.int32)

    def src_table(src):
        w, par, ptr, tt, vv = pack_sources(src)
        return (L.ptr(hold(w)), L.ptr(hold(par), C.c_double),
                L.ptr(hold(ptr)) if ptr is not None else None,
                L.ptr(hold(tt), C.c_double) if ptr is not None else None,
                L.ptr(hold(vv), C.c_double) if ptr is not None else None)

    d = L.CircuitDesc()
    d.n_total = topo.n_total
    d.n_vsources = m
    d.n_isources = k
    d.nse = nse
    d.indptr = L.ptr(hold(np.ascontiguousarray(topo.indptr, np.int32)))
    d.indices = L.ptr(hold(np.ascontiguousarray(topo.indices, np.int32)))
    d.slot_diag = L.ptr(hold(slot_diag))
    d.slot_b = L.ptr(hold(slot_b), C.c_double)
    d.vs_node_p = L.ptr(hold(np.ascontiguousarray(topo.vs_node_p, np.int32)))
    d.vs_node_n = L.ptr(hold(np.ascontiguousarray(topo.vs_node_n, np.int32)))
    d.vs_wave, d.vs_params, d.vs_pwl_ptr, d.vs_pwl_t, d.vs_pwl_v = src_table(net.vsources)
    d.is_wave, d.is_params, d.is_pwl_ptr, d.is_pwl_t, d.is_pwl_v = src_table(net.isources)
    d.is_dst = L.ptr(hold(np.ascontiguousarray(red.is_dst, np.int32)))
    d.n_jac_contrib = red.n_jac_contrib
    d.n_res_contrib = red.n_res_contrib
    d.jac_seg = L.ptr(hold(np.ascontiguousarray(red.jac_seg, np.int32)))
    d.res_seg = L.ptr(hold(np.ascontiguousarray(red.res_seg, np.int32)))
    d.br_ptr = L.ptr(hold(br_ptr))
    d.br_ent = L.ptr(hold(br_ent))
    # residual test mask: internal device nodes and NOI nodes are left out
    # (solver_factories.py:86-92); NOI rows for the solve
    skip = [topo.node_names[n] - 1 for n in list(net.internal_nodes) + list(net.noi_nodes)]
    if skip:
        mask = np.ones(topo.n_unknowns, dtype=np.int32)
        mask[skip] = 0
        d.conv_mask = L.ptr(hold(mask))
    else:
        d.conv_mask = None
    noi = sorted({topo.node_names[n] - 1 for n in net.noi_nodes})
    d.n_noi = len(noi)
    d.noi_rows = L.ptr(hold(np.asarray(noi, dtype=np.int32))) if noi else None
    models = list(topo.models.items())
    if len(models) > L.MAX_MODELS:
        raise ValueError("too many model groups")
    d.n_models = len(models)
    for i, (name, ms) in enumerate(models):
        md = d.models[i]
        md.model_id = L.MODEL_IDS[name]
        md.count = ms.count
        n_jac = ms.csr_pos.shape[1]
        n_res = ms.res_idx.shape[1]
        md.v_node1 = L.ptr(hold(_soa(ms.v_node1)))
        md.v_node2 = L.ptr(hold(_soa(ms.v_node2)))
        md.jac_dst = L.ptr(hold(_soa(red.jac_dst[name].reshape(ms.count, n_jac))))
        md.res_dst = L.ptr(hold(_soa(red.res_dst[name].reshape(ms.count, n_res))))
        md.params = L.ptr(hold(param_block(name, ms.params)), C.c_double)
    return d, keep


class DeviceCircuit:
    """A circuit resident on the GPU (vjx_circuit handle)."""

    def __init__(self, topo: Topology, net: Netlist):
        self.topo = topo
        self.net = net
        self.red = build_reduction_map(topo)
        d, keep = circuit_desc(topo, net, self.red)
        h = C.c_void_p()
        L.check(L.lib.vjx_circuit_create(C.byref(d), C.byref(h)), "vjx_circuit_create")
        self.h = h
        del keep

    def __del__(self):
        lib = getattr(L, "lib", None) if L is not None else None  # gone at shutdown
        if getattr(self, "h", None) and lib is not None:
            lib.vjx_circuit_destroy(self.h)
            self.h = None

    @property
    def lu(self) -> SparseLU:
        ph = L.lib.vjx_circuit_pattern(self.h)
        return SparseLU(self.topo.indptr, self.topo.indices, _handle=ph, _owner=self)

    @property
    def n_limit_states(self) -> int:
        return int(L.lib.vjx_circuit_limit_states(self.h))

    def eval_blocks(self) -> Dict[str, int]:
        """Eval workgroup size per model group (256 unless the autotuner of
        a transient run with ``autotune`` chose 128 or 64)."""
        out = {}
        for g, name in enumerate(self.topo.models):
            b = int(L.lib.vjx_circuit_eval_block(self.h, g))
            L.check(min(b, 0), "vjx_circuit_eval_block")
            out[name] = b
        return out

    def source_values(self, t: float, device="cuda"):
        """Source values at time t evaluated on the device (vjx_eval_sources):
        (vsource_vals, isource_vals) as torch tensors."""
        import torch

        m, k = self.topo.n_vsources, self.net.isources.count
        v = torch.zeros(max(m, 1), dtype=torch.float64, device=device)
        i = torch.zeros(max(k, 1), dtype=torch.float64, device=device)
        torch.cuda.synchronize()
        L.check(L.lib.vjx_eval_sources(self.h, float(t), C.c_void_p(v.data_ptr()),
                                       C.c_void_p(i.data_ptr())), "vjx_eval_sources")
        return v[:m], i[:k]

    def build_system(self, X, vsrc, isrc, Q_prev=None, dQdt_prev=None, Q_prev2=None, c0=0.0,
                     c1=0.0, d1=0.0, c2=0.0, gmin=1e-12, gshunt=0.0, want_jacobian=True,
                     limit_state=None):
        """Device build_system on torch CUDA tensors; returns a dict of tensors
        (with ``limit_state_out`` when the circuit has limited devices;
        ``limit_state`` None means zeros)."""
        import torch

        t = self.topo
        dev = X.device
        nl = self.n_limit_states
        out = {
            "csr_data": torch.empty(t.nse, dtype=torch.float64, device=dev) if want_jacobian else None,
            "f": torch.empty(t.n_augmented, dtype=torch.float64, device=dev),
            "Q": torch.empty(t.n_unknowns, dtype=torch.float64, device=dev),
            "I_vsource": torch.empty(max(t.n_vsources, 1), dtype=torch.float64, device=dev),
            "max_res_contrib": torch.empty(t.n_unknowns, dtype=torch.float64, device=dev),
            "limit_state_out": torch.empty(nl, dtype=torch.float64, device=dev) if nl else None,
        }

        def p(x):
            return C.c_void_p(x.data_ptr()) if x is not None else None

        torch.cuda.synchronize()
        L.check(L.lib.vjx_build_system(self.h, p(X), p(vsrc), p(isrc), p(Q_prev), p(dQdt_prev),
                                       p(Q_prev2), c0, c1, d1, c2, gmin, gshunt,
                                       p(out["csr_data"]), p(out["f"]), p(out["Q"]),
                                       p(out["I_vsource"]), p(out["max_res_contrib"]),
                                       p(limit_state), p(out["limit_state_out"])),
                "vjx_build_system")
        return out

    def eval_stamp(self, model, X, vsrc, isrc=None, c0=0.0, gmin=1e-12, limit_state=None,
                   want_jacobian=True):
        """One model group's fused eval + stamp (vjx_eval_stamp): ``model`` is a
        group name of the topology or its index, or None for every group.
        Returns a dict of CUDA tensors (csr_data, f, Q, max_res_contrib,
        limit_state_out)."""
        import torch

        t = self.topo
        names = list(t.models)
        idx = -1 if model is None else (names.index(model) if isinstance(model, str) else int(model))
        dev = X.device
        nl = self.n_limit_states
        out = {
            "csr_data": torch.empty(t.nse, dtype=torch.float64, device=dev) if want_jacobian else None,
            "f": torch.empty(t.n_augmented, dtype=torch.float64, device=dev),
            "Q": torch.empty(t.n_unknowns, dtype=torch.float64, device=dev),
            "max_res_contrib": torch.empty(t.n_unknowns, dtype=torch.float64, device=dev),
            "limit_state_out": torch.zeros(nl, dtype=torch.float64, device=dev) if nl else None,
        }

        def p(x):
            return C.c_void_p(x.data_ptr()) if x is not None and x.numel() else None

        torch.cuda.synchronize()
        L.check(L.lib.vjx_eval_stamp(self.h, idx, p(X), p(vsrc), p(isrc), float(c0), float(gmin),
                                     p(out["csr_data"]), p(out["f"]), p(out["Q"]),
                                     p(out["max_res_contrib"]), p(limit_state),
                                     p(out["limit_state_out"])), "vjx_eval_stamp")
        return out

    def nr_solve(self, X_init, vsource_vals, isource_vals, Q_prev, integ_c0, gmin=1e-12,
                 gshunt=0.0, integ_c1=0.0, integ_d1=0.0, dQdt_prev=None, integ_c2=0.0,
                 Q_prev2=None, limit_state_in=None, res_tol_floor=None,
                 options: Optional[SimulationOptions] = None, precision: str = "f64",
                 refine_steps: Optional[int] = None, use_graph: bool = False):
        """The reference's nr_solve (solver_factories.py:604-727) on the device:
        torch CUDA tensors in, ``(X, iterations, converged, max_f, Q, dQdt,
        I_vsource, limit_state, max_res_contrib)`` out (tensors and Python
        scalars; ``limit_state`` None when no device is limited)."""
        import torch

        t = self.topo
        o = options or SimulationOptions()
        dev = X_init.device
        nl = self.n_limit_states
        tc = L.TranConfig()
        tc.reltol, tc.abstol, tc.vntol = o.reltol, o.abstol, o.vntol
        tc.gmin = o.gmin  # node-row diagonal regularisation; `gmin` below is $gmin
        tc.nr_damping, tc.tran_itl = o.nr_damping, int(o.tran_itl)
        tc.repivot_growth = float(o.repivot_growth)
        tc.precision = L.VJX_PREC_F32 if precision == "f32" else L.VJX_PREC_F64
        tc.refine_steps = (1 if precision == "f32" else 0) if refine_steps is None else refine_steps
        tc.use_graph = 1 if use_graph else 0

        def p(x):
            return C.c_void_p(x.data_ptr()) if x is not None else None

        def vec(n):
            return torch.empty(n, dtype=torch.float64, device=dev)

        out = {"X": vec(t.n_total + t.n_vsources), "Q": vec(t.n_unknowns),
               "dQdt": vec(t.n_unknowns), "I_vsource": vec(max(t.n_vsources, 1)),
               "limit_state": vec(nl) if nl else None, "max_res_contrib": vec(t.n_unknowns)}
        a = L.NewtonArgs()
        a.X_init, a.vsource_vals, a.isource_vals = p(X_init), p(vsource_vals), p(isource_vals)
        a.Q_prev, a.dQdt_prev, a.Q_prev2 = p(Q_prev), p(dQdt_prev), p(Q_prev2)
        a.limit_state_in, a.res_tol_floor = p(limit_state_in), p(res_tol_floor)
        a.c0, a.c1, a.d1, a.c2 = integ_c0, integ_c1, integ_d1, integ_c2
        a.gmin, a.gshunt = gmin, gshunt
        r = L.NewtonResult()
        r.X, r.Q, r.dQdt, r.I_vsource = p(out["X"]), p(out["Q"]), p(out["dQdt"]), p(out["I_vsource"])
        r.limit_state, r.max_res_contrib = p(out["limit_state"]), p(out["max_res_contrib"])
        torch.cuda.synchronize()
        L.check(L.lib.vjx_newton_solve(self.h, C.byref(tc), C.byref(a), C.byref(r)),
                "vjx_newton_solve")
        return (out["X"], int(r.iterations), bool(r.converged), float(r.max_f), out["Q"],
                out["dQdt"], out["I_vsource"][:t.n_vsources], out["limit_state"],
                out["max_res_contrib"])

    def make_nr_solve(self, options: Optional[SimulationOptions] = None,
                      max_iterations: Optional[int] = None, precision: str = "f64",
                      refine_steps: Optional[int] = None, use_graph: bool = False):
        """A Newton solver with the reference's positional signature
        ``nr_solve(X, vsource_vals, isource_vals, Q_prev, integ_c0,
        device_arrays, gmin, gshunt, integ_c1, integ_d1, dQdt_prev, integ_c2,
        Q_prev2, limit_state_in=None, res_tol_floor=None)`` -> 9-tuple
        (solver_factories.py:604-620), capped at ``max_iterations`` (default
        the options' tran_itl); numpy or CUDA-tensor inputs, CUDA tensors
        out.  ``device_arrays`` is accepted for the signature and unused: the
        device data lives in the circuit handle."""
        import torch

        o = replace(options or SimulationOptions())
        if max_iterations is not None:
            o.tran_itl = int(max_iterations)

        def dev(x):
            if x is None or isinstance(x, torch.Tensor):
                return x
            return torch.as_tensor(np.ascontiguousarray(x, dtype=np.float64), device="cuda")

        def nr_solve(X, vsource_vals, isource_vals, Q_prev, integ_c0, device_arrays=None,
                     gmin=1e-12, gshunt=0.0, integ_c1=0.0, integ_d1=0.0, dQdt_prev=None,
                     integ_c2=0.0, Q_prev2=None, limit_state_in=None, res_tol_floor=None):
            return self.nr_solve(dev(X), dev(vsource_vals), dev(isource_vals), dev(Q_prev),
                                 float(integ_c0), gmin=float(gmin), gshunt=float(gshunt),
                                 integ_c1=float(integ_c1), integ_d1=float(integ_d1),
                                 dQdt_prev=dev(dQdt_prev), integ_c2=float(integ_c2),
                                 Q_prev2=dev(Q_prev2), limit_state_in=dev(limit_state_in),
                                 res_tol_floor=dev(res_tol_floor), options=o,
                                 precision=precision, refine_steps=refine_steps,
                                 use_graph=use_graph)

        return nr_solve


def init_mid_rail(topo: Topology, net: Netlist) -> np.ndarray:
    """Initial guess for the DC solve (full_mna.py:694-745, sources.py:553-568)."""
    vs = net.vsources
    vdd = 0.0
    for i in range(vs.count):
        if vs.params["dc"][i] > vdd:
            vdd = float(vs.params["dc"][i])
    V = np.full(topo.n_total, vdd / 2.0)
    V[0] = 0.0
    for name, idx in topo.node_names.items():
        low = name.lower()
        if "vdd" in low or "vcc" in low:
            V[idx] = vdd
        elif low in ("gnd", "vss", "0"):
            V[idx] = 0.0
    for i in range(vs.count):
        p, n = int(topo.vs_node_p[i]), int(topo.vs_node_n[i])
        if n == 0 and p > 0:
            V[p] = float(vs.params["dc"][i])
    return V


class CircuitEngine:
    """Parse / prepare / run_transient on a flattened netlist."""

    def __init__(self, netlist):
        """``netlist``: a :class:`Netlist`, or the path of a VACASK ``.sim``
        file (parsed, flattened and mapped by :mod:`vajax_amd.vacask`, as
        ``CircuitEngine(sim_path)`` in engine.py:304)."""
        if isinstance(netlist, (str, bytes)) or hasattr(netlist, "__fspath__"):
            from .vacask import load

            self.sim_path = netlist
            netlist = load(netlist)
        if not isinstance(netlist, Netlist):
            raise TypeError("CircuitEngine expects a Netlist or a .sim path")
        self.netlist = netlist
        self.options = SimulationOptions()
        self.analysis_params: Dict[str, Any] = dict(netlist.analysis)
        for k in ("icmode", "reltol", "abstol", "vntol", "gmin", "tran_itl"):
            if k in netlist.options:
                cur = getattr(self.options, k)
                v = netlist.options[k]
                setattr(self.options, k, type(cur)(v) if not isinstance(cur, str) else str(v))
        self.topology: Optional[Topology] = None
        self.device: Optional[DeviceCircuit] = None
        self.node_names: Dict[str, int] = {}
        self._prepared = False

    def parse(self):
        self.topology = build_topology(self.netlist)
        self.node_names = dict(self.topology.node_names)
        self.num_nodes = self.topology.n_total
        self.device = DeviceCircuit(self.topology, self.netlist)
        self._prepared = False
        return self

    def prepare(self, *, t_stop: Optional[float] = None, dt: Optional[float] = None,
                adaptive_config: Optional[AdaptiveConfig] = None, precision: str = "f64",
                refine_steps: Optional[int] = None, max_steps: Optional[int] = None,
                use_graph: bool = False, band_width: Optional[int] = None,
                autotune: bool = False, nd_leaf: Optional[int] = None,
                checkpoint_interval: Optional[int] = None, replay_times=None,
                window_from_step: int = 0, **_ignored) -> None:
        """``checkpoint_interval``: output rows kept on the device before they
        are copied to the host (full_mna.py:784-1000 checkpoint mode): None
        keeps all rows resident unless they would take more than
        ``OUTPUT_RESIDENT_BYTES``, then chunks of that size; <= 0 keeps all
        rows resident; k > 0 chunks of k rows.  ``replay_times``: accepted
        times of another run to step through exactly (test harness, adaptive
        loop).  ``window_from_step``: the statistics ``window_*`` cover the
        steps from this accepted-step count on."""
        if self.topology is None:
            self.parse()
        if nd_leaf is not None:
            # nested-dissection leaf size of the LU ordering (fill versus
            # depth of the elimination); a change re-runs the analysis
            self.device.lu.set_ordering("mlnd", nd_leaf)
        if band_width is not None:
            # launch plan of the triangular solves (levels per band; 0 = one
            # launch per level); a change re-runs the symbolic analysis
            self.device.lu.set_band_width(band_width)
        if t_stop is None:
            t_stop = float(self.analysis_params.get("stop", 1e-3))
        if dt is None:
            dt = float(self.analysis_params.get("step", 1e-6))
        if max_steps is None:
            max_steps = int(t_stop / dt * 1.1) + 10
        kw = {}
        opts = self.netlist.options
        mapping = {"tran_lteratio": "lte_ratio", "tran_redofactor": "redo_factor",
                   "tran_gshunt": "gshunt_init", "reltol": "reltol", "abstol": "abstol",
                   "tran_fs": "tran_fs", "tran_ft": "tran_ft", "tran_minpts": "tran_minpts",
                   "maxstep": "max_dt", "tran_method": "integration_method"}
        for src, dst in mapping.items():
            if src in opts:
                kw[dst] = opts[src]
        cfg = AdaptiveConfig(**kw)
        if adaptive_config is not None:
            defaults = AdaptiveConfig()
            over = {f.name: getattr(adaptive_config, f.name) for f in fields(AdaptiveConfig)
                    if getattr(adaptive_config, f.name) != getattr(defaults, f.name)}
            if over:
                cfg = replace(cfg, **over)
        if precision not in ("f64", "f32"):
            raise ValueError("precision must be 'f64' or 'f32'")
        topo = self.topology
        row_bytes = 8 * (topo.n_total + max(topo.n_vsources, 1))
        if checkpoint_interval is None:
            fits = max_steps * row_bytes <= OUTPUT_RESIDENT_BYTES
            chunk = 0 if fits else max(1, OUTPUT_RESIDENT_BYTES // row_bytes)
        else:
            chunk = max(int(checkpoint_interval), 0)
        self._prep = dict(output_chunk=int(chunk), window_from_step=int(window_from_step))
        if replay_times is not None:
            rt = np.ascontiguousarray(replay_times, dtype=np.float64)
            self._prep["replay_times"] = rt
            max_steps = max(int(max_steps), len(rt))
        self._prep.update(t_stop=float(t_stop), dt=float(dt), cfg=cfg, precision=precision,
                          refine_steps=(1 if precision == "f32" else 0) if refine_steps is None
                          else int(refine_steps), max_steps=int(max_steps), use_graph=use_graph,
                          autotune=bool(autotune))
        self._prepared = True

    def _tran_config(self) -> L.TranConfig:
        p = self._prep
        cfg: AdaptiveConfig = p["cfg"]
        o = self.options
        tc = L.TranConfig()
        tc.t_stop = p["t_stop"]
        tc.dt = p["dt"]
        tc.max_steps = p["max_steps"]
        tc.method = L.VJX_METHOD[method_name(cfg.integration_method)]
        tc.fixed_step = 1 if o.fixed_step_transient else 0
        tc.uic = 1 if str(o.icmode).lower() == "uic" else 0
        tc.precision = L.VJX_PREC_F32 if p["precision"] == "f32" else L.VJX_PREC_F64
        tc.refine_steps = p["refine_steps"]
        tc.tran_itl = o.tran_itl
        tc.use_graph = 1 if p["use_graph"] else 0
        tc.autotune = 1 if p.get("autotune") else 0
        tc.repivot_growth = float(o.repivot_growth)
        tc.output_chunk = p.get("output_chunk", 0)
        tc.window_from_step = p.get("window_from_step", 0)
        rt = p.get("replay_times")
        if rt is not None:
            tc.replay_times_host = L.ptr(rt, C.c_double)
            tc.n_replay = len(rt)
        tc.reltol = o.reltol
        tc.abstol = o.abstol
        tc.vntol = o.vntol
        tc.gmin = o.gmin
        tc.nr_damping = o.nr_damping
        tc.lte_ratio = cfg.lte_ratio
        tc.redo_factor = cfg.redo_factor
        tc.min_dt = cfg.min_dt
        tc.max_dt = cfg.max_dt
        tc.tran_minpts = cfg.tran_minpts
        tc.warmup_steps = cfg.warmup_steps
        tc.max_order = cfg.max_order
        tc.grow_factor = cfg.grow_factor
        tc.gshunt_init = cfg.gshunt_init
        tc.gshunt_target = cfg.gshunt_target
        tc.gshunt_steps = cfg.gshunt_steps
        tc.tran_fs = cfg.tran_fs
        tc.tran_ft = cfg.tran_ft
        tc.max_consecutive_rejects = cfg.max_consecutive_rejects
        if cfg.reltol != o.reltol or cfg.abstol != o.abstol:
            # the LTE uses the adaptive config's tolerances, NR the options';
            # one pair is carried, so they must agree
            raise ValueError("AdaptiveConfig and SimulationOptions tolerances differ")
        return tc

    def compute_dc_operating_point(self, precision: str = "f64", use_graph: bool = False):
        """DC operating point on the device: direct Newton (op_itl iterations)
        from the starting voltages of dc_operating_point.py:20-122, then the
        homotopy chain of the options when it fails
        (dc_operating_point.py:125-242, engine.py:1195-1228).  Returns the
        node voltages [n_total]; ``self.dc_result`` keeps the full solution
        and how it was reached."""
        from .dc import compute_dc_operating_point, get_dc_source_values, get_vdd_value
        from .dc import initialize_dc_voltages

        if self.topology is None:
            self.parse()
        topo = self.topology
        V0, _ = initialize_dc_voltages(topo, self.netlist, get_vdd_value(self.netlist))
        X0 = np.zeros(topo.n_total + topo.n_vsources)
        X0[:topo.n_total] = V0
        vs, isr = get_dc_source_values(self.netlist)
        nr = self.device.make_nr_solve(self.options, self.options.op_itl, precision=precision,
                                       use_graph=use_graph)
        X, res = compute_dc_operating_point(nr, X0, vs, isr, np.zeros(topo.n_unknowns),
                                            self.options)
        X = X.cpu().numpy() if hasattr(X, "cpu") else np.asarray(X)
        res.V = X
        self.dc_result = res
        return X[:topo.n_total]

    def run_transient(self) -> TransientResult:
        if not self._prepared:
            self.prepare()
        topo = self.topology
        tc = self._tran_config()
        ms = tc.max_steps
        n_total, m = topo.n_total, topo.n_vsources
        times = np.zeros(ms)
        V = np.zeros((ms, n_total))
        I = np.zeros((ms, max(m, 1)))
        v0 = np.ascontiguousarray(init_mid_rail(topo, self.netlist))
        tc.v_init_host = L.ptr(v0, C.c_double)
        res = L.TranResult()
        res.times = L.ptr(times, C.c_double)
        res.V_out = L.ptr(V, C.c_double)
        res.I_out = L.ptr(I, C.c_double)
        rc = L.check(L.lib.vjx_transient_run(self.device.h, C.byref(tc), C.byref(res)),
                     "vjx_transient_run")
        n = res.n_steps
        stats = {
            "n_steps": n,
            "total_timesteps": n,
            "accepted_steps": n if not tc.fixed_step else n - 1,
            "rejected_steps": res.rejected_steps,
            "total_nr_iterations": int(res.total_nr_iterations),
            "total_evals": int(res.total_evals),
            "avg_nr_iterations": res.total_nr_iterations / max(n, 1),
            "wall_time": res.wall_time,
            "time_per_step_ms": res.wall_time / n * 1000 if n else 0.0,
            "min_dt_used": res.min_dt_used,
            "max_dt_used": res.max_dt_used,
            "convergence_rate": n / max(n + res.rejected_steps, 1),
            "dc_converged": bool(res.dc_converged),
            "dc_iterations": res.dc_iterations,
            "dc_time": res.dc_time,
            "repivots": res.repivots,
            "window_steps": res.window_steps,
            "window_rejected": res.window_rejected,
            "window_nr_iterations": int(res.window_nr_iterations),
            "window_time": res.window_time,
            "strategy": "fixed_step_full_mna" if tc.fixed_step else "adaptive_full_mna",
            "precision": self._prep["precision"],
            "return_code": rc,
            "V_out": V[:n],
            "I_out": I[:n],
        }
        voltages = {name: V[:n, idx] for name, idx in topo.node_names.items() if idx > 0}
        currents = {name: I[:n, i] for i, name in enumerate(self.netlist.vsources.names)}
        return TransientResult(times=times[:n], voltages=voltages, currents=currents, stats=stats)
