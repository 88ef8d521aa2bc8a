"""DC operating point: initial voltages, a direct Newton solve and, when it
fails, the homotopy chain (vajax/analysis/dc_operating_point.py:20-242,
vajax/analysis/sources.py:519-568)."""

from __future__ import annotations

from typing import Callable, Optional, Tuple

import numpy as np

from .homotopy import HomotopyConfig, HomotopyResult, run_homotopy_chain
from .netlist import Netlist


def get_dc_source_values(net: Netlist) -> Tuple[np.ndarray, np.ndarray]:
    """DC values of the sources (sources.py:519-550): a vsource's ``dc``
    parameter whatever its waveform; an isource's ``val0`` when it is a pulse,
    else its ``dc``."""
    vs, isr = net.vsources, net.isources
    v = np.array([float(vs.params["dc"][i]) for i in range(vs.count)], dtype=np.float64)
    i = np.array([float(isr.params["val0"][j] if isr.wave[j] else isr.params["dc"][j])
                  for j in range(isr.count)], dtype=np.float64)
    return v, i


def get_vdd_value(net: Netlist) -> float:
    """Largest vsource ``dc`` value, at least 0 (sources.py:553-568)."""
    vdd = 0.0
    for x in net.vsources.params["dc"]:
        if x > vdd:
            vdd = float(x)
    return vdd


def initialize_dc_voltages(topo, net: Netlist, vdd: float) -> Tuple[np.ndarray, Optional[np.ndarray]]:
    """Starting voltages (dc_operating_point.py:20-122): every node at vdd/2,
    ground 0, names containing vdd/vcc at vdd, gnd/vss at 0, the positive node
    of each vsource at its negative node's voltage plus its ``dc`` value (in
    source order), NOI nodes at 0.  Returns (V[n_total], NOI node indices)."""
    V = np.full(topo.n_total, vdd / 2.0)
    V[0] = 0.0
    for name, idx in topo.node_names.items():
        low = name.lower()
        if "vdd" in low or "vcc" in low:
            V[idx] = vdd
        elif low in ("gnd", "vss", "0"):
            V[idx] = 0.0
    vs = net.vsources
    for i in range(vs.count):
        p, n = int(topo.vs_node_p[i]), int(topo.vs_node_n[i])
        if p > 0:
            V[p] = float(vs.params["dc"][i]) + (0.0 if n == 0 else V[n])
    noi = [topo.node_names[n] for n in net.noi_nodes if n in topo.node_names]
    if noi:
        V[noi] = 0.0
        return V, np.array(noi, dtype=np.int64)
    return V, None


def homotopy_config_from_options(options) -> HomotopyConfig:
    """dc_operating_point.py:189-205: the chain's parameters from the options."""
    return HomotopyConfig(
        gmin=options.gmin,
        gdev_start=options.homotopy_startgmin,
        gdev_target=options.homotopy_mingmin,
        gmin_factor=options.homotopy_gminfactor,
        gmin_factor_min=options.homotopy_mingminfactor,
        gmin_factor_max=options.homotopy_maxgminfactor,
        gmin_max=options.homotopy_maxgmin,
        gmin_max_steps=options.homotopy_gminsteps,
        source_step=options.homotopy_srcstep,
        source_step_min=options.homotopy_minsrcstep,
        source_scale=options.homotopy_srcscale,
        source_max_steps=options.homotopy_srcsteps,
        chain=tuple(options.op_homotopy),
        max_iterations=options.op_itlcont,
        abstol=options.abstol,
        debug=0,
    )


def compute_dc_operating_point(nr_solve: Callable, X_init, vsource_dc_vals, isource_dc_vals,
                               Q_prev, options, device_arrays=None):
    """Direct Newton from X_init; on failure the homotopy chain, whose best
    point is kept even when it fails (dc_operating_point.py:125-242).  Returns
    (X, HomotopyResult) -- the result records the direct solve as method
    "direct" when no homotopy ran."""
    out = nr_solve(X_init, vsource_dc_vals, isource_dc_vals, Q_prev, 0.0, device_arrays,
                   options.gmin, 0.0, 0.0, 0.0, None, 0.0, None)
    X, iters, conv = out[0], int(out[1]), bool(out[2])
    if conv:
        return X, HomotopyResult(True, X, "direct", iters, 0, options.gmin, 1.0)
    cfg = homotopy_config_from_options(options)
    r = run_homotopy_chain(nr_solve, X_init, vsource_dc_vals, isource_dc_vals, Q_prev,
                           device_arrays, cfg)
    r.iterations += iters
    return r.V, r
