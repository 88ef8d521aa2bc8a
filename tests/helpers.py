"""Shared helpers: run the same circuit through the oracle and the GPU path."""

from __future__ import annotations

import numpy as np

from oracle import mna as omna
from oracle import transient as otr


def oracle_configs(net, **over):
    method = str(net.options.get("tran_method", "trap"))
    cfg = otr.AdaptiveConfig(integration_method=otr.method_of(method))
    for k, v in over.items():
        if hasattr(cfg, k):
            setattr(cfg, k, v)
    opts = otr.Options(icmode=str(net.options.get("icmode", "op")))
    return cfg, opts


def run_oracle(net, t_stop, dt, fixed=False, tran_fs=None, dense=False, max_steps=None,
               tran_itl=None, **over):
    cir = omna.setup(net)
    cfg, opts = oracle_configs(net, **over)
    if tran_itl is not None:
        opts.tran_itl = int(tran_itl)
    if tran_fs is not None:
        cfg.tran_fs = tran_fs
    if fixed:
        out = otr.run_fixed(cir, t_stop, dt, cfg.integration_method, opts, tran_fs=cfg.tran_fs,
                            dense=dense)
    else:
        out = otr.run_adaptive(cir, t_stop, dt, cfg, opts, dense=dense, max_steps=max_steps)
    return cir, out


def run_gpu(net, t_stop, dt, fixed=False, tran_fs=None, precision="f64", refine_steps=None,
            use_graph=False, band_width=None, autotune=False, max_steps=None, tran_itl=None,
            repivot_growth=None, checkpoint_interval=None, replay_times=None,
            window_from_step=0, **over):
    from vajax_amd.engine import AdaptiveConfig, CircuitEngine

    eng = CircuitEngine(net)
    eng.parse()
    eng.options.fixed_step_transient = fixed
    if tran_itl is not None:
        eng.options.tran_itl = int(tran_itl)
    if repivot_growth is not None:
        eng.options.repivot_growth = float(repivot_growth)
    cfg = AdaptiveConfig(integration_method=str(net.options.get("tran_method", "trap")))
    if tran_fs is not None:
        cfg.tran_fs = tran_fs
    for k, v in over.items():
        setattr(cfg, k, v)
    eng.prepare(t_stop=t_stop, dt=dt, adaptive_config=cfg, precision=precision,
                refine_steps=refine_steps, use_graph=use_graph, band_width=band_width,
                autotune=autotune, max_steps=max_steps, checkpoint_interval=checkpoint_interval,
                replay_times=replay_times, window_from_step=window_from_step)
    return eng, eng.run_transient()


def waveform_error(times_a, V_a, times_b, V_b):
    """max |V_a - V_b| after interpolating b onto a's time points (per node)."""
    err = 0.0
    for j in range(V_a.shape[1]):
        vb = np.interp(times_a, times_b, V_b[:, j])
        err = max(err, float(np.max(np.abs(V_a[:, j] - vb))))
    return err


def grid_deviation(times, ref_times):
    """Largest relative difference between two accepted time grids of the
    same length (0 where both are 0)."""
    t = np.asarray(times, dtype=np.float64)
    r = np.asarray(ref_times, dtype=np.float64)
    scale = np.maximum(np.abs(r), 1e-300)
    return float(np.max(np.abs(t - r) / scale))
