"""Homotopy continuation for the DC operating point: gmin stepping (device
$gmin "gdev" or shunt "gshunt") and source stepping, chained gdev -> gshunt
-> src (vajax/analysis/homotopy.py:36-567, after VACASK lib/hmtpgmin.cpp
and lib/hmtpsrc.cpp).

The step control runs on the host; every continuation step is one call of a
Newton solver with the reference's ``nr_solve`` positional signature
``nr_solve(X, vsource_vals, isource_vals, Q_prev, integ_c0, device_arrays,
gmin, gshunt, integ_c1, integ_d1, dQdt_prev, integ_c2, Q_prev2)`` returning
the 9-tuple ``(X, iterations, converged, max_f, Q, dQdt, I_vsource,
limit_state, max_res_contrib)``.  ``DeviceCircuit.make_nr_solve`` gives the
device one (``vjx_newton_solve``): vectors stay on the GPU between steps and
only the iteration count and the converged flag come back.
"""

from __future__ import annotations

import math
from dataclasses import dataclass
from typing import Any, Callable, Tuple

# Newton iterations at or below this fraction of max_iterations count as fast
# convergence, above the second as slow (homotopy.py:28-32)
FAST_FRACTION = 0.25
SLOW_FRACTION = 0.75


@dataclass
class HomotopyConfig:
    """Step-control parameters (homotopy.py:36-69; VACASK lib/options.cpp)."""

    gmin: float = 1e-12
    gdev_start: float = 1e-3
    gdev_target: float = 1e-13
    gmin_factor: float = 10.0
    gmin_factor_min: float = 1.1
    gmin_factor_max: float = 100.0
    gmin_max: float = 1.0
    gmin_max_steps: int = 100
    source_step: float = 0.1
    source_step_min: float = 0.001
    source_scale: float = 2.0
    source_max_steps: int = 100
    chain: Tuple[str, ...] = ("gdev", "gshunt", "src")
    max_iterations: int = 100  # the Newton cap the fast/slow thresholds refer to
    abstol: float = 1e-9
    debug: int = 0


@dataclass
class HomotopyResult:
    """homotopy.py:72-81."""

    converged: bool
    V: Any
    method: str = ""
    iterations: int = 0
    homotopy_steps: int = 0
    final_gmin: float = 0.0
    final_source_scale: float = 1.0


def _say(cfg: HomotopyConfig, level: int, msg: str) -> None:
    if cfg.debug >= level:
        print(msg, flush=True)


def _dc_solve(nr_solve, X, vs, isv, Q_prev, device_arrays, gmin, gshunt):
    """One DC Newton solve (integ_c0 = 0, no history); returns (X, iters, conv)."""
    out = nr_solve(X, vs, isv, Q_prev, 0.0, device_arrays, gmin, gshunt, 0.0, 0.0, None, 0.0, None)
    return out[0], int(out[1]), bool(out[2])


def _speed(cfg: HomotopyConfig, iters: int) -> int:
    """-1 fast, +1 slow, 0 otherwise, against the Newton cap of the config."""
    if iters <= int(cfg.max_iterations * FAST_FRACTION):
        return -1
    if iters > int(cfg.max_iterations * SLOW_FRACTION):
        return 1
    return 0


def gmin_stepping(nr_solve: Callable, V_init, vsource_vals, isource_vals, Q_prev,
                  device_arrays, source_scale: float, config: HomotopyConfig,
                  mode: str = "gdev") -> HomotopyResult:
    """Adaptive gmin stepping (homotopy.py:90-276).

    The extra conductance starts at ``gdev_start`` and falls by ``factor``
    per converged step down to ``gdev_target``; slow convergence shrinks the
    factor to its square root, a failure before the first success raises the
    conductance (giving up above ``gmin_max``), a failure after it backtracks
    to the last good point with ``factor ** 0.25`` (giving up below
    ``gmin_factor_min``).  ``mode`` "gdev" adds it to the device simparam
    $gmin, "gshunt" as a shunt from every node to ground.  A final solve at
    the plain gmin follows any success.
    """
    g = config.gdev_start
    target = config.gdev_target
    factor = config.gmin_factor
    vs = vsource_vals * source_scale
    isv = isource_vals * source_scale
    V, V_good, g_good = V_init, V_init, g
    reached = False
    iters_total = 0
    steps = 0
    _say(config, 1, f"Homotopy: {mode} stepping from {g:.2e}")
    for _ in range(config.gmin_max_steps):
        steps += 1
        if mode == "gdev":
            gmin, gshunt = config.gmin + g, 0.0
        else:
            gmin, gshunt = config.gmin, g
        X, iters, conv = _dc_solve(nr_solve, V, vs, isv, Q_prev, device_arrays, gmin, gshunt)
        iters_total += iters
        if conv:
            reached = True
            V_good, g_good = X, g
            _say(config, 1, f"Homotopy: {mode}={g:.2e} converged in {iters} iterations")
            if g <= target:
                break
            if _speed(config, iters) > 0:
                factor = max(math.sqrt(factor), config.gmin_factor_min)
            if g / factor < target:
                factor = g / target
                g = target
            else:
                g = g / factor
            V = X
            continue
        _say(config, 1, f"Homotopy: {mode}={g:.2e} failed in {iters} iterations")
        if not reached:
            g = g * factor
            if g > config.gmin_max:
                return HomotopyResult(False, V_init, f"{mode}_stepping", iters_total, steps,
                                      final_gmin=g)
        else:
            factor = factor ** 0.25
            if factor < config.gmin_factor_min:
                break
            V, g = V_good, g_good
    if not reached:
        return HomotopyResult(False, V_good, f"{mode}_stepping", iters_total, steps, final_gmin=g)
    X, iters, conv = _dc_solve(nr_solve, V_good, vs, isv, Q_prev, device_arrays, config.gmin, 0.0)
    iters_total += iters
    steps += 1
    return HomotopyResult(conv, X if conv else V_good, f"{mode}_stepping", iters_total, steps,
                          final_gmin=config.gmin)


def source_stepping(nr_solve: Callable, V_init, vsource_vals, isource_vals, Q_prev,
                    device_arrays, config: HomotopyConfig) -> HomotopyResult:
    """Adaptive source stepping with a gmin fallback at zero sources
    (homotopy.py:279-467).

    Solves with all sources off (falling back to gdev, then gshunt stepping
    there), then raises the source factor by ``raise`` per converged step up
    to 1: fast convergence multiplies ``raise`` by ``source_scale``, slow
    convergence divides it (not below ``source_step_min``), a failure halves
    it (giving up below ``source_step_min``).
    """
    step = config.source_step
    good = 0.0
    iters_total = 0
    steps = 0
    _say(config, 1, "Homotopy: source stepping")
    X, iters, conv = _dc_solve(nr_solve, V_init, vsource_vals * 0.0, isource_vals * 0.0, Q_prev,
                               device_arrays, config.gmin, 0.0)
    iters_total += iters
    steps += 1
    if conv:
        V_good = X
    else:
        r = gmin_stepping(nr_solve, V_init, vsource_vals, isource_vals, Q_prev, device_arrays,
                          0.0, config, "gdev")
        iters_total += r.iterations
        steps += r.homotopy_steps
        if not r.converged:
            r = gmin_stepping(nr_solve, V_init, vsource_vals, isource_vals, Q_prev,
                              device_arrays, 0.0, config, "gshunt")
            iters_total += r.iterations
            steps += r.homotopy_steps
        if not r.converged:
            return HomotopyResult(False, V_init, "source_stepping", iters_total, steps,
                                  final_source_scale=0.0)
        V_good = r.V
    for _ in range(config.source_max_steps):
        f = min(good + step, 1.0)
        X, iters, conv = _dc_solve(nr_solve, V_good, vsource_vals * f, isource_vals * f, Q_prev,
                                   device_arrays, config.gmin, 0.0)
        iters_total += iters
        steps += 1
        if conv:
            V_good, good = X, f
            _say(config, 1, f"Homotopy: srcfact={f:.4f} converged in {iters} iterations")
            if good >= 1.0:
                return HomotopyResult(True, V_good, "source_stepping", iters_total, steps,
                                      final_source_scale=1.0)
            speed = _speed(config, iters)
            if speed < 0:
                step *= config.source_scale
            elif speed > 0:
                step = max(step / config.source_scale, config.source_step_min)
        else:
            _say(config, 1, f"Homotopy: srcfact={f:.4f} failed in {iters} iterations")
            step *= 0.5
            if step < config.source_step_min:
                break
    return HomotopyResult(good >= 1.0, V_good, "source_stepping", iters_total, steps,
                          final_source_scale=good)


def run_homotopy_chain(nr_solve: Callable, V_init, vsource_vals, isource_vals, Q_prev,
                       device_arrays, config: HomotopyConfig) -> HomotopyResult:
    """Try the algorithms of ``config.chain`` in order, each from the best
    point of the one before (homotopy.py:470-567)."""
    V = V_init
    iters_total = 0
    steps = 0
    for algo in config.chain:
        if algo in ("gdev", "gshunt"):
            r = gmin_stepping(nr_solve, V, vsource_vals, isource_vals, Q_prev, device_arrays,
                              1.0, config, algo)
        elif algo == "src":
            r = source_stepping(nr_solve, V, vsource_vals, isource_vals, Q_prev, device_arrays,
                                config)
        else:
            _say(config, 1, f"Homotopy: unknown algorithm {algo!r} skipped")
            continue
        iters_total += r.iterations
        steps += r.homotopy_steps
        if r.converged:
            return HomotopyResult(True, r.V, f"chain_{r.method}", iters_total, steps,
                                  r.final_gmin, r.final_source_scale)
        V = r.V
    return HomotopyResult(False, V, "chain_failed", iters_total, steps)
