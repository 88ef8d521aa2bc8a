"""Newton solver and adaptive / fixed-step transient loops (oracle).

* ``nr_solve`` -- while-loop NR of solver_factories.py:230-325 and 574-727:
  build -> residual test (|f| < max(max_res_contrib*reltol, floor) on node
  rows, < vntol on branch rows) -> solve J d = -f -> delta test
  (|d| < max(|X|*reltol, vntol | abstol), skipped at iteration 0) -> X += d;
  stop when both tests pass or after ``max_iterations``; then one more build
  at X_final for Q, I_vsource and max_res_contrib, and
  dQdt = c0*Q + c1*Q_prev + d1*dQdt_prev + c2*Q_prev2.  Internal device nodes
  are left out of the residual test and of max_f; noise-only internal (NOI)
  nodes get identity rows and columns and a zero right-hand side in the
  solve, no delta test, and are clamped to 0 V after the update
  (solver_factories.py:37-137, 271-313, 1268-1276).  With limited devices
  the limit state of each iteration's build feeds the next, and convergence
  also needs max|ls_out - ls_in| < max(max|ls_in|*reltol, vntol) and an
  iteration index >= 1 (solver_factories.py:317-323); the final build reads
  the last iteration's state.
* ``predict`` / ``lte_timestep`` -- adaptive.py:94-218 and 221-339.
* ``run_adaptive`` -- FullMNAStrategy.run (full_mna.py:780-1196) and body_fn
  (full_mna.py:1840-2194), including tran_minpts / tran_fs handling, the
  first-step zero history, warm-up, NR-failure cuts, min_dt force-accept,
  LTE rejection with the consecutive-reject cap, and the output rule.  The
  limit state starts at zero (full_mna.py:1028-1033; the DC solve's is not
  kept) and advances only with a new solution (full_mna.py:2162-2167).
* ``run_fixed`` -- _run_fixed_step (full_mna.py:1198-1441).

The linear solve is SuperLU (scipy, full partial pivoting) on the CSR matrix -- the sparse paths of
the reference add no regularisation (solver_factories.py:1225-1297) -- or a
dense LU with the dense path's 1e-14*I (solver_factories.py:784-788) when
``dense=True``.
"""

from __future__ import annotations

import math
import time
from dataclasses import dataclass, field, replace
from typing import Dict, Optional

import numpy as np
import scipy.sparse.linalg as spla

from . import mna
from .sources import source_values

BE, TRAP, GEAR2 = "be", "trap", "gear2"
_METHODS = {"be": BE, "euler": BE, "backward_euler": BE, "trap": TRAP, "trapezoidal": TRAP,
            "am2": TRAP, "gear2": GEAR2, "gear": GEAR2, "bdf2": GEAR2, "bdf": GEAR2}


@dataclass
class Options:
    """SimulationOptions defaults used on this path (options.py:35-192)."""

    reltol: float = 1e-3
    abstol: float = 1e-12
    vntol: float = 1e-6
    gmin: float = 1e-12
    tran_itl: int = 20
    nr_damping: float = 1.0
    icmode: str = "op"


@dataclass
class AdaptiveConfig:
    """adaptive.py:37-91 defaults."""

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
    gshunt_init: float = 0.0
    gshunt_steps: int = 5
    gshunt_target: float = 0.0
    tran_fs: float = 0.25
    tran_ft: float = 0.25
    integration_method: str = TRAP
    max_consecutive_rejects: int = 5


def method_of(name: str) -> str:
    return _METHODS[str(name).lower()]


def integration_coeffs(method: str, dt_cur: float, dt_prev: float):
    """(c0, c1, d1, c2, error_coeff) of the adaptive loop (full_mna.py:1852-1891).

    dQ/dt ~ c0*Q_new + c1*Q_prev + d1*dQdt_prev + c2*Q_prev2; gear2 is the
    variable-step BDF2 with omega = dt_cur / dt_prev.
    """
    inv = 1.0 / dt_cur
    if method == BE:
        return inv, -inv, 0.0, 0.0, -0.5
    if method == TRAP:
        return 2.0 * inv, -2.0 * inv, -1.0, 0.0, 1.0 / 24.0
    om = dt_cur / dt_prev if dt_prev > 0 else 1.0
    op1 = 1.0 + om
    c0 = (1.0 + 2.0 * om) / (op1 * dt_cur)
    c1 = -op1 / dt_cur
    c2 = (om * om) / (op1 * dt_cur)
    return c0, c1, 0.0, c2, -2.0 / 9.0


class LinearSolver:
    def __init__(self, cir, dense: bool = False):
        self.cir = cir
        self.dense = dense

    def solve(self, csr_data, f):
        """delta = -J^-1 f.  A non-finite or singular system gives a NaN
        update instead of an exception: UMFPACK's singular-matrix warning is
        tolerated and its solution is returned (umfpack_ffi.cpp:297, 320), so
        the iteration simply fails the convergence tests."""
        if not (np.all(np.isfinite(csr_data)) and np.all(np.isfinite(f))):
            return np.full(f.shape, np.nan)
        if self.dense:
            J = mna.csr_matrix(self.cir, csr_data).toarray()
            J = J + 1e-14 * np.eye(J.shape[0])
            try:
                return np.linalg.solve(J, -f)
            except np.linalg.LinAlgError:
                return np.full(f.shape, np.nan)
        A = mna.csr_matrix(self.cir, csr_data).tocsc()
        # full partial pivoting: CMOS DC points leave cut-off nodes tied only
        # by gmin, where a 0.1 diagonal threshold can meet an exact zero pivot
        try:
            lu = spla.splu(A, permc_spec="COLAMD", diag_pivot_thresh=1.0)
        except RuntimeError:
            if A.shape[0] > 6000:
                raise
            try:
                return np.linalg.solve(A.toarray(), -f)
            except np.linalg.LinAlgError:
                return np.full(f.shape, np.nan)
        return lu.solve(-f)


@dataclass
class NRResult:
    X: np.ndarray
    iterations: int
    converged: bool
    max_f: float
    Q: np.ndarray
    dQdt: np.ndarray
    I_vsource: np.ndarray
    max_res_contrib: np.ndarray
    lim: np.ndarray = None  # limit state after the last iteration


def nr_solve(cir, solver: LinearSolver, opts: Options, X_init, vsrc, isrc, Q_prev, c0,
             gshunt=0.0, c1=0.0, d1=0.0, dQdt_prev=None, c2=0.0, Q_prev2=None,
             res_tol_floor=None, max_iterations=None, lim_in=None) -> NRResult:
    n_unk, m = cir.n_unk, cir.m
    n_total = cir.n_total
    max_it = opts.tran_itl if max_iterations is None else max_iterations
    dQ = dQdt_prev if dQdt_prev is not None else np.zeros(n_unk)
    Q2 = Q_prev2 if Q_prev2 is not None else np.zeros(n_unk)
    floor = res_tol_floor if res_tol_floor is not None else np.full(n_unk, opts.abstol)
    delta_abs = np.concatenate([np.full(n_unk, opts.vntol), np.full(m, opts.abstol)])
    X = np.array(X_init, dtype=np.float64)
    lim = np.zeros(cir.n_lim) if lim_in is None else np.array(lim_in, dtype=np.float64)
    it = 0
    converged = False
    max_f = math.inf
    while (not converged) and it < max_it:
        csr, f, Q, _, mrc, lim_out = mna.build_system_lim(cir, X, vsrc, isrc, Q_prev, c0, gshunt,
                                                          c1, d1, dQ, c2, Q2, lim)
        res_tol = np.concatenate([np.maximum(mrc * opts.reltol, floor), np.full(m, opts.vntol)])
        f_check = f
        if cir.conv_mask is not None and not cir.conv_mask.all():
            f_check = np.where(np.concatenate([cir.conv_mask, np.ones(m, bool)]), f, 0.0)
        max_f = float(np.max(np.abs(f_check))) if f.size else 0.0
        res_ok = bool(np.all(np.abs(f_check) < res_tol))
        noi = cir.noi_res if cir.noi_res is not None and len(cir.noi_res) else None
        if noi is not None:
            csr, f = _apply_noi(cir, csr, f, noi)
        delta = solver.solve(csr, f)
        conv_delta = np.concatenate([delta[:n_unk] * opts.nr_damping, delta[n_unk:]])
        if noi is not None:
            conv_delta[noi] = 0.0
        X_ref = np.concatenate([X[1:n_total], X[n_total:]])
        tol = np.maximum(np.abs(X_ref) * opts.reltol, delta_abs)
        delta_ok = (it == 0) or bool(np.all(np.abs(conv_delta) < tol))
        Vd = delta[:n_unk]
        max_vd = float(np.max(np.abs(Vd))) if n_unk else 0.0
        v_scale = min(1.0, 1e30 / max(max_vd, 1e-30))
        X = X.copy()
        X[1:n_total] += Vd * v_scale * opts.nr_damping
        X[n_total:] += delta[n_unk:]
        if noi is not None:
            X[noi + 1] = 0.0
        converged = res_ok and delta_ok
        if cir.n_lim:
            lim_delta = float(np.max(np.abs(lim_out - lim)))
            lim_ref = max(float(np.max(np.abs(lim))) * opts.reltol, opts.vntol)
            converged = converged and lim_delta < lim_ref and it >= 1
        lim = lim_out
        it += 1
    _, _, Qf, I_vs, mrc, _ = mna.build_system_lim(cir, X, vsrc, isrc, Q_prev, c0, gshunt, c1, d1,
                                                  dQ, c2, Q2, lim)
    dQdt = c0 * Qf + c1 * Q_prev + d1 * dQ + c2 * Q2
    return NRResult(X, it, converged, max_f, Qf, dQdt, I_vs, mrc, lim)


def _apply_noi(cir, csr, f, noi):
    """NOI rows and columns zeroed, their diagonals 1, their right-hand side
    0 (solver_factories.py:1268-1276, masks of :105-135: columns are masked
    in node rows only, branch rows keep theirs)."""
    csr = csr.copy()
    f = f.copy()
    noi_set = set(int(r) for r in noi)
    cols = np.asarray(cir.indices)
    for r in range(cir.n_unk):
        for e in range(int(cir.indptr[r]), int(cir.indptr[r + 1])):
            if r in noi_set or int(cols[e]) in noi_set:
                csr[e] = 0.0
    for r in noi:
        for e in range(int(cir.indptr[r]), int(cir.indptr[r + 1])):
            if int(cols[e]) == r:
                csr[e] = 1.0
    f[noi] = 0.0
    return csr, f


def predict(V_hist, dt_hist, hist_count, new_dt, max_order):
    """adaptive.py:94-218 (order <= 2 Lagrange extrapolation)."""
    order = max(min(hist_count - 1, max_order), 0)
    cs = np.cumsum(dt_hist)
    tau = -(new_dt + np.concatenate([[0.0], cs[:-1]])) / new_dt
    ratio1 = new_dt / max(dt_hist[0], 1e-30)
    a1 = [1.0 + ratio1, -ratio1, 0.0]
    err1 = -ratio1 * (1 + ratio1) / 2
    tau0, tau1, tau2 = -1.0, tau[1], tau[2]
    den0 = (tau0 - tau1) * (tau0 - tau2)
    den1 = (tau1 - tau0) * (tau1 - tau2)
    den2 = (tau2 - tau0) * (tau2 - tau1)
    den0 = 1e-30 if abs(den0) < 1e-30 else den0
    den1 = 1e-30 if abs(den1) < 1e-30 else den1
    den2 = 1e-30 if abs(den2) < 1e-30 else den2
    l0 = (0 - tau1) * (0 - tau2) / den0
    l1 = (0 - tau0) * (0 - tau2) / den1
    l2 = (0 - tau0) * (0 - tau1) / den2
    a2 = [l0, l1, l2]
    err2 = -tau0 * tau1 * tau2 / 6
    if order == 0:
        a, err = [1.0, 0.0, 0.0], 1.0
    elif order == 1:
        a, err = a1, err1
    else:
        a, err = a2, err2
    a = [a[i] if i <= order else 0.0 for i in range(3)]
    V_pred = a[0] * V_hist[0] + a[1] * V_hist[1] + a[2] * V_hist[2]
    return V_pred, err, a


def lte_timestep(V_new, V_pred, pred_err, dt, hist_count, cfg: AdaptiveConfig, err_integ, V_max_hist):
    """adaptive.py:221-339 (relrefAlllocal tolerance)."""
    lte = (V_new - V_pred) * (err_integ / (err_integ - pred_err))
    V_cur = np.maximum(np.abs(V_new), np.abs(V_pred))
    V_scale = np.maximum(V_max_hist, V_cur)
    tol = cfg.reltol * V_scale + cfg.abstol
    lte_norm = float(np.max(np.abs(lte) / tol))
    order = min(hist_count - 1, cfg.max_order)
    factor = math.pow(cfg.lte_ratio / max(lte_norm, 1e-10), 1.0 / (order + 1))
    factor = min(max(factor, 0.1), cfg.grow_factor)
    return dt * factor, lte_norm


def _init_mid_rail(cir) -> np.ndarray:
    """full_mna.py:694-745 with get_vdd_value (sources.py:553-568)."""
    net = cir.net
    vs = net.vsources
    vdd = 0.0
    for i in range(vs.count):
        if vs.params["dc"][i] > vdd:
            vdd = float(vs.params["dc"][i])
    V = np.full(cir.n_total, vdd / 2.0)
    V[0] = 0.0
    for name, idx in cir.node_names.items():
        low = name.lower()
        if "vdd" in low or "vcc" in low:
            V[idx] = vdd
        elif low in ("gnd", "vss", "0"):
            V[idx] = 0.0
    for i in range(vs.count):
        p, n = int(cir.vs_p[i]), int(cir.vs_n[i])
        if n == 0 and p > 0:
            V[p] = float(vs.params["dc"][i])
    return V


@dataclass
class TransientOut:
    times: np.ndarray
    V: np.ndarray  # [n_steps, n_total]
    I: np.ndarray  # [n_steps, m]
    stats: Dict = field(default_factory=dict)


def initial_state(cir, solver, opts: Options):
    """DC operating point or UIC start (full_mna.py:871-957)."""
    n_total, n_unk, m = cir.n_total, cir.n_unk, cir.m
    X0 = np.zeros(n_total + m)
    X0[:n_total] = _init_mid_rail(cir)
    vs0 = source_values(cir.net.vsources, 0.0)
    is0 = source_values(cir.net.isources, 0.0)
    if opts.icmode == "uic":
        X0[:n_total] = 0.0
        _, _, Q_init, I_dc, _ = mna.build_system(cir, X0, vs0, is0, np.zeros(n_unk), 0.0)
        mrc = np.zeros(n_unk)
        dc_info = None
    else:
        r = nr_solve(cir, solver, opts, X0, vs0, is0, np.zeros(n_unk), 0.0)
        X0, Q_init, I_dc, mrc = r.X, r.Q, r.I_vsource, r.max_res_contrib
        dc_info = r
    return X0, Q_init, I_dc, mrc, dc_info


def run_adaptive(cir, t_stop: float, dt: float, cfg: AdaptiveConfig, opts: Options,
                 max_steps: Optional[int] = None, dense: bool = False, step_limit=None) -> TransientOut:
    n_total, n_unk, m = cir.n_total, cir.n_unk, cir.m
    if max_steps is None:
        max_steps = int(t_stop / dt * 1.1) + 10
    if cfg.tran_minpts > 0:
        eff = min(cfg.max_dt, t_stop / cfg.tran_minpts)
        if eff < cfg.max_dt:
            cfg = replace(cfg, max_dt=eff)
    if cfg.tran_fs != 1.0:
        dt = dt * cfg.tran_fs
    solver = LinearSolver(cir, dense=dense)
    X, Q_prev, I_dc, hist_mrc, _ = initial_state(cir, solver, opts)
    max_hist = cfg.max_order + 2
    V_hist = np.zeros((max_hist, n_total))
    V_hist[0] = X[:n_total]
    dt_hist = np.full(max_hist, dt)
    hist_count = 1
    times = [0.0]
    Vout = [X[:n_total].copy()]
    Iout = [I_dc[:m].copy() if m else np.zeros(1)]
    t = 0.0
    step_idx = 1
    warmup = 0
    dQdt_prev = np.zeros(n_unk)
    Q_prev2 = np.zeros(n_unk)
    total_nr = rejected = consec = 0
    min_dt_used = max_dt_used = dt
    V_max_hist = np.abs(X[:n_total])
    lim = np.zeros(cir.n_lim)
    method = method_of(cfg.integration_method)
    trace = []  # (t_next, dt, Newton iterations) of every step attempt
    t0 = time.perf_counter()
    iters = 0
    while t < t_stop and step_idx < max_steps:
        if step_limit is not None and iters >= step_limit:
            break
        iters += 1
        dt_cur = dt
        if dt_cur <= 0.0:
            # full_mna.py:2076 clamps the next dt to t_stop - t_next also after a
            # rejected step, so rejecting a step that ends on t_stop leaves dt = 0.
            # The reference then builds with 1/dt = inf: every Newton iteration is
            # non-finite, the step is a Newton failure at min_dt, and it is
            # force-accepted with the previous solution and zero source currents
            # (full_mna.py:2059-2085).  Restated without the non-finite arithmetic.
            total_nr += opts.tran_itl
            consec = 0
            last_good = dt_hist[0] if hist_count > 0 else cfg.min_dt * 16.0
            new_dt = max(last_good / 8.0, cfg.min_dt * 2.0)
            new_dt = min(max(new_dt, cfg.min_dt), cfg.max_dt)
            new_dt = min(new_dt, t_stop - t)
            times.append(t)
            Vout.append(X[:n_total].copy())
            Iout.append(np.zeros(max(m, 1)))
            step_idx += 1
            min_dt_used = min(min_dt_used, dt_cur)
            max_dt_used = max(max_dt_used, dt_cur)
            dt = new_dt
            continue
        c0, c1, d1, c2, e_int = integration_coeffs(method, dt_cur, dt_hist[0])
        t_next = t + dt_cur
        vs = source_values(cir.net.vsources, t_next)
        isv = source_values(cir.net.isources, t_next)
        can_predict = (warmup >= cfg.warmup_steps) and (hist_count >= 2)
        V_pred, pred_err, _ = predict(V_hist, dt_hist, hist_count, dt_cur, cfg.max_order)
        X_init = X.copy()
        if can_predict:
            X_init[:n_total] = V_pred
        ramp = min(max(step_idx / cfg.gshunt_steps, 0.0), 1.0) if cfg.gshunt_steps > 0 else 1.0
        gshunt = cfg.gshunt_init + ramp * (cfg.gshunt_target - cfg.gshunt_init)
        floor = np.maximum(hist_mrc * cfg.reltol, cfg.abstol)
        r = nr_solve(cir, solver, opts, X_init, vs, isv, Q_prev, c0, gshunt, c1, d1, dQdt_prev, c2,
                     Q_prev2, res_tol_floor=floor, lim_in=lim)
        total_nr += r.iterations
        trace.append((t_next, dt_cur, r.iterations))
        nr_failed = not r.converged
        at_min = dt_cur <= cfg.min_dt
        fail_min = nr_failed and at_min
        V_new = r.X[:n_total]
        dt_lte, lte_norm = lte_timestep(V_new, V_pred, pred_err, dt_cur, hist_count, cfg, e_int, V_max_hist)
        lte_reject = (dt_cur / dt_lte > cfg.redo_factor) and can_predict and r.converged
        nr_reject = nr_failed and not at_min
        force = consec >= cfg.max_consecutive_rejects
        reject = (lte_reject or nr_reject) and not force
        accept = not reject
        if accept:
            consec = 0
        elif lte_reject:
            consec += 1
        dt_from_lte = dt_lte if can_predict else dt_cur
        dt_halved = max(dt_cur * cfg.tran_ft, cfg.min_dt)
        last_good = dt_hist[0] if hist_count > 0 else cfg.min_dt * 16.0
        dt_rec = max(last_good / 8.0, cfg.min_dt * 2.0)
        new_dt = (dt_rec if at_min else dt_halved) if nr_failed else dt_from_lte
        new_dt = min(max(new_dt, cfg.min_dt), cfg.max_dt)
        new_dt = min(new_dt, t_stop - t_next)
        use_new = accept and not fail_min
        if accept:
            t = t_next
        if use_new:
            X = r.X.copy()
            lim = r.lim
            Q_prev2 = Q_prev
            Q_prev = r.Q
            dQdt_prev = r.dQdt
            V_hist = np.roll(V_hist, 1, axis=0)
            V_hist[0] = V_new
            dt_hist = np.roll(dt_hist, 1)
            dt_hist[0] = dt_cur
            hist_count = min(hist_count + 1, max_hist)
            warmup += 1
            V_max_hist = np.maximum(V_max_hist, np.abs(V_new))
            hist_mrc = np.maximum(hist_mrc, r.max_res_contrib)
        if accept:
            times.append(t_next)
            Vout.append(X[:n_total].copy())
            Iout.append(np.zeros(max(m, 1)) if fail_min else (r.I_vsource[:m].copy() if m else np.zeros(1)))
            step_idx += 1
            min_dt_used = min(min_dt_used, dt_cur)
            max_dt_used = max(max_dt_used, dt_cur)
        else:
            rejected += 1
        dt = new_dt
    wall = time.perf_counter() - t0
    n_steps = step_idx
    stats = {
        "n_steps": n_steps,
        "accepted_steps": n_steps,
        "rejected_steps": rejected,
        "total_nr_iterations": total_nr,
        "wall_time": wall,
        "time_per_step_ms": wall / n_steps * 1000 if n_steps else 0.0,
        "min_dt_used": min_dt_used,
        "max_dt_used": max_dt_used,
        "body_iterations": iters,
        "trace": trace,
    }
    return TransientOut(np.array(times), np.array(Vout), np.array(Iout), stats)


def run_fixed(cir, t_stop: float, dt: float, method: str, opts: Options, tran_fs: float = 0.25,
              max_steps: Optional[int] = None, dense: bool = False, n_steps: Optional[int] = None,
              abstol: float = 1e-12) -> TransientOut:
    """Fixed-step loop of full_mna.py:1198-1441 (constant coefficients)."""
    n_total, n_unk, m = cir.n_total, cir.n_unk, cir.m
    if max_steps is None:
        max_steps = int(t_stop / dt * 1.1) + 10
    dt = dt * tran_fs if tran_fs != 1.0 else dt
    if n_steps is None:
        n_steps = max(min(int(t_stop / dt), max_steps - 1), 1)
    solver = LinearSolver(cir, dense=dense)
    X, Q_prev, I_dc, _, _ = initial_state(cir, solver, opts)
    inv = 1.0 / dt
    method = method_of(method)
    c0, c1, d1, c2 = inv, -inv, 0.0, 0.0
    if method == TRAP:
        c0, c1, d1 = 2.0 * inv, -2.0 * inv, -1.0
    elif method == GEAR2:
        c0, c1, c2 = 1.5 * inv, -2.0 * inv, 0.5 * inv
    times = np.zeros(n_steps + 1)
    Vout = np.zeros((n_steps + 1, n_total))
    Iout = np.zeros((n_steps + 1, max(m, 1)))
    Vout[0] = X[:n_total]
    dQdt_prev = np.zeros(n_unk)
    Q_prev2 = np.zeros(n_unk)
    floor = np.full(n_unk, abstol)
    lim = np.zeros(cir.n_lim)
    total_nr = 0
    t0 = time.perf_counter()
    for i in range(n_steps):
        t_next = (float(i) + 1.0) * dt
        vs = source_values(cir.net.vsources, t_next)
        isv = source_values(cir.net.isources, t_next)
        r = nr_solve(cir, solver, opts, X, vs, isv, Q_prev, c0, 0.0, c1, d1, dQdt_prev, c2, Q_prev2,
                     res_tol_floor=floor, lim_in=lim)
        if r.converged:  # full_mna.py:1349-1353
            lim = r.lim
        total_nr += r.iterations
        times[i + 1] = t_next
        Vout[i + 1] = r.X[:n_total]
        if m:
            Iout[i + 1] = r.I_vsource[:m]
        X = r.X
        Q_prev2 = Q_prev
        Q_prev = r.Q
        dQdt_prev = r.dQdt
    wall = time.perf_counter() - t0
    stats = {"n_steps": n_steps + 1, "accepted_steps": n_steps, "total_nr_iterations": total_nr,
             "wall_time": wall, "time_per_step_ms": wall / n_steps * 1000}
    return TransientOut(times, Vout, Iout, stats)
