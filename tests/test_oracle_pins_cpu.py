"""Known answers for the oracle's adaptive step loop and its node / CSR setup.

These tests pin ``oracle.transient.run_adaptive`` and ``oracle.mna.setup``
(and the product's ``vajax_amd.topology``) against answers derived
independently of the oracle's code:

* a single-node circuit (capacitor C to ground, current source, the diagonal
  gmin) simulated by a scalar restatement of the reference's step control
  (vajax/analysis/transient/full_mna.py:1840-2194,
  vajax/analysis/transient/adaptive.py:94-339) whose numerics come from
  textbook constructions, not from the oracle's formulas: the BDF2
  derivative weights and the predictor's extrapolation weights are solved
  from Vandermonde systems, and the order-2 predictor error coefficient is
  the extrapolation error of tau^3/6.  The first three steps are also
  checked against values computed by hand (fractions).  Three scenarios
  drive the loop through dt growth, LTE rejection with the consecutive-reject
  cap, Newton failure cuts and min_dt force-accepts;
* a single RC section against the closed form 1 - exp(-t/RC);
* node numbering and the CSR pattern of a small netlist written out by hand
  from vajax/analysis/engine.py:459-468 and
  vajax/analysis/transient/full_mna.py:461-479.
"""

from __future__ import annotations

import math
from fractions import Fraction

import numpy as np
import pytest

from oracle import mna as omna
from oracle import transient as otr
from vajax_amd import topology
from vajax_amd.netlist import NetlistBuilder

C_F = 1e-12  # capacitance of the single node
GMIN = 1e-12  # the options' diagonal gmin (mna_builder.py:384)


# --------------------------------------------------------------------------
# independent numerics
# --------------------------------------------------------------------------

def _vander_weights(nodes, target, deriv):
    """Weights w with sum_i w_i p(nodes[i]) = p(target) (deriv 0) or
    p'(target) (deriv 1) for every polynomial of degree < len(nodes)."""
    k = len(nodes)
    A = np.array([[x ** j for x in nodes] for j in range(k)], dtype=np.float64)
    if deriv == 0:
        rhs = np.array([target ** j for j in range(k)], dtype=np.float64)
    else:
        rhs = np.array([j * target ** (j - 1) if j else 0.0 for j in range(k)], dtype=np.float64)
    return np.linalg.solve(A, rhs)


def _bdf2_weights(h, hp):
    """d/dt at the new point of the quadratic through (0, -h, -h-hp)."""
    return _vander_weights([0.0, -h, -h - hp], 0.0, 1)


def _pulse(t, v0, v1, rise, fall, width, period, delay):
    if t < delay:
        return v0
    tin = math.fmod(t - delay, period)
    if tin < rise:
        return v0 + (v1 - v0) * tin / rise
    if tin < rise + width:
        return v1
    if tin < rise + width + fall:
        return v1 - (v1 - v0) * (tin - rise - width) / fall
    return v0


def mini_run(I_of_t, t_stop, dt, body_limit, tran_itl=20, min_dt=1e-15, reltol=1e-3,
             abstol=1e-12, vntol=1e-6, lte_ratio=3.5, redo=2.5, grow=2.0, tran_fs=0.25,
             tran_ft=0.25, warmup_steps=2, max_rejects=5, tran_minpts=50, lte_reltol=None,
             G=0.0, dc_start=False):
    """Scalar gear2 restatement of the adaptive loop for
    C dV/dt + (G + gmin) V = I(t), one node plus ground: uic start (V = 0) or
    the DC point V = I(0) / (G + gmin).  lte_reltol is the step control's
    reltol (LTE tolerance and Newton floor; default: the Newton reltol)."""
    lte_reltol = reltol if lte_reltol is None else lte_reltol
    max_dt = min(math.inf, t_stop / tran_minpts)
    dt = dt * tran_fs
    V = I_of_t(0.0) / (G + GMIN) if dc_start else 0.0
    # residual contributions at the node: the source's |I| and the resistor's |G V|
    hist_mrc = max(abs(I_of_t(0.0)), abs(G * V)) if dc_start else 0.0
    V_hist = [V, 0.0, 0.0, 0.0]  # t_n, t_{n-1}, ...
    dt_hist = [dt, dt, dt, dt]
    hist_count, warmup, consec = 1, 0, 0
    t = 0.0
    Qp, Q2 = C_F * V, 0.0  # charge at t_n and t_{n-1}; the first step's Q2 is zero
    vmax = abs(V)
    times, Vs, Is = [0.0], [V], []
    total_nr = rejected = 0
    log = []
    for _ in range(body_limit):
        if not (t < t_stop):
            break
        h = dt
        hp = dt_hist[0]
        w0, w1, w2 = _bdf2_weights(h, hp)
        t_next = t + h
        I = I_of_t(t_next)
        can_predict = warmup >= warmup_steps and hist_count >= 2
        order = max(min(hist_count - 1, 2), 0)
        taus = [-1.0, -(h + dt_hist[0]) / h, -(h + dt_hist[0] + dt_hist[1]) / h]
        if order == 2:
            a = _vander_weights(taus, 0.0, 0)
            pred_err = 0.0 - sum(a[i] * taus[i] ** 3 / 6.0 for i in range(3))
        elif order == 1:
            a = _vander_weights(taus[:2], 0.0, 0)
            r1 = h / dt_hist[0]
            pred_err = -r1 * (1 + r1) / 2  # adaptive.py:141 (the reference's constant)
        else:
            a, pred_err = [1.0], 1.0
        V_pred = sum(a[i] * V_hist[i] for i in range(len(a)))
        V_init = V_pred if can_predict else V
        # linear Newton: f(V) = C (w0 V + w1 Qp/C + w2 Q2/C) + gmin V - I
        A = w0 * C_F + G + GMIN
        f0 = w0 * C_F * V_init + w1 * Qp + w2 * Q2 + (G + GMIN) * V_init - I
        # the floor from the historic maxima uses the step control's reltol
        # (full_mna.py:1814, 1925-1928), the test itself Newton's
        res_tol = max(max(abs(I), abs(G * V_init)) * reltol, max(hist_mrc * lte_reltol, abstol))
        V1 = V_init - f0 / A
        if abs(f0) < res_tol:
            iters, conv = 1, True
        elif tran_itl >= 2:
            iters, conv = 2, True
        else:
            iters, conv = 1, False
        total_nr += iters
        V_new = V1
        at_min = h <= min_dt
        fail_min = (not conv) and at_min
        e_int = -2.0 / 9.0
        lte = (V_new - V_pred) * (e_int / (e_int - pred_err))
        tol = lte_reltol * max(vmax, abs(V_new), abs(V_pred)) + abstol
        lte_norm = abs(lte) / tol
        lorder = min(hist_count - 1, 2)
        factor = min(max((lte_ratio / max(lte_norm, 1e-10)) ** (1.0 / (lorder + 1)), 0.1), grow)
        dt_lte = h * factor
        lte_reject = (h / dt_lte > redo) and can_predict and conv
        nr_reject = (not conv) and not at_min
        reject = (lte_reject or nr_reject) and not (consec >= max_rejects)
        accept = not reject
        if accept:
            consec = 0
        elif lte_reject:
            consec += 1
        if not conv:
            new_dt = max(dt_hist[0] / 8.0, 2 * min_dt) if at_min else max(h * tran_ft, min_dt)
        else:
            new_dt = dt_lte if can_predict else h
        new_dt = min(max(new_dt, min_dt), max_dt)
        new_dt = min(new_dt, t_stop - t_next)
        log.append(dict(t_next=t_next, h=h, iters=iters, conv=conv, accept=accept,
                        lte_norm=lte_norm if can_predict else None))
        if accept:
            t = t_next
            if not fail_min:
                V = V_new
                Q2, Qp = Qp, C_F * V_new
                V_hist = [V_new] + V_hist[:-1]
                dt_hist = [h] + dt_hist[:-1]
                hist_count = min(hist_count + 1, 4)
                warmup += 1
                vmax = max(vmax, abs(V_new))
                hist_mrc = max(hist_mrc, abs(I), abs(G * V_new))
            times.append(t_next)
            Vs.append(V)
            Is.append(0.0 if fail_min else I)
        else:
            rejected += 1
        dt = new_dt
    return dict(times=np.array(times), V=np.array(Vs), total_nr=total_nr, rejected=rejected,
                log=log)


def _cap_netlist(kind, R=None, **wave):
    b = NetlistBuilder("single_cap")
    # current I enters node a (p = ground, n = a: +I stamped at p, -I at n)
    b.source("isource", "i1", "0", "a", type=kind, **wave)
    b.device("capacitor", "c1", ["a", "0"], c=C_F)
    if R is not None:
        b.device("resistor", "r1", ["a", "0"], r=R)
    return b.build(options={"tran_method": "gear2", "icmode": "uic"})


def _run_oracle(net, t_stop, dt, body_limit, tran_itl=20, icmode="uic", **cfg_over):
    cir = omna.setup(net)
    cfg = otr.AdaptiveConfig(integration_method="gear2", **cfg_over)
    opts = otr.Options(icmode=icmode, tran_itl=tran_itl)
    return cir, otr.run_adaptive(cir, t_stop, dt, cfg, opts, step_limit=body_limit)


def _compare(out, cir, ref):
    a = cir.node_names["a"]
    assert len(out.times) == len(ref["times"]), (len(out.times), len(ref["times"]))
    np.testing.assert_allclose(out.times, ref["times"], rtol=1e-10, atol=1e-27)
    np.testing.assert_allclose(out.V[:, a], ref["V"], rtol=1e-9, atol=1e-15)
    assert out.stats["total_nr_iterations"] == ref["total_nr"]
    assert out.stats["rejected_steps"] == ref["rejected"]


I_ON = 1e-3


def test_first_steps_by_hand():
    """Constant 1 mA into 1 pF from V = 0 (uic), dt0 = 1 ps * tran_fs.

    With u = I dt0 / C (= 0.25 mV): the first gear2 step sees a zero Q_prev2,
    so 3/(2 dt0) C V1 = I gives V1 = 2u/3; the second (3 V2 - 4 V1)/(2 dt0) C
    = I gives V2 = 14u/9; the third is the first predicted one: V3 = 68u/27,
    quadratic extrapolation 3 V2 - 3 V1 + V0 = 8u/3, LTE factor
    (-2/9) / (-2/9 - 1) = 2/11 (uniform steps: tau = -1, -2, -3, predictor
    error 1), so lte = (-4u/27)(2/11) and tol = 1e-3 * 8u/3 (+ abstol).
    gmin (1e-12 S) shifts these by ~1e-12 relative."""
    net = _cap_netlist("dc", dc=I_ON)
    cir, out = _run_oracle(net, 1e-9, 1e-12, body_limit=3)
    a = cir.node_names["a"]
    u = I_ON * 0.25e-12 / C_F
    expect = [Fraction(0), Fraction(2, 3), Fraction(14, 9), Fraction(68, 27)]
    assert len(out.times) == 4
    np.testing.assert_allclose(out.times, [0.0, 0.25e-12, 0.5e-12, 0.75e-12], rtol=1e-15)
    np.testing.assert_allclose(out.V[:, a], [float(e) * u for e in expect], rtol=1e-9, atol=0)
    lte = abs(Fraction(-4, 27) * Fraction(2, 11))
    lte_norm = float(lte) * u / (1e-3 * float(Fraction(8, 3)) * u + 1e-12)
    factor = min(max((3.5 / lte_norm) ** (1.0 / 3.0), 0.1), 2.0)
    # the fourth step is taken with dt3 * factor and accepted (1/factor < 2.5)
    cir, out4 = _run_oracle(net, 1e-9, 1e-12, body_limit=4)
    assert out4.stats["rejected_steps"] == 0
    np.testing.assert_allclose(out4.times[4] - out4.times[3], 0.25e-12 * factor, rtol=1e-9)
    # every step takes two Newton iterations (a linear solve, then the check)
    assert out4.stats["total_nr_iterations"] == 8


def test_step_sequence_constant_current():
    """dt growth under the LTE rule up to max_dt = t_stop / 50."""
    net = _cap_netlist("dc", dc=I_ON)
    ref = mini_run(lambda t: I_ON, 1e-9, 1e-12, body_limit=40)
    cir, out = _run_oracle(net, 1e-9, 1e-12, body_limit=40)
    _compare(out, cir, ref)
    steps = np.diff(ref["times"])
    assert steps.max() > 4 * steps[0]  # the rule grew dt


@pytest.mark.parametrize("cap,lte_reltol", [(5, 1e-3), (1, 1e-3), (5, 1e-5)])
def test_step_sequence_lte_reject_and_cap(cap, lte_reltol):
    """A current step at 2 ps after V = 0: dt doubles (zero LTE) until the
    step across the edge is rejected for LTE, then shrinks by the LTE factor
    and regrows.  With the consecutive-reject cap at 1 the second of two
    consecutive LTE rejections is force-accepted; with a step-control reltol
    of 1e-5 the LTE factor falls below its 0.1 clip."""
    wave = dict(val0=0.0, val1=I_ON, rise=1e-15, fall=1e-15, width=1e-6, period=2e-6,
                delay=2e-12)
    net = _cap_netlist("pulse", **wave)
    ref = mini_run(lambda t: _pulse(t, 0.0, I_ON, 1e-15, 1e-15, 1e-6, 2e-6, 2e-12), 1e-9,
                   1e-12, body_limit=40, max_rejects=cap, lte_reltol=lte_reltol)
    log = ref["log"]
    assert any(not e["accept"] and e["conv"] for e in log)  # LTE rejection
    if cap == 1:
        forced = [e for e in log if e["accept"] and e["lte_norm"] and e["lte_norm"] > 3.5 * 2.5 ** 3]
        assert forced, "no force-accepted LTE rejection"
    else:
        assert ref["rejected"] >= 3
    if lte_reltol < 1e-3:
        assert any(e["lte_norm"] and e["lte_norm"] > 3.5 * 10 ** 3 for e in log)  # clipped
    cir, out = _run_oracle(net, 1e-9, 1e-12, body_limit=40, max_consecutive_rejects=cap,
                           reltol=lte_reltol)
    _compare(out, cir, ref)


def test_step_sequence_newton_failure_and_min_dt():
    """tran_itl = 1: Newton converges only where its starting point already
    satisfies the residual test (the flat part); after the current step every
    step fails, dt is cut by tran_ft down to min_dt, force-accepted there with
    the previous solution and restarted from last_good / 8."""
    wave = dict(val0=0.0, val1=I_ON, rise=1e-15, fall=1e-15, width=1e-6, period=2e-6,
                delay=2e-12)
    net = _cap_netlist("pulse", **wave)
    ref = mini_run(lambda t: _pulse(t, 0.0, I_ON, 1e-15, 1e-15, 1e-6, 2e-6, 2e-12), 1e-9,
                   1e-12, body_limit=30, tran_itl=1, min_dt=5e-14)
    log = ref["log"]
    assert any((not e["conv"]) and (not e["accept"]) for e in log)  # NR failure cut
    assert any((not e["conv"]) and e["accept"] for e in log)  # min_dt force-accept
    cir, out = _run_oracle(net, 1e-9, 1e-12, body_limit=30, tran_itl=1, min_dt=5e-14)
    _compare(out, cir, ref)


def test_step_sequence_dc_start_rise_and_decay():
    """R || C from its DC point (0.2 V) through a 0.2 -> 1 mA current pulse
    and back: the first step's zero Q_prev2 kicks V (the reference's
    first-step history), the historic maxima of |V| (LTE tolerance) and of
    the residual contributions (Newton floor) govern the decay."""
    R = 1e3
    wave = dict(val0=0.2 * I_ON, val1=I_ON, rise=5e-12, fall=5e-12, width=100e-12,
                period=2e-6, delay=20e-12)
    b = NetlistBuilder("single_rc_node")
    b.source("isource", "i1", "0", "a", type="pulse", **wave)
    b.device("capacitor", "c1", ["a", "0"], c=C_F)
    b.device("resistor", "r1", ["a", "0"], r=R)
    net = b.build(options={"tran_method": "gear2", "icmode": "op"})
    I_of_t = lambda t: _pulse(t, 0.2 * I_ON, I_ON, 5e-12, 5e-12, 100e-12, 2e-6, 20e-12)
    ref = mini_run(I_of_t, 1e-9, 1e-12, body_limit=120, G=1.0 / R, dc_start=True)
    # the first step moves V away from its DC value (zero Q_prev2)
    assert abs(ref["V"][1] - ref["V"][0]) > 1e-3
    assert ref["V"][-1] < ref["V"].max() - 0.05  # decayed after the pulse
    cir, out = _run_oracle(net, 1e-9, 1e-12, body_limit=120, icmode="op")
    _compare(out, cir, ref)


# --------------------------------------------------------------------------
# closed form: single RC section
# --------------------------------------------------------------------------

def test_single_rc_closed_form():
    """in -> R -> out, C to ground, 0 -> 1 V step (1 fs rise): the adaptive
    gear2 run follows 1 - exp(-t/RC) within the LTE tolerance (reltol 1e-3
    per step, accumulated over the run)."""
    R, C = 1e3, 1e-12
    b = NetlistBuilder("single_rc")
    b.source("vsource", "v1", "in", "0", type="pulse", val0=0.0, val1=1.0, rise=1e-15,
             fall=1e-15, width=1e-6, period=2e-6, delay=0.0)
    b.device("resistor", "r1", ["in", "out"], r=R)
    b.device("capacitor", "c1", ["out", "0"], c=C)
    net = b.build(options={"tran_method": "gear2", "icmode": "uic"})
    cir = omna.setup(net)
    cfg = otr.AdaptiveConfig(integration_method="gear2")
    out = otr.run_adaptive(cir, 5e-9, 1e-12, cfg, otr.Options(icmode="uic"))
    tau = R * C
    t = out.times
    v = out.V[:, cir.node_names["out"]]
    exact = 1.0 - np.exp(-np.maximum(t - 1e-15, 0.0) / tau)
    err = np.max(np.abs(v - exact))
    assert t[-1] == pytest.approx(5e-9, rel=1e-12)
    assert err < 5e-3, err
    # and the error is at the LTE level, not zero (the integration is real)
    assert err > 1e-6
    # end value within 1e-3 of the settled closed form
    assert abs(v[-1] - exact[-1]) < 1e-3


# --------------------------------------------------------------------------
# node numbering and CSR pattern written out by hand
# --------------------------------------------------------------------------

def _kat_netlist():
    b = NetlistBuilder("kat", ground="0")
    b.source("vsource", "v1", "in", "0", type="dc", dc=1.0)
    b.device("resistor", "r1", ["in", "n10"], r=1e3)
    b.device("resistor", "r2", ["n10", "n2"], r=2e3)
    b.device("capacitor", "c1", ["n2", "0"], c=1e-12)
    b.source("vsource", "v2", "n2", "n10", type="dc", dc=0.5)
    b.source("isource", "i1", "n2", "0", type="dc", dc=1e-6)
    return b.build(options={"tran_method": "gear2"})


# ground 0; sorted names: "in" < "n10" < "n2"; MNA row = index - 1; branch
# rows of v1, v2 follow (4, 5 -> MNA 3, 4)
KAT_NODES = {"0": 0, "in": 1, "n10": 2, "n2": 3}
# entries (row, col):
#   r1 (in, n10)      -> (0,0) (0,1) (1,0) (1,1)
#   r2 (n10, n2)      -> (1,1) (1,2) (2,1) (2,2)
#   c1 (n2, ground)   -> (2,2); ground terms fold onto (0,0)
#   v1 (in, ground)   -> B (0,3), B^T (3,0); ground side folds onto (0,0)
#   v2 (n2, n10)      -> B (2,4) (1,4), B^T (4,2) (4,1)
#   node-row diagonals (0,0) (1,1) (2,2); branch rows have none
KAT_INDPTR = [0, 3, 7, 10, 11, 13]
KAT_INDICES = [0, 1, 3,  0, 1, 2, 4,  1, 2, 4,  0,  1, 2]


def test_node_numbering_and_csr_pattern_kat():
    net = _kat_netlist()
    cir = omna.setup(net)
    assert cir.node_names == KAT_NODES
    assert (cir.n_total, cir.n_unk, cir.m, cir.n_aug) == (4, 3, 2, 5)
    assert cir.indptr.tolist() == KAT_INDPTR
    assert cir.indices.tolist() == KAT_INDICES
    assert cir.diag_pos.tolist() == [0, 4, 8]
    # B / B^T slots in stamping order: all b_p, all b_n, all bt_p, all bt_n
    assert cir.b_pos == [(2, 1.0), (9, 1.0), (6, -1.0), (10, 1.0), (12, 1.0), (11, -1.0)]
    # capacitor stamp positions: (n2,n2) = slot 8, the ground terms dropped
    cap = [om for om in cir.models if om.name == "capacitor"][0]
    assert cap.csr_pos[0].tolist() == [8, -1, -1, -1]
    res = [om for om in cir.models if om.name == "resistor"][0]
    assert res.csr_pos.tolist() == [[0, 1, 3, 4], [4, 5, 7, 8]]

    # the product's setup gives the same answer
    topo = topology.build_topology(net)
    assert topo.node_names == KAT_NODES
    assert topo.indptr.tolist() == KAT_INDPTR
    assert topo.indices.tolist() == KAT_INDICES
    assert topo.diag_pos.tolist() == [0, 4, 8]
    assert topo.models["capacitor"].csr_pos[0].tolist()[0] == 8
    assert topo.models["resistor"].csr_pos.tolist() == [[0, 1, 3, 4], [4, 5, 7, 8]]


def test_kat_dense_system():
    """The KAT circuit's Jacobian at X = 0 (c0 = 0) written out by hand."""
    net = _kat_netlist()
    cir = omna.setup(net)
    X = np.zeros(cir.n_total + cir.m)
    vs = np.array([1.0, 0.5])
    isv = np.array([1e-6])
    csr, f, Q, I_vs, _ = omna.build_system(cir, X, vs, isv, np.zeros(cir.n_unk), 0.0)
    J = omna.csr_matrix(cir, csr).toarray()
    g1, g2, g = 1e-3, 0.5e-3, GMIN
    expect = np.array([
        [g1 + g, -g1, 0, 1, 0],
        [-g1, g1 + g2 + g, -g2, 0, -1],
        [0, -g2, g2 + g, 0, 1],
        [1, 0, 0, 0, 0],
        [0, -1, 1, 0, 0],
    ])
    np.testing.assert_allclose(J, expect, rtol=1e-15, atol=0)
    # f at X = 0: the current source stamps +I at p (n2), branch rows -E
    np.testing.assert_allclose(f, [0.0, 0.0, 1e-6, -1.0, -0.5], rtol=1e-15, atol=0)


def test_oracle_solver_returns_nan_for_singular_or_nonfinite_systems():
    # UMFPACK's singular-matrix warning is tolerated and its non-finite
    # solution returned (umfpack_ffi.cpp:297, 320): the oracle's solve gives
    # a NaN update instead of raising, so Newton fails its tests
    from types import SimpleNamespace

    from oracle.transient import LinearSolver

    cir = SimpleNamespace(n_aug=2, indptr=np.array([0, 2, 4]), indices=np.array([0, 1, 0, 1]))
    f = np.array([1.0, 2.0])
    for dense in (False, True):
        solver = LinearSolver(cir, dense=dense)
        x = solver.solve(np.array([2.0, 1.0, 1.0, 3.0]), f)
        assert np.allclose(np.array([[2.0, 1.0], [1.0, 3.0]]) @ x, -f)
        assert np.all(np.isnan(solver.solve(np.array([1.0, np.inf, 1.0, 1.0]), f)))
    # an exactly singular system: NaN from the dense fallback of the sparse path
    x = LinearSolver(cir).solve(np.array([1.0, 1.0, 1.0, 1.0]), f)
    assert np.all(np.isnan(x)) or not np.all(np.isfinite(x))
