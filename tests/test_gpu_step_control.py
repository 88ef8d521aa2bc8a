"""GPU parity of the adaptive step control on the known-answer scenarios of
tests/test_oracle_pins_cpu.py (whose independent scalar restatement pins
the oracle): dt growth, LTE rejection with the consecutive-reject cap,
Newton-failure cuts with min_dt force-accepts, and a DC start with the
first step's zero charge history.  The HIP path must take the same accepted
time points, iteration totals and rejections as the oracle."""

import numpy as np
import pytest

from vajax_amd.netlist import NetlistBuilder

from .helpers import run_gpu, run_oracle, grid_deviation

pytestmark = pytest.mark.gpu

I_ON = 1e-3


def _cap(kind, R=None, icmode="uic", **wave):
    b = NetlistBuilder("single_cap")
    b.source("isource", "i1", "0", "a", type=kind, **wave)
    b.device("capacitor", "c1", ["a", "0"], c=1e-12)
    if R is not None:
        b.device("resistor", "r1", ["a", "0"], r=R)
    return b.build(options={"tran_method": "gear2", "icmode": icmode})


STEP = dict(val0=0.0, val1=I_ON, rise=1e-15, fall=1e-15, width=1e-6, period=2e-6, delay=2e-12)

CASES = {
    "constant": (lambda: _cap("dc", dc=I_ON), {}),
    "lte_reject": (lambda: _cap("pulse", **STEP), {}),
    "reject_cap": (lambda: _cap("pulse", **STEP), {"max_consecutive_rejects": 1}),
    "newton_fail_min_dt": (lambda: _cap("pulse", **STEP), {"tran_itl": 1, "min_dt": 5e-14}),
    "dc_start_pulse": (lambda: _cap("pulse", R=1e3, icmode="op", val0=0.2 * I_ON, val1=I_ON,
                                    rise=5e-12, fall=5e-12, width=100e-12, period=2e-6,
                                    delay=20e-12), {}),
}


@pytest.mark.parametrize("case", sorted(CASES))
def test_step_control_matches_oracle(case):
    make, over = CASES[case]
    net = make()
    over = dict(over)
    t_stop = 0.3e-9 if case == "newton_fail_min_dt" else 1e-9
    _, ref = run_oracle(net, t_stop, 1e-12, **over)
    _, res = run_gpu(net, t_stop, 1e-12, **over)
    assert len(res.times) == len(ref.times), (len(res.times), len(ref.times))
    dev = grid_deviation(res.times, ref.times)
    assert dev <= 1e-12, dev
    err = float(np.max(np.abs(res.stats["V_out"] - ref.V)))
    assert err <= 1e-12, err
    assert res.stats["total_nr_iterations"] == ref.stats["total_nr_iterations"]
    assert res.stats["rejected_steps"] == ref.stats["rejected_steps"]


def _overflowing_dag():
    """A 20-gate DAG whose second primary input rises to 1e200 V at 30 ps
    (the other inputs stay low): on every Newton iterate after that edge the
    MOSFET currents, which grow as the square of that voltage, overflow, so
    the Jacobian holds infinite entries."""
    import dataclasses

    from vajax_amd import generators as G

    net = G.gate_dag(20, seed=1)
    vs = net.vsources
    params = {k: np.array(v, dtype=np.float64) for k, v in vs.params.items()}
    params["val1"][2:] = 0.0
    params["val1"][1] = 1e200
    params["delay"][1] = 30e-12
    return dataclasses.replace(net, vsources=dataclasses.replace(vs, params=params))


def test_overflowing_iterate_is_a_newton_failure_not_an_error():
    """A non-finite Jacobian is not a pivoting problem: the pivot monitor
    leaves it alone, the iteration fails the convergence tests (the
    reference's UMFPACK returns the non-finite solution of such a system,
    umfpack_ffi.cpp:297, 320), and the step control rejects the step and
    cuts dt down to min_dt and its force-accepts, taking the oracle's
    decisions step for step instead of the run ending in an error."""
    net = _overflowing_dag()
    _, ref = run_oracle(net, 0.2e-9, 1e-12, max_steps=80)
    _, res = run_gpu(net, 0.2e-9, 1e-12, max_steps=80)
    assert ref.stats["rejected_steps"] >= 30 and ref.stats["min_dt_used"] <= 1e-15
    assert len(res.times) == len(ref.times)
    dev = grid_deviation(res.times, ref.times)
    assert dev <= 1e-12, dev
    assert res.stats["rejected_steps"] == ref.stats["rejected_steps"]
    assert res.stats["total_nr_iterations"] == ref.stats["total_nr_iterations"]
    assert np.all(np.isfinite(res.stats["V_out"]))
    err = float(np.max(np.abs(res.stats["V_out"] - ref.V)))
    assert err <= 1e-9, err
