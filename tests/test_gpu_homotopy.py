"""GPU parity of the DC operating point with the homotopy chain: the same
host step control driven by the device Newton solve (vjx_newton_solve) and by
the oracle's, on a diode clamp whose DC point needs source stepping after
gdev and gshunt stepping give up, one where the whole chain fails, and the
engine's compute_dc_operating_point.

Bars: the same method, Newton iteration total and continuation step count
(every converged/failed decision agrees); final node voltages within 1e-9 V.
"""

import numpy as np
import pytest

from vajax_amd import dc as D
from vajax_amd import generators as G
from vajax_amd.engine import CircuitEngine, SimulationOptions

from .test_homotopy_cpu import _clamp_start, oracle_nr_solve

pytestmark = pytest.mark.gpu


def _both(net, cap, options):
    cir, X0, vs, isv = _clamp_start(net)
    Xo, ro = D.compute_dc_operating_point(oracle_nr_solve(cir, cap), X0, vs, isv,
                                          np.zeros(cir.n_unk), options)
    eng = CircuitEngine(net)
    eng.parse()
    nr = eng.device.make_nr_solve(options, cap)
    Xg, rg = D.compute_dc_operating_point(nr, X0, vs, isv, np.zeros(cir.n_unk), options)
    Xg = Xg.cpu().numpy() if hasattr(Xg, "cpu") else np.asarray(Xg)
    return cir, (Xo, ro), (Xg, rg)


@pytest.mark.parametrize("cap,method", [(12, "chain_source_stepping"), (10, "chain_failed"),
                                        (20, "direct")])
def test_clamp_dc_chain_matches_oracle(cap, method):
    net = G.diode_clamp_chain()
    cir, (Xo, ro), (Xg, rg) = _both(net, cap, SimulationOptions())
    assert ro.method == method
    assert (rg.method, rg.converged) == (ro.method, ro.converged)
    assert (rg.iterations, rg.homotopy_steps) == (ro.iterations, ro.homotopy_steps)
    err = float(np.max(np.abs(Xg[:cir.n_total] - Xo[:cir.n_total])))
    assert err <= 1e-9, err


def test_gshunt_only_chain_matches_oracle():
    net = G.diode_clamp_chain(stages=2, r=10.0)
    o = SimulationOptions()
    o.op_homotopy = ("gshunt",)
    cir, (Xo, ro), (Xg, rg) = _both(net, 12, o)
    assert ro.converged and ro.method == "chain_gshunt_stepping"
    assert (rg.method, rg.converged, rg.iterations, rg.homotopy_steps) == \
        (ro.method, ro.converged, ro.iterations, ro.homotopy_steps)
    err = float(np.max(np.abs(Xg[:cir.n_total] - Xo[:cir.n_total])))
    assert err <= 1e-9, err


def test_engine_dc_operating_point():
    net = G.diode_clamp_chain()
    eng = CircuitEngine(net)
    eng.options.op_itl = 12
    V = eng.compute_dc_operating_point()
    assert eng.dc_result.converged and eng.dc_result.method == "chain_source_stepping"
    cir, X0, vs, isv = _clamp_start(net)
    direct = oracle_nr_solve(cir, 100)(X0, vs, isv, np.zeros(cir.n_unk), 0.0)
    assert float(np.max(np.abs(V - direct[0][:cir.n_total]))) <= 1e-9


@pytest.mark.parametrize("which", ["ring", "dag"])
def test_ekv26_dc_operating_points_match_oracle(which):
    """DC operating points of EKV 2.6 circuits: an 11-stage ring (plain
    Newton from the mid-rail guess) and a 30-gate DAG (its stack nodes held
    only by gmin: the chain's gshunt stepping), the device Newton solve
    against the oracle's under the same host step control."""
    net = (G.ring_oscillator(11, seed=2, model="ekv26") if which == "ring"
           else G.gate_dag(30, seed=3, model="ekv26"))
    cir, (Xo, ro), (Xg, rg) = _both(net, 50, SimulationOptions())
    assert ro.converged and ro.method == ("direct" if which == "ring" else "chain_gshunt_stepping")
    assert (rg.method, rg.converged) == (ro.method, ro.converged)
    assert (rg.iterations, rg.homotopy_steps) == (ro.iterations, ro.homotopy_steps)
    err = float(np.max(np.abs(Xg[:cir.n_total] - Xo[:cir.n_total])))
    assert err <= 1e-9, err
