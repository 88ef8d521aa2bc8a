"""The HIP path against the committed golden fixtures (tests/golden/, made by
the oracle and checked by tests/test_golden_cpu.py): waveforms, Newton
iteration totals and the homotopy DC point, without running the oracle.

Bars: the RC ladder within 1e-12 V with equal iteration counts; the gate DAG
within 1e-9 V; the diode ladder within 1e-9 V over the common prefix of the
adaptive time grids; the DC point within 1e-9 V with the same method,
iteration total and continuation steps.
"""

import os

import numpy as np
import pytest

from vajax_amd import dc as D
from vajax_amd import generators as G
from vajax_amd.engine import CircuitEngine, SimulationOptions
from vajax_amd.topology import build_topology

from .golden import make_golden as MG
from .helpers import run_gpu

pytestmark = pytest.mark.gpu

HERE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def _load(name):
    with np.load(os.path.join(HERE, name + ".npz"), allow_pickle=False) as z:
        return {k: z[k] for k in z.files}


@pytest.mark.parametrize("name", sorted(MG.CASES))
def test_hip_path_matches_golden_transients(name):
    make, t_stop, dt, fixed = MG.CASES[name]
    want = _load(name)
    _, res = run_gpu(make(), t_stop, dt, fixed=fixed, tran_fs=1.0 if fixed else None,
                     use_graph=not fixed)
    V = res.stats["V_out"]
    if fixed:
        assert np.array_equal(res.times, want["times"])
        tol = 1e-12 if name.startswith("rc") else 1e-9
        if name.startswith("rc"):
            assert res.stats["total_nr_iterations"] == want["total_nr_iterations"][0]
        assert float(np.max(np.abs(V - want["V"]))) <= tol
        return
    n = min(len(res.times), len(want["times"]))
    same = np.isclose(res.times[:n], want["times"][:n], rtol=1e-9, atol=0)
    k = n if same.all() else int(np.argmin(same))
    assert k >= 0.5 * len(want["times"]), (k, len(want["times"]))
    assert float(np.max(np.abs(V[:k] - want["V"][:k]))) <= 1e-9


def test_hip_homotopy_matches_golden_dc_point():
    want = _load("dc_clamp_homotopy")
    net = G.diode_clamp_chain()
    eng = CircuitEngine(net)
    eng.parse()
    topo = build_topology(net)
    V0, _ = D.initialize_dc_voltages(topo, net, D.get_vdd_value(net))
    X0 = np.zeros(topo.n_total + topo.n_vsources)
    X0[:topo.n_total] = V0
    vs, isv = D.get_dc_source_values(net)
    X, r = D.compute_dc_operating_point(eng.device.make_nr_solve(SimulationOptions(), 12), X0,
                                        vs, isv, np.zeros(topo.n_unknowns), SimulationOptions())
    X = X.cpu().numpy()[:topo.n_total]
    assert r.method.encode() == bytes(want["method"])
    assert (r.iterations, r.homotopy_steps) == (want["iterations"][0], want["homotopy_steps"][0])
    assert float(np.max(np.abs(X - want["X"]))) <= 1e-9
