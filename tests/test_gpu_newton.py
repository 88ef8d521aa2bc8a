"""GPU parity of the single Newton solve (vjx_newton_solve, the reference's
nr_solve 9-tuple of solver_factories.py:604-727) against the oracle's
nr_solve: a DC solve of the diode ladder from the mid-rail guess, a
backward-Euler solve of the gate DAG, a BDF2-like transient solve of the diode
ladder with a limit state and a residual floor, and an iteration cap that
stops short of convergence.

Bars: equal iteration count and converged flag; X, Q, dQdt, I_vsource,
limit state and max_res_contrib within 1e-9 relative to each vector's scale
(the LU's pivot order differs from the oracle's, so the rounding does).
"""

import numpy as np
import pytest

from oracle import mna as omna
from oracle import transient as otr
from vajax_amd import generators as G
from vajax_amd.engine import CircuitEngine, SimulationOptions

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def torch():
    import torch as _t

    assert _t.cuda.is_available()
    return _t


def _dev(torch, a):
    if a is None:
        return None
    return torch.as_tensor(np.ascontiguousarray(a, dtype=np.float64), device="cuda")


def _close(got, want, rel=1e-9):
    got = got.cpu().numpy()[:len(want)] if hasattr(got, "cpu") else np.asarray(got)
    scale = max(float(np.max(np.abs(want))) if len(want) else 0.0, 1e-300)
    return float(np.max(np.abs(got - want))) <= rel * scale if len(want) else True


def _compare(torch, net, X0, vs, isv, Qp, c0, c1=0.0, d1=0.0, dQ=None, c2=0.0, Q2=None,
             lim=None, floor=None, tran_itl=20):
    cir = omna.setup(net)
    opts = otr.Options()
    opts.tran_itl = tran_itl
    want = otr.nr_solve(cir, otr.LinearSolver(cir), opts, X0, vs, isv, Qp, c0, 0.0, c1, d1, dQ,
                        c2, Q2, res_tol_floor=floor, lim_in=lim)
    eng = CircuitEngine(net)
    eng.parse()
    o = SimulationOptions()
    o.tran_itl = tran_itl
    got = eng.device.nr_solve(_dev(torch, X0), _dev(torch, vs), _dev(torch, isv),
                              _dev(torch, Qp), c0, gmin=1e-12, gshunt=0.0, integ_c1=c1,
                              integ_d1=d1, dQdt_prev=_dev(torch, dQ), integ_c2=c2,
                              Q_prev2=_dev(torch, Q2), limit_state_in=_dev(torch, lim),
                              res_tol_floor=_dev(torch, floor), options=o)
    X, iters, conv, max_f, Q, dQdt, Ivs, ls, mrc = got
    assert iters == want.iterations and conv == want.converged
    # max_f of a converged solve is rounding noise (~1e-11 A here): absolute bar
    assert max_f == pytest.approx(want.max_f, rel=1e-6, abs=1e-12)
    for name, g, w in (("X", X, want.X), ("Q", Q, want.Q), ("dQdt", dQdt, want.dQdt),
                       ("I_vsource", Ivs, want.I_vsource), ("mrc", mrc, want.max_res_contrib)):
        assert _close(g, w), name
    if cir.n_lim:
        assert _close(ls, want.lim), "limit_state"
    return want


def test_dc_newton_solve_of_the_diode_ladder(torch):
    net = G.diode_ladder(60, seed=3)
    cir = omna.setup(net)
    X0 = np.zeros(cir.n_total + cir.m)
    X0[:cir.n_total] = otr._init_mid_rail(cir)
    vs = otr.source_values(net.vsources, 0.0)
    want = _compare(torch, net, X0, vs, np.zeros(0), np.zeros(cir.n_unk), 0.0)
    assert want.converged and want.iterations >= 3


def test_transient_newton_solve_of_the_gate_dag(torch):
    net = G.gate_dag(400, seed=3)
    cir = omna.setup(net)
    rng = np.random.default_rng(8)
    X0 = np.zeros(cir.n_total + cir.m)
    X0[1:cir.n_total] = rng.uniform(0.0, 1.0, cir.n_unk)
    vs = otr.source_values(net.vsources, 0.5e-9)
    Qp = omna.build_system(cir, X0, vs, np.zeros(0), np.zeros(cir.n_unk), 0.0)[2]
    h = 1e-12
    want = _compare(torch, net, X0, vs, np.zeros(0), Qp, 1.0 / h, -1.0 / h)
    assert want.converged and want.iterations >= 2


def test_transient_newton_solve_of_the_diode_ladder(torch):
    net = G.diode_ladder(50, seed=1)
    cir = omna.setup(net)
    rng = np.random.default_rng(4)
    X0 = np.zeros(cir.n_total + cir.m)
    X0[1:cir.n_total] = rng.uniform(-0.3, 0.9, cir.n_unk)
    vs = np.array([1.7])
    Qp = omna.build_system(cir, X0, vs, np.zeros(0), np.zeros(cir.n_unk), 0.0)[2]
    dQ = rng.uniform(-1e-4, 1e-4, cir.n_unk)
    Q2 = Qp * 0.98
    h = 2e-12
    c0, c1, c2 = 1.5 / h, -2.0 / h, 0.5 / h
    lim = rng.uniform(-0.2, 0.8, cir.n_lim)
    floor = np.full(cir.n_unk, 1e-11)
    want = _compare(torch, net, X0, vs, np.zeros(0), Qp, c0, c1, 0.0, dQ, c2, Q2, lim, floor)
    assert want.converged


def test_newton_solve_iteration_cap(torch):
    net = G.diode_ladder(30, seed=2)
    cir = omna.setup(net)
    X0 = np.zeros(cir.n_total + cir.m)
    want = _compare(torch, net, X0, np.array([2.0]), np.zeros(0), np.zeros(cir.n_unk), 0.0,
                    tran_itl=2)
    assert not want.converged and want.iterations == 2
