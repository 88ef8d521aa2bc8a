"""GPU parity of device limiting against the oracle: the limiting functions,
the diode's limited eval (lim_rhs, limit state), the system build with a
limit state, and DC / transient runs of diode circuits with pnjlim active.

Bars: fetlim bit-exact; pnjlim, the diode eval and the build within 1e-13
relative (device exp/log vs numpy); waveforms within 1e-9 V over the run
(fixed step) or over the common prefix of the adaptive time grids.
"""

import ctypes as C

import numpy as np
import pytest

from oracle import limiting as OL
from oracle import mna as omna
from oracle import models as omod
from vajax_amd import generators as G

from .helpers import run_gpu, run_oracle

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def torch():
    import torch as _t

    assert _t.cuda.is_available()
    return _t


def _dev(torch, a):
    return torch.as_tensor(np.ascontiguousarray(a, dtype=np.float64), device="cuda")


@pytest.mark.parametrize("kind", ["pnjlim", "fetlim"])
def test_limit_functions_match_oracle(torch, kind):
    from vajax_amd import _lib as L

    rng = np.random.default_rng(7)
    n = 20000
    vnew = rng.uniform(-4, 4, n)
    vold = rng.uniform(-1.5, 1.5, n)
    vold[:100] = 0.0  # the zero initial state
    out = torch.empty(n, dtype=torch.float64, device="cuda")
    vn, vo = _dev(torch, vnew), _dev(torch, vold)
    if kind == "pnjlim":
        p1, p2 = 0.0287, 0.61
        want = OL.pnjlim(vnew, vold, p1, p2)
    else:
        p1, p2 = 0.45, 0.0
        want = OL.fetlim(vnew, vold, p1)
    rc = L.lib.vjx_limit_eval(0 if kind == "pnjlim" else 1, n, C.c_void_p(vn.data_ptr()),
                              C.c_void_p(vo.data_ptr()), p1, p2, C.c_void_p(out.data_ptr()))
    assert rc == 0, L.last_error()
    got = out.cpu().numpy()
    if kind == "fetlim":
        assert np.array_equal(got, want)
    else:
        assert np.allclose(got, want, rtol=1e-14, atol=1e-16)
        assert np.mean(got != vnew) > 0.3  # the limiter was active


def test_diode_eval_contract_matches_oracle(torch):
    from vajax_amd import _lib as L
    from vajax_amd.models import param_block

    rng = np.random.default_rng(3)
    n = 4099
    p = {"is": rng.uniform(1e-16, 1e-12, n), "n": rng.uniform(1.0, 2.0, n),
         "cj": rng.uniform(0.0, 1e-14, n)}
    vd = rng.uniform(-3, 3, n)
    vold = rng.uniform(-1, 1, n)
    vold[:50] = 0.0
    ref = omod.eval_diode(vd, vold, omod.diode_cache(p))
    par = _dev(torch, param_block("diode", p))
    vbd, lid = _dev(torch, vd[None, :]), _dev(torch, vold)
    outs = [torch.empty((n, 2), dtype=torch.float64, device="cuda") for _ in range(2)]
    outs += [torch.empty((n, 4), dtype=torch.float64, device="cuda") for _ in range(2)]
    lims = [torch.empty((n, 2), dtype=torch.float64, device="cuda") for _ in range(2)]
    lo = torch.empty((n, 1), dtype=torch.float64, device="cuda")
    ptrs = [C.c_void_p(o.data_ptr()) for o in outs + lims + [lo]]
    rc = L.lib.vjx_model_eval(L.MODEL_IDS["diode"], n, C.c_void_p(vbd.data_ptr()),
                              C.c_void_p(par.data_ptr()), C.c_void_p(lid.data_ptr()), *ptrs)
    assert rc == 0, L.last_error()
    for got, want in zip(outs + lims + [lo], ref):
        got = got.cpu().numpy()
        scale = np.max(np.abs(want), axis=1, keepdims=True)
        assert np.all(np.abs(got - want) <= 1e-13 * scale + 1e-300)


def test_build_system_with_limit_state_matches_oracle(torch):
    from vajax_amd.engine import CircuitEngine

    net = G.diode_ladder(60, seed=5)
    cir = omna.setup(net)
    eng = CircuitEngine(net)
    eng.parse()
    dc = eng.device
    assert dc.n_limit_states == cir.n_lim > 0
    rng = np.random.default_rng(2)
    X = np.zeros(cir.n_total + cir.m)
    X[1:cir.n_total] = rng.uniform(-2, 2, cir.n_unk)
    X[cir.n_total:] = rng.uniform(-1e-3, 1e-3, cir.m)
    lim = rng.uniform(-1, 1, cir.n_lim)
    lim[::7] = 0.0
    vs = rng.uniform(-1, 2, cir.m)
    Qp, dQ, Q2 = (rng.uniform(-1e-14, 1e-14, cir.n_unk) for _ in range(3))
    c0, c1, d1, c2 = 1.5e12, -2e12, 0.0, 0.5e12
    want = omna.build_system_lim(cir, X, vs, [], Qp, c0, 0.0, c1, d1, dQ, c2, Q2, lim)
    out = dc.build_system(_dev(torch, X), _dev(torch, vs), None, _dev(torch, Qp),
                          _dev(torch, dQ), _dev(torch, Q2), c0=c0, c1=c1, d1=d1, c2=c2,
                          gmin=1e-12, gshunt=0.0, limit_state=_dev(torch, lim))
    for name, w in zip(["csr_data", "f", "Q", "I_vsource", "max_res_contrib", "limit_state_out"],
                       want):
        g = out[name].cpu().numpy()[:len(w)]
        scale = np.max(np.abs(w)) if len(w) else 1.0
        assert np.max(np.abs(g - w)) <= 1e-12 * scale, name


def test_diode_divider_dc_on_gpu():
    from vajax_amd.netlist import NetlistBuilder

    b = NetlistBuilder("divider")
    b.source("vsource", "v1", "in", "0", type="dc", dc=3.0)
    b.device("resistor", "r1", ["in", "d"], r=470.0)
    b.device("diode", "d1", ["d", "0"], **{"is": 1e-14, "n": 1.0, "cj": 1e-15})
    net = b.build(analysis={"stop": 20e-12, "step": 1e-12},
                  options={"icmode": "op", "tran_method": "gear2"})
    _, ref = run_oracle(net, 20e-12, 1e-12, fixed=True, tran_fs=1.0)
    _, res = run_gpu(net, 20e-12, 1e-12, fixed=True, tran_fs=1.0)
    assert res.stats["dc_converged"]
    err = float(np.max(np.abs(res.stats["V_out"] - ref.V)))
    assert err <= 1e-9, err


@pytest.mark.parametrize("use_graph", [False, True])
def test_diode_ladder_fixed_step_matches_oracle(use_graph):
    net = G.diode_ladder(40, seed=2)
    _, ref = run_oracle(net, 60e-12, 1e-12, fixed=True, tran_fs=1.0)
    _, res = run_gpu(net, 60e-12, 1e-12, fixed=True, tran_fs=1.0, use_graph=use_graph)
    assert np.array_equal(res.times, ref.times)
    assert res.stats["total_nr_iterations"] == ref.stats["total_nr_iterations"]
    err = float(np.max(np.abs(res.stats["V_out"] - ref.V)))
    assert err <= 1e-9, err


def test_diode_ladder_adaptive_matches_oracle():
    net = G.diode_ladder(40, seed=4)
    _, ref = run_oracle(net, 1.5e-9, 1e-12)
    _, res = run_gpu(net, 1.5e-9, 1e-12, use_graph=True)
    n = min(len(ref.times), len(res.times))
    same = np.isclose(res.times[:n], ref.times[:n], rtol=1e-9, atol=0)
    k = n if same.all() else int(np.argmin(same))
    assert k >= 0.5 * len(ref.times), (k, len(ref.times))
    err = float(np.max(np.abs(res.stats["V_out"][:k] - ref.V[:k])))
    assert err <= 1e-9, err
