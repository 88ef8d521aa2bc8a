"""GPU parity of the per-group fused eval + stamp (vjx_eval_stamp) against
the oracle's build restricted to that group (same CSR pattern, the other
groups and the current sources removed), for every group of a diode ladder
(resistor, capacitor, limited diode) and of a ring oscillator (MOSFETs, with
a current source), and the all-groups form against build_system.

Bar: within 1e-12 of each vector's scale (the R/C stamps bit for bit).
"""

import dataclasses

import numpy as np
import pytest

from oracle import mna as omna
from vajax_amd import generators as G
from vajax_amd.engine import CircuitEngine

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def torch():
    import torch as _t

    assert _t.cuda.is_available()
    return _t


def _dev(torch, a):
    return torch.as_tensor(np.ascontiguousarray(a, dtype=np.float64), device="cuda")


def _close(got, want, rel=1e-12):
    got = got.cpu().numpy()[:len(want)]
    scale = max(float(np.max(np.abs(want))), 1e-300) if len(want) else 1.0
    return float(np.max(np.abs(got - want))) <= rel * scale if len(want) else True


@pytest.mark.parametrize("gen", ["diode", "ring"])
def test_eval_stamp_per_group_matches_oracle(torch, gen):
    net = G.diode_ladder(40, seed=3) if gen == "diode" else G.ring_oscillator(11, seed=2)
    cir = omna.setup(net)
    eng = CircuitEngine(net)
    eng.parse()
    dc = eng.device
    rng = np.random.default_rng(5)
    X = np.zeros(cir.n_total + cir.m)
    X[1:cir.n_total] = rng.uniform(0.0, 1.0, cir.n_unk)  # the models' physical range
    X[cir.n_total:] = rng.uniform(-1e-3, 1e-3, cir.m)
    vs = rng.uniform(0.5, 1.5, cir.m)
    isv = rng.uniform(-1e-5, 1e-5, net.isources.count)
    lim = rng.uniform(-0.5, 0.8, cir.n_lim) if cir.n_lim else np.zeros(0)
    c0 = 2e12
    names = list(eng.topology.models)
    assert len(names) >= 2
    for g, name in enumerate(names + [None]):
        if name is None:
            sub, iw = cir, isv
        else:
            sub = dataclasses.replace(cir, models=[om for om in cir.models if om.name == name])
            iw = np.zeros(0)
        want = omna.build_system_lim(sub, X, vs, iw, np.zeros(cir.n_unk), c0, 0.0, 0.0, 0.0,
                                     None, 0.0, None, lim)
        out = dc.eval_stamp(name, _dev(torch, X), _dev(torch, vs), _dev(torch, isv), c0=c0,
                            gmin=1e-12, limit_state=_dev(torch, lim) if cir.n_lim else None)
        for key, w in (("csr_data", want[0]), ("f", want[1]), ("Q", want[2]),
                       ("max_res_contrib", want[4])):
            assert _close(out[key], w), (name, key)
        if cir.n_lim and name in (None, "diode"):
            assert _close(out["limit_state_out"], want[5]), (name, "limit_state")


@pytest.mark.parametrize("model", ["mosfet", "ekv26"])
def test_eval_workgroup_autotuner_keeps_results(model):
    """The transient autotuner also picks each large model group's eval
    workgroup size (256, 128 or 64 lanes): one lane per device, so the run
    with the tuned plan equals the run without tuning bit for bit."""
    from .helpers import run_gpu

    net = G.gate_dag(2500, seed=5, model=model)  # 10 000 MOSFETs: tuned
    eng, a = run_gpu(net, 30e-12, 1e-12, autotune=True, use_graph=True)
    blocks = eng.device.eval_blocks()
    assert blocks[model] in (64, 128, 256), blocks
    _, b = run_gpu(net, 30e-12, 1e-12, autotune=False, use_graph=True)
    assert np.array_equal(a.times, b.times)
    assert np.array_equal(a.stats["V_out"], b.stats["V_out"])
