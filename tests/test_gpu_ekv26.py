"""EKV 2.6 (vajax/devices/models/ekv/ekv.va) on the GPU against the oracle.

Bars: the eval within 1e-12 of each instance's largest output (device
exp/log/sqrt against numpy's, through the model's cancellations); the
assembled system within 1e-12 of its largest entry; node voltages within
1e-9 V on a fixed grid (FP64) and 1e-6 V with the FP32 LU and one FP64
refinement, from a 150-gate DAG up to the 20k-gate configuration; on the
oracle's adaptive grid (replayed) 1e-9 V + 1e-6 |V|.
"""

import ctypes as C
import sys

import numpy as np
import pytest

from oracle import ekv26 as E
from oracle import mna as omna
from vajax_amd import generators as G
from vajax_amd.netlist import EKV26_PARAMS

from .helpers import run_gpu, run_oracle

pytestmark = pytest.mark.gpu


def _say(msg):
    print(msg, file=sys.stderr, flush=True)


@pytest.fixture(scope="module")
def torch():
    import torch as _t

    assert _t.cuda.is_available()
    return _t


def _dev(torch, a):
    return torch.as_tensor(np.ascontiguousarray(a, dtype=np.float64), device="cuda")


def _devices(n, rng):
    typ = np.where(rng.random(n) < 0.5, 1.0, -1.0)
    p = {k: np.full(n, float(v)) for k, v in EKV26_PARAMS.items()}
    p.update(type=typ, w=rng.uniform(0.2e-6, 2e-6, n), l=rng.uniform(45e-9, 1e-6, n),
             vto=typ * rng.uniform(0.2, 0.5, n), cox=np.full(n, 0.015),
             xj=np.full(n, 30e-9), ucrit=rng.uniform(2e6, 8e6, n), hdif=np.full(n, 0.1e-6),
             xd_cj=np.full(n, 1e-3), xd_cjsw=np.full(n, 1e-10), xd_cjswg=np.full(n, 1e-10),
             xd_vts=np.full(n, 0.3), xd_vtssw=np.full(n, 0.2), xd_vtsswg=np.full(n, 0.25),
             xd_js=np.full(n, 1e-4), xd_gmin=np.full(n, 1e-12), xd_xjbv=np.full(n, 1.0),
             xd_bv=np.full(n, 1.5), trise=rng.uniform(-20, 60, n))
    # the THETA mobility form on a quarter of the devices
    p["e0"][: n // 4] = 0.0
    p["theta"][: n // 4] = 0.1
    return p


def test_ekv26_eval_matches_oracle(torch):
    from vajax_amd import _lib as L
    from vajax_amd.models import param_block

    _say("ekv26: raw eval of 8191 devices")
    rng = np.random.default_rng(21)
    n = 8191
    p = _devices(n, rng)
    vb = rng.uniform(-1.6, 1.6, (3, n))
    vb[:, :16] = 0.0  # VD = VS: the Mode switch point
    vb[0, 16:64] = -3.0  # deep weak inversion: the last interpolation pieces
    vb[2, 64:96] = vb[1, 64:96] - 1e-9  # just on the reverse side
    ref = E.eval_ekv26(vb[0], vb[1], vb[2], E.cache(p))
    par = _dev(torch, param_block("ekv26", p))
    vbd = _dev(torch, vb)
    outs = [torch.empty((n, 4), dtype=torch.float64, device="cuda") for _ in range(2)]
    outs += [torch.empty((n, 16), dtype=torch.float64, device="cuda") for _ in range(2)]
    rc = L.lib.vjx_model_eval(L.MODEL_IDS["ekv26"], n, C.c_void_p(vbd.data_ptr()),
                              C.c_void_p(par.data_ptr()), None,
                              *[C.c_void_p(o.data_ptr()) for o in outs], None, None, None)
    assert rc == 0, L.last_error()
    for name, got, want in zip(("res_resist", "res_react", "jac_resist", "jac_react"), outs, ref):
        got = got.cpu().numpy()
        assert np.all(np.isfinite(got)), name
        scale = np.max(np.abs(want), axis=1, keepdims=True)
        err = np.abs(got - want) / (scale + 1e-300)
        _say(f"ekv26 eval {name}: max error {float(err.max()):.2e} of the row scale")
        assert np.all(err <= 1e-12), (name, np.argwhere(err > 1e-12)[:5])


def test_ekv26_build_system_matches_oracle(torch):
    from vajax_amd.engine import CircuitEngine

    _say("ekv26: build_system of a 300-gate DAG")
    net = G.gate_dag(300, seed=2, model="ekv26")
    cir = omna.setup(net)
    eng = CircuitEngine(net)
    eng.parse()
    rng = np.random.default_rng(5)
    X = np.zeros(cir.n_total + cir.m)
    X[1:cir.n_total] = rng.uniform(-0.2, 1.2, cir.n_unk)
    X[cir.n_total:] = rng.uniform(-1e-3, 1e-3, cir.m)
    vs = rng.uniform(0, 1, cir.m)
    Qp, dQ, Q2 = (rng.uniform(-1e-15, 1e-15, cir.n_unk) for _ in range(3))
    c0, c1, d1, c2, gsh = 1.5e12, -2e12, 0.0, 0.5e12, 1e-9
    want = omna.build_system(cir, X, vs, np.zeros(0), Qp, c0, gsh, c1, d1, dQ, c2, Q2)
    out = eng.device.build_system(_dev(torch, X), _dev(torch, vs), None, _dev(torch, Qp),
                                  _dev(torch, dQ), _dev(torch, Q2), c0=c0, c1=c1, d1=d1, c2=c2,
                                  gmin=1e-12, gshunt=gsh)
    for name, w in zip(["csr_data", "f", "Q", "I_vsource", "max_res_contrib"], want):
        g = out[name].cpu().numpy()[: len(w)]
        scale = np.max(np.abs(w)) if len(w) else 1.0
        assert np.max(np.abs(g - w)) <= 1e-12 * scale, name


def test_ekv26_gate_dag_fixed_step_fp64_and_mixed():
    net = G.gate_dag(150, seed=7, model="ekv26")
    _say("ekv26 dag 150: oracle 150 fixed steps")
    _, ref = run_oracle(net, 0.15e-9, 1e-12, fixed=True, tran_fs=1.0)
    _say("ekv26 dag 150: gpu fp64")
    _, r64 = run_gpu(net, 0.15e-9, 1e-12, fixed=True, tran_fs=1.0, use_graph=True)
    assert np.array_equal(r64.times, ref.times)
    e64 = float(np.max(np.abs(r64.stats["V_out"] - ref.V)))
    _say(f"ekv26 dag 150: fp64 max |dV| {e64:.3e} V")
    assert e64 <= 1e-9, e64
    _, r32 = run_gpu(net, 0.15e-9, 1e-12, fixed=True, tran_fs=1.0, precision="f32",
                     use_graph=True)
    e32 = float(np.max(np.abs(r32.stats["V_out"] - ref.V)))
    _say(f"ekv26 dag 150: mixed max |dV| {e32:.3e} V")
    assert e32 <= 1e-6, e32


def test_ekv26_gate_dag_adaptive_on_the_oracle_grid():
    """The adaptive Gear2 run through the first input edges, replayed on the
    oracle's accepted grid (each step's Newton solve and LTE estimate are
    the GPU's own; only the step sizes are the oracle's)."""
    net = G.gate_dag(60, seed=1, t_stop=0.6e-9, model="ekv26")
    _say("ekv26 dag 60: oracle adaptive 0.6 ns")
    _, ref = run_oracle(net, 0.6e-9, 1e-12)
    assert ref.stats["rejected_steps"] > 0  # the LTE control was exercised
    _say(f"ekv26 dag 60: gpu replaying {len(ref.times)} points")
    _, rr = run_gpu(net, ref.times[-1], 1e-12, use_graph=True, replay_times=ref.times)
    assert np.array_equal(rr.times, ref.times)
    assert rr.stats["rejected_steps"] == 0
    err = np.abs(rr.stats["V_out"] - ref.V)
    _say(f"ekv26 dag 60: max |dV| {float(err.max()):.3e} V")
    assert float(np.max(err - 1e-6 * np.abs(ref.V))) <= 1e-9
    # Free-running, the two step sequences part within a few dozen steps
    # (decisions within rounding of their thresholds: the oracle against
    # itself with 1e-15 relative noise in its solves parts after 28 steps),
    # and the waveforms, sampled on different grids across 20 ps edges,
    # cannot be compared point by point.  The bar is the logic: every node
    # crosses VDD/2 the same number of times, at times within 5 ps (the
    # oracle against its perturbed self: 1.1 ps, one node with a glitch).
    _, res = run_gpu(net, 0.6e-9, 1e-12, use_graph=True)
    ca = _crossings(ref.times, ref.V)
    cb = _crossings(res.times, res.stats["V_out"])
    same = [len(a) == len(b) for a, b in zip(ca, cb)]
    dt = max((float(np.max(np.abs(a - b))) for a, b, ok in zip(ca, cb, same) if ok and len(a)),
             default=0.0)
    _say(f"ekv26 dag 60: {sum(len(a) for a in ca)} crossings, {same.count(False)} nodes "
         f"with another count, max shift {dt:.3e} s")
    assert sum(len(a) for a in ca) > 50
    assert same.count(False) <= 2 and dt <= 5e-12


def _crossings(times, V, level=0.5):
    """Times at which each node crosses ``level`` (linear interpolation)."""
    out = []
    for j in range(V.shape[1]):
        v = V[:, j] - level
        s = np.nonzero(np.sign(v[:-1]) * np.sign(v[1:]) < 0)[0]
        out.append(times[s] - v[s] * (times[s + 1] - times[s]) / (v[s + 1] - v[s]))
    return out


def _period(times, v, level=0.5):
    up = np.nonzero((v[:-1] < level) & (v[1:] >= level))[0]
    t = times[up] + (level - v[up]) * (times[up + 1] - times[up]) / (v[up + 1] - v[up])
    return np.diff(t)


def test_ekv26_ring_oscillator():
    """11-stage ring of EKV 2.6 inverters: on the oracle's grid to the FP64
    bar, and free-running with the same period to 1 %."""
    net = G.ring_oscillator(11, seed=2, t_stop=2e-9, model="ekv26")
    _say("ekv26 ring 11: oracle adaptive 2 ns")
    cir, ref = run_oracle(net, 2e-9, 1e-12)
    _, rr = run_gpu(net, ref.times[-1], 1e-12, use_graph=True, replay_times=ref.times)
    assert np.array_equal(rr.times, ref.times) and rr.stats["rejected_steps"] == 0
    err = np.abs(rr.stats["V_out"] - ref.V)
    _say(f"ekv26 ring 11: max |dV| on the oracle grid {float(err.max()):.3e} V")
    assert float(np.max(err - 1e-6 * np.abs(ref.V))) <= 1e-9
    _, res = run_gpu(net, 2e-9, 1e-12, use_graph=True)
    j = cir.node_names["out003"]
    pr = _period(ref.times, ref.V[:, j])
    pg = _period(res.times, res.stats["V_out"][:, j])
    _say(f"ekv26 ring 11: periods oracle {pr} gpu {pg}")
    assert len(pr) >= 2 and len(pg) == len(pr)
    assert np.all(np.abs(pg - pr) <= 1e-2 * pr)


def test_ekv26_packed_batch_fixed_step_fp64_and_mixed():
    """Monte Carlo copies and a ragged topology of EKV 2.6 DAGs packed into
    one block-diagonal system, on a fixed 1 ps grid against the oracle run of
    the same packed netlist."""
    net = G.batch_mc(3, 60, ragged=(40, 90), seed=3, model="ekv26")
    _say("ekv26 batch: oracle 100 fixed steps")
    _, ref = run_oracle(net, 0.1e-9, 1e-12, fixed=True, tran_fs=1.0)
    _, r64 = run_gpu(net, 0.1e-9, 1e-12, fixed=True, tran_fs=1.0, use_graph=True)
    assert np.array_equal(r64.times, ref.times)
    e64 = float(np.max(np.abs(r64.stats["V_out"] - ref.V)))
    _, r32 = run_gpu(net, 0.1e-9, 1e-12, fixed=True, tran_fs=1.0, precision="f32",
                     use_graph=True)
    e32 = float(np.max(np.abs(r32.stats["V_out"] - ref.V)))
    _say(f"ekv26 batch: max |dV| fp64 {e64:.3e} V, mixed {e32:.3e} V")
    assert e64 <= 1e-9 and e32 <= 1e-6, (e64, e32)


def test_ekv26_config_dag_20k_fixed_fp64_and_mixed():
    """The 20k-gate DAG with every MOSFET the EKV 2.6 model (80 000 devices,
    40 130 unknowns), 25 fixed steps through the supply ramp, against the
    oracle: FP64 within 1e-9 V, FP32 LU + FP64 refinement within 1e-6 V."""
    net = G.gate_dag(20000, seed=0, model="ekv26")
    _say("ekv26 dag 20k: oracle 25 fixed steps (about a minute)")
    _, ref = run_oracle(net, 25e-12, 1e-12, fixed=True, tran_fs=1.0)
    _say("ekv26 dag 20k: gpu fp64")
    _, r64 = run_gpu(net, 25e-12, 1e-12, fixed=True, tran_fs=1.0, use_graph=True)
    assert np.array_equal(r64.times, ref.times)
    e64 = float(np.max(np.abs(r64.stats["V_out"] - ref.V)))
    _say("ekv26 dag 20k: gpu mixed")
    _, r32 = run_gpu(net, 25e-12, 1e-12, fixed=True, tran_fs=1.0, precision="f32",
                     use_graph=True)
    e32 = float(np.max(np.abs(r32.stats["V_out"] - ref.V)))
    _say(f"ekv26 dag 20k: max |dV| fp64 {e64:.3e} V, mixed {e32:.3e} V")
    assert e64 <= 1e-9 and e32 <= 1e-6, (e64, e32)


def test_ekv26_array_multiplier_4x4():
    """The reference's multiplier testbench topology at 4x4 (the generator of
    its mul64 benchmark) with EKV 2.6 MOSFETs: the oracle's adaptive grid
    replayed to the FP64 bar, and the free-running run's node voltages
    within the rails."""
    from vajax_amd import vacask

    net = G.with_ekv26(vacask.load("", text=G.array_multiplier_testbench(
        4, [1, 1, 0, 1], [0, 1, 1, 0], t_stop=2e-9)))
    assert net.groups["ekv26"].count > 0 and "mosfet" not in net.groups
    _say("ekv26 mul4: oracle adaptive 2 ns")
    _, ref = run_oracle(net, 2e-9, 2e-12, lte_ratio=1.5)
    _, rr = run_gpu(net, ref.times[-1], 2e-12, use_graph=True, replay_times=ref.times,
                    lte_ratio=1.5)
    assert np.array_equal(rr.times, ref.times) and rr.stats["rejected_steps"] == 0
    err = np.abs(rr.stats["V_out"] - ref.V)
    _say(f"ekv26 mul4: max |dV| on the oracle grid {float(err.max()):.3e} V")
    assert float(np.max(err - 1e-6 * np.abs(ref.V))) <= 1e-9
    _, res = run_gpu(net, 2e-9, 2e-12, use_graph=True, lte_ratio=1.5)
    V = res.stats["V_out"]
    assert abs(res.times[-1] - 2e-9) <= 1e-15 and np.all(np.isfinite(V))
    assert float(V.min()) >= -0.5 and float(V.max()) <= 1.2 + 0.5
