"""GPU parity at the benchmark configurations' full sizes.

Adaptive runs of nonlinear circuits take LTE decisions on voltages that the
GPU and CPU Newton loops reproduce only to rounding, so their accepted grids
can part after many steps.  Where a whole window is compared, the GPU
replays the oracle's accepted time grid (same coefficients, predictor and
history; vjx_tran_config.replay_times_host) and the node voltages are
compared at every step.  For the gate DAGs the oracle is its C++ restatement
(oracle/ctran.cpp, equal to the numpy oracle's grid and within 1.1e-13 V of
its voltages over the first 24 steps of the 65k-gate DAG).

* RC ladder, 200 sections, 10 000 fixed BDF2 steps: the whole waveform
  against the oracle (a linear recurrence: agreement to rounding).
* 101-stage ring oscillator, adaptive from the kick-start: common prefix of
  accepted time points against the oracle, the oscillation period (to 1 %)
  over 40 ns, and the whole 40 ns on the oracle's replayed grid.
* 20k-gate DAG: fixed 1 ps grid through the supply ramp, FP64 and mixed;
  200 adaptive steps on the oracle's grid, FP64 and mixed.
* 65k-gate DAG (~130k unknowns): the adaptive run's common prefix with the
  oracle, and the benchmark's whole run (300 untimed and 100 timed accepted
  steps) on the oracle's grid, FP64 and mixed.
* Batched: the packed Monte Carlo + ragged batch (260 instances, ~1.07 M
  unknowns) on a fixed grid, FP64 and mixed, instance by instance against
  standalone runs and the oracle; the adaptive batch (one common grid) against
  standalone runs replaying that grid.
"""

import sys

import numpy as np
import pytest

from oracle import mna as omna
from oracle.ctran import CTran
from vajax_amd import generators as G

from .helpers import oracle_configs, run_gpu, run_oracle

pytestmark = pytest.mark.gpu


def _say(msg):
    print(msg, file=sys.stderr, flush=True)


def test_config_rc_ladder_10k_fixed_steps():
    net = G.rc_ladder(200, seed=0)
    _say("rc ladder 200 x 10k steps: oracle")
    _, ref = run_oracle(net, 10e-9, 1e-12, fixed=True, tran_fs=1.0)
    _say("rc ladder: gpu")
    _, res = run_gpu(net, 10e-9, 1e-12, fixed=True, tran_fs=1.0, use_graph=True)
    assert len(ref.times) == 10001
    assert np.array_equal(res.times, ref.times)
    err = float(np.max(np.abs(res.stats["V_out"] - ref.V)))
    assert err <= 1e-12, err
    assert res.stats["total_nr_iterations"] == ref.stats["total_nr_iterations"]


def _ctran_adaptive(net, t_stop, dt, max_steps):
    cir = omna.setup(net)
    cfg, opts = oracle_configs(net)
    out = CTran(cir).run_adaptive(t_stop, dt, cfg, opts, max_steps=max_steps, threads=16)
    return cir, out


def _within(a, b, tag):
    """FP64 bar of the transient core: |dV| <= 1e-9 V + 1e-6 |V| at every node
    and step (Newton's own tolerance is 1e-6 V + 1e-3 |V|)."""
    err = np.abs(a - b)
    worst = float(np.max(err - 1e-6 * np.abs(b)))
    _say(f"{tag}: max |dV| = {float(np.max(err)):.3e} V")
    assert worst <= 1e-9, (tag, float(np.max(err)), worst)
    return float(np.max(err))


def _prefix(res, ref):
    n = min(len(ref.times), len(res.times))
    same = np.isclose(res.times[:n], ref.times[:n], rtol=1e-9, atol=0)
    return n if same.all() else int(np.argmin(same))


def _period(t, v, level=0.5):
    """Mean period from rising crossings of `level` (linear interpolation)."""
    s = v - level
    idx = np.where((s[:-1] < 0) & (s[1:] >= 0))[0]
    tc = t[idx] - s[idx] * (t[idx + 1] - t[idx]) / (s[idx + 1] - s[idx])
    return np.diff(tc)


def test_config_ring_101_adaptive():
    net = G.ring_oscillator(101, seed=0)
    t_stop = 40e-9
    _say("ring 101: oracle")
    cir, ref = run_oracle(net, t_stop, 1e-12)
    _say(f"ring 101: {len(ref.times)} oracle steps; gpu")
    eng, res = run_gpu(net, t_stop, 1e-12, use_graph=True)
    n = min(len(ref.times), len(res.times))
    same = np.isclose(res.times[:n], ref.times[:n], rtol=1e-9, atol=0)
    k = n if same.all() else int(np.argmin(same))
    assert k >= 50, k
    # on the common grid the two Newton loops may stop at different iterates
    # inside the Newton voltage tolerance (vntol = 1e-6 V); over thousands of
    # steps of a free-running oscillator that is the bar
    err = float(np.max(np.abs(res.stats["V_out"][:k] - ref.V[:k])))
    assert err <= 1e-6, (k, err)
    # the oscillation itself: same period to 1 % (both runs keep their own
    # LTE-controlled grids after the common prefix)
    j = cir.node_names[net.nodes[3]]
    pr = _period(ref.times, ref.V[:, j])
    pg = _period(res.times, res.stats["V_out"][:, j])
    assert len(pr) >= 2 and len(pg) >= 2, (len(pr), len(pg))
    assert abs(np.mean(pg[1:]) - np.mean(pr[1:])) <= 1e-2 * np.mean(pr[1:])
    # the whole run on the oracle's grid
    _say("ring 101: gpu replaying the oracle grid")
    _, rr = run_gpu(net, t_stop, 1e-12, use_graph=True, replay_times=ref.times)
    assert len(rr.times) == len(ref.times)
    assert rr.stats["rejected_steps"] == 0  # every replayed Newton solve converged
    err = float(np.max(np.abs(rr.stats["V_out"] - ref.V)))
    assert err <= 1e-6, err


def test_config_dag_20k_fixed_fp64_and_mixed():
    net = G.gate_dag(20000, seed=0)
    _say("dag 20k: oracle 25 fixed steps")
    _, ref = run_oracle(net, 25e-12, 1e-12, fixed=True, tran_fs=1.0)
    _say("dag 20k: gpu fp64")
    _, r64 = run_gpu(net, 25e-12, 1e-12, fixed=True, tran_fs=1.0, use_graph=True)
    assert np.array_equal(r64.times, ref.times)
    e64 = float(np.max(np.abs(r64.stats["V_out"] - ref.V)))
    assert e64 <= 1e-9, e64
    _say("dag 20k: gpu mixed")
    _, r32 = run_gpu(net, 25e-12, 1e-12, fixed=True, tran_fs=1.0, precision="f32",
                     use_graph=True)
    e32 = float(np.max(np.abs(r32.stats["V_out"] - ref.V)))
    assert e32 <= 1e-6, e32


def _replay_fp64_and_mixed(net, ref, cir, tag):
    _say(f"{tag}: gpu fp64 on the oracle grid ({len(ref.times)} points)")
    eng, r64 = run_gpu(net, ref.times[-1], 1e-12, use_graph=True, replay_times=ref.times)
    assert np.array_equal(r64.times, ref.times)
    assert r64.stats["rejected_steps"] == 0
    e64 = _within(r64.stats["V_out"], ref.V, f"{tag} fp64 replay")
    _say(f"{tag}: gpu mixed on the oracle grid")
    _, r32 = run_gpu(net, ref.times[-1], 1e-12, use_graph=True, replay_times=ref.times,
                     precision="f32")
    assert r32.stats["rejected_steps"] == 0
    e32 = float(np.max(np.abs(r32.stats["V_out"] - ref.V)))
    assert e32 <= 1e-6, e32
    _say(f"{tag}: max |dV| fp64 {e64:.3e} V, mixed {e32:.3e} V")
    return r64, r32


def test_config_dag_20k_200_adaptive_steps_fp64_and_mixed():
    net = G.gate_dag(20000, seed=0)
    _say("dag 20k: oracle (C++) 200 adaptive steps")
    cir, ref = _ctran_adaptive(net, 1e-9, 1e-12, max_steps=201)
    assert len(ref.times) == 201
    _say("dag 20k: gpu adaptive")
    _, res = run_gpu(net, 1e-9, 1e-12, max_steps=201, use_graph=True)
    k = _prefix(res, ref)
    assert k >= 20, k
    _within(res.stats["V_out"][:k], ref.V[:k], f"dag 20k adaptive, common prefix {k}")
    _replay_fp64_and_mixed(net, ref, cir, "dag 20k")


def test_config_dag_65k_bench_steps_fp64_and_mixed():
    # the benchmark's whole run: 300 untimed accepted steps and the 100 timed
    # ones after them
    net = G.gate_dag(65000, seed=0)
    _say("dag 65k: oracle (C++) 400 adaptive steps")
    cir, ref = _ctran_adaptive(net, 1e-9, 1e-12, max_steps=401)
    assert len(ref.times) == 401
    _say("dag 65k: gpu adaptive")
    eng, res = run_gpu(net, 1e-9, 1e-12, max_steps=401, use_graph=True, autotune=True)
    assert eng.topology.n_augmented > 130000
    k = _prefix(res, ref)
    assert k >= 10, k
    _within(res.stats["V_out"][:k], ref.V[:k], f"dag 65k adaptive, common prefix {k}")
    _replay_fp64_and_mixed(net, ref, cir, "dag 65k")


def _cols(names_b, n1, net, b):
    cols_b = [names_b[f"i{b:03d}.{nm}"] for nm in net.nodes if nm != net.ground]
    cols_1 = [n1[nm] for nm in net.nodes if nm != net.ground]
    return cols_b, cols_1


def test_config_batch_fixed_grid_matches_standalone():
    nets = [G.monte_carlo(G.gate_dag(2000, seed=0), 1 + b, 0.03) for b in range(256)]
    ragged = [G.gate_dag(g, seed=1000 + j) for j, g in enumerate((500, 1000, 2000, 4000))]
    batch = G.pack(nets + ragged)
    _say("batch 260 instances: gpu packed, 6 fixed steps, fp64 and mixed")
    eng, rb = run_gpu(batch, 6e-12, 1e-12, fixed=True, tran_fs=1.0, use_graph=True)
    _, rb32 = run_gpu(batch, 6e-12, 1e-12, fixed=True, tran_fs=1.0, use_graph=True,
                      precision="f32")
    info = eng.device.lu.info()
    assert info["n_block_patterns"] < 20 and info["n_blocks"] >= 260
    names = eng.topology.node_names
    for b, net in ((0, nets[0]), (1, nets[1]), (255, nets[255]), (259, ragged[3])):
        _say(f"batch: standalone instance {b} and its oracle")
        e1, r1 = run_gpu(net, 6e-12, 1e-12, fixed=True, tran_fs=1.0)
        assert np.array_equal(r1.times, rb.times)
        cols_b, cols_1 = _cols(names, e1.topology.node_names, net, b)
        err = float(np.max(np.abs(rb.stats["V_out"][:, cols_b] - r1.stats["V_out"][:, cols_1])))
        assert err <= 1e-9, (b, err)
        cir = omna.setup(net)
        _, opts = oracle_configs(net)
        ref = CTran(cir).run_fixed(6e-12, 1e-12, "gear2", opts, tran_fs=1.0, threads=16)
        oc = [cir.node_names[nm] for nm in net.nodes if nm != net.ground]
        e64 = float(np.max(np.abs(rb.stats["V_out"][:, cols_b] - ref.V[:, oc])))
        e32 = float(np.max(np.abs(rb32.stats["V_out"][:, cols_b] - ref.V[:, oc])))
        _say(f"batch instance {b}: fp64 {e64:.3e} V, mixed {e32:.3e} V against the oracle")
        assert e64 <= 1e-9, (b, e64)
        assert e32 <= 1e-6, (b, e32)


def test_config_batch_adaptive_common_grid():
    """The adaptive batch advances every instance on one grid (the smallest
    LTE step over the batch); each instance then equals its standalone run on
    that grid, in FP64 and mixed precision."""
    batch = G.batch_mc(256, 2000, ragged=(500, 1000, 2000, 4000), seed=0)
    _say("adaptive batch: 20 steps fp64 and mixed")
    eng, rb = run_gpu(batch, 1e-9, 1e-12, max_steps=21, use_graph=True)
    _, rb32 = run_gpu(batch, 1e-9, 1e-12, max_steps=21, use_graph=True, precision="f32")
    assert len(rb.times) == 21
    n32 = min(len(rb32.times), len(rb.times))
    names = eng.topology.node_names
    mc0 = G.monte_carlo(G.gate_dag(2000, seed=0), 1, 0.03)
    rg = G.gate_dag(4000, seed=1003)
    for b, net in ((0, mc0), (259, rg)):
        _say(f"adaptive batch: instance {b} standalone on the batch grid")
        e1, r1 = run_gpu(net, 1e-9, 1e-12, use_graph=True, replay_times=rb.times)
        assert r1.stats["rejected_steps"] == 0
        cols_b, cols_1 = _cols(names, e1.topology.node_names, net, b)
        err = float(np.max(np.abs(rb.stats["V_out"][:, cols_b] - r1.stats["V_out"][:, cols_1])))
        assert err <= 1e-9, (b, err)
        e32 = float(np.max(np.abs(rb32.stats["V_out"][:n32, cols_b]
                                  - r1.stats["V_out"][:n32, cols_1])))
        assert e32 <= 1e-6, (b, e32)
