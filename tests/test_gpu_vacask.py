"""Netlists through CircuitEngine(sim_path) on the GPU: an RC ladder written
as a VACASK netlist gives the generator's waveforms bit for bit, and a CMOS
inverter chain built from subcircuits over PSP103 model cards (run with the
EKV substitute) matches the oracle on the same mapped circuit within 1e-9 V.
The reference's benchmark circuits: the SRAM testbench's flattened netlist
(pwl clocks and inputs; a fixture, the reference is not on this machine)
against the oracle, and the array multiplier of the mul64 benchmark as 4x4
and 16x16 testbenches against the oracle (and the 4x4 against its
topology's Boolean evaluation).
"""

import sys

import numpy as np
import pytest

from vajax_amd import generators as G
from vajax_amd import vacask as V
from vajax_amd.engine import AdaptiveConfig, CircuitEngine

from .helpers import run_gpu, run_oracle, grid_deviation
from .test_vacask_cpu import _rc_ladder_sim

pytestmark = pytest.mark.gpu


def _progress(msg):
    print(msg, file=sys.stderr, flush=True)

CHAIN_SIM = """Inverter chain
ground 0
global vdd
model v vsource
model c capacitor
model psp103n psp103va (
    type=1
    tr=27.0
)
model psp103p psp103va (
    type=-1
    tr=27.0
)
parameters vsup=1.0

subckt inv(out in)
  parameters wn=0.5u wp=1u
  mp (out in vdd vdd) psp103p w=wp l=0.1u
  mn (out in 0 0) psp103n w=wn l=0.1u
  cl (out 0) c c=2f
ends

vdd (vdd 0) v dc=vsup
vin (a0 0) v type="pulse" val0=0 val1=vsup delay=20p rise=10p fall=10p width=200p period=1
x1 (a1 a0) inv
x2 (a2 a1) inv wn=1u wp=2u
x3 (a3 a2) inv
x4 (a4 a3) inv wp=3u

control
  options tran_method="gear2"
  analysis t tran step=1p stop=150p icmode="op"
endc
"""


def _run(eng, t_stop, dt):
    eng.parse()
    eng.options.fixed_step_transient = True
    cfg = AdaptiveConfig(integration_method="gear2", tran_fs=1.0)
    eng.prepare(t_stop=t_stop, dt=dt, adaptive_config=cfg)
    return eng.run_transient()


def test_rc_ladder_sim_file_matches_the_generator(tmp_path):
    p = tmp_path / "rc.sim"
    p.write_text(_rc_ladder_sim(60, 4))
    a = _run(CircuitEngine(p), 50e-12, 1e-12)
    b = _run(CircuitEngine(G.rc_ladder(60, seed=4)), 50e-12, 1e-12)
    assert np.array_equal(a.times, b.times)
    assert np.array_equal(a.stats["V_out"], b.stats["V_out"])


def test_inverter_chain_dc_needs_the_homotopy_chain(tmp_path):
    """Plain Newton from mid-rail diverges on this chain; the homotopy chain
    (gdev gives up, gshunt succeeds) finds the logic levels, on the GPU as
    with the oracle's Newton solve."""
    from vajax_amd import dc as D
    from vajax_amd.engine import SimulationOptions

    from .test_homotopy_cpu import _clamp_start, oracle_nr_solve

    p = tmp_path / "chain.sim"
    p.write_text(CHAIN_SIM)
    net = V.load(p)
    cir, X0, vs, isv = _clamp_start(net)
    Xo, ro = D.compute_dc_operating_point(oracle_nr_solve(cir, 100), X0, vs, isv,
                                          np.zeros(cir.n_unk), SimulationOptions())
    eng = CircuitEngine(p)
    Vg = eng.compute_dc_operating_point()
    rg = eng.dc_result
    assert ro.converged and ro.method == "chain_gshunt_stepping"
    assert (rg.method, rg.converged, rg.iterations, rg.homotopy_steps) == \
        (ro.method, ro.converged, ro.iterations, ro.homotopy_steps)
    assert float(np.max(np.abs(Vg - Xo[:cir.n_total]))) <= 1e-9
    lv = {n: Vg[cir.node_names[n]] for n in ("a1", "a2", "a3", "a4")}
    assert lv["a1"] > 0.99 and lv["a2"] < 0.01 and lv["a3"] > 0.99 and lv["a4"] < 0.01


def test_inverter_chain_sim_matches_oracle(tmp_path):
    # supply ramped from zero initial conditions, as the gate DAG does
    text = CHAIN_SIM.replace("vdd (vdd 0) v dc=vsup",
                             'vdd (vdd 0) v type="pulse" val0=0 val1=vsup rise=20p width=1 '
                             "period=2").replace('icmode="op"', 'icmode="uic"')
    p = tmp_path / "chain.sim"
    p.write_text(text)
    net = V.load(p)
    assert net.options["icmode"] == "uic"
    assert net.groups["mosfet"].count == 8
    assert sorted(net.groups["mosfet"].params["type"]) == [-1.0] * 4 + [1.0] * 4
    _, ref = run_oracle(net, 150e-12, 1e-12, fixed=True, tran_fs=1.0)
    res = _run(CircuitEngine(p), 150e-12, 1e-12)
    assert np.array_equal(res.times, ref.times)
    err = float(np.max(np.abs(res.stats["V_out"] - ref.V)))
    assert err <= 1e-9, err


def _sim_file(tmp_path, text, name):
    p = tmp_path / name
    p.write_text(text)
    return str(p)


def test_tb_dp_sram_testbench_matches_oracle():
    import os

    from .golden import netlist_npz

    net = netlist_npz.load(os.path.join(os.path.dirname(__file__), "golden",
                                        "tb_dp512x8_netlist.npz"))
    _, ref = run_oracle(net, 20e-9, 1e-10)
    _, res = run_gpu(net, 20e-9, 1e-10, use_graph=True)
    assert len(res.times) == len(ref.times)
    dev = grid_deviation(res.times, ref.times)
    assert dev <= 1e-12, dev
    err = float(np.max(np.abs(res.stats["V_out"] - ref.V)))
    assert err <= 1e-9, err


def test_array_multiplier_4x4_testbench(tmp_path):
    from .test_vacask_cpu import _mul_logic

    a, b = [1, 1, 0, 1], [0, 1, 1, 0]
    path = _sim_file(tmp_path, G.array_multiplier_testbench(4, a, b, t_stop=6e-9), "mul4.sim")
    eng = CircuitEngine(path)
    eng.parse()
    eng.prepare(t_stop=6e-9, dt=2e-12, use_graph=True)
    res = eng.run_transient()
    cir, ref = run_oracle(V.load(path), 6e-9, 2e-12, lte_ratio=1.5)
    n = min(len(ref.times), len(res.times))
    same = np.isclose(res.times[:n], ref.times[:n], rtol=1e-9, atol=0)
    k = n if same.all() else int(np.argmin(same))
    assert k >= 20, k
    assert float(np.max(np.abs(res.stats["V_out"][:k] - ref.V[:k]))) <= 1e-6
    _, rr = run_gpu(V.load(path), 6e-9, 2e-12, use_graph=True, replay_times=ref.times)
    assert rr.stats["rejected_steps"] == 0
    d = np.abs(rr.stats["V_out"] - ref.V)
    assert float(np.max(d - 1e-6 * np.abs(ref.V))) <= 1e-9, float(np.max(d))
    got = [int(res.voltage(f"top.p{i}")[-1] > 0.6) for i in range(7)]
    assert got == _mul_logic(4, a, b)


def test_array_multiplier_16x16_window(tmp_path):
    """The 16x16 multiplier of the mul64 arrangement (7 352 unknowns, a dense
    trailing block in its LU) through CircuitEngine(path): 12 adaptive steps
    against the oracle, then the oracle's grid replayed in FP64 and mixed
    precision."""
    # The Newton iterate of the third solve diverges (Jacobian entries of
    # 1e37 and beyond); the pivot monitor's bounded re-analysis (a fill limit
    # of four times the factor, one attempt per solve) turns that into a
    # Newton failure and a cut dt, as in the oracle, in well under a second.
    path = _sim_file(tmp_path, G.array_multiplier_testbench(16), "mul16.sim")
    eng = CircuitEngine(path)
    eng.parse()
    net = eng.netlist
    assert net.groups["mosfet"].count == G.array_multiplier_mosfets(16)
    _progress("mul16: adaptive run on the device")
    eng.prepare(t_stop=2e-9, dt=2e-12, max_steps=13, use_graph=True)
    res = eng.run_transient()
    _progress(f"mul16: device run done, {len(res.times)} rows, "
              f"{res.stats['repivots']} re-pivots; oracle run")
    _, ref = run_oracle(net, 2e-9, 2e-12, max_steps=13, lte_ratio=1.5)
    assert len(ref.times) == 13
    n = min(len(ref.times), len(res.times))
    same = np.isclose(res.times[:n], ref.times[:n], rtol=1e-9, atol=0)
    k = n if same.all() else int(np.argmin(same))
    assert k >= 6, k
    for prec, bar in (("f64", 1e-9), ("f32", 1e-6)):
        _progress(f"mul16: replay of the oracle grid in {prec}")
        _, rr = run_gpu(net, 2e-9, 2e-12, use_graph=True, replay_times=ref.times, precision=prec)
        assert rr.stats["rejected_steps"] == 0
        d = np.abs(rr.stats["V_out"] - ref.V)
        worst = float(np.max(d - 1e-6 * np.abs(ref.V)))
        assert worst <= bar, (prec, float(np.max(d)))


def test_array_multiplier_64x64_whole_run_properties():
    """The reference's mul64 benchmark circuit at full size (the 64x64
    multiplier testbench, 266 408 MOSFETs; too large for the oracle): its
    whole 2 ns transient completes with the dense trailing block on the
    matrix cores, every node voltage stays finite and within half a volt of
    the supply rails, and mixed precision on the FP64 run's grid stays within
    1e-6 V."""
    net = V.load("", text=G.array_multiplier_testbench(64))
    assert net.groups["mosfet"].count == G.array_multiplier_mosfets(64) == 266408
    _progress("mul64: FP64 adaptive run over 2 ns")
    eng, res = run_gpu(net, 2e-9, 2e-12, use_graph=True, autotune=True)
    info = eng.device.lu.info()
    assert info["dense_k"] >= 96 and info["dense_m"] > 0, info
    assert abs(res.times[-1] - 2e-9) <= 1e-15
    V64 = res.stats["V_out"]
    assert np.all(np.isfinite(V64))
    # within half a volt of the rails (edges couple through the terminal
    # charges: -0.25 V undershoot measured); a diverging run would leave them
    _progress(f"mul64: node voltages in [{float(V64.min()):.3f}, {float(V64.max()):.3f}] V")
    assert float(V64.min()) >= -0.5 and float(V64.max()) <= 1.2 + 0.5
    assert res.stats["rejected_steps"] <= len(res.times) // 4
    _progress(f"mul64: {len(res.times)} rows; mixed precision on the same grid")
    _, r32 = run_gpu(net, 2e-9, 2e-12, use_graph=True, replay_times=res.times, precision="f32")
    assert r32.stats["rejected_steps"] == 0
    err = float(np.max(np.abs(r32.stats["V_out"] - V64)))
    _progress(f"mul64: max |dV| mixed vs FP64 = {err:.3e} V")
    assert err <= 1e-6, err
