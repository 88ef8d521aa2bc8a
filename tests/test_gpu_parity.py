"""GPU parity: the HIP path against the oracle on the same seeded inputs.

Bars: bit-exact for the CSR pattern and for resistor/capacitor stamping (same
FP64 operation order, no FMA contraction); FP64 LU bit-exact against the CPU
execution of the same numeric program; MOSFET eval within 1e-13 relative
(device exp/log1p vs numpy); node-voltage waveforms within 1e-9 V absolute
for FP64 and 1e-6 V for FP32-LU + FP64 refinement.
"""

import numpy as np
import pytest

from oracle import mna as omna
from oracle import models as omod
from vajax_amd import generators as G

from .helpers import run_gpu, run_oracle, waveform_error, grid_deviation

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def torch():
    import torch as _t

    assert _t.cuda.is_available()
    return _t


def _dev(torch, a):
    return torch.as_tensor(np.ascontiguousarray(a), device="cuda")


def _engine(net):
    from vajax_amd.engine import CircuitEngine

    e = CircuitEngine(net)
    e.parse()
    return e


@pytest.mark.parametrize("model", ["resistor", "capacitor", "mosfet"])
def test_model_eval_contract(torch, model):
    import ctypes as C

    from vajax_amd import _lib as L
    from vajax_amd.models import param_block

    rng = np.random.default_rng(11)
    n = 4099
    if model == "mosfet":
        typ = np.where(rng.random(n) < 0.5, 1.0, -1.0)
        p = {"type": typ, "w": rng.uniform(0.2e-6, 2e-6, n), "l": rng.uniform(45e-9, 90e-9, n),
             "vto": rng.uniform(0.2, 0.5, n), "gamma": rng.uniform(0.2, 0.7, n),
             "phi": rng.uniform(0.5, 0.9, n), "kp": rng.uniform(5e-5, 3e-4, n),
             "theta": rng.uniform(0.0, 0.3, n), "cox": np.full(n, 0.02),
             "cov": np.full(n, 2e-10), "cj": np.full(n, 5e-10)}
        vb = rng.uniform(-1.5, 1.5, (3, n))
        vb[:, :8] = [[0.0] * 8, [0.0] * 8, [0.0] * 8]
        ref = omod.eval_mosfet(vb[0], vb[1], vb[2], omod.mosfet_cache(p))
        nv, nr, nj = 3, 4, 16
    else:
        key = "r" if model == "resistor" else "c"
        p = {key: rng.uniform(1e2, 1e5, n) if model == "resistor" else rng.uniform(1e-15, 1e-12, n)}
        vb = rng.uniform(-2, 2, (1, n))
        fn = omod.eval_resistor if model == "resistor" else omod.eval_capacitor
        ref = fn(vb[0], p[key])
        nv, nr, nj = 1, 2, 4
    par = _dev(torch, param_block(model, p))
    vbd = _dev(torch, vb)
    outs = [torch.empty((n, nr), dtype=torch.float64, device="cuda") for _ in range(2)]
    outs += [torch.empty((n, nj), dtype=torch.float64, device="cuda") for _ in range(2)]
    rc = L.lib.vjx_model_eval(L.MODEL_IDS[model], n, C.c_void_p(vbd.data_ptr()),
                              C.c_void_p(par.data_ptr()), None,
                              *[C.c_void_p(o.data_ptr()) for o in outs], None, None, None)
    assert rc == 0, L.last_error()
    for got, want in zip(outs, ref):
        got = got.cpu().numpy()
        if model == "mosfet":
            # entries such as gb = -(gg+gs+gd) cancel: compare against the
            # largest magnitude of the same instance's output row
            scale = np.max(np.abs(want), axis=1, keepdims=True)
            assert np.all(np.abs(got - want) <= 1e-13 * scale + 1e-300), model
        else:
            assert np.array_equal(got, want), model


def _random_state(cir, rng, spread=1.0):
    X = np.zeros(cir.n_total + cir.m)
    X[1:cir.n_total] = rng.uniform(-spread, spread, cir.n_unk)
    X[cir.n_total:] = rng.uniform(-1e-3, 1e-3, cir.m)
    return X


@pytest.mark.parametrize("which", ["rc", "ring", "dag"])
def test_build_system_matches_oracle(torch, which):
    net = {"rc": lambda: G.rc_ladder(200, seed=2),
           "ring": lambda: G.ring_oscillator(11, seed=2),
           "dag": lambda: G.gate_dag(300, seed=2)}[which]()
    cir = omna.setup(net)
    eng = _engine(net)
    topo = eng.topology
    assert np.array_equal(topo.indptr, cir.indptr) and np.array_equal(topo.indices, cir.indices)
    rng = np.random.default_rng(5)
    X = _random_state(cir, rng)
    vs = rng.uniform(0, 1, cir.m)
    isv = rng.uniform(-1e-4, 1e-4, net.isources.count)
    Qp, dQ, Q2 = (rng.uniform(-1e-13, 1e-13, cir.n_unk) for _ in range(3))
    c0, c1, d1, c2, gsh = 1.5e12, -2e12, 0.0, 0.5e12, 1e-9
    want = omna.build_system(cir, X, vs, isv, Qp, c0, gsh, c1, d1, dQ, c2, Q2)
    out = eng.device.build_system(_dev(torch, X), _dev(torch, vs),
                                  _dev(torch, isv) if len(isv) else None,
                                  _dev(torch, Qp), _dev(torch, dQ), _dev(torch, Q2),
                                  c0=c0, c1=c1, d1=d1, c2=c2, gmin=1e-12, gshunt=gsh)
    names = ["csr_data", "f", "Q", "I_vsource", "max_res_contrib"]
    for name, w in zip(names, want):
        g = out[name].cpu().numpy()[: len(w)]
        if which == "rc":
            assert np.array_equal(g, w), name
        else:
            scale = np.max(np.abs(w)) if len(w) else 1.0
            assert np.max(np.abs(g - w)) <= 1e-12 * scale, name


def test_lu_gpu_matches_cpu_program(torch):
    import scipy.sparse as sp

    net = G.gate_dag(2000, seed=4)
    cir = omna.setup(net)
    rng = np.random.default_rng(1)
    X = _random_state(cir, rng, 0.5)
    csr, *_ = omna.build_system(cir, X, np.zeros(cir.m), np.zeros(0), np.zeros(cir.n_unk), 1e13)
    eng = _engine(net)
    lu = eng.device.lu
    lu.analyze(csr)
    d = _dev(torch, csr)
    assert lu.refactor(d, 0) == 0
    got = lu.download()
    want = lu.refactor_cpu(csr, 0)
    assert np.array_equal(got, want)
    A = sp.csr_matrix((csr, cir.indices, cir.indptr), shape=(cir.n_aug, cir.n_aug))
    b = rng.standard_normal(cir.n_aug)
    bd = _dev(torch, b)
    x = torch.empty(cir.n_aug, dtype=torch.float64, device="cuda")
    lu.solve(d, bd, x, 0)
    xg = x.cpu().numpy()
    assert np.array_equal(xg, lu.solve_cpu(csr, b, 0))
    assert np.linalg.norm(A @ xg - b) / np.linalg.norm(b) < 1e-12
    # FP32 factorisation + one FP64 refinement step
    assert lu.refactor(d, 1) == 0
    lu.solve(d, bd, x, 1)
    xg = x.cpu().numpy()
    assert np.linalg.norm(A @ xg - b) / np.linalg.norm(b) < 1e-9


@pytest.mark.parametrize("form", ["lds", "hbm"])
def test_solve_band_forms_match_cpu_program(torch, monkeypatch, form):
    """The two forms of a solve band task -- everything it reads in LDS
    (default) and HBM only -- run the same sums in the same order (long rows
    as one-wavefront sums): each is bit-identical to the CPU execution of the
    program, on a 20k-gate DAG whose bands range from 2048 tasks to one."""
    from vajax_amd.sparse import SparseLU

    monkeypatch.setenv("VJX_SOLVE_LDS", "0" if form == "hbm" else "1")
    net = G.gate_dag(20000, seed=2)
    cir = omna.setup(net)
    rng = np.random.default_rng(3)
    X = _random_state(cir, rng, 0.5)
    csr, *_ = omna.build_system(cir, X, np.zeros(cir.m), np.zeros(0), np.zeros(cir.n_unk), 1e13)
    lu = SparseLU(cir.indptr, cir.indices)
    lu.analyze(csr)
    d = _dev(torch, csr)
    assert lu.refactor(d, 0) == 0
    b = rng.standard_normal(cir.n_aug)
    x = torch.empty(cir.n_aug, dtype=torch.float64, device="cuda")
    for _ in range(2):  # the second solve reuses the uploaded programs
        lu.solve(d, _dev(torch, b), x, 0)
        assert np.array_equal(x.cpu().numpy(), lu.solve_cpu(csr, b, 0))
    # FP32 factors and solves (the same kernels at 4-byte values) + one FP64
    # refinement step
    import scipy.sparse as sp

    assert lu.refactor(d, 1) == 0
    lu.solve(d, _dev(torch, b), x, 1)
    A = sp.csr_matrix((csr, cir.indices, cir.indptr), shape=(cir.n_aug, cir.n_aug))
    xg = x.cpu().numpy()
    assert np.linalg.norm(A @ xg - b) / np.linalg.norm(b) < 1e-9


def test_solve_host_matches_scipy():
    import scipy.sparse as sp
    import scipy.sparse.linalg as spla

    from vajax_amd.sparse import solve

    net = G.ring_oscillator(21, seed=1)
    cir = omna.setup(net)
    rng = np.random.default_rng(3)
    csr, *_ = omna.build_system(cir, _random_state(cir, rng), np.ones(cir.m), np.zeros(1),
                                np.zeros(cir.n_unk), 1e12)
    b = rng.standard_normal(cir.n_aug)
    x = solve(cir.indptr, cir.indices, csr, b)
    A = sp.csr_matrix((csr, cir.indices, cir.indptr), shape=(cir.n_aug, cir.n_aug)).tocsc()
    xs = spla.spsolve(A, b)
    assert np.allclose(x, xs, rtol=1e-10, atol=1e-14)


def test_solve_transpose_and_dot_match_scipy():
    """solve_transpose / dot of the linear-solver FFI (umfpack_ffi.cpp:345-481)."""
    import scipy.sparse as sp
    import scipy.sparse.linalg as spla
    import torch

    from vajax_amd import _lib as L
    from vajax_amd.sparse import SparseLU, dot, solve_transpose

    net = G.gate_dag(300, seed=2)
    cir = omna.setup(net)
    rng = np.random.default_rng(5)
    for it in range(2):  # the second call reuses the cached A^T handle
        X = np.zeros(cir.n_total + cir.m)
        X[1:cir.n_total] = rng.uniform(0.0, 1.0, cir.n_unk)  # an operating-point-like state
        csr, *_ = omna.build_system(cir, X, np.ones(cir.m), np.zeros(0), np.zeros(cir.n_unk), 1e12)
        A = sp.csr_matrix((csr, cir.indices, cir.indptr), shape=(cir.n_aug, cir.n_aug))
        assert (A - A.T).count_nonzero() > 0  # the test means something
        b = rng.standard_normal(cir.n_aug)
        xt = solve_transpose(cir.indptr, cir.indices, csr, b)
        want = spla.spsolve(A.T.tocsc(), b)
        assert np.allclose(xt, want, rtol=1e-9, atol=1e-14), it
        assert np.allclose(dot(cir.indptr, cir.indices, csr, b), A @ b, rtol=1e-14, atol=0)
    lu = SparseLU(cir.indptr, cir.indices)
    d = torch.as_tensor(csr, device="cuda")
    bd = torch.as_tensor(b, device="cuda")
    for prec, refine, tol in ((L.VJX_PREC_F64, 0, 1e-9), (L.VJX_PREC_F32, 2, 1e-8)):
        x = torch.empty_like(bd)
        lu.solve_transpose(d, bd, x, prec, refine)
        res = np.linalg.norm(A.T @ x.cpu().numpy() - b) / np.linalg.norm(b)
        assert res <= tol, (prec, res)


@pytest.mark.parametrize("fixed", [True, False])
def test_chunked_output_matches_resident_output(fixed):
    """Output rows copied to the host in chunks (checkpoint mode) give the
    same waveforms bit for bit as all rows kept on the device."""
    net = G.diode_ladder(30, seed=6)
    t_stop = 80e-12 if fixed else 1.2e-9
    _, a = run_gpu(net, t_stop, 1e-12, fixed=fixed, tran_fs=1.0 if fixed else None,
                   checkpoint_interval=0)
    _, b = run_gpu(net, t_stop, 1e-12, fixed=fixed, tran_fs=1.0 if fixed else None,
                   checkpoint_interval=7)
    assert len(a.times) > 3 * 7
    assert np.array_equal(a.times, b.times)
    assert np.array_equal(a.stats["V_out"], b.stats["V_out"])
    assert np.array_equal(a.stats["I_out"], b.stats["I_out"])
    # and the chunked run against the oracle (same accepted grid; the
    # limited junctions keep the bar at 1e-9 V)
    _, ref = run_oracle(net, t_stop, 1e-12, fixed=fixed, tran_fs=1.0 if fixed else None)
    n = min(len(ref.times), len(b.times))
    same = np.isclose(b.times[:n], ref.times[:n], rtol=1e-9, atol=0)
    k = n if same.all() else int(np.argmin(same))
    assert k >= min(n, 3 * 7), (k, n)
    assert float(np.max(np.abs(b.stats["V_out"][:k] - ref.V[:k]))) <= 1e-9
    assert float(np.max(np.abs(b.stats["I_out"][:k, :ref.I.shape[1]] - ref.I[:k]))) <= 1e-9


def test_transient_rc_fixed_step():
    net = G.rc_ladder(200, seed=0)
    _, ref = run_oracle(net, 1e-9, 1e-12, fixed=True, tran_fs=1.0)
    _, res = run_gpu(net, 1e-9, 1e-12, fixed=True, tran_fs=1.0)
    assert res.stats["n_steps"] == len(ref.times)
    assert np.array_equal(res.times, ref.times)
    V = res.stats["V_out"]
    assert np.max(np.abs(V - ref.V)) <= 1e-12
    assert res.stats["total_nr_iterations"] == ref.stats["total_nr_iterations"]


def test_transient_rc_adaptive():
    net = G.rc_ladder(200, seed=0)
    _, ref = run_oracle(net, 10e-9, 1e-12)
    _, res = run_gpu(net, 10e-9, 1e-12)
    assert res.stats["n_steps"] == len(ref.times)
    assert res.stats["rejected_steps"] == ref.stats["rejected_steps"]
    dev = grid_deviation(res.times, ref.times)
    assert dev <= 1e-12, dev
    assert np.max(np.abs(res.stats["V_out"] - ref.V)) <= 1e-9


def test_transient_ring_uic():
    net = G.ring_oscillator(11, seed=0)
    _, ref = run_oracle(net, 0.3e-9, 1e-12)
    _, res = run_gpu(net, 0.3e-9, 1e-12)
    err = waveform_error(ref.times, ref.V, res.times, res.stats["V_out"])
    assert err <= 1e-6, err
    assert abs(res.stats["n_steps"] - len(ref.times)) <= 2


def test_transient_gate_dag_fixed_step_fp64_and_mixed():
    """Fixed 1 ps grid through the supply ramp: same time points by construction."""
    net = G.gate_dag(200, seed=7)
    _, ref = run_oracle(net, 0.15e-9, 1e-12, fixed=True, tran_fs=1.0)
    _, r64 = run_gpu(net, 0.15e-9, 1e-12, fixed=True, tran_fs=1.0)
    assert np.array_equal(r64.times, ref.times)
    e64 = float(np.max(np.abs(r64.stats["V_out"] - ref.V)))
    assert e64 <= 1e-9, e64
    _, r32 = run_gpu(net, 0.15e-9, 1e-12, fixed=True, tran_fs=1.0, precision="f32")
    e32 = float(np.max(np.abs(r32.stats["V_out"] - ref.V)))
    assert e32 <= 1e-6, e32


@pytest.mark.parametrize("precision", ["f64", "f32"])
def test_hip_graph_replays_match_eager(precision):
    """The captured Newton-iteration graph is replayed once per iteration with
    host work (scalar upload, status read-back, step control) in between; every
    replay must give exactly the eager result."""
    net = G.gate_dag(300, seed=9)
    _, a = run_gpu(net, 0.05e-9, 1e-12, precision=precision, use_graph=True)
    _, b = run_gpu(net, 0.05e-9, 1e-12, precision=precision, use_graph=False)
    assert a.stats["total_nr_iterations"] > 20
    assert np.array_equal(a.times, b.times)
    assert np.array_equal(a.stats["V_out"], b.stats["V_out"])
    assert np.array_equal(a.stats["I_out"], b.stats["I_out"])


def test_transient_gate_dag_adaptive():
    """Adaptive BDF2 with LTE rejections.

    Accept/reject and the Newton stopping test are discontinuous in the
    solution, so two correct implementations whose LU round differently
    follow the same step sequence only until the first decision that lies
    within rounding of its threshold.  Bars: the common prefix of accepted
    time points covers at least 30 % of the run with waveforms within
    1e-7 V on it (ten times below the Newton voltage tolerance vntol = 1e-6,
    where each run's Newton loop is free to stop), and the whole waveforms
    agree to the LTE tolerance scale (reltol * lte_ratio * 1 V = 3.5e-3 V)
    after interpolation."""
    net = G.gate_dag(200, seed=7)
    _, ref = run_oracle(net, 0.2e-9, 1e-12)
    _, res = run_gpu(net, 0.2e-9, 1e-12)
    n = min(len(ref.times), len(res.times))
    same = np.isclose(res.times[:n], ref.times[:n], rtol=1e-9, atol=0)
    k = n if same.all() else int(np.argmin(same))
    assert k >= 0.3 * len(ref.times), (k, len(ref.times), len(res.times))
    err = float(np.max(np.abs(res.stats["V_out"][:k] - ref.V[:k])))
    assert err <= 1e-7, err
    e_all = waveform_error(ref.times, ref.V, res.times, res.stats["V_out"])
    assert e_all <= 3.5e-3, e_all
    assert res.times[-1] == pytest.approx(0.2e-9, rel=1e-12)


def test_packed_batch_fixed_step_fp64_and_mixed():
    """Monte Carlo copies plus ragged topologies packed into one block-diagonal
    system (one analysis per distinct block pattern, one launch per level for
    all instances): same time grid, same bars as a single circuit, and each
    instance's waveform equals the packed oracle's."""
    nets = [G.monte_carlo(G.gate_dag(120, seed=3), 40 + b, 0.03) for b in range(3)]
    nets += [G.gate_dag(50, seed=11), G.gate_dag(80, seed=12)]
    net = G.pack(nets)
    _, ref = run_oracle(net, 0.12e-9, 1e-12, fixed=True, tran_fs=1.0)
    eng, r64 = run_gpu(net, 0.12e-9, 1e-12, fixed=True, tran_fs=1.0)
    info = eng.device.lu.info()
    assert info["n_block_patterns"] < info["n_blocks"]
    assert np.array_equal(r64.times, ref.times)
    e64 = float(np.max(np.abs(r64.stats["V_out"] - ref.V)))
    assert e64 <= 1e-9, e64
    _, r32 = run_gpu(net, 0.12e-9, 1e-12, fixed=True, tran_fs=1.0, precision="f32")
    e32 = float(np.max(np.abs(r32.stats["V_out"] - ref.V)))
    assert e32 <= 1e-6, e32
    # instances differ (device variation), so the comparison is not vacuous
    y0 = r64.voltage("i000.g000119y")
    y1 = r64.voltage("i001.g000119y")
    assert np.max(np.abs(y0 - y1)) > 1e-6


@pytest.mark.parametrize("precision", ["f64", "f32"])
def test_band_programs_match_level_programs(precision):
    """Refactorisation and solves as band programs (one launch per band of
    levels, one workgroup per group of connected pivots/rows, plus one wide
    Schur-update launch per band) against the one-launch-per-level programs,
    both against the oracle on the same fixed time grid."""
    net = G.gate_dag(300, seed=13)
    _, ref = run_oracle(net, 0.08e-9, 1e-12, fixed=True, tran_fs=1.0)
    eng_a, a = run_gpu(net, 0.08e-9, 1e-12, fixed=True, tran_fs=1.0, precision=precision,
                       band_width=16)
    eng_b, b = run_gpu(net, 0.08e-9, 1e-12, fixed=True, tran_fs=1.0, precision=precision,
                       band_width=0)
    ia, ib = eng_a.device.lu.info(), eng_b.device.lu.info()
    assert ia["fwd_bands"] > 0 and ib["fwd_bands"] == 0
    assert ia["fact_launches"] < ib["fact_launches"]
    tol = 1e-9 if precision == "f64" else 1e-6
    for r in (a, b):
        assert np.array_equal(r.times, ref.times)
        err = float(np.max(np.abs(r.stats["V_out"] - ref.V)))
        assert err <= tol, err


def test_autotuned_launch_plan_gives_the_same_waveforms():
    """The autotuner changes launch plans only.  The solves' band width and
    the LU task pipeline keep every sum and its order (bit-identical
    results); a different refactorisation band width reorders the Schur sums
    (rounding-level differences), so the waveforms agree to 1e-12 V and are
    bit-identical when the tuned band width is the default one."""
    net = G.gate_dag(400, seed=17)
    eng_a, a = run_gpu(net, 0.05e-9, 1e-12, autotune=True, use_graph=True)
    eng_b, b = run_gpu(net, 0.05e-9, 1e-12, autotune=False, use_graph=True)
    ia = eng_a.device.lu.info()
    assert ia["tuned_solve_ms"] > 0 and ia["solve_band_width"] in (8, 12, 16, 24, 32, 48, 64)
    assert ia["tuned_factor_ms"] > 0 and ia["band_width"] in (12, 16, 24)
    assert ia["lu_pipeline"] in (0, 1, 2)
    assert np.array_equal(a.times, b.times)
    if ia["band_width"] == 16:
        assert np.array_equal(a.stats["V_out"], b.stats["V_out"])
    else:
        assert float(np.max(np.abs(a.stats["V_out"] - b.stats["V_out"]))) <= 1e-12


def test_lu_pipelines_are_bit_identical():
    """The three LU band-task pipelines run the same sums in the same order."""
    import torch

    from vajax_amd.engine import DeviceCircuit
    from vajax_amd.topology import build_topology

    net = G.gate_dag(3000, seed=5)
    topo = build_topology(net)
    dc = DeviceCircuit(topo, net)
    n_total, m = topo.n_total, topo.n_vsources
    rng = np.random.default_rng(3)
    X = torch.tensor(np.concatenate([[0.0], rng.uniform(0, 1, n_total - 1), np.zeros(m)]),
                     device="cuda")
    out = dc.build_system(X, torch.ones(m, dtype=torch.float64, device="cuda"),
                          torch.zeros(0, dtype=torch.float64, device="cuda"), c0=1e12)
    lu = dc.lu
    facs = []
    for mode in ("0", "1", "2"):
        import os

        os.environ["VJX_LU_DEEP"] = mode
        try:
            lu.set_band_width(0)
            lu.set_band_width(16)  # re-analyse: the pipeline is read at upload
            lu.refactor(out["csr_data"])
            facs.append(lu.download())
        finally:
            del os.environ["VJX_LU_DEEP"]
    assert np.array_equal(facs[0], facs[1]) and np.array_equal(facs[0], facs[2])


def test_repivot_path_keeps_parity():
    """Force the pivot-growth monitor from the host (growth limit 1e-300):
    the first iteration of every Newton solve re-runs the symbolic analysis
    on its matrix, drops the captured graph and redoes the iteration from the
    saved iterate; the redo still exceeds the limit, so the monitor stands
    down for the rest of that solve (one re-pivot per solve: each of the 40
    steps; the DAG starts from zero initial conditions, without a DC solve).
    The waveform must not change."""
    net = G.gate_dag(150, seed=23)
    _, ref = run_oracle(net, 40e-12, 1e-12, fixed=True, tran_fs=1.0)
    _, res = run_gpu(net, 40e-12, 1e-12, fixed=True, tran_fs=1.0, use_graph=True,
                     repivot_growth=1e-300)
    assert res.stats["repivots"] == len(res.times) - 1 == 40
    assert np.array_equal(res.times, ref.times)
    err = float(np.max(np.abs(res.stats["V_out"] - ref.V)))
    assert err <= 1e-9, err


def test_newton_failure_path_matches_oracle():
    """Newton capped at 2 iterations: failed steps are retried at dt/4 and
    the step control must make the same decisions as the oracle.  This case
    also rejects a step that ends on t_stop, which leaves dt = 0 and a
    force-accepted repeat of the previous time point (full_mna.py:2076)."""
    net = G.gate_dag(150, seed=29)
    _, ref = run_oracle(net, 0.1e-9, 1e-12, tran_itl=2)
    _, res = run_gpu(net, 0.1e-9, 1e-12, tran_itl=2)
    assert ref.stats["rejected_steps"] > 0
    assert ref.stats["min_dt_used"] == 0.0
    assert res.stats["rejected_steps"] == ref.stats["rejected_steps"]
    assert res.stats["min_dt_used"] == 0.0
    n = min(len(ref.times), len(res.times))
    same = np.isclose(res.times[:n], ref.times[:n], rtol=1e-9, atol=0)
    k = n if same.all() else int(np.argmin(same))
    assert k >= 0.3 * len(ref.times), (k, len(ref.times))
    err = float(np.max(np.abs(res.stats["V_out"][:k] - ref.V[:k])))
    assert err <= 1e-6, err
