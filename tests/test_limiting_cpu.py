"""Device limiting in the oracle: pnjlim / fetlim, the diode's limited eval,
lim_rhs and the limit state in Newton (vajax/analysis/limiting.py:22-107,
openvaf_jax/codegen/function_builder.py:1126-1212,
vajax/analysis/solver_factories.py:317-323).

The limiting functions are pinned branch by branch with values computed by
hand from the formulas; the diode with limiting is pinned by a closed-form
DC point (a resistor-diode divider solved by bisection) and by the lim_rhs
correction, which must equal the linearisation of the diode current from the
limited voltage back to the raw one.
"""

import math

import numpy as np
import pytest
from scipy.optimize import brentq

from oracle import limiting as L
from oracle import mna
from oracle import models as M
from oracle import transient as T
from vajax_amd.netlist import NetlistBuilder

VT = M.VT


@pytest.mark.parametrize("vnew,vold,want", [
    # forward, large rise from a forward-biased old value: vold + vt*log(1 + d/vt)
    (0.9, 0.6, 0.6 + 0.026 * math.log(1 + 0.3 / 0.026)),
    # forward, large fall (vnew still above vcrit): vold - vt*log(1 - d/vt)
    (0.7, 1.2, 1.2 - 0.026 * math.log(1 + 0.5 / 0.026)),
    # from a non-positive old value, vnew > 0: vt*log(vnew/vt)
    (1.5, 0.0, 0.026 * math.log(1.5 / 0.026)),
    (1.5, -0.3, 0.026 * math.log(1.5 / 0.026)),
    # small step above vcrit: unchanged
    (0.62, 0.61, 0.62),
    # below vcrit, positive: unchanged
    (0.3, 0.0, 0.3),
    # negative with a positive old value: max(vnew, -vold - 1)
    (-3.0, 0.5, -1.5),
    (-0.2, 0.5, -0.2),
    # negative with a non-positive old value: max(vnew, 2*vold - 1)
    (-3.0, -0.4, -1.8),
    (-1.0, -0.4, -1.0),
])
def test_pnjlim_branches(vnew, vold, want):
    got = float(L.pnjlim(vnew, vold, 0.026, 0.6))
    assert got == pytest.approx(want, rel=1e-15, abs=1e-15)


def test_pnjlim_is_elementwise_and_limits_towards_vold():
    rng = np.random.default_rng(0)
    vnew = rng.uniform(-3, 3, 1000)
    vold = rng.uniform(-1, 1.2, 1000)
    got = L.pnjlim(vnew, vold, 0.03, 0.65)
    for a, b, g in zip(vnew, vold, got):
        assert g == float(L.pnjlim(a, b, 0.03, 0.65))
        # a limited value lies between vold and vnew, or is the
        # reverse-recovery value vcrit / vt*log(vnew/vt) (vold <= 0)
        if g != a and b > 0:
            assert min(a, b) - 1e-12 <= g <= max(a, b) + 1e-12


@pytest.mark.parametrize("vnew,vold,vto,want", [
    (3.0, 1.0, 0.5, 1.0 + min(2.0, 2 * 0.5 + 2)),   # on: step 2|vold-vto|+2 = 3 -> 3.0
    (6.0, 1.0, 0.5, 4.0),                          # on: clipped at +3
    (-4.0, 1.0, 0.5, -2.0),                        # on: clipped at -3
    (0.9, 0.2, 0.5, 0.7),                          # off: clipped at +0.5
    (0.0, 0.2, 0.5, 0.0),                          # off: within 0.5
    (-1.0, 0.2, 0.5, -0.3),                        # off: clipped at -0.5
])
def test_fetlim_branches(vnew, vold, vto, want):
    assert float(L.fetlim(vnew, vold, vto)) == pytest.approx(want, abs=1e-15)


def _divider(vs=2.0, r=1e3, is_=1e-14, n=1.0):
    b = NetlistBuilder("divider")
    b.source("vsource", "v1", "in", "0", type="dc", dc=vs)
    b.device("resistor", "r1", ["in", "d"], r=r)
    b.device("diode", "d1", ["d", "0"], **{"is": is_, "n": n, "cj": 0.0})
    return b.build(analysis={"stop": 1e-9, "step": 1e-12}, options={"icmode": "op"})


@pytest.mark.parametrize("vs,r,is_,n", [(2.0, 1e3, 1e-14, 1.0), (5.0, 220.0, 1e-12, 1.8),
                                        (0.3, 1e4, 1e-15, 1.0)])
def test_diode_divider_dc_matches_closed_form(vs, r, is_, n):
    net = _divider(vs, r, is_, n)
    cir = mna.setup(net)
    assert cir.n_lim == 1
    solver = T.LinearSolver(cir)
    X0, _, _, _, info = T.initial_state(cir, solver, T.Options())
    assert info.converged
    nvt = n * VT
    gmin = cir.gmin

    def kcl(v):  # current into node d: from the resistor, minus diode and gmin
        return (vs - v) / r - is_ * (math.exp(v / nvt) - 1.0) - gmin * v

    v_exact = brentq(kcl, -1.0, vs, xtol=1e-15, rtol=1e-15)
    vd = X0[cir.node_names["d"]]
    assert abs(vd - v_exact) < 1e-9, (vd, v_exact)
    # converged with the limit settled on the solution
    assert abs(info.lim[0] - vd) < 1e-6
    # limiting was exercised: from the mid-rail start the first step is limited
    if vs > 1.0:
        assert info.iterations >= 3


def test_lim_rhs_is_the_linearisation_back_to_the_raw_voltage():
    net = _divider(2.0, 1e3, 1e-14, 1.0)
    cir = mna.setup(net)
    d = cir.node_names["d"]
    X = np.zeros(cir.n_total + cir.m)
    X[cir.node_names["in"]] = 2.0
    X[d] = 1.6  # raw junction voltage far above vcrit
    z = np.zeros(cir.n_unk)
    csr, f, Q, I, mrc, lim_out = mna.build_system_lim(cir, X, [2.0], [], z, 0.0, lim_in=[0.0])
    nvt = VT
    vl = nvt * math.log(1.6 / nvt)  # pnjlim from vold = 0
    assert lim_out[0] == pytest.approx(vl, rel=1e-15)
    e = math.exp(vl / nvt)
    i_l = 1e-14 * (e - 1.0)
    g_l = 1e-14 * e / nvt
    # node d row: resistor current + diode current at the limited voltage,
    # minus lim_rhs = g(vl) * (vl - vraw): the diode's tangent at vl
    # evaluated at the raw voltage
    row = d - 1
    want = (1.6 - 2.0) / 1e3 + (i_l - g_l * (vl - 1.6)) + cir.gmin * 1.6
    assert f[row] == pytest.approx(want, rel=1e-12)
    # the Jacobian is the limited one
    pos = cir.diag_pos[row]
    assert csr[pos] == pytest.approx(1e-3 + g_l + cir.gmin, rel=1e-12)


def test_limited_transient_with_rejections_stays_clamped():
    from vajax_amd import generators as G

    net = G.diode_ladder(12, seed=3)
    cir = mna.setup(net)
    cfg = T.AdaptiveConfig(integration_method="gear2")
    out = T.run_adaptive(cir, 2e-9, 1e-12, cfg, T.Options())
    assert out.stats["rejected_steps"] > 0
    # clamp diodes hold the ladder nodes near a forward drop
    idx = [i for name, i in cir.node_names.items() if name.startswith("n")]
    assert out.V[:, idx].max() < 1.2 and out.V[:, idx].min() > -1.2
    # every node voltage stays finite and the run reaches t_stop
    assert np.isfinite(out.V).all() and out.times[-1] == pytest.approx(2e-9)


# ---- internal device nodes and noise-only internal (NOI) nodes ----------

def _noi_chain(noi=("a",), internal=()):
    b = NetlistBuilder("noi_chain")
    b.source("vsource", "v1", "in", "0", type="dc", dc=1.0)
    b.device("resistor", "r1", ["in", "a"], r=1e3)
    b.device("resistor", "r2", ["a", "b"], r=2e3)
    b.device("resistor", "r3", ["b", "0"], r=3e3)
    b.device("capacitor", "c1", ["b", "0"], c=1e-13)
    return b.build(analysis={"stop": 1e-9, "step": 1e-12}, options={"icmode": "op"},
                   internal_nodes=internal, noi_nodes=noi)


def test_noi_node_is_held_at_zero_and_decoupled():
    net = _noi_chain()
    cir = mna.setup(net)
    assert list(cir.noi_res) == [cir.node_names["a"] - 1]
    assert not cir.conv_mask[cir.node_names["a"] - 1]
    X0, _, I_dc, _, info = T.initial_state(cir, T.LinearSolver(cir), T.Options())
    assert info.converged
    # a is clamped to 0 V; b then sits between two grounded resistors; the
    # source drives r1 into the clamped node
    assert X0[cir.node_names["a"]] == 0.0
    assert abs(X0[cir.node_names["b"]]) < 1e-12
    assert I_dc[0] == pytest.approx(-1.0 / 1e3, rel=1e-9)


def test_noi_solve_is_identity_on_noi_rows():
    net = _noi_chain()
    cir = mna.setup(net)
    rng = np.random.default_rng(0)
    csr = rng.uniform(-1, 1, cir.nse)
    f = rng.uniform(-1, 1, cir.n_aug)
    A0 = mna.csr_matrix(cir, csr).toarray()
    csr2, f2 = T._apply_noi(cir, csr, f, cir.noi_res)
    A = mna.csr_matrix(cir, csr2).toarray()
    r = cir.noi_res[0]
    want_row = np.zeros(cir.n_aug)
    want_row[r] = 1.0
    assert np.array_equal(A[r], want_row)
    # the column is zeroed in node rows, kept in branch rows
    assert np.all(A[:cir.n_unk, r][np.arange(cir.n_unk) != r] == 0.0)
    assert np.array_equal(A[cir.n_unk:, r], A0[cir.n_unk:, r])
    assert f2[r] == 0.0 and np.array_equal(np.delete(f2, r), np.delete(f, r))


def test_internal_nodes_are_left_out_of_the_residual_test():
    net = _noi_chain(noi=(), internal=("b",))
    cir = mna.setup(net)
    assert not cir.conv_mask[cir.node_names["b"] - 1] and len(cir.noi_res) == 0
    X0, _, _, _, info = T.initial_state(cir, T.LinearSolver(cir), T.Options())
    assert info.converged
    # the linear divider's solution is unchanged by the mask (gmin = 1e-12 S
    # on each node moves it by ~1e-9 relative)
    want_b = 1.0 * 3e3 / (1e3 + 2e3 + 3e3)
    assert X0[cir.node_names["b"]] == pytest.approx(want_b, rel=1e-8)
    plain = mna.setup(_noi_chain(noi=(), internal=()))
    ref = T.initial_state(plain, T.LinearSolver(plain), T.Options())
    assert X0[cir.node_names["b"]] == ref[0][cir.node_names["b"]]
