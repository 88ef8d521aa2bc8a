"""Host side of the sparse LU (ordering, pivoting, level program) on the CPU.

The CPU execution of the numeric program (``vjx_lu_refactor_cpu`` /
``vjx_lu_solve_cpu``) runs exactly the update lists the GPU kernels run, so
these tests pin the symbolic analysis without a GPU: solutions against
SciPy's SuperLU and the residual norm, valid permutations, the LU pattern
containing A, and L*U reproducing P*A*Q.
"""

import numpy as np
import pytest
import scipy.sparse as sp
import scipy.sparse.linalg as spla

from oracle import mna
from vajax_amd import generators as G
from vajax_amd.sparse import SparseLU


def _system(net, c0=1e12, seed=0):
    cir = mna.setup(net)
    rng = np.random.default_rng(seed)
    X = np.zeros(cir.n_total + cir.m)
    # node voltages inside the rails (supply at 1 V): the EKV guard branch for
    # gate-bulk voltages far below the pinch-off range is not exercised here
    X[1:cir.n_total] = rng.uniform(0.2, 0.8, cir.n_unk)
    if "vdd" in cir.node_names:
        X[cir.node_names["vdd"]] = 1.0
    csr, *_ = mna.build_system(cir, X, np.zeros(cir.m), np.zeros(net.isources.count),
                               np.zeros(cir.n_unk), c0)
    A = sp.csr_matrix((csr, cir.indices, cir.indptr), shape=(cir.n_aug, cir.n_aug))
    return cir, csr, A


@pytest.mark.parametrize("net", [G.rc_ladder(200), G.ring_oscillator(31), G.gate_dag(800, seed=2)],
                         ids=["rc", "ring", "dag"])
def test_solve_matches_superlu(net):
    cir, csr, A = _system(net)
    lu = SparseLU(cir.indptr, cir.indices)
    lu.analyze(csr)
    b = np.random.default_rng(1).standard_normal(cir.n_aug)
    x = lu.solve_cpu(csr, b)
    xs = spla.spsolve(A.tocsc(), b)
    assert np.linalg.norm(A @ x - b) <= 1e-12 * np.linalg.norm(b)
    assert np.allclose(x, xs, rtol=1e-9, atol=1e-12 * np.max(np.abs(xs)))
    x32 = lu.solve_cpu(csr, b, 1)
    assert np.linalg.norm(A @ x32 - b) <= 1e-5 * np.linalg.norm(b)


def test_factors_reproduce_permuted_matrix():
    net = G.gate_dag(300, seed=5)
    cir, csr, A = _system(net)
    lu = SparseLU(cir.indptr, cir.indices)
    lu.analyze(csr)
    p, q = lu.perm()
    assert sorted(p) == list(range(cir.n_aug)) and sorted(q) == list(range(cir.n_aug))
    rp, col = lu.lu_pattern()
    vals = lu.refactor_cpu(csr)
    n = cir.n_aug
    F = sp.csr_matrix((vals, col, rp), shape=(n, n)).toarray()
    Lm = np.tril(F, -1) + np.eye(n)
    Um = np.triu(F)
    PAQ = A.toarray()[p][:, q]
    assert np.allclose(Lm @ Um, PAQ, rtol=0, atol=1e-12 * np.max(np.abs(PAQ)))
    # every A entry is inside the LU pattern
    pinv = np.argsort(p)
    qinv = np.argsort(q)
    lu_set = set(zip(np.repeat(np.arange(n), np.diff(rp)).tolist(), col.tolist()))
    rows = np.repeat(np.arange(n), np.diff(cir.indptr))
    for r, c in zip(rows.tolist(), cir.indices.tolist()):
        assert (int(pinv[r]), int(qinv[c])) in lu_set


def test_zero_diagonal_branch_rows_are_pivoted():
    # voltage-source branch rows have no diagonal entry (mna.py:650-667)
    net = G.rc_ladder(10)
    cir, csr, A = _system(net)
    assert A.diagonal()[-1] == 0.0
    lu = SparseLU(cir.indptr, cir.indices)
    lu.analyze(csr)
    info = lu.info()
    assert info["n_offdiag_pivots"] >= 1
    b = np.ones(cir.n_aug)
    x = lu.solve_cpu(csr, b)
    assert np.linalg.norm(A @ x - b) <= 1e-12 * np.linalg.norm(b)


def test_level_program_statistics():
    net = G.gate_dag(500, seed=1)
    cir, csr, A = _system(net)
    lu = SparseLU(cir.indptr, cir.indices)
    lu.analyze(csr)
    info = lu.info()
    assert info["nnz_l"] + info["nnz_u"] >= A.nnz
    assert info["fact_levels"] >= 1 and info["fwd_levels"] >= 1
    assert info["n_updates"] > 0 and info["flops"] >= 2 * info["n_updates"]
    assert info["n_dense_rows"] == 1  # the supply rail


def test_packed_batch_shares_block_analysis():
    # Monte Carlo copies + ragged topologies packed block-diagonally: one
    # analysis per distinct block pattern, laid out block after block
    nets = [G.monte_carlo(G.gate_dag(150, seed=4), 10 + b, 0.03) for b in range(3)]
    nets += [G.gate_dag(60, seed=7), G.gate_dag(90, seed=8)]
    packed = G.pack(nets)
    cir, csr, A = _system(packed, seed=3)
    lu = SparseLU(cir.indptr, cir.indices)
    lu.analyze(csr)
    info = lu.info()
    assert info["n_blocks"] > len(nets)  # instances + isolated unused-input branches
    # 3 MC copies share one pattern; the two ragged DAGs add two more
    assert info["n_block_patterns"] <= info["n_blocks"] - 2
    b = np.random.default_rng(2).standard_normal(cir.n_aug)
    x = lu.solve_cpu(csr, b)
    assert np.linalg.norm(A @ x - b) <= 1e-12 * np.linalg.norm(b)
    x32 = lu.solve_cpu(csr, b, 1)
    assert np.linalg.norm(A @ x32 - b) <= 1e-5 * np.linalg.norm(b)
    # the merged level program runs as many levels as the deepest block
    single = []
    for net in nets:
        c1, d1, _ = _system(net, seed=3)
        l1 = SparseLU(c1.indptr, c1.indices)
        l1.analyze(d1)
        single.append(l1.info())
    assert info["fact_levels"] <= max(s["fact_levels"] for s in single)
    assert info["nnz_l"] + info["nnz_u"] >= sum(s["nnz_l"] + s["nnz_u"] for s in single) * 0.5


def test_packed_batch_block_equals_standalone():
    # the first Monte Carlo instance's block of the packed LU equals the
    # standalone LU of that instance bit for bit (same pattern, same values,
    # same pivot order, same fixed-order update program)
    nets = [G.monte_carlo(G.gate_dag(120, seed=9), 30 + b, 0.03) for b in range(3)]
    packed = G.pack(nets)
    cir = mna.setup(packed)
    c0 = mna.setup(nets[0])
    rng = np.random.default_rng(5)
    X0 = np.zeros(c0.n_total + c0.m)
    X0[1:c0.n_total] = rng.uniform(0.2, 0.8, c0.n_unk)
    # same node voltages for every instance, mapped through the prefixed names
    X = np.zeros(cir.n_total + cir.m)
    for name, k in c0.node_names.items():
        if k == 0:
            continue  # ground is shared
        for b in range(3):
            X[cir.node_names[f"i{b:03d}.{name}"]] = X0[k]
    d0, *_ = mna.build_system(c0, X0, np.zeros(c0.m), np.zeros(0), np.zeros(c0.n_unk), 1e12)
    dp, *_ = mna.build_system(cir, X, np.zeros(cir.m), np.zeros(0), np.zeros(cir.n_unk), 1e12)
    l0 = SparseLU(c0.indptr, c0.indices)
    l0.analyze(d0)
    v0 = l0.refactor_cpu(d0)
    lp = SparseLU(cir.indptr, cir.indices)
    lp.analyze(dp)
    vp = lp.refactor_cpu(dp)
    # instance 0 owns the lowest-numbered vertices, so its block comes first
    i0 = lp.info()
    assert i0["n_blocks"] >= 3
    assert np.array_equal(vp[: len(v0)], v0)


@pytest.mark.parametrize("width", [0, 1, 4, 16, 64])
def test_band_programs_respect_dependencies(width):
    # bands of the triangular solves: every row once, after the rows it reads;
    # the host solve under that program still matches SuperLU
    net = G.gate_dag(900, seed=6)
    cir, csr, A = _system(net)
    lu = SparseLU(cir.indptr, cir.indices)
    lu.set_band_width(width)
    lu.analyze(csr)
    lu.check_schedules()
    b = np.random.default_rng(4).standard_normal(cir.n_aug)
    x = lu.solve_cpu(csr, b)
    assert np.linalg.norm(A @ x - b) <= 1e-12 * np.linalg.norm(b)


def test_band_program_with_hub_row():
    # a supply row longer than one hub chunk (4096 entries) is finished by a
    # chunked reduction after its band; the solve stays exact
    net = G.gate_dag(4500, seed=2)
    cir, csr, A = _system(net)
    lu = SparseLU(cir.indptr, cir.indices)
    lu.analyze(csr)
    lu.check_schedules()
    rp, col = lu.lu_pattern()
    assert np.max(np.diff(rp)) > 4096
    b = np.random.default_rng(8).standard_normal(cir.n_aug)
    x = lu.solve_cpu(csr, b)
    assert np.linalg.norm(A @ x - b) <= 1e-12 * np.linalg.norm(b)


@pytest.mark.parametrize("width", [1, 3, 16])
def test_band_refactorisation_program(width):
    # refactorisation in bands: each update reads only final slots and no two
    # tasks of one launch write the same slot; L*U reproduces P*A*Q
    net = G.gate_dag(400, seed=21)
    cir, csr, A = _system(net)
    lu = SparseLU(cir.indptr, cir.indices)
    lu.set_band_width(width)
    lu.analyze(csr)
    lu.check_schedules()
    info = lu.info()
    assert info["fact_launches"] <= 2 * ((info["fact_levels"] + width - 1) // width)
    p, q = lu.perm()
    rp, col = lu.lu_pattern()
    vals = lu.refactor_cpu(csr)
    n = cir.n_aug
    F = sp.csr_matrix((vals, col, rp), shape=(n, n)).toarray()
    PAQ = A.toarray()[p][:, q]
    LU = (np.tril(F, -1) + np.eye(n)) @ np.triu(F)
    assert np.allclose(LU, PAQ, rtol=0, atol=1e-12 * np.max(np.abs(PAQ)))


def test_packed_batch_band_programs_checked():
    nets = [G.monte_carlo(G.gate_dag(150, seed=4), 10 + b, 0.03) for b in range(3)]
    nets += [G.gate_dag(60, seed=7), G.gate_dag(90, seed=8)]
    cir, csr, A = _system(G.pack(nets), seed=3)
    lu = SparseLU(cir.indptr, cir.indices)
    lu.analyze(csr)
    lu.check_schedules()


@pytest.mark.parametrize("packed", [False, True], ids=["single", "packed"])
def test_fill_limit_abandons_the_analysis_and_keeps_the_previous_one(packed):
    # the transient loop's pivot monitor re-analyses under a fill limit: an
    # analysis beyond it fails with VJX_ERR_FILL and the previous pivots stay
    import ctypes as C

    from vajax_amd import _lib as L

    if packed:
        net = G.pack([G.monte_carlo(G.gate_dag(150, seed=4), 10 + b, 0.03) for b in range(3)])
    else:
        net = G.gate_dag(300, seed=5)
    cir, csr, A = _system(net)
    lu = SparseLU(cir.indptr, cir.indices)
    lu.analyze(csr)
    info = lu.info()
    nnz = info["nnz_l"] + info["nnz_u"]
    vals = lu.refactor_cpu(csr)
    p0, q0 = lu.perm()
    _, csr2, A2 = _system(net, seed=7)
    d2 = np.ascontiguousarray(csr2)
    lu.set_fill_limit(nnz // 2)
    assert L.lib.vjx_lu_analyze(lu.h, L.ptr(d2, C.c_double)) == L.VJX_ERR_FILL
    assert "fill limit" in L.last_error()
    after = lu.info()
    assert after["nnz_l"] + after["nnz_u"] == nnz
    p1, q1 = lu.perm()
    assert np.array_equal(p0, p1) and np.array_equal(q0, q1)
    assert np.array_equal(lu.refactor_cpu(csr), vals)
    # without a limit the same matrix analyses and solves
    lu.set_fill_limit(0)
    lu.analyze(d2)
    b = np.random.default_rng(3).standard_normal(cir.n_aug)
    x = lu.solve_cpu(d2, b)
    assert np.linalg.norm(A2 @ x - b) <= 1e-12 * np.linalg.norm(b)
    with pytest.raises(L.VjxError):
        lu.set_fill_limit(-1)


def test_multilevel_dissection_beats_level_structures_and_is_deterministic():
    # multilevel separators (the default ordering) against BFS level-structure
    # separators on a gate DAG and an array multiplier (whose BFS separators
    # are wide): fewer LU updates and fewer pivot levels, the same solution,
    # and the same permutation on every analysis
    from vajax_amd import vacask

    nets = [G.gate_dag(2000, seed=0),
            vacask.load("", text=G.array_multiplier_testbench(16))]
    for net in nets:
        cir, csr, A = _system(net)
        stats = {}
        for order in ("nd", "mlnd"):
            lu = SparseLU(cir.indptr, cir.indices)
            lu.set_ordering(order)
            lu.analyze(csr)
            lu.check_schedules()
            stats[order] = lu.info()
            b = np.random.default_rng(4).standard_normal(cir.n_aug)
            x = lu.solve_cpu(csr, b)
            assert np.linalg.norm(A @ x - b) <= 1e-12 * np.linalg.norm(b)
            if order == "mlnd":
                p1, q1 = lu.perm()
                again = SparseLU(cir.indptr, cir.indices)
                again.analyze(csr)
                p2, q2 = again.perm()
                assert np.array_equal(p1, p2) and np.array_equal(q1, q2)
        assert stats["mlnd"]["n_updates"] < 0.6 * stats["nd"]["n_updates"], stats
        assert stats["mlnd"]["fact_levels"] < stats["nd"]["fact_levels"], stats


def test_automatic_leaf_breaks_chains_on_small_systems():
    """Ordering 3 with the automatic leaf: a 200-section RC ladder (a chain,
    one pivot level per row inside a 256-row minimum-degree leaf) takes
    4-row leaves and a logarithmic level count; a 2k-gate DAG, whose levels
    come from its separators, keeps 256-row leaves; an explicit leaf is
    taken as given.  Every program solves its system."""
    for net, leaf, max_levels in ((G.rc_ladder(200), 4, 16), (G.gate_dag(2000, seed=0), 256, None)):
        cir, csr, A = _system(net)
        lu = SparseLU(cir.indptr, cir.indices)
        lu.analyze(csr)
        lu.check_schedules()
        info = lu.info()
        assert info["nd_leaf"] == leaf, info["nd_leaf"]
        if max_levels:
            assert info["fact_levels"] <= max_levels, info["fact_levels"]
            fixed = SparseLU(cir.indptr, cir.indices)
            fixed.set_ordering("mlnd", 256)
            fixed.analyze(csr)
            assert fixed.info()["nd_leaf"] == 256 and fixed.info()["fact_levels"] >= cir.n_aug // 2
        b = np.random.default_rng(1).standard_normal(cir.n_aug)
        x = lu.solve_cpu(csr, b)
        assert np.linalg.norm(A @ x - b) <= 1e-12 * np.linalg.norm(b)
