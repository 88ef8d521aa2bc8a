"""The dense trailing block of the LU refactorisation on the FP64 matrix
cores: the MFMA matrix product against numpy (which also pins the operand
layout of v_mfma_f64_16x16x4f64), the blocked dense LU against L U = A, and
the refactorisation of an array multiplier's Jacobian (whose top separators
fill in densely) against the CPU execution of the same program and SciPy."""

import ctypes as C
import sys

import numpy as np
import pytest

from vajax_amd import _lib as L
from vajax_amd import generators as G
from vajax_amd import vacask

pytestmark = pytest.mark.gpu


def _dev(torch, a):
    return torch.as_tensor(np.ascontiguousarray(a, dtype=np.float64), device="cuda")


def _p(t):
    return C.c_void_p(t.data_ptr())


@pytest.mark.parametrize("M,N,Kd", [(16, 16, 4), (70, 45, 37), (128, 192, 64), (300, 257, 129)])
def test_mfma_product(M, N, Kd):
    import torch

    rng = np.random.default_rng(M + N + Kd)
    A, B, Cm = rng.standard_normal((M, Kd)), rng.standard_normal((Kd, N)), rng.standard_normal((M, N))
    # operands inside larger arrays (leading dimensions above the widths)
    Ab = np.zeros((M, Kd + 3)); Ab[:, :Kd] = A
    Bb = np.zeros((Kd, N + 5)); Bb[:, :N] = B
    Cb = np.zeros((M, N + 2)); Cb[:, :N] = Cm
    dA, dB, dC = _dev(torch, Ab), _dev(torch, Bb), _dev(torch, Cb)
    rc = L.lib.vjx_dense_gemm_sub(M, N, Kd, _p(dA), Kd + 3, _p(dB), N + 5, _p(dC), N + 2)
    assert rc == 0, L.last_error()
    got = dC.cpu().numpy()
    want = Cm - A @ B
    scale = np.abs(A) @ np.abs(B) + np.abs(Cm)
    assert np.all(np.abs(got[:, :N] - want) <= 1e-14 * scale)
    assert np.all(got[:, N:] == 0.0)  # nothing written outside the tile


@pytest.mark.parametrize("K", [1, 31, 33, 150, 333])
def test_dense_lu(K):
    import torch

    rng = np.random.default_rng(K)
    A = rng.standard_normal((K, K)) + K * np.eye(K)
    dA = _dev(torch, A)
    assert L.lib.vjx_dense_lu(K, _p(dA)) == 0, L.last_error()
    F = dA.cpu().numpy()
    Lm = np.tril(F, -1) + np.eye(K)
    Um = np.triu(F)
    assert np.max(np.abs(Lm @ Um - A)) <= 1e-12 * np.max(np.abs(A))


def _multiplier_jacobian(n):
    from oracle import mna as omna
    from oracle.sources import source_values
    from vajax_amd.topology import build_topology

    net = vacask.load("", text=G.array_multiplier_testbench(n))
    cir = omna.setup(net)
    topo = build_topology(net)
    rng = np.random.default_rng(1)
    X = np.zeros(cir.n_total + cir.m)
    X[1:cir.n_total] = rng.uniform(0, 1.2, cir.n_total - 1)
    csr, *_ = omna.build_system(cir, X, source_values(net.vsources, 5e-13), np.zeros(0),
                                np.zeros(cir.n_unk), 1e12)
    return cir, topo, omna.csr_matrix(cir, csr), csr


@pytest.mark.parametrize("precision", [L.VJX_PREC_F64, L.VJX_PREC_F32])
@pytest.mark.parametrize("ordering", ["mlnd", "nd"])
def test_multiplier_refactorisation_with_dense_block(ordering, precision):
    import torch

    from vajax_amd.sparse import SparseLU

    # the 16x16 multiplier keeps a dense trailing block of 144 rows under the
    # default (multilevel) ordering and of 518 under BFS level structures
    print(f"dense: 16x16 multiplier, ordering {ordering}", file=sys.stderr, flush=True)
    cir, topo, A, csr = _multiplier_jacobian(16)
    lu = SparseLU(topo.indptr, topo.indices)
    lu.set_ordering(ordering)
    lu.analyze(csr)
    lu.check_schedules()
    info = lu.info()
    assert info["dense_k"] >= 96 and info["dense_m"] > 0, info
    d = _dev(torch, csr)
    lu.refactor(d, precision)
    got = lu.download()
    want = lu.refactor_cpu(csr, precision)
    tol = 1e-12 if precision == L.VJX_PREC_F64 else 1e-5
    assert np.all(np.abs(got - want) <= tol * (np.abs(want) + 1.0))
    b = np.random.default_rng(2).standard_normal(cir.n_aug)
    x = torch.zeros(cir.n_aug, dtype=torch.float64, device="cuda")
    lu.solve(d, _dev(torch, b), x, refine_steps=1 if precision == L.VJX_PREC_F32 else 0)
    xs = x.cpu().numpy()
    assert np.linalg.norm(A @ xs - b) / np.linalg.norm(b) < 1e-12
