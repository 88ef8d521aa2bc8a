"""vjx_lu_solve_noalloc (the allocation-free form of vjx_lu_solve) against
vjx_lu_solve through the C ABI: the same kernels on the same factors, so the
results must be identical bit for bit, in FP64 and in FP32 with FP64
refinement, on two calls in a row of which the second is a smaller system.
The size limit is checked on the host before any device work."""

import ctypes as C

import numpy as np
import pytest

from vajax_amd import _lib as L
from vajax_amd import generators as G
from vajax_amd.engine import DeviceCircuit
from vajax_amd.sparse import SparseLU
from vajax_amd.topology import build_topology


def _p(t):
    return C.c_void_p(t.data_ptr())


def test_noalloc_rejects_oversized_system_on_host():
    n = (1 << 19) + 1
    indptr = np.arange(n + 1, dtype=np.int32)
    indices = np.arange(n, dtype=np.int32)
    lu = SparseLU(indptr, indices)
    rc = L.lib.vjx_lu_solve_noalloc(lu.h, None, C.c_void_p(8), C.c_void_p(8), 0)
    assert rc == L.VJX_ERR_ARG
    assert "scratch limit" in L.last_error()


def _jacobian(net):
    import torch

    topo = build_topology(net)
    dc = DeviceCircuit(topo, net)
    rng = np.random.default_rng(5)
    n_total, m = topo.n_total, topo.n_vsources
    X = torch.tensor(np.concatenate([[0.0], rng.uniform(0, 1, n_total - 1), np.zeros(m)]),
                     device="cuda")
    vs = torch.tensor(np.full(max(m, 1), 1.0)[:m], device="cuda")
    isv = torch.zeros(net.isources.count, dtype=torch.float64, device="cuda")
    out = dc.build_system(X, vs, isv, c0=1e12)
    return topo, out["csr_data"], out["f"]


@pytest.mark.gpu
@pytest.mark.parametrize("precision,refine", [(L.VJX_PREC_F64, 0), (L.VJX_PREC_F32, 1),
                                              (L.VJX_PREC_F32, 2)])
def test_noalloc_solve_equals_solve(precision, refine):
    import torch

    torch.cuda.set_device(0)
    for net in (G.gate_dag(3000, seed=4), G.gate_dag(200, seed=9)):  # second call: smaller
        topo, csr, f = _jacobian(net)
        lu = SparseLU(topo.indptr, topo.indices)
        lu.refactor(csr, precision)
        b = -f
        x_ref = torch.zeros_like(b)
        x_new = torch.full_like(b, float("nan"))
        lu.solve(csr, b, x_ref, refine_steps=refine)
        rc = L.lib.vjx_lu_solve_noalloc(lu.h, _p(csr), _p(b), _p(x_new), refine)
        assert rc == 0, L.last_error()
        assert torch.equal(x_new, x_ref), float((x_new - x_ref).abs().max())
        # residual sanity: the solve is a solve
        y = torch.zeros_like(b)
        lu.spmv(csr, x_new, y)
        rel = float((y - b).abs().max() / b.abs().max())
        assert rel < (1e-10 if precision == L.VJX_PREC_F64 else 1e-5), rel
