"""Sparse LU front end (drop-in for the reference's ``vajax.sparse`` solve).

``solve(indptr, indices, data, b)`` keeps the semantics of
vajax/sparse/umfpack_jax.py:85 / umfpack_ffi.cpp:329-338 on host buffers.
:class:`SparseLU` exposes the handle API of include/vjx.h: one analysis per
pattern, GPU refactorisation per matrix, FP64 or FP32 + FP64-refinement
solves.  Device buffers are torch CUDA tensors (torch is used only to hold
HBM memory).
"""

from __future__ import annotations

import ctypes as C

import numpy as np

from . import _lib as L


def _i32(a):
    return np.ascontiguousarray(a, dtype=np.int32)


def _f64(a):
    return np.ascontiguousarray(a, dtype=np.float64)


class SparseLU:
    """LU handle for one CSR pattern (vjx_pattern)."""

    def __init__(self, indptr, indices, n: int | None = None, _handle=None, _owner=None):
        self.indptr = _i32(indptr)
        self.indices = _i32(indices)
        self.n = len(self.indptr) - 1 if n is None else int(n)
        self.nnz = len(self.indices)
        self._owner = _owner
        if _handle is not None:
            self.h = C.c_void_p(_handle)
            self._owned = False
        else:
            h = C.c_void_p()
            L.check(L.lib.vjx_pattern_create(self.n, self.nnz, L.ptr(self.indptr),
                                              L.ptr(self.indices), C.byref(h)),
                    "vjx_pattern_create")
            self.h = h
            self._owned = True

    def __del__(self):
        # at interpreter shutdown the module globals may already be gone
        lib = getattr(L, "lib", None) if L is not None else None
        if getattr(self, "_owned", False) and self.h and lib is not None:
            lib.vjx_pattern_destroy(self.h)
            self.h = None

    # ---- host-side analysis ----
    def analyze(self, data) -> None:
        d = _f64(data)
        L.check(L.lib.vjx_lu_analyze(self.h, L.ptr(d, C.c_double)), "vjx_lu_analyze")

    def set_band_width(self, width: int) -> None:
        """Levels per triangular-solve band (0: one launch per level); applies
        from the next analysis."""
        L.check(L.lib.vjx_pattern_set_band_width(self.h, int(width)), "vjx_pattern_set_band_width")

    def set_ordering(self, ordering: str = "mlnd", nd_leaf: int = 0) -> None:
        """Fill-reducing order of the next analysis: "md" (minimum degree),
        "natural", "nd" (nested dissection with BFS level-structure
        separators) or "mlnd" (nested dissection with multilevel separators,
        the default); minimum-degree leaves of at most ``nd_leaf`` rows (0 =
        256 for "mlnd", max(48, n / 128) per diagonal block for "nd")."""
        code = {"md": 0, "natural": 1, "nd": 2, "mlnd": 3}[ordering]
        L.check(L.lib.vjx_pattern_set_ordering(self.h, code, int(nd_leaf)),
                "vjx_pattern_set_ordering")

    def set_fill_limit(self, max_lu_nnz: int) -> None:
        """Later analyses stop with VJX_ERR_FILL (the previous analysis stays
        in effect) once L + U would hold more than ``max_lu_nnz`` entries;
        0 = no limit."""
        L.check(L.lib.vjx_pattern_set_fill_limit(self.h, int(max_lu_nnz)),
                "vjx_pattern_set_fill_limit")

    def set_dense(self, min_rows: int = 96, density: float = 0.4) -> None:
        """Dense trailing block of the next analysis: the largest trailing
        block of at least ``min_rows`` rows with LU density >= ``density``
        (``min_rows`` 0: none)."""
        L.check(L.lib.vjx_pattern_set_dense(self.h, int(min_rows), float(density)),
                "vjx_pattern_set_dense")

    def check_schedules(self) -> None:
        L.check(L.lib.vjx_lu_check_schedules(self.h), "vjx_lu_check_schedules")

    def info(self) -> dict:
        info = L.LuInfo()
        L.check(L.lib.vjx_lu_get_info(self.h, C.byref(info)), "vjx_lu_get_info")
        return info.as_dict()

    def perm(self):
        p = np.zeros(self.n, np.int32)
        q = np.zeros(self.n, np.int32)
        L.check(L.lib.vjx_lu_get_perm(self.h, L.ptr(p), L.ptr(q)), "vjx_lu_get_perm")
        return p, q

    def lu_pattern(self):
        info = self.info()
        rp = np.zeros(self.n + 1, np.int32)
        col = np.zeros(int(info["nnz_l"] + info["nnz_u"]), np.int32)
        L.check(L.lib.vjx_lu_get_pattern(self.h, L.ptr(rp), L.ptr(col)), "vjx_lu_get_pattern")
        return rp, col

    def refactor_cpu(self, data, precision: int = L.VJX_PREC_F64) -> np.ndarray:
        d = _f64(data)
        info = self.info() if self._analyzed() else None
        if info is None:
            self.analyze(d)
            info = self.info()
        out = np.zeros(int(info["nnz_l"] + info["nnz_u"]))
        L.check(L.lib.vjx_lu_refactor_cpu(self.h, L.ptr(d, C.c_double), precision,
                                          L.ptr(out, C.c_double)), "vjx_lu_refactor_cpu")
        return out

    def solve_cpu(self, data, b, precision: int = L.VJX_PREC_F64) -> np.ndarray:
        d = _f64(data)
        bb = _f64(b)
        x = np.zeros(self.n)
        L.check(L.lib.vjx_lu_solve_cpu(self.h, L.ptr(d, C.c_double), precision,
                                       L.ptr(bb, C.c_double), L.ptr(x, C.c_double)),
                "vjx_lu_solve_cpu")
        return x

    def _analyzed(self) -> bool:
        info = L.LuInfo()
        return L.lib.vjx_lu_get_info(self.h, C.byref(info)) == 0

    # ---- device path (torch CUDA tensors) ----
    def refactor(self, data_dev, precision: int = L.VJX_PREC_F64) -> int:
        return L.check(L.lib.vjx_lu_refactor(self.h, C.c_void_p(data_dev.data_ptr()), precision),
                       "vjx_lu_refactor")

    def solve(self, data_dev, b_dev, x_dev, refine_steps: int = 0) -> None:
        L.check(L.lib.vjx_lu_solve(self.h, C.c_void_p(data_dev.data_ptr()),
                                   C.c_void_p(b_dev.data_ptr()), C.c_void_p(x_dev.data_ptr()),
                                   refine_steps), "vjx_lu_solve")

    def solve_transpose(self, data_dev, b_dev, x_dev, precision: int = L.VJX_PREC_F64,
                        refine_steps: int = 0) -> int:
        """A^T x = b: factorises A^T (a second handle kept by this one) and solves."""
        return L.check(L.lib.vjx_lu_solve_transpose(self.h, C.c_void_p(data_dev.data_ptr()),
                                                    C.c_void_p(b_dev.data_ptr()),
                                                    C.c_void_p(x_dev.data_ptr()), precision,
                                                    refine_steps), "vjx_lu_solve_transpose")

    def spmv(self, data_dev, x_dev, y_dev) -> None:
        L.check(L.lib.vjx_spmv(self.h, C.c_void_p(data_dev.data_ptr()),
                               C.c_void_p(x_dev.data_ptr()), C.c_void_p(y_dev.data_ptr())),
                "vjx_spmv")

    def download(self) -> np.ndarray:
        info = self.info()
        out = np.zeros(int(info["nnz_l"] + info["nnz_u"]))
        L.check(L.lib.vjx_lu_download(self.h, L.ptr(out, C.c_double)), "vjx_lu_download")
        return out


def solve(indptr, indices, data, b) -> np.ndarray:
    """x = A^{-1} b for a CSR matrix on host buffers (GPU factorisation)."""
    ip, ix = _i32(indptr), _i32(indices)
    d, bb = _f64(data), _f64(b)
    n = len(ip) - 1
    x = np.zeros(n)
    L.check(L.lib.vjx_solve_host(n, len(ix), L.ptr(ip), L.ptr(ix), L.ptr(d, C.c_double),
                                 L.ptr(bb, C.c_double), L.ptr(x, C.c_double)), "vjx_solve_host")
    return x


def _host_op(fn, name, indptr, indices, data, v) -> np.ndarray:
    ip, ix = _i32(indptr), _i32(indices)
    d, vv = _f64(data), _f64(v)
    n = len(ip) - 1
    out = np.zeros(n)
    L.check(fn(n, len(ix), L.ptr(ip), L.ptr(ix), L.ptr(d, C.c_double), L.ptr(vv, C.c_double),
               L.ptr(out, C.c_double)), name)
    return out


def solve_transpose(indptr, indices, data, b) -> np.ndarray:
    """x = A^{-T} b (vajax/sparse/umfpack_jax.py:148-177) on host buffers."""
    return _host_op(L.lib.vjx_solve_transpose_host, "vjx_solve_transpose_host", indptr, indices,
                    data, b)


def dot(indptr, indices, data, x) -> np.ndarray:
    """b = A x (vajax/sparse/umfpack_jax.py:118-145) on host buffers."""
    return _host_op(L.lib.vjx_dot_host, "vjx_dot_host", indptr, indices, data, x)


def clear_cache() -> None:
    L.lib.vjx_clear_cache()
