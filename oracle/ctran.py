"""ctypes front end of the C++ CPU restatement (``oracle/ctran.cpp``).

Test infrastructure and the CPU baseline: the same algorithm as
``oracle.transient`` (numpy), restated in C++ with OpenMP over device
instances and a serial KLU-style sparse LU (Gilbert-Peierls factorisation
with threshold pivoting, then refactorisation on the fixed pivot order).

Setup runs here, in Python, and is excluded from timing as the reference
excludes JIT and symbolic analysis: the circuit arrays come from
``oracle.mna.setup``; the fill-reducing column order is SuperLU's minimum
degree on A+A^T (scipy ``splu(permc_spec="MMD_AT_PLUS_A")``), the ordering
family KLU and UMFPACK use for circuit matrices.

    cir = oracle.mna.setup(net)
    ct = CTran(cir)
    out = ct.run_adaptive(t_stop, dt, AdaptiveConfig(...), Options(...), threads=8)

``out`` is an ``oracle.transient.TransientOut`` like the numpy loops return.
"""

from __future__ import annotations

import ctypes as C
import os
import subprocess
from pathlib import Path

import numpy as np

from vajax_amd.netlist import SOURCE_DEFAULTS

from . import mna
from .transient import (GEAR2, TRAP, AdaptiveConfig, Options, TransientOut, _init_mid_rail,
                        method_of)

HERE = Path(__file__).resolve().parent
SRC = HERE / "ctran.cpp"
LIB = HERE / "libctran.so"

i64p = C.POINTER(C.c_int64)
i32p = C.POINTER(C.c_int32)
f64p = C.POINTER(C.c_double)

_SRC_KEYS = ("dc", "val0", "val1", "rise", "fall", "width", "period", "delay")
_KIND = {"resistor": 0, "capacitor": 1, "mosfet": 2}
_MOS_KEYS = ("sgn", "vgoff", "phi", "gamma", "beta0", "theta", "cgs", "cgd", "cgb", "cdb", "csb")
_METHOD = {"be": 0, TRAP: 1, GEAR2: 2}


class _Model(C.Structure):
    _fields_ = [("kind", C.c_int32), ("count", C.c_int32), ("nodes", i64p), ("csr_pos", i64p),
                ("res_idx", i64p), ("par", f64p)]


class _Sources(C.Structure):
    _fields_ = [("count", C.c_int32), ("p", i64p), ("n", i64p), ("wave", i32p), ("par", f64p)]


class _Desc(C.Structure):
    _fields_ = [("n_total", C.c_int32), ("n_unk", C.c_int32), ("m", C.c_int32),
                ("n_models", C.c_int32), ("nse", C.c_int64), ("indptr", i64p), ("indices", i64p),
                ("diag_pos", i64p), ("nb", C.c_int32), ("b_pos", i64p), ("b_val", f64p),
                ("vs", _Sources), ("is_", _Sources), ("models", _Model * 4),
                ("col_order", i64p), ("gmin", C.c_double)]


class _Options(C.Structure):
    _fields_ = [("reltol", C.c_double), ("abstol", C.c_double), ("vntol", C.c_double),
                ("tran_itl", C.c_int32), ("nr_damping", C.c_double), ("uic", C.c_int32)]


class _Adaptive(C.Structure):
    _fields_ = [("lte_ratio", C.c_double), ("redo_factor", C.c_double), ("reltol", C.c_double),
                ("abstol", C.c_double), ("min_dt", C.c_double), ("max_dt", C.c_double),
                ("tran_minpts", C.c_int32), ("warmup_steps", C.c_int32),
                ("max_order", C.c_int32), ("grow_factor", C.c_double),
                ("gshunt_init", C.c_double), ("gshunt_steps", C.c_int32),
                ("gshunt_target", C.c_double), ("tran_fs", C.c_double), ("tran_ft", C.c_double),
                ("method", C.c_int32), ("max_consecutive_rejects", C.c_int32)]


class _Stats(C.Structure):
    _fields_ = [("n_steps", C.c_int64), ("rejected", C.c_int64), ("total_nr", C.c_int64),
                ("lu_factors", C.c_int64), ("lu_refactors", C.c_int64),
                ("body_iterations", C.c_int64), ("wall", C.c_double),
                ("min_dt_used", C.c_double), ("max_dt_used", C.c_double),
                ("t_build", C.c_double), ("t_lu", C.c_double), ("t_solve", C.c_double),
                ("dc_converged", C.c_int32), ("dc_iterations", C.c_int32),
                ("threads", C.c_int32)]


def build(force: bool = True) -> Path:
    """Compile ``ctran.cpp`` with g++ (-O3, OpenMP, no FMA contraction)."""
    if not force and LIB.exists():
        return LIB
    tmp = LIB.with_suffix(".so.tmp")
    cmd = ["g++", "-O3", "-std=c++17", "-fPIC", "-shared", "-fopenmp", "-ffp-contract=off",
           str(SRC), "-o", str(tmp)]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"oracle C++ build failed:\n{r.stderr}")
    os.replace(tmp, LIB)
    return LIB


_lib = None


def _load():
    global _lib
    if _lib is None:
        if not LIB.exists():
            build()
        lib = C.CDLL(str(LIB))
        lib.ct_create.argtypes = [C.POINTER(_Desc), C.POINTER(C.c_void_p)]
        lib.ct_destroy.argtypes = [C.c_void_p]
        lib.ct_destroy.restype = None
        lib.ct_last_error.restype = C.c_char_p
        lib.ct_build_system.argtypes = [C.c_void_p] + [f64p] * 4 + [C.c_double] * 4 + [f64p] + \
            [C.c_double] + [f64p] * 6
        lib.ct_lu_solve.argtypes = [C.c_void_p, f64p, f64p, f64p, C.c_int32, i64p]
        lib.ct_run_adaptive.argtypes = [C.c_void_p, C.c_double, C.c_double, C.POINTER(_Options),
                                        C.POINTER(_Adaptive), C.c_int64, f64p, C.c_int32, f64p,
                                        f64p, f64p, C.POINTER(_Stats)]
        lib.ct_run_fixed.argtypes = [C.c_void_p, C.c_double, C.c_int32, C.POINTER(_Options),
                                     C.c_int64, C.c_double, f64p, C.c_int32, f64p, f64p, f64p,
                                     C.POINTER(_Stats)]
        _lib = lib
    return _lib


def _p(a, t=f64p):
    return a.ctypes.data_as(t)


def column_order(cir) -> np.ndarray:
    """q[k] = original column in pivot position k: SuperLU's minimum degree on
    A+A^T, read from ``splu``'s ``perm_c`` (column i of A moves to position
    perm_c[i])."""
    import scipy.sparse as sp
    import scipy.sparse.linalg as spla

    n = cir.n_aug
    pat = sp.csr_matrix((np.ones(cir.nse), cir.indices, cir.indptr), shape=(n, n)).tocsc()
    # a structurally nonsingular stand-in with the pattern of A: the order
    # depends on the pattern only
    A = pat + sp.identity(n, format="csc") * float(n)
    lu = spla.splu(A, permc_spec="MMD_AT_PLUS_A", diag_pivot_thresh=0.0,
                   options={"SymmetricMode": True})
    q = np.empty(n, dtype=np.int64)
    q[lu.perm_c] = np.arange(n)
    return q


class CTran:
    def __init__(self, cir: mna.OracleCircuit, col_order=None):
        lib = _load()
        self.cir = cir
        self._keep = []

        def arr(a, dt):
            a = np.ascontiguousarray(a, dtype=dt)
            self._keep.append(a)
            return a

        d = _Desc()
        d.n_total, d.n_unk, d.m = cir.n_total, cir.n_unk, cir.m
        d.nse = cir.nse
        d.indptr = _p(arr(cir.indptr, np.int64), i64p)
        d.indices = _p(arr(cir.indices, np.int64), i64p)
        d.diag_pos = _p(arr(cir.diag_pos, np.int64), i64p)
        d.nb = len(cir.b_pos)
        d.b_pos = _p(arr([p for p, _ in cir.b_pos] or [0], np.int64), i64p)
        d.b_val = _p(arr([v for _, v in cir.b_pos] or [0.0], np.float64))

        def sources(grp, p, n):
            if grp.count and np.any(np.asarray(grp.wave) > 1):
                raise ValueError("the C++ restatement evaluates dc and pulse sources only")
            s = _Sources()
            s.count = grp.count
            s.p = _p(arr(p if grp.count else [0], np.int64), i64p)
            s.n = _p(arr(n if grp.count else [0], np.int64), i64p)
            s.wave = _p(arr(grp.wave if grp.count else [0], np.int32), i32p)
            par = np.stack([np.asarray(grp.params.get(k, np.full(grp.count, SOURCE_DEFAULTS[k])),
                                       dtype=np.float64) for k in _SRC_KEYS], axis=1)
            s.par = _p(arr(par if grp.count else np.zeros(8), np.float64))
            return s

        d.vs = sources(cir.net.vsources, cir.vs_p, cir.vs_n)
        d.is_ = sources(cir.net.isources, cir.is_p, cir.is_n)
        if len(cir.models) > 4:
            raise ValueError("at most 4 model groups")
        d.n_models = len(cir.models)
        for i, om in enumerate(cir.models):
            md = d.models[i]
            md.kind = _KIND[om.name]
            md.count = om.nodes.shape[0]
            md.nodes = _p(arr(om.nodes, np.int64), i64p)
            md.csr_pos = _p(arr(om.csr_pos, np.int64), i64p)
            md.res_idx = _p(arr(om.res_idx, np.int64), i64p)
            if om.name == "mosfet":
                par = np.stack([np.asarray(om.cache[k], dtype=np.float64) for k in _MOS_KEYS], axis=1)
            else:
                par = np.asarray(om.params["r" if om.name == "resistor" else "c"], dtype=np.float64)
            md.par = _p(arr(par, np.float64))
        self.col_order = arr(column_order(cir) if col_order is None else col_order, np.int64)
        d.col_order = _p(self.col_order, i64p)
        d.gmin = cir.gmin
        h = C.c_void_p()
        if lib.ct_create(C.byref(d), C.byref(h)) != 0:
            raise RuntimeError(lib.ct_last_error().decode())
        self.h = h
        self._keep = []  # the library copied every array

    def __del__(self):
        if getattr(self, "h", None) and _lib is not None:
            _lib.ct_destroy(self.h)
            self.h = None

    # -- single operations (tests) --
    def build_system(self, X, vsrc, isrc, Q_prev, c0, gshunt=0.0, c1=0.0, d1=0.0, dQdt_prev=None,
                     c2=0.0, Q_prev2=None):
        cir = self.cir
        z = np.zeros(cir.n_unk)
        f64 = lambda a: np.ascontiguousarray(a, dtype=np.float64)  # noqa: E731
        X, vsrc, isrc, Q_prev = f64(X), f64(vsrc if len(vsrc) else [0.0]), \
            f64(isrc if len(isrc) else [0.0]), f64(Q_prev)
        dQ = f64(z if dQdt_prev is None else dQdt_prev)
        Q2 = f64(z if Q_prev2 is None else Q_prev2)
        csr, f = np.zeros(cir.nse), np.zeros(cir.n_aug)
        Q, mrc, Ivs = np.zeros(cir.n_unk), np.zeros(cir.n_unk), np.zeros(max(cir.m, 1))
        _lib.ct_build_system(self.h, _p(X), _p(vsrc), _p(isrc), _p(Q_prev), c0, gshunt, c1, d1,
                             _p(dQ), c2, _p(Q2), _p(csr), _p(f), _p(Q), _p(Ivs), _p(mrc))
        return csr, f, Q, Ivs[:cir.m], mrc

    def lu_solve(self, csr_data, b, fresh=False):
        """x = A^{-1} b; returns (x, counts = [factorisations, refactorisations,
        nnz(L) off-diagonal, nnz(U) off-diagonal])."""
        x = np.zeros(self.cir.n_aug)
        counts = np.zeros(4, dtype=np.int64)
        rc = _lib.ct_lu_solve(self.h, _p(np.ascontiguousarray(csr_data, dtype=np.float64)),
                              _p(np.ascontiguousarray(b, dtype=np.float64)), _p(x),
                              1 if fresh else 0, _p(counts, i64p))
        if rc != 0:
            raise RuntimeError(_lib.ct_last_error().decode())
        return x, counts

    # -- transient loops --
    def _x_init(self):
        cir = self.cir
        X0 = np.zeros(cir.n_total + cir.m)
        X0[:cir.n_total] = _init_mid_rail(cir)
        return X0

    @staticmethod
    def _opts(opts: Options):
        o = _Options()
        o.reltol, o.abstol, o.vntol = opts.reltol, opts.abstol, opts.vntol
        o.tran_itl, o.nr_damping = int(opts.tran_itl), opts.nr_damping
        o.uic = 1 if opts.icmode == "uic" else 0
        return o

    def _out(self, rc, times, V, I, st, n):
        if rc != 0:
            raise RuntimeError(_lib.ct_last_error().decode() or f"ctran failed ({rc})")
        stats = {k: getattr(st, k) for k, _ in _Stats._fields_}
        stats["wall_time"] = st.wall
        stats["total_nr_iterations"] = st.total_nr
        stats["rejected_steps"] = st.rejected
        stats["accepted_steps"] = n
        acc = max(n - 1, 1)
        stats["time_per_step_ms"] = st.wall / acc * 1e3
        return TransientOut(times[:n].copy(), V[:n].copy(), I[:n, :max(self.cir.m, 1)].copy(), stats)

    def run_adaptive(self, t_stop, dt, cfg: AdaptiveConfig, opts: Options, max_steps=None,
                     threads: int = 0) -> TransientOut:
        cir = self.cir
        if max_steps is None:
            max_steps = int(t_stop / dt * 1.1) + 10
        a = _Adaptive()
        for k, _ in _Adaptive._fields_:
            if k == "method":
                a.method = _METHOD[method_of(cfg.integration_method)]
            else:
                setattr(a, k, getattr(cfg, k))
        mo = max(cir.m, 1)
        times = np.zeros(max_steps)
        V = np.zeros((max_steps, cir.n_total))
        I = np.zeros((max_steps, mo))
        st = _Stats()
        X0 = self._x_init()
        rc = _lib.ct_run_adaptive(self.h, t_stop, dt, C.byref(self._opts(opts)), C.byref(a),
                                  max_steps, _p(X0), threads, _p(times), _p(V), _p(I), C.byref(st))
        return self._out(rc, times, V, I, st, int(st.n_steps))

    def run_fixed(self, t_stop, dt, method, opts: Options, tran_fs=0.25, max_steps=None,
                  n_steps=None, abstol=1e-12, threads: int = 0) -> TransientOut:
        cir = self.cir
        if max_steps is None:
            max_steps = int(t_stop / dt * 1.1) + 10
        dt = dt * tran_fs if tran_fs != 1.0 else dt
        if n_steps is None:
            n_steps = max(min(int(t_stop / dt), max_steps - 1), 1)
        mo = max(cir.m, 1)
        times = np.zeros(n_steps + 1)
        V = np.zeros((n_steps + 1, cir.n_total))
        I = np.zeros((n_steps + 1, mo))
        st = _Stats()
        X0 = self._x_init()
        rc = _lib.ct_run_fixed(self.h, dt, _METHOD[method_of(method)], C.byref(self._opts(opts)),
                               n_steps, abstol, _p(X0), threads, _p(times), _p(V), _p(I),
                               C.byref(st))
        return self._out(rc, times, V, I, st, n_steps + 1)
