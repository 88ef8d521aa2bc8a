"""ctypes binding of ``libvjx.so`` (declared in include/vjx.h).

The library is required: importing this module raises if the shared object is
missing or lacks a symbol, so no code path can silently fall back to a CPU
implementation.
"""

from __future__ import annotations

import ctypes as C
import os
from pathlib import Path

LIB_PATH = Path(os.environ.get("VJX_LIB", Path(__file__).resolve().parent / "libvjx.so"))

VJX_OK = 0
VJX_WARN_PIVOT = 1
VJX_WARN_NOCONV = 2
VJX_ERR_ARG = -1
VJX_ERR_HIP = -2
VJX_ERR_SINGULAR = -3
VJX_ERR_STATE = -4
VJX_ERR_FILL = -5
VJX_PREC_F64 = 0
VJX_PREC_F32 = 1
VJX_METHOD = {"be": 0, "trap": 1, "gear2": 2}
MODEL_IDS = {"resistor": 0, "capacitor": 1, "mosfet": 2, "diode": 3, "ekv26": 4}
MAX_MODELS = 8

i32p = C.POINTER(C.c_int32)
f64p = C.POINTER(C.c_double)
vp = C.c_void_p


class LuInfo(C.Structure):
    _fields_ = [
        ("n", C.c_int32), ("nnz_a", C.c_int32), ("nnz_l", C.c_int64), ("nnz_u", C.c_int64),
        ("n_updates", C.c_int64), ("flops", C.c_double),
        ("fact_levels", C.c_int32), ("fwd_levels", C.c_int32), ("bwd_levels", C.c_int32),
        ("fact_launches", C.c_int32), ("fwd_launches", C.c_int32), ("bwd_launches", C.c_int32),
        ("n_dense_rows", C.c_int32), ("n_offdiag_pivots", C.c_int32),
        ("min_pivot_ratio", C.c_double),
        ("n_blocks", C.c_int32),
        ("n_block_patterns", C.c_int32),
        ("fwd_bands", C.c_int32), ("bwd_bands", C.c_int32),
        ("fwd_hub_rows", C.c_int32), ("bwd_hub_rows", C.c_int32),
        ("fact_task_updates", C.c_int64), ("fact_schur_updates", C.c_int64),
        ("fact_task_targets", C.c_int64), ("fact_schur_targets", C.c_int64),
        ("analyze_ms", C.c_double * 5),
        ("band_width", C.c_int32), ("solve_band_width", C.c_int32),
        ("tuned_solve_ms", C.c_double), ("tuned_factor_ms", C.c_double),
        ("lu_pipeline", C.c_int32),
        ("dense_k", C.c_int32), ("dense_m", C.c_int32), ("dense_flops", C.c_double),
        ("nd_leaf", C.c_int32),
    ]

    def as_dict(self):
        d = {k: getattr(self, k) for k, _ in self._fields_ if k != "analyze_ms"}
        d["analyze_ms"] = list(self.analyze_ms)
        return d


class ModelDesc(C.Structure):
    _fields_ = [
        ("model_id", C.c_int32), ("count", C.c_int32),
        ("v_node1", i32p), ("v_node2", i32p), ("jac_dst", i32p), ("res_dst", i32p),
        ("params", f64p),
    ]


class CircuitDesc(C.Structure):
    _fields_ = [
        ("n_total", C.c_int32), ("n_vsources", C.c_int32), ("n_isources", C.c_int32),
        ("nse", C.c_int32),
        ("indptr", i32p), ("indices", i32p), ("slot_diag", i32p), ("slot_b", f64p),
        ("vs_node_p", i32p), ("vs_node_n", i32p), ("vs_wave", i32p), ("vs_params", f64p),
        ("is_wave", i32p), ("is_params", f64p),
        ("vs_pwl_ptr", i32p), ("vs_pwl_t", f64p), ("vs_pwl_v", f64p),
        ("is_pwl_ptr", i32p), ("is_pwl_t", f64p), ("is_pwl_v", f64p),
        ("is_dst", i32p),
        ("n_jac_contrib", C.c_int32), ("n_res_contrib", C.c_int32),
        ("jac_seg", i32p), ("res_seg", i32p), ("br_ptr", i32p), ("br_ent", i32p),
        ("conv_mask", i32p),
        ("n_models", C.c_int32),
        ("models", ModelDesc * MAX_MODELS),
        ("n_noi", C.c_int32),
        ("noi_rows", i32p),
    ]


class EvalSplit(C.Structure):
    _fields_ = [("n_device_params", C.c_int32), ("n_device_cache", C.c_int32),
                ("voltage_positions", i32p), ("cache_src", i32p), ("mfactor_index", C.c_int32)]


class TranConfig(C.Structure):
    _fields_ = [
        ("t_stop", C.c_double), ("dt", C.c_double),
        ("max_steps", C.c_int32), ("method", C.c_int32), ("fixed_step", C.c_int32),
        ("uic", C.c_int32), ("precision", C.c_int32), ("refine_steps", C.c_int32),
        ("tran_itl", C.c_int32), ("use_graph", C.c_int32),
        ("reltol", C.c_double), ("abstol", C.c_double), ("vntol", C.c_double),
        ("gmin", C.c_double), ("nr_damping", C.c_double), ("lte_ratio", C.c_double),
        ("redo_factor", C.c_double), ("min_dt", C.c_double), ("max_dt", C.c_double),
        ("tran_minpts", C.c_int32), ("warmup_steps", C.c_int32), ("max_order", C.c_int32),
        ("grow_factor", C.c_double), ("gshunt_init", C.c_double), ("gshunt_target", C.c_double),
        ("gshunt_steps", C.c_int32),
        ("tran_fs", C.c_double), ("tran_ft", C.c_double),
        ("max_consecutive_rejects", C.c_int32),
        ("v_init_host", f64p),
        ("autotune", C.c_int32),
        ("repivot_growth", C.c_double),
        ("output_chunk", C.c_int32),
        ("replay_times_host", f64p), ("n_replay", C.c_int32),
        ("window_from_step", C.c_int32),
    ]


class TranResult(C.Structure):
    _fields_ = [
        ("times", f64p), ("V_out", f64p), ("I_out", f64p),
        ("n_steps", C.c_int32), ("rejected_steps", C.c_int32),
        ("total_nr_iterations", C.c_int64), ("total_evals", C.c_int64),
        ("dc_converged", C.c_int32), ("dc_iterations", C.c_int32),
        ("min_dt_used", C.c_double), ("max_dt_used", C.c_double),
        ("wall_time", C.c_double), ("dc_time", C.c_double),
        ("repivots", C.c_int32),
        ("window_steps", C.c_int32), ("window_rejected", C.c_int32),
        ("window_nr_iterations", C.c_int64), ("window_time", C.c_double),
    ]


class NewtonArgs(C.Structure):
    _fields_ = [
        ("X_init", vp), ("vsource_vals", vp), ("isource_vals", vp), ("Q_prev", vp),
        ("dQdt_prev", vp), ("Q_prev2", vp), ("limit_state_in", vp), ("res_tol_floor", vp),
        ("c0", C.c_double), ("c1", C.c_double), ("d1", C.c_double), ("c2", C.c_double),
        ("gmin", C.c_double), ("gshunt", C.c_double),
    ]


class NewtonResult(C.Structure):
    _fields_ = [
        ("X", vp), ("Q", vp), ("dQdt", vp), ("I_vsource", vp), ("limit_state", vp),
        ("max_res_contrib", vp), ("iterations", C.c_int32), ("converged", C.c_int32),
        ("max_f", C.c_double), ("repivots", C.c_int32),
    ]


class PhaseTimes(C.Structure):
    _fields_ = [("build_ms", C.c_double), ("lu_ms", C.c_double), ("solve_ms", C.c_double),
                ("n_iterations", C.c_int64), ("kernel_ms", C.c_double * 9),
                ("kernel_launches", C.c_int64 * 9)]

    KERNEL_CLASSES = ("eval", "stamp", "lu_gather", "lu_tasks", "lu_schur", "lu_scale",
                      "solve_tasks", "solve_hub", "lu_dense")

    def as_dict(self):
        d = {k: getattr(self, k) for k in ("build_ms", "lu_ms", "solve_ms", "n_iterations")}
        d["kernels"] = {name: {"ms": self.kernel_ms[i], "launches": self.kernel_launches[i]}
                        for i, name in enumerate(self.KERNEL_CLASSES)}
        return d


_PROTOS = {
    "vjx_version": (C.c_int, []),
    "vjx_last_error": (C.c_char_p, []),
    "vjx_device_count": (C.c_int, []),
    "vjx_pattern_create": (C.c_int, [C.c_int32, C.c_int32, i32p, i32p, C.POINTER(vp)]),
    "vjx_pattern_destroy": (None, [vp]),
    "vjx_lu_analyze": (C.c_int, [vp, f64p]),
    "vjx_lu_refactor": (C.c_int, [vp, vp, C.c_int32]),
    "vjx_lu_solve": (C.c_int, [vp, vp, vp, vp, C.c_int32]),
    "vjx_lu_solve_noalloc": (C.c_int, [vp, vp, vp, vp, C.c_int32]),
    "vjx_spmv": (C.c_int, [vp, vp, vp, vp]),
    "vjx_solve_host": (C.c_int, [C.c_int32, C.c_int32, i32p, i32p, f64p, f64p, f64p]),
    "vjx_lu_solve_transpose": (C.c_int, [vp, vp, vp, vp, C.c_int32, C.c_int32]),
    "vjx_solve_transpose_host": (C.c_int, [C.c_int32, C.c_int32, i32p, i32p, f64p, f64p, f64p]),
    "vjx_dot_host": (C.c_int, [C.c_int32, C.c_int32, i32p, i32p, f64p, f64p, f64p]),
    "vjx_clear_cache": (None, []),
    "vjx_lu_get_info": (C.c_int, [vp, C.POINTER(LuInfo)]),
    "vjx_lu_get_perm": (C.c_int, [vp, i32p, i32p]),
    "vjx_lu_get_pattern": (C.c_int, [vp, i32p, i32p]),
    "vjx_lu_check_schedules": (C.c_int, [vp]),
    "vjx_pattern_set_band_width": (C.c_int, [vp, C.c_int32]),
    "vjx_pattern_set_ordering": (C.c_int, [vp, C.c_int32, C.c_int32]),
    "vjx_pattern_set_fill_limit": (C.c_int, [vp, C.c_int64]),
    "vjx_pattern_set_dense": (C.c_int, [vp, C.c_int32, C.c_double]),
    "vjx_lu_autotune": (C.c_int, [vp, vp]),
    "vjx_lu_refactor_cpu": (C.c_int, [vp, f64p, C.c_int32, f64p]),
    "vjx_lu_solve_cpu": (C.c_int, [vp, f64p, C.c_int32, f64p, f64p]),
    "vjx_lu_download": (C.c_int, [vp, f64p]),
    "vjx_circuit_create": (C.c_int, [C.POINTER(CircuitDesc), C.POINTER(vp)]),
    "vjx_circuit_destroy": (None, [vp]),
    "vjx_circuit_pattern": (vp, [vp]),
    "vjx_circuit_limit_states": (C.c_int, [vp]),
    "vjx_circuit_eval_block": (C.c_int, [vp, C.c_int32]),
    "vjx_build_system": (C.c_int, [vp, vp, vp, vp, vp, vp, vp, C.c_double, C.c_double,
                                   C.c_double, C.c_double, C.c_double, C.c_double, vp, vp, vp,
                                   vp, vp, vp, vp]),
    "vjx_dense_lu": (C.c_int, [C.c_int32, vp]),
    "vjx_dense_gemm_sub": (C.c_int, [C.c_int32, C.c_int32, C.c_int32, vp, C.c_int32, vp, C.c_int32,
                                     vp, C.c_int32]),
    "vjx_model_eval_split": (C.c_int, [C.c_int32, C.c_int32, C.POINTER(EvalSplit)] + [vp] * 15),
    "vjx_model_eval": (C.c_int, [C.c_int32, C.c_int32, vp, vp, vp, vp, vp, vp, vp, vp, vp, vp]),
    "vjx_eval_stamp": (C.c_int, [vp, C.c_int32, vp, vp, vp, C.c_double, C.c_double, vp, vp, vp, vp,
                                 vp, vp]),
    "vjx_limit_eval": (C.c_int, [C.c_int32, C.c_int32, vp, vp, C.c_double, C.c_double, vp]),
    "vjx_eval_sources": (C.c_int, [vp, C.c_double, vp, vp]),
    "vjx_transient_run": (C.c_int, [vp, C.POINTER(TranConfig), C.POINTER(TranResult)]),
    "vjx_newton_solve": (C.c_int, [vp, C.POINTER(TranConfig), C.POINTER(NewtonArgs),
                                   C.POINTER(NewtonResult)]),
    "vjx_profile_enable": (C.c_int, [C.c_int32]),
    "vjx_profile_reset": (C.c_int, []),
    "vjx_profile_read": (C.c_int, [C.POINTER(PhaseTimes)]),
}

EXPORTED = tuple(_PROTOS)


def _load():
    if not LIB_PATH.exists():
        raise ImportError(
            f"vajax_amd native library not found at {LIB_PATH}; build it with "
            "`python -m vajax_amd.build` (there is no CPU fallback)")
    # One HIP runtime per process: torch ships its own libamdhip64 with the
    # same soname as /opt/rocm's.  Importing torch first makes libvjx bind to
    # that copy, so device pointers and streams from torch tensors are valid
    # here; loading libvjx first would pull in a second runtime, which then
    # reports no device.
    try:
        import torch  # noqa: F401
    except ImportError:
        pass
    lib = C.CDLL(str(LIB_PATH))
    for name, (res, args) in _PROTOS.items():
        fn = getattr(lib, name)  # AttributeError if the symbol is missing
        fn.restype = res
        fn.argtypes = args
    return lib


lib = _load()


class VjxError(RuntimeError):
    pass


def last_error() -> str:
    msg = lib.vjx_last_error()
    return msg.decode() if msg else ""


def check(rc: int, what: str) -> int:
    if rc < 0:
        raise VjxError(f"{what} failed ({rc}): {last_error()}")
    return rc


def ptr(a, ctype=C.c_int32):
    """ctypes pointer to a contiguous numpy array (None -> NULL)."""
    if a is None:
        return None
    return a.ctypes.data_as(C.POINTER(ctype))
