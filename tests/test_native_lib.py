"""The C-ABI library loads and exports every function include/vjx.h declares."""

import ctypes
import os
import re

from vajax_amd import _lib

HEADER = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "include",
                      "vjx.h")


def declared_functions():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(vjx_[a-z0-9_]+)\s*\(", text)))


def test_header_declares_functions():
    names = declared_functions()
    assert "vjx_lu_refactor" in names and "vjx_transient_run" in names
    assert len(names) >= 20


def test_library_exports_every_declared_symbol():
    lib = ctypes.CDLL(str(_lib.LIB_PATH))
    missing = [n for n in declared_functions() if not hasattr(lib, n)]
    assert not missing, missing


def test_binding_covers_header():
    assert set(declared_functions()) == set(_lib.EXPORTED)


def test_no_device_is_reported_cleanly():
    # never fails; 0 without a GPU
    assert _lib.lib.vjx_device_count() >= 0
    assert _lib.lib.vjx_version() >= 100


def test_pattern_validation_errors():
    import numpy as np

    h = ctypes.c_void_p()
    indptr = np.array([0, 2, 3], np.int32)
    indices = np.array([1, 0, 1], np.int32)  # row 0 not sorted
    rc = _lib.lib.vjx_pattern_create(2, 3, _lib.ptr(indptr), _lib.ptr(indices), ctypes.byref(h))
    assert rc == -1
    assert "sorted" in _lib.last_error()
    indices = np.array([0, 1, 5], np.int32)
    rc = _lib.lib.vjx_pattern_create(2, 3, _lib.ptr(indptr), _lib.ptr(indices), ctypes.byref(h))
    assert rc == -1 and "range" in _lib.last_error()


def _desc_case(mutate):
    """A valid descriptor of a small circuit, mutated; vjx_circuit_create must
    reject it before any device work (so this runs without a GPU)."""
    import numpy as np

    from vajax_amd import generators as G
    from vajax_amd.engine import circuit_desc
    from vajax_amd.topology import build_topology

    net = G.gate_dag(30, seed=2)
    topo = build_topology(net)
    d, keep = circuit_desc(topo, net)
    extra = mutate(d, topo, np)
    h = ctypes.c_void_p()
    rc = _lib.lib.vjx_circuit_create(ctypes.byref(d), ctypes.byref(h))
    del keep, extra
    return rc, _lib.last_error()


def _arr(np, a, dt):
    a = np.ascontiguousarray(a, dtype=dt)
    return a, _lib.ptr(a, ctypes.c_double if dt == np.float64 else ctypes.c_int32)


def test_circuit_desc_validation():
    def set_array(d, np, field, values, dt):
        a, p = _arr(np, values, dt)
        setattr(d, field, p)
        return a

    def seg_copy(d, topo, np, field, n, fn):
        src = getattr(d, field)
        vals = np.array([src[i] for i in range(n + 1)], dtype=np.int32)
        fn(vals)
        a, p = _arr(np, vals, np.int32)
        setattr(d, field, p)
        return a

    cases = {
        # jac_seg not starting at 0 / decreasing / ending elsewhere
        "jac_seg": lambda d, t, np: seg_copy(d, t, np, "jac_seg", t.nse,
                                              lambda v: v.__setitem__(0, 1)),
        "jac_seg_dec": lambda d, t, np: seg_copy(
            d, t, np, "jac_seg", t.nse,
            lambda v: v.__setitem__(2, v[3] + 1) if t.nse > 3 else None),
        "res_seg_end": lambda d, t, np: seg_copy(d, t, np, "res_seg", t.n_unknowns,
                                                 lambda v: v.__setitem__(-1, v[-1] + 5)),
        "br_ptr_start": lambda d, t, np: seg_copy(d, t, np, "br_ptr", t.n_unknowns,
                                                  lambda v: v.__setitem__(0, 1)),
        "br_ptr_dec": lambda d, t, np: seg_copy(d, t, np, "br_ptr", t.n_unknowns,
                                                lambda v: v.__setitem__(-1, -1)),
        "null_slot_diag": lambda d, t, np: setattr(d, "slot_diag", None),
        "null_jac_seg": lambda d, t, np: setattr(d, "jac_seg", None),
        "null_model": lambda d, t, np: setattr(d.models[0], "jac_dst", None),
        "bad_wave": lambda d, t, np: set_array(d, np, "vs_wave", np.full(t.n_vsources, 9),
                                               np.int32),
        "model_id": lambda d, t, np: setattr(d.models[0], "model_id", 17),
        "negative_counts": lambda d, t, np: setattr(d, "n_jac_contrib", -1),
    }
    for name, mutate in cases.items():
        holder = []

        def mut(d, t, np, mutate=mutate):
            r = mutate(d, t, np)
            holder.append(r)
            return r

        rc, err = _desc_case(mut)
        assert rc == _lib.VJX_ERR_ARG, (name, rc, err)
        assert err.startswith("vjx_circuit_create"), (name, err)


def _header_struct_fields(name):
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    m = re.search(r"typedef struct\s*\{([^{}]*)\}\s*" + name + r"\s*;", text)
    assert m, name
    fields = []
    for decl in m.group(1).split(";"):
        decl = decl.strip()
        if not decl:
            continue
        ty_and_names = re.sub(r"\[[^\]]*\]", "", decl)
        head, *rest = [x.strip() for x in ty_and_names.split(",")]
        fields.append(re.findall(r"(\w+)\s*$", head)[0])
        fields.extend(re.findall(r"(\w+)\s*$", r)[0] for r in rest)
    return fields


def test_ctypes_structs_match_header():
    pairs = {"vjx_circuit_desc": _lib.CircuitDesc, "vjx_model_desc": _lib.ModelDesc,
             "vjx_tran_config": _lib.TranConfig, "vjx_tran_result": _lib.TranResult,
             "vjx_lu_info": _lib.LuInfo}
    for cname, py in pairs.items():
        assert [f[0] for f in py._fields_] == _header_struct_fields(cname), cname
