"""The openvaf_jax eval_fn contract in full through the C ABI
(vjx_model_eval_split; openvaf_jax/codegen/function_builder.py:967-996,
vmapped with in_axes (None, 0, None, 0, None, 0) as in
vajax/analysis/openvaf_models.py:996-997): shared / per-device parameter and
cache arrays, simparams with $mfactor, limit state in; all nine outputs
checked against the oracle's model functions."""

import ctypes as C

import numpy as np
import pytest

from oracle import models as omod
from vajax_amd import _lib as L
from vajax_amd.models import param_block

pytestmark = pytest.mark.gpu

N = 3001
SIMPARAMS = np.array([2.0, 2.0, 1e-12])  # [$analysis_type, $mfactor, gmin]


def _case(model, rng):
    if model == "mosfet":
        typ = np.where(rng.random(N) < 0.5, 1.0, -1.0)
        p = {"type": typ, "w": rng.uniform(0.2e-6, 2e-6, N), "l": rng.uniform(45e-9, 90e-9, N),
             "vto": np.full(N, 0.35), "gamma": np.full(N, 0.4), "phi": rng.uniform(0.5, 0.9, N),
             "kp": rng.uniform(5e-5, 3e-4, N), "theta": np.full(N, 0.2),
             "cox": np.full(N, 0.02), "cov": np.full(N, 2e-10), "cj": np.full(N, 5e-10)}
        vb = rng.uniform(-1.5, 1.5, (3, N))
        ref = omod.eval_mosfet(vb[0], vb[1], vb[2], omod.mosfet_cache(p))
        ref = tuple(ref) + (None, None, None)
    elif model == "ekv26":
        from oracle import ekv26 as E
        from vajax_amd.netlist import EKV26_PARAMS

        typ = np.where(rng.random(N) < 0.5, 1.0, -1.0)
        p = {k: np.full(N, float(v)) for k, v in EKV26_PARAMS.items()}
        p.update(type=typ, w=rng.uniform(0.2e-6, 2e-6, N), l=rng.uniform(45e-9, 1e-6, N),
                 vto=typ * 0.35, cox=np.full(N, 0.015), xj=np.full(N, 30e-9),
                 hdif=np.full(N, 0.1e-6), xd_cj=np.full(N, 1e-3))
        vb = rng.uniform(-1.2, 1.2, (3, N))
        ref = tuple(E.eval_ekv26(vb[0], vb[1], vb[2], E.cache(p))) + (None, None, None)
    elif model == "diode":
        p = {"is": rng.uniform(1e-15, 1e-13, N), "n": np.full(N, 1.2), "cj": rng.uniform(1e-15, 1e-14, N)}
        vb = rng.uniform(-1.0, 1.0, (1, N))
        vold = rng.uniform(-0.5, 0.8, N)
        ref = omod.eval_diode(vb[0], vold, omod.diode_cache(p))
        return p, vb, vold, ref
    else:
        key = "r" if model == "resistor" else "c"
        p = {key: rng.uniform(1e2, 1e5, N) if key == "r" else rng.uniform(1e-15, 1e-12, N)}
        vb = rng.uniform(-2, 2, (1, N))
        ref = (omod.eval_resistor if key == "r" else omod.eval_capacitor)(vb[0], p[key])
        ref = tuple(ref) + (None, None, None)
    return p, vb, None, ref


@pytest.mark.parametrize("model", ["resistor", "capacitor", "mosfet", "diode", "ekv26"])
def test_split_eval_nine_outputs(model):
    import torch

    rng = np.random.default_rng(21)
    p, vb, vold, ref = _case(model, rng)
    cache = param_block(model, p)  # [n_par][N]
    n_par, n_v = cache.shape[0], vb.shape[0]
    # columns constant over the devices go to shared_cache, the others to
    # device_cache in a shuffled order
    shared = [j for j in range(n_par) if np.all(cache[j] == cache[j, 0])]
    varying = [j for j in range(n_par) if j not in shared]
    rng.shuffle(varying)
    cache_src = np.zeros(n_par, np.int32)
    for k, j in enumerate(shared):
        cache_src[j] = k
    for k, j in enumerate(varying):
        cache_src[j] = ~k
    shared_cache = np.array([cache[j, 0] for j in shared] or [0.0])
    device_cache = np.stack([cache[j] for j in varying], axis=1) if varying else np.zeros((N, 1))
    # device_params: two unrelated parameter columns, then the voltages in
    # reverse order
    n_dp = 2 + n_v
    device_params = np.zeros((N, n_dp))
    device_params[:, 0] = rng.random(N)
    device_params[:, 1] = rng.random(N)
    vpos = np.array([2 + (n_v - 1 - k) for k in range(n_v)], np.int32)
    for k in range(n_v):
        device_params[:, vpos[k]] = vb[k]
    sp = L.EvalSplit()
    sp.n_device_params = n_dp
    sp.n_device_cache = device_cache.shape[1]
    sp.voltage_positions = L.ptr(vpos)
    sp.cache_src = L.ptr(cache_src)
    sp.mfactor_index = 1
    dev = lambda a: torch.as_tensor(np.ascontiguousarray(a, dtype=np.float64), device="cuda")  # noqa
    d_dp, d_sc, d_dc, d_sim = dev(device_params), dev(shared_cache), dev(device_cache), dev(SIMPARAMS)
    n_res, n_jac = ref[0].shape[1], ref[2].shape[1]
    n_lim = 1 if model == "diode" else 0
    d_li = dev(vold[:, None]) if model == "diode" else None
    outs = [torch.full((N, n_res), np.nan, dtype=torch.float64, device="cuda") for _ in range(2)]
    outs += [torch.full((N, n_jac), np.nan, dtype=torch.float64, device="cuda") for _ in range(2)]
    outs += [torch.full((N, n_res), np.nan, dtype=torch.float64, device="cuda") for _ in range(4)]
    lo = torch.full((N, max(n_lim, 1)), np.nan, dtype=torch.float64, device="cuda")
    ptr = lambda t: C.c_void_p(t.data_ptr()) if t is not None else None  # noqa
    rc = L.lib.vjx_model_eval_split(L.MODEL_IDS[model], N, C.byref(sp), None, ptr(d_dp),
                                    ptr(d_sc), ptr(d_dc), ptr(d_sim), ptr(d_li),
                                    *[ptr(o) for o in outs], ptr(lo) if n_lim else None)
    assert rc == 0, L.last_error()
    mf = SIMPARAMS[1]
    got = [o.cpu().numpy() for o in outs]
    names = ["res_resist", "res_react", "jac_resist", "jac_react", "lim_rhs_resist",
             "lim_rhs_react", "ss_resist", "ss_react"]
    for k in range(6):
        want = ref[k] * mf if ref[k] is not None else np.zeros_like(got[k])
        scale = np.max(np.abs(want), axis=1, keepdims=True) + 1e-300
        # EKV 2.6: device exp/log/sqrt against numpy's through the model's
        # cancellations (3e-14 of the row measured)
        tol = 1e-12 if model == "ekv26" else 1e-13
        assert np.all(np.abs(got[k] - want) <= tol * scale), (model, names[k])
    assert np.all(got[6] == 0.0) and np.all(got[7] == 0.0)
    if n_lim:
        np.testing.assert_allclose(lo.cpu().numpy(), ref[6], rtol=1e-14, atol=0)
