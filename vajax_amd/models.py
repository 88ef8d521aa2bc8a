"""Per-model parameter columns handed to the HIP eval kernels.

Each model's kernel reads a structure-of-arrays parameter block
``[n_par][count]`` (float64).  For the resistor and capacitor this is the
instance parameter itself (vajax/devices/models/resistor.va:19,
capacitor.va:17); for the MOSFET it is the per-instance init cache computed
once at setup, the analogue of the cache that the reference's
``prepare_static_inputs`` builds from the model's init function
(vajax/analysis/openvaf_models.py:889-921):

    sgn   = type                         (+1 NMOS, -1 PMOS)
    vgoff = (phi - vto) + gamma*sqrt(phi)   so that VG' = VG + vgoff
    phi, gamma, beta0 = kp*(w/l), theta
    cgs = cgd = cox*w*l/2 + cov*w,  cgb = 0,  cdb = csb = cj*w

and for the diode its init constants is, nvt = n*vt, the pnjlim critical
voltage vcrit = nvt*ln(nvt/(sqrt(2)*is)), and cj.
"""

from __future__ import annotations

from typing import Dict

import numpy as np

MOSFET_CACHE_COLUMNS = ("sgn", "vgoff", "phi", "gamma", "beta0", "theta",
                        "cgs", "cgd", "cgb", "cdb", "csb")


def mosfet_cache(p: Dict[str, np.ndarray]) -> np.ndarray:
    w, l = p["w"], p["l"]
    cgs = 0.5 * p["cox"] * w * l + p["cov"] * w
    cols = {
        "sgn": p["type"],
        "vgoff": (p["phi"] - p["vto"]) + p["gamma"] * np.sqrt(p["phi"]),
        "phi": p["phi"],
        "gamma": p["gamma"],
        "beta0": p["kp"] * (w / l),
        "theta": p["theta"],
        "cgs": cgs,
        "cgd": cgs,
        "cgb": np.zeros_like(w),
        "cdb": p["cj"] * w,
        "csb": p["cj"] * w,
    }
    return np.ascontiguousarray(np.stack([np.asarray(cols[c], dtype=np.float64)
                                          for c in MOSFET_CACHE_COLUMNS]))


# thermal voltage at the reference's default 27 C (k T / q, exact SI constants)
VT = 1.380649e-23 * 300.15 / 1.602176634e-19
DIODE_COLUMNS = ("is", "nvt", "vcrit", "cj")


def diode_cache(p: Dict[str, np.ndarray]) -> np.ndarray:
    nvt = p["n"] * VT
    cols = {"is": p["is"], "nvt": nvt, "vcrit": nvt * np.log(nvt / (np.sqrt(2.0) * p["is"])),
            "cj": p["cj"]}
    return np.ascontiguousarray(np.stack([np.asarray(cols[c], dtype=np.float64)
                                          for c in DIODE_COLUMNS]))


def param_block(model: str, params: Dict[str, np.ndarray]) -> np.ndarray:
    if model == "resistor":
        return np.ascontiguousarray(params["r"][None, :], dtype=np.float64)
    if model == "capacitor":
        return np.ascontiguousarray(params["c"][None, :], dtype=np.float64)
    if model == "mosfet":
        return mosfet_cache(params)
    if model == "diode":
        return diode_cache(params)
    raise ValueError(f"no eval kernel for model {model!r}")
