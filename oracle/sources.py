"""Independent-source waveforms (oracle).

``pulse`` follows vajax/analysis/sources.py:31-63 term by term:
t_in_period = (t - delay) mod period; rising = v0 + (v1-v0)*t_in/rise;
falling = v1 - (v1-v0)*(t_in - rise - width)/fall; value v0 before the
delay.  ``dc`` returns the dc parameter (sources.py:25-28).  Pinned by the
reference's known-answer tests (tests/test_sources.py:100-129) restated in
tests/test_oracle_units.py.
"""

from __future__ import annotations

import math

import numpy as np

from vajax_amd.netlist import WAVE_DC, WAVE_PULSE, SourceGroup


def pulse(t: float, v0, v1, rise, fall, width, period, delay):
    t_in = math.fmod(t - delay, period) if t >= delay else 0.0
    if t_in < 0.0:
        t_in += period
    if t < delay:
        return v0
    if t_in < rise:
        return v0 + (v1 - v0) * t_in / rise
    if t_in < rise + width:
        return v1
    if t_in < rise + width + fall:
        return v1 - (v1 - v0) * (t_in - rise - width) / fall
    return v0


def source_values(src: SourceGroup, t: float) -> np.ndarray:
    out = np.zeros(src.count, dtype=np.float64)
    p = src.params
    for i in range(src.count):
        w = int(src.wave[i])
        if w == WAVE_DC:
            out[i] = p["dc"][i]
        elif w == WAVE_PULSE:
            out[i] = pulse(t, p["val0"][i], p["val1"][i], p["rise"][i], p["fall"][i],
                           p["width"][i], p["period"][i], p["delay"][i])
        else:
            raise ValueError(f"unsupported waveform code {w}")
    return out
