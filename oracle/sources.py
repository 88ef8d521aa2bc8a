"""Independent-source waveforms (oracle; test infrastructure only).

Each waveform follows vajax/analysis/sources.py term by term:

* ``dc`` -- the dc parameter (sources.py:25-28);
* ``pulse`` -- sources.py:31-63: t_in_period = (t - delay) mod period;
  rising = v0 + (v1-v0)*t_in/rise; falling = v1 - (v1-v0)*(t_in-rise-width)/fall;
  v0 before the delay;
* ``sine`` -- sources.py:66-78: sinedc + ampl*sin(2*pi*freq*t + phase)
  (phase added as given, no delay);
* ``pwl`` -- sources.py:81-106: times = wave[0::2]*stretch,
  values = wave[1::2]*scale + offset, t mod pwlperiod when pwlperiod > 0,
  then jnp.interp (clamped at both ends; the segment is found by a
  right-sided search, so at a repeated time the later value holds);
* ``exp`` -- sources.py:109-125: rise with tau1 from delay, fall with tau2
  from delay + td2 starting at the risen value;
* ``am`` / ``fm`` -- sources.py:128-186, phases in degrees, the value
  before the delay as stated there.

Pinned by the reference's pulse known answers (tests/test_sources.py:100-129,
restated in tests/test_oracle_units.py) and by closed forms of every other
waveform in tests/test_sources_cpu.py.
"""

from __future__ import annotations

import math

import numpy as np

from vajax_amd.netlist import (WAVE_AM, WAVE_DC, WAVE_EXP, WAVE_FM, WAVE_PULSE, WAVE_PWL,
                               WAVE_SINE, SourceGroup)


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


def sine(t, sinedc, ampl, freq, phase):
    return sinedc + ampl * math.sin(2.0 * math.pi * freq * t + phase)


def pwl_points(wave, offset, scale, stretch):
    w = np.asarray(wave, dtype=np.float64)
    return w[0::2] * stretch, w[1::2] * scale + offset


def interp(t, xp, fp):
    """jnp.interp at one point: clamped ends, right-sided segment search."""
    n = len(xp)
    if t < xp[0]:
        return float(fp[0])
    if t > xp[n - 1]:
        return float(fp[n - 1])
    if n == 1:
        return float(fp[0])
    i = int(np.searchsorted(xp, t, side="right"))
    i = min(max(i, 1), n - 1)
    dx = xp[i] - xp[i - 1]
    if abs(dx) <= np.spacing(np.finfo(np.float64).eps):
        return float(fp[i - 1])
    return float(fp[i - 1] + ((t - xp[i - 1]) / dx) * (fp[i] - fp[i - 1]))


def pwl(t, xp, fp, period):
    if period > 0:
        t = t % period
    return interp(t, xp, fp)


def exp_wave(t, v0, v1, delay, td2, tau1, tau2):
    t_fall = delay + td2
    if t < delay:
        return v0
    if t < t_fall:
        return v0 + (v1 - v0) * (1.0 - math.exp(-(t - delay) / tau1))
    risen = v0 + (v1 - v0) * (1.0 - math.exp(-td2 / tau1))
    return risen + (v0 - risen) * (1.0 - math.exp(-(t - t_fall) / tau2))


def am(t, dc, a, f, ph, d, mf, mph, mi):
    if t < d:
        return dc + a * math.sin(ph * math.pi / 180) * (1 + mi * math.sin(mph * math.pi / 180))
    carrier = math.sin(2 * math.pi * f * (t - d) + ph * math.pi / 180)
    modulation = 1 + mi * math.sin(2 * math.pi * mf * (t - d) + mph * math.pi / 180)
    return dc + a * carrier * modulation


def fm(t, dc, a, f, ph, d, mf, mph, mi):
    if t < d:
        return dc + a * math.sin(ph * math.pi / 180 + mi * math.sin(mph * math.pi / 180))
    carrier_phase = 2 * math.pi * f * (t - d) + ph * math.pi / 180
    mod = mi * math.sin(2 * math.pi * mf * (t - d) + mph * math.pi / 180)
    return dc + a * math.sin(carrier_phase + mod)


def source_value(src: SourceGroup, i: int, t: float) -> float:
    p = src.params
    w = int(src.wave[i])
    g = lambda k: float(p[k][i])  # noqa: E731
    if w == WAVE_DC:
        return g("dc")
    if w == WAVE_PULSE:
        return pulse(t, g("val0"), g("val1"), g("rise"), g("fall"), g("width"), g("period"),
                     g("delay"))
    if w == WAVE_SINE:
        return sine(t, g("sinedc"), g("ampl"), g("freq"), g("phase"))
    if w == WAVE_PWL:
        xp, fp = pwl_points(src.pwl_points(i), g("offset"), g("scale"), g("stretch"))
        return pwl(t, xp, fp, g("pwlperiod"))
    if w == WAVE_EXP:
        return exp_wave(t, g("val0"), g("val1"), g("delay"), g("td2"), g("tau1"), g("tau2"))
    if w in (WAVE_AM, WAVE_FM):
        fn = am if w == WAVE_AM else fm
        return fn(t, g("sinedc"), g("ampl"), g("freq"), g("phase"), g("delay"), g("modfreq"),
                  g("modphase"), g("modindex"))
    raise ValueError(f"unsupported waveform code {w}")


def source_values(src: SourceGroup, t: float) -> np.ndarray:
    out = np.zeros(src.count, dtype=np.float64)
    for i in range(src.count):
        out[i] = source_value(src, i, t)
    return out
