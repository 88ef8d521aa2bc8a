"""Closed forms of the oracle's source waveforms (vajax/analysis/sources.py:25-294)
and the C-ABI packing of their parameters."""

from __future__ import annotations

import math

import numpy as np
import pytest

from oracle import sources as osrc
from vajax_amd.netlist import (SRC_NPAR, WAVE_AM, WAVE_EXP, WAVE_FM, WAVE_PULSE, WAVE_PWL,
                               WAVE_SINE, NetlistBuilder, pack_sources)


def _one(kind, **p):
    b = NetlistBuilder("s")
    b.source("vsource", "v1", "a", "0", type=kind, **p)
    return b.build().vsources


def test_sine_closed_form():
    src = _one("sine", sinedc=0.5, ampl=2.0, freq=1e9, phase=0.3)
    assert osrc.source_value(src, 0, 0.0) == pytest.approx(0.5 + 2.0 * math.sin(0.3), rel=1e-15)
    src = _one("sine", sinedc=0.5, ampl=2.0, freq=1e9, phase=0.0)
    assert osrc.source_value(src, 0, 0.25e-9) == pytest.approx(2.5, rel=1e-14)
    assert osrc.source_value(src, 0, 0.75e-9) == pytest.approx(-1.5, rel=1e-14)
    # defaults: sinedc 0, ampl 1, freq 1 kHz
    src = _one("sine")
    assert osrc.source_value(src, 0, 0.25e-3) == pytest.approx(1.0, rel=1e-12)


def test_exp_closed_form():
    src = _one("exp", val0=0.2, val1=1.2, delay=1e-9, td2=5e-9, tau1=1e-9, tau2=2e-9)
    assert osrc.source_value(src, 0, 0.5e-9) == 0.2
    assert osrc.source_value(src, 0, 2e-9) == pytest.approx(0.2 + 1.0 * (1 - math.exp(-1)), rel=1e-14)
    risen = 0.2 + 1.0 * (1 - math.exp(-5))
    assert osrc.source_value(src, 0, 6e-9) == pytest.approx(risen, rel=1e-14)  # continuous
    assert osrc.source_value(src, 0, 8e-9) == pytest.approx(
        risen + (0.2 - risen) * (1 - math.exp(-1)), rel=1e-14)


def test_am_fm_closed_form():
    src = _one("am", sinedc=0.1, ampl=1.0, freq=1e9, phase=90.0, delay=1e-9, modfreq=1e8,
               modphase=90.0, modindex=0.5)
    # before the delay: dc + a sin(90 deg) (1 + mi sin(90 deg)); continuous at the delay
    assert osrc.source_value(src, 0, 0.0) == pytest.approx(0.1 + 1.5, rel=1e-15)
    assert osrc.source_value(src, 0, 1e-9) == pytest.approx(0.1 + 1.5, rel=1e-15)
    # a quarter carrier period later the carrier (phase 90 deg) is cos(pi/2) = 0
    assert osrc.source_value(src, 0, 1.25e-9) == pytest.approx(0.1, abs=1e-12)
    src = _one("fm", sinedc=0.0, ampl=1.0, freq=1e9, phase=0.0, delay=0.0, modfreq=1e8,
               modphase=0.0, modindex=0.7)
    t = 2.5e-9  # modulation phase 2*pi*0.25 -> sin = 1
    assert osrc.source_value(src, 0, t) == pytest.approx(
        math.sin(2 * math.pi * 1e9 * t + 0.7), abs=1e-12)
    # sinephase is accepted for phase (sources.py:268, 280)
    src = _one("fm", sinephase=30.0, delay=1.0, modindex=0.0)
    assert osrc.source_value(src, 0, 0.0) == pytest.approx(0.5, rel=1e-14)


def test_pwl_interpolation():
    wave = [0.0, 0.0, 1e-9, 1.0, 2e-9, 1.0, 2e-9, 0.0, 3e-9, 0.5]
    src = _one("pwl", wave=wave, scale=2.0, offset=0.1)
    v = lambda t: osrc.source_value(src, 0, t)  # noqa: E731
    assert v(-1e-9) == 0.1  # before the first point
    assert v(0.5e-9) == pytest.approx(0.1 + 1.0, rel=1e-15)
    assert v(1.5e-9) == pytest.approx(2.1, rel=1e-15)
    # at a repeated time the later value holds (right-sided segment search)
    assert v(2e-9) == pytest.approx(0.1, rel=1e-15)
    assert v(2.5e-9) == pytest.approx(0.1 + 0.5, rel=1e-15)
    assert v(9e-9) == pytest.approx(1.1, rel=1e-15)  # after the last point
    src = _one("pwl", wave=wave, pwlperiod=3e-9, stretch=1.0)
    assert osrc.source_value(src, 0, 3.5e-9) == pytest.approx(0.5, rel=1e-12)
    src = _one("pwl", wave=wave, stretch=2.0)
    assert osrc.source_value(src, 0, 1e-9) == pytest.approx(0.5, rel=1e-15)


def test_pwl_validation():
    b = NetlistBuilder("s")
    with pytest.raises(ValueError):
        b.source("vsource", "v1", "a", "0", type="pwl", wave=[0.0, 0.0, 1e-9])
    with pytest.raises(ValueError):
        b.source("vsource", "v1", "a", "0", type="pwl", wave=[1e-9, 0.0, 0.0, 1.0])
    with pytest.raises(ValueError):
        b.source("vsource", "v1", "a", "0", type="dc", wave=[0.0, 1.0])


def test_pack_sources_layout():
    b = NetlistBuilder("s")
    b.source("vsource", "v0", "a", "0", type="dc", dc=0.7)
    b.source("vsource", "v1", "b", "0", type="pulse", val0=0.1, val1=0.9, rise=1e-11,
             fall=2e-11, width=1e-9, period=3e-9, delay=5e-10)
    b.source("vsource", "v2", "c", "0", type="sine", sinedc=0.2, ampl=0.3, freq=1e8, phase=0.4)
    b.source("vsource", "v3", "d", "0", type="pwl", wave=[0.0, 1.0, 1e-9, 2.0], scale=3.0,
             offset=-1.0, stretch=2.0, pwlperiod=5e-9)
    b.source("vsource", "v4", "e", "0", type="exp", val0=0.0, val1=1.0, delay=1e-9, td2=2e-9,
             tau1=3e-10, tau2=4e-10)
    b.source("vsource", "v5", "f", "0", type="am", sinedc=0.1, ampl=0.2, freq=1e9, phase=10.0,
             delay=1e-10, modfreq=1e7, modphase=20.0, modindex=0.3)
    b.source("vsource", "v6", "g", "0", type="fm")
    src = b.build().vsources
    wave, par, ptr, tt, vv = pack_sources(src)
    assert wave.tolist() == [0, WAVE_PULSE, WAVE_SINE, WAVE_PWL, WAVE_EXP, WAVE_AM, WAVE_FM]
    assert par.shape == (SRC_NPAR, 7)
    assert par[0, 0] == 0.7
    np.testing.assert_array_equal(par[1:8, 1], [0.1, 0.9, 1e-11, 2e-11, 1e-9, 3e-9, 5e-10])
    np.testing.assert_array_equal(par[1:5, 2], [0.2, 0.3, 1e8, 0.4])
    assert par[1, 3] == 5e-9
    np.testing.assert_array_equal(par[1:7, 4], [0.0, 1.0, 1e-9, 2e-9, 3e-10, 4e-10])
    np.testing.assert_array_equal(par[1:9, 5], [0.1, 0.2, 1e9, 10.0, 1e-10, 1e7, 20.0, 0.3])
    np.testing.assert_array_equal(par[1:9, 6], [0.0, 1.0, 1e3, 0.0, 0.0, 1e3, 0.0, 0.5])
    assert ptr.tolist() == [0, 0, 0, 0, 2, 2, 2, 2]
    np.testing.assert_array_equal(tt, [0.0, 2e-9])
    np.testing.assert_array_equal(vv, [2.0, 5.0])
