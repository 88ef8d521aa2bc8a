"""GPU parity of the source waveforms (vajax/analysis/sources.py:25-294):
device evaluation (vjx_eval_sources) against the oracle at many time points,
and adaptive transients of RC sections driven by every waveform type against
the oracle (waveforms within 1e-9 V, the same accepted time grid)."""

import numpy as np
import pytest

from oracle import sources as osrc
from vajax_amd.engine import DeviceCircuit
from vajax_amd.netlist import NetlistBuilder
from vajax_amd.topology import build_topology

from .helpers import run_gpu, run_oracle

pytestmark = pytest.mark.gpu

PWL = [0.0, 0.0, 0.2e-9, 1.0, 0.5e-9, 1.0, 0.5e-9, 0.3, 1.1e-9, -0.4, 1.6e-9, 0.0]


def _all_waveforms(with_loads):
    b = NetlistBuilder("waves")
    srcs = [
        ("dc", dict(dc=0.7)),
        ("pulse", dict(val0=0.1, val1=0.9, rise=30e-12, fall=50e-12, width=0.3e-9,
                       period=1.1e-9, delay=0.2e-9)),
        ("sine", dict(sinedc=0.2, ampl=0.5, freq=1.3e9, phase=0.4)),
        ("pwl", dict(wave=PWL, scale=0.8, offset=0.05, stretch=1.2, pwlperiod=2.1e-9)),
        ("exp", dict(val0=0.0, val1=1.0, delay=0.1e-9, td2=0.6e-9, tau1=0.15e-9, tau2=0.25e-9)),
        ("am", dict(sinedc=0.1, ampl=0.4, freq=2e9, phase=30.0, delay=0.15e-9, modfreq=3e8,
                    modphase=10.0, modindex=0.6)),
        ("fm", dict(sinedc=-0.1, ampl=0.6, freq=1.5e9, phase=45.0, delay=0.05e-9, modfreq=4e8,
                    modphase=0.0, modindex=1.2)),
    ]
    for i, (kind, p) in enumerate(srcs):
        b.source("vsource", f"v{i}", f"in{i}", "0", type=kind, **p)
        if with_loads:
            b.device("resistor", f"ra{i}", [f"in{i}", f"m{i}"], r=1e3 + 100 * i)
            b.device("capacitor", f"ca{i}", [f"m{i}", "0"], c=0.2e-12)
            b.device("resistor", f"rb{i}", [f"m{i}", f"o{i}"], r=2e3)
            b.device("capacitor", f"cb{i}", [f"o{i}", "0"], c=0.1e-12)
    # a pwl current source into a resistor
    b.source("isource", "i0", "0", "ci", type="pwl", wave=[0.0, 0.0, 0.4e-9, 1e-3, 1.0e-9, -5e-4])
    b.device("resistor", "rci", ["ci", "0"], r=500.0)
    if with_loads:
        b.device("capacitor", "cci", ["ci", "0"], c=0.3e-12)
    return b.build(options={"tran_method": "gear2", "icmode": "op"})


def test_device_source_values_match_oracle():
    net = _all_waveforms(False)
    dc = DeviceCircuit(build_topology(net), net)
    for t in np.concatenate([[0.0], np.linspace(0.0, 5e-9, 257)[1:], [0.5e-9, 0.6e-9, 2.52e-9]]):
        v, i = dc.source_values(float(t))
        ref_v = osrc.source_values(net.vsources, float(t))
        ref_i = osrc.source_values(net.isources, float(t))
        np.testing.assert_allclose(v.cpu().numpy(), ref_v, rtol=1e-13, atol=1e-15,
                                   err_msg=f"t={t}")
        np.testing.assert_allclose(i.cpu().numpy(), ref_i, rtol=1e-13, atol=1e-18)


def test_transient_driven_by_every_waveform():
    net = _all_waveforms(True)
    _, ref = run_oracle(net, 3e-9, 1e-12)
    _, res = run_gpu(net, 3e-9, 1e-12, use_graph=True)
    assert len(res.times) == len(ref.times)
    np.testing.assert_allclose(res.times, ref.times, rtol=1e-9, atol=0)
    err = float(np.max(np.abs(res.stats["V_out"] - ref.V)))
    assert err <= 1e-9, err
    assert res.stats["rejected_steps"] == ref.stats["rejected_steps"]
