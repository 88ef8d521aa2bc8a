"""Raw-file and CSV writers: ngspice raw layout checked byte by byte against
an independent packing (struct, little-endian doubles, point-major), round
trips through the readers, and the CSV column order of the reference
(csv_writer.py:36-51: x first, then node names sorted)."""

import struct
from types import SimpleNamespace

import numpy as np
import pytest

from vajax_amd import io as vio


def _result(n=17, freq=False):
    rng = np.random.default_rng(0)
    volts = {"out": rng.standard_normal(n), "in": rng.standard_normal(n) * 1e-7,
             "n010": rng.uniform(-1, 1, n)}
    x = np.cumsum(rng.uniform(1e-12, 2e-12, n))
    if freq:
        return SimpleNamespace(frequencies=x * 1e21, voltages=volts)
    return SimpleNamespace(times=x, voltages=volts)


def test_binary_rawfile_layout(tmp_path):
    r = _result()
    p = tmp_path / "t.raw"
    vio.write_rawfile(r, p, binary=True)
    raw = p.read_bytes()
    head = ("Title: vajax_amd simulation\nDate: \nPlotname: Transient Analysis\nFlags: real\n"
            "No. Variables: 4\nNo. Points: 17\nVariables:\n\t0\ttime\ttime\n\t1\tout\tvoltage\n"
            "\t2\tin\tvoltage\n\t3\tn010\tvoltage\nBinary:\n").encode("ascii")
    assert raw[:len(head)] == head
    body = b"".join(struct.pack("<dddd", r.times[i], r.voltages["out"][i], r.voltages["in"][i],
                                r.voltages["n010"][i]) for i in range(17))
    assert raw[len(head):] == body


@pytest.mark.parametrize("binary", [True, False])
@pytest.mark.parametrize("freq", [False, True])
def test_rawfile_round_trip(tmp_path, binary, freq):
    r = _result(freq=freq)
    p = tmp_path / "t.raw"
    vio.write_rawfile(r, p, binary=binary)
    info, table = vio.read_rawfile(p)
    assert info["binary"] == binary and info["n_vars"] == 4 and info["n_points"] == 17
    assert info["plotname"] == ("AC Analysis" if freq else "Transient Analysis")
    assert [v[0] for v in info["variables"]] == ["frequency" if freq else "time", "out", "in",
                                                "n010"]
    x = r.frequencies if freq else r.times
    want = np.column_stack([x, r.voltages["out"], r.voltages["in"], r.voltages["n010"]])
    if binary:
        assert np.array_equal(table, want)
    else:
        assert np.allclose(table, want, rtol=1e-14, atol=0)
    assert vio.read_rawfile_header(p)["variables"] == info["variables"]


def test_csv_round_trip_and_order(tmp_path):
    r = _result()
    p = tmp_path / "t.csv"
    vio.write_csv(r, p, precision=12)
    first = p.read_text().splitlines()[0]
    assert first == "time,in,n010,out"
    back = vio.read_csv(p)
    assert np.allclose(back["times"], r.times, rtol=1e-12, atol=0)
    for k, v in r.voltages.items():
        assert np.allclose(back["voltages"][k], v, rtol=1e-12, atol=0)
