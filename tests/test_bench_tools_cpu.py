"""bench.py's profile summarisers on synthetic rocprofv3 CSVs: kernel names
lose their return type, namespace and arguments; the trace summary covers
one timed window (between a pair of k_window_mark dispatches); the counter
summary scales FETCH_SIZE / WRITE_SIZE by the copy calibration."""

import csv
import json

import bench


def _write(path, fields, rows):
    with open(path, "w", newline="") as fh:
        w = csv.DictWriter(fh, fieldnames=fields)
        w.writeheader()
        for r in rows:
            w.writerow(r)


def test_kernel_name():
    raw = ("void vjx::(anonymous namespace)::k_lu_band_lds<double, 4, 3, 8>(int, int, "
           "vjx::(anonymous namespace)::BandProg, double*)")
    assert bench._kernel_name(raw) == "k_lu_band_lds<double, 4, 3, 8>"
    assert bench._kernel_name("__amd_rocclr_copyBuffer") == "__amd_rocclr_copyBuffer"
    assert bench._kernel_name("vjx::k_status_reset(vjx::DevStatus*)") == "k_status_reset"


def test_summarize_trace_picks_one_window(tmp_path, capsys):
    fields = ["Kernel_Name", "Start_Timestamp", "End_Timestamp"]
    rows, t = [], 0

    def add(name, dur):
        nonlocal t
        rows.append({"Kernel_Name": name, "Start_Timestamp": t, "End_Timestamp": t + dur})
        t += dur + 10

    mark = "void vjx::(anonymous namespace)::k_window_mark(int)"
    add("k_a(int)", 1000)            # before any window
    add(mark, 5)
    add("k_a(int)", 300)             # window 0: 2 x k_a, 1 x k_b
    add("k_a(int)", 500)
    add("void k_b<double>(int)", 200)
    add(mark, 5)
    add("k_b(int)", 7000)            # between windows
    add(mark, 5)
    add("k_c(int)", 50)              # window 1
    add(mark, 5)
    p = tmp_path / "trace.csv"
    _write(p, fields, rows)
    bench.summarize_trace(str(p), 0)
    out = capsys.readouterr().out
    assert "3 dispatches" in out and "1.000 ms of kernel time" not in out
    assert "| `k_a` | 2 | 0.001 | 0.40 |" in out
    assert "| `k_b<double>` | 1 |" in out and "k_c" not in out
    bench.summarize_trace(str(p), 1)
    out = capsys.readouterr().out
    assert "1 dispatches" in out and "| `k_c` | 1 |" in out


def test_summarize_pmc_scales_by_the_copy(tmp_path, capsys):
    fields = ["Dispatch_Id", "Kernel_Name", "Counter_Name", "Counter_Value"]
    k = "void vjx::(anonymous namespace)::k_lu_wide<double, true>(int)"
    _write(tmp_path / "f.csv", fields,
           [{"Dispatch_Id": i, "Kernel_Name": k, "Counter_Name": "FETCH_SIZE",
             "Counter_Value": 100.0 + 2 * i} for i in range(3)])
    _write(tmp_path / "w.csv", fields,
           [{"Dispatch_Id": i, "Kernel_Name": k, "Counter_Name": "WRITE_SIZE",
             "Counter_Value": 10.0} for i in range(3)])
    cp = "__amd_rocclr_copyBuffer"
    _write(tmp_path / "cf.csv", fields,
           [{"Dispatch_Id": i, "Kernel_Name": cp, "Counter_Name": "FETCH_SIZE",
             "Counter_Value": 524288.0} for i in range(5)])
    _write(tmp_path / "cw.csv", fields,
           [{"Dispatch_Id": i, "Kernel_Name": cp, "Counter_Name": "WRITE_SIZE",
             "Counter_Value": 1048576.0} for i in range(5)])
    out_json = tmp_path / "t.json"
    bench.summarize_pmc([str(tmp_path / n) for n in ("f.csv", "w.csv", "cf.csv", "cw.csv")],
                        str(out_json))
    d = json.loads(out_json.read_text())
    assert d["fetch_factor"] == 2.0 and d["write_factor"] == 1.0
    v = d["kernels"]["k_lu_wide<double, true>"]
    assert v["dispatches"] == 3 and v["fetch_kib"] == 102.0 and v["write_kib"] == 10.0
    assert v["traffic_bytes"] == (102.0 * 2.0 + 10.0) * 1024.0
    assert "k_lu_wide<double, true>" in capsys.readouterr().out
