"""Result writers: ngspice raw files (binary and ASCII) and CSV
(vajax/io/rawfile_writer.py:18-182, vajax/io/csv_writer.py:13-94).

Raw files carry the x variable (time, or frequency for results that have
``frequencies``) then every entry of ``result.voltages`` in its order, as
real doubles; the binary form stores each point's values contiguously in
native byte order after a ``Binary:`` line, the ASCII form one value per line
after ``Values:``.  CSV has a header row of the x name and the node names in
sorted order, and values in scientific notation.
"""

from __future__ import annotations

import csv
from pathlib import Path
from typing import Any, Dict, List, Tuple, Union

import numpy as np

PathLike = Union[str, Path]
TITLE = "vajax_amd simulation"


def _x_axis(result: Any) -> Tuple[str, str, np.ndarray, str]:
    if hasattr(result, "frequencies"):
        return "frequency", "frequency", np.asarray(result.frequencies, np.float64), "AC Analysis"
    return "time", "time", np.asarray(result.times, np.float64), "Transient Analysis"


def _header(plot: str, variables: List[Tuple[str, str]], n_points: int) -> List[str]:
    lines = [f"Title: {TITLE}", "Date: ", f"Plotname: {plot}", "Flags: real",
             f"No. Variables: {len(variables)}", f"No. Points: {n_points}", "Variables:"]
    lines += [f"\t{i}\t{name}\t{kind}" for i, (name, kind) in enumerate(variables)]
    return lines


def write_rawfile(result: Any, output_path: PathLike, binary: bool = True) -> None:
    """Write ``result`` (``times`` or ``frequencies``, ``voltages`` dict) as an
    ngspice raw file (rawfile_writer.py:18-57)."""
    x_name, x_kind, x, plot = _x_axis(result)
    names = list(result.voltages.keys())
    variables = [(x_name, x_kind)] + [(n, "voltage") for n in names]
    table = np.empty((len(x), len(variables)), dtype=np.float64)
    table[:, 0] = x
    for j, n in enumerate(names):
        table[:, j + 1] = np.asarray(result.voltages[n], np.float64)
    head = _header(plot, variables, len(x))
    path = Path(output_path)
    if binary:
        with open(path, "wb") as f:
            f.write(("\n".join(head + ["Binary:"]) + "\n").encode("ascii"))
            f.write(np.ascontiguousarray(table).tobytes())
        return
    with open(path, "w") as f:
        f.write("\n".join(head + ["Values:"]) + "\n")
        for i in range(table.shape[0]):
            f.write(f" {i}\t{table[i, 0]:.15e}\n")
            f.write("".join(f"\t{v:.15e}\n" for v in table[i, 1:]))


def read_rawfile_header(path: PathLike) -> Dict[str, Any]:
    """Title, plotname, variable and point counts, variables, binary flag
    (rawfile_writer.py:134-182)."""
    return _read(path, with_data=False)[0]


def read_rawfile(path: PathLike) -> Tuple[Dict[str, Any], np.ndarray]:
    """Header and the [n_points, n_vars] value table of a real raw file."""
    return _read(path, with_data=True)


def _read(path: PathLike, with_data: bool):
    info: Dict[str, Any] = {"variables": [], "binary": False}
    raw = Path(path).read_bytes()
    pos = 0
    in_vars = False
    while pos < len(raw):
        end = raw.find(b"\n", pos)
        end = len(raw) if end < 0 else end
        text = raw[pos:end].decode("ascii").strip()
        pos = end + 1
        low = text.lower()
        if low.startswith("title:"):
            info["title"] = text[6:].strip()
        elif low.startswith("plotname:"):
            info["plotname"] = text[9:].strip()
        elif low.startswith("no. variables:"):
            info["n_vars"] = int(text[14:])
        elif low.startswith("no. points:"):
            info["n_points"] = int(text[11:])
        elif low == "variables:":
            in_vars = True
        elif low in ("binary:", "values:"):
            info["binary"] = low == "binary:"
            break
        elif in_vars and text:
            parts = text.split()
            if len(parts) >= 3:
                info["variables"].append((parts[1], parts[2]))
    if not with_data:
        return info, None
    nv, npnt = info["n_vars"], info["n_points"]
    if info["binary"]:
        data = np.frombuffer(raw[pos:pos + 8 * nv * npnt], dtype=np.float64)
    else:
        vals = []
        for line in raw[pos:].decode("ascii").splitlines():
            parts = line.split()
            if parts:
                vals.append(float(parts[-1]))
        data = np.array(vals, dtype=np.float64)
    return info, data.reshape(npnt, nv).copy()


def write_csv(result: Any, output_path: PathLike, precision: int = 9) -> None:
    """x column then the node voltages in sorted name order
    (csv_writer.py:13-59)."""
    x_name, _, x, _ = _x_axis(result)
    names = sorted(result.voltages.keys())
    cols = [np.asarray(result.voltages[n], np.float64) for n in names]
    fmt = f"{{:.{int(precision)}e}}"
    with open(Path(output_path), "w", newline="") as f:
        w = csv.writer(f)
        w.writerow([x_name] + names)
        for i in range(len(x)):
            w.writerow([fmt.format(x[i])] + [fmt.format(c[i]) for c in cols])


def read_csv(input_path: PathLike) -> Dict[str, Any]:
    """``times`` (or ``frequencies``) and ``voltages`` (csv_writer.py:62-94)."""
    with open(Path(input_path), newline="") as f:
        rows = list(csv.reader(f))
    header, body = rows[0], rows[1:]
    table = np.array([[float(v) for v in r] for r in body], dtype=np.float64).reshape(
        len(body), len(header))
    key = "frequencies" if header[0] == "frequency" else "times"
    return {key: table[:, 0].copy(),
            "voltages": {n: table[:, j + 1].copy() for j, n in enumerate(header[1:])}}
