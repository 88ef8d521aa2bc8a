"""A flattened :class:`~vajax_amd.netlist.Netlist` as a numpy ``.npz`` file
(data only: node names, terminal indices, parameter columns, source
waveforms and pwl points), so a netlist the front end loaded from the
reference's benchmark data can travel as a test fixture."""

from __future__ import annotations

import numpy as np

from vajax_amd.netlist import SOURCE_DEFAULTS, DeviceGroup, Netlist, SourceGroup


def save(net: Netlist, path: str) -> None:
    d = {"title": np.array(net.title), "nodes": np.array(net.nodes), "ground": np.array(net.ground),
         "groups": np.array(list(net.groups))}
    for name, g in net.groups.items():
        d[f"g.{name}.terminals"] = g.terminals
        for k, v in g.params.items():
            d[f"g.{name}.p.{k}"] = v
    for kind, s in (("vs", net.vsources), ("is", net.isources)):
        d[f"{kind}.names"] = np.array(s.names)
        d[f"{kind}.terminals"] = s.terminals
        d[f"{kind}.wave"] = s.wave
        for k in SOURCE_DEFAULTS:
            d[f"{kind}.p.{k}"] = s.params[k]
        lens = [len(s.pwl_points(i)) for i in range(s.count)]
        d[f"{kind}.pwl_len"] = np.array(lens, dtype=np.int64)
        d[f"{kind}.pwl"] = (np.concatenate([s.pwl_points(i) for i in range(s.count)])
                            if s.count else np.zeros(0))
    for k, v in net.analysis.items():
        d[f"analysis.{k}"] = np.array(float(v))
    for k, v in net.options.items():
        d[f"options.{k}"] = np.array(v)
    np.savez_compressed(path, **d)


def load(path: str) -> Netlist:
    z = np.load(path, allow_pickle=False)
    groups = {}
    for name in [str(x) for x in z["groups"]]:
        pre = f"g.{name}.p."
        params = {k[len(pre):]: z[k] for k in z.files if k.startswith(pre)}
        groups[name] = DeviceGroup(model=name, terminals=z[f"g.{name}.terminals"], params=params)

    def src(kind, label):
        names = [str(x) for x in z[f"{kind}.names"]]
        lens = z[f"{kind}.pwl_len"]
        flat = z[f"{kind}.pwl"]
        offs = np.concatenate([[0], np.cumsum(lens)])
        pwl = [flat[offs[i]:offs[i + 1]] for i in range(len(names))]
        wave = z[f"{kind}.wave"].astype(np.int32)
        return SourceGroup(kind=label, names=names, terminals=z[f"{kind}.terminals"], wave=wave,
                           params={k: z[f"{kind}.p.{k}"] for k in SOURCE_DEFAULTS},
                           pwl=pwl if np.any(lens > 0) else None)

    analysis = {k[9:]: float(z[k]) for k in z.files if k.startswith("analysis.")}
    options = {}
    for k in z.files:
        if k.startswith("options."):
            v = z[k]
            options[k[8:]] = str(v) if v.dtype.kind in "US" else float(v)
    return Netlist(title=str(z["title"]), nodes=[str(x) for x in z["nodes"]],
                   ground=str(z["ground"]), groups=groups, vsources=src("vs", "vsource"),
                   isources=src("is", "isource"), analysis=analysis, options=options)
