"""Writes tests/golden/tb_dp512x8_netlist.npz: the reference's SRAM
testbench netlist (vajax/benchmarks/data/tb_dp/tb_dp512x8_klu.sim) as this
build's front end flattens it (resistors, capacitors, dc and pwl sources;
the transistor wrappers hold only ``@if`` blocks, which the reference's
parser skips, vajax/netlist/parser.py:604-645), so the GPU tests can run it
where the reference is absent.

    python -m tests.golden.make_tb_dp
"""

import os

from vajax_amd import vacask

from . import netlist_npz

HERE = os.path.dirname(os.path.abspath(__file__))
SIM = "/root/reference/vajax/benchmarks/data/tb_dp/tb_dp512x8_klu.sim"
OUT = os.path.join(HERE, "tb_dp512x8_netlist.npz")


def main():
    net = vacask.load(SIM)
    netlist_npz.save(net, OUT)
    print(OUT, {k: g.count for k, g in net.groups.items()}, net.vsources.count)


if __name__ == "__main__":
    main()
