"""vajax_amd: MI355X-native transient core (device eval, stamping, sparse LU,
Newton / BDF2 / LTE loop) behind a vajax-compatible API.

Pure-Python modules (netlist, generators, topology, models) import without the
native library; everything that computes (engine, sparse) loads ``libvjx.so``
on first use and raises if it is missing.
"""

__all__ = ["CircuitEngine", "TransientResult", "AdaptiveConfig", "SimulationOptions",
           "SparseLU", "solve"]


def __getattr__(name):
    if name in ("CircuitEngine", "TransientResult", "AdaptiveConfig", "SimulationOptions"):
        from . import engine

        return getattr(engine, name)
    if name in ("SparseLU", "solve"):
        from . import sparse

        return getattr(sparse, name)
    raise AttributeError(name)
