"""CPU oracle for the transient core -- test infrastructure only.

This package restates, in numpy/scipy, the reference's per-timestep transient
algorithm: device evaluation, CSR-direct stamping and residual assembly
(vajax/analysis/mna_builder.py, vajax/analysis/mna.py), the Newton-Raphson
loop (vajax/analysis/solver_factories.py), the predictor / LTE / step control
(vajax/analysis/transient/adaptive.py) and the adaptive and fixed-step time
loops (vajax/analysis/transient/full_mna.py).  Each function cites the
reference lines it follows.

It is the checker, never the product: only ``tests/``,
``__graft_entry__.smoke()`` and the ``cpu_baseline`` leg of ``bench.py`` may
import it.  The product path (``vajax_amd``) must not import anything from
here and fails loudly when its HIP library is missing.

Parity pins (what fixes each function independently of this restatement) are
listed in the docstrings of the test modules under ``tests/``; functions
without such a pin say "parity unpinned" in their docstring.
"""
