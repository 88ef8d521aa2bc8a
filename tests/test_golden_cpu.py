"""The committed golden fixtures (tests/golden/) are what the oracle computes
today: a guard against drift of the checker the GPU tests rely on."""

import os

import numpy as np
import pytest

from .golden import make_golden as MG
from .test_homotopy_cpu import oracle_nr_solve

HERE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def _load(name):
    with np.load(os.path.join(HERE, name + ".npz"), allow_pickle=False) as z:
        return {k: z[k] for k in z.files}


@pytest.mark.parametrize("name", sorted(MG.CASES))
def test_oracle_reproduces_golden_transients(name):
    want = _load(name)
    got = MG.transient_case(name)
    assert np.array_equal(got["times"], want["times"])
    assert got["total_nr_iterations"][0] == want["total_nr_iterations"][0]
    assert np.allclose(got["V"], want["V"], rtol=0, atol=1e-13)


def test_oracle_reproduces_golden_dc_homotopy():
    want = _load("dc_clamp_homotopy")
    got = MG.dc_clamp_case(oracle_nr_solve)
    assert bytes(got["method"]) == bytes(want["method"]) == b"chain_source_stepping"
    assert (got["iterations"][0], got["homotopy_steps"][0]) == \
        (want["iterations"][0], want["homotopy_steps"][0])
    assert np.allclose(got["X"], want["X"], rtol=0, atol=1e-13)
