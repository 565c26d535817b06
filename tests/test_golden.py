"""Committed golden fixtures (tests/golden/*.npz, written by
tests/golden/make_golden.py from the FP64 oracle).

CPU: the input generators still produce the inputs the fixtures were made
from (SHA-256 of every input array) and the oracle reproduces the traces.
GPU: the HIP path, through the C ABI, reproduces them within 1e-5 relative
L2 with either kernel schedule that applies."""
import os
import sys

import numpy as np
import pytest

from conftest import rel_l2

HERE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
sys.path.insert(0, HERE)
NAMES = ["iso32_fs", "ortho40_sep", "ortho36_w3d_fs"]


def _load(name):
    import make_golden

    wl = make_golden.cases()[name]
    fx = np.load(os.path.join(HERE, f"{name}.npz"))
    return wl, fx, make_golden


@pytest.mark.parametrize("name", NAMES)
def test_fixture_inputs_and_oracle(fwe, name):
    wl, fx, mg = _load(name)
    assert str(fx["digest"]) == mg.input_digest(wl)
    tr = mg.oracle_traces(wl)
    assert np.abs(fx["traces"]).max() > 0
    assert rel_l2(tr, fx["traces"]) < 1e-12


@pytest.mark.gpu
@pytest.mark.parametrize("name", NAMES)
def test_fixture_hip(gpu, name):
    from elanfmt_amd import workloads as W

    wl, fx, _ = _load(name)
    kinds = [gpu.KERNEL_TWO_PASS, gpu.KERNEL_FUSED]
    for kind in kinds:
        print(f"fixture {name} kernel {kind}", flush=True)
        s = W.make_solver(wl)
        s.set_kernel(kind)
        s.run(wl.nt)
        got = s.traces()
        s.close()
        assert rel_l2(got, fx["traces"]) < 1e-5, kind
