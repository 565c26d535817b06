"""fwe_combine_traces_noalloc: the trace combination staged through the
library's static device buffer gives the same bits as fwe_combine_traces,
which stages through per-call device allocations.

Two calls in a row on one device, the second smaller (fewer sources, fewer
wavefields, receivers and samples), so the second reads a buffer the first
left filled; then a request too large for the buffer is refused.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _batched_run(W, nx, ny, nz, nt, nrec_side, nb):
    wl = W.ortho_layered(nx=nx, ny=ny, nz=nz, L=5, nt=nt, f0=12.0, nrec_side=nrec_side, elementary=True,
                         m0=1e15)
    s = W.make_solver(wl)
    assert s.cfg.nbatch == 6
    s.run(wl.nt)
    return s, wl


def test_noalloc_combine_equals_allocating_combine(gpu):
    from elanfmt_amd import workloads as W

    rng = np.random.default_rng(7)
    # first call: 6 wavefields, 3x3x3 receivers, 80 samples, 40 sources
    s1, _ = _batched_run(W, 40, 36, 32, 80, 3, 6)
    print("noalloc: call 1, 40 sources over 6 wavefields", flush=True)
    c1 = rng.standard_normal((40, 6)).astype(np.float32)
    ref1 = s1.combine_traces(c1)
    got1 = s1.combine_traces(c1, noalloc=True)
    assert np.abs(ref1).max() > 0
    assert got1.shape == ref1.shape and np.array_equal(got1, ref1)

    # second call, smaller: 3 sources, 2x2x3 receivers, 50 samples
    s2, _ = _batched_run(W, 36, 32, 30, 50, 2, 6)
    print("noalloc: call 2, 3 sources, smaller traces", flush=True)
    c2 = rng.standard_normal((3, 6)).astype(np.float32)
    ref2 = s2.combine_traces(c2)
    got2 = s2.combine_traces(c2, noalloc=True)
    assert np.abs(ref2).max() > 0
    assert got2.shape == ref2.shape and np.array_equal(got2, ref2)
    # and the first context again, after the second call overwrote the buffer
    assert np.array_equal(s1.combine_traces(c1[:5], noalloc=True), ref1[:5])
    s1.close()
    s2.close()


def test_noalloc_combine_refuses_an_oversized_source(gpu):
    """One combined source larger than the static buffer is an argument
    error, raised before anything is copied."""
    from elanfmt_amd import fwe

    nt = 300_000
    s = fwe.Solver(8, 8, 8, 50.0, 50.0, 50.0, 1e-3, nt, mode=fwe.MODE_ORTHO, pml=fwe.PML_NONE,
                   free_surface=fwe.FS_NONE)
    # 30 receivers x 300000 samples = 9.0M floats > 8M floats (32 MiB)
    s.set_receivers([(fwe.REC_VZ, 2 + (r % 4), 2 + (r // 4) % 4, 3, 0) for r in range(30)])
    with pytest.raises(fwe.FweError) as ei:
        s.combine_traces(np.ones((1, 1), np.float32), noalloc=True)
    assert ei.value.status == fwe.ERR_ARG
    s.close()
