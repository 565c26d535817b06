"""Parity of the HIP path (through the C ABI) with the FP64 oracle.

Tolerance: relative L2 over all receiver traces <= 1e-5 (FP32 GPU vs FP64
oracle), and the same bound on whole fields where they are compared.
"""
import numpy as np
import pytest

from conftest import rel_l2

pytestmark = pytest.mark.gpu

TOL = 1e-5


def _run_both(wl, steps=None):
    from elanfmt_amd import workloads as W
    from oracle import Oracle

    steps = wl.nt if steps is None else steps
    print(f"case {wl.name} mode={wl.mode} pml={wl.pml} fs={wl.free_surface} nb={wl.nbatch} steps={steps}",
          flush=True)
    s = W.make_solver(wl)
    s.run(steps)
    g = s.traces().astype(np.float64)
    fields = [s.get_field(f) for f in range(9)]
    s.close()
    o = Oracle(wl.nx, wl.ny, wl.nz, wl.dx, wl.dy, wl.dz, wl.dt, wl.nt, mode=wl.mode, pml=wl.pml,
               free_surface=wl.free_surface)
    W.apply(o, wl)
    o.run(steps)
    return g, o.traces(), fields, [o.get_field(f) for f in range(9)]


def test_iso_small_profile(gpu):
    from elanfmt_amd import workloads as W

    wl = W.iso_homogeneous(n=40, L=6, nt=80, rec_depth=3, nrec_side=3)
    g, r, gf, rf = _run_both(wl)
    assert np.abs(r).max() > 0
    assert rel_l2(g, r) < TOL
    for f in range(3):
        assert rel_l2(gf[f], rf[f]) < TOL, f


def test_ortho_small_separable(gpu):
    from elanfmt_amd import workloads as W

    wl = W.ortho_layered(nx=72, ny=40, nz=36, L=6, nt=70, nrec_side=3)
    g, r, gf, rf = _run_both(wl)
    assert rel_l2(g, r) < TOL
    for f in range(9):
        assert rel_l2(gf[f], rf[f]) < TOL, f
