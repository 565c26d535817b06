"""Size-independent properties of the HIP path at the benchmark sizes, where
the FP64 oracle would take too long: linearity in the moment tensor on the
384x384x192 layered orthorhombic case, and batch members equal to single
runs on the 512x512x256 six-wavefield case."""
import numpy as np
import pytest

from conftest import rel_l2

pytestmark = pytest.mark.gpu


def test_linearity_in_the_moment_tensor_at_384x384x192(gpu):
    """traces(m1 + m2) = traces(m1) + traces(m2) to FP32 round-off on the
    full 384x384x192 grid (400 steps): the update is linear in the source."""
    from elanfmt_amd import fwe, workloads as W

    wl = W.config(1, nt=400)
    xs, ys, zs = wl.nx // 2 - 1, wl.ny // 2 - 1, wl.nz // 2 - 1
    m1 = np.array(wl.info["mt"]) if "mt" in wl.info else W.mt_lune_random(11, 1e15)[0]
    m2 = W.mt_lune_random(12, 1e15)[0]
    print(f"case ortho384x384x192 linearity, {wl.nt} steps", flush=True)
    tr = []
    for m in (m1, m2, np.asarray(m1) + np.asarray(m2)):
        wl.injections = fwe.mt_injections(fwe.SRC_LAYOUT_OK4, xs, ys, zs, list(m))
        s = W.make_solver(wl)
        s.run(wl.nt)
        tr.append(s.traces().astype(np.float64))
        s.close()
    assert np.abs(tr[0]).max() > 0 and np.abs(tr[1]).max() > 0
    assert rel_l2(tr[2], tr[0] + tr[1]) < 1e-5


def test_batch_members_equal_single_runs_at_512x512x256(gpu):
    """The six elementary-tensor wavefields advanced as one batch give each
    wavefield exactly the traces of its own single run (60 steps on the full
    512x512x256 medium)."""
    from elanfmt_amd import workloads as W

    wl = W.config(2, nt=60)
    xs, ys, zs = wl.nx // 2 - 1, wl.ny // 2 - 1, wl.nz // 2 - 1
    # receivers the first 60 steps reach: a few cells around the source
    wl.receivers = [(q, xs + dx, ys + dy, zs + dz, b) for b in range(wl.nbatch) for q in range(3)
                    for (dx, dy, dz) in ((4, 0, 0), (0, -5, 2), (3, 3, -3))]
    print(f"case ortho512x512x256 x{wl.nbatch}, {wl.nt} steps", flush=True)
    s = W.make_solver(wl)
    s.run(wl.nt)
    batch = s.traces()
    s.close()
    nr = len(batch) // wl.nbatch
    stf, inj, rec = wl.stf, wl.injections, wl.receivers
    for b in (0, 3, 5):
        wl.nbatch = 1
        wl.stf = [stf[b]]
        wl.injections = [(q, i, j, k, 0, c) for (q, i, j, k, bb, c) in inj if bb == b]
        wl.receivers = [(q, i, j, k, 0) for (q, i, j, k, bb) in rec if bb == b]
        s = W.make_solver(wl)
        s.run(wl.nt)
        single = s.traces()
        s.close()
        assert np.abs(single).max() > 0
        assert np.array_equal(single, batch[b * nr:(b + 1) * nr]), b
