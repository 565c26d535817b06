"""The convolutional PML on the GPU (two-pass kernels, every law and medium
encoding) against the FP64 oracle's CPML, through the C ABI.

Tolerance: relative L2 <= 1e-5 on receiver traces and on all nine fields
(FP32 GPU vs FP64 oracle).  Kernel forms (tilings, workgroup orders, batch
layouts, graph replay, restart with the memory variables) are compared bit
for bit.
"""
import numpy as np
import pytest

from conftest import rel_l2

pytestmark = pytest.mark.gpu

TOL = 1e-5


def _run_both(wl, steps=None, setup=None):
    from elanfmt_amd import workloads as W
    from oracle import Oracle

    steps = wl.nt if steps is None else steps
    print(f"case {wl.name} mode={wl.mode} cpml={wl.cpml_layer} fs={wl.free_surface} steps={steps}", flush=True)
    s = W.make_solver(wl)
    if setup:
        setup(s)
    s.run(steps)
    g = s.traces().astype(np.float64)
    fields = [s.get_field(f) for f in range(9)]
    s.close()
    o = Oracle(wl.nx, wl.ny, wl.nz, wl.dx, wl.dy, wl.dz, wl.dt, wl.nt, mode=wl.mode, pml=wl.pml,
               free_surface=wl.free_surface)
    W.apply(o, wl)
    o.run(steps)
    return g, o.traces(), fields, [o.get_field(f) for f in range(9)]


def _check(g, r, gf, rf):
    assert np.abs(r).max() > 0
    assert rel_l2(g, r) < TOL
    for f in range(9):
        assert rel_l2(gf[f], rf[f]) < TOL, f


@pytest.mark.parametrize("enc", [0, 1])
def test_cpml_ortho_layered_against_the_oracle(gpu, enc):
    """Layered orthorhombic medium through the material table (enc 0, the
    default) and the twelve arrays (enc 1); receivers inside the layer."""
    from elanfmt_amd import fwe, workloads as W

    wl = W.use_cpml(W.ortho_layered(nx=70, ny=44, nz=38, L=6, nt=90, nrec_side=3), 6)
    wl.receivers += [(q, i, j, k, 0) for q in range(3) for (i, j, k) in ((2, 20, 19), (66, 3, 30), (35, 42, 1))]
    g, r, gf, rf = _run_both(wl, setup=lambda s: s.set_medium_encoding(
        fwe.MEDIUM_AUTO if enc == 0 else fwe.MEDIUM_ARRAYS))
    _check(g, r, gf, rf)


@pytest.mark.parametrize("enc", [0, 1])
def test_cpml_iso_law_against_the_oracle(gpu, enc):
    from elanfmt_amd import fwe, workloads as W

    wl = W.use_cpml(W.iso_homogeneous(n=40, L=6, nt=90, rec_depth=1, nrec_side=3), 6)
    g, r, gf, rf = _run_both(wl, setup=lambda s: s.set_medium_encoding(
        fwe.MEDIUM_AUTO if enc == 0 else fwe.MEDIUM_ARRAYS))
    _check(g, r, gf, rf)


def test_cpml_isotropic_arrays_against_the_oracle(gpu):
    """A random-field isotropic cube: the kernels read C11 and C12 for the
    six normal moduli (the isotropic-array form)."""
    from elanfmt_amd import fwe, workloads as W

    wl = W.use_cpml(W.iso_random(40, seed=3, L=6, nt=80), 5)
    g, r, gf, rf = _run_both(wl, setup=lambda s: s.set_medium_encoding(fwe.MEDIUM_ISOTROPIC))
    _check(g, r, gf, rf)


@pytest.mark.parametrize("fs", [1, 2])
def test_cpml_ragged_grid_with_a_free_surface(gpu, fs):
    """BC1 (fs 1) or the vacuum-layer top (fs 2) with the z end at k = 0
    undamped, CPML on the other faces, a grid that fills no tile evenly,
    source and receivers near the top."""
    from elanfmt_amd import fwe, workloads as W

    wl = W.ortho_layered(nx=37, ny=29, nz=23, L=5, nt=70, nrec_side=2)
    wl.free_surface = fs
    W.use_cpml(wl, 5)
    wl.injections = fwe.mt_injections(fwe.SRC_LAYOUT_OK4, 17, 13, 3, [1e15, -5e14, 2e14, 3e14, -1e14, 4e14])
    wl.receivers = [(q, i, j, k, 0) for q in range(3) for (i, j, k) in ((17, 13, 1), (3, 4, 2), (33, 26, 20))]
    g, r, gf, rf = _run_both(wl)
    _check(g, r, gf, rf)


def test_cpml_tilings_orders_batches_and_graphs_agree(gpu):
    """Every tiling and workgroup order, a batch of three wavefields with
    three and one per workgroup, and graph replay: the same bits as one
    plain run per wavefield."""
    from elanfmt_amd import fwe, workloads as W

    base = W.use_cpml(W.ortho_layered(nx=70, ny=30, nz=26, L=5, nt=40, nrec_side=2), 5)
    rec0 = [t for t in base.receivers if t[4] == 0]
    ms = [[1e15, 0, 0, 0, 0, 0], [0, 0, 0, 3e14, -2e14, 0], [1e14, 2e14, -3e14, 0, 0, 5e14]]
    singles = []
    for m in ms:
        wl = W.use_cpml(W.ortho_layered(nx=70, ny=30, nz=26, L=5, nt=40, nrec_side=2), 5)
        wl.injections = fwe.mt_injections(fwe.SRC_LAYOUT_OK4, 31, 14, 12, m)
        wl.receivers = rec0
        ref = None
        for tx, ty, kc, order in ((64, 4, 8, 0), (64, 2, 3, 1), (128, 8, 26, 2), (128, 2, 1, 0), (64, 8, 5, 2)):
            print(f"case cpml70x30x26 m={m} tiling={tx}x{ty}x{kc} order={order}", flush=True)
            s = W.make_solver(wl)
            s.set_tile_cols(tx)
            s.set_tiling(ty, kc)
            s.set_wg_order(order)
            s.run(wl.nt)
            t = s.traces()
            s.close()
            if ref is None:
                ref = t
            assert np.array_equal(t, ref), (tx, ty, kc, order)
        singles.append(ref)
    wl = W.use_cpml(W.ortho_layered(nx=70, ny=30, nz=26, L=5, nt=40, nrec_side=2), 5)
    wl.nbatch = 3
    wl.stf = wl.stf * 3
    wl.injections = sum((fwe.mt_injections(fwe.SRC_LAYOUT_OK4, 31, 14, 12, m, batch=b) for b, m in enumerate(ms)), [])
    wl.receivers = [t[:4] + (b,) for b in range(3) for t in rec0]
    n0 = len(rec0)
    for bt, graph in ((3, 0), (1, 0), (3, 16), (1, 16)):
        print(f"case cpml batch of 3, batch_tile={bt} graph_steps={graph}", flush=True)
        s = W.make_solver(wl)
        s.set_batch_tile(bt)
        s.set_graph_steps(graph)
        s.run(25)
        s.run(15)
        tb = s.traces()
        s.close()
        for b in range(3):
            assert np.array_equal(tb[b * n0:(b + 1) * n0], singles[b]), (bt, graph, b)


def test_cpml_restart_with_memory_variables(gpu):
    """Fields, memory variables and the step saved after N steps and
    restored into a fresh context continue exactly as the uninterrupted run;
    without the memory variables they do not."""
    from elanfmt_amd import workloads as W

    wl = W.use_cpml(W.ortho_layered(nx=48, ny=40, nz=36, L=6, nt=80, nrec_side=2), 6)
    a = W.make_solver(wl)
    a.run(80)
    fa = [a.get_field(f) for f in range(9)]
    a.close()
    b = W.make_solver(wl)
    b.run(40)
    saved = ([b.get_field(f) for f in range(9)], b.cpml_state(), b.step())
    b.close()
    for with_psi in (True, False):
        c = W.make_solver(wl)
        for f in range(9):
            c.set_field(f, saved[0][f])
        if with_psi:
            c.set_cpml_state(saved[1])
        c.set_step(saved[2])
        c.run(40)
        fc = [c.get_field(f) for f in range(9)]
        c.close()
        same = all(np.array_equal(fa[f], fc[f]) for f in range(9))
        assert same == with_psi


def test_cpml_reset_and_refusals(gpu):
    """reset() clears the memory variables (two runs from rest agree bit for
    bit); the single-pass kernel and groups refuse the CPML; the autotuner
    keeps the two-pass kernels."""
    from elanfmt_amd import fwe, workloads as W

    wl = W.use_cpml(W.ortho_layered(nx=48, ny=40, nz=36, L=6, nt=60, nrec_side=2), 6)
    s = W.make_solver(wl)
    s.run(60)
    t1 = s.traces()
    s.reset()
    s.run(60)
    assert np.array_equal(s.traces(), t1)
    with pytest.raises(fwe.FweError):
        s.set_kernel(fwe.KERNEL_FUSED)
    s.autotune(2)
    assert s.kernel() == fwe.KERNEL_TWO_PASS
    s.run(60)
    assert np.array_equal(s.traces(), t1)
    with pytest.raises(fwe.FweError):
        fwe.Group([s])
    s.close()


def test_cpml_absorbs_on_the_gpu(gpu):
    """A 6-cell CPML against a far-padded reference on a 60^3 interior of
    the layered medium at 3 Hz: receiver error < 3e-3, against > 0.05 for
    the 6-cell and > 0.015 for the 24-cell plain (separable) layer."""
    from elanfmt_amd import learned as Lr, workloads as W

    n_int, nt, f0, L = 60, 400, 3.0, 6
    pad = int(np.ceil(0.5 / np.sqrt(3) * nt / 2 + 62))
    n_ref = n_int + 2 * pad
    med = Lr.layered_medium(n_ref, 7)
    dt = W.cfl_dt(med["vmax"], Lr.DX, Lr.DX, Lr.DX, 0.5)
    m, _, _ = W.mt_lune_random(218, 1e15)
    c = (n_int // 2 - 3, n_int // 2 + 2, n_int // 2 + 1)
    rec_int = Lr.face_receivers(n_int, 0, inset=2)

    def run(damp, kind, R=1e-8):
        n = n_int + 2 * damp if kind != "ref" else n_ref
        sh = (n - n_int) // 2
        rec = [(q, i + sh, j + sh, k + sh, b) for (q, i, j, k, b) in rec_int]
        wl = Lr.workload(f"g{n}", n, 24 if kind == "ref" else damp, Lr.crop(med, (n_ref - n) // 2, n), dt,
                         tuple(v + sh for v in c), m, f0, nt, rec, R=R)
        if kind == "cpml":
            W.use_cpml(wl, damp, vmax=med["vmax"], alpha_max=np.pi * f0)
        print(f"case absorption {kind} {n}^3", flush=True)
        s = W.make_solver(wl)
        s.run(nt)
        t = s.traces().astype(np.float64)
        s.close()
        return t

    ref = run(0, "ref")
    e_c = rel_l2(run(L, "cpml"), ref)
    e_p6 = rel_l2(run(L, "plain", R=1e-3), ref)
    e_p24 = rel_l2(run(24, "plain"), ref)
    print(f"errors: cpml6 {e_c:.2e} plain6 {e_p6:.2e} plain24 {e_p24:.2e}", flush=True)
    assert e_c < 3e-3
    assert e_p6 > 0.05 and e_p24 > 0.015


def test_cpml_layer_set_refused_and_changed(gpu):
    """A CPML context runs only once its layer is set; a layer that does not
    fit the grid is refused and leaves the set layer in place; setting a
    layer of another width reallocates the memory variables (the state size
    follows) and the run then equals a fresh context with that layer, bit
    for bit."""
    import copy

    from elanfmt_amd import fwe, workloads as W

    def wl_of(layer):
        return W.use_cpml(W.ortho_layered(nx=40, ny=36, nz=30, L=5, nt=40, nrec_side=2), layer)

    def traces(wl):
        s = W.make_solver(wl)
        s.run(wl.nt)
        t = s.traces()
        s.close()
        return t

    wl5, wl7 = wl_of(5), wl_of(7)
    t5, t7 = traces(wl5), traces(wl7)
    print("case cpml layer checks 40x36x30", flush=True)
    s = fwe.Solver(wl5.nx, wl5.ny, wl5.nz, wl5.dx, wl5.dy, wl5.dz, wl5.dt, wl5.nt, mode=wl5.mode, pml=fwe.PML_CPML)
    bare = copy.copy(wl5)
    bare.pml = fwe.PML_NONE  # everything but the damping
    W.apply(s, bare)
    with pytest.raises(fwe.FweError):
        s.run(1)
    s.set_cpml(5, *wl5.cpml)
    with pytest.raises(fwe.FweError):
        s.set_cpml(16, *wl5.cpml)  # 2 x 16 > nz
    s.reset()
    s.run(wl5.nt)
    assert np.array_equal(s.traces(), t5)
    s.set_cpml(7, *wl7.cpml)
    n7 = 6 * 14 * (wl7.ny * wl7.nz + wl7.nx * wl7.nz + wl7.nx * wl7.ny)
    assert s.cpml_state().size == n7
    s.reset()
    s.run(wl7.nt)
    assert np.array_equal(s.traces(), t7)
    s.close()
