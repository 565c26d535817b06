"""The single-pass sweep kernel (FWE_KERNEL_FUSED) against the two-pass
kernels (bit for bit: the same cell formulas, written with explicit fmaf)
and against the FP64 oracle (relative L2 <= 1e-5), on grids whose sizes are
not multiples of the sweep tile (62 columns x 2 or 4 rows x z chunk), with
injections and receivers on tile, halo and chunk edges, for every law,
medium encoding, damping form and top boundary.
"""
import numpy as np
import pytest

from conftest import rel_l2

pytestmark = pytest.mark.gpu

TOL = 1e-5


def _edge_points(wl, rows, kc, seed=0):
    """Injections and receivers at the grid's first and last columns (shear
    cells there are outside the update range and only accumulate their
    injections), at the sweep tiles' edge and halo columns (61..63,
    123..125) and at 64-column boundaries, rows around multiples of `rows`,
    planes around multiples of `kc`."""
    rng = np.random.default_rng(seed)
    nx, ny, nz = wl.nx, wl.ny, wl.nz
    ic = [i for i in (0, 1, 61, 62, 63, 64, 123, 124, 125, 128, nx - 2, nx - 1) if 0 <= i < nx]
    jc = [j for j in (1, rows - 1, rows, rows + 1, 2 * rows, ny - 2) if 0 <= j < ny]
    kcs = [k for k in (1, kc - 1, kc, kc + 1, 2 * kc - 1, nz - 2) if 0 <= k < nz]
    pts = sorted({(i, j, k) for i in ic for j in jc for k in kcs})
    inj, rec = [], []
    for n, (i, j, k) in enumerate(pts):
        if n % 3 == 0:
            inj.append((n % 6, i, j, k, 0, float(1e14 * rng.uniform(0.5, 2.0))))
        rec.append((n % 3, i, j, k, 0))
    # two injections into one cell and component: summed in list order
    inj.append((2, ic[1], jc[1], kcs[1], 0, 3e14))
    inj.append((2, ic[1], jc[1], kcs[1], 0, -1e14))
    return inj, rec


def _run(wl, kind, sweep=None, medium=None, steps=None):
    from elanfmt_amd import fwe, workloads as W

    s = W.make_solver(wl)
    if medium is not None:
        s.set_medium_encoding(medium)
    s.set_kernel(kind)
    if sweep is not None:
        s.set_sweep_tiling(*sweep)
    s.run(wl.nt if steps is None else steps)
    assert s.kernel() == kind
    tr = s.traces()
    f = [[s.get_field(q, b) for q in range(9)] for b in range(wl.nbatch)]
    enc = s.medium_encoding()[0]
    s.close()
    return tr, f, enc


def _oracle(wl, src_after_bc=None):
    from elanfmt_amd import workloads as W
    from oracle import Oracle

    o = Oracle(wl.nx, wl.ny, wl.nz, wl.dx, wl.dy, wl.dz, wl.dt, wl.nt, mode=wl.mode, pml=wl.pml,
               free_surface=wl.free_surface, src_after_bc=src_after_bc)
    W.apply(o, wl)
    o.run(wl.nt)
    return o


def _case(name):
    from elanfmt_amd import fwe, workloads as W

    if name == "orthoiso_sep":
        # 40^3 random-field cube: the ortho law on isotropic arrays
        return W.iso_random(40, 3, nt=40, L=5), fwe.MEDIUM_ISOTROPIC
    if name.startswith("ortho"):
        pml = {"sep": fwe.PML_SEPARABLE, "w3d": fwe.PML_WEIGHTS3D, "none": fwe.PML_SEPARABLE}[name.split("_")[2]]
        wl = W.ortho_layered(nx=131, ny=37, nz=29, L=5, nt=40, nrec_side=2, pml=pml)
        if name.endswith("none"):
            wl.pml, wl.prof = fwe.PML_NONE, None
        medium = fwe.MEDIUM_TABLE if name.split("_")[1] == "table" else fwe.MEDIUM_ARRAYS
        return wl, medium
    wl = W.iso_homogeneous(n=40, L=6, nt=40, rec_depth=2, nrec_side=2)
    if name.endswith("none"):
        wl.pml, wl.prof = fwe.PML_NONE, None
    medium = fwe.MEDIUM_TABLE if "table" in name else fwe.MEDIUM_ARRAYS
    return wl, medium


CASES = ["ortho_table_sep", "ortho_arrays_w3d", "ortho_table_none", "ortho_arrays_sep", "orthoiso_sep",
         "iso_table_profile", "iso_arrays_profile", "iso_arrays_none"]


@pytest.mark.parametrize("name", CASES)
@pytest.mark.parametrize("rows,kc", [(2, 5), (4, 7), (2, 29), (4, 1)])
def test_sweep_equals_two_pass_and_oracle(gpu, name, rows, kc):
    from elanfmt_amd import fwe

    wl, medium = _case(name)
    kc = min(kc, wl.nz)
    wl.injections, wl.receivers = _edge_points(wl, rows, kc, seed=rows * 100 + kc)
    print(f"sweep {name} rows={rows} kchunk={kc} grid={wl.nx}x{wl.ny}x{wl.nz}", flush=True)
    ts, fs, es = _run(wl, fwe.KERNEL_FUSED, (rows, kc), medium)
    t2, f2, e2 = _run(wl, fwe.KERNEL_TWO_PASS, None, medium)
    assert es == e2 == medium
    assert np.abs(t2).max() > 0
    assert np.array_equal(ts, t2)
    for q in range(9):
        assert np.array_equal(fs[0][q], f2[0][q]), q
    o = _oracle(wl)
    r = o.traces()
    assert rel_l2(ts, r) < TOL
    for q in range(9):
        assert rel_l2(fs[0][q], o.get_field(q)) < TOL, q


@pytest.mark.parametrize("fs_mode,after", [(1, 0), (1, 1), (2, 1)])
@pytest.mark.parametrize("rows,kc", [(2, 3), (4, 4)])
def test_sweep_free_surface(gpu, fs_mode, after, rows, kc):
    """BC1 in both injection orders and the vacuum-layer top, with
    injections in planes 0..3 (plane 1's Sxz/Syz are set from plane 2)."""
    from elanfmt_amd import fwe, workloads as W

    wl = W.ortho_layered(nx=70, ny=23, nz=17, L=4, nt=40, nrec_side=2)
    wl.free_surface = fs_mode
    inj = []
    for k in range(4):
        for comp in range(6):
            inj.append((comp, 30 + comp, 10 + k, k, 0, 1e14 * (1 + comp + k)))
    inj.append((fwe.SXZ, 62, rows, 1, 0, 2e14))
    inj.append((fwe.SXZ, 63, rows, 1, 0, 1e14))
    inj.append((fwe.SYZ, 61, rows - 1, 1, 0, -2e14))
    wl.injections = inj
    wl.receivers = [(q, i, j, k, 0) for q in range(3) for i in (30, 61, 62, 63, 64) for j in (rows, 11)
                    for k in (1, 2, 3)]
    print(f"sweep free surface {fs_mode} after_bc={after} rows={rows} kchunk={kc}", flush=True)

    def run(kind, sweep):
        s = fwe.Solver(wl.nx, wl.ny, wl.nz, wl.dx, wl.dy, wl.dz, wl.dt, wl.nt, mode=wl.mode, pml=wl.pml,
                       free_surface=fs_mode, src_after_bc=after)
        W.apply(s, wl)
        s.set_kernel(kind)
        if sweep:
            s.set_sweep_tiling(*sweep)
        s.run(wl.nt)
        out = s.traces(), [s.get_field(q) for q in range(9)]
        s.close()
        return out

    ts, fsw = run(fwe.KERNEL_FUSED, (rows, kc))
    t2, f2 = run(fwe.KERNEL_TWO_PASS, None)
    assert np.abs(t2).max() > 0
    assert np.array_equal(ts, t2)
    for q in range(9):
        assert np.array_equal(fsw[q], f2[q]), q
    o = _oracle(wl, src_after_bc=after)
    assert rel_l2(ts, o.traces()) < TOL


def test_sweep_batch_and_graph(gpu):
    """Six elementary-tensor wavefields in one sweep launch, replayed from a
    captured 6-step graph, split over two run calls: every wavefield equals
    the two-pass run bit for bit."""
    from elanfmt_amd import fwe, workloads as W

    wl = W.ortho_layered(nx=70, ny=30, nz=26, L=4, nt=41, nrec_side=2, elementary=True)
    t2, f2, _ = _run(wl, fwe.KERNEL_TWO_PASS)
    s = W.make_solver(wl)
    s.set_kernel(fwe.KERNEL_FUSED)
    s.set_sweep_tiling(2, 8)
    s.set_graph_steps(6)
    s.run(20)
    s.run(21)
    ts = s.traces()
    fb = [[s.get_field(q, b) for q in range(9)] for b in range(6)]
    s.close()
    assert np.abs(t2).max() > 0 and np.array_equal(ts, t2)
    for b in range(6):
        for q in range(9):
            assert np.array_equal(fb[b][q], f2[b][q]), (b, q)


def test_sweep_restart_and_switching(gpu):
    """Fields saved after a sweep run and loaded into a new context continue
    as the uninterrupted run; switching schedules between run calls keeps
    the state."""
    from elanfmt_amd import fwe, workloads as W

    wl = W.ortho_layered(nx=64, ny=24, nz=20, L=4, nt=40, nrec_side=2)
    t2, f2, _ = _run(wl, fwe.KERNEL_TWO_PASS)
    s = W.make_solver(wl)
    s.set_kernel(fwe.KERNEL_FUSED)
    s.run(13)
    s.set_kernel(fwe.KERNEL_TWO_PASS)
    s.run(7)
    s.set_kernel(fwe.KERNEL_FUSED)
    s.run(5)
    saved = [s.get_field(q) for q in range(9)]
    head = s.traces()
    s.close()
    r = W.make_solver(wl)
    r.set_kernel(fwe.KERNEL_FUSED)
    for q in range(9):
        r.set_field(q, saved[q])
    r.set_step(25)
    r.run(15)
    tail = r.traces()
    fr = [r.get_field(q) for q in range(9)]
    r.close()
    assert np.array_equal(head[:, :25], t2[:, :25])
    assert np.array_equal(tail[:, 25:], t2[:, 25:])
    for q in range(9):
        assert np.array_equal(fr[q], f2[0][q]), q


def test_autotune_may_pick_the_sweep(gpu):
    """Under AUTO the autotuner times both schedules; whichever it keeps,
    the run equals the two-pass run bit for bit."""
    from elanfmt_amd import fwe, workloads as W

    wl = W.ortho_layered(nx=96, ny=64, nz=48, L=6, nt=30, nrec_side=2)
    t2, _, _ = _run(wl, fwe.KERNEL_TWO_PASS)
    s = W.make_solver(wl)
    assert s.kernel() == fwe.KERNEL_TWO_PASS
    assert s.autotune(2) > 0
    kind = s.kernel()
    print(f"autotune kept kernel {kind}, sweep tiling {s.sweep_tiling()}", flush=True)
    s.run(wl.nt)
    assert np.array_equal(s.traces(), t2)
    s.close()
    s = W.make_solver(wl)
    s.set_kernel(fwe.KERNEL_FUSED)
    s.autotune(2)
    assert s.kernel() == fwe.KERNEL_FUSED
    s.run(wl.nt)
    assert np.array_equal(s.traces(), t2)
    with pytest.raises(fwe.FweError):
        s.set_sweep_tiling(3, 4)
    s.close()
