"""The ortho medium read through a material table (one byte per cell naming
one of the distinct twelve-value tuples) against the twelve streamed arrays:
the same values reach the same arithmetic, so every trace and field must
agree bit for bit, on every damping form, free-surface mode, batch size,
tiling and the captured time loop.  The arrays path itself is held to the
FP64 oracle by the other parity tests and by the check below."""
import numpy as np
import pytest

from conftest import rel_l2

pytestmark = pytest.mark.gpu


def _run(wl, enc, steps=None, tiling=None, graph=0):
    from elanfmt_amd import workloads as W

    s = W.make_solver(wl)
    s.set_medium_encoding(enc)
    if tiling:
        s.set_tile_cols(tiling[0])
        s.set_tiling(tiling[1], tiling[2])
        s.set_wg_order(tiling[3])
    if graph:
        s.set_graph_steps(graph)
    s.run(wl.nt if steps is None else steps)
    out = s.traces(), [s.get_field(f, b) for b in range(wl.nbatch) for f in range(9)], s.medium_encoding()
    s.close()
    return out


def _distinct_tuples(wl):
    """Distinct (C[0..8], B[0..2]) tuples over the (nx+1)(ny+1)(nz+1) box,
    each array zero beyond its staggered extent (the padded layout adds only
    more all-zero elements, and the corner element is all zero already)."""
    n = (wl.nx + 1, wl.ny + 1, wl.nz + 1)
    cols = []
    for a in list(wl.cij) + list(wl.buoy):
        p = np.zeros(n, np.float32, order="F")
        p[:a.shape[0], :a.shape[1], :a.shape[2]] = a
        cols.append(p.ravel(order="F"))
    t = np.ascontiguousarray(np.stack(cols, -1)).view(np.dtype((np.void, 48))).ravel()
    return len(np.unique(t))


@pytest.mark.parametrize("pml,fs,nb", [(3, 0, 1), (3, 1, 3), (3, 2, 1), (2, 1, 1), (2, 0, 2), (0, 2, 3), (0, 0, 1)])
def test_table_equals_arrays(gpu, pml, fs, nb):
    from elanfmt_amd import fwe, workloads as W

    wl = W.ortho_layered(nx=66, ny=34, nz=28, L=5, nt=45, nrec_side=2, pml=pml if pml else fwe.PML_SEPARABLE)
    if pml == 0:
        wl.pml, wl.prof, wl.w3 = fwe.PML_NONE, None, None
    wl.free_surface = fs
    m = [1e15, -2e14, 3e14, 4e14, -5e14, 2e14]
    wl.nbatch = nb
    wl.stf = [wl.stf[0] * (1.0 + 0.5 * b) for b in range(nb)]
    wl.injections = sum((fwe.mt_injections(fwe.SRC_LAYOUT_ANISO, 30, 15, 2 + 5 * b, m, batch=b) for b in range(nb)), [])
    wl.receivers = [(q, 30 + a, 15, k, b) for b in range(nb) for q in range(3) for a in (0, 4) for k in (1, 2, 9)]
    print(f"case ortho66x34x28 pml={pml} fs={fs} nb={nb}", flush=True)
    ta, fa, ea = _run(wl, fwe.MEDIUM_ARRAYS)
    tt, ft, et = _run(wl, fwe.MEDIUM_AUTO)
    assert ea == (fwe.MEDIUM_ARRAYS, et[1])
    assert et[0] == fwe.MEDIUM_TABLE and et[1] == _distinct_tuples(wl) and 0 < et[1] <= 256
    assert np.abs(ta).max() > 0
    assert np.array_equal(tt, ta)
    for q, (x, y) in enumerate(zip(ft, fa)):
        assert np.array_equal(x, y), q


def test_table_tilings_and_graph_equal_arrays(gpu):
    from elanfmt_amd import fwe, workloads as W

    wl = W.ortho_layered(nx=70, ny=30, nz=26, L=4, nt=30, nrec_side=2)
    ref, _, _ = _run(wl, fwe.MEDIUM_ARRAYS)
    for tiling in ((64, 2, 1, 0), (64, 8, 26, 1), (128, 4, 3, 0), (128, 2, 1, 1), (128, 2, 2, 2), (64, 4, 1, 2)):
        print(f"case ortho70x30x26 table tiling={tiling}", flush=True)
        t, _, e = _run(wl, fwe.MEDIUM_TABLE, tiling=tiling)
        assert e[0] == fwe.MEDIUM_TABLE
        assert np.array_equal(t, ref), tiling
    t, _, _ = _run(wl, fwe.MEDIUM_TABLE, graph=7)
    assert np.array_equal(t, ref)


def test_table_matches_the_oracle(gpu):
    from elanfmt_amd import fwe, workloads as W
    from oracle import Oracle

    wl = W.ortho_layered(nx=72, ny=40, nz=36, L=6, nt=70, nrec_side=3)
    wl.free_surface = 1
    t, f, e = _run(wl, fwe.MEDIUM_TABLE)
    assert e[0] == fwe.MEDIUM_TABLE
    o = Oracle(wl.nx, wl.ny, wl.nz, wl.dx, wl.dy, wl.dz, wl.dt, wl.nt, mode=wl.mode, pml=wl.pml, free_surface=1)
    W.apply(o, wl)
    o.run(wl.nt)
    assert np.abs(o.traces()).max() > 0
    assert rel_l2(t.astype(np.float64), o.traces()) < 1e-5
    for q in range(9):
        assert rel_l2(f[q], o.get_field(q)) < 1e-5, q


@pytest.mark.parametrize("nvalues", [60, 400])
def test_table_limit_and_heterogeneous_fallback(gpu, nvalues):
    """C11 drawn per cell from nvalues levels: the table holds every distinct
    tuple up to 256 entries; beyond that the arrays are used, and asking for
    the table is refused."""
    from elanfmt_amd import fwe, workloads as W

    wl = W.ortho_layered(nx=40, ny=36, nz=30, L=4, nt=25, nrec_side=2)
    rng = np.random.default_rng(nvalues)
    levels = np.linspace(0.9, 1.0, nvalues).astype(np.float32)
    wl.cij = list(wl.cij)
    wl.cij[0] = (wl.cij[0] * levels[rng.integers(0, nvalues, wl.cij[0].shape)]).astype(np.float32)
    n = _distinct_tuples(wl)
    ta, fa, _ = _run(wl, fwe.MEDIUM_ARRAYS)
    s = W.make_solver(wl)
    enc = s.medium_encoding()
    if n <= 256:
        assert enc == (fwe.MEDIUM_TABLE, n)
        s.set_medium_encoding(fwe.MEDIUM_TABLE)
    else:
        assert enc == (fwe.MEDIUM_ARRAYS, 0)
        with pytest.raises(fwe.FweError):
            s.set_medium_encoding(fwe.MEDIUM_TABLE)
    s.run(wl.nt)
    assert np.array_equal(s.traces(), ta)
    s.close()


@pytest.mark.parametrize("fs,pml", [(0, 1), (1, 1), (0, 0)])
def test_iso_table_equals_arrays(gpu, fs, pml):
    """The iso law reads (c11, c12) through the table: a two-layer medium
    gives the bits of the two streamed arrays, on both damping forms and
    free-surface modes, and matches the oracle."""
    from elanfmt_amd import fwe, workloads as W
    from oracle import Oracle

    wl = W.iso_homogeneous(n=36, L=5, nt=60, rec_depth=3, nrec_side=2, free_surface=fs)
    # a faster lower half: two tuples in the cells plus the zero padding
    wl.c11 = wl.c11.copy()
    wl.c12 = wl.c12.copy()
    wl.c11[:, :, 18:] *= np.float32(1.3)
    wl.c12[:, :, 18:] *= np.float32(1.2)
    if pml == 0:
        wl.pml, wl.prof = fwe.PML_NONE, None
    print(f"case iso36 two layers fs={fs} pml={pml}", flush=True)
    ta, fa, ea = _run(wl, fwe.MEDIUM_ARRAYS)
    tt, ft, et = _run(wl, fwe.MEDIUM_AUTO)
    assert ea == (fwe.MEDIUM_ARRAYS, 3) and et == (fwe.MEDIUM_TABLE, 3)
    assert np.abs(ta).max() > 0
    assert np.array_equal(tt, ta)
    for q, (x, y) in enumerate(zip(ft, fa)):
        assert np.array_equal(x, y), q
    o = Oracle(wl.nx, wl.ny, wl.nz, wl.dx, wl.dy, wl.dz, wl.dt, wl.nt, mode=wl.mode, pml=wl.pml, free_surface=fs)
    W.apply(o, wl)
    o.run(wl.nt)
    assert rel_l2(tt.astype(np.float64), o.traces()) < 1e-5


def test_iso_table_limit_groups_and_refusals(gpu):
    """An iso medium with more distinct (c11, c12) pairs than the table holds
    streams the arrays and refuses the table; iso members on the table get
    the same bits in a group (groups read the arrays); the ortho law's
    isotropic form is refused for the iso law."""
    from elanfmt_amd import fwe, workloads as W

    wl = W.iso_homogeneous(n=24, L=4, nt=30, rec_depth=3, nrec_side=2)
    wl.c11 = wl.c11 * (1.0 + 0.01 * np.arange(24, dtype=np.float32))[:, None, None]
    wl.c11 = wl.c11 * (1.0 + 0.001 * np.arange(24, dtype=np.float32))[None, :, None]
    s = W.make_solver(wl)
    assert s.medium_encoding() == (fwe.MEDIUM_ARRAYS, 0)
    with pytest.raises(fwe.FweError):
        s.set_medium_encoding(fwe.MEDIUM_TABLE)
    with pytest.raises(fwe.FweError):
        s.set_medium_encoding(fwe.MEDIUM_ISOTROPIC)
    s.close()
    wls = [W.iso_homogeneous(n=n, L=4, nt=40, rec_depth=3, nrec_side=2) for n in (28, 36)]
    ref = []
    for w in wls:
        t, _, e = _run(w, fwe.MEDIUM_AUTO)
        assert e[0] == fwe.MEDIUM_TABLE
        ref.append(t)
    print("case iso group of two table members", flush=True)
    sol = [W.make_solver(w) for w in wls]
    g = fwe.Group(sol)
    g.run(40)
    for s, r in zip(sol, ref):
        assert np.array_equal(s.traces(), r)
    g.close()
    for s in sol:
        s.close()


def test_group_reads_the_arrays_and_equals_table_runs(gpu):
    """Groups read the medium arrays; members whose separate runs use the
    material table get the same bits from the group."""
    from elanfmt_amd import fwe, workloads as W

    wls = [W.ortho_layered(nx=n, ny=n - 6, nz=n - 10, L=4, nt=50, nrec_side=2, seed=3 + q)
           for q, n in enumerate((36, 44, 52))]
    ref = []
    for wl in wls:
        print(f"case {wl.name} table single run", flush=True)
        t, _, e = _run(wl, fwe.MEDIUM_AUTO)
        assert e[0] == fwe.MEDIUM_TABLE
        ref.append(t)
    sol = [W.make_solver(wl) for wl in wls]
    g = fwe.Group(sol)
    g.run(50)
    for s, r in zip(sol, ref):
        assert np.abs(r).max() > 0
        assert np.array_equal(s.traces(), r)
    g.close()
    for s in sol:
        s.close()


@pytest.mark.parametrize("fs", [0, 1])
def test_isotropic_arrays_equal_the_twelve(gpu, fs):
    """An isotropic random-field medium through the ortho law (C11 = C22 =
    C33, C12 = C13 = C23 in every cell, too many distinct tuples for a
    table): the eight-array form gives the bits of the twelve arrays, and
    both match the oracle."""
    from elanfmt_amd import fwe, workloads as W
    from oracle import Oracle

    wl = W.iso_random(40, seed=5, L=5, nt=60, f0=12.0, nrec_side=2)
    wl.free_surface = fs
    print(f"case isorand40 fs={fs}", flush=True)
    ta, fa, ea = _run(wl, fwe.MEDIUM_ARRAYS)
    ti, fi, ei = _run(wl, fwe.MEDIUM_AUTO)
    assert ea == (fwe.MEDIUM_ARRAYS, 0) and ei == (fwe.MEDIUM_ISOTROPIC, 0)
    assert np.abs(ta).max() > 0
    assert np.array_equal(ti, ta)
    for q, (x, y) in enumerate(zip(fi, fa)):
        assert np.array_equal(x, y), q
    o = Oracle(wl.nx, wl.ny, wl.nz, wl.dx, wl.dy, wl.dz, wl.dt, wl.nt, mode=wl.mode, pml=wl.pml, free_surface=fs)
    W.apply(o, wl)
    o.run(wl.nt)
    assert rel_l2(ti.astype(np.float64), o.traces()) < 1e-5


def test_isotropic_arrays_in_groups_and_refusals(gpu):
    """Groups of isotropic members read the eight arrays and equal the
    members' own runs; a member held to the twelve arrays puts the group on
    the twelve, with the same bits; a layered orthorhombic medium refuses
    the isotropic form."""
    from elanfmt_amd import fwe, workloads as W

    wls = [W.iso_random(n, seed=q, L=4, nt=40, f0=15.0, nrec_side=2) for q, n in enumerate((32, 40))]
    ref = []
    for wl in wls:
        t, _, e = _run(wl, fwe.MEDIUM_ISOTROPIC)
        assert e[0] == fwe.MEDIUM_ISOTROPIC
        ref.append(t)
    for held in (None, 1):
        print(f"case isorand group, member {held} on the twelve arrays", flush=True)
        sol = [W.make_solver(wl) for wl in wls]
        if held is not None:
            sol[held].set_medium_encoding(fwe.MEDIUM_ARRAYS)
        g = fwe.Group(sol)
        g.run(40)
        for s, r in zip(sol, ref):
            assert np.array_equal(s.traces(), r)
        g.close()
        for s in sol:
            s.close()
    wl = W.ortho_layered(nx=40, ny=36, nz=30, L=4, nt=5, nrec_side=2)
    s = W.make_solver(wl)
    with pytest.raises(fwe.FweError):
        s.set_medium_encoding(fwe.MEDIUM_ISOTROPIC)
    s.close()
