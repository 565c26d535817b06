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


def test_iso_free_surface_source_near_top(gpu):
    """BC1 with injections in planes 1 and 2 (OK4 order: inject, then BC1)."""
    from elanfmt_amd import workloads as W

    wl = W.iso_homogeneous(n=36, L=5, nt=60, rec_depth=2, nrec_side=3, free_surface=1)
    c = 18
    wl.injections = [(0, c, c, 1, 0, 3e14), (4, c, c, 1, 0, 2e14), (4, c + 1, c, 2, 0, -1e14),
                     (5, c, c + 2, 2, 0, 5e13), (2, c, c, 0, 0, 1e14), (3, c, c, 2, 0, 1e14)]
    wl.receivers = [(q, c + a, c + b, 1, 0) for q in range(3) for a, b in ((0, 0), (3, 1), (-4, 2))]
    g, r, gf, rf = _run_both(wl)
    assert np.abs(r).max() > 0
    assert rel_l2(g, r) < TOL
    for f in range(9):
        assert rel_l2(gf[f], rf[f]) < TOL, f


@pytest.mark.parametrize("pml", [0, 2, 3])
def test_ortho_free_surface_all_damping_forms(gpu, pml):
    """ANISO order (BC1, then inject) with weights.dat, separable and no
    damping; injections in planes 0..2 exercise the BC1/injection order."""
    from elanfmt_amd import fwe, workloads as W

    wl = W.ortho_layered(nx=70, ny=36, nz=30, L=5, nt=60, nrec_side=3,
                         pml=fwe.PML_WEIGHTS3D if pml == 2 else fwe.PML_SEPARABLE)
    if pml == 0:
        wl.pml, wl.prof, wl.w3 = fwe.PML_NONE, None, None
    wl.free_surface = 1
    c = 17
    wl.injections += [(4, c, c, 1, 0, 2e14), (5, c, c, 1, 0, -2e14), (2, c, c, 0, 0, 1e14),
                      (4, c + 3, c, 2, 0, 1e14)]
    g, r, gf, rf = _run_both(wl)
    assert rel_l2(g, r) < TOL
    for f in range(9):
        assert rel_l2(gf[f], rf[f]) < TOL, f


@pytest.mark.parametrize("law,fs", [("ortho", 0), ("ortho", 1), ("iso", 0), ("iso", 1)])
def test_injections_into_shear_cells_outside_the_update_range_accumulate(gpu, law, fs):
    """The reference updates Sxy only at i, j >= 1 (Sxz at i, k >= 1, Syz at
    j, k >= 1) and never writes the other shear cells, so injections there
    accumulate from step to step; the kernels keep those cells' values."""
    from elanfmt_amd import fwe, workloads as W

    if law == "ortho":
        wl = W.ortho_layered(nx=40, ny=34, nz=30, L=4, nt=40, nrec_side=2)
    else:
        wl = W.iso_homogeneous(n=36, L=4, nt=40, rec_depth=2, nrec_side=2)
    wl.free_surface = fs
    wl.injections = [(fwe.SXY, 0, 9, 9, 0, 2e14), (fwe.SXY, 9, 0, 9, 0, -1e14), (fwe.SXZ, 0, 9, 10, 0, 1e14),
                     (fwe.SXZ, 11, 8, 0, 0, 3e14), (fwe.SYZ, 8, 0, 11, 0, 2e14), (fwe.SYZ, 8, 11, 0, 0, -2e14),
                     (fwe.SXX, 12, 12, 12, 0, 1e15)]
    wl.receivers = [(q, i, j, k, 0) for q in range(3) for (i, j, k) in ((1, 9, 9), (9, 1, 9), (11, 8, 1), (20, 20, 20))]
    g, r, gf, rf = _run_both(wl)
    for comp, (i, j, k) in ((fwe.SXY, (0, 9, 9)), (fwe.SXZ, (11, 8, 0)), (fwe.SYZ, (8, 0, 11))):
        f = 3 + comp  # fields: Vx Vy Vz, then the six stresses
        assert abs(rf[f][i, j, k]) > 0
        assert abs(gf[f][i, j, k] - rf[f][i, j, k]) <= 1e-5 * abs(rf[f][i, j, k])
    assert rel_l2(g, r) < TOL
    for f in range(9):
        assert rel_l2(gf[f], rf[f]) < TOL, f


def test_iso_vacuum_layer_top(gpu):
    """Vacuum-layer top (free_surface=2, the BC2 variant,
    diff_BC2_OK3/FastWaveED_GPU3D_v1.cu:309-314): empty top cell layer and
    shear modulus 1e-5 on plane k=1, BC1 off; source and receivers near the
    top, including shear injections on plane 1."""
    from elanfmt_amd import workloads as W

    wl = W.iso_homogeneous(n=36, L=5, nt=60, rec_depth=2, nrec_side=3, free_surface=2)
    wl.c11[:, :, 0] = 0
    wl.c12[:, :, 0] = 0
    c = 18
    wl.injections = [(2, c, c, 2, 0, 3e14), (4, c, c, 1, 0, 2e14), (5, c, c + 1, 1, 0, -1e14), (3, c, c, 1, 0, 1e14)]
    wl.receivers = [(q, c + a, c + b, k, 0) for q in range(3) for a, b in ((0, 0), (3, 1)) for k in (1, 2)]
    g, r, gf, rf = _run_both(wl)
    assert np.abs(r).max() > 0
    assert rel_l2(g, r) < TOL
    for f in range(9):
        assert rel_l2(gf[f], rf[f]) < TOL, f


@pytest.mark.parametrize("nb", [1, 3])
def test_ortho_vacuum_layer_top(gpu, nb):
    """The vacuum-layer top on the orthorhombic law (its three shear moduli on
    plane k=1), single and batched wavefields, against the oracle."""
    from elanfmt_amd import fwe, workloads as W

    wl = W.ortho_layered(nx=66, ny=34, nz=28, L=5, nt=50, nrec_side=2)
    wl.free_surface = 2
    wl.injections = fwe.mt_injections(fwe.SRC_LAYOUT_ANISO, 30, 15, 2, [1e15, -2e14, 3e14, 4e14, -5e14, 2e14])
    wl.receivers = [(q, 30 + a, 15, k, 0) for q in range(3) for a in (0, 4) for k in (1, 2)]
    g0, r, gf, rf = _run_both(wl)
    assert np.abs(r).max() > 0
    assert rel_l2(g0, r) < TOL
    for f in range(9):
        assert rel_l2(gf[f], rf[f]) < TOL, f
    if nb > 1:
        wl.nbatch = nb
        wl.stf = [wl.stf[0]] * nb
        wl.injections = [(q, i, j, k, b, c) for b in range(nb) for (q, i, j, k, _, c) in wl.injections]
        wl.receivers = [(q, i, j, k, b) for b in range(nb) for (q, i, j, k, _) in wl.receivers]
        s = W.make_solver(wl)
        s.run(wl.nt)
        g = s.traces()
        s.close()
        n1 = len(g) // nb
        for b in range(nb):
            assert np.array_equal(g[b * n1:(b + 1) * n1], g0.astype(np.float32)), b


def test_free_surface_mode_is_checked(gpu):
    from elanfmt_amd import fwe

    with pytest.raises(fwe.FweError):
        fwe.Solver(nx=20, ny=20, nz=20, dx=50.0, dy=50.0, dz=50.0, dt=1e-3, nt=10, mode=fwe.MODE_ISO,
                   pml=fwe.PML_PROFILE, free_surface=3)


def test_batch_members_equal_single_runs(gpu):
    """nbatch wavefields sharing one medium give exactly the traces of
    independent single-wavefield runs."""
    from elanfmt_amd import fwe, workloads as W

    wl = W.ortho_layered(nx=66, ny=34, nz=28, L=5, nt=50, nrec_side=2)
    ms = [[1e15, 0, 0, 0, 0, 0], [0, 0, 0, 3e14, -2e14, 0], [1e14, 2e14, -3e14, 0, 0, 5e14]]
    singles = []
    for m in ms:
        wl.nbatch = 1
        wl.injections = fwe.mt_injections(fwe.SRC_LAYOUT_OK4, 30, 15, 12, m)
        rec0 = [t for t in wl.receivers if t[4] == 0]
        wl.receivers = rec0
        wl.stf = wl.stf[:1]
        s = W.make_solver(wl)
        s.run(wl.nt)
        singles.append(s.traces())
        s.close()
    wl.nbatch = len(ms)
    wl.injections = sum((fwe.mt_injections(fwe.SRC_LAYOUT_OK4, 30, 15, 12, m, batch=b) for b, m in enumerate(ms)), [])
    wl.receivers = [t[:4] + (b,) for b in range(len(ms)) for t in rec0]
    wl.stf = wl.stf * len(ms)
    s = W.make_solver(wl)
    s.set_tile_cols(128)  # the batch on 128-column tiles, the singles on 64
    s.run(wl.nt)
    tb = s.traces()
    s.close()
    n0 = len(rec0)
    for b in range(len(ms)):
        assert np.array_equal(tb[b * n0:(b + 1) * n0], singles[b]), b


@pytest.mark.parametrize("enc", [1, 2])
def test_batch_layouts_agree(gpu, enc):
    """Six wavefields advanced with 6, 3, 2 or 1 of them per workgroup (the
    rest side by side in the grid), in every workgroup order, with the medium
    arrays (enc 1) and the material table (enc 2): the same bits."""
    from elanfmt_amd import fwe, workloads as W

    wl = W.ortho_layered(nx=66, ny=34, nz=28, L=5, nt=40, nrec_side=2)
    nb = 6
    rng = np.random.default_rng(4)
    rec0 = [t for t in wl.receivers if t[4] == 0]
    wl.nbatch, wl.stf = nb, [wl.stf[0] * (1 + 0.25 * b) for b in range(nb)]
    wl.injections = sum((fwe.mt_injections(fwe.SRC_LAYOUT_OK4, 20 + 3 * b, 15, 10 + b, rng.normal(0, 1e15, 6), batch=b)
                         for b in range(nb)), [])
    wl.receivers = [t[:4] + (b,) for b in range(nb) for t in rec0]
    ref = None
    for bt, order in ((6, 0), (3, 2), (2, 1), (1, 0), (1, 2), (0, 0)):
        print(f"case ortho66x34x28 x6 enc={enc} batch_tile={bt} order={order}", flush=True)
        s = W.make_solver(wl)
        s.set_medium_encoding(enc)
        s.set_batch_tile(bt)
        s.set_wg_order(order)
        assert s.batch_tile() == (bt if bt else (1 if enc == fwe.MEDIUM_TABLE else 6))
        s.run(wl.nt)
        t = s.traces()
        s.close()
        if ref is None:
            ref = t
            assert np.abs(ref).max() > 0
        assert np.array_equal(t, ref), (bt, order)
    s = W.make_solver(wl)
    for bad in (4, 7, -1):
        with pytest.raises(fwe.FweError):
            s.set_batch_tile(bad)
    s.autotune(1)
    assert s.batch_tile() in (1, 6)
    s.run(wl.nt)
    assert np.array_equal(s.traces(), ref)
    s.close()


def test_restart_and_field_roundtrip(gpu):
    """Saving all nine fields after N steps, resetting, restoring them and
    the step counter, and running N more equals 2N steps bit for bit."""
    from elanfmt_amd import workloads as W

    wl = W.iso_homogeneous(n=32, L=4, nt=40, rec_depth=2, nrec_side=2)
    a = W.make_solver(wl)
    a.run(40)
    fa = [a.get_field(f) for f in range(9)]
    a.close()
    b = W.make_solver(wl)
    b.run(20)
    saved = [b.get_field(f) for f in range(9)]
    b.reset()
    assert b.step() == 0 and all(np.all(b.get_field(f) == 0) for f in range(9))
    for f in range(9):
        b.set_field(f, saved[f])
        assert np.array_equal(b.get_field(f), saved[f])
    b.set_step(20)
    b.run(20)
    for f in range(9):
        assert np.array_equal(b.get_field(f), fa[f]), f
    b.close()


def test_error_statuses(gpu):
    from elanfmt_amd import fwe

    s = fwe.Solver(20, 20, 20, 1.0, 1.0, 1.0, 1e-3, 10, mode=fwe.MODE_ORTHO, pml=fwe.PML_NONE)
    with pytest.raises(fwe.FweError) as e:
        s.run(1)
    assert e.value.status == fwe.ERR_STATE
    with pytest.raises(fwe.FweError) as e:
        s.set_receivers([(0, 20, 3, 3, 0)])
    assert e.value.status == fwe.ERR_ARG
    with pytest.raises(fwe.FweError) as e:
        s.set_injections([(6, 3, 3, 3, 0, 1.0)])
    assert e.value.status == fwe.ERR_ARG
    s.close()
    with pytest.raises(fwe.FweError) as e:
        fwe.Solver(20, 20, 20, 1.0, 1.0, 1.0, 1e-3, 10, mode=fwe.MODE_ISO, pml=fwe.PML_WEIGHTS3D)
    assert e.value.status == fwe.ERR_UNSUPPORTED


def test_c0_full_size_parity(gpu):
    """The 128^3 isotropic configuration, 600 steps: receiver traces within
    1e-5 relative L2 of the FP64 oracle."""
    from elanfmt_amd import workloads as W

    wl = W.iso_homogeneous()
    g, r, gf, rf = _run_both(wl)
    err = rel_l2(g, r)
    print(f"C0 128^3 x 600 rel-L2 traces = {err:.3e}", flush=True)
    assert err < TOL


def test_legacy_drop_in_binary(gpu, tmp_path):
    """fwe_run reads the legacy .dat set in its working directory and writes
    0_infos.inf and 0_0_*.res like the reference's a.exe."""
    import subprocess

    from elanfmt_amd import fwe, legacy, workloads as W
    from oracle import Oracle

    wl = W.iso_homogeneous(n=40, L=6, nt=70, free_surface=1)
    m = np.array(wl.info["mt"], np.float32)
    legacy.write_inputs(str(tmp_path), wl, mts=m)
    xs = 18
    wl.injections = fwe.mt_injections(fwe.SRC_LAYOUT_OK4, xs, xs, xs, m)
    recs = [(2, 25, 21, 1, 0), (0, 25, 21, 1, 0), (1, 25, 24, 1, 0), (1, 25, 23, 1, 0)]
    wl.receivers = recs
    spec = ",".join(f"{'xyz'[c]}:{i}:{j}:{k}:{n}" for (c, i, j, k, _), n in
                    zip(recs, ["Rec", "Rec_x", "Rec_y", "Rec_y2"]))
    r = subprocess.run([fwe.RUN_PATH, "nx=40", "ny=40", "nz=40", "nt=70", "mode=iso", "fs=1",
                        f"xsrc={xs}", f"ysrc={xs}", f"zsrc={xs}", f"rec={spec}", "NPARS1=10", "NPARS2=13",
                        "NPARS3=12"], cwd=str(tmp_path), capture_output=True, text=True, timeout=120)
    print(r.stdout[-500:], r.stderr[-500:], flush=True)
    assert r.returncode == 0
    assert legacy.read_infos(str(tmp_path)) == (4, 40, 40, 40, 1, 1)
    o = Oracle(40, 40, 40, wl.dx, wl.dy, wl.dz, np.float32(wl.dt), 70, mode=0, pml=1, free_surface=1)
    W.apply(o, wl)
    o.run(70)
    ref = o.traces()
    got = np.stack([legacy.read_res(str(tmp_path), n) for n in ["Rec", "Rec_x", "Rec_y", "Rec_y2"]])
    assert rel_l2(got, ref) < TOL
    assert np.array_equal(legacy.read_res(str(tmp_path), "Src"), wl.stf[0])
    vx = legacy.read_res(str(tmp_path), "Vx", (41, 40, 40))
    assert rel_l2(vx, o.get_field(0)) < TOL
    assert "(autotuned)" in r.stdout
    # the default tiling with 9-step graph replays: the same files bit for bit
    first = {n: legacy.read_res(str(tmp_path), n) for n in ["Rec", "Rec_x", "Rec_y", "Rec_y2", "Vx"]}
    r = subprocess.run([fwe.RUN_PATH, "nx=40", "ny=40", "nz=40", "nt=70", "mode=iso", "fs=1",
                        f"xsrc={xs}", f"ysrc={xs}", f"zsrc={xs}", f"rec={spec}", "tune=0", "graph=9", "nout=35"],
                       cwd=str(tmp_path), capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    assert "9-step graphs" in r.stdout and "(autotuned)" not in r.stdout
    for n, v in first.items():
        assert np.array_equal(legacy.read_res(str(tmp_path), n), v), n


@pytest.mark.parametrize("cpml", [0, 6])
def test_legacy_binary_elementary_tensors(gpu, tmp_path, cpml):
    """fwe_run elementary=1: the six unit tensors advance as one batch; the
    usual outputs then hold the MTS.dat tensor synthesised from them (as a
    direct run of that tensor, to FP32 round-off), and 0_0_<name>_<Mc>.res
    each unit tensor's traces (against the oracle run of that tensor, with
    the weight files).  cpml=6: the same with a thin convolutional PML."""
    import subprocess

    from elanfmt_amd import fwe, legacy, workloads as W
    from oracle import Oracle

    wl = W.iso_homogeneous(n=40, L=6, nt=70, free_surface=1)
    m = np.array(wl.info["mt"], np.float32)
    a, b = tmp_path / "a", tmp_path / "b"
    for d in (a, b):
        d.mkdir()
        legacy.write_inputs(str(d), wl, mts=m)
    xs = 18
    recs = [(2, 25, 21, 1, 0), (0, 25, 21, 1, 0), (1, 25, 24, 1, 0), (1, 2, 23, 30, 0)]
    names = ["Rec", "Rec_x", "Rec_y", "Rec_y2"]
    spec = ",".join(f"{'xyz'[c]}:{i}:{j}:{k}:{n}" for (c, i, j, k, _), n in zip(recs, names))
    args = [fwe.RUN_PATH, "nx=40", "ny=40", "nz=40", "nt=70", "mode=iso", "fs=1", f"xsrc={xs}", f"ysrc={xs}",
            f"zsrc={xs}", f"rec={spec}"] + ([f"cpml={cpml}"] if cpml else [])
    r = subprocess.run(args, cwd=str(a), capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    print(f"case fwe_run elementary=1 on iso40, cpml={cpml}", flush=True)
    r = subprocess.run(args + ["elementary=1"], cwd=str(b), capture_output=True, text=True, timeout=120)
    print(r.stdout[-400:], r.stderr[-300:], flush=True)
    assert r.returncode == 0 and "six unit moment tensors" in r.stdout
    direct = np.stack([legacy.read_res(str(a), n) for n in names])
    synth = np.stack([legacy.read_res(str(b), n) for n in names])
    assert np.abs(direct).max() > 0
    assert rel_l2(synth, direct) < 1e-5
    assert rel_l2(legacy.read_res(str(b), "Vx", (41, 40, 40)), legacy.read_res(str(a), "Vx", (41, 40, 40))) < 1e-5
    assert np.array_equal(legacy.read_res(str(b), "Src"), legacy.read_res(str(a), "Src"))
    assert legacy.read_infos(str(b)) == legacy.read_infos(str(a))
    unit = {c: np.stack([legacy.read_res(str(b), f"{n}_{c}") for n in names])
            for c in ("Mxx", "Myy", "Mzz", "Mxy", "Mxz", "Myz")}
    comb = sum(np.float64(m[q]) * unit[c].astype(np.float64) for q, c in enumerate(unit))
    assert rel_l2(comb, synth) < 1e-6
    if cpml:
        return  # (the thin layer's traces against the oracle: test_legacy_binary_with_a_thin_cpml)
    # one unit tensor against the oracle
    e = np.zeros(6, np.float32)
    e[3] = 1.0
    wl.injections = fwe.mt_injections(fwe.SRC_LAYOUT_OK4, xs, xs, xs, e)
    wl.receivers = recs
    o = Oracle(40, 40, 40, wl.dx, wl.dy, wl.dz, np.float32(wl.dt), 70, mode=0, pml=1, free_surface=1)
    W.apply(o, wl)
    o.run(70)
    assert np.abs(o.traces()).max() > 0
    assert rel_l2(unit["Mxy"], o.traces()) < TOL
    # snapshots hold one wavefield: refused
    r = subprocess.run(args + ["elementary=1", "snap=30"], cwd=str(b), capture_output=True, text=True, timeout=60)
    assert r.returncode == 2 and "elementary" in r.stderr


def test_legacy_binary_snapshots_and_restart(gpu, tmp_path):
    """fwe_run snap=30 writes all nine fields every 30 steps; a second run
    with restart=1 in a directory holding the inputs and snapshot 1 ends
    with the same fields, and the same traces from the restart step on."""
    import shutil
    import subprocess

    from elanfmt_amd import fwe, legacy, workloads as W

    wl = W.iso_homogeneous(n=36, L=6, nt=70, free_surface=1)
    m = np.array(wl.info["mt"], np.float32)
    a, b = tmp_path / "a", tmp_path / "b"
    a.mkdir()
    b.mkdir()
    legacy.write_inputs(str(a), wl, mts=m)
    legacy.write_inputs(str(b), wl, mts=m)
    args = [fwe.RUN_PATH, "nx=36", "ny=36", "nz=36", "nt=70", "mode=iso", "fs=1", "xsrc=17", "ysrc=17", "zsrc=17",
            "rec=z:22:20:1:Rec,x:22:20:1:Rec_x,y:22:23:1:Rec_y,y:22:22:1:Rec_y2", "nout=0"]
    r = subprocess.run(args + ["snap=30"], cwd=str(a), capture_output=True, text=True, timeout=120)
    print(r.stdout[-300:], r.stderr[-300:], flush=True)
    assert r.returncode == 0
    assert (a / "1_0_step.inf").read_text().split() == ["30"] and (a / "2_0_step.inf").read_text().split() == ["60"]
    assert not (a / "3_0_step.inf").exists()
    for name in ["Vx", "Vy", "Vz", "Sxx", "Syy", "Szz", "Sxy", "Sxz", "Syz", "step"]:
        ext = "inf" if name == "step" else "res"
        shutil.copy(a / f"1_0_{name}.{ext}", b / f"1_0_{name}.{ext}")
    r = subprocess.run(args + ["restart=1"], cwd=str(b), capture_output=True, text=True, timeout=120)
    print(r.stdout[-300:], r.stderr[-300:], flush=True)
    assert r.returncode == 0
    shapes = {"Vx": (37, 36, 36), "Vy": (36, 37, 36), "Vz": (36, 36, 37)}
    for name, shp in shapes.items():
        assert np.array_equal(legacy.read_res(str(a), name, shp), legacy.read_res(str(b), name, shp)), name
    for name in ["Rec", "Rec_x", "Rec_y", "Rec_y2"]:
        ta, tb = legacy.read_res(str(a), name), legacy.read_res(str(b), name)
        assert np.array_equal(ta[30:], tb[30:]) and np.all(tb[:30] == 0), name
    assert np.abs(legacy.read_res(str(a), "Rec")).max() > 0
    # a missing snapshot is an error, not a crash
    r = subprocess.run(args + ["restart=7"], cwd=str(b), capture_output=True, text=True, timeout=120)
    assert r.returncode == 1 and "snapshot 7" in r.stderr


def test_graph_replay_equals_direct_launches(gpu):
    """The captured time loop (kernels reading the step index from device
    memory) gives the same traces and fields bit for bit, including a
    remainder shorter than one replay and a second call."""
    from elanfmt_amd import workloads as W

    wl = W.iso_homogeneous(n=36, L=5, nt=75, rec_depth=2, nrec_side=2)
    a = W.make_solver(wl)
    a.run(75)
    ta, fa = a.traces(), [a.get_field(f) for f in range(9)]
    a.close()
    b = W.make_solver(wl)
    b.set_graph_steps(16)
    b.run(40)   # 2 replays + 8 direct
    b.run(35)   # 2 replays + 3 direct
    assert b.step() == 75
    tb, fb = b.traces(), [b.get_field(f) for f in range(9)]
    b.close()
    assert np.array_equal(ta, tb)
    for f in range(9):
        assert np.array_equal(fa[f], fb[f]), f


def test_autotune_and_tilings_agree(gpu):
    """Every tiling gives the same result bit for bit; autotune picks one of
    them and leaves the context reset."""
    from elanfmt_amd import fwe, workloads as W

    wl = W.ortho_layered(nx=70, ny=30, nz=26, L=4, nt=30, nrec_side=2)
    ref = None
    for tx, ty, kc, order in ((64, 4, 8, 0), (64, 2, 3, 1), (64, 8, 26, 0), (64, 4, 5, 1), (64, 4, 1, 0),
                              (64, 4, 1, 1), (64, 8, 2, 0), (128, 4, 8, 1), (128, 8, 1, 0), (128, 2, 1, 1),
                              (128, 2, 26, 0), (64, 2, 1, 2), (64, 8, 3, 2), (128, 4, 26, 2), (128, 2, 2, 2)):
        s = W.make_solver(wl)
        s.set_tiling(ty, kc)
        s.set_tile_cols(tx)
        s.set_wg_order(order)
        assert s.tiling() == (ty, kc) and s.tile_cols() == tx and s.wg_order() == order
        s.run(30)
        t = s.traces()
        s.close()
        if ref is None:
            ref = t
        assert np.array_equal(t, ref), (ty, kc)
    s = W.make_solver(wl)
    ms = s.autotune(2)
    assert ms > 0 and s.step() == 0
    s.run(30)
    assert np.array_equal(s.traces(), ref)
    with pytest.raises(fwe.FweError):
        s.set_wg_order(3)
    s.close()
    # chunks below three planes are refused with a free surface
    wl.free_surface = 1
    s = W.make_solver(wl)
    with pytest.raises(fwe.FweError):
        s.set_tiling(4, 2)
    s.set_tiling(4, 3)
    s.close()


def test_autotune_on_a_three_plane_grid(gpu):
    """A grid of three planes with BC1 (chunks of at least three planes):
    the only z chunk is the whole column; autotune must find it."""
    from elanfmt_amd import fwe, workloads as W

    wl = W.ortho_layered(nx=40, ny=30, nz=3, nl=1, L=2, nt=12, nrec_side=2)
    wl.free_surface = 1
    wl.injections = fwe.mt_injections(fwe.SRC_LAYOUT_ANISO, 20, 15, 0, [1e15, -2e14, 3e14, 4e14, -5e14, 2e14])
    wl.receivers = [(q, 22, 15, 1, 0) for q in range(3)]
    s = W.make_solver(wl)
    s.run(wl.nt)
    ref = s.traces()
    s.close()
    s = W.make_solver(wl)
    assert s.autotune(1) > 0
    assert s.tiling()[1] == 3
    s.run(wl.nt)
    assert np.array_equal(s.traces(), ref) and np.abs(ref).max() > 0
    s.close()


@pytest.mark.parametrize("fs", [1, 2])
def test_group_with_a_free_surface_equals_separate_runs(gpu, fs):
    """Groups with BC1 or the vacuum-layer top (receivers and a shear
    injection near the top) give the members' separate traces bit for bit."""
    from elanfmt_amd import fwe, workloads as W

    wls = [W.iso_random(n, seed=q + 5, L=5, nt=60, f0=15.0, nrec_side=2) for q, n in enumerate((34, 40))]
    for wl in wls:
        wl.free_surface = fs
        c = wl.nx // 2
        wl.injections = wl.injections + [(fwe.SXZ, c, c, 1, 0, 2e14), (fwe.SZZ, c, c, 2, 0, 1e14)]
        wl.receivers = wl.receivers + [(q, c + 3, c, k, 0) for q in range(3) for k in (1, 2)]
    ref = []
    for wl in wls:
        s = W.make_solver(wl)
        s.run(wl.nt)
        ref.append(s.traces())
        s.close()
    sol = [W.make_solver(wl) for wl in wls]
    g = fwe.Group(sol)
    g.set_tiling(4, 3)
    g.run(60)
    for s, r in zip(sol, ref):
        assert np.abs(r).max() > 0
        assert np.array_equal(s.traces(), r)
    g.close()
    for s in sol:
        s.close()


def test_group_equals_separate_runs(gpu):
    """An uneven group of models advanced by one launch per half-step gives
    every member exactly the traces of its separate run; autotune over the
    group keeps that."""
    from elanfmt_amd import fwe, workloads as W

    wls = [W.iso_random(n, seed=q, L=5, nt=120, f0=15.0, nrec_side=2) for q, n in enumerate((36, 44, 52))]
    ref = []
    for wl in wls:
        s = W.make_solver(wl)
        s.run(wl.nt)
        ref.append(s.traces())
        s.close()
    sol = [W.make_solver(wl) for wl in wls]
    g = fwe.Group(sol)
    g.run(45)
    g.run(75)
    for s, r in zip(sol, ref):
        assert s.step() == 120
        assert np.abs(r).max() > 0
        assert np.array_equal(s.traces(), r)
    ms = g.autotune(2)
    assert ms > 0 and all(s.step() == 0 for s in sol)
    g.run(120)
    for s, r in zip(sol, ref):
        assert np.array_equal(s.traces(), r)
    # 128-column tiles for every member
    for s in sol:
        s.reset()
    g.set_tile_cols(128)
    g.set_tiling(4, 7)
    assert all(s.tile_cols() == 128 and s.tiling() == (4, 7) for s in sol)
    g.run(120)
    for s, r in zip(sol, ref):
        assert np.array_equal(s.traces(), r)
    with pytest.raises(fwe.FweError):
        g.set_tile_cols(96)
    g.close()
    for s in sol:
        s.close()


def test_group_of_many_small_models_equals_separate_runs(gpu):
    """48 cubes of 13..40 cells (ragged against every tile shape, some
    smaller than one tile) in one group: every member's traces and fields
    equal its separate run bit for bit, after the group's autotune too."""
    from elanfmt_amd import fwe, workloads as W

    wls = W.uneven_batch(nmodels=48, seed=5, edges=(13, 17, 24, 31, 40), nt=60, L=4)
    print(f"group of {len(wls)} models, edges {sorted(set(w.nx for w in wls))}", flush=True)
    ref = []
    for wl in wls:
        s = W.make_solver(wl)
        s.run(wl.nt)
        ref.append((s.traces(), [s.get_field(q) for q in range(9)]))
        s.close()
    sol = [W.make_solver(wl) for wl in wls]
    g = fwe.Group(sol)
    for tune in (False, True):
        if tune:
            assert g.autotune(2) > 0
        g.run(25)
        g.run(35)
        for s, (tr, f) in zip(sol, ref):
            assert np.abs(tr).max() > 0
            assert np.array_equal(s.traces(), tr)
            for q in range(9):
                assert np.array_equal(s.get_field(q), f[q]), q
        for s in sol:
            s.reset()
    g.close()
    for s in sol:
        s.close()


def test_combine_elementary_traces_equals_direct_runs(gpu):
    """Six elementary-tensor wavefields in one batched run, combined with the
    weights of a general tensor, reproduce the direct single-tensor run
    (linearity in the source) to FP32 round-off."""
    from conftest import rel_l2
    from elanfmt_amd import fwe, workloads as W

    m0 = 1e15
    wl = W.ortho_layered(nx=66, ny=40, nz=36, L=5, nt=60, f0=12.0, nrec_side=3, elementary=True, m0=m0)
    s = W.make_solver(wl)
    s.run(wl.nt)
    mts = [W.mt_lune_random(100 + q, m0)[0] for q in range(3)]
    comb = s.combine_traces(np.array(mts) / m0)
    s.close()
    nr = len(wl.receivers) // 6
    xs, ys, zs = wl.nx // 2 - 1, wl.ny // 2 - 1, wl.nz // 2 - 1
    for q, m in enumerate(mts):
        wl1 = W.ortho_layered(nx=66, ny=40, nz=36, L=5, nt=60, f0=12.0, nrec_side=3)
        wl1.injections = fwe.mt_injections(fwe.SRC_LAYOUT_OK4, xs, ys, zs, m)
        d = W.make_solver(wl1)
        d.run(wl1.nt)
        direct = d.traces()
        d.close()
        assert direct.shape == comb[q].shape == (nr, wl.nt)
        err = rel_l2(comb[q], direct)
        print(f"combine MT {q}: rel-L2 {err:.2e}", flush=True)
        assert np.abs(direct).max() > 0 and err < 1e-5


def test_generated_inputs_drive_the_binary(gpu, tmp_path):
    """fwe_gen_inputs writes an ortho input set that fwe_run runs as the
    reference's a.exe would (no MATLAB step)."""
    import os
    import subprocess

    from elanfmt_amd import fwe, legacy

    tool = os.path.join(os.path.dirname(fwe.RUN_PATH), "fwe_gen_inputs")
    r = subprocess.run([tool, f"dir={tmp_path}", "nx=64", "ny=48", "nz=40", "nt=80", "mode=ortho", "L=8", "seed=3",
                        "lune=4", "f0=12"], capture_output=True, text=True, timeout=60)
    assert r.returncode == 0, r.stderr
    r = subprocess.run([fwe.RUN_PATH, "nx=64", "ny=48", "nz=40", "nt=80", "mode=ortho", "pml=taper", "fs=0",
                        "xsrc=31", "ysrc=23", "zsrc=19", "rec=x:40:23:12:Rec,y:40:23:12:Rec_x,z:40:23:12:Rec_y",
                        "nout=0"], cwd=str(tmp_path), capture_output=True, text=True, timeout=120)
    print(r.stdout[-300:], r.stderr[-300:], flush=True)
    assert r.returncode == 0
    assert legacy.read_infos(str(tmp_path)) == (4, 64, 48, 40, 1, 1)
    tr = legacy.read_res(str(tmp_path), "Rec")
    assert tr.shape == (80,) and np.all(np.isfinite(tr)) and np.abs(tr).max() > 0


def test_isohet_file_set_runs_as_the_ortho_law(gpu, tmp_path):
    """The iso-het input set of the station solver (9km.cu:381-400:
    c11u_het, c12u_het, c44u_het{xy,xz,yz}, staggered buoyancy, damping)
    gives bit for bit the traces of the ortho file set that spells the same
    medium out as C11m=C22m=C33m, C12m=C13m=C23m and the three edge moduli."""
    import os
    import shutil
    import subprocess

    from elanfmt_amd import fwe, legacy

    tool = os.path.join(os.path.dirname(fwe.RUN_PATH), "fwe_gen_inputs")
    a, b = tmp_path / "isohet", tmp_path / "ortho"
    a.mkdir()
    r = subprocess.run([tool, f"dir={a}", "nx=56", "ny=40", "nz=36", "nt=70", "mode=isohet", "L=6", "seed=5",
                        "pml=weights", "lune=2", "f0=12"], capture_output=True, text=True, timeout=60)
    assert r.returncode == 0, r.stderr
    shutil.copytree(a, b)
    for name in ("c11u_het", "c12u_het", "c44u_hetxy", "c44u_hetxz", "c44u_hetyz"):
        os.remove(b / f"{name}.dat")
    for src, dst in (("c11u_het", ("C11m", "C22m", "C33m")), ("c12u_het", ("C12m", "C13m", "C23m")),
                     ("c44u_hetyz", ("C44myz",)), ("c44u_hetxz", ("C55mxz",)), ("c44u_hetxy", ("C66mxy",))):
        for d in dst:
            shutil.copy(a / f"{src}.dat", b / f"{d}.dat")
    traces = []
    for d, mode in ((a, "isohet"), (b, "ortho")):
        r = subprocess.run([fwe.RUN_PATH, "nx=56", "ny=40", "nz=36", "nt=70", f"mode={mode}", "pml=weights", "fs=1",
                            "xsrc=27", "ysrc=19", "zsrc=17", "rec=x:35:19:2:Rec,y:35:19:2:Rec_x,z:20:10:5:Rec_y",
                            "nout=0"], cwd=str(d), capture_output=True, text=True, timeout=120)
        print(mode, r.stdout[-200:], r.stderr[-200:], flush=True)
        assert r.returncode == 0
        traces.append(np.stack([legacy.read_res(str(d), n) for n in ("Rec", "Rec_x", "Rec_y")]))
    assert np.abs(traces[0]).max() > 0
    assert np.array_equal(traces[0], traces[1])


@pytest.mark.parametrize("dims,tile", [((67, 45, 23), (128, 8, 5)), ((50, 19, 17), (64, 2, 1)), ((130, 9, 12), (128, 4, 3))])
def test_ragged_grids_match_the_oracle(gpu, dims, tile):
    """Sizes that are not multiples of the tile (and narrower than one
    tile) on both laws, with free surface, against the FP64 oracle."""
    from elanfmt_amd import fwe, workloads as W

    nx, ny, nz = dims
    wl = W.ortho_layered(nx=nx, ny=ny, nz=nz, L=3, nt=40, f0=20.0, nrec_side=2)
    wl.free_surface = 1
    wl.injections = fwe.mt_injections(fwe.SRC_LAYOUT_ANISO, nx // 2, ny // 2, nz // 2, [1e15, -3e14, 2e14, 4e14, 1e14, -2e14])
    wl.receivers = [(q, i, j, k, 0) for q in range(3) for (i, j, k) in ((nx - 2, ny - 2, 1), (1, 1, nz - 2), (nx // 3, ny // 2, 2))]
    from oracle import Oracle

    s = W.make_solver(wl)
    s.set_tile_cols(tile[0])
    s.set_tiling(tile[1], max(3, tile[2]))
    s.run(wl.nt)
    g = s.traces().astype(np.float64)
    gf = [s.get_field(f) for f in range(9)]
    s.close()
    o = Oracle(nx, ny, nz, wl.dx, wl.dy, wl.dz, wl.dt, wl.nt, mode=wl.mode, pml=wl.pml, free_surface=1)
    W.apply(o, wl)
    o.run(wl.nt)
    assert np.abs(o.traces()).max() > 0
    assert rel_l2(g, o.traces()) < TOL
    for f in range(9):
        assert rel_l2(gf[f], o.get_field(f)) < TOL, f


def test_eight_wavefield_iso_batch_equals_single_runs(gpu):
    """The largest batch (8 wavefields) on the iso law: each member equals
    its single-wavefield run bit for bit."""
    from elanfmt_amd import fwe, workloads as W

    wl = W.iso_homogeneous(n=40, L=5, nt=45, rec_depth=3, nrec_side=2)
    rng = np.random.default_rng(21)
    ms = [rng.normal(0, 1e15, 6) for _ in range(8)]
    rec0 = list(wl.receivers)
    singles = []
    for m in ms:
        wl.nbatch, wl.stf, wl.receivers = 1, wl.stf[:1], rec0
        wl.injections = fwe.mt_injections(fwe.SRC_LAYOUT_OK4, 19, 19, 19, m)
        s = W.make_solver(wl)
        s.run(wl.nt)
        singles.append(s.traces())
        s.close()
    wl.nbatch = 8
    wl.injections = sum((fwe.mt_injections(fwe.SRC_LAYOUT_OK4, 19, 19, 19, m, batch=b) for b, m in enumerate(ms)), [])
    wl.receivers = [t[:4] + (b,) for b in range(8) for t in rec0]
    wl.stf = wl.stf * 8
    s = W.make_solver(wl)
    s.run(wl.nt)
    tb = s.traces()
    s.close()
    n0 = len(rec0)
    for b in range(8):
        assert np.array_equal(tb[b * n0:(b + 1) * n0], singles[b]), b


def test_legacy_binary_with_a_thin_cpml(gpu, tmp_path):
    """fwe_run cpml=6: the legacy input set with a 6-cell convolutional PML
    (sized by the medium's fastest speed, no damping on the BC1 top) in
    place of the weight files, against the oracle's CPML with the same
    coefficients; snapshots carry the memory variables, so a restart
    continues bit for bit, and a snapshot without them is refused."""
    import shutil
    import subprocess

    from elanfmt_amd import fwe, legacy, workloads as W
    from oracle import Oracle

    n, nt, L = 40, 80, 6
    wl = W.iso_homogeneous(n=n, L=6, nt=nt, free_surface=1)
    m = np.array(wl.info["mt"], np.float32)
    a, b = tmp_path / "a", tmp_path / "b"
    for d in (a, b):
        d.mkdir()
        legacy.write_inputs(str(d), wl, mts=m)
    xs = 18
    wl.injections = fwe.mt_injections(fwe.SRC_LAYOUT_OK4, xs, xs, xs, m)
    recs = [(2, 25, 21, 1, 0), (0, 36, 21, 1, 0), (1, 25, 38, 1, 0), (1, 2, 23, 30, 0)]
    wl.receivers = recs
    spec = ",".join(f"{'xyz'[c]}:{i}:{j}:{k}:{nm}" for (c, i, j, k, _), nm in
                    zip(recs, ["Rec", "Rec_x", "Rec_y", "Rec_y2"]))
    args = [fwe.RUN_PATH, f"nx={n}", f"ny={n}", f"nz={n}", f"nt={nt}", "mode=iso", "fs=1", f"xsrc={xs}",
            f"ysrc={xs}", f"zsrc={xs}", f"rec={spec}", "nout=0", f"cpml={L}"]
    r = subprocess.run(args + ["snap=40"], cwd=str(a), capture_output=True, text=True, timeout=120)
    print(r.stdout[-400:], r.stderr[-300:], flush=True)
    assert r.returncode == 0 and "convolutional PML" in r.stdout
    dt32 = float(np.float32(wl.dt))
    vmax = float(np.sqrt(np.float64(np.max(wl.c11)) * np.float64(max(np.float32(v) for v in wl.bscal))))
    t = [W.cpml_table(n, L, wl.dx, dt32, vmax, R=1e-4, kappa_max=1.0, alpha_max=0.0, power=2.0, lo=lo)
         for lo in (1, 1, 0)]
    o = Oracle(n, n, n, wl.dx, wl.dy, wl.dz, dt32, nt, mode=0, pml=fwe.PML_CPML, free_surface=1)
    wl.pml = fwe.PML_CPML
    wl.cpml_layer, wl.cpml = L, tuple(t)
    W.apply(o, wl)
    o.run(nt)
    got = np.stack([legacy.read_res(str(a), nm) for nm in ["Rec", "Rec_x", "Rec_y", "Rec_y2"]])
    assert np.abs(o.traces()).max() > 0
    assert rel_l2(got, o.traces()) < TOL
    assert rel_l2(legacy.read_res(str(a), "Vx", (n + 1, n, n)), o.get_field(0)) < TOL
    assert (a / "1_0_cpml.res").exists()
    for name in ["Vx", "Vy", "Vz", "Sxx", "Syy", "Szz", "Sxy", "Sxz", "Syz", "step"]:
        ext = "inf" if name == "step" else "res"
        shutil.copy(a / f"1_0_{name}.{ext}", b / f"1_0_{name}.{ext}")
    r = subprocess.run(args + ["restart=1"], cwd=str(b), capture_output=True, text=True, timeout=120)
    assert r.returncode == 1 and "cpml" in r.stderr  # the memory variables are missing
    shutil.copy(a / "1_0_cpml.res", b / "1_0_cpml.res")
    r = subprocess.run(args + ["restart=1"], cwd=str(b), capture_output=True, text=True, timeout=120)
    print(r.stdout[-300:], r.stderr[-300:], flush=True)
    assert r.returncode == 0
    for nm in ["Rec", "Rec_x", "Rec_y", "Rec_y2"]:
        ta, tb = legacy.read_res(str(a), nm), legacy.read_res(str(b), nm)
        assert np.array_equal(ta[40:], tb[40:]), nm
    assert np.array_equal(legacy.read_res(str(a), "Vx", (n + 1, n, n)), legacy.read_res(str(b), "Vx", (n + 1, n, n)))
