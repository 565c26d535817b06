"""The legacy input generator fwe_gen_inputs (host code, no GPU): its files
equal the inputs the Python workload builders write for the same
configuration, and the legacy parameter reader of libfwe.so accepts them."""
import os
import subprocess

import numpy as np
import pytest

from elanfmt_amd import fwe as F, legacy, workloads as W

TOOL = os.path.join(os.path.dirname(F.RUN_PATH), "fwe_gen_inputs")


def _run(tmp, *args):
    r = subprocess.run([TOOL, f"dir={tmp}", *args], capture_output=True, text=True, timeout=60)
    assert r.returncode == 0, r.stderr
    return r


def _f32(d, name):
    return np.fromfile(os.path.join(d, name), dtype=np.float32)


def test_iso_inputs_match_the_workload_builder(fwe, tmp_path):
    a, b = tmp_path / "tool", tmp_path / "py"
    a.mkdir()
    _run(a, "nx=40", "ny=40", "nz=40", "nt=50", "L=5", "f0=4", "sdr=30:60:90", "m0=1e15")
    wl = W.iso_homogeneous(n=40, L=5, nt=50)
    legacy.write_inputs(str(b), wl, mts=np.array(wl.info["mt"], np.float32))
    for name in ["pa1.dat", "pa2.dat", "pa3.dat", "MTS.dat", "c11u_het.dat", "c12u_het.dat", "weights_x.dat",
                 "weights_y.dat", "weights_z.dat"]:
        np.testing.assert_array_equal(_f32(a, name), _f32(b, name), err_msg=name)
    np.testing.assert_allclose(_f32(a, "Src.dat"), _f32(b, "Src.dat"), rtol=2e-7, atol=0)
    import ctypes as C

    p = fwe.LegacyParams()
    assert fwe.lib().fwe_read_legacy_params(str(a).encode(), C.byref(p)) == 0
    assert (p.n_pa1, p.n_pa2, p.n_pa3, p.has_mts) == (10, 13, 12, 1)


@pytest.mark.parametrize("pml", ["taper", "weights"])
def test_ortho_inputs_match_the_workload_builder(fwe, tmp_path, pml):
    nx, ny, nz, L, seed = 36, 32, 30, 6, 5
    _run(tmp_path, f"nx={nx}", f"ny={ny}", f"nz={nz}", "nt=40", "mode=ortho", f"seed={seed}", f"L={L}",
         f"pml={pml}", "lune=9")
    cells, bcell, _, vmax = W.layered_ortho(nx, ny, nz, 5, seed)
    cij, buoy = W.ortho_from_cells(cells, bcell)
    dt = W.cfl_dt(vmax, 50.0, 50.0, 50.0, 0.5)
    assert np.isclose(_f32(tmp_path, "pa1.dat")[3], np.float32(dt))
    names = ["C11m", "C22m", "C33m", "C12m", "C13m", "C23m", "C44myz", "C55mxz", "C66mxy"]
    for n, arr in zip(names, cij):
        np.testing.assert_array_equal(_f32(tmp_path, n + ".dat"), arr.ravel(order="F"), err_msg=n)
    for ax, arr in zip("xyz", buoy):
        np.testing.assert_array_equal(_f32(tmp_path, f"vrho_{ax}_11_hc.dat"), arr.ravel(order="F"))
    t = [W.cell_profile(W.pml_profile(q, L, dt)) for q in (nx, ny, nz)]
    if pml == "taper":
        for ax, tt in zip("xyz", t):
            np.testing.assert_array_equal(_f32(tmp_path, f"taper_{ax}.dat"), tt)
    else:
        w = (t[0][:, None, None] * t[1][None, :, None] * t[2][None, None, :]).astype(np.float32)
        np.testing.assert_array_equal(_f32(tmp_path, "weights.dat"), w.ravel(order="F"))
    m, _, _ = W.mt_lune_random(9, 1e15)
    np.testing.assert_array_equal(_f32(tmp_path, "MTS.dat"), m.astype(np.float32))


def test_damping_that_does_not_fit_is_refused(tmp_path):
    r = subprocess.run([TOOL, f"dir={tmp_path}", "nx=30", "ny=30", "nz=30", "nt=10"], capture_output=True, text=True)
    assert r.returncode == 2 and "does not fit" in r.stderr


@pytest.mark.parametrize("mode", ["iso", "ortho"])
def test_vacuum_top_empties_the_first_cell_layer(fwe, tmp_path, mode):
    """fs=2: zero moduli in the top cell layer (k=0) and nowhere else, as the
    BC2 driver writes them (vizme_FastBiotED_GPU3D_OK_2024.m:238,243), and
    an undamped top."""
    nx, ny, nz = 24, 20, 16
    for fs in (0, 2):
        d = tmp_path / f"fs{fs}"
        d.mkdir()
        _run(d, f"nx={nx}", f"ny={ny}", f"nz={nz}", "nt=10", f"mode={mode}", "L=4", f"fs={fs}")
    names = ["c11u_het.dat", "c12u_het.dat"] if mode == "iso" else ["C11m.dat", "C22m.dat", "C33m.dat", "C12m.dat",
                                                                    "C13m.dat", "C23m.dat"]
    for name in names:
        a = _f32(tmp_path / "fs2", name).reshape(nz, ny, nx)
        b = _f32(tmp_path / "fs0", name).reshape(nz, ny, nx)
        assert np.all(a[0] == 0) and np.all(b[0] > 0), name
        np.testing.assert_array_equal(a[1:], b[1:], err_msg=name)
    if mode == "iso":
        wz = _f32(tmp_path / "fs2", "weights_z.dat")
        assert np.all(wz[: nz // 2] == 1) and wz[-1] < 1
    else:
        tz = _f32(tmp_path / "fs2", "taper_z.dat")
        assert tz[0] == 1 and tz[-1] < 1
        assert _f32(tmp_path / "fs0", "taper_z.dat")[0] < 1


def test_bad_free_surface_key_is_refused(tmp_path):
    r = subprocess.run([TOOL, f"dir={tmp_path}", "nx=10", "ny=10", "nz=10", "nt=5", "fs=3"], capture_output=True,
                       text=True, timeout=60)
    assert r.returncode == 2


def test_isohet_set_is_isotropic_and_consistent(tmp_path):
    """mode=isohet writes the station solver's file set (9km.cu:381-400) for
    a layered isotropic medium: edge shear moduli equal (c11 - c12) / 2 of
    the cells inside a layer, and no ortho moduli files."""
    nx, ny, nz = 30, 26, 22
    _run(tmp_path, f"nx={nx}", f"ny={ny}", f"nz={nz}", "nt=10", "mode=isohet", "L=4", "seed=9")
    names = {p.name for p in tmp_path.iterdir()}
    assert {"c11u_het.dat", "c12u_het.dat", "c44u_hetxy.dat", "c44u_hetxz.dat", "c44u_hetyz.dat",
            "vrho_x_11_hc.dat", "vrho_y_11_hc.dat", "vrho_z_11_hc.dat"} <= names
    assert not any(n.startswith("C11m") for n in names)
    c11 = _f32(tmp_path, "c11u_het.dat").reshape(nz, ny, nx)
    c12 = _f32(tmp_path, "c12u_het.dat").reshape(nz, ny, nx)
    gxy = _f32(tmp_path, "c44u_hetxy.dat").reshape(nz, ny + 1, nx + 1)
    mu = (c11.astype(np.float64) - c12) / 2
    assert np.all(mu > 0) and np.all(c12 > 0)
    # x-y edges of plane k inside the grid average four cells of the same layer
    np.testing.assert_allclose(gxy[:, 1:ny, 1:nx], mu[:, 1:, 1:], rtol=2e-6)
