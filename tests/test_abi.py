"""The C ABI library loads, exports every declared symbol, and its host-only
helpers follow the reference (no GPU needed)."""
import os
import re

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _declared(header):
    txt = open(os.path.join(ROOT, "include", header)).read()
    txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
    return sorted(set(re.findall(r"\b(fwe_[a-z0-9_]+)\s*\(", txt)))


@pytest.mark.parametrize("header", ["fwe.h", "fwe_gen.h", "fwe_post.h"])
def test_every_declared_symbol_is_exported_and_bound(fwe, header):
    names = _declared(header)
    assert len(names) >= 3
    L = fwe.lib()
    for n in names:
        assert hasattr(L, n), n
        assert n in fwe.SIGNATURES, f"{n} has no ctypes signature"


def test_version_and_device_count(fwe):
    assert fwe.lib().fwe_version().decode().startswith("fwe 1")
    assert fwe.device_count() >= 0


def test_create_without_device_fails_loudly(fwe):
    if fwe.device_count() > 0:
        pytest.skip("a GPU is present")
    with pytest.raises(fwe.FweError) as e:
        fwe.Solver(16, 16, 16, 1.0, 1.0, 1.0, 0.1, 10)
    assert e.value.status == fwe.ERR_HIP


def test_config_default_matches_reference_loop(fwe):
    import ctypes as C

    cfg = fwe.Config()
    assert fwe.lib().fwe_config_default(C.byref(cfg), fwe.MODE_ISO) == 0
    # BC1 runs every step in both reference loops (OK4.cu:490, ANISO.cu:536)
    assert cfg.free_surface == 1 and cfg.src_after_bc == 0 and cfg.pml == fwe.PML_PROFILE
    assert fwe.lib().fwe_config_default(C.byref(cfg), fwe.MODE_ORTHO) == 0
    assert cfg.free_surface == 1 and cfg.src_after_bc == 1 and cfg.pml == fwe.PML_WEIGHTS3D
    assert fwe.lib().fwe_config_default(C.byref(cfg), 7) == fwe.ERR_ARG


def test_mt_injection_offsets_follow_reference_guards(fwe):
    m = [1, 2, 3, 4, 5, 6]
    inj = {t[0]: t for t in fwe.mt_injections(fwe.SRC_LAYOUT_OK4, 234, 384, 100, m)}
    # OK4.cu:336-342: guards a < i < a+2 select i = a+1; Syz uses ys+1<iy<ys+3,
    # xs+1<ix<xs+3 -> (xs+2, ys+2); Sxz and Sxy use +2 < . < +4 -> (xs+3, ys+3)
    for c in (fwe.SXX, fwe.SYY, fwe.SZZ):
        assert inj[c][1:4] == (235, 385, 101)
    assert inj[fwe.SYZ][1:4] == (236, 386, 101)
    assert inj[fwe.SXZ][1:4] == (237, 387, 101)
    assert inj[fwe.SXY][1:4] == (237, 387, 101)
    assert [inj[c][5] for c in range(6)] == m
    an = {t[0]: t for t in fwe.mt_injections(fwe.SRC_LAYOUT_ANISO, 180, 256, 200, m)}
    # ANISO.cu:233-247: Sxy ys+1<iy<ys+3, xs+1<ix<xs+3 -> (182, 258);
    # Syz iy -> ys+2, xs+2<ix<xs+4 -> (183, 258); Sxz ys<iy<ys+2, xs-2<ix<xs -> (179, 257)
    for c in (fwe.SXX, fwe.SYY, fwe.SZZ):
        assert an[c][1:4] == (181, 257, 201)
    assert an[fwe.SXY][1:4] == (182, 258, 201)
    assert an[fwe.SYZ][1:4] == (183, 258, 201)
    assert an[fwe.SXZ][1:4] == (179, 257, 201)


def test_legacy_params_reader(fwe, tmp_path):
    import ctypes as C

    pa1 = np.arange(1, 11, dtype=np.float32)
    pa2 = np.arange(20, 33, dtype=np.float32)
    pa3 = np.arange(40, 52, dtype=np.float32)
    for n, a in (("pa1", pa1), ("pa2", pa2), ("pa3", pa3)):
        a.tofile(tmp_path / f"{n}.dat")
    p = fwe.LegacyParams()
    assert fwe.lib().fwe_read_legacy_params(str(tmp_path).encode(), C.byref(p)) == 0
    assert (p.n_pa1, p.n_pa2, p.n_pa3, p.has_mts) == (10, 13, 12, 0)
    assert list(p.pa1) == list(pa1) and p.pa2[7] == 27 and p.pa3[2] == 42
    np.array([1, 2, 3, 4, 5, 6], np.float32).tofile(tmp_path / "MTS.dat")
    assert fwe.lib().fwe_read_legacy_params(str(tmp_path).encode(), C.byref(p)) == 0
    assert p.has_mts == 1 and list(p.mts) == [1, 2, 3, 4, 5, 6]
    # a missing input is an error status, not a crash (OK4.cu:190 freads NULL)
    os.remove(tmp_path / "pa2.dat")
    assert fwe.lib().fwe_read_legacy_params(str(tmp_path).encode(), C.byref(p)) == fwe.ERR_IO
    assert b"pa2.dat" in fwe.lib().fwe_last_error()


@pytest.mark.parametrize("field,value,status", [
    ("nx", 2, "ERR_ARG"), ("nz", 2, "ERR_ARG"), ("nt", -1, "ERR_ARG"), ("nbatch", 0, "ERR_ARG"),
    ("nbatch", 9, "ERR_ARG"), ("dx", 0.0, "ERR_ARG"), ("dt", -1.0, "ERR_ARG"), ("free_surface", 3, "ERR_ARG"),
    ("pml", 5, "ERR_ARG"), ("mode", 5, "ERR_ARG")])
def test_create_rejects_bad_configs(fwe, field, value, status):
    """Configuration errors are reported before any device is touched, with
    the status the header documents and no context returned."""
    import ctypes as C

    cfg = fwe.Config()
    assert fwe.lib().fwe_config_default(C.byref(cfg), fwe.MODE_ORTHO) == fwe.OK
    cfg.nx, cfg.ny, cfg.nz, cfg.nt = 16, 12, 10, 5
    cfg.dx = cfg.dy = cfg.dz = 10.0
    cfg.dt = 1e-3
    cfg.pml = fwe.PML_SEPARABLE
    setattr(cfg, field, value)
    h = C.c_void_p(1)
    assert fwe.lib().fwe_create(C.byref(cfg), C.byref(h)) == getattr(fwe, status)
    assert not h.value


def test_create_rejects_unsupported_law_damping_pairs(fwe):
    import ctypes as C

    for mode, pml in ((fwe.MODE_ISO, fwe.PML_WEIGHTS3D), (fwe.MODE_ISO, fwe.PML_SEPARABLE),
                      (fwe.MODE_ORTHO, fwe.PML_PROFILE)):
        cfg = fwe.Config()
        assert fwe.lib().fwe_config_default(C.byref(cfg), mode) == fwe.OK
        cfg.nx = cfg.ny = cfg.nz = 8
        cfg.nt = 3
        cfg.dx = cfg.dy = cfg.dz = 10.0
        cfg.dt = 1e-3
        cfg.pml = pml
        h = C.c_void_p()
        assert fwe.lib().fwe_create(C.byref(cfg), C.byref(h)) == fwe.ERR_UNSUPPORTED
    cfg.nx = cfg.ny = 4096
    cfg.nz = 1024  # 2^34 cells
    cfg.pml = fwe.PML_SEPARABLE
    assert fwe.lib().fwe_create(C.byref(cfg), C.byref(h)) == fwe.ERR_ARG


def test_host_arrays_are_checked_before_the_library_reads_them(fwe):
    """The Python mirror refuses arrays of the wrong length (the C ABI reads
    as many floats as the grid needs through the pointer), and media windows
    outside the generated grid."""
    import numpy as np

    from elanfmt_amd import learned as Lr

    assert fwe._sized(np.zeros((3, 4, 5)), 60, "a").size == 60
    with pytest.raises(ValueError):
        fwe._sized(np.zeros((3, 4, 5)), 61, "a")
    m = Lr.layered_medium((20, 20, 10), 1)
    assert Lr.crop(m, (2, 3, 1), (10, 12, 8))["bcell"].shape == (10, 12, 8)
    for o, n in (((0, 0, -1), (20, 20, 10)), ((0, 0, 0), (20, 20, 11)), ((5, 0, 0), (16, 20, 10))):
        with pytest.raises(ValueError):
            Lr.crop(m, o, n)


def test_drop_in_binary_refuses_bad_arguments_before_any_device_call(fwe, tmp_path):
    """fwe_run checks its arguments before touching inputs or the GPU:
    unknown keys, a missing grid size, and elementary=1 with snapshots or a
    restart (snapshots hold one wavefield) exit with status 2."""
    import subprocess

    cases = [(["nx=8", "ny=8", "nz=8", "bogus=1"], "unknown key"),
             (["nt=10"], "grid size not given"),
             (["nx=8", "ny=8", "nz=8", "elementary=1", "snap=5"], "elementary"),
             (["nx=8", "ny=8", "nz=8", "elementary=1", "restart=1"], "elementary")]
    for args, msg in cases:
        r = subprocess.run([fwe.RUN_PATH] + args, cwd=str(tmp_path), capture_output=True, text=True, timeout=60)
        assert r.returncode == 2, (args, r.stderr)
        assert msg in r.stderr, (args, r.stderr)
