"""The algorithmic byte model behind every roofline figure
(workloads.alg_bytes): each streamed 3-D array counts 4 B once per pass
(read and/or write), the material-table index 1 B, 1-D tables, the table
entries, sources and receivers 0.  Counted here array by array from the
kernels' loads and stores.  Also the lookup of the measured traffic that
bench.py reports beside it."""
from elanfmt_amd import fwe, workloads as W

F = 4  # bytes per FP32 element


def test_ortho_single_wavefield():
    # stress: R Vx Vy Vz Sxx..Syz + medium, W Sxx..Syz; velocity: R S6 V3 + buoyancy, W V3
    assert W.alg_bytes(fwe.MODE_ORTHO, fwe.PML_SEPARABLE) == (F * (3 + 6 + 9 + 6), F * (6 + 3 + 3 + 3))
    assert W.alg_bytes(fwe.MODE_ORTHO, fwe.PML_SEPARABLE) == (96, 60)
    assert W.alg_bytes(fwe.MODE_ORTHO, fwe.PML_WEIGHTS3D) == (100, 64)  # + the 3-D weight array, both passes
    assert W.alg_bytes(fwe.MODE_ORTHO, fwe.PML_SEPARABLE, medium=fwe.MEDIUM_TABLE) == (F * 15 + 1, F * 12 + 1)
    assert W.alg_bytes(fwe.MODE_ORTHO, fwe.PML_SEPARABLE, medium=fwe.MEDIUM_TABLE) == (61, 49)
    # isotropic arrays: C11, C12 and the three edge moduli in stress
    assert W.alg_bytes(fwe.MODE_ORTHO, fwe.PML_SEPARABLE, medium=fwe.MEDIUM_ISOTROPIC) == (F * 20, F * 15)


def test_iso_law():
    # stress: R V3 S6 c11 c12, W S6; velocity: R S6 V3, W V3 (scalar buoyancy)
    assert W.alg_bytes(fwe.MODE_ISO, fwe.PML_PROFILE) == (68, 48)
    # material table: the index byte replaces c11 and c12; velocity reads no medium
    assert W.alg_bytes(fwe.MODE_ISO, fwe.PML_PROFILE, medium=fwe.MEDIUM_TABLE) == (F * 15 + 1, F * 12)


def test_batches():
    # six wavefields, medium read once per workgroup group
    assert W.alg_bytes(fwe.MODE_ORTHO, fwe.PML_SEPARABLE, 6) == (F * (6 * 15 + 9), F * (6 * 12 + 3))
    assert W.alg_bytes(fwe.MODE_ORTHO, fwe.PML_SEPARABLE, 6, batch_tile=1) == (F * 6 * 24, F * 6 * 15)
    assert W.alg_bytes(fwe.MODE_ORTHO, fwe.PML_SEPARABLE, 6, medium=fwe.MEDIUM_TABLE, batch_tile=1) == (366, 294)
    assert W.alg_bytes(fwe.MODE_ORTHO, fwe.PML_SEPARABLE, 6, medium=fwe.MEDIUM_TABLE) == (361, 289)
    assert W.alg_bytes(fwe.MODE_ORTHO, fwe.PML_SEPARABLE, 6, batch_tile=3) == (F * (6 * 15 + 2 * 9), F * (6 * 12 + 2 * 3))


def test_single_pass():
    # fields R9 W9 per wavefield, medium (C9 b3 [w]) once
    assert W.alg_bytes(fwe.MODE_ORTHO, fwe.PML_SEPARABLE, fused=True) == (F * (18 + 12), 0)
    assert W.alg_bytes(fwe.MODE_ISO, fwe.PML_PROFILE, fused=True) == (F * (18 + 2), 0)


def _bench():
    import importlib.util
    import os

    root = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
    spec = importlib.util.spec_from_file_location("bench", os.path.join(root, "bench.py"))
    b = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(b)
    return b


def test_roofline_traffic_needs_the_exact_launch():
    """roofline.traffic is the counter traffic of exactly the kernel, order and
    z chunk the run used: another order or z chunk of the same kernel, or a
    kernel never profiled, gives None (no neighbouring entry stands in)."""
    b = _bench()
    k = "k_stress<2, 3, false, 64, 4, 1>"
    assert b._geom(False, "linear", (64, 4, 2), (2, 12)) == "linear kc2"
    assert b._geom(True, "linear", (64, 4, 2), (2, 12)) == "rows2 kc12"
    t = b._traffic(k, "linear kc2", "ortho384x384x192")
    assert t is not None and 1.0 < t / (61 * 384 * 384 * 192) < 1.5
    assert b._traffic(k, "xcd kc2", "ortho384x384x192") is None
    assert b._traffic(k, "linear kc4", "ortho384x384x192") is None
    assert b._traffic("k_stress<2, 3, false, 64, 8, 1>", "linear kc1", "ortho384x384x192") is None
    assert b._traffic(k, "linear kc2", "ortho512x512x256") is None
