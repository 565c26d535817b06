"""Edge cases of the time loop through the C ABI: no points at all, a
source-free run, zero steps, and grids at the smallest sizes the library
accepts (3 cells per axis, ragged against every tile shape)."""
import numpy as np
import pytest

from conftest import rel_l2

pytestmark = pytest.mark.gpu


def _iso(nx, ny, nz, nt, fs=0, pml=None):
    from elanfmt_amd import fwe, workloads as W

    dx = 50.0
    vp, vs, rho = 6000.0, 3464.0, 2700.0
    dt = W.cfl_dt(vp, dx, dx, dx, 0.5)
    rng = np.random.default_rng(nx * 100 + ny * 10 + nz)
    c11 = np.full((nx, ny, nz), rho * vp * vp, np.float32)
    c12 = np.full((nx, ny, nz), rho * (vp * vp - 2 * vs * vs), np.float32)
    stf = (W.ricker(nt, dt, 8.0, 0.15) * np.float32(dt / dx ** 3)).astype(np.float32)
    inj = [(c, nx // 2, ny // 2, nz // 2, 0, float(v)) for c, v in enumerate(rng.uniform(0.5, 1.5, 6) * 1e15)]
    rec = sorted({(q, i, j, k, 0) for q in range(3) for i in range(1, nx - 1) for j in (1, ny - 2) for k in (1, nz - 2)})
    return W.Workload(f"iso{nx}x{ny}x{nz}", nx, ny, nz, dx, dx, dx, dt, nt, fwe.MODE_ISO,
                      fwe.PML_NONE if pml is None else pml, free_surface=fs, c11=c11, c12=c12,
                      c66=rho * vs * vs, bscal=(1 / rho, 1 / rho, 1 / rho), stf=[stf], injections=inj,
                      receivers=rec)


def test_no_points_and_no_source(gpu):
    """No injections and no receivers: the fields stay exactly zero and the
    trace block is empty; receivers without a source record exact zeros."""
    from elanfmt_amd import workloads as W

    wl = _iso(20, 18, 16, 25)
    wl.injections, wl.receivers = [], []
    print(f"case {wl.name} without points", flush=True)
    s = W.make_solver(wl)
    s.run(wl.nt)
    assert s.traces().shape == (0, wl.nt)
    for f in range(9):
        assert not np.any(s.get_field(f, 0)), f
    s.close()
    wl = _iso(20, 18, 16, 25)
    wl.injections = []
    s = W.make_solver(wl)
    s.run(wl.nt)
    t = s.traces()
    assert t.shape == (len(wl.receivers), wl.nt) and not np.any(t)
    s.close()


def test_zero_steps_is_a_no_op(gpu):
    from elanfmt_amd import workloads as W

    wl = _iso(24, 20, 18, 30)
    print(f"case {wl.name} run(0) then run(nt)", flush=True)
    s = W.make_solver(wl)
    s.run(0)
    assert not np.any(s.traces())
    for f in range(9):
        assert not np.any(s.get_field(f, 0)), f
    s.run(wl.nt)
    a = s.traces()
    s.close()
    s = W.make_solver(wl)
    s.run(wl.nt)
    assert np.array_equal(s.traces(), a) and np.abs(a).max() > 0
    s.close()


@pytest.mark.parametrize("shape,fs", [((3, 3, 3), 0), ((5, 3, 4), 0), ((3, 7, 5), 1), ((33, 3, 6), 1),
                                      ((65, 9, 3), 0), ((4, 4, 4), 1)])
def test_smallest_grids_match_the_oracle(gpu, shape, fs):
    """Grids of 3 to 65 cells per axis (partial tiles in x and y, a single
    interior plane in z) against the oracle, with and without BC1, through
    the material table and through the arrays."""
    from elanfmt_amd import fwe, workloads as W
    from oracle import Oracle

    nx, ny, nz = shape
    wl = _iso(nx, ny, nz, 40, fs=fs)
    print(f"case {wl.name} fs={fs}", flush=True)
    o = Oracle(nx, ny, nz, wl.dx, wl.dy, wl.dz, wl.dt, wl.nt, mode=wl.mode, pml=wl.pml, free_surface=fs)
    W.apply(o, wl)
    o.run(wl.nt)
    ref = o.traces()
    out = []
    for enc in (fwe.MEDIUM_AUTO, fwe.MEDIUM_ARRAYS):
        s = W.make_solver(wl)
        s.set_medium_encoding(enc)
        s.run(wl.nt)
        out.append(s.traces())
        s.close()
    assert np.array_equal(out[0], out[1])
    if np.abs(ref).max() == 0:
        assert not np.any(out[0])
    else:
        assert rel_l2(out[0].astype(np.float64), ref) < 1e-5
