"""Physics checks of the HIP path that do not rely on the oracle."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _surface_speed(fwe, W, free_surface, vacuum_top=False):
    """Median phase speed of the vertical velocity between receivers 20 cells
    apart just under the top of a homogeneous Poisson solid, from the lag of
    the peak of their cross-correlation (refined by a parabola through the
    three samples around it)."""
    nx, ny, nz, dx, L = 320, 64, 80, 50.0, 12
    vp, vs, rho = 6000.0, 6000.0 / np.sqrt(3.0), 2700.0
    dt = W.cfl_dt(vp, dx, dx, dx, 0.5)
    nt, f0 = 2400, 2.0
    prof = (W.pml_profile(nx, L, dt), W.pml_profile(ny, L, dt),
            W.pml_profile(nz, L, dt, top_undamped=1))
    stf = W.ricker(nt, dt, f0, 1.2 / f0) * np.float32(dt / dx ** 3)
    xs, yc = 30, ny // 2
    xr = list(range(100, 261, 20))
    c11 = np.full((nx, ny, nz), rho * vp * vp, np.float32)
    c12 = np.full((nx, ny, nz), rho * (vp * vp - 2 * vs * vs), np.float32)
    if vacuum_top:
        c11[:, :, 0] = 0
        c12[:, :, 0] = 0
    wl = W.Workload("lamb", nx, ny, nz, dx, dx, dx, dt, nt, fwe.MODE_ISO, fwe.PML_PROFILE,
                    free_surface=free_surface, c11=c11, c12=c12, c66=rho * vs * vs,
                    bscal=(1 / rho, 1 / rho, 1 / rho), prof=prof, stf=[stf],
                    injections=[(fwe.SZZ, xs, yc, 3 if vacuum_top else 2, 0, 1e15)],
                    receivers=[(fwe.REC_VZ, x, yc, 2, 0) for x in xr])
    s = W.make_solver(wl)
    s.run(nt)
    tr = s.traces().astype(np.float64)
    s.close()
    speeds = []
    for a, b in zip(tr[:-1], tr[1:]):
        cc = np.correlate(b, a, mode="full")
        k = int(np.argmax(cc))
        y0, y1, y2 = cc[k - 1], cc[k], cc[k + 1]
        lag = k - (len(a) - 1) + 0.5 * (y0 - y2) / (y0 - 2 * y1 + y2)
        speeds.append(20 * dx / (lag * dt))
    return float(np.median(speeds)), vs


def test_rayleigh_wave_speed_along_the_bc1_free_surface(gpu):
    """Lamb's problem: in a homogeneous Poisson solid (vp/vs = sqrt 3) the
    surface wave along a traction-free top travels at 0.9194 vs.  With BC1
    (OK4.cu:271-283) the measured speed fits that within 3%; with a rigid
    undamped top the same measurement lands far from it (the P-wave
    dominates), so the agreement comes from the free surface."""
    from elanfmt_amd import fwe, workloads as W

    speed, vs = _surface_speed(fwe, W, free_surface=1)
    c_r = 0.9194 * vs
    print(f"BC1 top: surface wave speed {speed:.1f} m/s, Rayleigh {c_r:.1f} m/s ({speed / c_r - 1:+.2%})", flush=True)
    assert abs(speed / c_r - 1) < 0.03
    rigid, _ = _surface_speed(fwe, W, free_surface=0)
    print(f"rigid top: {rigid:.1f} m/s ({rigid / c_r - 1:+.2%})", flush=True)
    assert abs(rigid / c_r - 1) > 0.2


def test_rayleigh_wave_speed_under_a_vacuum_layer(gpu):
    """The vacuum-layer top (diff_BC2_OK3/FastWaveED_GPU3D_v1.cu:309-314: an
    empty top cell layer and shear modulus 1e-5 on plane k=1) approximates
    the traction-free surface less closely than BC1: the same measurement
    gives -2.9% on the oracle.  It must stay within 4% of 0.9194 vs (the
    rigid top is at +59%)."""
    from elanfmt_amd import fwe, workloads as W

    speed, vs = _surface_speed(fwe, W, free_surface=2, vacuum_top=True)
    c_r = 0.9194 * vs
    print(f"vacuum layer: surface wave speed {speed:.1f} m/s ({speed / c_r - 1:+.2%})", flush=True)
    assert abs(speed / c_r - 1) < 0.04
