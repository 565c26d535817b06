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


def test_double_couple_on_the_gpu_matches_the_whole_space_solution(gpu):
    """The HIP path (the iso law through its material table) against the
    closed-form whole-space field of a double couple Mxy: the bounds the
    oracle meets, at 10-18 cells from the source (receiver geometry and
    sign convention as in test_oracle_green)."""
    from elanfmt_amd import fwe, workloads as W
    from test_oracle_green import DC_F0, DC_M0, DC_N, DC_OFFS, DC_S, DX, RHO, VP, VS, _check, _ricker, \
        whole_space_velocity

    nt = 540
    dt = W.cfl_dt(VP, DX, DX, DX, 0.5)
    n, s = DC_N, DC_S
    wl = W.Workload("dc112", n, n, n, DX, DX, DX, dt, nt, fwe.MODE_ISO, fwe.PML_PROFILE,
                    c11=np.full((n, n, n), RHO * VP * VP, np.float32),
                    c12=np.full((n, n, n), RHO * (VP * VP - 2 * VS * VS), np.float32), c66=RHO * VS * VS,
                    bscal=(1 / RHO,) * 3, prof=tuple(W.pml_profile(n, 16, dt) for _ in range(3)),
                    stf=[(W.ricker(nt, dt, DC_F0, 1.2 / DC_F0) * np.float32(dt / DX ** 3)).astype(np.float32)],
                    injections=[(fwe.SXY, s, s, s, 0, DC_M0)],
                    receivers=[(1, s + m, s, s, 0) for m in DC_OFFS] + [(0, s, s + m, s, 0) for m in DC_OFFS])
    print("case dc112 whole-space double couple", flush=True)
    sol = W.make_solver(wl)
    assert sol.medium_encoding()[0] == fwe.MEDIUM_TABLE
    sol.run(nt)
    tr = sol.traces().astype(np.float64)
    sol.close()
    t = (np.arange(nt) + 1) * dt
    rate, drate = _ricker(DC_F0, 1.2 / DC_F0)
    mxy = np.zeros((3, 3))
    mxy[0, 1] = mxy[1, 0] = -DC_M0
    for q, m in enumerate(DC_OFFS):
        r = (m + 0.5) * DX
        _check(tr[q], whole_space_velocity(mxy, (1, 0, 0), r, t, rate, drate)[1], 0.06, 0.02)
        _check(tr[len(DC_OFFS) + q], whole_space_velocity(mxy, (0, 1, 0), r, t, rate, drate)[0], 0.06, 0.02)
