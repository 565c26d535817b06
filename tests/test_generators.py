"""Input generators (include/fwe_gen.h) against closed forms, invariants and
brute force."""
import ctypes as C
import math

import numpy as np
import pytest


def test_ricker_closed_form(fwe):
    from elanfmt_amd import workloads as W

    dt, f0, t0 = 1e-3, 5.0, 0.3
    s = W.ricker(700, dt, f0, t0, amp=2.0).astype(np.float64)
    assert np.argmax(s) == 300 and abs(s[300] - 2.0) < 1e-6          # peak amp at t0
    # zero crossings where (pi f0 (t-t0))^2 = 1/2
    tz = t0 + 1.0 / (math.pi * f0 * math.sqrt(2.0))
    iz = int(tz / dt)
    assert s[iz] > 0 > s[iz + 1]
    # minima value amp*(-2) exp(-3/2) at (pi f0 (t-t0))^2 = 3/2
    assert abs(s.min() - 2.0 * (-2.0) * math.exp(-1.5)) < 2e-3
    # reference pert() with t0 given in steps (OK4.cu:210-218) is the same curve
    A0, f0r, dtr = 1e9, 2e3, 1e-5
    t0s = 30.0 / 3e4 / dtr
    ref = [A0 * (1 - 2 * (math.pi * (i - t0s) * dtr * f0r) ** 2) * math.exp(-(math.pi * (i - t0s) * dtr * f0r) ** 2)
           for i in range(300)]
    got = W.ricker(300, dtr, f0r, t0s * dtr, A0)
    assert np.allclose(got, ref, rtol=1e-6, atol=1e-3)


def _mt_matrix(m):
    return np.array([[m[0], m[3], m[4]], [m[3], m[1], m[5]], [m[4], m[5], m[2]]])


@pytest.mark.parametrize("sdr", [(30, 60, 90), (0, 90, 0), (117, 33, -71), (250, 80, 170)])
def test_double_couple_invariants(fwe, sdr):
    from elanfmt_amd import workloads as W

    M = _mt_matrix(W.mt_sdr(*sdr, 1.0))
    ev = np.sort(np.linalg.eigvalsh(M))
    assert np.allclose(ev, [-1, 0, 1], atol=1e-12)
    assert abs(np.trace(M)) < 1e-12 and abs(np.linalg.det(M)) < 1e-12


def test_double_couple_known_cases(fwe):
    from elanfmt_amd import workloads as W

    # vertical strike-slip on a N-S plane: pure Mxy (Aki & Richards Box 4.4)
    assert np.allclose(W.mt_sdr(0, 90, 0, 1.0), [0, 0, 0, 1, 0, 0], atol=1e-12)
    # 45-degree thrust striking north: Mzz = 1, Myy = -1 (x north, y east, z down)
    assert np.allclose(W.mt_sdr(0, 45, 90, 1.0), [0, -1, 1, 0, 0, 0], atol=1e-12)
    # normal fault is the negative of the thrust
    assert np.allclose(W.mt_sdr(0, 45, -90, 1.0), [0, 1, -1, 0, 0, 0], atol=1e-12)
    # slip vector and normal: M = m0 (n s^T + s n^T) for strike 90, dip 90, rake 0
    assert np.allclose(W.mt_sdr(90, 90, 0, 1.0), [0, 0, 0, -1, 0, 0], atol=1e-12)


def _lune(g, d):
    from elanfmt_amd import fwe as F

    lam = np.zeros(3)
    F.lib().fwe_gen_lune_eigs(g, d, F.dptr(lam))
    return lam


def test_lune_eigenvalues(fwe):
    s2, s6, s3 = math.sqrt(2), math.sqrt(6), math.sqrt(3)
    assert np.allclose(_lune(0, 0), [1 / s2, 0, -1 / s2])                 # double couple
    assert np.allclose(_lune(0, 90), [1 / s3] * 3)                         # isotropic
    assert np.allclose(_lune(0, -90), [-1 / s3] * 3)
    assert np.allclose(_lune(30, 0), np.array([1, 1, -2]) / s6)           # CLVDs
    assert np.allclose(_lune(-30, 0), np.array([2, -1, -1]) / s6)
    for g, d in [(12.0, 33.0), (-25.0, -60.0)]:
        lam = _lune(g, d)
        assert abs(np.linalg.norm(lam) - 1) < 1e-12 and lam[0] >= lam[1] >= lam[2]
        # Tape & Tape (2012): delta = 90 - acos(tr/(sqrt3 |L|)), gamma = atan((-l1+2l2-l3)/(sqrt3 (l1-l3)))
        dd = 90 - math.degrees(math.acos(lam.sum() / s3))
        gg = math.degrees(math.atan((-lam[0] + 2 * lam[1] - lam[2]) / (s3 * (lam[0] - lam[2]))))
        assert abs(dd - d) < 1e-9 and abs(gg - g) < 1e-9


def test_random_rotation_is_proper_orthogonal(fwe):
    r = fwe.RNG()
    fwe.lib().fwe_rng_seed(C.byref(r), 5)
    u = np.zeros(9)
    for _ in range(20):
        fwe.lib().fwe_gen_random_rotation(C.byref(r), fwe.dptr(u))
        U = u.reshape(3, 3)
        assert np.allclose(U @ U.T, np.eye(3), atol=1e-12) and abs(np.linalg.det(U) - 1) < 1e-12


def test_lune_random_is_uniform_and_has_the_drawn_eigenvalues(fwe):
    r = fwe.RNG()
    fwe.lib().fwe_rng_seed(C.byref(r), 11)
    m = np.zeros(6)
    g, d = C.c_double(), C.c_double()
    vs, us = [], []
    for _ in range(4000):
        fwe.lib().fwe_gen_mt_lune_random(C.byref(r), 2.0, fwe.dptr(m), C.byref(g), C.byref(d))
        b = math.radians(90 - d.value)
        vs.append(math.sin(3 * math.radians(g.value)) / 3)
        us.append(0.75 * b - 0.5 * math.sin(2 * b) + math.sin(4 * b) / 16)
        if len(vs) <= 50:
            ev = np.sort(np.linalg.eigvalsh(_mt_matrix(m)))[::-1]
            assert np.allclose(ev, _lune(g.value, d.value) * math.sqrt(2) * 2.0, atol=1e-9)
    # v uniform on [-1/3,1/3], u uniform on [0, 3pi/4]: 10-bin chi^2 well below the 0.1% point (27.9)
    for x, lo, hi in ((vs, -1 / 3, 1 / 3), (us, 0, 0.75 * math.pi)):
        h, _ = np.histogram(x, bins=10, range=(lo, hi))
        chi2 = ((h - 400) ** 2 / 400).sum()
        assert chi2 < 27.9, chi2


def test_pml_profile_formula(fwe):
    from elanfmt_amd import workloads as W

    n, L, dt, R = 128, 20, 2.5e-3, 1e-8
    w = W.pml_profile(n, L, dt, R=R).astype(np.float64)
    assert w.shape == (n + 1,)
    # OK4.m:306: sigma(p) = 3/2/(L/n) log10(1/R) ((L-p+1)/L)^2; weights(p) and weights(n-p+2) (1-based)
    d1 = 1.5 * (n / L) * 8.0
    assert abs(w[0] - math.exp(-d1 * dt)) < 1e-7 and abs(w[n] - math.exp(-d1 * dt)) < 1e-7
    assert abs(w[L - 1] - math.exp(-d1 / L ** 2 * dt)) < 1e-7
    assert np.all(w[L:n - L + 1] == 1.0)
    assert np.allclose(w[:L + 1], w[::-1][:L + 1])
    wz = W.pml_profile(n, L, dt, R=R, top_undamped=1)
    assert np.all(wz[: n // 2] == 1.0) and wz[n] == w[n]    # OK4.m:338
    t = W.cell_profile(w)
    assert np.allclose(t, 0.5 * (w[1:] + w[:-1]), rtol=1e-7)


def test_taper_and_weights3d(fwe):
    from elanfmt_amd import workloads as W

    rate = 0.16 / 12
    t = W.taper(40, 12, 12, rate).astype(np.float64)
    # ANIS.m:474-506: depth 12..1 on the low side, 1..12 on the high side
    assert abs(t[0] - math.exp(-(12 * rate) ** 2)) < 1e-7 and t[11] == pytest.approx(math.exp(-rate ** 2))
    assert t[12] == 1.0 and t[27] == 1.0 and t[28] == pytest.approx(math.exp(-rate ** 2))
    w = W.weights3d(30, 26, 22, 7, 5, 0, 6, rate)  # no top margin (ANIS.m:484-486)
    tx, ty, tz = W.taper(30, 7, 7, rate), W.taper(26, 5, 5, rate), W.taper(22, 0, 6, rate)
    assert np.allclose(w, tx[:, None, None] * ty[None, :, None] * tz[None, None, :], rtol=3e-7)
    assert w[15, 13, 0] == 1.0


def test_stagger_and_harmonic_brute_force(fwe):
    from elanfmt_amd import workloads as W

    rng = np.random.default_rng(3)
    nx, ny, nz = 5, 4, 3
    b = rng.uniform(1, 2, (nx, ny, nz)).astype(np.float32)
    bx, by, bz = W.stagger_buoyancy(b)
    for i in range(nx + 1):
        for j in range(ny):
            for k in range(nz):
                e = (b[i - 1, j, k] / 2.0 + b[i, j, k] / 2.0) if 1 <= i <= nx - 1 else 0.0
                assert bx[i, j, k] == pytest.approx(e, rel=1e-6)
    assert np.all(by[:, 0] == 0) and np.all(by[:, ny] == 0) and np.all(bz[:, :, 0] == 0)
    assert by[2, 2, 1] == pytest.approx((b[2, 1, 1] + b[2, 2, 1]) / 2, rel=1e-6)
    g = rng.uniform(1, 3, (nx, ny, nz)).astype(np.float32)
    gyz, gxz, gxy = W.harmonic_edges(g, g, g)
    h = lambda *v: 4.0 / sum(1.0 / x for x in v)
    assert gyz[1, 2, 1] == pytest.approx(h(g[1, 1, 0], g[1, 2, 0], g[1, 1, 1], g[1, 2, 1]), rel=1e-6)
    assert gxz[3, 1, 2] == pytest.approx(h(g[2, 1, 1], g[3, 1, 1], g[2, 1, 2], g[3, 1, 2]), rel=1e-6)
    assert gxy[4, 3, 2] == pytest.approx(h(g[3, 2, 2], g[4, 2, 2], g[3, 3, 2], g[4, 3, 2]), rel=1e-6)
    assert np.all(gxy[0] == 0) and np.all(gxy[nx] == 0) and np.all(gyz[:, :, nz] == 0)
    # harmonic mean of equal values is the value
    gy, gz, gq = W.harmonic_edges(np.full((3, 3, 3), 7.0, np.float32), np.full((3, 3, 3), 7.0, np.float32),
                                  np.full((3, 3, 3), 7.0, np.float32))
    assert gq[1, 1, 0] == pytest.approx(7.0)


def _stiff(fwe, vp, vs, rho, p):
    c = np.zeros(9)
    pp = np.ascontiguousarray(p, np.float64)
    st = fwe.lib().fwe_gen_ortho_stiffness(vp, vs, rho, fwe.dptr(pp), fwe.dptr(c))
    return st, c


def test_ortho_stiffness_special_cases(fwe):
    vp, vs, rho = 5000.0, 2900.0, 2600.0
    st, c = _stiff(fwe, vp, vs, rho, [0] * 7)
    lam, mu = rho * (vp ** 2 - 2 * vs ** 2), rho * vs ** 2
    assert st == 0
    assert np.allclose(c, [lam + 2 * mu] * 3 + [lam] * 3 + [mu] * 3, rtol=1e-12)    # isotropic
    # VTI (Thomsen 1986) is eps1=eps2, delta1=delta2, gamma1=gamma2, delta3=0 (Tsvankin 1997)
    e, d, g = 0.1, 0.05, 0.08
    st, c = _stiff(fwe, vp, vs, rho, [e, e, d, d, 0.0, g, g])
    C11, C22, C33, C12, C13, C23, C44, C55, C66 = c
    assert st == 0
    assert C11 == pytest.approx(C33 * (1 + 2 * e)) and C22 == pytest.approx(C11)
    assert C44 == pytest.approx(C55) and C66 == pytest.approx(C44 * (1 + 2 * g))
    assert C12 == pytest.approx(C11 - 2 * C66)                              # delta3 = 0 -> horizontal isotropy
    assert C13 == pytest.approx(C23)
    assert ((C13 + C55) ** 2 - (C33 - C55) ** 2) / (2 * C33 * (C33 - C55)) == pytest.approx(d)
    # random draws in the benchmark range are positive definite
    rng = np.random.default_rng(0)
    for _ in range(50):
        st, c = _stiff(fwe, rng.uniform(4000, 6500), 3000.0, 2700.0, rng.uniform(0, 0.15, 7))
        if st == 0:
            M = np.array([[c[0], c[3], c[4]], [c[3], c[1], c[5]], [c[4], c[5], c[2]]])
            assert np.all(np.linalg.eigvalsh(M) > 0)


def test_layered_medium_and_cfl(fwe):
    from elanfmt_amd import workloads as W

    cells, b, tops, vmax = W.layered_ortho(8, 6, 40, 5, 9)
    assert tops[0] == 0 and np.all(np.diff(tops) >= 1) and tops[-1] < 40
    for t0, t1 in zip(list(tops), list(tops[1:]) + [40]):
        for c in cells + [b]:
            layer = c[:, :, t0:t1]
            assert np.all(layer == layer[0, 0, 0])
    rho = 1 / b[0, 0, 0]
    assert 2400 <= rho <= 2900
    assert vmax >= math.sqrt(cells[2][0, 0, 0] / rho)
    # same seed, same medium
    cells2, _, tops2, _ = W.layered_ortho(8, 6, 40, 5, 9)
    assert np.array_equal(tops, tops2) and np.array_equal(cells[0], cells2[0])
    assert W.cfl_dt(6000, 50, 50, 50, 0.5) == pytest.approx(0.5 * 50 / (6000 * math.sqrt(3)))


def test_random_field_statistics(fwe):
    from elanfmt_amd import fwe as F

    n = 48
    out = np.zeros(n ** 3, np.float32)
    F.lib().fwe_gen_random_field(n, n, n, 10.0, 4, F.fptr(out))
    f = out.reshape((n, n, n), order="F").astype(np.float64)
    assert abs(f.mean()) < 1e-6 and abs(f.std() - 1) < 1e-5
    # Gaussian covariance exp(-r^2/corr^2): lag-5 correlation ~ exp(-0.25) = 0.78
    c5 = np.mean(f[:-5] * f[5:]) / f.var()
    assert 0.6 < c5 < 0.9
    c20 = np.mean(f[:-20] * f[20:]) / f.var()
    assert abs(c20) < 0.25
