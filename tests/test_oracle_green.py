"""The oracle against the closed-form whole-space solution of a point moment
tensor in a homogeneous isotropic medium (Aki & Richards 2002, section 4.3:
the double-couple field of eq. 4.29 and its radiation patterns, written for
a general symmetric moment tensor): near-field, intermediate P and S and far-field P and S terms.
This pins what the discrete-dispersion checks leave open: the scale and
sign of the stress injection, the 1/dx^3 point-source normalisation, the
time level of the recorded samples, and the coupling of c11, c12 and c66 to
the P and S speeds through a full radiated wavefield.

Convention fixed here: the reference adds M_c * Src[it] to the stresses
(OK4.cu:336-342), i.e. dsigma/dt = C:grad v + M r(t) delta(x).  Taking
sigma' = sigma - M (int r) delta(x) gives rho dv/dt = div sigma' + div(M
delta), the equivalent body force of a moment tensor -M (A&R section 3.3:
a moment density m acts as the body force f = -div m).  So the radiated field is A&R's with moment
function M_AR(t) = -M int r.  Sample it is the velocity after the update
of step it, at t = (it + 1) dt.  Grids are large enough that the damping
layer's reflections reach no receiver inside the pulse."""
import functools

import numpy as np
import pytest

from elanfmt_amd import fwe, workloads as W
from oracle import Oracle

VP, VS, RHO, DX = 6000.0, 3464.0, 2700.0, 50.0


def _ricker(f0, t0):
    def r(t):
        a = np.pi * f0 * (t - t0)
        return (1 - 2 * a * a) * np.exp(-a * a)

    def dr(t):
        a = np.pi * f0 * (t - t0)
        return np.pi * f0 * (4 * a ** 3 - 6 * a) * np.exp(-a * a)

    return r, dr


def whole_space_velocity(m, gamma, r, t, rate, drate):
    """Velocity (3, nt) at distance r in direction gamma from a point source
    whose moment-rate function is m * rate(t) (m symmetric 3x3, A&R
    convention), in the medium (VP, VS, RHO).  The displacement terms of
    eq. 4.29 differentiated in time: moment M(t) -> rate, rate -> drate."""
    g = np.asarray(gamma, float)
    gpq = np.einsum("p,q,pq->", g, g, m)       # g_p g_q M_pq
    mg = m @ g                                  # M_nq g_q
    tr = np.trace(m)
    a_n = 15 * g * gpq - 3 * g * tr - 6 * mg     # near field
    a_ip = 6 * g * gpq - g * tr - 2 * mg         # intermediate P
    a_is = -(6 * g * gpq - g * tr - 3 * mg)      # intermediate S
    a_fp = g * gpq                               # far-field P
    a_fs = -(g * gpq - mg)                       # far-field S
    tau = np.linspace(r / VP, r / VS, 4001)
    near = np.trapezoid(tau[None, :] * rate(t[:, None] - tau[None, :]), tau, axis=1)
    tp, ts = t - r / VP, t - r / VS
    out = (np.outer(a_n, near) / r ** 4
           + np.outer(a_ip, rate(tp)) / (VP ** 2 * r ** 2)
           + np.outer(a_is, rate(ts)) / (VS ** 2 * r ** 2)
           + np.outer(a_fp, drate(tp)) / (VP ** 3 * r)
           + np.outer(a_fs, drate(ts)) / (VS ** 3 * r))
    return out / (4 * np.pi * RHO)


def _run(n, L, f0, nt, inj, rec):
    dt = W.cfl_dt(VP, DX, DX, DX, 0.5)
    c11 = np.full((n, n, n), RHO * VP * VP, np.float32)
    c12 = np.full((n, n, n), RHO * (VP * VP - 2 * VS * VS), np.float32)
    stf = (W.ricker(nt, dt, f0, 1.2 / f0) * np.float32(dt / DX ** 3)).astype(np.float32)
    wl = W.Workload("whole-space", n, n, n, DX, DX, DX, dt, nt, fwe.MODE_ISO, fwe.PML_PROFILE, c11=c11,
                    c12=c12, c66=RHO * VS * VS, bscal=(1 / RHO,) * 3,
                    prof=tuple(W.pml_profile(n, L, dt) for _ in range(3)), stf=[stf], injections=inj,
                    receivers=rec)
    o = Oracle(n, n, n, DX, DX, DX, dt, nt, mode=fwe.MODE_ISO, pml=fwe.PML_PROFILE, free_surface=0)
    W.apply(o, wl)
    o.run(nt)
    tr = o.traces()
    o.close()
    return tr, (np.arange(nt) + 1) * dt


DC_N, DC_S, DC_M0, DC_F0, DC_OFFS = 112, 56, 1e15, 2.5, (10, 14, 18)


@functools.lru_cache(maxsize=1)
def _dc_run():
    """One Sxy injection; Vy receivers along x, then Vx receivers along y."""
    s = DC_S
    inj = [(fwe.SXY, s, s, s, 0, DC_M0)]
    rec = [(1, s + m, s, s, 0) for m in DC_OFFS] + [(0, s, s + m, s, 0) for m in DC_OFFS]
    return _run(DC_N, 16, DC_F0, 540, inj, rec)


def _check(got, want, rel_tol, peak_tol):
    rel = np.linalg.norm(got - want) / np.linalg.norm(want)
    peak = np.abs(got).max() / np.abs(want).max()
    assert rel < rel_tol and abs(peak - 1) < peak_tol, (rel, peak)


def test_closed_form_formula_reduces_to_the_explosion():
    """Isotropic source: no near-field or S terms, and the P field is the
    Stokes solution u_r = (M(t-r/a)/r^2 + M'(t-r/a)/(a r)) / (4 pi rho a^2)."""
    r0, dr0 = _ricker(4.0, 0.3)
    t = np.linspace(0, 1, 500)
    v = whole_space_velocity(np.eye(3), (0.0, 0.6, 0.8), 900.0, t, r0, dr0)
    want = (r0(t - 900 / VP) / 900 ** 2 + dr0(t - 900 / VP) / (VP * 900)) / (4 * np.pi * RHO * VP ** 2)
    assert np.allclose(v, np.outer((0.0, 0.6, 0.8), want), rtol=0, atol=1e-12 * np.abs(want).max())


def test_explosion_matches_the_stokes_solution():
    """Equal injections on Sxx, Syy, Szz of one cell; Vx, Vy, Vz receivers
    on the three axes.  Vx(i) sits at x = i - 1/2 (between Sxx(i-1) and
    Sxx(i)), so the receiver m cells out is at r = (m - 1/2) dx."""
    n, s, m0, f0 = 72, 36, 1e15, 4.0
    offs = (12, 16, 20)
    inj = [(c, s, s, s, 0, m0) for c in (fwe.SXX, fwe.SYY, fwe.SZZ)]
    rec = [(c, s + m * (c == 0), s + m * (c == 1), s + m * (c == 2), 0) for c in range(3) for m in offs]
    tr, t = _run(n, 12, f0, 330, inj, rec)
    rate, drate = _ricker(f0, 1.2 / f0)
    for c in range(3):
        for q, m in enumerate(offs):
            gamma = np.eye(3)[c]
            want = whole_space_velocity(-m0 * np.eye(3), gamma, (m - 0.5) * DX, t, rate, drate)
            # measured: rel-L2 0.031-0.053, peak within 1.5%
            _check(tr[c * len(offs) + q], want[c], 0.08, 0.03)


def test_double_couple_matches_the_whole_space_solution():
    """One Sxy injection (a double couple Mxy = Myx).  Sxy(i, j) sits at
    (i - 1/2, j - 1/2), Vy(i, j) at (i, j - 1/2) and Vx(i, j) at
    (i - 1/2, j): a Vy receiver m cells out along x and a Vx receiver m
    cells out along y are both at r = (m + 1/2) dx on an axis, where only
    the transverse (S-polarised) component is non-zero and every term of
    the solution (near field, intermediate P and S, far-field S) adds."""
    tr, t = _dc_run()
    offs = DC_OFFS
    rate, drate = _ricker(DC_F0, 1.2 / DC_F0)
    mxy = np.zeros((3, 3))
    mxy[0, 1] = mxy[1, 0] = -DC_M0
    for q, m in enumerate(offs):
        r = (m + 0.5) * DX
        vx_axis = whole_space_velocity(mxy, (1, 0, 0), r, t, rate, drate)
        vy_axis = whole_space_velocity(mxy, (0, 1, 0), r, t, rate, drate)
        assert np.abs(vx_axis[0]).max() < 1e-12 * np.abs(vx_axis[1]).max()  # no radial P on the axis
        # measured: rel-L2 0.024-0.031, peak within 0.9%
        _check(tr[q], vx_axis[1], 0.06, 0.02)
        _check(tr[len(offs) + q], vy_axis[0], 0.06, 0.02)


def test_a_wrong_sign_or_speed_fails():
    """The check separates: a flipped source sign or a 5% wrong S speed in
    the closed form misses the oracle's traces by far more than the bound."""
    tr, t = _dc_run()
    got = tr[DC_OFFS.index(14)]
    rate, drate = _ricker(DC_F0, 1.2 / DC_F0)
    mxy = np.zeros((3, 3))
    mxy[0, 1] = mxy[1, 0] = DC_M0  # opposite sign
    want = whole_space_velocity(mxy, (1, 0, 0), 14.5 * DX, t, rate, drate)[1]
    with pytest.raises(AssertionError):
        _check(got, want, 0.06, 0.02)
    global VS
    vs = VS
    try:
        VS = vs * 1.05
        want = whole_space_velocity(-mxy, (1, 0, 0), 14.5 * DX, t, rate, drate)[1]
    finally:
        VS = vs
    with pytest.raises(AssertionError):
        _check(got, want, 0.06, 0.02)
