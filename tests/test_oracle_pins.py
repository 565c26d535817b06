"""Pins of the CPU oracle against the reference's stated semantics and
against properties the mathematics fixes.

Citations: OK4.cu = src/FastWaveED_GPU3D_v1_DC_Earthquake_FULL_OK4_final/
FastWaveED_GPU3D_v1.cu, ANISO.cu = src/FastWaveED_GPU3D_v1_9km_station_ANISO/
FastWaveED_GPU3D_v1_9kmst_ANISO.cu.
"""
import numpy as np
import pytest

from conftest import rel_l2

F_VX, F_VY, F_VZ, F_SXX, F_SYY, F_SZZ, F_SXY, F_SXZ, F_SYZ = range(9)


def _shapes(nx, ny, nz):
    return [(nx + 1, ny, nz), (nx, ny + 1, nz), (nx, ny, nz + 1), (nx, ny, nz), (nx, ny, nz), (nx, ny, nz),
            (nx + 1, ny + 1, nz), (nx + 1, ny, nz + 1), (nx, ny + 1, nz + 1)]


def _iso(nx, ny, nz, dt=1e-3, pml=0, fs=0, c=(3.0, 1.0, 1.0), b=1.0, nt=4):
    from oracle import Oracle

    o = Oracle(nx, ny, nz, 1.0, 1.0, 1.0, dt, nt, mode=0, pml=pml, free_surface=fs)
    o.set_material_iso(np.full((nx, ny, nz), c[0], np.float32), np.full((nx, ny, nz), c[1], np.float32), c[2],
                       b, b, b)
    return o


def _ortho(nx, ny, nz, dt=1e-3, pml=0, fs=0, rng=None, nt=4, src_after_bc=1):
    from oracle import Oracle

    rng = rng or np.random.default_rng(0)
    o = Oracle(nx, ny, nz, 1.0, 1.0, 1.0, dt, nt, mode=1, pml=pml, free_surface=fs, src_after_bc=src_after_bc)
    sh = _shapes(nx, ny, nz)
    cij = [rng.uniform(1, 2, (nx, ny, nz)) for _ in range(6)] + [rng.uniform(0.3, 0.6, sh[s]) for s in (8, 7, 6)]
    buoy = [rng.uniform(0.5, 1.0, sh[s]) for s in range(3)]
    o.set_material_ortho(cij, buoy)
    return o


def _randomize(o, nx, ny, nz, rng):
    st = []
    for f, s in enumerate(_shapes(nx, ny, nz)):
        a = rng.uniform(0.5, 1, s) * rng.choice([-1.0, 1.0], s)  # |a| >= 0.5
        o.set_field(f, a)
        st.append(a)
    return st


def _ranges(nx, ny, nz):
    """Reference update ranges as boolean masks per field (App. of the
    oracle: OK4.cu:296-329,351-378; ANISO.cu:306-352,374-386)."""
    sh = _shapes(nx, ny, nz)
    m = [np.zeros(s, bool) for s in sh]
    m[F_VX][1:nx, 1:ny - 1, 1:nz - 1] = True
    m[F_VY][1:nx - 1, 1:ny, 1:nz - 1] = True
    m[F_VZ][1:nx - 1, 1:ny - 1, 1:nz] = True
    for f in (F_SXX, F_SYY, F_SZZ):
        m[f][:] = True
    m[F_SXY][1:nx, 1:ny, :nz] = True
    m[F_SXZ][1:nx, :ny, 1:nz] = True
    m[F_SYZ][:nx, 1:ny, 1:nz] = True
    return m


def test_iso_profile_damping_indices():
    """With dt -> 0 every update reduces to (damping factor) x (old value),
    which exposes the profile indexing of OK4.cu:307-309 (normal stress:
    0.5*(w[i]+w[i+1]) along its own axis), OK4.cu:323-329 (shear:
    wx[i]*wy[j]*wz[k]) and OK4.cu:361,368,375 (velocity: w[i-1] along its
    own axis), and the update ranges (cells outside keep their value)."""
    nx, ny, nz = 9, 8, 7
    rng = np.random.default_rng(1)
    o = _iso(nx, ny, nz, dt=1e-13, pml=1)
    wx, wy, wz = (rng.uniform(0.5, 1.0, n + 1) for n in (nx, ny, nz))
    o.set_pml_profiles(wx, wy, wz)
    before = _randomize(o, nx, ny, nz, rng)
    o.run(1)
    m = _ranges(nx, ny, nz)
    w32 = [np.float32(w).astype(np.float64) for w in (wx, wy, wz)]
    wx, wy, wz = w32
    hx, hy, hz = 0.5 * (wx[1:] + wx[:-1]), 0.5 * (wy[1:] + wy[:-1]), 0.5 * (wz[1:] + wz[:-1])
    exp = {
        F_SXX: hx[:, None, None] * np.ones((1, ny, nz)),
        F_SYY: hy[None, :, None] * np.ones((nx, 1, nz)),
        F_SZZ: hz[None, None, :] * np.ones((nx, ny, 1)),
        F_SXY: wx[:, None, None] * wy[None, :, None] * wz[None, None, :nz],
        F_SXZ: wx[:, None, None] * wy[None, :ny, None] * wz[None, None, :],
        F_SYZ: wx[:nx, None, None] * wy[None, :, None] * wz[None, None, :],
        F_VX: np.broadcast_to(np.concatenate([[1.0], wx[:nx]])[:, None, None], (nx + 1, ny, nz)),
        F_VY: np.broadcast_to(np.concatenate([[1.0], wy[:ny]])[None, :, None], (nx, ny + 1, nz)),
        F_VZ: np.broadcast_to(np.concatenate([[1.0], wz[:nz]])[None, None, :], (nx, ny, nz + 1)),
    }
    for f in range(9):
        after = o.get_field(f)
        ratio = after / before[f]
        e = np.where(m[f], exp[f], 1.0)
        assert np.allclose(ratio, e, rtol=1e-9, atol=0), f


@pytest.mark.parametrize("pml", [2, 3])
def test_ortho_cell_weight_on_old_value(pml):
    """ANISO.cu:306-352,374-386: the cell-centred weight w(i,j,k) multiplies
    the old value of every stress and velocity component (3-D array, or the
    separable product tx[i]*ty[j]*tz[k])."""
    nx, ny, nz = 8, 7, 9
    rng = np.random.default_rng(2)
    o = _ortho(nx, ny, nz, dt=1e-13, pml=pml, rng=rng)
    if pml == 2:
        w = rng.uniform(0.5, 1.0, (nx, ny, nz))
        o.set_pml_weights(w)
        w = np.float32(w).astype(np.float64)
    else:
        t = [rng.uniform(0.5, 1.0, n) for n in (nx, ny, nz)]
        o.set_pml_profiles(*t)
        t = [np.float32(a).astype(np.float64) for a in t]
        w = t[0][:, None, None] * t[1][None, :, None] * t[2][None, None, :]
    before = _randomize(o, nx, ny, nz, rng)
    o.run(1)
    m = _ranges(nx, ny, nz)
    for f, s in enumerate(_shapes(nx, ny, nz)):
        ratio = o.get_field(f) / before[f]
        e = np.ones(s)
        e[:nx, :ny, :nz] = w  # the velocity/shear cell (i,j,k) uses w(i,j,k)
        assert np.allclose(ratio, np.where(m[f], e, 1.0), rtol=1e-9, atol=0), f


@pytest.mark.parametrize("mode", [0, 1])
def test_bc1_invariants(mode):
    """BC1 (OK4.cu:279-282): after the stress pass Szz(k=0)=0,
    Sxz(1<=i<nx, j, 1) = Sxz(i,j,2)/3, Syz(i, 1<=j<ny, 1) = Syz(i,j,2)/3;
    the velocity pass never writes stress."""
    nx, ny, nz = 8, 9, 7
    rng = np.random.default_rng(3)
    o = _iso(nx, ny, nz, pml=0, fs=1) if mode == 0 else _ortho(nx, ny, nz, pml=0, fs=1, rng=rng)
    _randomize(o, nx, ny, nz, rng)
    o.run(1)
    szz, sxz, syz = o.get_field(F_SZZ), o.get_field(F_SXZ), o.get_field(F_SYZ)
    assert np.all(szz[:, :, 0] == 0)
    assert np.allclose(sxz[1:nx, :, 1], sxz[1:nx, :, 2] / 3, rtol=1e-15)
    assert np.allclose(syz[:, 1:ny, 1], syz[:, 1:ny, 2] / 3, rtol=1e-15)


@pytest.mark.parametrize("mode", [0, 1])
def test_vacuum_layer_shear_plane(mode):
    """Vacuum-layer top (free_surface=2, diff_BC2_OK3/FastWaveED_GPU3D_v1.cu:
    309-314): after one step from a random state, every shear stress on the
    plane k=1 is old + dt*1e-5*(strain-rate sum) from the old velocities (no
    damping here), every other stress equals the step without a free
    surface, and BC1 is not applied (Szz at k=0 is not zeroed)."""
    nx, ny, nz = 8, 9, 7
    dt = 1e-3
    runs = {}
    for fs in (0, 2):
        rng = np.random.default_rng(7)
        o = _iso(nx, ny, nz, dt=dt, pml=0, fs=fs) if mode == 0 else _ortho(nx, ny, nz, dt=dt, pml=0, fs=fs, rng=rng)
        before = _randomize(o, nx, ny, nz, rng)
        o.run(1)
        runs[fs] = [o.get_field(f) for f in range(9)]
    vx, vy, vz = before[F_VX], before[F_VY], before[F_VZ]
    k = 1
    exp_xy = before[F_SXY][1:nx, 1:ny, k] + dt * 1e-5 * ((vy[1:nx, 1:ny, k] - vy[0:nx - 1, 1:ny, k]) +
                                                         (vx[1:nx, 1:ny, k] - vx[1:nx, 0:ny - 1, k]))
    exp_xz = before[F_SXZ][1:nx, :ny, k] + dt * 1e-5 * ((vz[1:nx, :, k] - vz[0:nx - 1, :, k]) +
                                                        (vx[1:nx, :, k] - vx[1:nx, :, k - 1]))
    exp_yz = before[F_SYZ][:nx, 1:ny, k] + dt * 1e-5 * ((vy[:, 1:ny, k] - vy[:, 1:ny, k - 1]) +
                                                        (vz[:, 1:ny, k] - vz[:, 0:ny - 1, k]))
    vac, free = runs[2], runs[0]
    assert np.allclose(vac[F_SXY][1:nx, 1:ny, k], exp_xy, rtol=1e-13, atol=0)
    assert np.allclose(vac[F_SXZ][1:nx, :ny, k], exp_xz, rtol=1e-13, atol=0)
    assert np.allclose(vac[F_SYZ][:nx, 1:ny, k], exp_yz, rtol=1e-13, atol=0)
    for f in (F_SXX, F_SYY, F_SZZ):
        assert np.array_equal(vac[f], free[f]), f
    for f in (F_SXY, F_SXZ, F_SYZ):
        a, b = vac[f].copy(), free[f].copy()
        assert not np.allclose(a[:, :, k], b[:, :, k])  # the medium's modulus is not 1e-5
        a[:, :, k] = 0
        b[:, :, k] = 0
        assert np.array_equal(a, b), f
    assert np.any(vac[F_SZZ][:, :, 0] != 0)


@pytest.mark.parametrize("mode", [0, 1])
def test_injection_from_rest(mode):
    """From rest, one step leaves exactly coef*Src[0] in each injected stress
    cell and zero in every other stress cell (OK4.cu:336-342)."""
    from elanfmt_amd import fwe

    nx, ny, nz = 12, 11, 10
    o = _iso(nx, ny, nz) if mode == 0 else _ortho(nx, ny, nz)
    o.set_stf([2.5, 0.0, 0.0, 0.0])
    m = [1.0, -2.0, 3.0, 4.0, -5.0, 6.0]
    layout = fwe.SRC_LAYOUT_OK4 if mode == 0 else fwe.SRC_LAYOUT_ANISO
    inj = fwe.mt_injections(layout, 4, 3, 3, m)
    o.set_injections(inj)
    o.run(1)
    for comp, i, j, k, _, coef in inj:
        a = o.get_field(F_SXX + comp)
        assert a[i, j, k] == pytest.approx(2.5 * coef, rel=1e-15)
        a[i, j, k] = 0
        assert np.all(a == 0)


def test_receivers_sample_the_named_cell():
    """Receivers read V(ix,iy,iz) of their component after the velocity
    update of the step (OK4.cu:383-386)."""
    nx, ny, nz = 7, 6, 5
    o = _iso(nx, ny, nz, c=(0.0, 0.0, 0.0), b=0.0, nt=3)
    for f in range(3):
        s = _shapes(nx, ny, nz)[f]
        i, j, k = np.meshgrid(*[np.arange(n) for n in s], indexing="ij")
        o.set_field(f, (f + 1) * 1e6 + i + 100 * j + 10000 * k)
    rec = [(0, 3, 2, 1, 0), (1, 1, 5, 3, 0), (2, 6, 0, 4, 0), (2, 2, 3, 2, 0)]
    o.set_receivers(rec)
    o.run(2)
    tr = o.traces()
    for r, (c, i, j, k, _) in enumerate(rec):
        assert tr[r, 0] == pytest.approx((c + 1) * 1e6 + i + 100 * j + 10000 * k, rel=1e-14)
        assert tr[r, 1] == tr[r, 0] and tr[r, 2] == 0  # third step not run yet


def _homog_source_run(nx, ny, nz, inj, nt=30, mode=0, rec=None, pml=0):
    from elanfmt_amd import workloads as W
    from oracle import Oracle

    vp, vs, rho, dx = 3000.0, 1700.0, 2500.0, 10.0
    dt = W.cfl_dt(vp, dx, dx, dx, 0.5)
    o = Oracle(nx, ny, nz, dx, dx, dx, dt, nt, mode=mode, pml=pml)
    # scalars rounded as the float32 arrays of the ortho law are
    c11, c12, c66 = rho * vp ** 2, rho * (vp ** 2 - 2 * vs ** 2), float(np.float32(rho * vs ** 2))
    b = float(np.float32(1 / rho))
    if mode == 0:
        o.set_material_iso(np.full((nx, ny, nz), c11, np.float32), np.full((nx, ny, nz), c12, np.float32), c66,
                           b, b, b)
    else:
        sh = _shapes(nx, ny, nz)
        cij = [np.full((nx, ny, nz), v, np.float32) for v in (c11, c11, c11, c12, c12, c12)]
        cij += [np.full(sh[s], c66, np.float32) for s in (8, 7, 6)]
        o.set_material_ortho(cij, [np.full(sh[s], 1 / rho, np.float32) for s in range(3)])
    o.set_stf(W.ricker(nt, dt, 40.0, 0.03))
    o.set_injections(inj)
    if rec:
        o.set_receivers(rec)
    o.run(nt)
    return o


def test_iso_equals_ortho_with_isotropic_constants():
    """The iso law (OK4) and the ortho law (ANISO) agree when C11=C22=C33,
    C12=C13=C23, C44=C55=C66 and there is no damping."""
    n = 14
    inj = [(0, 6, 6, 6, 0, 1e9), (3, 7, 7, 6, 0, -5e8), (4, 6, 7, 7, 0, 3e8), (5, 7, 6, 7, 0, 2e8)]
    a = _homog_source_run(n, n, n, inj, mode=0)
    b = _homog_source_run(n, n, n, inj, mode=1)
    for f in range(9):
        assert rel_l2(a.get_field(f), b.get_field(f)) < 1e-13, f


@pytest.mark.parametrize("swap", ["xy", "xz"])
def test_mirror_symmetry(swap):
    """Homogeneous isotropic cube: a Mxx source and a Myy (or Mzz) source at
    the same cell give wavefields that are mirror images under the x<->y (or
    x<->z) swap; pins the three component paths against transposition."""
    n = 13
    c = 6
    if swap == "xy":
        a = _homog_source_run(n, n, n, [(0, c, c, c, 0, 1e9)])
        b = _homog_source_run(n, n, n, [(1, c, c, c, 0, 1e9)])
        perm = (1, 0, 2)
        pairs = [(F_VX, F_VY), (F_VZ, F_VZ), (F_SXX, F_SYY), (F_SZZ, F_SZZ), (F_SXY, F_SXY), (F_SXZ, F_SYZ)]
    else:
        a = _homog_source_run(n, n, n, [(0, c, c, c, 0, 1e9)])
        b = _homog_source_run(n, n, n, [(2, c, c, c, 0, 1e9)])
        perm = (2, 1, 0)
        pairs = [(F_VX, F_VZ), (F_VY, F_VY), (F_SXX, F_SZZ), (F_SYY, F_SYY), (F_SXZ, F_SXZ), (F_SXY, F_SYZ)]
    for fa, fb in pairs:
        A = a.get_field(fa)
        B = np.transpose(b.get_field(fb), perm)
        assert np.abs(A).max() > 0
        assert rel_l2(A, B) < 1e-13, (fa, fb)


def test_superposition():
    """Linearity in the moment tensor: run(M1 + M2) = run(M1) + run(M2)."""
    from elanfmt_amd import fwe

    n = 12
    m1, m2 = [1e9, 0, -3e8, 2e8, 0, 5e8], [0, 7e8, 1e8, 0, -4e8, 0]
    rec = [(c, 3, 4, 5, 0) for c in range(3)] + [(c, 8, 2, 6, 0) for c in range(3)]
    runs = []
    for m in (m1, m2, list(np.add(m1, m2))):
        inj = fwe.mt_injections(fwe.SRC_LAYOUT_OK4, 4, 4, 4, m)
        runs.append(_homog_source_run(n, n, n, inj, rec=rec).traces())
    assert rel_l2(runs[0] + runs[1], runs[2]) < 1e-13
    assert np.abs(runs[2]).max() > 0


def test_zero_source_stays_zero_and_cfl_limit():
    """No source -> fields stay exactly zero; the staggered O(2,2) scheme is
    stable for vp*dt*sqrt(3)/dx < 1 and grows without bound above it."""
    from oracle import Oracle

    n = 10
    o = _homog_source_run(n, n, n, [])
    assert all(np.all(o.get_field(f) == 0) for f in range(9))
    growth = []
    for cfl in (0.95, 1.05):
        vp, rho, dx = 3000.0, 2500.0, 10.0
        dt = cfl * dx / (vp * np.sqrt(3))
        o = Oracle(n, n, n, dx, dx, dx, dt, 4, mode=0, pml=0)
        o.set_material_iso(np.full((n, n, n), rho * vp ** 2, np.float32), np.full((n, n, n), rho * vp ** 2 / 3,
                           np.float32), rho * vp ** 2 / 3, 1 / rho, 1 / rho, 1 / rho)
        rng = np.random.default_rng(4)
        for f in range(3):
            o.set_field(f, rng.uniform(-1, 1, _shapes(n, n, n)[f]))
        e0 = o.sum_sq_velocity()
        o.run(400)
        growth.append(o.sum_sq_velocity() / e0)
    assert growth[0] < 10 and growth[1] > 1e6


def _single_element_medium(nx, ny, nz, where):
    """Uniform ortho medium whose nine moduli and three buoyancies have
    distinct background values, with one element of every array raised by
    its own factor at staggered index `where`."""
    sh = _shapes(nx, ny, nz)
    # C11 C22 C33 C12 C13 C23 on cells, C44yz C55xz C66xy on edges
    base_c = [2.0, 2.5, 3.0, 0.5, 0.7, 0.9, 1.1, 1.3, 1.7]
    base_b = [0.8, 0.6, 0.4]
    cell_of_c = [(nx, ny, nz)] * 6 + [sh[F_SYZ], sh[F_SXZ], sh[F_SXY]]
    cij = [np.full(s, v) for s, v in zip(cell_of_c, base_c)]
    buoy = [np.full(sh[f], v) for f, v in zip((F_VX, F_VY, F_VZ), base_b)]
    for n, a in enumerate(cij):
        a[where] *= 1.5 + 0.25 * n
    for n, a in enumerate(buoy):
        a[where] *= 3.0 + 0.5 * n
    return cij, buoy


def test_ortho_heterogeneous_staggered_medium_indexing():
    """Which element of each staggered medium array a cell update reads.

    ANISO.cu:327-352: Sxy(i,j,k) reads C66xy at its own edge index
    i+j(nx+1)+k(nx+1)(ny+1) (:333), Sxz C55xz at i+j(nx+1)+k(nx+1)ny (:342),
    Syz C44yz at i+j nx+k nx(ny+1) (:351); ANISO.cu:374-385: Vx, Vy, Vz read
    vrhox11het/vrhoy11het/vrhoz11het at the velocity's own staggered index.
    Each array holds a uniform background value with one element at (3,4,2)
    raised by its own factor.  A linear velocity field (constant strain
    rates) then makes every shear stress dt*C*(rate) with C the background,
    except exactly the element at (3,4,2) of the stress whose modulus was
    raised; a linear stress field does the same for the velocities through
    the buoyancies.  Reading an edge array with the cell index, or swapping
    two moduli or buoyancies, moves or rescales the changed element."""
    nx, ny, nz = 9, 8, 7
    dx, dy, dz, dt = 2.0, 3.0, 5.0, 1e-3
    where = (3, 4, 2)
    from oracle import Oracle

    cij, buoy = _single_element_medium(nx, ny, nz, where)
    sh = _shapes(nx, ny, nz)
    o = Oracle(nx, ny, nz, dx, dy, dz, dt, 2, mode=1, pml=0, free_surface=0)
    o.set_material_ortho(cij, buoy)
    # float32 images of the medium, as the oracle stores them
    cij = [np.float32(a).astype(np.float64) for a in cij]
    buoy = [np.float32(a).astype(np.float64) for a in buoy]

    # --- shear moduli: V linear, S zero ---------------------------------
    ax, az = 0.3, -0.2    # Vx = ax*y + az*z
    by_, bz_ = 0.7, 0.1   # Vy = by_*x + bz_*z
    cx, cy = -0.4, 0.9    # Vz = cx*x + cy*y
    gx = lambda s: np.arange(s[0])[:, None, None] * dx
    gy = lambda s: np.arange(s[1])[None, :, None] * dy
    gz = lambda s: np.arange(s[2])[None, None, :] * dz
    o.set_field(F_VX, np.broadcast_to(ax * gy(sh[F_VX]) + az * gz(sh[F_VX]), sh[F_VX]))
    o.set_field(F_VY, np.broadcast_to(by_ * gx(sh[F_VY]) + bz_ * gz(sh[F_VY]), sh[F_VY]))
    o.set_field(F_VZ, np.broadcast_to(cx * gx(sh[F_VZ]) + cy * gy(sh[F_VZ]), sh[F_VZ]))
    for f in range(F_SXX, F_SYZ + 1):
        o.set_field(f, np.zeros(sh[f]))
    o.run(1)
    m = _ranges(nx, ny, nz)
    # (field, modulus array index, background modulus, strain-rate sum)
    for f, ci, rate in ((F_SXY, 8, by_ + ax), (F_SXZ, 7, cx + az), (F_SYZ, 6, bz_ + cy)):
        got = o.get_field(f)
        bg = dt * cij[ci].flat[0 if where != (0, 0, 0) else 1] * rate
        want = np.where(m[f], bg, 0.0)
        want[where] = dt * cij[ci][where] * rate
        assert np.allclose(got, want, rtol=1e-12, atol=1e-18), f
        # exactly one element differs from the background
        changed = np.argwhere(np.abs(got - bg) > 1e-9 * abs(bg) + 1e-18)
        changed = [tuple(c) for c in changed if m[f][tuple(c)]]
        assert changed == [where], (f, changed)
    # the normal stresses see no strain: they stay zero
    for f in (F_SXX, F_SYY, F_SZZ):
        assert np.all(o.get_field(f) == 0.0), f

    # --- buoyancies: V zero, normal stresses linear ----------------------
    o.reset()
    p, q, r = 0.6, -0.8, 1.2  # Sxx = p*x, Syy = q*y, Szz = r*z
    for f in (F_VX, F_VY, F_VZ, F_SXY, F_SXZ, F_SYZ):
        o.set_field(f, np.zeros(sh[f]))
    o.set_field(F_SXX, np.broadcast_to(p * gx(sh[F_SXX]), sh[F_SXX]))
    o.set_field(F_SYY, np.broadcast_to(q * gy(sh[F_SYY]), sh[F_SYY]))
    o.set_field(F_SZZ, np.broadcast_to(r * gz(sh[F_SZZ]), sh[F_SZZ]))
    o.run(1)  # stress pass: zero strain rates leave S unchanged (no damping)
    for f, bi, grad in ((F_VX, 0, p), (F_VY, 1, q), (F_VZ, 2, r)):
        got = o.get_field(f)
        want = np.where(m[f], buoy[bi] * grad * dt, 0.0)
        assert np.allclose(got, want, rtol=1e-12, atol=1e-18), f
        bg = buoy[bi].flat[0] * grad * dt
        changed = [tuple(c) for c in np.argwhere(np.abs(got - bg) > 1e-9 * abs(bg)) if m[f][tuple(c)]]
        assert changed == [where], (f, changed)
