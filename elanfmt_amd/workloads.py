"""Synthetic workloads of the benchmark configurations, assembled from the C
generators of libfwe.so (include/fwe_gen.h).

Each builder returns a Workload: the solver configuration plus every input
array in the reference's staggered shapes (numpy, x-fastest = Fortran order)
and the injection/receiver lists.  ``apply`` loads a Workload into anything
with the fwe.Solver call pattern (the HIP solver, or the oracle in tests).

Source scaling: the reference adds M_ij * Src[it] to the stress
(OK4.cu:336-342).  Here Src[it] = r(t) * dt / (dx*dy*dz) with r a unit-peak
Ricker moment-rate, and the injection coefficients are the moment tensor in
N*m, so one step adds the moment released in dt, spread over one cell.
"""
from __future__ import annotations

import ctypes as C
import math
from dataclasses import dataclass, field

import numpy as np

from . import fwe


@dataclass
class Workload:
    name: str
    nx: int
    ny: int
    nz: int
    dx: float
    dy: float
    dz: float
    dt: float
    nt: int
    mode: int
    pml: int
    free_surface: int = 0
    nbatch: int = 1
    # iso medium
    c11: np.ndarray | None = None
    c12: np.ndarray | None = None
    c66: float = 0.0
    bscal: tuple = (0.0, 0.0, 0.0)
    # ortho medium
    cij: list | None = None
    buoy: list | None = None
    # damping
    prof: tuple | None = None
    w3: np.ndarray | None = None
    # convolutional PML: layer cells per face and the three axis tables
    cpml_layer: int = 0
    cpml: tuple | None = None
    # source and receivers
    stf: list = field(default_factory=list)  # one per batch
    injections: list = field(default_factory=list)
    receivers: list = field(default_factory=list)
    info: dict = field(default_factory=dict)

    @property
    def cells(self):
        return self.nx * self.ny * self.nz


def _fa(n):
    return np.zeros(n, np.float32)


def ricker(nt, dt, f0, t0, amp=1.0):
    out = _fa(nt)
    fwe.lib().fwe_gen_ricker(nt, dt, f0, t0, amp, fwe.fptr(out))
    return out


def mt_sdr(strike, dip, rake, m0):
    m = np.zeros(6)
    fwe.lib().fwe_gen_mt_sdr(strike, dip, rake, m0, fwe.dptr(m))
    return m


def mt_lune_random(seed, m0):
    r = fwe.RNG()
    fwe.lib().fwe_rng_seed(C.byref(r), seed)
    m = np.zeros(6)
    g, d = C.c_double(), C.c_double()
    fwe.lib().fwe_gen_mt_lune_random(C.byref(r), m0, fwe.dptr(m), C.byref(g), C.byref(d))
    return m, g.value, d.value


def pml_profile(n, L, dt, R=1e-8, lo=1, hi=1, top_undamped=0):
    w = _fa(n + 1)
    fwe.lib().fwe_gen_pml_profile(n, L, R, dt, lo, hi, top_undamped, fwe.fptr(w))
    return w


def cell_profile(w):
    n = len(w) - 1
    t = _fa(n)
    fwe.lib().fwe_gen_cell_profile(n, fwe.fptr(np.ascontiguousarray(w, np.float32)), fwe.fptr(t))
    return t


def taper(n, m_lo, m_hi, rate):
    t = _fa(n)
    fwe.lib().fwe_gen_taper(n, m_lo, m_hi, rate, fwe.fptr(t))
    return t


def cpml_table(n, L, h, dt, vmax, R=1e-3, kappa_max=1.0, alpha_max=0.0, power=2.0, lo=1, hi=1):
    """Convolutional PML coefficients of one axis (fwe_gen_cpml_table):
    6 x (n+1) floats, b, a, 1/kappa at integer then half nodes."""
    t = _fa(6 * (n + 1))
    if fwe.lib().fwe_gen_cpml_table(n, L, h, dt, vmax, R, kappa_max, alpha_max, power, lo, hi, fwe.fptr(t)) != 0:
        raise ValueError("bad CPML parameters")
    return t.reshape(6, n + 1)


def medium_vmax(wl: Workload):
    """Fastest wave speed bound of a workload's medium: sqrt(max modulus x
    max buoyancy) (moduli and buoyancy are the physical values)."""
    if wl.mode == fwe.MODE_ISO:
        return float(np.sqrt(np.max(wl.c11) * max(wl.bscal)))
    return float(np.sqrt(max(float(np.max(c)) for c in wl.cij[:3]) * max(float(np.max(b)) for b in wl.buoy)))


def use_cpml(wl: Workload, layer, R=1e-4, kappa_max=1.0, alpha_max=None, power=2.0, vmax=None):
    """Switch a workload's damping to a convolutional PML of `layer` cells on
    every face (the top end of z undamped under a free surface).  alpha_max
    defaults to pi f0 of the workload's wavelet (0 when unknown)."""
    if alpha_max is None:
        alpha_max = np.pi * float(wl.info.get("f0", 0.0))
    if vmax is None:
        vmax = wl.info.get("vmax") or medium_vmax(wl)
    top = 0 if wl.free_surface else 1
    wl.cpml = (cpml_table(wl.nx, layer, wl.dx, wl.dt, vmax, R, kappa_max, alpha_max, power),
               cpml_table(wl.ny, layer, wl.dy, wl.dt, vmax, R, kappa_max, alpha_max, power),
               cpml_table(wl.nz, layer, wl.dz, wl.dt, vmax, R, kappa_max, alpha_max, power, lo=top))
    wl.cpml_layer = layer
    wl.pml = fwe.PML_CPML
    wl.prof = None
    wl.w3 = None
    return wl


def weights3d(nx, ny, nz, mx, my, mz_lo, mz_hi, rate):
    w = _fa(nx * ny * nz)
    fwe.lib().fwe_gen_weights3d(nx, ny, nz, mx, mx, my, my, mz_lo, mz_hi, rate, fwe.fptr(w))
    return w.reshape((nx, ny, nz), order="F")


def stagger_buoyancy(bcell):
    nx, ny, nz = bcell.shape
    b = fwe.f32(bcell)
    bx, by, bz = _fa((nx + 1) * ny * nz), _fa(nx * (ny + 1) * nz), _fa(nx * ny * (nz + 1))
    fwe.lib().fwe_gen_stagger_buoyancy(nx, ny, nz, fwe.fptr(b), fwe.fptr(bx), fwe.fptr(by), fwe.fptr(bz))
    return [bx.reshape((nx + 1, ny, nz), order="F"), by.reshape((nx, ny + 1, nz), order="F"),
            bz.reshape((nx, ny, nz + 1), order="F")]


def harmonic_edges(c44, c55, c66):
    nx, ny, nz = c44.shape
    a, b, c = fwe.f32(c44), fwe.f32(c55), fwe.f32(c66)
    gyz, gxz, gxy = _fa(nx * (ny + 1) * (nz + 1)), _fa((nx + 1) * ny * (nz + 1)), _fa((nx + 1) * (ny + 1) * nz)
    fwe.lib().fwe_gen_harmonic_edges(nx, ny, nz, fwe.fptr(a), fwe.fptr(b), fwe.fptr(c), fwe.fptr(gyz),
                                     fwe.fptr(gxz), fwe.fptr(gxy))
    return [gyz.reshape((nx, ny + 1, nz + 1), order="F"), gxz.reshape((nx + 1, ny, nz + 1), order="F"),
            gxy.reshape((nx + 1, ny + 1, nz), order="F")]


def layered_ortho(nx, ny, nz, nl, seed, vp=(4000.0, 6500.0), ratio=(1.65, 1.85), rho=(2400.0, 2900.0),
                  aniso=0.15):
    """Cell Cij (9) + cell buoyancy of a seeded layered orthorhombic medium."""
    n = nx * ny * nz
    cells = [_fa(n) for _ in range(9)]
    b = _fa(n)
    tops = np.zeros(nl, np.int32)
    vmax = C.c_double()
    st = fwe.lib().fwe_gen_layered_ortho(nx, ny, nz, nl, seed, vp[0], vp[1], ratio[0], ratio[1], rho[0],
                                         rho[1], aniso, (fwe._FP * 9)(*[fwe.fptr(c) for c in cells]),
                                         fwe.fptr(b), tops.ctypes.data_as(fwe._IP), C.byref(vmax))
    if st != 0:
        raise ValueError("layered medium generation failed")
    sh = (nx, ny, nz)
    return [c.reshape(sh, order="F") for c in cells], b.reshape(sh, order="F"), tops, vmax.value


def ortho_from_cells(cells, bcell):
    """Reference ortho inputs (ANISO.cu:411-457) from cell moduli:
    C11..C23 cell arrays, harmonic edge C44yz/C55xz/C66xy, staggered buoyancy."""
    gyz, gxz, gxy = harmonic_edges(cells[6], cells[7], cells[8])
    return list(cells[:6]) + [gyz, gxz, gxy], stagger_buoyancy(bcell)


def cfl_dt(vmax, dx, dy, dz, cfl=0.5):
    return fwe.lib().fwe_gen_cfl_dt(vmax, dx, dy, dz, cfl)


def apply(solver, wl: Workload):
    """Load a workload into a Solver-like object."""
    if wl.mode == fwe.MODE_ISO:
        solver.set_material_iso(wl.c11, wl.c12, wl.c66, *wl.bscal)
    else:
        solver.set_material_ortho(wl.cij, wl.buoy)
    if wl.pml in (fwe.PML_PROFILE, fwe.PML_SEPARABLE):
        solver.set_pml_profiles(*wl.prof)
    elif wl.pml == fwe.PML_WEIGHTS3D:
        solver.set_pml_weights(wl.w3)
    elif wl.pml == fwe.PML_CPML:
        solver.set_cpml(wl.cpml_layer, *wl.cpml)
    for b, s in enumerate(wl.stf):
        solver.set_stf(s, b)
    solver.set_injections(wl.injections)
    solver.set_receivers(wl.receivers)


def make_solver(wl: Workload, device=0):
    s = fwe.Solver(wl.nx, wl.ny, wl.nz, wl.dx, wl.dy, wl.dz, wl.dt, wl.nt, mode=wl.mode, pml=wl.pml,
                   free_surface=wl.free_surface, nbatch=wl.nbatch, device=device)
    apply(s, wl)
    return s


def _rec_grid(xs, ys, iz, comps=(0, 1, 2), batch=0):
    return [(c, int(x), int(y), int(iz), batch) for y in ys for x in xs for c in comps]


def iso_homogeneous(n=128, dx=50.0, vp=6000.0, vs=3464.0, rho=2700.0, L=20, nt=600, f0=4.0,
                    sdr=(30.0, 60.0, 90.0), m0=1e15, nrec_side=4, rec_depth=10, free_surface=0):
    """Isotropic homogeneous cube (default: 128^3, 20-cell PML on all faces,
    double couple strike 30 dip 60 rake 90 at the centre, Ricker 4 Hz,
    600 steps, 16 receivers x (vx,vy,vz) on the plane 10 cells below the top
    PML)."""
    nx = ny = nz = n
    dt = cfl_dt(vp, dx, dx, dx, 0.5)
    c11 = np.full((nx, ny, nz), rho * vp * vp, np.float32)
    c12 = np.full((nx, ny, nz), rho * (vp * vp - 2 * vs * vs), np.float32)
    c66 = rho * vs * vs
    b = 1.0 / rho
    prof = tuple(pml_profile(m, L, dt) for m in (nx, ny, nz))
    stf = ricker(nt, dt, f0, 1.2 / f0) * np.float32(dt / (dx * dx * dx))
    m = mt_sdr(*sdr, m0)
    c = n // 2
    inj = fwe.mt_injections(fwe.SRC_LAYOUT_OK4, c - 1, c - 1, c - 1, m)
    lo, hi = L + 2 * (n // 16), n - L - 2 * (n // 16)
    xs = np.linspace(lo, hi, nrec_side).round().astype(int)
    rec = _rec_grid(xs, xs, L + rec_depth)
    return Workload(f"iso{n}", nx, ny, nz, dx, dx, dx, dt, nt, fwe.MODE_ISO, fwe.PML_PROFILE,
                    free_surface=free_surface, c11=c11, c12=c12, c66=c66, bscal=(b, b, b), prof=prof,
                    stf=[stf], injections=inj, receivers=rec,
                    info=dict(vp=vp, vs=vs, rho=rho, mt=m.tolist(), f0=f0, pml_cells=L))


def ortho_layered(nx=384, ny=384, nz=192, dx=50.0, nl=5, seed=1, L=24, nt=3000, f0=3.0, m0=1e15,
                  nrec_side=8, pml=fwe.PML_SEPARABLE, nbatch=1, elementary=False, comps=(0, 1, 2)):
    """Layered orthorhombic box (default: 384x384x192, 5 seeded layers,
    24-cell damping on all faces, lune-uniform MT, Ricker 3 Hz, 3000 steps,
    8x8 receivers x (vx,vy,vz) two cells below the top damping layer).
    elementary=True: nbatch = 6 wavefields, one per elementary tensor
    (Mxx, Myy, Mzz, Mxy, Mxz, Myz), sharing the medium."""
    cells, bcell, tops, vmax = layered_ortho(nx, ny, nz, nl, seed)
    cij, buoy = ortho_from_cells(cells, bcell)
    dt = cfl_dt(vmax, dx, dx, dx, 0.5)
    if pml == fwe.PML_SEPARABLE:
        prof = tuple(cell_profile(pml_profile(m, L, dt)) for m in (nx, ny, nz))
        w3 = None
    else:
        prof = None
        t = [cell_profile(pml_profile(m, L, dt)) for m in (nx, ny, nz)]
        w3 = (t[0][:, None, None] * t[1][None, :, None] * t[2][None, None, :]).astype(np.float32)
    stf = ricker(nt, dt, f0, 1.2 / f0) * np.float32(dt / (dx * dx * dx))
    m, g, d = mt_lune_random(seed * 7919 + 1, m0)
    xs_, ys_, zs_ = nx // 2 - 1, ny // 2 - 1, nz // 2 - 1
    if elementary:
        nbatch = 6
        inj = []
        for e in range(6):
            me = np.zeros(6)
            me[e] = m0
            inj += [t for t in fwe.mt_injections(fwe.SRC_LAYOUT_OK4, xs_, ys_, zs_, me, batch=e) if t[5] != 0]
    else:
        inj = []
        for bb in range(nbatch):
            inj += fwe.mt_injections(fwe.SRC_LAYOUT_OK4, xs_, ys_, zs_, m, batch=bb)
    mx = 8 if nx > 64 else 2
    xs = np.linspace(L + mx, nx - L - mx, nrec_side).round().astype(int)
    ys = np.linspace(L + mx, ny - L - mx, nrec_side).round().astype(int)
    rec = []
    for bb in range(nbatch):
        rec += _rec_grid(xs, ys, L + 2, comps=comps, batch=bb)
    return Workload(f"ortho{nx}x{ny}x{nz}", nx, ny, nz, dx, dx, dx, dt, nt, fwe.MODE_ORTHO, pml,
                    nbatch=nbatch, cij=cij, buoy=buoy, prof=prof, w3=w3, stf=[stf] * nbatch,
                    injections=inj, receivers=rec,
                    info=dict(layer_tops=tops.tolist(), vmax=vmax, mt=m.tolist(), gamma=g, delta=d, f0=f0,
                              pml_cells=L, seed=seed))


def config(idx, **kw):
    """The named benchmark configurations at full size."""
    if idx == 0:
        return iso_homogeneous(**kw)
    if idx == 1:
        return ortho_layered(**kw)
    if idx == 2:
        kw.setdefault("nx", 512), kw.setdefault("ny", 512), kw.setdefault("nz", 256)
        kw.setdefault("nt", 4000), kw.setdefault("nrec_side", 16)
        return ortho_layered(elementary=True, **kw)
    if idx == 3:
        return uneven_batch(**kw)
    raise ValueError(idx)


# algorithmic bytes per cell-update of each kernel formulation: every
# streamed 3-D array counted once per pass (read and/or write), FP32.
def alg_bytes(mode, pml, nbatch=1, fused=False, medium=fwe.MEDIUM_ARRAYS, batch_tile=None):
    """(stress_bytes, velocity_bytes) per cell per step, all wavefields; for
    the single-pass kernel the whole step is in the first entry.  medium:
    MEDIUM_ARRAYS, the arrays (ortho twelve, iso c11 and c12); MEDIUM_TABLE,
    one index byte per cell and pass that reads the medium (the table
    entries stay in the caches; the iso velocity pass reads no medium);
    MEDIUM_ISOTROPIC (ortho law), C11 and C12 for the six normal moduli.
    batch_tile:
    wavefields per workgroup (default all); the medium is read once per
    workgroup group of wavefields, nbatch / batch_tile times per cell."""
    w = 1 if pml == fwe.PML_WEIGHTS3D else 0
    reads = nbatch // (batch_tile or nbatch)  # medium reads per cell and pass
    if medium == fwe.MEDIUM_TABLE and not fused and mode == fwe.MODE_ISO:
        # stress: R V3 S6 + index byte, W S6 ; velocity: R S6 V3, W V3
        return 4 * nbatch * 15 + reads, 4 * nbatch * 12
    if medium == fwe.MEDIUM_TABLE and not fused:
        # stress: R V3 S6 [w] + index byte, W S6 ; velocity: R S6 V3 [w] + index byte, W V3
        return 4 * (nbatch * 15 + w * reads) + reads, 4 * (nbatch * 12 + w * reads) + reads
    if fused:
        # single-pass sweep: fields R9 W9 per wavefield and the medium once
        # per wavefield (each wavefront advances one wavefield): the index
        # byte (table), iso c11 c12, ortho C9 b3, isotropic arrays C11 C12
        # C44 C55 C66 b3; plus the 3-D weight
        if medium == fwe.MEDIUM_TABLE:
            return nbatch * (4 * (18 + w) + 1), 0
        if medium == fwe.MEDIUM_ISOTROPIC:
            return 4 * nbatch * (18 + 8 + w), 0
        med = 2 if mode == fwe.MODE_ISO else 12 + w
        return 4 * nbatch * (18 + med), 0
    if mode == fwe.MODE_ISO:
        # stress: R V3 S6 c11 c12, W S6 ; velocity: R S6 V3, W V3
        return 4 * (nbatch * 15 + 2 * reads), 4 * (nbatch * 12)
    if medium == fwe.MEDIUM_ISOTROPIC:
        # stress: R V3 S6 C11 C12 C44 C55 C66 [w], W S6 ; velocity as below
        return 4 * (nbatch * 15 + (5 + w) * reads), 4 * (nbatch * 12 + (3 + w) * reads)
    # stress: R V3 S6 C9 [w], W S6 ; velocity: R S6 V3 b3 [w], W V3
    return 4 * (nbatch * 15 + (9 + w) * reads), 4 * (nbatch * 12 + (3 + w) * reads)


def random_field(nx, ny, nz, corr, seed):
    out = _fa(nx * ny * nz)
    fwe.lib().fwe_gen_random_field(nx, ny, nz, corr, seed, fwe.fptr(out))
    return out.reshape((nx, ny, nz), order="F")


def iso_random(n, seed, vp0=5500.0, vs0=3175.0, rho0=2650.0, pert=0.08, corr=10.0, dx=50.0, L=16, nt=1500,
               f0=None, nrec_side=4):
    """Heterogeneous isotropic cube: background vp/vs/rho times (1 + pert*g),
    g a unit-variance Gaussian random field of correlation length `corr`
    cells; run through the ortho law (C11=C22=C33, C12=C13=C23, harmonic
    edge shear moduli, staggered buoyancy) as the reference's iso-het
    station variants are.  Own lune MT, Ricker f0 ~ U[2,5] Hz, separable
    damping on all faces."""
    rng = np.random.default_rng(seed)
    g = random_field(n, n, n, corr, seed * 7 + 3).astype(np.float64)
    s = 1.0 + pert * g
    vp, vs, rho = vp0 * s, vs0 * s, rho0 * s
    c11 = (rho * vp * vp).astype(np.float32)
    c12 = (rho * (vp * vp - 2 * vs * vs)).astype(np.float32)
    c44 = (rho * vs * vs).astype(np.float32)
    cij, buoy = ortho_from_cells([c11, c11, c11, c12, c12, c12, c44, c44, c44], (1.0 / rho).astype(np.float32))
    dt = cfl_dt(float(vp.max()), dx, dx, dx, 0.5)
    f0 = float(rng.uniform(2.0, 5.0)) if f0 is None else f0
    prof = tuple(cell_profile(pml_profile(n, L, dt)) for _ in range(3))
    stf = ricker(nt, dt, f0, 1.2 / f0) * np.float32(dt / dx ** 3)
    m, gam, dl = mt_lune_random(seed * 131 + 7, 1e15)
    c = n // 2 - 1
    inj = fwe.mt_injections(fwe.SRC_LAYOUT_OK4, c, c, c, m)
    xs = np.linspace(L + 4, n - L - 4, nrec_side).round().astype(int)
    rec = _rec_grid(xs, xs, L + 2)
    return Workload(f"isorand{n}", n, n, n, dx, dx, dx, dt, nt, fwe.MODE_ORTHO, fwe.PML_SEPARABLE,
                    cij=cij, buoy=buoy, prof=prof, stf=[stf], injections=inj, receivers=rec,
                    info=dict(f0=f0, mt=m.tolist(), seed=seed, pml_cells=L))


def uneven_batch(nmodels=32, seed=11, edges=(96, 128, 160, 192, 256), nt=1500, **kw):
    """An uneven batch of independent heterogeneous isotropic cubes with
    seeded edge lengths drawn from `edges`."""
    rng = np.random.default_rng(seed)
    sizes = rng.choice(np.asarray(edges), size=nmodels)
    return [iso_random(int(n), seed * 1000 + q, nt=nt, **kw) for q, n in enumerate(sizes)]
