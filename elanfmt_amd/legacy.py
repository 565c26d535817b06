"""Writers/readers of the reference's on-disk contract.

Inputs (raw float32, x fastest; OK4.m:348-429, ANIS.m:378-533):
  pa1.dat  [dx dy dz dt Lx Ly Lz lamx lamy lamz]
  pa2.dat  [c11u c33u c13u c12u c23u c44u c55u c66u alpha1 alpha2 alpha3 M1 c22u]
  pa3.dat  [vrho_x_11 vrho_y_11 vrho_z_11 vrho_x_12 .. eta_k3]   (12 values)
  MTS.dat  [Mxx Myy Mzz Mxy Mxz Myz]
  Src.dat  (nt)
  iso:   c11u_het.dat c12u_het.dat weights_x.dat weights_y.dat weights_z.dat
  ortho: C11m C22m C33m C12m C13m C23m C44myz C55mxz C66mxy .dat,
         vrho_x_11_hc vrho_y_11_hc vrho_z_11_hc .dat, weights.dat
         (or taper_x/y/z.dat for the separable damping, which has no
         reference counterpart)
Outputs (OK4.cu:153-164,521-532): 0_infos.inf, 0_0_<name>.res.
"""
from __future__ import annotations

import os

import numpy as np

from . import fwe


def _w(d, name, a):
    np.asarray(a, np.float32).ravel(order="F").tofile(os.path.join(d, name))


def write_inputs(d, wl, mts=None):
    """Write the legacy input set of a single-wavefield Workload into d."""
    os.makedirs(d, exist_ok=True)
    Lx, Ly, Lz = (wl.nx - 1) * wl.dx, (wl.ny - 1) * wl.dy, (wl.nz - 1) * wl.dz
    _w(d, "pa1.dat", [wl.dx, wl.dy, wl.dz, wl.dt, Lx, Ly, Lz, 1.0, 1.0, 1.0])
    if wl.mode == fwe.MODE_ISO:
        c11 = float(wl.c11.ravel()[0])
        c12 = float(wl.c12.ravel()[0])
        pa2 = [c11, c11, c12, c12, c12, wl.c66, wl.c66, wl.c66, 0, 0, 0, 0, c11]
        pa3 = list(wl.bscal) + [0.0] * 9
    else:
        pa2 = [0.0] * 13
        pa3 = [0.0] * 12
    _w(d, "pa2.dat", pa2)
    _w(d, "pa3.dat", pa3)
    if mts is not None:
        _w(d, "MTS.dat", mts)
    _w(d, "Src.dat", wl.stf[0])
    if wl.mode == fwe.MODE_ISO:
        _w(d, "c11u_het.dat", wl.c11)
        _w(d, "c12u_het.dat", wl.c12)
        for ax, w in zip("xyz", wl.prof):
            _w(d, f"weights_{ax}.dat", w)
    else:
        names = ["C11m", "C22m", "C33m", "C12m", "C13m", "C23m", "C44myz", "C55mxz", "C66mxy"]
        for n, a in zip(names, wl.cij):
            _w(d, n + ".dat", a)
        for ax, b in zip("xyz", wl.buoy):
            _w(d, f"vrho_{ax}_11_hc.dat", b)
        if wl.pml == fwe.PML_WEIGHTS3D:
            _w(d, "weights.dat", wl.w3)
        elif wl.pml == fwe.PML_SEPARABLE:
            for ax, t in zip("xyz", wl.prof):
                _w(d, f"taper_{ax}.dat", t)


def read_infos(d):
    """0_infos.inf -> (PRECIS, nx, ny, nz, NB_PARAMS, NB_EVOL) (OK4.m:458)."""
    with open(os.path.join(d, "0_infos.inf")) as f:
        return tuple(int(v) for v in f.read().split())


def read_res(d, name, shape=None):
    a = np.fromfile(os.path.join(d, f"0_0_{name}.res"), dtype=np.float32)
    return a.reshape(shape, order="F") if shape is not None else a
