"""ctypes binding of the CPU oracle (liboracle.so / liboracle32.so).

TEST INFRASTRUCTURE ONLY (see oracle/__init__.py).
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_FP = C.POINTER(C.c_float)
_DP = C.POINTER(C.c_double)
_IP = C.POINTER(C.c_int32)
_libs = {}


def lib_path(fp32=False):
    return os.path.join(_HERE, "liboracle32.so" if fp32 else "liboracle.so")


def _lib(fp32=False):
    if fp32 not in _libs:
        p = lib_path(fp32)
        if not os.path.exists(p):
            raise RuntimeError(f"{p} not built (make -C oracle)")
        L = C.CDLL(p)
        sig = {
            "oracle_create": (C.c_void_p, [C.c_int] * 4 + [C.c_double] * 4 + [C.c_int] * 4),
            "oracle_destroy": (None, [C.c_void_p]),
            "oracle_set_iso": (C.c_int, [C.c_void_p, _FP, _FP] + [C.c_double] * 4),
            "oracle_set_ortho": (C.c_int, [C.c_void_p, C.POINTER(_FP), C.POINTER(_FP)]),
            "oracle_set_profiles": (C.c_int, [C.c_void_p, _FP, _FP, _FP]),
            "oracle_set_weights3d": (C.c_int, [C.c_void_p, _FP]),
            "oracle_set_cpml": (C.c_int, [C.c_void_p, C.c_int, _FP, _FP, _FP]),
            "oracle_set_stf": (C.c_int, [C.c_void_p, _FP, C.c_int]),
            "oracle_set_injections": (C.c_int, [C.c_void_p, C.c_int, _IP, _IP, _FP]),
            "oracle_set_receivers": (C.c_int, [C.c_void_p, C.c_int, _IP, _IP]),
            "oracle_reset": (None, [C.c_void_p]),
            "oracle_run": (None, [C.c_void_p, C.c_int, C.c_int]),
            "oracle_step_index": (C.c_int, [C.c_void_p]),
            "oracle_get_traces": (None, [C.c_void_p, _DP]),
            "oracle_get_field": (C.c_int, [C.c_void_p, C.c_int, _DP]),
            "oracle_set_field": (C.c_int, [C.c_void_p, C.c_int, _DP]),
            "oracle_field_size": (C.c_int64, [C.c_void_p, C.c_int]),
            "oracle_sum_sq_velocity": (C.c_double, [C.c_void_p]),
            "oracle_real_bytes": (C.c_int, []),
        }
        for n, (r, a) in sig.items():
            f = getattr(L, n)
            f.restype, f.argtypes = r, a
        _libs[fp32] = L
    return _libs[fp32]


def _f32(a):
    return np.ascontiguousarray(np.asarray(a, np.float32).ravel(order="F"))


_SHAPES = lambda nx, ny, nz: [(nx + 1, ny, nz), (nx, ny + 1, nz), (nx, ny, nz + 1), (nx, ny, nz),
                              (nx, ny, nz), (nx, ny, nz), (nx + 1, ny + 1, nz), (nx + 1, ny, nz + 1),
                              (nx, ny + 1, nz + 1)]


class Oracle:
    """Same call pattern as elanfmt_amd.fwe.Solver, on the CPU, FP64 (or FP32)."""

    def __init__(self, nx, ny, nz, dx, dy, dz, dt, nt, mode=0, pml=1, free_surface=0,
                 src_after_bc=None, fp32=False):
        self.L = _lib(fp32)
        if src_after_bc is None:
            src_after_bc = 0 if mode == 0 else 1
        self.nx, self.ny, self.nz, self.nt = nx, ny, nz, nt
        self._keep = []
        self.h = self.L.oracle_create(mode, nx, ny, nz, dx, dy, dz, dt, pml, free_surface,
                                      src_after_bc, nt)
        if not self.h:
            raise ValueError("oracle_create failed")
        self.nrec = 0

    def close(self):
        if self.h:
            self.L.oracle_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def set_material_iso(self, c11, c12, c66, bx, by, bz):
        a, b = _f32(c11), _f32(c12)
        self.L.oracle_set_iso(self.h, a.ctypes.data_as(_FP), b.ctypes.data_as(_FP), c66, bx, by, bz)

    def set_material_ortho(self, cij, buoy):
        cs = [_f32(c) for c in cij]
        bs = [_f32(b) for b in buoy]
        self.L.oracle_set_ortho(self.h, (_FP * 9)(*[c.ctypes.data_as(_FP) for c in cs]),
                                (_FP * 3)(*[b.ctypes.data_as(_FP) for b in bs]))

    def set_pml_profiles(self, wx, wy, wz):
        a, b, c = _f32(wx), _f32(wy), _f32(wz)
        self.L.oracle_set_profiles(self.h, a.ctypes.data_as(_FP), b.ctypes.data_as(_FP),
                                   c.ctypes.data_as(_FP))

    def set_pml_weights(self, w):
        a = _f32(w)
        self.L.oracle_set_weights3d(self.h, a.ctypes.data_as(_FP))

    def set_cpml(self, layer, tx, ty, tz):
        """pml = 4: convolutional PML of `layer` cells per face; per axis the
        6 x (n+1) coefficient table of fwe_gen_cpml_table."""
        t = [np.ascontiguousarray(np.asarray(a, np.float32).ravel()) for a in (tx, ty, tz)]
        for a, n in zip(t, (self.nx, self.ny, self.nz)):
            if a.size != 6 * (n + 1):
                raise ValueError("a CPML table holds 6 x (n+1) floats")
        if self.L.oracle_set_cpml(self.h, layer, *[a.ctypes.data_as(_FP) for a in t]) != 0:
            raise ValueError("oracle_set_cpml failed")

    def set_stf(self, stf, batch=0):
        a = _f32(stf)
        self.L.oracle_set_stf(self.h, a.ctypes.data_as(_FP), len(a))

    def set_injections(self, inj):
        inj = [t for t in inj]
        comp = np.array([t[0] for t in inj], np.int32)
        ijk = np.array([[t[1], t[2], t[3]] for t in inj], np.int32).reshape(-1)
        coef = np.array([t[5] for t in inj], np.float32)
        self.L.oracle_set_injections(self.h, len(inj), comp.ctypes.data_as(_IP),
                                     ijk.ctypes.data_as(_IP), coef.ctypes.data_as(_FP))

    def set_receivers(self, rec):
        comp = np.array([t[0] for t in rec], np.int32)
        ijk = np.array([[t[1], t[2], t[3]] for t in rec], np.int32).reshape(-1)
        self.L.oracle_set_receivers(self.h, len(rec), comp.ctypes.data_as(_IP),
                                    ijk.ctypes.data_as(_IP))
        self.nrec = len(rec)

    def reset(self):
        self.L.oracle_reset(self.h)

    def run(self, nsteps, nthreads=0):
        self.L.oracle_run(self.h, nsteps, nthreads)

    def step(self):
        return self.L.oracle_step_index(self.h)

    def traces(self):
        out = np.zeros(max(1, self.nrec * self.nt), np.float64)
        self.L.oracle_get_traces(self.h, out.ctypes.data_as(_DP))
        return out[: self.nrec * self.nt].reshape(self.nrec, self.nt)

    def get_field(self, f):
        shp = _SHAPES(self.nx, self.ny, self.nz)[f]
        out = np.zeros(int(np.prod(shp)), np.float64)
        self.L.oracle_get_field(self.h, f, out.ctypes.data_as(_DP))
        return out.reshape(shp, order="F")

    def set_field(self, f, arr):
        a = np.ascontiguousarray(np.asarray(arr, np.float64).ravel(order="F"))
        self.L.oracle_set_field(self.h, f, a.ctypes.data_as(_DP))

    def sum_sq_velocity(self):
        return self.L.oracle_sum_sq_velocity(self.h)
