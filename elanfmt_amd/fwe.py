"""ctypes binding of libfwe.so (include/fwe.h, include/fwe_gen.h and include/fwe_post.h).

A thin Python mirror of the C ABI, used by the tests and bench.py.  All
compute happens in libfwe.so on the GPU; there is no CPU path here: when the
library or a HIP device is missing, the calls raise.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libfwe.so")
RUN_PATH = os.path.join(_HERE, "fwe_run")

# status codes (fwe_status)
OK, ERR_ARG, ERR_IO, ERR_HIP, ERR_STATE, ERR_NOMEM, ERR_UNSUPPORTED, ERR_NONFINITE = range(8)
KERNEL_AUTO, KERNEL_TWO_PASS, KERNEL_FUSED = range(3)
ORDER_LINEAR, ORDER_XCD, ORDER_PLANE_SPLIT = 0, 1, 2
MEDIUM_AUTO, MEDIUM_ARRAYS, MEDIUM_TABLE, MEDIUM_ISOTROPIC = 0, 1, 2, 3
FS_NONE, FS_BC1, FS_VACUUM = 0, 1, 2
MODE_ISO, MODE_ORTHO = 0, 1
PML_NONE, PML_PROFILE, PML_WEIGHTS3D, PML_SEPARABLE, PML_CPML = 0, 1, 2, 3, 4
LB_DISSIPATIVE, LB_FREE = 0, 1
SXX, SYY, SZZ, SXY, SXZ, SYZ = range(6)
F_VX, F_VY, F_VZ, F_SXX, F_SYY, F_SZZ, F_SXY, F_SXZ, F_SYZ = range(9)
REC_VX, REC_VY, REC_VZ = range(3)
SRC_LAYOUT_OK4, SRC_LAYOUT_ANISO = 0, 1
MAX_BATCH = 8

FIELD_NAMES = ["Vx", "Vy", "Vz", "Sxx", "Syy", "Szz", "Sxy", "Sxz", "Syz"]


class Config(C.Structure):
    _fields_ = [("nx", C.c_int32), ("ny", C.c_int32), ("nz", C.c_int32), ("nt", C.c_int32),
                ("mode", C.c_int32), ("pml", C.c_int32), ("free_surface", C.c_int32),
                ("src_after_bc", C.c_int32), ("nbatch", C.c_int32), ("device", C.c_int32),
                ("dx", C.c_double), ("dy", C.c_double), ("dz", C.c_double), ("dt", C.c_double)]


class Injection(C.Structure):
    _fields_ = [("comp", C.c_int32), ("ix", C.c_int32), ("iy", C.c_int32), ("iz", C.c_int32),
                ("batch", C.c_int32), ("coef", C.c_float)]


class Receiver(C.Structure):
    _fields_ = [("comp", C.c_int32), ("ix", C.c_int32), ("iy", C.c_int32), ("iz", C.c_int32),
                ("batch", C.c_int32)]


class LegacyParams(C.Structure):
    _fields_ = [("pa1", C.c_double * 10), ("pa2", C.c_double * 13), ("pa3", C.c_double * 12),
                ("mts", C.c_double * 6), ("n_pa1", C.c_int32), ("n_pa2", C.c_int32),
                ("n_pa3", C.c_int32), ("has_mts", C.c_int32)]


class RNG(C.Structure):
    _fields_ = [("s", C.c_uint64)]


class FweError(RuntimeError):
    def __init__(self, status, msg):
        super().__init__(f"fwe status {status}: {msg}")
        self.status = status


_lib = None

_FP = C.POINTER(C.c_float)
_DP = C.POINTER(C.c_double)
_IP = C.POINTER(C.c_int32)

# name -> (restype, argtypes); every symbol of include/fwe.h and fwe_gen.h
SIGNATURES = {
    "fwe_version": (C.c_char_p, []),
    "fwe_last_error": (C.c_char_p, []),
    "fwe_device_count": (C.c_int, [_IP]),
    "fwe_config_default": (C.c_int, [C.POINTER(Config), C.c_int32]),
    "fwe_create": (C.c_int, [C.POINTER(Config), C.POINTER(C.c_void_p)]),
    "fwe_destroy": (C.c_int, [C.c_void_p]),
    "fwe_get_config": (C.c_int, [C.c_void_p, C.POINTER(Config)]),
    "fwe_set_material_iso": (C.c_int, [C.c_void_p, _FP, _FP, C.c_float, C.c_float, C.c_float, C.c_float]),
    "fwe_set_material_ortho": (C.c_int, [C.c_void_p, C.POINTER(_FP), C.POINTER(_FP)]),
    "fwe_set_pml_profiles": (C.c_int, [C.c_void_p, _FP, _FP, _FP]),
    "fwe_set_pml_weights": (C.c_int, [C.c_void_p, _FP]),
    "fwe_set_cpml": (C.c_int, [C.c_void_p, C.c_int32, _FP, _FP, _FP]),
    "fwe_get_medium_vmax": (C.c_int, [C.c_void_p, C.POINTER(C.c_double)]),
    "fwe_cpml_state_size": (C.c_int, [C.c_void_p, C.POINTER(C.c_int64)]),
    "fwe_get_cpml_state": (C.c_int, [C.c_void_p, C.c_int32, _FP]),
    "fwe_set_cpml_state": (C.c_int, [C.c_void_p, C.c_int32, _FP]),
    "fwe_set_stf": (C.c_int, [C.c_void_p, C.c_int32, _FP, C.c_int32]),
    "fwe_set_injections": (C.c_int, [C.c_void_p, C.c_int32, C.POINTER(Injection)]),
    "fwe_set_receivers": (C.c_int, [C.c_void_p, C.c_int32, C.POINTER(Receiver)]),
    "fwe_mt_injections": (C.c_int, [C.c_int32, C.c_int32, C.c_int32, C.c_int32, _FP, C.c_int32,
                                    C.POINTER(Injection)]),
    "fwe_reset": (C.c_int, [C.c_void_p]),
    "fwe_run": (C.c_int, [C.c_void_p, C.c_int32]),
    "fwe_run_timed": (C.c_int, [C.c_void_p, C.c_int32, _DP, _DP, _DP]),
    "fwe_synchronize": (C.c_int, [C.c_void_p]),
    "fwe_set_graph_steps": (C.c_int, [C.c_void_p, C.c_int32]),
    "fwe_set_kernel": (C.c_int, [C.c_void_p, C.c_int32]),
    "fwe_get_kernel": (C.c_int, [C.c_void_p, _IP]),
    "fwe_set_tiling": (C.c_int, [C.c_void_p, C.c_int32, C.c_int32]),
    "fwe_set_sweep_tiling": (C.c_int, [C.c_void_p, C.c_int32, C.c_int32]),
    "fwe_get_sweep_tiling": (C.c_int, [C.c_void_p, _IP, _IP]),
    "fwe_get_tiling": (C.c_int, [C.c_void_p, _IP, _IP]),
    "fwe_autotune": (C.c_int, [C.c_void_p, C.c_int32, _DP]),
    "fwe_get_step": (C.c_int, [C.c_void_p, _IP]),
    "fwe_set_step": (C.c_int, [C.c_void_p, C.c_int32]),
    "fwe_get_traces": (C.c_int, [C.c_void_p, _FP]),
    "fwe_combine_traces": (C.c_int, [C.c_void_p, C.c_int32, _FP, _FP]),
    "fwe_combine_traces_noalloc": (C.c_int, [C.c_void_p, C.c_int32, _FP, _FP]),
    "fwe_get_field": (C.c_int, [C.c_void_p, C.c_int32, C.c_int32, _FP]),
    "fwe_set_field": (C.c_int, [C.c_void_p, C.c_int32, C.c_int32, _FP]),
    "fwe_read_legacy_params": (C.c_int, [C.c_char_p, C.POINTER(LegacyParams)]),
    "fwe_load_legacy_inputs": (C.c_int, [C.c_void_p, C.c_char_p, C.POINTER(LegacyParams)]),
    "fwe_write_legacy_outputs": (C.c_int, [C.c_void_p, C.c_char_p, C.POINTER(C.c_char_p), C.c_int32]),
    "fwe_copy_probe": (C.c_int, [C.c_int32, C.c_int64, C.c_int32, _DP]),
    "fwe_group_create": (C.c_int, [C.c_int32, C.POINTER(C.c_void_p), C.POINTER(C.c_void_p)]),
    "fwe_group_destroy": (C.c_int, [C.c_void_p]),
    "fwe_group_run": (C.c_int, [C.c_void_p, C.c_int32]),
    "fwe_group_run_timed": (C.c_int, [C.c_void_p, C.c_int32, _DP]),
    "fwe_group_set_tiling": (C.c_int, [C.c_void_p, C.c_int32, C.c_int32]),
    "fwe_group_autotune": (C.c_int, [C.c_void_p, C.c_int32, _DP]),
    "fwe_set_tile_cols": (C.c_int, [C.c_void_p, C.c_int32]),
    "fwe_get_tile_cols": (C.c_int, [C.c_void_p, _IP]),
    "fwe_lb_param_count": (C.c_int, [C.c_int32, _IP]),
    "fwe_lb_create": (C.c_int, [C.c_int32, C.c_int32, C.POINTER(C.c_void_p)]),
    "fwe_lb_destroy": (C.c_int, [C.c_void_p]),
    "fwe_lb_init": (C.c_int, [C.c_void_p, C.c_uint64]),
    "fwe_lb_set_correction_form": (C.c_int, [C.c_void_p, C.c_int32]),
    "fwe_lb_get_correction_form": (C.c_int, [C.c_void_p, C.POINTER(C.c_int32)]),
    "fwe_lb_set_params": (C.c_int, [C.c_void_p, C.c_int32, _FP]),
    "fwe_lb_get_params": (C.c_int, [C.c_void_p, C.c_int32, _FP]),
    "fwe_lb_eval": (C.c_int, [C.c_void_p, C.c_int32, C.c_int32, _FP, _IP, _FP, _FP, _FP, _FP, _DP]),
    "fwe_lb_train_step": (C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p, C.c_int32, C.c_float, C.c_float, C.c_int32,
                                    _DP]),
    "fwe_set_learned_boundary": (C.c_int, [C.c_void_p, C.c_void_p, C.c_float]),
    "fwe_set_wg_order": (C.c_int, [C.c_void_p, C.c_int32]),
    "fwe_group_set_tile_cols": (C.c_int, [C.c_void_p, C.c_int32]),
    "fwe_get_wg_order": (C.c_int, [C.c_void_p, _IP]),
    "fwe_set_medium_encoding": (C.c_int, [C.c_void_p, C.c_int32]),
    "fwe_set_batch_tile": (C.c_int, [C.c_void_p, C.c_int32]),
    "fwe_get_batch_tile": (C.c_int, [C.c_void_p, _IP]),
    "fwe_get_medium_encoding": (C.c_int, [C.c_void_p, _IP, _IP]),
    # generators (fwe_gen.h)
    "fwe_rng_seed": (None, [C.POINTER(RNG), C.c_uint64]),
    "fwe_rng_next": (C.c_uint64, [C.POINTER(RNG)]),
    "fwe_rng_uniform": (C.c_double, [C.POINTER(RNG)]),
    "fwe_rng_normal": (C.c_double, [C.POINTER(RNG)]),
    "fwe_gen_ricker": (None, [C.c_int32, C.c_double, C.c_double, C.c_double, C.c_double, _FP]),
    "fwe_gen_mt_sdr": (None, [C.c_double, C.c_double, C.c_double, C.c_double, _DP]),
    "fwe_gen_lune_eigs": (None, [C.c_double, C.c_double, _DP]),
    "fwe_gen_mt_from_eigs": (None, [_DP, _DP, C.c_double, _DP]),
    "fwe_gen_random_rotation": (None, [C.POINTER(RNG), _DP]),
    "fwe_gen_mt_lune_random": (None, [C.POINTER(RNG), C.c_double, _DP, _DP, _DP]),
    "fwe_gen_pml_profile": (None, [C.c_int32, C.c_int32, C.c_double, C.c_double, C.c_int32, C.c_int32,
                                   C.c_int32, _FP]),
    "fwe_gen_cell_profile": (None, [C.c_int32, _FP, _FP]),
    "fwe_gen_taper": (None, [C.c_int32, C.c_int32, C.c_int32, C.c_double, _FP]),
    "fwe_gen_cpml_table": (C.c_int32, [C.c_int32, C.c_int32] + [C.c_double] * 7 + [C.c_int32, C.c_int32, _FP]),
    "fwe_gen_weights3d": (None, [C.c_int32] * 9 + [C.c_double, _FP]),
    "fwe_gen_stagger_buoyancy": (None, [C.c_int32] * 3 + [_FP] * 4),
    "fwe_gen_harmonic_edges": (None, [C.c_int32] * 3 + [_FP] * 6),
    "fwe_gen_ortho_stiffness": (C.c_int, [C.c_double, C.c_double, C.c_double, _DP, _DP]),
    "fwe_gen_layered_ortho": (C.c_int, [C.c_int32] * 4 + [C.c_uint64] + [C.c_double] * 7 +
                              [C.POINTER(_FP), _FP, _IP, _DP]),
    "fwe_gen_random_field": (None, [C.c_int32] * 3 + [C.c_double, C.c_uint64, _FP]),
    "fwe_gen_cfl_dt": (C.c_double, [C.c_double] * 5),
    "fwe_gen_material_table": (C.c_int32, [C.c_int64, C.c_int32, C.POINTER(_FP), C.c_int32,
                                           C.POINTER(C.c_uint8), _FP]),
    # post-processing (fwe_post.h)
    "fwe_post_rotate_mt": (None, [C.c_double, _DP, _DP]),
    "fwe_post_butter_bandpass": (C.c_int32, [C.c_int32, C.c_double, C.c_double, C.c_double, _DP]),
    "fwe_post_sosfilt": (C.c_int32, [C.c_int32, _DP, _FP, C.c_int64, _FP]),
}


def lib():
    """Load libfwe.so (built by __graft_entry__.build()); raise if absent."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise FweError(ERR_STATE, f"{LIB_PATH} is not built; run __graft_entry__.build()")
        L = C.CDLL(LIB_PATH)
        for name, (res, args) in SIGNATURES.items():
            fn = getattr(L, name)
            fn.restype = res
            fn.argtypes = args
        _lib = L
    return _lib


def _check(st):
    if st != OK:
        raise FweError(st, lib().fwe_last_error().decode())


def fptr(a: np.ndarray):
    assert a.dtype == np.float32 and a.flags.c_contiguous
    return a.ctypes.data_as(_FP)


def dptr(a: np.ndarray):
    assert a.dtype == np.float64 and a.flags.c_contiguous
    return a.ctypes.data_as(_DP)


def f32(a) -> np.ndarray:
    return np.ascontiguousarray(np.asarray(a, dtype=np.float32).ravel(order="F"))


def _sized(a, n, what):
    """a as a float32 x-fastest vector of exactly n elements (the library
    reads n floats through the pointer)."""
    v = f32(a)
    if v.size != n:
        raise ValueError(f"{what}: {v.size} values, the grid needs {n}")
    return v


def device_count() -> int:
    n = C.c_int32(0)
    _check(lib().fwe_device_count(C.byref(n)))
    return n.value


def rotate_mt(azimuth_deg, m):
    """Moment tensor (Mxx, Myy, Mzz, Mxy, Mxz, Myz) rotated by the station
    azimuth, Q M Q^T (fwe_post_rotate_mt)."""
    a = np.ascontiguousarray(m, dtype=np.float64)
    out = np.zeros(6)
    lib().fwe_post_rotate_mt(float(azimuth_deg), dptr(a), dptr(out))
    return out


def butter_bandpass(order, flo, fhi, fs):
    """Second-order sections (order x 6) of the Butterworth band-pass
    (fwe_post_butter_bandpass)."""
    sos = np.zeros((order, 6))
    if lib().fwe_post_butter_bandpass(int(order), float(flo), float(fhi), float(fs), dptr(sos)) != 0:
        raise FweError(ERR_ARG, f"band {flo}..{fhi} Hz, order {order} at {fs} Hz is not a valid band-pass")
    return sos


def sosfilt(sos, x):
    """Causal cascade of the sections over float32 samples (fwe_post_sosfilt)."""
    sos = np.ascontiguousarray(sos, dtype=np.float64)
    x = np.ascontiguousarray(x, dtype=np.float32)
    y = np.empty_like(x)
    if lib().fwe_post_sosfilt(len(sos), dptr(sos), fptr(x), x.size, fptr(y)) != 0:
        raise FweError(ERR_ARG, "bad sections")
    return y


def field_shape(nx, ny, nz, field):
    """Reference shape (x, y, z) of a field id (include/fwe.h)."""
    return [(nx + 1, ny, nz), (nx, ny + 1, nz), (nx, ny, nz + 1), (nx, ny, nz), (nx, ny, nz),
            (nx, ny, nz), (nx + 1, ny + 1, nz), (nx + 1, ny, nz + 1), (nx, ny + 1, nz + 1)][field]


class Solver:
    """One fwe_ctx.  Arrays are numpy in the reference's x-fastest layout:
    a 3-D numpy array of shape (nx', ny', nz') is passed Fortran-ordered."""

    def __init__(self, nx, ny, nz, dx, dy, dz, dt, nt, mode=MODE_ISO, pml=None,
                 free_surface=0, src_after_bc=None, nbatch=1, device=0):
        L = lib()
        cfg = Config()
        _check(L.fwe_config_default(C.byref(cfg), mode))
        cfg.nx, cfg.ny, cfg.nz, cfg.nt = nx, ny, nz, nt
        cfg.dx, cfg.dy, cfg.dz, cfg.dt = dx, dy, dz, dt
        if pml is not None:
            cfg.pml = pml
        cfg.free_surface = free_surface
        if src_after_bc is not None:
            cfg.src_after_bc = src_after_bc
        cfg.nbatch = nbatch
        cfg.device = device
        self.cfg = cfg
        self._h = C.c_void_p()
        _check(L.fwe_create(C.byref(cfg), C.byref(self._h)))
        self.nrec = 0

    # context management
    def close(self):
        if self._h:
            lib().fwe_destroy(self._h)  # also detaches a learned boundary
            self._h = C.c_void_p()
        self._lb = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    # medium and damping
    def _dims(self):
        return self.cfg.nx, self.cfg.ny, self.cfg.nz

    def set_material_iso(self, c11, c12, c66, bx, by, bz):
        nx, ny, nz = self._dims()
        a, b = _sized(c11, nx * ny * nz, "c11"), _sized(c12, nx * ny * nz, "c12")
        _check(lib().fwe_set_material_iso(self._h, fptr(a), fptr(b), c66, bx, by, bz))

    def set_material_ortho(self, cij, buoy):
        nx, ny, nz = self._dims()
        nc = nx * ny * nz
        csz = [nc] * 6 + [nx * (ny + 1) * (nz + 1), (nx + 1) * ny * (nz + 1), (nx + 1) * (ny + 1) * nz]
        bsz = [(nx + 1) * ny * nz, nx * (ny + 1) * nz, nx * ny * (nz + 1)]
        cs = [_sized(c, n, f"cij[{q}]") for q, (c, n) in enumerate(zip(cij, csz))]
        bs = [_sized(b, n, f"buoy[{q}]") for q, (b, n) in enumerate(zip(buoy, bsz))]
        if len(cs) != 9 or len(bs) != 3:
            raise ValueError("nine moduli and three buoyancy arrays")
        carr = (_FP * 9)(*[fptr(c) for c in cs])
        barr = (_FP * 3)(*[fptr(b) for b in bs])
        _check(lib().fwe_set_material_ortho(self._h, carr, barr))

    def set_pml_profiles(self, wx, wy, wz):
        add = 1 if self.cfg.pml == PML_PROFILE else 0
        a, b, c = (_sized(w, n + add, "profile") for w, n in zip((wx, wy, wz), self._dims()))
        _check(lib().fwe_set_pml_profiles(self._h, fptr(a), fptr(b), fptr(c)))

    def set_pml_weights(self, w):
        nx, ny, nz = self._dims()
        a = _sized(w, nx * ny * nz, "weights")
        _check(lib().fwe_set_pml_weights(self._h, fptr(a)))

    def set_cpml(self, layer, tx, ty, tz):
        """Convolutional PML of `layer` cells per face; per axis the 6 x (n+1)
        table of workloads.cpml_table (fwe_gen_cpml_table)."""
        t = [np.ascontiguousarray(np.asarray(a, np.float32).ravel()) for a in (tx, ty, tz)]
        for a, n in zip(t, (self.cfg.nx, self.cfg.ny, self.cfg.nz)):
            if a.size != 6 * (n + 1):
                raise ValueError("a CPML table holds 6 x (n+1) floats")
        _check(lib().fwe_set_cpml(self._h, layer, *[fptr(a) for a in t]))

    def medium_vmax(self):
        """sqrt(largest normal modulus x largest buoyancy) of the medium."""
        v = C.c_double(0)
        _check(lib().fwe_get_medium_vmax(self._h, C.byref(v)))
        return v.value

    def cpml_state(self, batch=0):
        """Memory variables of one wavefield (flat, floats)."""
        n = C.c_int64(0)
        _check(lib().fwe_cpml_state_size(self._h, C.byref(n)))
        out = np.zeros(n.value, np.float32)
        _check(lib().fwe_get_cpml_state(self._h, batch, fptr(out)))
        return out

    def set_cpml_state(self, state, batch=0):
        a = np.ascontiguousarray(np.asarray(state, np.float32).ravel())
        n = C.c_int64(0)
        _check(lib().fwe_cpml_state_size(self._h, C.byref(n)))
        if a.size != n.value:
            raise ValueError(f"CPML state holds {n.value} floats")
        _check(lib().fwe_set_cpml_state(self._h, batch, fptr(a)))

    # source and receivers
    def set_stf(self, stf, batch=0):
        a = f32(stf)
        _check(lib().fwe_set_stf(self._h, batch, fptr(a), len(a)))

    def set_injections(self, inj):
        """inj: list of (comp, ix, iy, iz, batch, coef)."""
        arr = (Injection * max(1, len(inj)))(*[Injection(*t) for t in inj])
        _check(lib().fwe_set_injections(self._h, len(inj), arr))

    def set_receivers(self, rec):
        """rec: list of (comp, ix, iy, iz, batch)."""
        arr = (Receiver * max(1, len(rec)))(*[Receiver(*t) for t in rec])
        _check(lib().fwe_set_receivers(self._h, len(rec), arr))
        self.nrec = len(rec)

    # time loop
    def reset(self):
        _check(lib().fwe_reset(self._h))

    def run(self, nsteps):
        _check(lib().fwe_run(self._h, nsteps))

    def run_timed(self, nsteps):
        a, b, w = C.c_double(), C.c_double(), C.c_double()
        _check(lib().fwe_run_timed(self._h, nsteps, C.byref(a), C.byref(b), C.byref(w)))
        return a.value, b.value, w.value

    def synchronize(self):
        _check(lib().fwe_synchronize(self._h))

    def set_graph_steps(self, g):
        _check(lib().fwe_set_graph_steps(self._h, g))

    def set_kernel(self, kind):
        _check(lib().fwe_set_kernel(self._h, kind))

    def kernel(self):
        k = C.c_int32()
        _check(lib().fwe_get_kernel(self._h, C.byref(k)))
        return k.value

    def set_tiling(self, tile_rows, zchunk):
        _check(lib().fwe_set_tiling(self._h, tile_rows, zchunk))

    def tiling(self):
        a, b = C.c_int32(), C.c_int32()
        _check(lib().fwe_get_tiling(self._h, C.byref(a), C.byref(b)))
        return a.value, b.value

    def set_sweep_tiling(self, rows, zchunk):
        _check(lib().fwe_set_sweep_tiling(self._h, rows, zchunk))

    def sweep_tiling(self):
        a, b = C.c_int32(), C.c_int32()
        _check(lib().fwe_get_sweep_tiling(self._h, C.byref(a), C.byref(b)))
        return a.value, b.value

    def set_tile_cols(self, cols):
        _check(lib().fwe_set_tile_cols(self._h, cols))

    def tile_cols(self):
        c = C.c_int32()
        _check(lib().fwe_get_tile_cols(self._h, C.byref(c)))
        return c.value

    def set_wg_order(self, order):
        _check(lib().fwe_set_wg_order(self._h, order))

    def wg_order(self):
        c = C.c_int32()
        _check(lib().fwe_get_wg_order(self._h, C.byref(c)))
        return c.value

    def set_batch_tile(self, wavefields):
        """Wavefields per workgroup (a divisor of nbatch; 0 = default)."""
        _check(lib().fwe_set_batch_tile(self._h, wavefields))

    def batch_tile(self):
        c = C.c_int32()
        _check(lib().fwe_get_batch_tile(self._h, C.byref(c)))
        return c.value

    def set_medium_encoding(self, enc):
        """MEDIUM_AUTO (material table when the medium has one, either law;
        else the isotropic arrays when they apply), MEDIUM_ARRAYS (all the
        arrays), MEDIUM_TABLE or MEDIUM_ISOTROPIC (fail when they do not
        apply)."""
        _check(lib().fwe_set_medium_encoding(self._h, enc))

    def medium_encoding(self):
        """(encoding fwe_run uses: MEDIUM_ARRAYS, MEDIUM_TABLE or MEDIUM_ISOTROPIC, table entries)"""
        e, n = C.c_int32(), C.c_int32()
        _check(lib().fwe_get_medium_encoding(self._h, C.byref(e), C.byref(n)))
        return e.value, n.value

    def autotune(self, trial_steps=5):
        ms = C.c_double()
        _check(lib().fwe_autotune(self._h, trial_steps, C.byref(ms)))
        return ms.value

    def step(self):
        it = C.c_int32()
        _check(lib().fwe_get_step(self._h, C.byref(it)))
        return it.value

    def set_step(self, it):
        _check(lib().fwe_set_step(self._h, it))

    # results
    def traces(self):
        out = np.zeros(max(1, self.nrec * self.cfg.nt), np.float32)
        st = lib().fwe_get_traces(self._h, fptr(out))
        if st not in (OK, ERR_NONFINITE):
            _check(st)
        return out[: self.nrec * self.cfg.nt].reshape(self.nrec, self.cfg.nt)

    def combine_traces(self, coef, noalloc=False):
        """Traces of len(coef) sources, each a linear combination (coef row,
        one weight per wavefield) of the nbatch elementary-source wavefields;
        shape (n, receivers per wavefield, nt).  noalloc: stage through the
        library's static device buffer (fwe_combine_traces_noalloc) instead
        of per-call device allocations."""
        cf = np.ascontiguousarray(np.atleast_2d(coef), dtype=np.float32)  # row m = source m
        n, nb = cf.shape
        if nb != self.cfg.nbatch:
            raise ValueError(f"coef needs {self.cfg.nbatch} columns")
        nr = self.nrec // self.cfg.nbatch
        out = np.empty(max(1, n * nr * self.cfg.nt), np.float32)
        fn = lib().fwe_combine_traces_noalloc if noalloc else lib().fwe_combine_traces
        _check(fn(self._h, n, fptr(cf), fptr(out)))
        return out[: n * nr * self.cfg.nt].reshape(n, nr, self.cfg.nt)

    def get_field(self, field, batch=0):
        shp = field_shape(self.cfg.nx, self.cfg.ny, self.cfg.nz, field)
        out = np.zeros(int(np.prod(shp)), np.float32)
        _check(lib().fwe_get_field(self._h, batch, field, fptr(out)))
        return out.reshape(shp, order="F")

    def set_field(self, field, arr, batch=0):
        shp = field_shape(self.cfg.nx, self.cfg.ny, self.cfg.nz, field)
        a = _sized(arr, int(np.prod(shp)), f"field {field}")
        _check(lib().fwe_set_field(self._h, batch, field, fptr(a)))


def mt_injections(layout, xs, ys, zs, m, batch=0):
    out = (Injection * 6)()
    mm = np.ascontiguousarray(np.asarray(m, np.float32))
    _check(lib().fwe_mt_injections(layout, xs, ys, zs, fptr(mm), batch, out))
    return [(o.comp, o.ix, o.iy, o.iz, o.batch, o.coef) for o in out]


def copy_probe(nbytes, reps=20, device=0):
    ms = C.c_double()
    _check(lib().fwe_copy_probe(device, nbytes, reps, C.byref(ms)))
    return ms.value


class Group:
    """fwe_group: several single-wavefield Solvers advanced by one launch per
    half-step.  The solvers must stay alive while the group is used."""

    def __init__(self, solvers):
        self.solvers = list(solvers)
        arr = (C.c_void_p * len(self.solvers))(*[s._h.value for s in self.solvers])
        self._h = C.c_void_p()
        _check(lib().fwe_group_create(len(self.solvers), arr, C.byref(self._h)))

    def close(self):
        if self._h:
            lib().fwe_group_destroy(self._h)
            self._h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def run(self, nsteps):
        _check(lib().fwe_group_run(self._h, nsteps))

    def run_timed(self, nsteps):
        ms = C.c_double()
        _check(lib().fwe_group_run_timed(self._h, nsteps, C.byref(ms)))
        return ms.value

    def set_tiling(self, tile_rows, zchunk):
        _check(lib().fwe_group_set_tiling(self._h, tile_rows, zchunk))

    def set_tile_cols(self, cols):
        _check(lib().fwe_group_set_tile_cols(self._h, cols))

    def autotune(self, trial_steps=3):
        ms = C.c_double()
        _check(lib().fwe_group_autotune(self._h, trial_steps, C.byref(ms)))
        return ms.value


class LearnedBoundary:
    """fwe_lb: per-face networks of a thin absorbing layer (see
    include/fwe.h).  Faces 0..5 = x-, x+, y-, y+, z-, z+."""

    def __init__(self, layer=6, device=0):
        self.layer = layer
        n = C.c_int32()
        _check(lib().fwe_lb_param_count(layer, C.byref(n)))
        self.nparams = n.value
        self._h = C.c_void_p()
        _check(lib().fwe_lb_create(device, layer, C.byref(self._h)))

    def close(self):
        """Free the networks; refused (FweError) while a solver still has
        the boundary attached (detach it, or close that solver, first)."""
        if self._h:
            _check(lib().fwe_lb_destroy(self._h))
            self._h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def init(self, seed):
        _check(lib().fwe_lb_init(self._h, seed))

    def get_params(self, face):
        out = np.zeros(self.nparams, np.float32)
        _check(lib().fwe_lb_get_params(self._h, face, fptr(out)))
        return out

    def set_params(self, face, p):
        a = np.ascontiguousarray(p, dtype=np.float32)
        if a.size != self.nparams:
            raise ValueError(f"need {self.nparams} parameters")
        _check(lib().fwe_lb_set_params(self._h, face, fptr(a)))

    def eval(self, face, x, depth, target=None, mask=None):
        """(y [n,9], loss, grad [nparams]) of one face's network on explicit
        samples; loss and grad are None without target/mask."""
        x = np.ascontiguousarray(x, dtype=np.float32)
        n = x.shape[0]
        d = np.ascontiguousarray(depth, dtype=np.int32)
        y = np.zeros((n, 9), np.float32)
        if target is None:
            _check(lib().fwe_lb_eval(self._h, face, n, fptr(x), d.ctypes.data_as(_IP), None, None, fptr(y), None,
                                     None))
            return y, None, None
        t = np.ascontiguousarray(target, dtype=np.float32)
        m = np.ascontiguousarray(mask, dtype=np.float32)
        g = np.zeros(self.nparams, np.float32)
        loss = C.c_double()
        _check(lib().fwe_lb_eval(self._h, face, n, fptr(x), d.ctypes.data_as(_IP), fptr(t), fptr(m), fptr(y),
                                 fptr(g), C.byref(loss)))
        return y, loss.value, g

    def set_correction_form(self, form):
        """LB_DISSIPATIVE (the default) or LB_FREE."""
        _check(lib().fwe_lb_set_correction_form(self._h, form))

    def correction_form(self):
        f = C.c_int32()
        _check(lib().fwe_lb_get_correction_form(self._h, C.byref(f)))
        return f.value

    def train_step(self, thick, thin, offset, impedance, lr, resync=True):
        loss = C.c_double()
        _check(lib().fwe_lb_train_step(self._h, thick._h, thin._h, offset, impedance, lr, int(resync),
                                       C.byref(loss)))
        return loss.value

    def attach(self, solver, impedance):
        _check(lib().fwe_set_learned_boundary(solver._h, self._h, impedance))
        solver._lb = self  # the solver's steps read these networks: keep them alive

    @staticmethod
    def detach(solver):
        _check(lib().fwe_set_learned_boundary(solver._h, None, 1.0))
        solver._lb = None
