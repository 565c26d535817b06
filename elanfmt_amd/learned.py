"""Training and evaluation of the learned thin boundary (fwe_lb_* in
include/fwe.h) on synthetic runs of the solver itself.

Teacher: a run whose grid extends THICK - L cells beyond the thin grid on
every face, with exactly the thin layer's damping over the thin window and a
smooth damping ramp beyond it.  Its window over the thin grid is copied into
the thin run; the networks learn the difference between the teacher step and
the thin run's plain step in the thin layer (fwe_lb_train_step), which is
the missing continuation at the edge of the thin grid.  Media are generated on the largest grid of an
experiment and cropped, so every run sees the same medium on its overlap.

Evaluation: receiver traces of the thin run with and without the networks,
and of a thicker plain layer, against a reference run on a grid padded far
enough that no wave returns from its own boundary within the time window.
"""
from __future__ import annotations

from dataclasses import dataclass

import numpy as np

from . import fwe, workloads as W

THICK = 32  # damping cells of the teacher runs
DX = 50.0
# convolutional PML of the thin layer (fwe_gen_cpml_table arguments), and of
# the references the thin layers are judged against
CPML_THIN = dict(R=1e-4, kappa_max=1.0, alpha_max=0.0, power=2.0)
CPML_REF = dict(R=1e-6, kappa_max=1.0, power=2.0)  # alpha_max = pi f0


def impedance(cells, bcell):
    """Z = sqrt(mean C11 * mean rho): stresses are divided by Z in the
    network inputs."""
    return float(np.sqrt(np.mean(cells[0], dtype=np.float64) * np.mean(1.0 / bcell, dtype=np.float64)))


def _dims(n):
    return (n, n, n) if np.isscalar(n) else tuple(int(v) for v in n)


def layered_medium(n, seed, nl=5):
    """Layered orthorhombic medium on an n^3 (or nx x ny x nz) grid, as the
    384x384x192 workload: cell moduli, cell buoyancy, staggered inputs."""
    nx, ny, nz = _dims(n)
    cells, bcell, _, vmax = W.layered_ortho(nx, ny, nz, nl, seed)
    cij, buoy = W.ortho_from_cells(cells, bcell)
    return dict(cells=cells, bcell=bcell, cij=cij, buoy=buoy, vmax=vmax)


def extended_medium(core, grid, seed, nl=5):
    """A layered medium generated on the `core` grid (an interior) and
    continued unchanged outward to `grid` (centred): every cell outside the
    core takes the value of the nearest core cell, then the staggered
    arrays are built on the whole grid.  Every absorbing layer emulates the
    grid's own medium continued to infinity, so a reference run on the
    larger grid measures the boundaries alone, not the medium the smaller
    grids leave out."""
    core, grid = _dims(core), _dims(grid)
    m = layered_medium(core, seed, nl)
    pad = [((g - c) // 2, g - c - (g - c) // 2) for g, c in zip(grid, core)]
    if any(a < 0 or b < 0 for a, b in pad):
        raise ValueError(f"grid {grid} smaller than the core {core}")
    cells = [np.asfortranarray(np.pad(c, pad, mode="edge")) for c in m["cells"]]
    bcell = np.asfortranarray(np.pad(m["bcell"], pad, mode="edge"))
    vmax = m["vmax"]  # no new material outside the core
    del m
    cij, buoy = W.ortho_from_cells(cells, bcell)
    return dict(cells=cells, bcell=bcell, cij=cij, buoy=buoy, vmax=vmax)


_CIJ_EXTRA = [(0, 0, 0)] * 6 + [(0, 1, 1), (1, 0, 1), (1, 1, 0)]
_B_EXTRA = [(1, 0, 0), (0, 1, 0), (0, 0, 1)]


def crop(med, o, n):
    """Window [o, o+n) per axis of a medium (staggered arrays keep their +1);
    o and n are scalars or per-axis triples."""
    o3, n3 = _dims(o), _dims(n)
    full = med["bcell"].shape
    if any(a < 0 or a + b > f for a, b, f in zip(o3, n3, full)):
        raise ValueError(f"window {n3} at {o3} leaves the medium's grid {full}")

    def sl(a, ex):
        return a[o3[0]:o3[0] + n3[0] + ex[0], o3[1]:o3[1] + n3[1] + ex[1], o3[2]:o3[2] + n3[2] + ex[2]]

    return dict(cells=[sl(c, (0, 0, 0)) for c in med["cells"]], bcell=sl(med["bcell"], (0, 0, 0)),
                cij=[sl(c, e) for c, e in zip(med["cij"], _CIJ_EXTRA)],
                buoy=[sl(b, e) for b, e in zip(med["buoy"], _B_EXTRA)], vmax=med["vmax"])


def thin_taper(n, L, dt, R):
    """Per-step cell taper of an L-cell plain layer on an n-cell axis."""
    return W.cell_profile(W.pml_profile(n, L, dt, R=R))


def teacher_taper(n_thin, L, off, dt, vmax, R_thin, R_ext=1e-6):
    """Taper of the teacher axis (n_thin + 2 off cells): exactly the thin
    layer's taper over the thin window, continued outward by the thin edge
    value times a quadratic damping ramp over the `off` extra cells (round
    trip reflection target R_ext).  Teacher and thin steps then differ only
    where the thin grid ends."""
    t = thin_taper(n_thin, L, dt, R_thin)
    d0 = 3.0 * vmax * np.log(1.0 / R_ext) / (2.0 * off * DX)
    q = np.arange(1, off + 1)
    ext = np.exp(-d0 * (q / off) ** 2 * dt)
    out = np.empty(n_thin + 2 * off, np.float32)
    out[off:off + n_thin] = t
    out[off - q] = t[0] * ext
    out[off + n_thin - 1 + q] = t[-1] * ext
    return out


def workload(name, n, damp, med, dt, src, m, f0, nt, receivers=(), R=1e-8, taper=None, nbatch=1):
    """Ortho workload on an n^3 (or nx x ny x nz) grid with `damp` cells of
    separable damping (reflection target R) on every face, or the given 1-D
    taper on all three axes (cubes).  With nbatch > 1, `m` is a list of
    nbatch tensors, one per wavefield, and receivers repeat per wavefield."""
    nx, ny, nz = _dims(n)
    if taper is not None:
        prof = tuple(np.asarray(taper, np.float32) for _ in range(3))
    else:
        prof = tuple(W.cell_profile(W.pml_profile(q, damp, dt, R=R)) for q in (nx, ny, nz))
    stf = W.ricker(nt, dt, f0, 1.2 / f0) * np.float32(dt / DX ** 3)
    ms = [m] if nbatch == 1 else list(m)
    inj = sum((fwe.mt_injections(fwe.SRC_LAYOUT_OK4, *src, mb, batch=b) for b, mb in enumerate(ms)), [])
    rec = [t[:4] + (b,) for b in range(nbatch) for t in receivers]
    return W.Workload(name, nx, ny, nz, DX, DX, DX, dt, nt, fwe.MODE_ORTHO, fwe.PML_SEPARABLE, nbatch=nbatch,
                      cij=med["cij"], buoy=med["buoy"], prof=prof, stf=[stf] * nbatch, injections=inj,
                      receivers=rec)


def teacher_cpml_table(thin_table, off):
    """CPML table of a teacher axis (the thin axis plus `off` cells at both
    ends): the thin table over the thin window, and beyond it the values of
    the thin layer's outer edge (normalised depth 1, the thin table's half
    node 0), so the teacher's layer is the thin one continued at full
    damping.  Teacher and thin steps then differ only where the thin grid
    ends."""
    t = np.asarray(thin_table, np.float32)
    n = t.shape[1] - 1
    edge = t[3:6, 0]
    out = np.empty((6, n + 2 * off + 1), np.float32)
    out[:, off:off + n + 1] = t
    for r in range(3):
        out[r, :off] = edge[r]
        out[3 + r, :off] = edge[r]
        out[r, off + n + 1:] = edge[r]
        out[3 + r, off + n + 1:] = edge[r]
    return out


def cpml_workload(wl, layer, vmax, tables=None, **cp):
    """Switch a workload to the convolutional PML: `layer` cells with the
    generator's tables (cp: fwe_gen_cpml_table arguments), or the given
    per-axis tables."""
    if tables is None:
        return W.use_cpml(wl, layer, vmax=vmax, **cp)
    wl.cpml, wl.cpml_layer, wl.pml, wl.prof, wl.w3 = tuple(tables), layer, fwe.PML_CPML, None, None
    return wl


@dataclass
class TrainLog:
    runs: int = 0
    steps: int = 0
    first_loss: float = float("nan")
    last_loss: float = float("nan")
    losses: list = None


def train(lb: fwe.LearnedBoundary, runs=200, n_int=40, warm=None, steps=None, lr=5e-4, seed0=1000, f0=(3.0, 8.0),
          R_thin=1e-3, resync_every=8, log_every=0, progress_every=0, base="plain", extend=False):
    """Training on `runs` thick-damping runs over seeded random layered
    media and lune moment tensors.  n_int is the interior edge, or a
    sequence of edges drawn per run.  Each run: `warm` plain steps (default
    1.5 n_int, until the first wavefront reaches the layer), then `steps`
    training steps (default 5 n_int; one Adam step per face each); the thin
    run is re-synchronised to the thick run every `resync_every` steps and
    rolls out with its own corrections in between.  base: the thin layer's
    own damping, "plain" (the reference's profile) or "cpml" (CPML_THIN;
    the teacher continues it, teacher_cpml_table).  extend: media generated
    on the interior and continued unchanged outward (always with "cpml"),
    instead of one medium generated on the teacher grid."""
    L = lb.layer
    off = THICK - L
    sizes = [int(n_int)] if np.isscalar(n_int) else [int(v) for v in n_int]
    log = TrainLog(losses=[])
    for r in range(runs):
        seed = seed0 + r
        rng = np.random.default_rng(seed)
        ni = sizes[int(rng.integers(len(sizes)))]
        n_thin, n_thick = ni + 2 * L, ni + 2 * THICK
        wu = warm if warm is not None else max(60, int(1.5 * ni))
        nsteps = steps if steps is not None else 5 * ni
        med = extended_medium(ni, n_thick, seed) if (extend or base == "cpml") else layered_medium(n_thick, seed)
        dt = W.cfl_dt(med["vmax"], DX, DX, DX, 0.5)
        z = impedance(med["cells"], med["bcell"])
        ff = float(rng.uniform(*f0))
        m, _, _ = W.mt_lune_random(seed * 17 + 5, 1e15)
        c = ni // 2 + rng.integers(-ni // 6, ni // 6 + 1, 3)
        nt = wu + nsteps + 2
        taper = teacher_taper(n_thin, L, off, dt, med["vmax"], R_thin)
        wthick = workload("thick", n_thick, THICK, med, dt, tuple(int(v) + THICK for v in c), m, ff, nt, taper=taper)
        wthin = workload("thin", n_thin, L, crop(med, off, n_thin), dt, tuple(int(v) + L for v in c), m, ff, nt,
                         R=R_thin)
        if base == "cpml":
            tt = W.cpml_table(n_thin, L, DX, dt, med["vmax"], **CPML_THIN)
            cpml_workload(wthin, L, med["vmax"], tables=(tt, tt, tt))
            tk = teacher_cpml_table(tt, off)
            cpml_workload(wthick, THICK, med["vmax"], tables=(tk, tk, tk))
        thick = W.make_solver(wthick)
        thin = W.make_solver(wthin)
        thick.run(wu)
        for t in range(nsteps):
            loss = lb.train_step(thick, thin, off, z, lr, resync=(t % resync_every == 0))
            if log.steps == 0:
                log.first_loss = loss
            log.steps += 1
            log.last_loss = loss
            if log_every and log.steps % log_every == 0:
                log.losses.append((log.steps, loss))
        thick.close()
        thin.close()
        log.runs += 1
        if progress_every and log.runs % progress_every == 0:
            print(f"train: {log.runs}/{runs} runs, loss {log.last_loss:.3e}", flush=True)
    return log


def face_receivers(n_thin, L, inset=2, per_side=2):
    """Receivers (vx, vy, vz) `inset` cells inside the layer on each face
    of the thin grid: per_side^2 points per face."""
    d = L + inset
    ts = np.linspace(n_thin // 4, n_thin - 1 - n_thin // 4, per_side).round().astype(int).tolist()
    pts = []
    for a in ts:
        for b in ts:
            pts += [(d, a, b), (n_thin - 1 - d, a, b), (a, d, b), (a, n_thin - 1 - d, b), (a, b, d),
                    (a, b, n_thin - 1 - d)]
    return [(comp, i, j, k, 0) for (i, j, k) in pts for comp in range(3)]


def evaluate(lb: fwe.LearnedBoundary | None, seed=7, n_int=40, nt=400, f0=5.0, fallback=24, R_thin=1e-3,
             cfl=0.5, base="plain", extend=False):
    """Relative L2 error of receiver traces near the boundary against a
    padded reference, for the thin layer (base "plain" or "cpml"), the thin
    layer with the networks (if lb), and a `fallback`-cell plain layer.
    extend (always with "cpml"): the medium generated on the interior and
    continued unchanged outward, and the reference bounded by a 32-cell
    CPML, so the errors measure the boundaries alone."""
    L = lb.layer if lb is not None else 6
    n_thin = n_int + 2 * L
    rng = np.random.default_rng(seed)
    # padding so that nothing returns from the reference's own boundary
    cells_per_step = cfl / np.sqrt(3.0)
    pad = int(np.ceil(cells_per_step * nt / 2 + 24 + 8))
    n_ref = n_int + 2 * pad
    ext = extend or base == "cpml"
    med = extended_medium(n_int, n_ref, seed) if ext else layered_medium(n_ref, seed)
    dt = W.cfl_dt(med["vmax"], DX, DX, DX, cfl)
    z = impedance(med["cells"], med["bcell"])
    m, _, _ = W.mt_lune_random(seed * 31 + 1, 1e15)
    c = n_int // 2 + rng.integers(-n_int // 8, n_int // 8 + 1, 3)
    rec_thin = face_receivers(n_thin, L)

    def run(n, damp, R=1e-8, with_lb=False, cp=None):
        o = (n_ref - n) // 2
        shift = (n - n_thin) // 2
        rec = [(q, i + shift, j + shift, k + shift, b) for (q, i, j, k, b) in rec_thin]
        wl = workload(f"g{n}", n, damp, crop(med, o, n), dt, tuple(int(v) + (n - n_int) // 2 for v in c), m, f0,
                      nt, rec, R=R)
        if cp is not None:
            cpml_workload(wl, damp, med["vmax"], **cp)
        s = W.make_solver(wl)
        if with_lb:
            lb.attach(s, z)
        s.run(nt)
        tr = s.traces().astype(np.float64)
        s.close()
        return tr

    ref = run(n_ref, 32, cp=dict(CPML_REF, alpha_max=np.pi * f0)) if ext else run(n_ref, 24)
    out = {"reference_grid": n_ref, "thin_grid": n_thin, "steps": nt, "base": base}

    def err(t):
        return float(np.linalg.norm(t - ref) / np.linalg.norm(ref))

    cp = CPML_THIN if base == "cpml" else None
    out["thin_plain" if base == "plain" else "thin_cpml"] = err(run(n_thin, L, R=R_thin, cp=cp))
    if lb is not None:
        out["thin_learned"] = err(run(n_thin, L, R=R_thin, with_lb=True, cp=cp))
    out[f"plain_{fallback}"] = err(run(n_int + 2 * fallback, fallback))
    return out


def save(lb: fwe.LearnedBoundary, path, **meta):
    """Parameters of the six faces (FP32 masters) and metadata, .npz."""
    import json

    p = np.stack([lb.get_params(f) for f in range(6)])
    np.savez(path, params=p, layer=np.int32(lb.layer), meta=np.array(json.dumps(meta)))


def load(path, device=0):
    """A LearnedBoundary with the parameters saved by `save` (and the
    correction form recorded there), and its metadata."""
    import json

    with np.load(path, allow_pickle=False) as d:
        lb = fwe.LearnedBoundary(int(d["layer"]), device=device)
        for f in range(6):
            lb.set_params(f, d["params"][f])
        meta = json.loads(str(d["meta"]))
        lb.set_correction_form(int(meta.get("form", fwe.LB_DISSIPATIVE)))
        return lb, meta
