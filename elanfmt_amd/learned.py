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


_CIJ_EXTRA = [(0, 0, 0)] * 6 + [(0, 1, 1), (1, 0, 1), (1, 1, 0)]
_B_EXTRA = [(1, 0, 0), (0, 1, 0), (0, 0, 1)]


def crop(med, o, n):
    """Window [o, o+n) per axis of a medium (staggered arrays keep their +1);
    o and n are scalars or per-axis triples."""
    o3, n3 = _dims(o), _dims(n)

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


@dataclass
class TrainLog:
    runs: int = 0
    steps: int = 0
    first_loss: float = float("nan")
    last_loss: float = float("nan")
    losses: list = None


def train(lb: fwe.LearnedBoundary, runs=200, n_int=40, warm=None, steps=None, lr=5e-4, seed0=1000, f0=(3.0, 8.0),
          R_thin=1e-3, resync_every=8, log_every=0, progress_every=0):
    """Training on `runs` thick-damping runs over seeded random layered
    media and lune moment tensors.  n_int is the interior edge, or a
    sequence of edges drawn per run.  Each run: `warm` plain steps (default
    1.5 n_int, until the first wavefront reaches the layer), then `steps`
    training steps (default 5 n_int; one Adam step per face each); the thin
    run is re-synchronised to the thick run every `resync_every` steps and
    rolls out with its own corrections in between."""
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
        med = layered_medium(n_thick, seed)
        dt = W.cfl_dt(med["vmax"], DX, DX, DX, 0.5)
        z = impedance(med["cells"], med["bcell"])
        ff = float(rng.uniform(*f0))
        m, _, _ = W.mt_lune_random(seed * 17 + 5, 1e15)
        c = ni // 2 + rng.integers(-ni // 6, ni // 6 + 1, 3)
        nt = wu + nsteps + 2
        taper = teacher_taper(n_thin, L, off, dt, med["vmax"], R_thin)
        thick = W.make_solver(workload("thick", n_thick, THICK, med, dt, tuple(int(v) + THICK for v in c), m, ff, nt,
                                       taper=taper))
        thin = W.make_solver(workload("thin", n_thin, L, crop(med, off, n_thin), dt, tuple(int(v) + L for v in c),
                                      m, ff, nt, R=R_thin))
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
             cfl=0.5):
    """Relative L2 error of receiver traces near the boundary against a
    padded reference, for the thin plain layer, the thin layer with the
    networks (if lb), and a `fallback`-cell plain layer."""
    L = lb.layer if lb is not None else 6
    n_thin = n_int + 2 * L
    rng = np.random.default_rng(seed)
    # padding so that nothing returns from the reference's own boundary
    cells_per_step = cfl / np.sqrt(3.0)
    pad = int(np.ceil(cells_per_step * nt / 2 + 24 + 8))
    n_ref = n_int + 2 * pad
    med = layered_medium(n_ref, seed)
    dt = W.cfl_dt(med["vmax"], DX, DX, DX, cfl)
    z = impedance(med["cells"], med["bcell"])
    m, _, _ = W.mt_lune_random(seed * 31 + 1, 1e15)
    c = n_int // 2 + rng.integers(-n_int // 8, n_int // 8 + 1, 3)
    rec_thin = face_receivers(n_thin, L)

    def run(n, damp, R=1e-8, with_lb=False):
        o = (n_ref - n) // 2
        shift = (n - n_thin) // 2
        rec = [(q, i + shift, j + shift, k + shift, b) for (q, i, j, k, b) in rec_thin]
        wl = workload(f"g{n}", n, damp, crop(med, o, n), dt, tuple(int(v) + (n - n_int) // 2 for v in c), m, f0,
                      nt, rec, R=R)
        s = W.make_solver(wl)
        if with_lb:
            lb.attach(s, z)
        s.run(nt)
        tr = s.traces().astype(np.float64)
        s.close()
        return tr

    ref = run(n_ref, 24)
    out = {"reference_grid": n_ref, "thin_grid": n_thin, "steps": nt}

    def err(t):
        return float(np.linalg.norm(t - ref) / np.linalg.norm(ref))

    out["thin_plain"] = err(run(n_thin, L, R=R_thin))
    if lb is not None:
        out["thin_learned"] = err(run(n_thin, L, R=R_thin, with_lb=True))
    out[f"plain_{fallback}"] = err(run(n_int + 2 * fallback, fallback))
    return out


def save(lb: fwe.LearnedBoundary, path, **meta):
    """Parameters of the six faces (FP32 masters) and metadata, .npz."""
    import json

    p = np.stack([lb.get_params(f) for f in range(6)])
    np.savez(path, params=p, layer=np.int32(lb.layer), meta=np.array(json.dumps(meta)))


def load(path, device=0):
    """A LearnedBoundary with the parameters saved by `save`, and its metadata."""
    import json

    with np.load(path, allow_pickle=False) as d:
        lb = fwe.LearnedBoundary(int(d["layer"]), device=device)
        for f in range(6):
            lb.set_params(f, d["params"][f])
        return lb, json.loads(str(d["meta"]))
