"""Benchmark of the FDTD time loop on one MI355X.

Default workload (--config 1): the layered orthorhombic 384x384x192 FP32
box with 24-cell separable damping on all faces, a lune-uniform moment
tensor, Ricker 3 Hz and 64 three-component receivers; one step = one full
velocity-stress update of all 28.3M cells.  Inputs are resident in HBM
before the timed region.  Prints one JSON line.

  python bench.py [--gpus 1] [--steps K] [--warmup 50] [--config 1]
                  [--kernel auto|two_pass|fused] [--cpu-steps 100] [--no-cpu]

Other configurations:
  --config 0  isotropic homogeneous 128^3, 600 steps (a parity case; fits
              in the last-level cache, so not a bandwidth figure)
  --config 2  six elementary-tensor wavefields over one 512x512x256 medium
              in one batched launch per half-step, against six plain runs
              of the same medium; then 1000 lune tensors synthesised by
              linear combination and 10 of them checked against direct runs
  --config 3  32 isotropic random-field cubes (edges 96..256) advanced as
              one group (one launch per half-step over all models), against
              the same models run one after the other; also 64 cubes of
              24..48 and 32 of 48..96 cells (--small-steps steps) both ways
  --config 4  the learned 6-cell boundary on a 512x512x256 grid: receiver
              errors against a 2x-padded reference (with the plain 6-cell
              and 24-cell layers beside it), and six elementary tensors with
              the learned layer against six plain 24-cell runs
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

PEAK_HBM_GBS = 8000.0
METRIC = "Gcell-updates/s + % of 8 TB/s HBM; batched-6-MT vs 6 plain runs at equal error"
DEFAULT_STEPS = {0: 600, 1: 3000, 2: 1000, 3: 1500, 4: 200}


def _traffic(kernel_name, geom, workload):
    """HBM bytes per launch measured with rocprofv3 --pmc (FETCH_SIZE and
    WRITE_SIZE, each scaled by the same counter on the copy probe), kept in
    profiles/traffic.json under "<kernel name> <geometry>", for exactly this
    kernel (template arguments included) launched with exactly this
    workgroup order and z chunk (geometry "linear kc2"; the sweep: "rows2
    kc12") on this workload; None when that launch was not profiled."""
    p = os.path.join(ROOT, "profiles", "traffic.json")
    try:
        with open(p) as f:
            d = json.load(f)[workload]
    except Exception:
        return None
    v = d.get(f"{kernel_name} {geom}")
    return None if v is None else v["bytes_per_launch"]


def _geom(fused, order, tiling, sweep):
    """Launch geometry that traffic.json keys carry besides the kernel name:
    workgroup order and z chunk (two-pass), rows per lane and z chunk (sweep)."""
    return f"rows{sweep[0]} kc{sweep[1]}" if fused else f"{order} kc{tiling[2]}"


def cpu_baseline(wl, steps, threads, fp32=False, repeats=3):
    """The FP64 oracle (CPU restatement of the reference step, OpenMP over z
    planes), or its FP32 build, timed on `steps` steps of one wavefield of
    the workload, `repeats` times; returns the median rate (Gcell/s) and
    the seconds of each timing (time loop only, no setup or I/O)."""
    from elanfmt_amd import workloads as W
    from oracle import Oracle

    o = Oracle(wl.nx, wl.ny, wl.nz, wl.dx, wl.dy, wl.dz, wl.dt, max(wl.nt, steps + 2), mode=wl.mode,
               pml=wl.pml, free_surface=wl.free_surface, fp32=fp32)
    W.apply(o, wl)
    o.run(2, threads)  # first touch
    els = []
    for _ in range(repeats):
        o.reset()
        t0 = time.perf_counter()
        o.run(steps, threads)
        els.append(time.perf_counter() - t0)
    o.close()
    return wl.cells * steps / float(np.median(els)) / 1e9, els


def host_threads():
    """Threads for the CPU baseline: the host CPUs this process may use,
    limited to OMP_NUM_THREADS when that is set (a GPU host's share of its
    CPUs); returns (threads, CPUs visible)."""
    visible = len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else (os.cpu_count() or 1)
    omp = int(os.environ.get("OMP_NUM_THREADS", "0") or 0)
    return (min(visible, omp) if omp > 0 else visible), visible


def _cpu_entry(args, wl_single, name):
    """CPU baseline beside the GPU figure: the FP64 oracle on `cpu_steps`
    steps of this workload (median of three timings) and its FP32 build;
    for the 384x384x192 case also the 128^3 isotropic case in full (600
    steps)."""
    if args.no_cpu or args.cpu_steps <= 0:
        return None
    from elanfmt_amd import workloads as W

    threads, visible = host_threads()
    v, els = cpu_baseline(wl_single, args.cpu_steps, threads)
    v32, els32 = cpu_baseline(wl_single, args.cpu_steps, threads, fp32=True)
    out = {"value": round(v, 5), "unit": "Gcell/s", "cores": threads, "host_cpus_visible": visible, "kind": "port",
           "sample": f"FP64 oracle, {args.cpu_steps} steps of {name}, median of 3 timings "
                     f"({', '.join(f'{e:.2f}' for e in els)} s), OpenMP over z planes",
           "fp32_value": round(v32, 5),
           "fp32_sample": f"FP32 build of the same oracle, {args.cpu_steps} steps, median of 3 "
                          f"({', '.join(f'{e:.2f}' for e in els32)} s)"}
    if args.config == 1:
        c0 = W.config(0)
        v0, els0 = cpu_baseline(c0, c0.nt, threads)
        out["config0_full"] = {"value": round(v0, 5), "unit": "Gcell/s", "steps": c0.nt,
                               "sample": f"FP64 oracle, {c0.name} all {c0.nt} steps, median of 3 "
                                         f"({', '.join(f'{e:.2f}' for e in els0)} s)"}
    return out


GRAPH_STEPS = 25  # steps per captured time-loop graph when replay is faster


def _wall(solver, steps):
    solver.reset()
    solver.synchronize()
    t0 = time.perf_counter()
    solver.run(steps)
    solver.synchronize()
    return time.perf_counter() - t0


def _timed_run(solver, steps, warmup, autotune, graph=True):
    """autotune (library autotuner: every tile shape timed, fastest kept,
    state reset) -> warmup -> per-kernel durations from HIP events around
    every launch, on a run of min(steps, 1000) steps of its own -> launch
    mode (plain launches, or replay of a captured 25-step graph when a
    200-step trial says it is faster) -> the timed run: `steps` steps from
    rest, no events inside, synchronised on both sides.  Returns wall
    seconds, the per-kernel event times (ms, scaled to `steps` launches),
    the autotuner's step time and the graph length used (0: none)."""
    tune = solver.autotune(3) if autotune else None
    if warmup:
        solver.run(warmup)
    n_ev = max(1, min(steps, 1000))
    solver.reset()
    solver.synchronize()
    ms_a, ms_b, _ = solver.run_timed(n_ev)
    gsteps = 0
    if graph and steps >= 2 * GRAPH_STEPS:
        trial = min(steps, 200)
        plain = _wall(solver, trial)
        solver.set_graph_steps(GRAPH_STEPS)
        _wall(solver, GRAPH_STEPS)  # capture
        if _wall(solver, trial) < plain:
            gsteps = GRAPH_STEPS
        else:
            solver.set_graph_steps(0)
    el = _wall(solver, steps)
    return el, ms_a * steps / n_ev, ms_b * steps / n_ev, tune, gsteps


def _kernel_entries(wl, fused, tiling, ms_a, ms_b, steps, cells, medium=1, batch_tile=None, sweep=None):
    """Per-kernel durations and algorithmic bandwidth, and the dominant
    kernel (name as rocprofv3 prints it).  fused: the single-pass sweep,
    sweep = its (rows per lane, z chunk)."""
    from elanfmt_amd import workloads as W

    ba, bb = W.alg_bytes(wl.mode, wl.pml, wl.nbatch, fused=fused, medium=medium, batch_tile=batch_tile)
    bt = batch_tile or wl.nbatch
    # kernel law template argument: 2 ortho through the material table, 3 ortho on
    # isotropic arrays, 4 iso through the material table
    km = wl.mode
    if medium in (2, 3):
        km = 4 if wl.mode == 0 else medium
    fs = "true" if wl.free_surface == 1 else "false"
    if fused:
        kernels = {"sweep": {"ms_per_launch": round(ms_a / steps, 5), "alg_bytes_per_cell": ba,
                             "GBs": round(ba * cells / (ms_a / steps * 1e-3) / 1e9, 1)}}
        return kernels, ("sweep", ms_a, ba, f"k_sweep<{km}, {wl.pml}, {fs}, {sweep[0]}>"), ba + bb
    kernels = {
        "stress": {"ms_per_launch": round(ms_a / steps, 5), "alg_bytes_per_cell": ba,
                   "GBs": round(ba * cells / (ms_a / steps * 1e-3) / 1e9, 1)},
        "velocity": {"ms_per_launch": round(ms_b / steps, 5), "alg_bytes_per_cell": bb,
                     "GBs": round(bb * cells / (ms_b / steps * 1e-3) / 1e9, 1)},
    }
    if ms_a >= ms_b:
        dom = ("stress", ms_a, ba, f"k_stress<{km}, {wl.pml}, {fs}, {tiling[0]}, {tiling[1]}, {bt}>")
    else:
        dom = ("velocity", ms_b, bb, f"k_velocity<{km}, {wl.pml}, {tiling[0]}, {tiling[1]}, {bt}>")
    return kernels, dom, ba + bb


def _roofline(dom, steps, cells, wname, geom):
    name, ms, b, kname = dom
    ach = b * cells / (ms / steps * 1e-3) / 1e9
    traffic = _traffic(kname, geom, wname)
    return {"bound": "hbm", "kernel": name, "achieved": round(ach, 1), "peak": PEAK_HBM_GBS, "unit": "GB/s",
            "frac": round(ach / PEAK_HBM_GBS, 4), "traffic": traffic, "kernel_name": kname, "launch": geom,
            "alg_bytes_per_launch": b * cells}


MEDIUM = {"auto": 0, "arrays": 1, "table": 2, "isotropic": 3}  # fwe.MEDIUM_*
# the reference's only published figure for this path (Performance/PerfWE.m:3):
# 0.0244 s per iteration at 512^3 on an A100, precision and variant not
# stated; another grid than any line here, so vs_baseline stays null
PUBLISHED_A100 = {"value": 5.50, "unit": "Gcell/s", "hardware": "NVIDIA A100", "grid": [512, 512, 512],
                  "source": "Performance/PerfWE.m:3 (0.0244 s/iteration; Gcell/s = cells / time)"}
ORDER_NAMES = ("linear", "xcd", "plane split")  # fwe.ORDER_*


def _medium_desc(wl, medium, ntab):
    from elanfmt_amd import fwe

    if wl.mode == fwe.MODE_ISO:
        if medium == fwe.MEDIUM_TABLE:
            return f"material table of (c11, c12) ({ntab} entries, one index byte per cell)"
        return "c11, c12 arrays"
    if medium == fwe.MEDIUM_TABLE:
        return f"material table ({ntab} entries, one index byte per cell)"
    if medium == fwe.MEDIUM_ISOTROPIC:
        return "isotropic: C11, C12, three edge moduli and three buoyancy arrays"
    return "twelve streamed arrays"


# what each configuration's line measures, beside the metric's full text
MEASURES = {0: "single-wavefield Gcell-updates/s and % of 8 TB/s HBM (a cache-resident parity case)",
            1: "single-wavefield Gcell-updates/s and % of 8 TB/s HBM (the plain-stencil half of the metric)",
            2: "batched six elementary-MT wavefields against six plain runs (wavefield Gcell-updates/s)",
            3: "uneven batch of 32 models: Gcell-updates/s of the whole batch",
            4: "thin absorbing layer: batched-6 interior cell-updates/s on the 6-cell CPML grid, errors of every "
               "layer form against a padded reference, and six plain runs at equal error"}


def _line(value, steps, warmup, el, config, **extra):
    out = {"metric": METRIC, "value": round(value, 3), "unit": "Gcell/s", "n_gpus": 1, "steps": steps,
           "warmup": warmup, "ms_per_step": round(el / steps * 1e3, 4), "higher_is_better": True,
           "scaling": "weak", "vs_baseline": None, "dtype": "f32", "data": "synthetic", "config": config}
    out.update({k: v for k, v in extra.items() if v is not None})
    return out


def bench_plain(args, steps):
    """One wavefield, the plain two-pass (or single-pass) kernel."""
    from elanfmt_amd import fwe, workloads as W

    wl = W.config(args.config, nt=max(steps, 1))
    kind = {"auto": fwe.KERNEL_AUTO, "two_pass": fwe.KERNEL_TWO_PASS, "fused": fwe.KERNEL_FUSED}[args.kernel]
    copy_ms = fwe.copy_probe(1 << 30, reps=10)
    copy_gbs = 2 * (1 << 30) / (copy_ms * 1e-3) / 1e9
    s = W.make_solver(wl)
    s.set_kernel(kind)
    s.set_medium_encoding(MEDIUM[args.medium])
    el, ms_a, ms_b, tune, gsteps = _timed_run(s, steps, args.warmup, not args.no_autotune)
    # the schedule the run used (under AUTO the autotuner times both)
    fused = s.kernel() == fwe.KERNEL_FUSED
    sweep = s.sweep_tiling()
    finite = bool(np.isfinite(s.traces()).all())
    tiling = (s.tile_cols(),) + tuple(s.tiling())
    order = ORDER_NAMES[s.wg_order()]
    enc, ntab = s.medium_encoding()
    tr_main = s.traces()
    s.close()
    cells = wl.cells
    kernels, dom, bstep = _kernel_entries(wl, fused, tiling, ms_a, ms_b, steps, cells, enc, sweep=sweep)
    # the same workload with the medium streamed from its arrays, in the
    # same process (the encoding changes which bytes stream, not results)
    arrays_run = None
    if enc in (fwe.MEDIUM_TABLE, fwe.MEDIUM_ISOTROPIC):
        n2 = min(steps, 1000)
        s2 = W.make_solver(wl)
        s2.set_kernel(kind)
        s2.set_medium_encoding(fwe.MEDIUM_ARRAYS)
        el2, a2, b2, _, _ = _timed_run(s2, n2, min(args.warmup, 20), not args.no_autotune)
        nr = len(wl.receivers)
        tr2 = s2.traces()
        # the first n2 samples of every trace, bit for bit
        same = bool(np.array_equal(tr2.reshape(nr, -1)[:, :n2], tr_main.reshape(nr, -1)[:, :n2]))
        t2 = (s2.tile_cols(),) + tuple(s2.tiling())
        f2 = s2.kernel() == fwe.KERNEL_FUSED
        sw2 = s2.sweep_tiling()
        s2.close()
        _, _, bst2 = _kernel_entries(wl, f2, t2, a2, b2, n2, cells, fwe.MEDIUM_ARRAYS, sweep=sw2)
        arrays_run = {"value": round(cells * wl.nbatch * n2 / el2 / 1e9, 3), "unit": "Gcell/s", "steps": n2,
                      "ms_per_step": round(el2 / n2 * 1e3, 4), "kernel": "sweep" if f2 else "two-pass",
                      "tiling": list(sw2) if f2 else list(t2), "alg_bytes_per_cell": bst2,
                      "hbm_pct_step": round(100 * bst2 * cells * n2 / el2 / 1e9 / PEAK_HBM_GBS, 2),
                      "traces_identical_to_main_run": same}
    desc = (f"{wl.name}: layered orthorhombic FP32, {wl.info.get('pml_cells')}-cell separable damping, "
            f"lune MT, Ricker {wl.info.get('f0')} Hz, {len(wl.receivers)} receiver traces"
            if wl.mode == fwe.MODE_ORTHO else
            f"{wl.name}: isotropic homogeneous FP32, {wl.info.get('pml_cells')}-cell profile damping, "
            f"double couple, Ricker {wl.info.get('f0')} Hz, {len(wl.receivers)} receiver traces")
    config = {"workload": desc, "grid": [wl.nx, wl.ny, wl.nz], "wavefields": wl.nbatch, "nt": wl.nt,
              "kernel": "single-pass sweep" if fused else "two-pass",
              "tiling": ({"rows_per_lane": sweep[0], "zchunk": sweep[1]} if fused else list(tiling)),
              "wg_order": None if fused else order,
              "medium": _medium_desc(wl, enc, ntab),
              "autotune_ms_per_step": None if tune is None else round(tune, 4), "graph_steps": gsteps,
              "parallelism": "single GPU"}
    return _line(cells * wl.nbatch * steps / el / 1e9, steps, args.warmup, el, config,
                 hbm_pct_step=round(100 * bstep * cells * steps / el / 1e9 / PEAK_HBM_GBS, 2),
                 kernels=kernels, roofline=_roofline(dom, steps, cells, wl.name, _geom(fused, order, tiling, sweep)),
                 copy_probe_GBs=round(copy_gbs, 1), traces_finite=finite, medium_arrays_run=arrays_run,
                 published_reference=PUBLISHED_A100,
                 cpu_baseline=_cpu_entry(args, W.config(args.config, nt=args.cpu_steps + 2), wl.name))


def bench_batched(args, steps):
    """Six elementary-tensor wavefields sharing the 512x512x256 medium in one
    batched launch per half-step, against six plain single-wavefield runs
    over the same medium (one per elementary tensor, run one after the
    other); then 1000 seeded lune tensors by linear combination, 10 checked
    against direct single-tensor runs."""
    from elanfmt_amd import fwe, workloads as W

    m0 = 1e15
    wl = W.config(2, nt=max(steps, 1), m0=m0)
    cells = wl.cells
    print(f"# batched: {wl.name} x{wl.nbatch}, {steps} steps", file=sys.stderr, flush=True)
    s = W.make_solver(wl)
    s.set_medium_encoding(MEDIUM[args.medium])
    el_b, ms_a, ms_b, tune, gsteps = _timed_run(s, steps, args.warmup, not args.no_autotune)
    tiling = (s.tile_cols(),) + tuple(s.tiling())
    order = ORDER_NAMES[s.wg_order()]
    enc, ntab = s.medium_encoding()
    bt = s.batch_tile()
    # the autotuner may keep the single-pass sweep (one launch per step)
    fused = s.kernel() == fwe.KERNEL_FUSED
    sweep = s.sweep_tiling()
    kernels, dom, bstep = _kernel_entries(wl, fused, tiling, ms_a, ms_b, steps, cells, enc, bt, sweep=sweep)
    # 1000 lune tensors by linear combination of the six elementary runs
    mts = np.array([W.mt_lune_random(50000 + q, m0)[0] for q in range(1000)])
    t0 = time.perf_counter()
    synth = s.combine_traces(mts / m0)
    el_syn = time.perf_counter() - t0
    s.close()

    # six plain runs, one per elementary tensor
    nr = len(wl.receivers) // 6
    el_p, plain_tiling, plain_kernel = 0.0, None, None
    xs, ys, zs = wl.nx // 2 - 1, wl.ny // 2 - 1, wl.nz // 2 - 1
    base = W.config(2, nt=max(steps, 1), m0=m0)
    base.nbatch, base.stf = 1, base.stf[:1]
    base.receivers = [t for t in base.receivers if t[4] == 0]
    tuned = None
    for e in range(6):
        me = np.zeros(6)
        me[e] = m0
        base.injections = [t for t in fwe.mt_injections(fwe.SRC_LAYOUT_OK4, xs, ys, zs, me) if t[5] != 0]
        p = W.make_solver(base)
        p.set_medium_encoding(MEDIUM[args.medium])
        if tuned is None:
            p.autotune(3)
            tuned = (p.tile_cols(),) + tuple(p.tiling()) + (p.wg_order(),)
        else:
            p.set_tiling(*tuned[1:3])
            p.set_tile_cols(tuned[0])
            p.set_wg_order(tuned[3])
        el, _, _, _, _ = _timed_run(p, steps, args.warmup if e == 0 else 0, False)
        el_p += el
        plain_tiling = (p.tile_cols(),) + tuple(p.tiling())
        plain_kernel = "single-pass sweep" if p.kernel() == fwe.KERNEL_FUSED else "two-pass"
        p.close()
    print(f"# batched {el_b:.2f} s, six plain {el_p:.2f} s", file=sys.stderr, flush=True)

    # 10 direct single-tensor runs against the synthesised traces
    errs = []
    for q in range(10):
        base.injections = fwe.mt_injections(fwe.SRC_LAYOUT_OK4, xs, ys, zs, mts[q])
        d = W.make_solver(base)
        d.set_tiling(*tuned[1:3])
        d.set_tile_cols(tuned[0])
        d.run(steps)
        tr = d.traces().astype(np.float64)
        d.close()
        sy = synth[q].astype(np.float64)
        errs.append(float(np.linalg.norm(sy - tr) / max(np.linalg.norm(tr), 1e-300)))
    config = {"workload": f"{wl.name}: layered orthorhombic FP32, 24-cell separable damping, six elementary "
                          f"moment tensors (Mxx..Myz) as one batch of wavefields, {nr} receiver traces each",
              "grid": [wl.nx, wl.ny, wl.nz], "wavefields": 6, "nt": wl.nt,
              "kernel": "single-pass sweep, all six wavefields per launch" if fused else "two-pass batched",
              "tiling": ({"rows_per_lane": sweep[0], "zchunk": sweep[1]} if fused else list(tiling)),
              "wg_order": None if fused else order, "medium": _medium_desc(wl, enc, ntab),
              "wavefields_per_workgroup": None if fused else bt,
              "autotune_ms_per_step": None if tune is None else round(tune, 4), "graph_steps": gsteps,
              "parallelism": "single GPU"}
    return _line(cells * 6 * steps / el_b / 1e9, steps, args.warmup, el_b, config,
                 hbm_pct_step=round(100 * bstep * cells * steps / el_b / 1e9 / PEAK_HBM_GBS, 2),
                 kernels=kernels, roofline=_roofline(dom, steps, cells, wl.name + "x6", _geom(fused, order, tiling, sweep)),
                 batched_vs_plain={"batched_s": round(el_b, 3), "six_plain_runs_s": round(el_p, 3),
                                   "speedup": round(el_p / el_b, 3), "plain_tiling": list(plain_tiling),
                                   "plain_kernel": plain_kernel},
                 synthesis={"tensors": len(mts), "seconds": round(el_syn, 3),
                            "checked_direct_runs": len(errs), "max_rel_l2": float(f"{max(errs):.3e}")})


def bench_uneven(args, steps):
    """32 random-field cubes as one group (one stress and one velocity launch
    per step over all models' tiles) against the same models run one after
    the other, each with its own autotuned tiling; the value is the faster
    of the two."""
    from elanfmt_amd import fwe, workloads as W

    wls = W.config(3, nt=max(steps, 1))
    cells = sum(w.cells for w in wls)
    print(f"# uneven: {len(wls)} models, {cells / 1e6:.1f} M cells, {steps} steps", file=sys.stderr, flush=True)
    sol = [W.make_solver(w) for w in wls]
    for s in sol:
        s.set_medium_encoding(MEDIUM[args.medium])
    enc = sol[0].medium_encoding()[0] if all(s.medium_encoding()[0] == fwe.MEDIUM_ISOTROPIC for s in sol) \
        else fwe.MEDIUM_ARRAYS
    g = fwe.Group(sol)
    tune = g.autotune(3) if not args.no_autotune else None
    g.run(args.warmup)
    for s in sol:
        s.reset()
        s.synchronize()
    t0 = time.perf_counter()
    g.run_timed(steps)
    for s in sol:
        s.synchronize()
    el_g = time.perf_counter() - t0
    ms_grp = [0.0, 0.0]
    tr_g = [s.traces() for s in sol]
    g.close()
    tiling = sol[0].tiling()
    el_seq = 0.0
    for s, w in zip(sol, wls):
        el, a, b, _, _ = _timed_run(s, steps, 0, not args.no_autotune)
        el_seq += el
        ms_grp[0] += a
        ms_grp[1] += b
    same = all(np.array_equal(s.traces(), t) for s, t in zip(sol, tr_g))
    for s in sol:
        s.close()
    from elanfmt_amd import workloads as Wm

    ba, bb = Wm.alg_bytes(fwe.MODE_ORTHO, fwe.PML_SEPARABLE, medium=enc)
    config = {"workload": f"uneven batch: {len(wls)} isotropic random-field cubes (edges "
                          f"{sorted(set(w.nx for w in wls))}), ortho law, 16-cell separable damping, own lune MT "
                          f"and Ricker f0 each", "models": len(wls), "cells": cells, "nt": steps,
              "kernel": "two-pass group", "tiling": list(tiling), "medium": _medium_desc(wls[0], enc, 0),
              "alg_bytes_per_cell": ba + bb,
              "autotune_ms_per_step": None if tune is None else round(tune, 4), "parallelism": "single GPU"}
    small = [_small_group(n, e, args.small_steps) for n, e in ((64, (24, 32, 40, 48)), (32, (48, 64, 80, 96)))]
    # the job's throughput is that of the faster way to run it; both are reported
    el = el_g
    if el_seq < el_g:
        el = el_seq
        config.update({"kernel": "two-pass, the models one after the other, each autotuned", "tiling": "per model"})
    return _line(cells * steps / el / 1e9, steps, args.warmup, el, config,
                 hbm_pct_step=round(100 * (ba + bb) * cells * steps / el / 1e9 / PEAK_HBM_GBS, 2),
                 group_vs_sequential={"group_s": round(el_g, 3), "sequential_s": round(el_seq, 3),
                                      "speedup": round(el_seq / el_g, 3), "traces_identical": same,
                                      "sequential_stress_ms": round(ms_grp[0], 2),
                                      "sequential_velocity_ms": round(ms_grp[1], 2)},
                 group_vs_sequential_small_models=small)


def _small_group(nmodels, edges, steps):
    """Many small isotropic random-field cubes (6-cell separable damping) as
    one group against the same models one after the other, each autotuned
    and replaying a captured graph when that is faster: where one model
    cannot fill the GPU, one launch per half-step for all of them pays."""
    from elanfmt_amd import fwe, workloads as W

    wls = W.uneven_batch(nmodels=nmodels, seed=12, edges=edges, nt=steps + 60, L=6)
    cells = sum(w.cells for w in wls)
    print(f"# uneven, small: {nmodels} models, edges {edges}, {steps} steps", file=sys.stderr, flush=True)
    sol = [W.make_solver(w) for w in wls]
    g = fwe.Group(sol)
    g.autotune(3)
    g.run(20)
    for s in sol:
        s.reset()
        s.synchronize()
    t0 = time.perf_counter()
    g.run(steps)
    for s in sol:
        s.synchronize()
    el_g = time.perf_counter() - t0
    tr_g = [s.traces() for s in sol]
    g.close()
    el_s, graphs = 0.0, 0
    for s in sol:
        el, _, _, _, gsteps = _timed_run(s, steps, 20, True)
        el_s += el
        graphs += gsteps > 0
    same = all(np.array_equal(s.traces(), t) for s, t in zip(sol, tr_g))
    for s in sol:
        s.close()
    return {"models": nmodels, "edges": list(edges), "cells": cells, "steps": steps,
            "group_s": round(el_g, 4), "sequential_s": round(el_s, 4), "speedup": round(el_s / el_g, 2),
            "group_gcell_s": round(cells * steps / el_g / 1e9, 2), "sequential_graph_replays": graphs,
            "traces_identical": same}


def bench_learned(args, steps):
    """Thin absorbing layers at full size.  A 512x512x256 grid whose outer 6
    cells are the thin layer (interior 500x500x244) with: the plain 6-cell
    layer (the reference's damping profile), the learned networks over it,
    and the convolutional PML; plain layers of 24 and more cells over the
    same interior; all against a 2x-padded reference (1024x1024x512) bounded
    by a 32-cell convolutional PML, receiver errors over `eval_steps`.
    Throughput: the six elementary tensors as one batch on the thin grid
    against six plain runs (timed as one plain run x 6) at 24 cells and at
    the width whose error first reaches the thin layer's."""
    from elanfmt_amd import learned as Lr
    from elanfmt_amd import workloads as W

    L, L8, fb, Lref = 6, 8, 24, 32
    thin, inner = (512, 512, 256), (500, 500, 244)
    thin8 = tuple(q + 2 * L8 for q in inner)
    ref = (1024, 1024, 512)
    widths = sorted({fb} | {int(w) for w in args.plain_widths.split(",") if w})
    lb, meta = Lr.load(os.path.join(ROOT, "elanfmt_amd", "data", "learned_boundary_L6.npz"))
    seed, f0, nt = 4242, 3.0, args.eval_steps
    # CPML coefficients (fwe_gen_cpml_table): the 6-cell layer as measured
    # best at this size over R, power, the frequency shift and kappa
    cp_thin = dict(R=1e-4, kappa_max=1.0, alpha_max=0.0, power=2.0)
    cp_8 = dict(R=1e-4, kappa_max=1.0, alpha_max=np.pi * f0, power=2.0)
    cp_ref = dict(R=1e-6, kappa_max=1.0, alpha_max=np.pi * f0, power=2.0)
    # one medium for every run: generated on the interior and continued
    # unchanged outward (each absorbing layer emulates that continuation),
    # on a grid that holds the reference and the widest plain grid; each run
    # takes the window centred on the interior
    big = tuple(max(r, q + 2 * max(widths)) for r, q in zip(ref, inner))
    print(f"# learned: medium {inner} continued to {big} on the host", file=sys.stderr, flush=True)
    med = Lr.extended_medium(inner, big, seed)
    dt = W.cfl_dt(med["vmax"], Lr.DX, Lr.DX, Lr.DX, 0.5)
    tw = tuple(slice((b - t) // 2, (b - t) // 2 + t) for b, t in zip(big, thin))
    z = Lr.impedance([med["cells"][0][tw]], med["bcell"][tw])
    m, _, _ = W.mt_lune_random(seed, 1e15)
    # 128 receivers (8 x 16) on the plane 10 cells inside the thin layer, x/y/z velocity
    xs = np.linspace(L + 10, thin[0] - L - 11, 8).round().astype(int).tolist()
    ys = np.linspace(L + 10, thin[1] - L - 11, 16).round().astype(int).tolist()
    rec_thin = [(c, x, y, L + 10, 0) for y in ys for x in xs for c in range(3)]
    src_thin = (thin[0] // 2 - 1, thin[1] // 2 - 1, thin[2] // 2 - 1)

    def grid_run(g, damp, kind, R=1e-8, nsteps=nt, nbatch=1, timed=False):
        """kind: "plain" (the reference's profile), "learned" (networks over
        the plain layer) or "cpml"."""
        o = tuple((b - q) // 2 for b, q in zip(big, g))
        sh = tuple((q - t) // 2 for q, t in zip(g, thin))
        rec = [(c, i + sh[0], j + sh[1], k + sh[2], b) for (c, i, j, k, b) in rec_thin]
        src = tuple(a + b for a, b in zip(src_thin, sh))
        mm = m if nbatch == 1 else [np.eye(6)[e] * 1e15 for e in range(nbatch)]
        wl = Lr.workload(f"g{g}", g, damp, Lr.crop(med, o, g), dt, src, mm, f0, nsteps, rec, R=R, nbatch=nbatch)
        if kind == "cpml":
            W.use_cpml(wl, damp, vmax=med["vmax"], **(cp_ref if g == ref else (cp_8 if damp == L8 else cp_thin)))
        sv = W.make_solver(wl)
        if kind == "learned":
            lb.attach(sv, z)
        if timed:
            # autotuned (unless --no-autotune), then the median of three
            # timed runs from rest
            if not args.no_autotune:
                sv.autotune(3)
            sv.run(10)
            ts = []
            for _ in range(3):
                sv.reset()
                sv.synchronize()
                t0 = time.perf_counter()
                sv.run(nsteps)
                sv.synchronize()
                ts.append(time.perf_counter() - t0)
            sv.close()
            return float(np.median(ts))
        sv.run(nsteps)
        tr = sv.traces().astype(np.float64)
        sv.close()
        return tr

    def plain_grid(w):
        return tuple(q + 2 * w for q in inner)

    errs = None
    w_eq = w_bar = None
    if nt > 0:  # --eval-steps 0: timing only
        print(f"# learned: reference run {ref} (32-cell CPML) x {nt} steps", file=sys.stderr, flush=True)
        tref = grid_run(ref, Lref, "cpml")
        nref = np.linalg.norm(tref)
        errs = {}
        runs = [("thin_plain_6", thin, L, "plain", meta["R_thin"]), ("thin_learned_6", thin, L, "learned", meta["R_thin"]),
                ("thin_cpml_6", thin, L, "cpml", 1e-8), ("thin_cpml_8", thin8, L8, "cpml", 1e-8)]
        runs += [(f"plain_{w}", plain_grid(w), w, "plain", 1e-8) for w in widths]
        for name, g, damp, kind, R in runs:
            print(f"# learned: {name} {g}", file=sys.stderr, flush=True)
            t = grid_run(g, damp, kind, R)
            errs[name] = float(np.linalg.norm(t - tref) / nref)
        ok = [w for w in widths if errs[f"plain_{w}"] <= errs["thin_cpml_6"]]
        w_eq = min(ok) if ok else None
        ok3 = [w for w in widths if errs[f"plain_{w}"] <= 1e-3]
        w_bar = min(ok3) if ok3 else None
    print(f"# learned: timing {steps} steps", file=sys.stderr, flush=True)
    t_b6 = grid_run(thin, L, "cpml", nsteps=steps, nbatch=6, timed=True)
    t_b8 = grid_run(thin8, L8, "cpml", nsteps=steps, nbatch=6, timed=True)
    t_b6_lb = grid_run(thin, L, "learned", meta["R_thin"], nsteps=steps, nbatch=6, timed=True)
    t_c1 = grid_run(thin, L, "cpml", nsteps=steps, timed=True)
    t_t1 = grid_run(thin, L, "plain", meta["R_thin"], nsteps=steps, timed=True)
    w_cmp = w_eq if w_eq is not None else (max(widths) if errs is not None else fb)
    w_cmp3 = w_bar if w_bar is not None else (max(widths) if errs is not None else fb)
    # the widest plain layer still less accurate than the thin CPML: with
    # w_eq it brackets the plain width of equal error (a lower bound on the
    # speedup there)
    below = [w for w in widths if errs is not None and w_eq is not None and w < w_eq]
    w_lo = max(below) if below else None
    t_p = {w: grid_run(plain_grid(w), w, "plain", nsteps=steps, timed=True)
           for w in sorted({fb, w_cmp, w_cmp3} | ({w_lo} if w_lo else set()))}
    lb.close()
    cells_in = inner[0] * inner[1] * inner[2]
    value = cells_in * 6 * steps / t_b6 / 1e9
    config = {"workload": f"6-cell convolutional PML on 512x512x256 (interior 500x500x244), layered orthorhombic "
                          f"FP32 (generated on the interior, continued unchanged outward), six elementary moment "
                          f"tensors in one batch, Ricker {f0} Hz, 128 receivers x 3 components; also the 8-cell "
                          f"CPML and the 6-cell plain layer with and without the learned networks "
                          f"(elanfmt_amd/train_lb.py, {meta['runs']} runs)",
              "grid": list(thin), "interior": list(inner), "wavefields": 6, "eval_steps": nt,
              "reference": {"grid": list(ref), "boundary": f"{Lref}-cell convolutional PML", **cp_ref},
              "thin_cpml": cp_thin, "thin_cpml_8": cp_8, "parallelism": "single GPU"}
    eq = None
    if errs is not None:
        eq = {"plain_width_at_thin_cpml_error": w_eq,
              "compared_width": w_cmp,
              "note": ("smallest plain width whose error is at most the thin CPML's" if w_eq is not None else
                       f"no plain width up to {max(widths)} reaches the thin CPML's error; the widest is "
                       f"compared, so its speedup is a lower bound"),
              "bar_1e-3": {k: v <= 1e-3 for k, v in errs.items()},
              "plain_width_at_1e-3": w_bar, "compared_width_at_1e-3": w_cmp3}
    return _line(value, steps, 0, t_b6, config, unit_note="interior cell-updates/s (6 wavefields)",
                 trace_error_vs_padded_reference=errs, equal_error=eq,
                 throughput={"batched6_cpml_s": round(t_b6, 3),
                             "six_plain24_runs_s": round(6 * t_p[fb], 3),
                             "speedup_vs_six_plain24": round(6 * t_p[fb] / t_b6, 3),
                             f"six_plain{w_cmp}_runs_s": round(6 * t_p[w_cmp], 3),
                             "speedup_at_equal_error": round(6 * t_p[w_cmp] / t_b6, 3),
                             **({f"six_plain{w_lo}_runs_s": round(6 * t_p[w_lo], 3),
                                 f"speedup_vs_six_plain{w_lo}": round(6 * t_p[w_lo] / t_b6, 3),
                                 "bracket_note": f"plain {w_lo} cells: error {errs[f'plain_{w_lo}']:.2e}, above the "
                                                 f"thin CPML's {errs['thin_cpml_6']:.2e}, so the plain width of "
                                                 f"equal error lies between {w_lo} and {w_cmp} cells"}
                                if w_lo else {}),
                             "batched6_cpml8_s": round(t_b8, 3),
                             f"six_plain{w_cmp3}_runs_s": round(6 * t_p[w_cmp3], 3),
                             "speedup_cpml8_at_1e-3": round(6 * t_p[w_cmp3] / t_b8, 3),
                             "batched6_learned_plain6_s": round(t_b6_lb, 3),
                             "speedup_learned_vs_six_plain24": round(6 * t_p[fb] / t_b6_lb, 3),
                             "note": "six plain runs timed as 6 x one measured run; every run autotuned, then "
                                     "the median of three timings",
                             "cpml_thin_ms_per_step": round(t_c1 / steps * 1e3, 3),
                             "plain_thin_ms_per_step": round(t_t1 / steps * 1e3, 3),
                             **{f"plain{w}_ms_per_step": round(t / steps * 1e3, 3) for w, t in t_p.items()}})


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=None)
    ap.add_argument("--warmup", type=int, default=50)
    ap.add_argument("--config", type=int, default=1, choices=[0, 1, 2, 3, 4])
    ap.add_argument("--eval-steps", type=int, default=4000, help="config 4: steps of the error runs")
    ap.add_argument("--plain-widths", default="48,96,144,192,240",
                    help="config 4: plain-layer widths (besides 24) whose errors bracket the thin layer's")
    ap.add_argument("--kernel", default="auto", choices=["auto", "two_pass", "fused"])
    ap.add_argument("--cpu-steps", type=int, default=100)
    ap.add_argument("--small-steps", type=int, default=400, help="config 3: steps of the small-model groups")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-autotune", action="store_true")
    ap.add_argument("--medium", default="auto", choices=sorted(MEDIUM),
                    help="ortho medium: the material table when it fits, else the isotropic arrays when they "
                         "apply (auto); always the twelve arrays; the table; the isotropic arrays")
    args = ap.parse_args()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) > 1:
        print("bench.py measures one GPU only: the library has no multi-GPU path "
              "(models and wavefields would be placed on GPUs whole, a single large model split in z-slabs); "
              "run with --gpus 1", file=sys.stderr)
        sys.exit(2)

    from elanfmt_amd import fwe

    if fwe.device_count() < 1:
        print("no HIP device visible", file=sys.stderr)
        sys.exit(1)
    steps = args.steps if args.steps is not None else DEFAULT_STEPS[args.config]
    if args.config in (0, 1):
        out = bench_plain(args, steps)
    elif args.config == 2:
        out = bench_batched(args, steps)
    elif args.config == 3:
        out = bench_uneven(args, steps)
    else:
        out = bench_learned(args, steps)
    out["measures"] = MEASURES[args.config]
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
