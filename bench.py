"""Benchmark of the FDTD time loop on one MI355X.

Workload (default): the layered orthorhombic 384x384x192 FP32 box with
24-cell separable damping on all faces, a lune-uniform moment tensor,
Ricker 3 Hz and 64 three-component receivers; one step = one full
velocity-stress update of all 28.3M cells.  Inputs are resident in HBM
before the timed region.  Prints one JSON line.

  python bench.py [--gpus 1] [--steps 3000] [--warmup 50] [--config 1]
                  [--kernel auto|two_pass|fused] [--cpu-steps 300] [--no-cpu]
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

PEAK_HBM_GBS = 8000.0


def _traffic(kernel_name, workload):
    """HBM bytes per launch measured with rocprofv3 --pmc (FETCH_SIZE and
    WRITE_SIZE, each scaled by the same counter on the copy probe), kept in
    profiles/traffic.json."""
    p = os.path.join(ROOT, "profiles", "traffic.json")
    try:
        with open(p) as f:
            d = json.load(f)
        return d[workload][kernel_name]["bytes_per_launch"]
    except Exception:
        return None


def cpu_baseline(wl_builder, steps, threads):
    """The FP64 oracle (CPU restatement of the reference step, OpenMP over z
    planes) on a bounded number of steps of the same workload."""
    from elanfmt_amd import workloads as W
    from oracle import Oracle

    wl = wl_builder()
    o = Oracle(wl.nx, wl.ny, wl.nz, wl.dx, wl.dy, wl.dz, wl.dt, wl.nt, mode=wl.mode, pml=wl.pml,
               free_surface=wl.free_surface)
    W.apply(o, wl)
    o.run(2, threads)  # first touch
    t0 = time.perf_counter()
    o.run(steps, threads)
    el = time.perf_counter() - t0
    o.close()
    return wl.cells * steps / el / 1e9, el


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3000)
    ap.add_argument("--warmup", type=int, default=50)
    ap.add_argument("--config", type=int, default=1)
    ap.add_argument("--kernel", default="auto", choices=["auto", "two_pass", "fused"])
    ap.add_argument("--cpu-steps", type=int, default=300)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-autotune", action="store_true")
    args = ap.parse_args()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) > 1:
        print("bench.py measures one GPU only: this build has no multi-GPU path "
              "(the z-slab decomposition it would use is described in DESIGN.md); run with --gpus 1",
              file=sys.stderr)
        sys.exit(2)

    from elanfmt_amd import fwe, workloads as W

    if fwe.device_count() < 1:
        print("no HIP device visible", file=sys.stderr)
        sys.exit(1)
    steps = args.steps
    wl = W.config(args.config, nt=max(steps, 1))
    kind = {"auto": fwe.KERNEL_AUTO, "two_pass": fwe.KERNEL_TWO_PASS, "fused": fwe.KERNEL_FUSED}[args.kernel]
    # copy probe: achievable stream bandwidth on this device (1 GiB copy)
    copy_ms = fwe.copy_probe(1 << 30, reps=10)
    copy_gbs = 2 * (1 << 30) / (copy_ms * 1e-3) / 1e9

    s = W.make_solver(wl)
    s.set_kernel(kind)
    fused = s.kernel() == fwe.KERNEL_FUSED
    tune_ms = None
    if not args.no_autotune:
        # library autotuner: times each tile shape, keeps the fastest, resets
        tune_ms = s.autotune(3)
    s.run(args.warmup)
    s.synchronize()
    s.reset()
    s.synchronize()
    t0 = time.perf_counter()
    ms_a, ms_b, ms_w = s.run_timed(steps)
    s.synchronize()
    el = time.perf_counter() - t0
    tr = s.traces()
    finite = bool(__import__("numpy").isfinite(tr).all())
    tiling = s.tiling()
    s.close()

    cells = wl.cells
    ba, bb = W.alg_bytes(wl.mode, wl.pml, wl.nbatch, fused=fused)
    gcells = cells * wl.nbatch * steps / el / 1e9
    if fused:
        kernels = {"fused": {"ms_per_step": round(ms_a / steps, 5), "alg_bytes_per_cell": ba,
                             "GBs": round(ba * cells / (ms_a / steps * 1e-3) / 1e9, 1)}}
        dom, ms_dom, b_dom = "fused", ms_a, ba
        kname = f"k_fused<{wl.mode}, {wl.pml}, {tiling[0]}>"
    else:
        kernels = {
            "stress": {"ms_per_launch": round(ms_a / steps, 5), "alg_bytes_per_cell": ba,
                       "GBs": round(ba * cells / (ms_a / steps * 1e-3) / 1e9, 1)},
            "velocity": {"ms_per_launch": round(ms_b / steps, 5), "alg_bytes_per_cell": bb,
                         "GBs": round(bb * cells / (ms_b / steps * 1e-3) / 1e9, 1)},
        }
        dom, ms_dom, b_dom = ("stress", ms_a, ba) if ms_a >= ms_b else ("velocity", ms_b, bb)
        kname = (f"k_stress<{wl.mode}, {wl.pml}, {'true' if wl.free_surface else 'false'}, {tiling[0]}>"
                 if dom == "stress" else f"k_velocity<{wl.mode}, {wl.pml}, {tiling[0]}>")
    per_launch_s = ms_dom / steps * 1e-3
    ach = b_dom * cells / per_launch_s / 1e9
    step_gbs = (ba + bb) * cells * steps / el / 1e9
    wname = wl.name
    out = {
        "metric": "Gcell-updates/s + % of 8 TB/s HBM; batched-6-MT vs 6 plain runs at equal error",
        "value": round(gcells, 3),
        "unit": "Gcell/s",
        "n_gpus": 1,
        "steps": steps,
        "warmup": args.warmup,
        "ms_per_step": round(el / steps * 1e3, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {
            "workload": (f"{wname}: layered orthorhombic FP32, {wl.info.get('pml_cells')}-cell separable damping, "
                         f"lune MT, Ricker {wl.info.get('f0')} Hz, {len(wl.receivers)} receiver traces")
            if wl.mode == fwe.MODE_ORTHO else f"{wname}: isotropic homogeneous FP32",
            "grid": [wl.nx, wl.ny, wl.nz], "wavefields": wl.nbatch, "nt": wl.nt,
            "kernel": "fused single-pass" if fused else "two-pass", "tiling": list(tiling),
            "autotune_ms_per_step": None if tune_ms is None else round(tune_ms, 4),
            "parallelism": "single GPU",
        },
        "hbm_pct_step": round(100 * step_gbs / PEAK_HBM_GBS, 2),
        "kernels": kernels,
        "roofline": {"bound": "hbm", "kernel": dom, "achieved": round(ach, 1), "peak": PEAK_HBM_GBS,
                     "unit": "GB/s", "frac": round(ach / PEAK_HBM_GBS, 4),
                     "traffic": _traffic(kname, wname)},
        "copy_probe_GBs": round(copy_gbs, 1),
        "traces_finite": finite,
    }
    if not args.no_cpu and args.cpu_steps > 0:
        threads = min(16, os.cpu_count() or 1)
        v, el_cpu = cpu_baseline(lambda: W.config(args.config, nt=args.cpu_steps + 2), args.cpu_steps, threads)
        out["cpu_baseline"] = {"value": round(v, 5), "unit": "Gcell/s", "cores": threads, "kind": "port",
                               "sample": f"FP64 oracle, {args.cpu_steps} steps of {wname} "
                                         f"({el_cpu:.1f} s), OpenMP over z planes"}
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
