"""Train the learned thin boundary on the GPU and save its parameters.

  python -m elanfmt_amd.train_lb [--runs 200] [--out elanfmt_amd/data/learned_boundary_L6.npz]

Each run: a seeded random layered orthorhombic medium and lune moment
tensor on an interior of one of the --sizes edges, a teacher run on the
interior plus 32 cells per face (the thin layer's damping over the thin window, a damping ramp
beyond), and a thin run with a 6-cell layer; see elanfmt_amd/learned.py.
After training, the layer is evaluated on held-out media against a padded
reference, and the numbers are stored with the parameters.
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)


def main():
    from elanfmt_amd import fwe, learned

    ap = argparse.ArgumentParser()
    ap.add_argument("--runs", type=int, default=200)
    ap.add_argument("--layer", type=int, default=6)
    ap.add_argument("--lr", type=float, default=5e-4)
    ap.add_argument("--resync", type=int, default=8)
    ap.add_argument("--r-thin", type=float, default=1e-3)
    ap.add_argument("--seed0", type=int, default=1000)
    ap.add_argument("--sizes", default="40", help="interior edges drawn per run, comma separated")
    ap.add_argument("--f0", default="3,8", help="Ricker peak-frequency range (Hz) drawn per run, lo,hi")
    ap.add_argument("--init", default=None, help="start from the parameters of this .npz instead of a fresh init")
    ap.add_argument("--base", default="plain", choices=["plain", "cpml"],
                    help="the thin layer's own damping: the reference's profile, or the convolutional PML")
    ap.add_argument("--extend", action="store_true",
                    help="media generated on the interior and continued outward (always with --base cpml)")
    ap.add_argument("--form", default="dissipative", choices=["dissipative", "free"],
                    help="corrections applied in their dissipative part only, or as predicted")
    ap.add_argument("--eval-sizes", default="80:500:5,120:700:4", help="interior:steps:f0 of the larger evaluations")
    ap.add_argument("--out", default=None)
    a = ap.parse_args()
    if a.out is None:
        a.out = os.path.join(ROOT, "elanfmt_amd", "data",
                             "learned_boundary_L6.npz" if a.base == "plain" else "learned_boundary_cpml_L6.npz")
    if a.init:
        lb, _ = learned.load(a.init)
    else:
        lb = fwe.LearnedBoundary(a.layer)
    form = fwe.LB_FREE if a.form == "free" else fwe.LB_DISSIPATIVE
    lb.set_correction_form(form)
    t0 = time.time()
    sizes = [int(v) for v in a.sizes.split(",")]
    f0_range = tuple(float(v) for v in a.f0.split(","))
    log = learned.train(lb, runs=a.runs, n_int=sizes, lr=a.lr, resync_every=a.resync, R_thin=a.r_thin, f0=f0_range,
                        seed0=a.seed0, log_every=200, progress_every=10, base=a.base, extend=a.extend)
    t_train = time.time() - t0
    print(f"trained {log.runs} runs, {log.steps} steps in {t_train:.1f} s; loss {log.first_loss:.3e} -> "
          f"{log.last_loss:.3e}", flush=True)
    evals = {}
    ev = dict(R_thin=a.r_thin, base=a.base, extend=a.extend)
    for sd in (7, 8, 9):
        evals[sd] = learned.evaluate(lb, seed=sd, **ev)
        print(f"held-out medium {sd}: {evals[sd]}", flush=True)
    for spec in a.eval_sizes.split(","):
        n_int, nt, f0 = (t(v) for t, v in zip((int, int, float), spec.split(":")))
        evals[f"interior{n_int}"] = learned.evaluate(lb, seed=7, n_int=n_int, nt=nt, f0=f0, **ev)
        print(f"held-out medium 7, interior {n_int}: {evals[f'interior{n_int}']}", flush=True)
    learned.save(lb, a.out, runs=a.runs, layer=a.layer, lr=a.lr, resync_every=a.resync, R_thin=a.r_thin,
                 seed0=a.seed0, sizes=sizes, f0=list(f0_range), train_seconds=round(t_train, 1), steps=log.steps,
                 loss_curve=[(s, float(f"{l:.4e}")) for s, l in log.losses], evaluation=evals, base=a.base,
                 extend=bool(a.extend or a.base == "cpml"), cpml=learned.CPML_THIN if a.base == "cpml" else None,
                 form=form)
    print(json.dumps({"out": a.out, "evaluation": evals}), flush=True)
    lb.close()


if __name__ == "__main__":
    main()
