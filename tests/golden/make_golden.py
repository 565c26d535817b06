"""Write the golden fixtures of tests/golden/ from the FP64 oracle.

The reference ships no golden vectors (its FK synthetics are absent, its one
recorded point check has no configuration), so these fixtures are produced
by the oracle -- itself pinned by tests/test_oracle_*.py -- on small seeded
workloads.  Each fixture stores the receiver traces, a SHA-256 of every
input array (to catch generator drift) and the case parameters.

  python tests/golden/make_golden.py
"""
import hashlib
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))


def cases():
    from elanfmt_amd import fwe, workloads as W

    out = {}
    # iso, OK4 damping profiles, free surface, OK4 injection order
    wl = W.iso_homogeneous(n=32, L=5, nt=60, rec_depth=2, nrec_side=2, free_surface=1, f0=20.0)
    out["iso32_fs"] = wl
    # ortho layered, separable damping, no free surface
    wl = W.ortho_layered(nx=40, ny=36, nz=32, L=5, nt=120, nrec_side=2, seed=3, f0=20.0)
    out["ortho40_sep"] = wl
    # ortho layered, 3-D weights, free surface, ANISO injection layout/order
    wl = W.ortho_layered(nx=36, ny=34, nz=30, L=5, nt=120, nrec_side=2, seed=5, pml=fwe.PML_WEIGHTS3D, f0=20.0)
    wl.free_surface = 1
    wl.injections = fwe.mt_injections(fwe.SRC_LAYOUT_ANISO, 16, 15, 3, wl.info["mt"])
    out["ortho36_w3d_fs"] = wl
    return out


def input_digest(wl):
    h = hashlib.sha256()
    arrs = []
    if wl.c11 is not None:
        arrs += [wl.c11, wl.c12, np.array([wl.c66, *wl.bscal])]
    if wl.cij is not None:
        arrs += list(wl.cij) + list(wl.buoy)
    if wl.prof is not None:
        arrs += list(wl.prof)
    if wl.w3 is not None:
        arrs.append(wl.w3)
    arrs += list(wl.stf)
    arrs.append(np.array([t[:5] for t in wl.injections], np.int64))
    arrs.append(np.array([t[5] for t in wl.injections], np.float32))
    arrs.append(np.array(wl.receivers, np.int64))
    for a in arrs:
        h.update(np.ascontiguousarray(np.asarray(a, np.float32 if np.asarray(a).dtype.kind == "f" else np.int64)).tobytes())
    h.update(np.float64(wl.dt).tobytes())
    return h.hexdigest()


def oracle_traces(wl):
    from elanfmt_amd import workloads as W
    from oracle import Oracle

    o = Oracle(wl.nx, wl.ny, wl.nz, wl.dx, wl.dy, wl.dz, wl.dt, wl.nt, mode=wl.mode, pml=wl.pml,
               free_surface=wl.free_surface)
    W.apply(o, wl)
    o.run(wl.nt)
    return o.traces()


def main():
    for name, wl in cases().items():
        tr = oracle_traces(wl)
        np.savez_compressed(os.path.join(HERE, f"{name}.npz"), traces=tr, digest=input_digest(wl),
                            shape=np.array([wl.nx, wl.ny, wl.nz, wl.nt]), mode=wl.mode, pml=wl.pml,
                            free_surface=wl.free_surface)
        print(name, tr.shape, float(np.abs(tr).max()))


if __name__ == "__main__":
    main()
