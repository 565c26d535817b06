"""Test oracle: CPU restatement of the reference's FDTD step.

TEST INFRASTRUCTURE ONLY.  Imported by tests/, __graft_entry__.smoke() and
bench.py's cpu_baseline leg, as the checker; the product (libfwe.so and
elanfmt_amd/) never imports it.

Parity status of each oracle function (what pins it; see tests/):
  stress_iso / stress_ortho / velocity -- spectral (FFT) re-derivation of the
      same staggered scheme in a periodic box (tests/test_oracle_spectral.py),
      iso == ortho cross-check, Mxx<->Myy mirror symmetry, superposition,
      second-order convergence of the plane-wave speed.
  damping indexing (PROFILE w[i-1] vs mean(w[i],w[i+1]); WEIGHTS3D on the
      old value) -- single-entry probes (tests/test_oracle_pins.py).
  bc1 -- invariants Szz(k=0)=0, Sxz(k=1)=Sxz(k=2)/3 (test_oracle_pins.py).
  inject -- one step from rest puts coef*stf exactly at the OK4 / ANISO
      offsets (test_oracle_pins.py).
  record -- index identity on a synthetic field (test_oracle_pins.py).
The reference ships no golden vectors (its FK synthetics are absent and its
one point check has no recorded configuration), so no function is pinned to
reference output; the reference's CUDA sources cannot be built here (no
nvcc), so there is no oracle/_ref.
"""
from .pyoracle import Oracle, lib_path  # noqa: F401
