// fwe_stencil.h -- the per-cell update formulas shared by every kernel form
// (two-pass single, batched, multi-model group and single-pass).  Every
// multiply-add is written out with fmaf and the library is compiled with
// -ffp-contract=off, so the rounding of a cell update never depends on how
// the compiler schedules the surrounding kernel: all forms give the same
// bits for the same inputs.
#pragma once
#include <hip/hip_runtime.h>

namespace fwe {

// a*ra + b*rb (one strain-rate pair, e.g. dVy/dx + dVx/dy)
__device__ __forceinline__ float rate2(float a, float ra, float b, float rb) { return fmaf(a, ra, b * rb); }

// a*ra + b*rb + c*rc (a stress-divergence row, or C:eps for one component)
__device__ __forceinline__ float rate3(float a, float ra, float b, float rb, float c, float rc) {
  return fmaf(c, rc, fmaf(b, rb, a * ra));
}

// isotropic normal stress, OK4.cu:307-309: f * (s + (cx ex + cy ey + cz ez))
__device__ __forceinline__ float iso_normal(float f, float s, float cx, float ex, float cy, float ey, float cz,
                                            float ez) {
  return f * (s + rate3(cx, ex, cy, ey, cz, ez));
}

// isotropic shear stress, OK4.cu:323-329: w * (s + c d)
__device__ __forceinline__ float iso_shear(float w, float s, float c, float d) { return w * fmaf(c, d, s); }

// orthorhombic normal stress, ANISO.cu:311-324: s w + (cx ex + cy ey + cz ez)
__device__ __forceinline__ float ortho_normal(float w, float s, float cx, float ex, float cy, float ey, float cz,
                                              float ez) {
  return fmaf(s, w, rate3(cx, ex, cy, ey, cz, ez));
}

// orthorhombic shear stress, ANISO.cu:326-352: s w + c d
__device__ __forceinline__ float ortho_shear(float w, float s, float c, float d) { return fmaf(s, w, c * d); }

// velocity, OK4.cu:351-378 (iso: w (v + b d)) and ANISO.cu:374-386 (ortho: v w + b d)
__device__ __forceinline__ float iso_vel(float w, float v, float b, float d) { return w * fmaf(b, d, v); }
__device__ __forceinline__ float ortho_vel(float w, float v, float b, float d) { return fmaf(v, w, b * d); }

}  // namespace fwe
