// fwe_kernels.hip -- launch dispatch of the step kernels (instantiated per
// law and damping form in fwe_step.hip) and the small kernels of libfwe.so:
// the graph step counter, the memory-copy probe and the trace combination.
#include "fwe_internal.h"

namespace fwe {

// Step counter of a captured time-loop graph: one lane advances it after the
// G steps of one replay (graph nodes run in order on one stream).
__global__ void k_advance(int32_t* it_base, int32_t g, int32_t set) {
  if (threadIdx.x == 0 && blockIdx.x == 0) *it_base = set ? g : *it_base + g;
}

// Plain vector copy, the memory-copy probe (Performance/PerfMC.m).
// Each workgroup copies 4 x 256 16-byte vectors: all four loads are issued
// before the first store, so every wave has 4 KiB in flight.
typedef float v4f __attribute__((ext_vector_type(4)));
__global__ __launch_bounds__(256) void k_copy(const v4f* __restrict__ src, v4f* __restrict__ dst,
                                              int64_t n4) {
  const int64_t base = (int64_t)blockIdx.x * 1024 + threadIdx.x;
  v4f a[4];
#pragma unroll
  for (int u = 0; u < 4; ++u) {
    const int64_t q = base + u * 256;
    if (q < n4) a[u] = __builtin_nontemporal_load(src + q);
  }
#pragma unroll
  for (int u = 0; u < 4; ++u) {
    const int64_t q = base + u * 256;
    if (q < n4) __builtin_nontemporal_store(a[u], dst + q);
  }
}

// Seismograms of arbitrary moment tensors from the traces of the elementary
// ones, by linearity of the wave equation in the source:
//   out[m][q] = sum_e coef[m][e] * tr[e][q],  q < per (receivers x samples).
// Each thread reads its nb elementary samples once and writes mc outputs,
// so the kernel is bound by the output stream.
__global__ __launch_bounds__(256) void k_combine(const float* __restrict__ tr, int32_t nb, int64_t per,
                                                 const float* __restrict__ coef, int32_t mc,
                                                 float* __restrict__ out) {
  const int64_t q = (int64_t)blockIdx.x * 256 + threadIdx.x;
  if (q >= per) return;
  float e[FWE_MAX_BATCH];
#pragma unroll
  for (int b = 0; b < FWE_MAX_BATCH; ++b) e[b] = b < nb ? tr[(int64_t)b * per + q] : 0.f;
  for (int m = 0; m < mc; ++m) {
    float s = 0.f;
#pragma unroll
    for (int b = 0; b < FWE_MAX_BATCH; ++b)
      if (b < nb) s = fmaf(coef[m * nb + b], e[b], s);
    __builtin_nontemporal_store(s, out + (int64_t)m * per + q);
  }
}

// ---------------------------------------------------------------------------
// dispatch: law x damping form -> instantiation unit
// ---------------------------------------------------------------------------
hipError_t launch_stress(const StepArgs& a, int mode, int pml, int fs, hipStream_t s) {
  if (pml == PML_CPML) {
    switch (mode) {
      case MODE_ISO: return stress_mp<MODE_ISO, PML_CPML>(a, fs, s);
      case MODE_ISO_TAB: return stress_mp<MODE_ISO_TAB, PML_CPML>(a, fs, s);
      case MODE_ORTHO: return stress_mp<MODE_ORTHO, PML_CPML>(a, fs, s);
      case MODE_ORTHO_TAB: return stress_mp<MODE_ORTHO_TAB, PML_CPML>(a, fs, s);
      case MODE_ORTHO_ISO: return stress_mp<MODE_ORTHO_ISO, PML_CPML>(a, fs, s);
      default: return hipErrorInvalidValue;
    }
  }
  if (mode == MODE_ISO_TAB) {
    if (pml == PML_PROFILE) return stress_mp<MODE_ISO_TAB, PML_PROFILE>(a, fs, s);
    if (pml == PML_NONE) return stress_mp<MODE_ISO_TAB, PML_NONE>(a, fs, s);
    return hipErrorInvalidValue;
  }
  if (mode == MODE_ISO) {
    if (pml == PML_PROFILE) return stress_mp<MODE_ISO, PML_PROFILE>(a, fs, s);
    if (pml == PML_NONE) return stress_mp<MODE_ISO, PML_NONE>(a, fs, s);
    return hipErrorInvalidValue;
  }
  if (mode == MODE_ORTHO_ISO) {
    switch (pml) {
      case PML_NONE: return stress_mp<MODE_ORTHO_ISO, PML_NONE>(a, fs, s);
      case PML_W3D: return stress_mp<MODE_ORTHO_ISO, PML_W3D>(a, fs, s);
      case PML_SEP: return stress_mp<MODE_ORTHO_ISO, PML_SEP>(a, fs, s);
      default: return hipErrorInvalidValue;
    }
  }
  if (mode == MODE_ORTHO_TAB) {
    switch (pml) {
      case PML_NONE: return stress_mp<MODE_ORTHO_TAB, PML_NONE>(a, fs, s);
      case PML_W3D: return stress_mp<MODE_ORTHO_TAB, PML_W3D>(a, fs, s);
      case PML_SEP: return stress_mp<MODE_ORTHO_TAB, PML_SEP>(a, fs, s);
      default: return hipErrorInvalidValue;
    }
  }
  switch (pml) {
    case PML_NONE: return stress_mp<MODE_ORTHO, PML_NONE>(a, fs, s);
    case PML_W3D: return stress_mp<MODE_ORTHO, PML_W3D>(a, fs, s);
    case PML_SEP: return stress_mp<MODE_ORTHO, PML_SEP>(a, fs, s);
    default: return hipErrorInvalidValue;
  }
}

hipError_t launch_velocity(const StepArgs& a, int mode, int pml, hipStream_t s) {
  if (pml == PML_CPML) {
    switch (mode) {
      case MODE_ISO: return velocity_mp<MODE_ISO, PML_CPML>(a, s);
      case MODE_ISO_TAB: return velocity_mp<MODE_ISO_TAB, PML_CPML>(a, s);
      case MODE_ORTHO: return velocity_mp<MODE_ORTHO, PML_CPML>(a, s);
      case MODE_ORTHO_TAB: return velocity_mp<MODE_ORTHO_TAB, PML_CPML>(a, s);
      case MODE_ORTHO_ISO: return velocity_mp<MODE_ORTHO_ISO, PML_CPML>(a, s);
      default: return hipErrorInvalidValue;
    }
  }
  if (mode == MODE_ISO_TAB) {
    if (pml == PML_PROFILE) return velocity_mp<MODE_ISO_TAB, PML_PROFILE>(a, s);
    if (pml == PML_NONE) return velocity_mp<MODE_ISO_TAB, PML_NONE>(a, s);
    return hipErrorInvalidValue;
  }
  if (mode == MODE_ISO) {
    if (pml == PML_PROFILE) return velocity_mp<MODE_ISO, PML_PROFILE>(a, s);
    if (pml == PML_NONE) return velocity_mp<MODE_ISO, PML_NONE>(a, s);
    return hipErrorInvalidValue;
  }
  if (mode == MODE_ORTHO_ISO) {
    switch (pml) {
      case PML_NONE: return velocity_mp<MODE_ORTHO_ISO, PML_NONE>(a, s);
      case PML_W3D: return velocity_mp<MODE_ORTHO_ISO, PML_W3D>(a, s);
      case PML_SEP: return velocity_mp<MODE_ORTHO_ISO, PML_SEP>(a, s);
      default: return hipErrorInvalidValue;
    }
  }
  if (mode == MODE_ORTHO_TAB) {
    switch (pml) {
      case PML_NONE: return velocity_mp<MODE_ORTHO_TAB, PML_NONE>(a, s);
      case PML_W3D: return velocity_mp<MODE_ORTHO_TAB, PML_W3D>(a, s);
      case PML_SEP: return velocity_mp<MODE_ORTHO_TAB, PML_SEP>(a, s);
      default: return hipErrorInvalidValue;
    }
  }
  switch (pml) {
    case PML_NONE: return velocity_mp<MODE_ORTHO, PML_NONE>(a, s);
    case PML_W3D: return velocity_mp<MODE_ORTHO, PML_W3D>(a, s);
    case PML_SEP: return velocity_mp<MODE_ORTHO, PML_SEP>(a, s);
    default: return hipErrorInvalidValue;
  }
}

hipError_t launch_sweep(const SweepArgs& a, int mode, int pml, int fs, hipStream_t s) {
  if (mode == MODE_ISO_TAB || mode == MODE_ISO) {
    const bool tab = mode == MODE_ISO_TAB;
    if (pml == PML_PROFILE) return tab ? sweep_mp<MODE_ISO_TAB, PML_PROFILE>(a, fs, s) : sweep_mp<MODE_ISO, PML_PROFILE>(a, fs, s);
    if (pml == PML_NONE) return tab ? sweep_mp<MODE_ISO_TAB, PML_NONE>(a, fs, s) : sweep_mp<MODE_ISO, PML_NONE>(a, fs, s);
    return hipErrorInvalidValue;
  }
  if (mode == MODE_ORTHO_ISO) {
    switch (pml) {
      case PML_NONE: return sweep_mp<MODE_ORTHO_ISO, PML_NONE>(a, fs, s);
      case PML_W3D: return sweep_mp<MODE_ORTHO_ISO, PML_W3D>(a, fs, s);
      case PML_SEP: return sweep_mp<MODE_ORTHO_ISO, PML_SEP>(a, fs, s);
      default: return hipErrorInvalidValue;
    }
  }
  if (mode == MODE_ORTHO_TAB) {
    switch (pml) {
      case PML_NONE: return sweep_mp<MODE_ORTHO_TAB, PML_NONE>(a, fs, s);
      case PML_W3D: return sweep_mp<MODE_ORTHO_TAB, PML_W3D>(a, fs, s);
      case PML_SEP: return sweep_mp<MODE_ORTHO_TAB, PML_SEP>(a, fs, s);
      default: return hipErrorInvalidValue;
    }
  }
  switch (pml) {
    case PML_NONE: return sweep_mp<MODE_ORTHO, PML_NONE>(a, fs, s);
    case PML_W3D: return sweep_mp<MODE_ORTHO, PML_W3D>(a, fs, s);
    case PML_SEP: return sweep_mp<MODE_ORTHO, PML_SEP>(a, fs, s);
    default: return hipErrorInvalidValue;
  }
}

hipError_t launch_group(const StepArgs* tab, const int32_t* wg, int nblk, int tx, int ty, int order, int mode, int pml,
                        int fs, bool vel, hipStream_t s) {
  if (mode == MODE_ORTHO_TAB || mode == MODE_ISO_TAB) return hipErrorInvalidValue;
  if (mode == MODE_ORTHO_ISO) {
    switch (pml) {
      case PML_NONE: return group_mp<MODE_ORTHO_ISO, PML_NONE>(tab, wg, nblk, tx, ty, order, fs, vel, s);
      case PML_W3D: return group_mp<MODE_ORTHO_ISO, PML_W3D>(tab, wg, nblk, tx, ty, order, fs, vel, s);
      case PML_SEP: return group_mp<MODE_ORTHO_ISO, PML_SEP>(tab, wg, nblk, tx, ty, order, fs, vel, s);
      default: return hipErrorInvalidValue;
    }
  }
  if (mode == MODE_ISO) {
    if (pml == PML_PROFILE) return group_mp<MODE_ISO, PML_PROFILE>(tab, wg, nblk, tx, ty, order, fs, vel, s);
    if (pml == PML_NONE) return group_mp<MODE_ISO, PML_NONE>(tab, wg, nblk, tx, ty, order, fs, vel, s);
    return hipErrorInvalidValue;
  }
  switch (pml) {
    case PML_NONE: return group_mp<MODE_ORTHO, PML_NONE>(tab, wg, nblk, tx, ty, order, fs, vel, s);
    case PML_W3D: return group_mp<MODE_ORTHO, PML_W3D>(tab, wg, nblk, tx, ty, order, fs, vel, s);
    case PML_SEP: return group_mp<MODE_ORTHO, PML_SEP>(tab, wg, nblk, tx, ty, order, fs, vel, s);
    default: return hipErrorInvalidValue;
  }
}

hipError_t launch_advance(int32_t* it_base, int32_t g, int32_t set, hipStream_t s) {
  hipLaunchKernelGGL(k_advance, dim3(1), dim3(64), 0, s, it_base, g, set);
  return hipGetLastError();
}

hipError_t launch_copy(const float* src, float* dst, int64_t n, hipStream_t s) {
  const int64_t n4 = n / 4;
  const int64_t blocks = (n4 + 1023) / 1024;
  if (blocks > 0x7fffffff) return hipErrorInvalidValue;
  hipLaunchKernelGGL(k_copy, dim3((unsigned)blocks), dim3(256), 0, s, reinterpret_cast<const v4f*>(src),
                     reinterpret_cast<v4f*>(dst), n4);
  return hipGetLastError();
}

hipError_t launch_combine(const float* tr, int nb, int64_t per, const float* coef, int mc, float* out,
                          hipStream_t s) {
  const int64_t blocks = (per + 255) / 256;
  if (nb < 1 || nb > FWE_MAX_BATCH || blocks > 0x7fffffff) return hipErrorInvalidValue;
  if (blocks == 0 || mc == 0) return hipSuccess;
  hipLaunchKernelGGL(k_combine, dim3((unsigned)blocks), dim3(256), 0, s, tr, nb, per, coef, mc, out);
  return hipGetLastError();
}

}  // namespace fwe
