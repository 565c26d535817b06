// fwe_lb.hip -- kernels of the learned thin boundary: a per-face MLP that
// corrects the nine fields of the cells of an L-cell absorbing layer after
// every step, trained on the GPU against runs with a thick damping layer.
//
// Network of one face (81 -> 64 -> 64 -> 9, SiLU):
//   x  = stencil of the cell, 9 fields at 9 cells (the cell, its 6 face
//        neighbours, and two cells out along the face normal), stresses
//        divided by the impedance Z, the whole vector divided by its RMS r
//   a1 = W1 x + b1[depth]      h1 = silu(a1)
//   a2 = W2 h1 + b2            h2 = silu(a2)
//   y  = W3 h2 + b3            correction: dV = r y[0:3], dS = Z r y[3:9]
// applied after the step in its dissipative part only (lb_shrink).
// W1, W2, W3 are used rounded to bfloat16 (kept as FP32 master copies for
// training); all sums accumulate in FP32.
//
// Training (k_lb_mlp): one wavefront evaluates one sample, lane h owns
// hidden unit h, weights staged in LDS as bfloat16 for the whole block.
// Inference (k_lb_infer): 16 cells per wavefront on the matrix cores.
#include <algorithm>

#include "fwe_internal.h"

namespace fwe {

__device__ __forceinline__ uint16_t to_bf16(float f) {
  uint32_t u = __float_as_uint(f);
  if ((u & 0x7f800000u) == 0x7f800000u) return (uint16_t)(u >> 16);  // inf / nan
  u += 0x7fffu + ((u >> 16) & 1u);                                   // round to nearest even
  return (uint16_t)(u >> 16);
}
__device__ __forceinline__ float bf(uint16_t b) { return __uint_as_float((uint32_t)b << 16); }
__device__ __forceinline__ float sigm(float a) { return 1.f / (1.f + __expf(-a)); }

constexpr int LB_WAVES = 4;  // samples in flight per block

struct LbShared {
  uint16_t w1[LB_IN * LB_H];   // [k][h]
  uint16_t w2[LB_H * LB_H];    // [m][h]
  uint16_t w2t[LB_H * LB_H];   // [h][m]
  uint16_t w3[LB_H * LB_OUT];  // [h][o]
  float b1[LB_MAXL * LB_H];
  float b2[LB_H];
  float b3[LB_OUT];
  float xs[LB_WAVES][LB_IN + 3];
  float h1s[LB_WAVES][LB_H];
  float h2s[LB_WAVES][LB_H];
  float gs[LB_WAVES][12];
  float d2s[LB_WAVES][LB_H];
  float red[256];
};

__device__ void lb_stage(LbShared& sh, const float* __restrict__ P, int L, bool transpose) {
  const int t = threadIdx.x;
  for (int q = t; q < LB_IN * LB_H; q += 256) sh.w1[q] = to_bf16(P[lb_off_w1() + q]);
  for (int q = t; q < LB_H * LB_H; q += 256) {
    const uint16_t v = to_bf16(P[lb_off_w2(L) + q]);
    sh.w2[q] = v;
    if (transpose) sh.w2t[(q % LB_H) * LB_H + q / LB_H] = v;
  }
  for (int q = t; q < LB_H * LB_OUT; q += 256) sh.w3[q] = to_bf16(P[lb_off_w3(L) + q]);
  for (int q = t; q < L * LB_H; q += 256) sh.b1[q] = P[lb_off_b1() + q];
  for (int q = t; q < LB_H; q += 256) sh.b2[q] = P[lb_off_b2(L) + q];
  for (int q = t; q < LB_OUT; q += 256) sh.b3[q] = P[lb_off_b3(L) + q];
}

// hidden layers of the sample in sh.xs[w]; returns a1, a2 of lane h and
// leaves h2 in sh.h2s[w].  Every thread of the block calls it (barriers).
__device__ __forceinline__ void lb_hidden(LbShared& sh, int w, int h, int d, bool valid, float* h1out,
                                          float* h2out, float& a1, float& a2) {
  a1 = 0.f;
  if (valid) {
    float s = sh.b1[d * LB_H + h];
    for (int k = 0; k < LB_IN; ++k) s = fmaf(bf(sh.w1[k * LB_H + h]), sh.xs[w][k], s);
    a1 = s;
    const float h1 = a1 * sigm(a1);
    sh.h1s[w][h] = h1;
    if (h1out) h1out[h] = h1;
  }
  __syncthreads();
  a2 = 0.f;
  if (valid) {
    float s = sh.b2[h];
    for (int m = 0; m < LB_H; ++m) s = fmaf(bf(sh.w2[m * LB_H + h]), sh.h1s[w][m], s);
    a2 = s;
    const float h2 = a2 * sigm(a2);
    sh.h2s[w][h] = h2;
    if (h2out) h2out[h] = h2;
  }
  __syncthreads();
}

__device__ __forceinline__ float lb_output(const LbShared& sh, int w, int o) {
  float s = sh.b3[o];
  for (int h = 0; h < LB_H; ++h) s = fmaf(bf(sh.w3[h * LB_OUT + o]), sh.h2s[w][h], s);
  return s;
}

__device__ __forceinline__ float dsilu(float a) {
  const float s = sigm(a);
  return s * (1.f + a * (1.f - s));
}

// Forward (and, with T != nullptr, the backward pass to the activation
// gradients) of n samples of one face.  Loss = sum M (y - T)^2 * inv_count.
// Outputs per sample: Y [9], H1, H2, D1 = dL/da1, D2 = dL/da2 [64], G =
// dL/dy [9]; per block: the partial loss.
__global__ __launch_bounds__(256) void k_lb_mlp(const float* __restrict__ P, int32_t L, int32_t n,
                                               const float* __restrict__ X, const int32_t* __restrict__ depth,
                                               const float* __restrict__ T, const float* __restrict__ M,
                                               float inv_count, float* __restrict__ Y, float* __restrict__ H1,
                                               float* __restrict__ H2, float* __restrict__ D1,
                                               float* __restrict__ D2, float* __restrict__ G,
                                               float* __restrict__ loss_part) {
  __shared__ LbShared sh;
  const bool train = T != nullptr;
  lb_stage(sh, P, L, train);
  __syncthreads();
  const int w = threadIdx.x >> 6, h = threadIdx.x & 63;
  float lsum = 0.f;
  for (int64_t base = (int64_t)blockIdx.x * LB_WAVES; base < n; base += (int64_t)gridDim.x * LB_WAVES) {
    const int64_t s = base + w;
    const bool valid = s < n;
    int d = 0;
    if (valid) {
      sh.xs[w][h] = X[s * LB_IN + h];
      if (h + 64 < LB_IN) sh.xs[w][h + 64] = X[s * LB_IN + h + 64];
      d = depth[s];
    }
    __syncthreads();
    float a1, a2;
    lb_hidden(sh, w, h, d, valid, H1 ? H1 + s * LB_H : nullptr, H2 ? H2 + s * LB_H : nullptr, a1, a2);
    if (valid && h < LB_OUT) {
      const float y = lb_output(sh, w, h);
      if (Y) Y[s * LB_OUT + h] = y;
      if (train) {
        const float mk = M[s * LB_OUT + h];
        const float e = y - T[s * LB_OUT + h];
        lsum = fmaf(mk * e, e, lsum);
        const float g = 2.f * mk * e * inv_count;
        sh.gs[w][h] = g;
        G[s * LB_OUT + h] = g;
      }
    }
    __syncthreads();
    if (!train) continue;
    if (valid) {
      float dh2 = 0.f;
      for (int o = 0; o < LB_OUT; ++o) dh2 = fmaf(bf(sh.w3[h * LB_OUT + o]), sh.gs[w][o], dh2);
      const float da2 = dh2 * dsilu(a2);
      sh.d2s[w][h] = da2;
      D2[s * LB_H + h] = da2;
    }
    __syncthreads();
    if (valid) {
      float dh1 = 0.f;
      for (int q = 0; q < LB_H; ++q) dh1 = fmaf(bf(sh.w2t[q * LB_H + h]), sh.d2s[w][q], dh1);
      D1[s * LB_H + h] = dh1 * dsilu(a1);
    }
  }
  if (loss_part) {
    sh.red[threadIdx.x] = lsum;
    __syncthreads();
    if (threadIdx.x == 0) {
      float t = 0.f;
      for (int q = 0; q < 256; ++q) t += sh.red[q];
      loss_part[blockIdx.x] = t;
    }
  }
}

// Partial sums over chunks of samples:  part[c][p][q] = sum_{s in chunk c}
// a(s,p) * B[s][q], with a(s,p) = A[s][p] (A != nullptr), the one-hot of
// depth[s] (onehot), or 1.  Reduced in a fixed order by k_lb_reduce.
__global__ __launch_bounds__(256) void k_lb_atb(const float* __restrict__ A, int32_t lda,
                                               const int32_t* __restrict__ onehot, int32_t P,
                                               const float* __restrict__ B, int32_t ldb, int32_t Q, int32_t n,
                                               int32_t chunk, float* __restrict__ part) {
  const int q = threadIdx.x, p = blockIdx.y * 4 + threadIdx.y;
  const int c = blockIdx.x;
  if (q >= Q || p >= P) return;
  const int64_t s0 = (int64_t)c * chunk, s1 = min((int64_t)n, s0 + chunk);
  float acc = 0.f;
  for (int64_t s = s0; s < s1; ++s) {
    const float a = A ? A[s * lda + p] : (onehot ? (onehot[s] == p ? 1.f : 0.f) : 1.f);
    acc = fmaf(a, B[s * ldb + q], acc);
  }
  part[((int64_t)c * P + p) * Q + q] = acc;
}

__global__ void k_lb_reduce(const float* __restrict__ part, int32_t nchunk, int32_t count, float* __restrict__ out) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= count) return;
  float s = 0.f;
  for (int c = 0; c < nchunk; ++c) s += part[(int64_t)c * count + i];
  out[i] = s;
}

// Adam on one face's parameters (bias corrections c1 = 1/(1-b1^t), c2 = 1/(1-b2^t)).
__global__ void k_lb_adam(float* __restrict__ P, float* __restrict__ m, float* __restrict__ v,
                          const float* __restrict__ g, int32_t n, float lr, float b1, float b2, float eps, float c1,
                          float c2) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const float gi = g[i];
  const float mi = b1 * m[i] + (1.f - b1) * gi;
  const float vi = b2 * v[i] + (1.f - b2) * gi * gi;
  m[i] = mi;
  v[i] = vi;
  P[i] -= lr * (mi * c1) / (sqrtf(vi * c2) + eps);
}

// ---------------------------------------------------------------------------
// stencil gather from a field set (one wavefield, padded layout)
// ---------------------------------------------------------------------------

// The nine stencil cells of a layer cell, q = 0..8: the cell itself; its
// neighbours at -x, +x, -y, +y, -z, +z; and the cells two steps along the
// face normal on the negative and positive side.
//
// Input e (0..80) of the cell is field e % 9 at stencil cell q = e / 9,
// computed without branches: e differs between the lanes of a wavefront,
// so a branch or a switch here would make every gathered load wait for the
// previous one, whereas with selects all of a lane's loads issue together.
// Neighbours outside the grid (and inputs with use = false) read element 0
// and yield 0.
__device__ __forceinline__ float lb_input(const LbGeom& g, const LbCell& c, int e, bool use = true) {
  const int q = e / 9, f = e - q * 9;
  const int ax = (c.fd >> 8) >> 1;
  const int axis = q >= 7 ? ax : (q - 1) >> 1;  // q = 0: -1, no axis
  const int mag = q >= 7 ? 2 : (q >= 1 ? 1 : 0);
  const int d = (q & 1) ? -mag : mag;  // odd q: the negative side
  const int i = c.i + (axis == 0 ? d : 0), j = c.j + (axis == 1 ? d : 0), k = c.k + (axis == 2 ? d : 0);
  const bool ok = use && i >= 0 && j >= 0 && k >= 0 && i < g.nx && j < g.ny && k < g.nz;
  // 32-bit byte offset from the field set's base (the set is below 4 GiB,
  // checked on the host), masked arithmetically: written as a select of the
  // offset, the compiler turns it back into a branch around the load
  const uint32_t off = ((uint32_t)f * (uint32_t)g.fstride + (uint32_t)i + (uint32_t)j * (uint32_t)g.sy +
                        (uint32_t)k * (uint32_t)g.sz) * 4u & (0u - (uint32_t)ok);
  const float v = *reinterpret_cast<const float*>(reinterpret_cast<const char*>(g.F[0]) + off);
  return ok ? (f >= 3 ? v * g.zinv : v) : 0.f;
}

// Training features of n cells: X [n][81] normalised, R [n] = RMS, depth.
__global__ __launch_bounds__(256) void k_lb_gather(const LbGeom g, const LbCell* __restrict__ cells, int32_t n,
                                                  float* __restrict__ X, float* __restrict__ R,
                                                  int32_t* __restrict__ depth) {
  const int64_t s = (int64_t)blockIdx.x * 256 + threadIdx.x;
  if (s >= n) return;
  const LbCell c = cells[s];
  float ss = 0.f;
  for (int e = 0; e < LB_IN; ++e) {
    const float v = lb_input(g, c, e);
    X[s * LB_IN + e] = v;
    ss = fmaf(v, v, ss);
  }
  const float r = sqrtf(ss * (1.f / LB_IN)) + 1e-30f;
  const float ri = 1.f / r;
  for (int e = 0; e < LB_IN; ++e) X[s * LB_IN + e] *= ri;
  R[s] = r;
  depth[s] = c.fd & 0xff;
}

// Shear components outside the stress update's range at this cell (the
// reference never updates them) are masked.
__device__ __forceinline__ bool lb_comp_live(const LbCell& c, int f) {
  if (f == 6) return c.i > 0 && c.j > 0;
  if (f == 7) return c.i > 0 && c.k > 0;
  if (f == 8) return c.j > 0 && c.k > 0;
  return true;
}

// Targets: T = (thick(n+1) at cell+off - thin step) / (r * unit), unit = 1
// for velocities and Z for stresses.  Weights M = (r / r_ref)^2 on live
// components of cells with r >= rmin (wscale = 1 / r_ref^2), so the loss is
// the squared error in physical units relative to the batch's RMS level.
__global__ __launch_bounds__(256) void k_lb_targets(const LbGeom thin, const LbGeom thick, int32_t off,
                                                   const LbCell* __restrict__ cells, int32_t n,
                                                   const float* __restrict__ R, float rmin, float wscale,
                                                   float* __restrict__ T, float* __restrict__ M) {
  const int64_t s = (int64_t)blockIdx.x * 256 + threadIdx.x;
  if (s >= n) return;
  const LbCell c = cells[s];
  const float r = R[s];
  const bool big = r >= rmin;
  const float ri = big ? 1.f / r : 0.f;
  const float wt = big ? r * r * wscale : 0.f;
  const int64_t a = c.i + c.j * thin.sy + (int64_t)c.k * thin.sz;
  const int64_t b = (c.i + off) + (c.j + off) * thick.sy + (int64_t)(c.k + off) * thick.sz;
  for (int f = 0; f < LB_OUT; ++f) {
    const bool live = big && lb_comp_live(c, f);
    const float d = thick.F[f][b] - thin.F[f][a];
    T[s * LB_OUT + f] = live ? d * ri * (f >= 3 ? thin.zinv : 1.f) : 0.f;
    M[s * LB_OUT + f] = live ? wt : 0.f;
  }
}


// The dissipative part of a correction: the stepped value v may move toward
// zero by at most its own magnitude and never grow, so the layer acts as a
// cell- and wave-dependent sponge (stable wherever plain damping is).
__device__ __forceinline__ float lb_shrink(float v, float d) {
  return v > 0.f ? v + fminf(fmaxf(d, -v), 0.f) : (v < 0.f ? v + fmaxf(fminf(d, -v), 0.f) : v);
}

// ---------------------------------------------------------------------------
// inference on the matrix cores
// ---------------------------------------------------------------------------
// k_lb_infer evaluates the networks of every layer cell of every wavefield
// in one launch, before the step.  A wavefront takes 16 cells of one face
// at a time and forms the transposed products
//   H1^T = W1^T X^T,   H2^T = W2^T H1^T,   Y^T = W3^T H2^T
// with v_mfma_f32_16x16x32_bf16, the weights as the A operand (rows: hidden
// units or outputs) and the cells as the B operand (columns).  Lane
// (j, g) = (lane % 16, lane / 16) holds eight k values of cell j in each
// B operand and receives rows 4g..4g+3 of each 16-row result tile, i.e.
// hidden units 16t + 4g + e of cell j: these are the next layer's B operand
// once the next weights are packed in the matching k order, so no
// activation leaves the registers.  Element m of lane group g may stand for
// any k as long as A and B use the same assignment: the instruction pairs
// A and B elements of equal (g, m).
//
// Layer-1 k order (lb_slot1): lane group g gathers stencil points g and
// g + 4 with all nine fields, and a share of point 8 (fields 0..2 for g = 0,
// fields 1+2g and 2+2g otherwise): 21 of its 24 slots, the other 3 zero.
// Layers 2 and 3 (lb_k2): slot n of group g is hidden unit
// 16 (n / 4) + 4 g + n % 4, which is where layer 1 (2) left it.
typedef __bf16 bf8 __attribute__((ext_vector_type(8)));
typedef float f4 __attribute__((ext_vector_type(4)));

__host__ __device__ constexpr int lb_slot1(int g, int n) {
  if (n < 9) return 9 * g + n;
  if (n < 18) return 9 * (g + 4) + (n - 9);
  const int r = n - 18, cnt = g == 0 ? 3 : 2, f0 = g == 0 ? 0 : 1 + 2 * g;
  return r < cnt ? 72 + f0 + r : -1;
}
__host__ __device__ constexpr int lb_k2(int g, int n) { return 16 * (n >> 2) + 4 * g + (n & 3); }

// every input lands in exactly one layer-1 slot and every hidden unit in
// exactly one layer-2/3 slot
constexpr bool lb_slots_are_bijections() {
  int seen1[LB_IN] = {}, seen2[LB_H] = {};
  for (int g = 0; g < 4; ++g)
    for (int n = 0; n < 24; ++n) {
      const int e = lb_slot1(g, n);
      if (e >= LB_IN) return false;
      if (e >= 0) ++seen1[e];
    }
  for (int g = 0; g < 4; ++g)
    for (int n = 0; n < 16; ++n) ++seen2[lb_k2(g, n)];
  for (int e = 0; e < LB_IN; ++e)
    if (seen1[e] != 1) return false;
  for (int h = 0; h < LB_H; ++h)
    if (seen2[h] != 1) return false;
  return true;
}
static_assert(lb_slots_are_bijections(), "inference operand order must cover each input and hidden unit once");

__global__ __launch_bounds__(256) void k_lb_pack(const float* __restrict__ P, int32_t np, int32_t L,
                                                uint8_t* __restrict__ pk) {
  constexpr int NW1 = 4 * 3 * 64 * 8, NW2 = 4 * 2 * 64 * 8, NW3 = 2 * 64 * 8;
  constexpr int NB = LB_MAXL * LB_H + LB_H + 16;
  const float* __restrict__ p = P + (int64_t)blockIdx.y * np;
  uint8_t* o = pk + (int64_t)blockIdx.y * LB_PK_BYTES;
  const int q = blockIdx.x * 256 + threadIdx.x;
  if (q < NW1 + NW2 + NW3) {
    // element m of lane (i, g), tile t, k-step s
    const int m = q & 7, lane = (q >> 3) & 63, i = lane & 15, g = lane >> 4;
    float v;
    if (q < NW1) {
      const int ts = q >> 9, t = ts / 3, s = ts - 3 * t;
      const int e = lb_slot1(g, 8 * s + m);
      v = e >= 0 ? p[lb_off_w1() + e * LB_H + 16 * t + i] : 0.f;
    } else if (q < NW1 + NW2) {
      const int ts = (q - NW1) >> 9, t = ts >> 1, s = ts & 1;
      v = p[lb_off_w2(L) + lb_k2(g, 8 * s + m) * LB_H + 16 * t + i];
    } else {
      const int s = (q - NW1 - NW2) >> 9;
      v = i < LB_OUT ? p[lb_off_w3(L) + lb_k2(g, 8 * s + m) * LB_OUT + i] : 0.f;
    }
    reinterpret_cast<uint16_t*>(o)[q] = to_bf16(v);
  } else if (q < NW1 + NW2 + NW3 + NB) {
    const int r = q - (NW1 + NW2 + NW3);
    float v;
    if (r < LB_MAXL * LB_H) v = r < L * LB_H ? p[lb_off_b1() + r] : 0.f;
    else if (r < LB_MAXL * LB_H + LB_H) v = p[lb_off_b2(L) + r - LB_MAXL * LB_H];
    else v = r - LB_MAXL * LB_H - LB_H < LB_OUT ? p[lb_off_b3(L) + r - LB_MAXL * LB_H - LB_H] : 0.f;
    reinterpret_cast<float*>(o + LB_PK_B1)[r] = v;
  }
}

// a * sigmoid(a) from the hardware exp2 and reciprocal (about 2 ulp before
// the result is rounded to bfloat16)
__device__ __forceinline__ float silu_hw(float a) {
  return a * __builtin_amdgcn_rcpf(1.f + __builtin_amdgcn_exp2f(-1.44269504f * a));
}

__global__ __launch_bounds__(256) void k_lb_infer(const LbInfer a) {
  __shared__ __align__(16) uint8_t wsh[LB_PK_BYTES];
  const int blk = blockIdx.x;
  int f = 0;
#pragma unroll
  for (int q = 1; q < 6; ++q) f += blk >= a.blk_off[q] ? 1 : 0;
  {
    const uint4* __restrict__ src = reinterpret_cast<const uint4*>(a.pk + (int64_t)f * LB_PK_BYTES);
    for (int q = threadIdx.x; q < LB_PK_BYTES / 16; q += 256) reinterpret_cast<uint4*>(wsh)[q] = src[q];
  }
  __syncthreads();
  // (w through readfirstlane: the compiler then knows it, and every loop
  // bound and branch derived from it, to be uniform in the wavefront)
  const int lane = threadIdx.x & 63, w = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  const int jl = lane & 15, g = lane >> 4;
  const int c0 = a.face_off[f], c1 = a.face_off[f + 1], nc = a.face_off[6];
  const int ngrp = (c1 - c0 + 15) >> 4;
  const int nw = (a.blk_off[f + 1] - a.blk_off[f]) * 4;
  // stencil point g of this lane group: the cell, -x, +x, -y; point g + 4:
  // +y, -z, +z, two cells in along the normal; point 8: two cells out.  A
  // workgroup's cells lie on one face, so every axis is a lane constant:
  // the coordinate it moves (0 i, 1 j, 2 k), the grid's extent along it and
  // the offset of the step.
  const int nax = f >> 1;
  const int ax0 = g == 3 ? 1 : 0, d0 = g == 0 ? 0 : (g == 2 ? 1 : -1);
  const int ax1 = g == 0 ? 1 : (g == 3 ? nax : 2), d1 = g == 0 ? 1 : (g == 1 ? -1 : (g == 2 ? 1 : -2));
  auto extent = [&](int ax) { return ax == 0 ? a.nx : (ax == 1 ? a.ny : a.nz); };
  auto pitch = [&](int ax) { return ax == 0 ? 1 : (ax == 1 ? a.sy : a.sz); };
  const int n0 = extent(ax0), n1 = extent(ax1), n2 = extent(nax);
  const uint32_t dd0 = (uint32_t)(d0 * pitch(ax0)), dd1 = (uint32_t)(d1 * pitch(ax1)), dd2 = (uint32_t)(2 * pitch(nax));
  const uint32_t fs4 = a.fstride * 4u;
  const uint32_t fo8 = (uint32_t)(g == 0 ? 0 : 1 + 2 * g) * fs4;  // first field of the point-8 share
  const bool vel8 = g == 0;                                      // that share: velocities or stresses
  const float zinv = a.zinv, z = 1.f / a.zinv;
  const bf8* __restrict__ w1 = reinterpret_cast<const bf8*>(wsh + LB_PK_W1) + lane;
  const bf8* __restrict__ w2 = reinterpret_cast<const bf8*>(wsh + LB_PK_W2) + lane;
  const bf8* __restrict__ w3 = reinterpret_cast<const bf8*>(wsh + LB_PK_W3) + lane;
  const f4* __restrict__ b2p = reinterpret_cast<const f4*>(wsh + LB_PK_B2) + g;
  const f4 b3v = reinterpret_cast<const f4*>(wsh + LB_PK_B3)[g];

  const int64_t drow = (int64_t)nc;

  // A wavefront's items are (16-cell group, wavefield) pairs, wavefields
  // innermost.  The gathers of the next item are issued before the current
  // item's network is evaluated, and the cell records of the next group one
  // group ahead, so the loads' latency hides behind the arithmetic.
  struct Geo {
    int cell, depth, live;  // live: outputs 6, 7, 8 kept (bits 0..2)
    bool valid, k0, k1, k2;
    uint32_t o0, o1, o2;    // byte offsets of the lane's stencil points
  };
  // (lanes past the face's end read its last record and are masked by
  // Geo::valid: an unconditional load is not waited for until it is used)
  auto cell_of = [&](int grp) { return a.cells[min(c0 + grp * 16 + jl, c1 - 1)]; };
  auto geo = [&](int grp, const LbCell& c) {
    Geo q;
    q.cell = c0 + grp * 16 + jl;
    q.valid = q.cell < c1;
    q.depth = c.fd & 0xff;
    q.live = (c.i > 0 && c.j > 0) | ((c.i > 0 && c.k > 0) << 1) | ((c.j > 0 && c.k > 0) << 2);
    const uint32_t base = (uint32_t)c.i + (uint32_t)c.j * (uint32_t)a.sy + (uint32_t)c.k * (uint32_t)a.sz;
    auto coord = [&](int ax) { return ax == 0 ? c.i : (ax == 1 ? c.j : c.k); };
    // byte offset of a stencil point from the wavefield's base, 0 off the
    // grid (the loads then read element 0 and are discarded)
    auto point = [&](int cc, int d, int n, uint32_t dd, bool& ok) -> uint32_t {
      ok = q.valid && (unsigned)(cc + d) < (unsigned)n;
      return ((base + dd) * 4u) & (0u - (uint32_t)ok);
    };
    q.o0 = point(coord(ax0), d0, n0, dd0, q.k0);
    q.o1 = point(coord(ax1), d1, n1, dd1, q.k1);
    q.o2 = point(coord(nax), 2, n2, dd2, q.k2);
    return q;
  };
  // buffer loads: the wavefield's base in scalar registers, the field's
  // offset as the scalar offset, the stencil point's as the vector one (the
  // nine fields of a wavefield span less than 4 GiB, checked on the host),
  // so no per-load address arithmetic
  auto gather = [&](const Geo& q, int b, float* x) {
    const __amdgpu_buffer_rsrc_t fb =
        __builtin_amdgcn_make_buffer_rsrc(a.F + (int64_t)b * a.wstride, (short)0, (int)(9u * fs4), 0x00020000);
    auto ld = [&](uint32_t voff, uint32_t soff) {
      return __uint_as_float(__builtin_amdgcn_raw_buffer_load_b32(fb, voff, soff, 0));
    };
#pragma unroll
    for (int m = 0; m < 9; ++m) {
      x[m] = ld(q.o0, m * fs4);
      x[9 + m] = ld(q.o1, m * fs4);
    }
    x[18] = ld(q.o2 + fo8, 0);
    x[19] = ld(q.o2 + fo8, fs4);
    x[20] = ld(q.o2, 2 * fs4);
  };
  auto infer = [&](const Geo& q, int b, float* x) {
    // the weights are read from LDS in every item: an offset the compiler
    // cannot see through keeps it from hoisting all 22 operand tiles (88
    // registers) out of the loop, which would halve the occupancy
    int zo;
    asm volatile("s_mov_b32 %0, 0" : "=s"(zo));
    const bf8* __restrict__ w1i = w1 + zo;
    const bf8* __restrict__ w2i = w2 + zo;
    const bf8* __restrict__ w3i = w3 + zo;
    const f4* __restrict__ b2i = b2p + zo;
#pragma unroll
    for (int m = 0; m < 9; ++m) {
      x[m] = q.k0 ? x[m] : 0.f;
      x[9 + m] = q.k1 ? x[9 + m] : 0.f;
    }
    x[18] = q.k2 ? x[18] : 0.f;
    x[19] = q.k2 ? x[19] : 0.f;
    x[20] = q.k2 && vel8 ? x[20] : 0.f;  // the third point-8 slot: group 0 only
    // sum of squares with stresses divided by the impedance
    float sv = 0.f, ss = 0.f;
#pragma unroll
    for (int m = 0; m < 9; ++m) {
      if (m < 3) {
        sv = fmaf(x[m], x[m], sv);
        sv = fmaf(x[9 + m], x[9 + m], sv);
      } else {
        ss = fmaf(x[m], x[m], ss);
        ss = fmaf(x[9 + m], x[9 + m], ss);
      }
    }
    float s8 = x[18] * x[18];
    s8 = fmaf(x[19], x[19], s8);
    s8 = fmaf(x[20], x[20], s8);
    float tot = fmaf(zinv * zinv, vel8 ? ss : ss + s8, vel8 ? sv + s8 : sv);
    tot += __shfl_xor(tot, 16);
    tot += __shfl_xor(tot, 32);
    const float r = sqrtf(tot * (1.f / LB_IN)) + 1e-30f;
    const float rv = __builtin_amdgcn_rcpf(r), rs = rv * zinv, r8 = vel8 ? rv : rs;
    bf8 xb[3];
#pragma unroll
    for (int n = 0; n < 24; ++n) {
      float v;
      if (n < 18) v = x[n] * ((n % 9) < 3 ? rv : rs);
      else if (n < 21) v = x[n] * r8;
      else v = 0.f;
      xb[n >> 3][n & 7] = (__bf16)v;
    }
    // layer 1: bias of the cell's depth as the accumulator's start
    f4 acc[4];
    const f4* __restrict__ b1p = reinterpret_cast<const f4*>(wsh + LB_PK_B1) + q.depth * 16 + g;
#pragma unroll
    for (int t = 0; t < 4; ++t) {
      acc[t] = b1p[4 * t];
#pragma unroll
      for (int s = 0; s < 3; ++s)
        acc[t] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(w1i[(t * 3 + s) * 64], xb[s], acc[t], 0, 0, 0);
    }
    bf8 h[2];
#pragma unroll
    for (int t = 0; t < 4; ++t)
#pragma unroll
      for (int e = 0; e < 4; ++e) h[t >> 1][(t & 1) * 4 + e] = (__bf16)silu_hw(acc[t][e]);
    // layer 2
#pragma unroll
    for (int t = 0; t < 4; ++t) {
      acc[t] = b2i[4 * t];
#pragma unroll
      for (int s = 0; s < 2; ++s)
        acc[t] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(w2i[(t * 2 + s) * 64], h[s], acc[t], 0, 0, 0);
    }
#pragma unroll
    for (int t = 0; t < 4; ++t)
#pragma unroll
      for (int e = 0; e < 4; ++e) h[t >> 1][(t & 1) * 4 + e] = (__bf16)silu_hw(acc[t][e]);
    // layer 3: outputs 4g .. 4g+3 of cell j
    f4 y = b3v;
#pragma unroll
    for (int s = 0; s < 2; ++s) y = __builtin_amdgcn_mfma_f32_16x16x32_bf16(w3i[s * 64], h[s], y, 0, 0, 0);
    // every lane stores (no branch, so the compiler can count the stores
    // and wait for the next gathers only): lanes without an output, past
    // the face's end or at outputs 9..15, write to the trash slot
    float* __restrict__ Db = a.D + (int64_t)b * LB_OUT * drow + q.cell;
#pragma unroll
    for (int e = 0; e < 4; ++e) {
      const int o = 4 * g + e;
      const bool keep = q.valid && o < LB_OUT;
      const bool live = o < 6 || ((q.live >> (min(o, 8) - 6)) & 1);
      *(keep ? Db + (int64_t)o * drow : a.trash + lane) = live ? y[e] * r * (o >= 3 ? z : 1.f) : 0.f;
    }
  };

  int grp = (blk - a.blk_off[f]) * 4 + w;
  if (grp >= ngrp) return;
  Geo cur = geo(grp, cell_of(grp));
  float x[21];
  gather(cur, 0, x);
  for (;;) {
    const int gn = grp + nw;
    const bool more = gn < ngrp;
    // the next group's cell records, in flight during this group
    const LbCell cn = cell_of(more ? gn : grp);
    Geo nxt = cur;
    for (int b = 0; b < a.nbatch; ++b) {
      float xn[21];
      if (b + 1 < a.nbatch) {
        gather(cur, b + 1, xn);
      } else if (more) {
        // opaque to the compiler, so the records are not used (and waited
        // for) right where they are loaded, in the loop's header
        LbCell cq = cn;
        asm volatile("" : "+v"(cq.i), "+v"(cq.j), "+v"(cq.k), "+v"(cq.fd));
        nxt = geo(gn, cq);
        gather(nxt, 0, xn);
      }
      infer(cur, b, x);
#pragma unroll
      for (int m = 0; m < 21; ++m) x[m] = xn[m];
    }
    if (!more) break;
    cur = nxt;
    grp = gn;
  }
}

// After the step: fields <- shrink(fields, D), or fields + D in the free
// form, every wavefield (blockIdx.y).
__global__ __launch_bounds__(256) void k_lb_add(const LbInfer a) {
  const int nc = a.face_off[6];
  const int s = blockIdx.x * 256 + threadIdx.x;
  if (s >= nc) return;
  const LbCell c = a.cells[s];
  float* F = a.F + (int64_t)blockIdx.y * a.wstride + ((uint32_t)c.i + (uint32_t)c.j * (uint32_t)a.sy +
                                                      (uint32_t)c.k * (uint32_t)a.sz);
  const float* D = a.D + (int64_t)blockIdx.y * LB_OUT * nc + s;
  // all loads first: the nine fields are distinct arrays, but the compiler
  // cannot know it and would otherwise wait for each load after each store
  float v[LB_OUT], d[LB_OUT];
#pragma unroll
  for (int q = 0; q < LB_OUT; ++q) {
    v[q] = F[(int64_t)q * a.fstride];
    d[q] = D[(int64_t)q * nc];
  }
#pragma unroll
  for (int q = 0; q < LB_OUT; ++q) F[(int64_t)q * a.fstride] = a.free_form ? v[q] + d[q] : lb_shrink(v[q], d[q]);
}

// Training rollout: the corrections predicted before the update, from Y
// [n][9] and the stencil RMS R, applied as k_lb_infer + k_lb_add do.
__global__ __launch_bounds__(256) void k_lb_apply_y(const LbGeom g, const LbCell* __restrict__ cells, int32_t n,
                                                   const float* __restrict__ Y, const float* __restrict__ R,
                                                   int32_t free_form) {
  const int64_t s = (int64_t)blockIdx.x * 256 + threadIdx.x;
  if (s >= n) return;
  const LbCell c = cells[s];
  const int64_t a = c.i + c.j * g.sy + (int64_t)c.k * g.sz;
  const float r = R[s], z = 1.f / g.zinv;
  for (int f = 0; f < LB_OUT; ++f)
    if (lb_comp_live(c, f)) {
      float* p = const_cast<float*>(g.F[f]) + a;
      const float d = Y[s * LB_OUT + f] * r * (f >= 3 ? z : 1.f);
      *p = free_form ? *p + d : lb_shrink(*p, d);
    }
}

// Copy the thin window [off, off+n] of a thick field set (every array,
// including the staggered +1 planes) into a thin field set.
__global__ __launch_bounds__(256) void k_lb_crop(const LbGeom thick, const LbGeom thin, int32_t off) {
  const int i = blockIdx.x * 64 + (threadIdx.x & 63);
  const int j = blockIdx.y * 4 + (threadIdx.x >> 6);
  const int k = blockIdx.z;
  if (i > thin.nx || j > thin.ny || k > thin.nz) return;
  const int64_t a = i + j * thin.sy + (int64_t)k * thin.sz;
  const int64_t b = (i + off) + (j + off) * thick.sy + (int64_t)(k + off) * thick.sz;
  for (int f = 0; f < 9; ++f) const_cast<float*>(thin.F[f])[a] = thick.F[f][b];
}

// Memory variables of the convolutional PML, thin window of the thick run
// -> thin run: thin slab element e (all three slabs, six variables each)
// names cell (i,j,k); the thick run holds the same variable at cell
// (i+off, j+off, k+off) of its own, wider slab.
__device__ __forceinline__ int64_t psi_slab_index(const PsiGeom& g, int d, int i, int j, int k) {
  const int L2 = 2 * g.L;
  if (d == 0) {
    const int xi = i < g.L ? i : i - g.nx + L2;
    return xi + (int64_t)L2 * (j + (int64_t)g.ny * k);
  }
  if (d == 1) {
    const int yj = j < g.L ? j : j - g.ny + L2;
    return i + (int64_t)g.nx * (yj + (int64_t)L2 * k);
  }
  const int zk = k < g.L ? k : k - g.nz + L2;
  return i + (int64_t)g.nx * (j + (int64_t)g.ny * zk);
}

__global__ __launch_bounds__(256) void k_lb_crop_psi(const PsiGeom thick, const PsiGeom thin, int32_t off) {
  const int64_t e = (int64_t)blockIdx.x * 256 + threadIdx.x;
  const int64_t n0 = 6 * thin.pvar[0], n1 = 6 * thin.pvar[1];
  if (e >= n0 + n1 + 6 * thin.pvar[2]) return;
  const int L2 = 2 * thin.L;
  int d, i, j, k;
  int64_t r, v;
  if (e < n0) {
    d = 0; v = e / thin.pvar[0]; r = e - v * thin.pvar[0];
    const int xi = (int)(r % L2);
    const int64_t t = r / L2;
    j = (int)(t % thin.ny); k = (int)(t / thin.ny);
    i = xi < thin.L ? xi : xi - L2 + thin.nx;
  } else if (e < n0 + n1) {
    d = 1; v = (e - n0) / thin.pvar[1]; r = e - n0 - v * thin.pvar[1];
    i = (int)(r % thin.nx);
    const int64_t t = r / thin.nx;
    const int yj = (int)(t % L2);
    k = (int)(t / L2);
    j = yj < thin.L ? yj : yj - L2 + thin.ny;
  } else {
    d = 2; v = (e - n0 - n1) / thin.pvar[2]; r = e - n0 - n1 - v * thin.pvar[2];
    i = (int)(r % thin.nx);
    const int64_t t = r / thin.nx;
    j = (int)(t % thin.ny);
    const int zk = (int)(t / thin.ny);
    k = zk < thin.L ? zk : zk - L2 + thin.nz;
  }
  int64_t tb = 0;
  for (int q = 0; q < d; ++q) tb += 6 * thick.pvar[q];
  thin.psi[e] = thick.psi[tb + v * thick.pvar[d] + psi_slab_index(thick, d, i + off, j + off, k + off)];
}

// ---------------------------------------------------------------------------
// launchers
// ---------------------------------------------------------------------------
static unsigned lb_blocks(int64_t n, int per) { return (unsigned)std::max<int64_t>(1, (n + per - 1) / per); }

hipError_t lb_launch_mlp(const float* P, int L, int n, const float* X, const int32_t* depth, const float* T,
                         const float* M, float inv_count, float* Y, float* H1, float* H2, float* D1, float* D2,
                         float* G, float* loss_part, int nblk, hipStream_t s) {
  if (n <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_lb_mlp, dim3(nblk), dim3(256), 0, s, P, L, n, X, depth, T, M, inv_count, Y, H1, H2, D1,
                     D2, G, loss_part);
  return hipGetLastError();
}

hipError_t lb_launch_atb(const float* A, int lda, const int32_t* onehot, int P, const float* B, int ldb, int Q,
                         int n, int chunk, float* part, float* out, hipStream_t s) {
  const int nchunk = std::max(1, (n + chunk - 1) / chunk);
  if (Q > 64) return hipErrorInvalidValue;
  hipLaunchKernelGGL(k_lb_atb, dim3(nchunk, (P + 3) / 4), dim3(64, 4), 0, s, A, lda, onehot, P, B, ldb, Q, n,
                     chunk, part);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return e;
  hipLaunchKernelGGL(k_lb_reduce, dim3(lb_blocks((int64_t)P * Q, 256)), dim3(256), 0, s, part, nchunk, P * Q, out);
  return hipGetLastError();
}

hipError_t lb_launch_adam(float* P, float* m, float* v, const float* g, int n, float lr, float b1, float b2,
                          float eps, float c1, float c2, hipStream_t s) {
  hipLaunchKernelGGL(k_lb_adam, dim3(lb_blocks(n, 256)), dim3(256), 0, s, P, m, v, g, n, lr, b1, b2, eps, c1, c2);
  return hipGetLastError();
}

hipError_t lb_launch_gather(const LbGeom& g, const LbCell* cells, int n, float* X, float* R, int32_t* depth,
                            hipStream_t s) {
  if (n <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_lb_gather, dim3(lb_blocks(n, 256)), dim3(256), 0, s, g, cells, n, X, R, depth);
  return hipGetLastError();
}

hipError_t lb_launch_targets(const LbGeom& thin, const LbGeom& thick, int off, const LbCell* cells, int n,
                             const float* R, float rmin, float wscale, float* T, float* M, hipStream_t s) {
  if (n <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_lb_targets, dim3(lb_blocks(n, 256)), dim3(256), 0, s, thin, thick, off, cells, n, R, rmin,
                     wscale, T, M);
  return hipGetLastError();
}


hipError_t lb_launch_pack(const float* P, int np, int L, uint8_t* pk, hipStream_t s) {
  constexpr int n = 4 * 3 * 64 * 8 + 4 * 2 * 64 * 8 + 2 * 64 * 8 + LB_MAXL * LB_H + LB_H + 16;
  hipLaunchKernelGGL(k_lb_pack, dim3(lb_blocks(n, 256), 6), dim3(256), 0, s, P, np, L, pk);
  return hipGetLastError();
}

hipError_t lb_launch_infer(LbInfer& a, hipStream_t s) {
  // workgroups per face in proportion to its 16-cell groups, about 8 per CU
  // in all; every wavefront then takes the same share of groups
  int64_t ng[6], tot = 0;
  for (int f = 0; f < 6; ++f) {
    ng[f] = (a.face_off[f + 1] - a.face_off[f] + 15) / 16;
    tot += ng[f];
  }
  a.blk_off[0] = 0;
  if (tot == 0 || a.nbatch <= 0) return hipSuccess;
  const int64_t want = std::min<int64_t>(2048, (tot + 3) / 4);
  for (int f = 0; f < 6; ++f) {
    const int64_t nb = ng[f] ? std::max<int64_t>(1, (want * ng[f] + tot / 2) / tot) : 0;
    a.blk_off[f + 1] = a.blk_off[f] + (int32_t)nb;
  }
  hipLaunchKernelGGL(k_lb_infer, dim3(a.blk_off[6]), dim3(256), 0, s, a);
  return hipGetLastError();
}

hipError_t lb_launch_add(const LbInfer& a, hipStream_t s) {
  if (a.face_off[6] <= 0 || a.nbatch <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_lb_add, dim3(lb_blocks(a.face_off[6], 256), a.nbatch), dim3(256), 0, s, a);
  return hipGetLastError();
}

hipError_t lb_launch_apply_y(const LbGeom& g, const LbCell* cells, int n, const float* Y, const float* R,
                             int free_form, hipStream_t s) {
  if (n <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_lb_apply_y, dim3(lb_blocks(n, 256)), dim3(256), 0, s, g, cells, n, Y, R, free_form);
  return hipGetLastError();
}

hipError_t lb_launch_crop_psi(const PsiGeom& thick, const PsiGeom& thin, int off, hipStream_t s) {
  const int64_t n = 6 * (thin.pvar[0] + thin.pvar[1] + thin.pvar[2]);
  hipLaunchKernelGGL(k_lb_crop_psi, dim3(lb_blocks(n, 256)), dim3(256), 0, s, thick, thin, off);
  return hipGetLastError();
}

hipError_t lb_launch_crop(const LbGeom& thick, const LbGeom& thin, int off, hipStream_t s) {
  hipLaunchKernelGGL(k_lb_crop, dim3((thin.nx + 64) / 64, (thin.ny + 4) / 4, thin.nz + 1), dim3(256), 0, s, thick,
                     thin, off);
  return hipGetLastError();
}

}  // namespace fwe
