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
// One wavefront evaluates one sample: lane h owns hidden unit h.  Weights
// are staged in LDS as bfloat16 for the whole block.
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

// Shear components that the stress pass keeps at zero for this cell are
// masked (a correction there would be overwritten).
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

// Inference on the matrix cores, before the step: corrections of n cells of
// one face into D [n][9].  A wavefront evaluates 16 cells at a time with
// v_mfma_f32_16x16x16_bf16: layer 1 is X[16 x 96] W1[96 x 64] (inputs padded
// 81 -> 96), layer 2 H1[16 x 64] W2[64 x 64], layer 3 H2[16 x 64] W3[64 x 16]
// (outputs padded 9 -> 16).  Operand layout of that instruction: A[i][k] in
// lane i%16 + 16*(k/4), element k%4; B[k][n] in lane n + 16*(k/4); the
// result D[m][n] in lane n + 16*(m/4), element m%4.  Inputs and hidden
// activations are rounded to bfloat16, products accumulate in FP32.  The
// weights are rounded to bfloat16 once per block into LDS, transposed so
// that a lane's B operand (four consecutive k of one n) is one 8-byte LDS
// read; held in registers instead they took 88 VGPRs and left room for one
// wavefront per SIMD.  The activations change layout through a per-wave LDS
// tile.
typedef short bf4 __attribute__((ext_vector_type(4)));
typedef float f4 __attribute__((ext_vector_type(4)));

__device__ __forceinline__ void wave_sync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

__device__ __forceinline__ float silu(float a) { return a * sigm(a); }

constexpr int LB_K1 = 104;  // LDS row pitch (bf16) of W1^T: 96 padded inputs + 8
constexpr int LB_K2 = 72;   // of W2^T and W3^T: 64 + 8

__global__ __launch_bounds__(256) void k_lb_predict_mfma(const float* __restrict__ P, int32_t L, const LbGeom g,
                                                        const LbCell* __restrict__ cells, int32_t n,
                                                        float* __restrict__ D) {
  __shared__ uint16_t hs[LB_WAVES][16][LB_H + 4];
  __shared__ __align__(16) uint16_t sw1[LB_H][LB_K1];
  __shared__ __align__(16) uint16_t sw2[LB_H][LB_K2];
  __shared__ __align__(16) uint16_t sw3[16][LB_K2];
  const int w = threadIdx.x >> 6, l = threadIdx.x & 63;
  const int row = l & 15, grp = l >> 4;
  // weights rounded to bfloat16, transposed: sw1[n][k] = W1[k][n]
  for (int q = threadIdx.x; q < LB_H * 96; q += 256) {
    const int nn = q / 96, k = q - nn * 96;
    sw1[nn][k] = k < LB_IN ? to_bf16(P[lb_off_w1() + k * LB_H + nn]) : 0;
  }
  for (int q = threadIdx.x; q < LB_H * LB_H; q += 256) {
    const int nn = q / LB_H, k = q - nn * LB_H;
    sw2[nn][k] = to_bf16(P[lb_off_w2(L) + k * LB_H + nn]);
  }
  for (int q = threadIdx.x; q < 16 * LB_H; q += 256) {
    const int nn = q / LB_H, k = q - nn * LB_H;
    sw3[nn][k] = nn < LB_OUT ? to_bf16(P[lb_off_w3(L) + k * LB_OUT + nn]) : 0;
  }
  __syncthreads();
  // B operand of this lane: k = 16 s + 4 grp + (0..3), n = 16 t + row
  auto b1op = [&](int s, int t) { return *reinterpret_cast<const bf4*>(&sw1[t * 16 + row][s * 16 + 4 * grp]); };
  auto b2op = [&](int s, int t) { return *reinterpret_cast<const bf4*>(&sw2[t * 16 + row][s * 16 + 4 * grp]); };
  auto b3op = [&](int s) { return *reinterpret_cast<const bf4*>(&sw3[row][s * 16 + 4 * grp]); };
  float b2v[4];
#pragma unroll
  for (int t = 0; t < 4; ++t) b2v[t] = P[lb_off_b2(L) + t * 16 + row];
  const float b3v = row < LB_OUT ? P[lb_off_b3(L) + row] : 0.f;
  const float* __restrict__ b1 = P + lb_off_b1();
  const float z = 1.f / g.zinv;

  for (int64_t base = ((int64_t)blockIdx.x * LB_WAVES + w) * 16; base < n;
       base += (int64_t)gridDim.x * LB_WAVES * 16) {
    // the A row of this lane: cell base + row
    const int64_t sa = base + row;
    const bool va = sa < n;
    const LbCell c = va ? cells[sa] : LbCell{0, 0, 0, 0};
    float v[6][4];
#pragma unroll
    for (int s = 0; s < 6; ++s)
#pragma unroll
      for (int j = 0; j < 4; ++j) {
        const int e = s * 16 + 4 * grp + j;
        v[s][j] = lb_input(g, c, e, va && e < LB_IN);
      }
    // all 24 gathers in flight before the first use (the scheduler would
    // otherwise place each use right after its load and wait for each)
    __builtin_amdgcn_sched_barrier(0);
    float ss = 0.f;
#pragma unroll
    for (int s = 0; s < 6; ++s)
#pragma unroll
      for (int j = 0; j < 4; ++j) ss = fmaf(v[s][j], v[s][j], ss);
    ss += __shfl_xor(ss, 16);
    ss += __shfl_xor(ss, 32);
    const float r = sqrtf(ss * (1.f / LB_IN)) + 1e-30f;
    const float ri = 1.f / r;
    bf4 a[6];
#pragma unroll
    for (int s = 0; s < 6; ++s)
#pragma unroll
      for (int j = 0; j < 4; ++j) a[s][j] = (short)to_bf16(v[s][j] * ri);
    // per result row m = 4*grp + j: depth, RMS and shear flags of cell base + m
    const int flags = (c.fd & 0xff) | ((c.i > 0) << 8) | ((c.j > 0) << 9) | ((c.k > 0) << 10);
    int fm[4];
    float rm[4];
#pragma unroll
    for (int j = 0; j < 4; ++j) {
      fm[j] = __shfl(flags, 4 * grp + j);
      rm[j] = __shfl(r, 4 * grp + j);
    }
    // layer 1
    f4 acc[4];
#pragma unroll
    for (int t = 0; t < 4; ++t) {
      acc[t] = f4{0.f, 0.f, 0.f, 0.f};
#pragma unroll
      for (int s = 0; s < 6; ++s) acc[t] = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a[s], b1op(s, t), acc[t], 0, 0, 0);
    }
#pragma unroll
    for (int t = 0; t < 4; ++t)
#pragma unroll
      for (int j = 0; j < 4; ++j)
        hs[w][4 * grp + j][t * 16 + row] = to_bf16(silu(acc[t][j] + b1[(fm[j] & 0xff) * LB_H + t * 16 + row]));
    wave_sync();
    bf4 a2[4];
#pragma unroll
    for (int s = 0; s < 4; ++s)
#pragma unroll
      for (int j = 0; j < 4; ++j) a2[s][j] = (short)hs[w][row][s * 16 + 4 * grp + j];
    wave_sync();
    // layer 2
#pragma unroll
    for (int t = 0; t < 4; ++t) {
      acc[t] = f4{0.f, 0.f, 0.f, 0.f};
#pragma unroll
      for (int s = 0; s < 4; ++s) acc[t] = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a2[s], b2op(s, t), acc[t], 0, 0, 0);
    }
#pragma unroll
    for (int t = 0; t < 4; ++t)
#pragma unroll
      for (int j = 0; j < 4; ++j) hs[w][4 * grp + j][t * 16 + row] = to_bf16(silu(acc[t][j] + b2v[t]));
    wave_sync();
#pragma unroll
    for (int s = 0; s < 4; ++s)
#pragma unroll
      for (int j = 0; j < 4; ++j) a2[s][j] = (short)hs[w][row][s * 16 + 4 * grp + j];
    wave_sync();
    // layer 3: lane holds y[m = 4*grp + j][o = row]
    f4 y = f4{0.f, 0.f, 0.f, 0.f};
#pragma unroll
    for (int s = 0; s < 4; ++s) y = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a2[s], b3op(s), y, 0, 0, 0);
    if (row < LB_OUT) {
#pragma unroll
      for (int j = 0; j < 4; ++j) {
        const int64_t sm = base + 4 * grp + j;
        if (sm < n) {
          const LbCell cm{(fm[j] >> 8) & 1, (fm[j] >> 9) & 1, (fm[j] >> 10) & 1, 0};
          D[sm * LB_OUT + row] = lb_comp_live(cm, row) ? (y[j] + b3v) * rm[j] * (row >= 3 ? z : 1.f) : 0.f;
        }
      }
    }
  }
}

// Inference, after the step: fields <- shrink(fields, D).
__global__ __launch_bounds__(256) void k_lb_add(const LbGeom g, const LbCell* __restrict__ cells, int32_t n,
                                               const float* __restrict__ D) {
  const int64_t s = (int64_t)blockIdx.x * 256 + threadIdx.x;
  if (s >= n) return;
  const LbCell c = cells[s];
  const int64_t a = c.i + c.j * g.sy + (int64_t)c.k * g.sz;
  // all loads first: the nine fields are distinct arrays, but the compiler
  // cannot know it and would otherwise wait for each load after each store
  float v[LB_OUT], d[LB_OUT];
#pragma unroll
  for (int f = 0; f < LB_OUT; ++f) {
    v[f] = g.F[f][a];
    d[f] = D[s * LB_OUT + f];
  }
#pragma unroll
  for (int f = 0; f < LB_OUT; ++f) const_cast<float*>(g.F[f])[a] = lb_shrink(v[f], d[f]);
}

// Training rollout: the corrections predicted before the update, from Y
// [n][9] and the stencil RMS R, applied as k_lb_predict_mfma + k_lb_add do.
__global__ __launch_bounds__(256) void k_lb_apply_y(const LbGeom g, const LbCell* __restrict__ cells, int32_t n,
                                                   const float* __restrict__ Y, const float* __restrict__ R) {
  const int64_t s = (int64_t)blockIdx.x * 256 + threadIdx.x;
  if (s >= n) return;
  const LbCell c = cells[s];
  const int64_t a = c.i + c.j * g.sy + (int64_t)c.k * g.sz;
  const float r = R[s], z = 1.f / g.zinv;
  for (int f = 0; f < LB_OUT; ++f)
    if (lb_comp_live(c, f)) {
      float* p = const_cast<float*>(g.F[f]) + a;
      *p = lb_shrink(*p, Y[s * LB_OUT + f] * r * (f >= 3 ? z : 1.f));
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


hipError_t lb_launch_predict_mfma(const float* P, int L, const LbGeom& g, const LbCell* cells, int n, float* D,
                                  hipStream_t s) {
  if (n <= 0) return hipSuccess;
  const int nblk = (int)std::min<int64_t>(2048, (n + 16 * LB_WAVES - 1) / (16 * LB_WAVES));
  hipLaunchKernelGGL(k_lb_predict_mfma, dim3(nblk), dim3(256), 0, s, P, L, g, cells, n, D);
  return hipGetLastError();
}

hipError_t lb_launch_add(const LbGeom& g, const LbCell* cells, int n, const float* D, hipStream_t s) {
  if (n <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_lb_add, dim3(lb_blocks(n, 256)), dim3(256), 0, s, g, cells, n, D);
  return hipGetLastError();
}

hipError_t lb_launch_apply_y(const LbGeom& g, const LbCell* cells, int n, const float* Y, const float* R,
                             hipStream_t s) {
  if (n <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_lb_apply_y, dim3(lb_blocks(n, 256)), dim3(256), 0, s, g, cells, n, Y, R);
  return hipGetLastError();
}

hipError_t lb_launch_crop(const LbGeom& thick, const LbGeom& thin, int off, hipStream_t s) {
  hipLaunchKernelGGL(k_lb_crop, dim3((thin.nx + 64) / 64, (thin.ny + 4) / 4, thin.nz + 1), dim3(256), 0, s, thick,
                     thin, off);
  return hipGetLastError();
}

}  // namespace fwe
