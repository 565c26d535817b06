// fwe_tiles.h -- gfx950 kernels of the staggered-grid velocity-stress FDTD
// step (stress + damping + BC1 + injection, velocity + damping + receivers)
// and their launch templates.  Included by the instantiation unit
// fwe_step.hip, which is compiled once per (law, damping form).
// Semantics follow the reference kernels
//   compute_StressPrf / update_Qxft   OK4.cu:285-388   (iso law)
//   compute_StressOrtho / update_VelocityFromStress / BC1 / source / record
//                                      ANISO.cu:228-389 (ortho law)
// Layout and schedule are MI355X-first:
//  * every 3-D array (fields, medium, 3-D weights) shares one padded layout,
//    x fastest, row pitch a multiple of 32 floats (128 B lines), so one index
//    serves all staggered arrays and a 64-lane wavefront reads 256 B
//    contiguous per array;
//  * a workgroup owns a TX x TY column tile (TX = 64 or 128: one or two
//    wavefronts per row) and marches a chunk of z planes,
//    carrying the z-neighbours it needs in registers;
//  * the order of workgroups over tiles is tuned: linear (the workgroups in
//    flight stream one stretch of every array), grouped per XCD
//    (neighbouring tiles share an L2) or plane split (both: every XCD owns
//    one eighth of each chunk of planes, and all eight walk the chunks
//    together);
//  * the reference's separate source/BC1/record launches are fused into the
//    two stencil kernels (per-tile injection lists, per-(tile, plane)
//    receiver lists);
//  * the ortho medium is read in one of three forms with the same values
//    (the law template argument): twelve arrays (MODE_ORTHO), a material
//    table indexed by one byte per cell (MODE_ORTHO_TAB), or the eight
//    arrays of an isotropic medium (MODE_ORTHO_ISO); the iso law's c11 and
//    c12 as two arrays (MODE_ISO) or through a table (MODE_ISO_TAB);
//  * a plane's loads are all issued before its arithmetic, one memory round
//    trip per plane.
// The stencil is HBM-bound (~0.8 flop/B), so there is no MFMA here.
#pragma once
#include <type_traits>
#include "fwe_internal.h"
#include "fwe_stencil.h"

namespace fwe {

// Workgroup w is dispatched to XCD (w mod 8).  xcd_remap gives each XCD a
// contiguous range of logical tiles so x/y neighbours share an L2; the
// linear order instead keeps the workgroups in flight on one contiguous
// stretch of every array, which HBM streams better.  Which wins depends on
// the tiling (one-plane chunks: linear, measured), so the order is a tuned
// parameter (StepArgs::xcd_order).
__device__ __forceinline__ int xcd_remap(int b, int nb) {
  const int q = nb >> 3, r = nb & 7;
  const int x = b & 7, w = b >> 3;
  return x * q + (x < r ? x : r) + w;
}

// Plane split: the P = ntx*nty tiles of one z chunk, padded to 8*s, are
// cut into eight contiguous runs of s tiles (x fastest, then y); XCD x
// takes run x of every chunk, chunk after chunk.  A tile's x and y
// neighbours (mostly) and its z neighbour (always) are then processed on
// the same XCD shortly before or after it, so its halo lines hit that L2,
// while the workgroups in flight still cover a few contiguous planes of
// every array.  Padding slots return -1.
__device__ __forceinline__ int plane_split(int b, int P, int s) {
  const int x = b & 7, q = b >> 3;
  const int kc = q / s;
  const int t = x * s + (q - kc * s);
  return t < P ? t + kc * P : -1;
}

// logical tile t of the launch -> tile of the grid (StepArgs::tiles, box)
__device__ __forceinline__ int box_tile(const StepArgs& a, int t) {
  if (a.tiles) return a.tiles[t];
  if (a.bnx == a.ntx && a.bny == a.nty) return t + (a.bk0 * a.nty) * a.ntx;  // whole planes of tiles
  const int bx = t % a.bnx;
  t /= a.bnx;
  const int by = t % a.bny;
  const int bk = t / a.bny;
  return (a.bx0 + bx) + (a.by0 + by + (a.bk0 + bk) * a.nty) * a.ntx;
}

__device__ __forceinline__ int tile_of(const StepArgs& a, int b, int nb) {
  if (a.xcd_order == 1) return box_tile(a, xcd_remap(b, nb));
  if (a.xcd_order == 2) return plane_split(b, a.ntx * a.nty, a.split);
  return box_tile(a, b);
}

// workgroups of one group of NB wavefields (plane split pads every chunk to
// 8*split tiles); a launch covers nbg such groups, one after the other
__host__ __device__ inline int launch_blocks(const StepArgs& a) {
  if (a.xcd_order == 2) return a.nkc * 8 * a.split;
  return a.tiles ? a.ntiles : a.bnx * a.bny * a.bnk;
}

// workgroup -> (wavefield group g, tile lb or -1 for a padding slot).  Each
// group's range is a multiple of 8 workgroups under plane split, so its
// workgroups keep their XCD assignment.
__device__ __forceinline__ void split_batch(const StepArgs& a, int& g, int& lb) {
  const int per = launch_blocks(a);
  g = a.nbg > 1 ? (int)blockIdx.x / per : 0;
  lb = tile_of(a, (int)blockIdx.x - g * per, per);
}

__device__ __forceinline__ float ld(const float* p) { return __builtin_nontemporal_load(p); }
__device__ __forceinline__ void st(float* p, float v) { __builtin_nontemporal_store(v, p); }

// Unconditional load of p[idx] that yields 0 when !use; lanes that do not
// need it read p[0] (always in bounds).  Keeping loads out of branches lets
// the compiler keep a whole plane's loads in flight instead of draining the
// memory counter at every control-flow join.  These are the neighbour loads
// (c-1, c-sy, c-sz), re-read by adjacent tiles, so they go through the
// caches; the streamed centre values use ld().
__device__ __forceinline__ float ldz(const float* p, bool use, int64_t idx) {
  const float v = p[use ? idx : 0];
  return use ? v : 0.f;
}

// Convolutional PML: derivative g through its memory variable, whose old
// value psi0 was loaded with the plane's other values and whose new value
// is stored through p: psi <- b psi + a g, returns g/kappa + psi.
__device__ __forceinline__ float cpml_v(float g, float psi0, float* p, float b, float a, float kinv) {
  const float ps = fmaf(b, psi0, a * g);
  *p = ps;
  return fmaf(g, kinv, ps);
}

// Coefficients of one axis index: b, a, 1/kappa at the integer node, then at
// the half node.  A lane's x index and a wavefront's row are fixed for its
// whole z-march, so those are loaded once, before the march.
struct CpmlCoef {
  float bi, ai, ki, bh, ah, kh;
};
__device__ __forceinline__ CpmlCoef cpml_coef(const float* t, int m, int q) {
  return CpmlCoef{t[q], t[m + q], t[2 * m + q], t[3 * m + q], t[4 * m + q], t[5 * m + q]};
}

// Slab membership and element offsets of cell (i,j,k) for the convolutional
// PML (see StepArgs::psi).
struct CpmlCell {
  bool x, y, z;
  int64_t ox, oy, oz;
};
__device__ __forceinline__ CpmlCell cpml_cell(const StepArgs& a, int i, int j, int k) {
  const int L = a.cL, L2 = 2 * a.cL;
  CpmlCell c;
  c.x = i < L || i >= a.nx - L;
  c.y = j < L || j >= a.ny - L;
  c.z = k < L || k >= a.nz - L;
  // slab indices clamped into the slab, so cells outside it name a valid
  // element too (their loads are issued with the slab lanes' and unused)
  const int xi = min(max(i < L ? i : i - a.nx + L2, 0), L2 - 1);
  const int yj = min(max(j < L ? j : j - a.ny + L2, 0), L2 - 1);
  const int zk = min(max(k < L ? k : k - a.nz + L2, 0), L2 - 1);
  c.ox = xi + (int64_t)L2 * (j + (int64_t)a.ny * k);
  c.oy = i + (int64_t)a.nx * (yj + (int64_t)L2 * k);
  c.oz = i + (int64_t)a.nx * (j + (int64_t)a.ny * zk);
  return c;
}

// ---------------------------------------------------------------------------
// Stress pass.  For cell (i,j,k) of the column tile:
//   normal:  S_ii <- damp * S_ii + sum_j C_ij * dt * dV_j/dx_j
//   shear:   S_xy <- damp * S_xy + C66 * dt * (dVy/dx + dVx/dy)   (i,j >= 1)
//            S_xz ... (i,k >= 1),  S_yz ... (j,k >= 1)
// iso law (OK4.cu:296-329): damp multiplies (old + increment); ortho law
// (ANISO.cu:306-352): damp multiplies the old value only.
// Injection and BC1 are applied in the reference order (see StepArgs).
// NB wavefields share the medium loads; for each of them the z-neighbours
// Vx(k-1), Vy(k-1) and Vz(k) are carried in registers from the previous
// plane, so every field value is read from memory once per plane.
// ---------------------------------------------------------------------------
template <int MODE, int PML, bool FS, int TX, int TY, int NB>
__device__ __forceinline__ void stress_tile(const StepArgs& a, int lb, const int wb) {
  // a wavefront holds one tile row (64 lanes of a 64- or 128-column row), so
  // the row is uniform per wavefront: saying so keeps row addressing in
  // scalar registers
  const int lane = threadIdx.x, row = __builtin_amdgcn_readfirstlane(threadIdx.y);
  const int tx = lb % a.ntx;
  lb /= a.ntx;
  const int tyi = lb % a.nty;
  const int kc = lb / a.nty;
  const int i = tx * TX + lane, j = tyi * TY + row;
  if (i >= a.nx || j >= a.ny) return;
  const int k0 = kc * a.kchunk;
  const int k1 = min(k0 + a.kchunk, a.nz);
  const int64_t sy = a.sy, sz = a.sz;
  const int tile = tx + tyi * a.ntx;
  const int e0 = a.inj_off[tile], e1 = a.inj_off[tile + 1];
  const bool im = i >= 1, jm = j >= 1;
  const float third = 1.0f / 3.0f;

  float fxh = 1.f, fyh = 1.f, wxi = 1.f, wyj = 1.f;
  if (PML == PML_PROFILE) { fxh = a.whx[i]; fyh = a.why[j]; }
  CpmlCoef qx{}, qy{};  // convolutional PML: this lane's x and this row's y coefficients
  bool xw = false;       // this wavefront holds x-slab cells
  if (PML == PML_CPML) {
    if (i < a.cL || i >= a.nx - a.cL) qx = cpml_coef(a.ctab[0], a.nx + 1, i);
    if (j < a.cL || j >= a.ny - a.cL) qy = cpml_coef(a.ctab[1], a.ny + 1, j);
    xw = __builtin_amdgcn_ballot_w64(i < a.cL || i >= a.nx - a.cL) != 0;
  }
  if (PML == PML_PROFILE || PML == PML_SEP) { wxi = a.wx[i]; wyj = a.wy[j]; }

  // carried z-neighbours, per wavefield: Vx(k-1), Vy(k-1), Vz(k)
  float cvx[NB], cvy[NB], cvz[NB];

  // The first plane of the chunk loads its z-neighbours with its other
  // loads (a prologue before the loop would cost the workgroup one more
  // memory round trip); later planes use the carries.
  auto plane = [&](const int k, auto first_tag) __attribute__((always_inline)) {
    constexpr bool FIRST = decltype(first_tag)::value;
    const int64_t c = (int64_t)i + j * sy + k * sz;
    const bool km = k >= 1;
    CpmlCell cc{};
    if (PML == PML_CPML) cc = cpml_cell(a, i, j, k);
    float wzk = 1.f, fzh = 1.f;
    if (PML == PML_PROFILE) { wzk = a.wz[k]; fzh = a.whz[k]; }
    if (PML == PML_SEP) wzk = a.wz[k];
    float w = 1.f;
    if (PML == PML_W3D) w = ld(a.w3 + c);
    if (PML == PML_SEP) w = wxi * wyj * wzk;
    // medium at this cell, shared by all wavefields of the batch
    float m0, m1, m2 = 0.f, m3 = 0.f, m4 = 0.f, m5 = 0.f, m6 = 0.f, m7 = 0.f, m8 = 0.f;
    if (MODE == MODE_ORTHO_TAB) {
      // one byte per cell selects the cell's twelve values in the table
      // (48 B per entry, 16-B aligned; a layered medium keeps the entries in
      // the caches)
      const float* t = a.mtab + 12u * (uint32_t)a.mid[c];
      const float4 q0 = *reinterpret_cast<const float4*>(t);
      const float4 q1 = *reinterpret_cast<const float4*>(t + 4);
      m0 = q0.x; m1 = q0.y; m2 = q0.z; m3 = q0.w;
      m4 = q1.x; m5 = q1.y; m6 = q1.z; m7 = q1.w;
      m8 = t[8];
    } else if (MODE == MODE_ISO_TAB) {
      const float2 q = *reinterpret_cast<const float2*>(a.mtab + 2u * (uint32_t)a.mid[c]);
      m0 = q.x; m1 = q.y;
    } else if (MODE == MODE_ORTHO_ISO) {
      // isotropic medium (as the station solver's c11/c12 arrays,
      // FastWaveED_GPU3D_v1_9kmst.cu:381-400): two normal moduli per cell
      m0 = ld(a.C[0] + c); m3 = ld(a.C[3] + c);
      m1 = m0; m2 = m0; m4 = m3; m5 = m3;
      m6 = ld(a.C[6] + c); m7 = ld(a.C[7] + c); m8 = ld(a.C[8] + c);
    } else {
      m0 = ld(a.C[0] + c); m1 = ld(a.C[1] + c);
      if (MODE == MODE_ORTHO) {
        m2 = ld(a.C[2] + c); m3 = ld(a.C[3] + c); m4 = ld(a.C[4] + c); m5 = ld(a.C[5] + c);
        m6 = ld(a.C[6] + c); m7 = ld(a.C[7] + c); m8 = ld(a.C[8] + c);
      }
    }
    // vacuum-layer top (diff_BC2_OK3/FastWaveED_GPU3D_v1.cu:309-314): the
    // shear moduli of plane vac_k are replaced (uniform per plane)
    const bool vac = k == a.vac_k;
    const float c66k = vac ? a.vac_cdt : a.c66dt;
    if (vac) { m6 = a.vac_cdt; m7 = a.vac_cdt; m8 = a.vac_cdt; }
#pragma unroll
    for (int b = 0; b < NB; ++b) {
      const int64_t fo = (int64_t)(wb + b) * a.fstride;
      const float* Vx = a.V[0] + fo;
      const float* Vy = a.V[1] + fo;
      const float* Vz = a.V[2] + fo;
      float* Sxx = a.S[0] + fo; float* Syy = a.S[1] + fo; float* Szz = a.S[2] + fo;
      float* Sxy = a.S[3] + fo; float* Sxz = a.S[4] + fo; float* Syz = a.S[5] + fo;
      // all loads of this cell first, branch-free
      const float vx = Vx[c], vy = Vy[c];
      const float vx_ip = Vx[c + 1], vy_jp = Vy[c + sy], vz_kp = Vz[c + sz];
      const float vz = FIRST ? Vz[c] : cvz[b];
      const float vx_km = FIRST ? ldz(Vx, k >= 1, c - sz) : cvx[b];
      const float vy_km = FIRST ? ldz(Vy, k >= 1, c - sz) : cvy[b];
      const float vx_jm = ldz(Vx, jm, c - sy);
      const float vy_im = ldz(Vy, im, c - 1);
      const float vz_im = ldz(Vz, im, c - 1);
      const float vz_jm = ldz(Vz, jm, c - sy);
      float sxx = ld(Sxx + c), syy = ld(Syy + c), szz = ld(Szz + c);
      const float oxy = ld(Sxy + c), oxz = ld(Sxz + c), oyz = ld(Syz + c);
      // memory variables of the slabs this wavefront touches, loaded with the
      // plane's other values (one memory round trip per plane)
      float px[3] = {0.f, 0.f, 0.f}, py[3] = {0.f, 0.f, 0.f}, pz[3] = {0.f, 0.f, 0.f};
      int64_t po = 0;
      if (PML == PML_CPML) {
        po = (int64_t)(wb + b) * a.pstride;
        if (xw)
          for (int v = 0; v < 3; ++v) px[v] = a.psi[0][po + cc.ox + v * a.pvar[0]];
        if (cc.y)
          for (int v = 0; v < 3; ++v) py[v] = a.psi[1][po + cc.oy + v * a.pvar[1]];
        if (cc.z)
          for (int v = 0; v < 3; ++v) pz[v] = a.psi[2][po + cc.oz + v * a.pvar[2]];
      }
      float sxy = oxy, sxz = oxz, syz = oyz;
      cvx[b] = vx; cvy[b] = vy; cvz[b] = vz_kp;
      float ex, ey, ez, dxy, dxz, dyz;
      if (PML == PML_CPML && (cc.x || cc.y || cc.z)) {
        // the nine strain-rate derivatives, each through its memory
        // variable where the cell lies in that axis' slab (cells in no slab
        // take the plain form below, bit for bit as the plain kernel)
        float gxx = (vx_ip - vx) * a.rdx, gyy = (vy_jp - vy) * a.rdy, gzz = (vz_kp - vz) * a.rdz;
        float gyx = (vy - vy_im) * a.rdx, gxy = (vx - vx_jm) * a.rdy;  // Sxy node
        float gzx = (vz - vz_im) * a.rdx, gxz = (vx - vx_km) * a.rdz;  // Sxz node
        float gyz = (vy - vy_km) * a.rdz, gzy = (vz - vz_jm) * a.rdy;  // Syz node
        {
          if (cc.x) {
            float* p = a.psi[0] + po + cc.ox;
            gxx = cpml_v(gxx, px[0], p, qx.bi, qx.ai, qx.ki);
            gyx = cpml_v(gyx, px[1], p + a.pvar[0], qx.bh, qx.ah, qx.kh);
            gzx = cpml_v(gzx, px[2], p + 2 * a.pvar[0], qx.bh, qx.ah, qx.kh);
          }
          if (cc.y) {
            float* p = a.psi[1] + po + cc.oy;
            gyy = cpml_v(gyy, py[0], p, qy.bi, qy.ai, qy.ki);
            gxy = cpml_v(gxy, py[1], p + a.pvar[1], qy.bh, qy.ah, qy.kh);
            gzy = cpml_v(gzy, py[2], p + 2 * a.pvar[1], qy.bh, qy.ah, qy.kh);
          }
          if (cc.z) {
            float* p = a.psi[2] + po + cc.oz;
            const float* t = a.ctab[2];
            const int m = a.nz + 1;
            gzz = cpml_v(gzz, pz[0], p, t[k], t[m + k], t[2 * m + k]);
            gxz = cpml_v(gxz, pz[1], p + a.pvar[2], t[3 * m + k], t[4 * m + k], t[5 * m + k]);
            gyz = cpml_v(gyz, pz[2], p + 2 * a.pvar[2], t[3 * m + k], t[4 * m + k], t[5 * m + k]);
          }
        }
        ex = gxx; ey = gyy; ez = gzz;
        dxy = gyx + gxy; dxz = gzx + gxz; dyz = gyz + gzy;
      } else {
        ex = (vx_ip - vx) * a.rdx; ey = (vy_jp - vy) * a.rdy; ez = (vz_kp - vz) * a.rdz;
        dxy = rate2(vy - vy_im, a.rdx, vx - vx_jm, a.rdy);
        dxz = rate2(vz - vz_im, a.rdx, vx - vx_km, a.rdz);
        dyz = rate2(vy - vy_km, a.rdz, vz - vz_jm, a.rdy);
      }
      if (iso_law(MODE)) {
        // OK4.cu:307-309,323-329 with c11,c12 arrays holding c*dt
        sxx = iso_normal(fxh, sxx, m0, ex, m1, ey, m1, ez);
        syy = iso_normal(fyh, syy, m1, ex, m0, ey, m1, ez);
        szz = iso_normal(fzh, szz, m1, ex, m1, ey, m0, ez);
        sxy = iso_shear(wzk * wyj * wxi, sxy, c66k, dxy);
        sxz = iso_shear(wyj * wzk * wxi, sxz, c66k, dxz);
        syz = iso_shear(wxi * wyj * wzk, syz, c66k, dyz);
      } else {
        // ANISO.cu:306-352 with Cij holding Cij*dt
        sxx = ortho_normal(w, sxx, m0, ex, m3, ey, m4, ez);
        syy = ortho_normal(w, syy, m3, ex, m1, ey, m5, ez);
        szz = ortho_normal(w, szz, m4, ex, m5, ey, m2, ez);
        sxy = ortho_shear(w, sxy, m8, dxy);
        sxz = ortho_shear(w, sxz, m7, dxz);
        syz = ortho_shear(w, syz, m6, dyz);
      }
      // shear cells outside the reference update ranges keep their value
      // (zero unless a point injection adds to them, as in the reference,
      // which never writes them otherwise)
      if (!(im && jm)) sxy = oxy;
      if (!(im && km)) sxz = oxz;
      if (!(jm && km)) syz = oyz;
      // point injections of this cell (OK4.cu:336-342 / ANISO.cu:228-252)
      float src[6] = {0.f, 0.f, 0.f, 0.f, 0.f, 0.f};
      float src_xz1 = 0.f, src_yz1 = 0.f;  // injections at (i,j,1), for BC1
      if (e1 > e0) {
        const int it = a.it + (a.it_base ? *a.it_base : 0);
        const float amp = (it < a.nstf) ? a.stf[(int64_t)(wb + b) * a.nstf + it] : 0.f;
        for (int e = e0; e < e1; ++e) {
          const PointEntry p = a.inj[e];
          if (p.i == i && p.j == j && p.batch == wb + b) {
            if (p.k == k) src[p.comp] = fmaf(p.coef, amp, src[p.comp]);
            if (FS && k == 2 && p.k == 1) {
              if (p.comp == 4) src_xz1 = fmaf(p.coef, amp, src_xz1);
              if (p.comp == 5) src_yz1 = fmaf(p.coef, amp, src_yz1);
            }
          }
        }
      }
      if (!a.src_after_bc) {
        sxx += src[0]; syy += src[1]; szz += src[2];
        sxy += src[3]; sxz += src[4]; syz += src[5];
      }
      if (FS) {
        // BC1 (OK4.cu:279-282): Szz(k=0)=0, Sxz/Syz(k=1) = Sxz/Syz(k=2)/3
        if (k == 0) szz = 0.f;
        if (k == 2) {
          if (im) Sxz[c - sz] = fmaf(third, sxz, a.src_after_bc ? src_xz1 : 0.f);
          if (jm) Syz[c - sz] = fmaf(third, syz, a.src_after_bc ? src_yz1 : 0.f);
        }
      }
      if (a.src_after_bc) {
        sxx += src[0]; syy += src[1]; szz += src[2];
        sxy += src[3]; sxz += src[4]; syz += src[5];
      }
      // streaming stores: the outputs are not read again in this pass
      st(Sxx + c, sxx); st(Syy + c, syy); st(Szz + c, szz);
      st(Sxy + c, sxy); st(Sxz + c, sxz); st(Syz + c, syz);
    }
  };
  plane(k0, std::true_type{});
  for (int k = k0 + 1; k < k1; ++k) plane(k, std::false_type{});
}

// ---------------------------------------------------------------------------
// Velocity pass (OK4.cu:351-378 iso, ANISO.cu:374-386 ortho) and receiver
// sampling after the update (OK4.cu:383-386).  Ranges:
//   Vx 1<=i<=nx-1, 1<=j<=ny-2, 1<=k<=nz-2
//   Vy 1<=i<=nx-2, 1<=j<=ny-1, 1<=k<=nz-2
//   Vz 1<=i<=nx-2, 1<=j<=ny-2, 1<=k<=nz-1
// Per wavefield, Szz(k-1), Sxz(k) and Syz(k) are carried from the previous
// plane.
// ---------------------------------------------------------------------------
template <int MODE, int PML, int TX, int TY, int NB>
__device__ __forceinline__ void velocity_tile(const StepArgs& a, int lb, const int wb) {
  // a wavefront holds one tile row (64 lanes of a 64- or 128-column row), so
  // the row is uniform per wavefront: saying so keeps row addressing in
  // scalar registers
  const int lane = threadIdx.x, row = __builtin_amdgcn_readfirstlane(threadIdx.y);
  const int tx = lb % a.ntx;
  lb /= a.ntx;
  const int tyi = lb % a.nty;
  const int kc = lb / a.nty;
  const int i = tx * TX + lane, j = tyi * TY + row;
  if (i >= a.nx || j >= a.ny) return;
  const int k0 = kc * a.kchunk;
  const int k1 = min(k0 + a.kchunk, a.nz);
  const int nx = a.nx, ny = a.ny, nz = a.nz;
  const int64_t sy = a.sy, sz = a.sz;
  const int tile = tx + tyi * a.ntx;
  const bool im = i >= 1, jm = j >= 1;
  const bool xin = im && jm;  // all three need i>=1, j>=1, k>=1
  const bool ux_ij = xin && j <= ny - 2;
  const bool uy_ij = xin && i <= nx - 2;
  const bool uz_ij = xin && i <= nx - 2 && j <= ny - 2;

  float wxm = 1.f, wym = 1.f, txi = 1.f, tyj = 1.f;
  CpmlCoef qx{}, qy{};  // convolutional PML: this lane's x and this row's y coefficients
  bool xw = false;       // this wavefront holds x-slab cells
  if (PML == PML_CPML) {
    if (i < a.cL || i >= nx - a.cL) qx = cpml_coef(a.ctab[0], nx + 1, i);
    if (j < a.cL || j >= ny - a.cL) qy = cpml_coef(a.ctab[1], ny + 1, j);
    xw = __builtin_amdgcn_ballot_w64(i < a.cL || i >= nx - a.cL) != 0;
  }
  if (PML == PML_PROFILE) {
    wxm = im ? a.wx[i - 1] : 1.f;
    wym = jm ? a.wy[j - 1] : 1.f;
  }
  if (PML == PML_SEP) { txi = a.wx[i]; tyj = a.wy[j]; }

  float cszz[NB], csxz[NB], csyz[NB];  // Szz(k-1), Sxz(k), Syz(k)

  // The first plane of the chunk reads its z-neighbours from memory in the
  // same batch as its other loads (a prologue load before the loop would
  // make the loop head wait for it, and with one-plane chunks that doubles
  // the memory round trips of the workgroup); later planes use the carries.
  auto plane = [&](const int k, auto first_tag) __attribute__((always_inline)) {
    constexpr bool FIRST = decltype(first_tag)::value;
    const int64_t c = (int64_t)i + j * sy + k * sz;
    CpmlCell cc{};
    if (PML == PML_CPML) cc = cpml_cell(a, i, j, k);
    const bool ux = ux_ij && k >= 1 && k <= nz - 2;
    const bool uy = uy_ij && k >= 1 && k <= nz - 2;
    const bool uz = uz_ij && k >= 1;
    float w = 1.f;
    if (PML == PML_W3D) w = ld(a.w3 + c);
    const float wzk = PML == PML_SEP ? a.wz[k] : 1.f;
    const int64_t tk = (int64_t)tile * nz + k;
    const int r0 = a.rec_koff[tk], r1 = a.rec_koff[tk + 1];
    float bx = a.bxdt, by = a.bydt, bz = a.bzdt;
    if (MODE == MODE_ORTHO || MODE == MODE_ORTHO_ISO) {
      // buoyancy through the caches (measured faster than streaming here)
      bx = a.B[0][c]; by = a.B[1][c]; bz = a.B[2][c];
    }
    // material-table index first: its entry is loaded after the field
    // loads of the first wavefield, so those are in flight meanwhile
    const uint32_t mi = MODE == MODE_ORTHO_TAB ? (uint32_t)a.mid[c] : 0u;
    float wzm = 1.f;
    if (PML == PML_PROFILE) wzm = k >= 1 ? a.wz[k - 1] : 1.f;
#pragma unroll
    for (int b = 0; b < NB; ++b) {
      const int64_t fo = (int64_t)(wb + b) * a.fstride;
      float* Vx = a.V[0] + fo; float* Vy = a.V[1] + fo; float* Vz = a.V[2] + fo;
      const float* Sxx = a.S[0] + fo; const float* Syy = a.S[1] + fo;
      const float* Szz = a.S[2] + fo; const float* Sxy = a.S[3] + fo;
      const float* Sxz = a.S[4] + fo; const float* Syz = a.S[5] + fo;
      // all loads of this cell first, branch-free (c+1, c+sy, c+sz stay
      // inside the padded extents; c-1, c-sy, c-sz only when valid)
      float vx = ld(Vx + c), vy = ld(Vy + c), vz = ld(Vz + c);
      const float sxx = Sxx[c], sxx_im = ldz(Sxx, im, c - 1);
      const float syy = Syy[c], syy_jm = ldz(Syy, jm, c - sy);
      const float szz = Szz[c];
      const float szz_km = FIRST ? ldz(Szz, k >= 1, c - sz) : cszz[b];
      const float sxy = Sxy[c], sxy_jp = Sxy[c + sy], sxy_ip = Sxy[c + 1];
      const float sxz = FIRST ? Sxz[c] : csxz[b];
      const float sxz_kp = Sxz[c + sz], sxz_ip = Sxz[c + 1];
      const float syz = FIRST ? Syz[c] : csyz[b];
      const float syz_kp = Syz[c + sz], syz_jp = Syz[c + sy];
      // memory variables of the slabs this wavefront touches, loaded with the
      // plane's other values (one memory round trip per plane)
      float px[3] = {0.f, 0.f, 0.f}, py[3] = {0.f, 0.f, 0.f}, pz[3] = {0.f, 0.f, 0.f};
      int64_t po = 0;
      if (PML == PML_CPML) {
        po = (int64_t)(wb + b) * a.pstride;
        if (xw)
          for (int v = 0; v < 3; ++v) px[v] = a.psi[0][po + cc.ox + (3 + v) * a.pvar[0]];
        if (cc.y)
          for (int v = 0; v < 3; ++v) py[v] = a.psi[1][po + cc.oy + (3 + v) * a.pvar[1]];
        if (cc.z)
          for (int v = 0; v < 3; ++v) pz[v] = a.psi[2][po + cc.oz + (3 + v) * a.pvar[2]];
      }
      cszz[b] = szz; csxz[b] = sxz_kp; csyz[b] = syz_kp;
      if (MODE == MODE_ORTHO_TAB && b == 0) {
        // entry words 8..11: C66xy, bx, by, bz
        const float4 q = *reinterpret_cast<const float4*>(a.mtab + 12u * mi + 8);
        bx = q.y; by = q.z; bz = q.w;
      }
      // one memory round trip per plane: nothing below is scheduled before
      // every load above is issued
      if (NB == 1) __builtin_amdgcn_sched_barrier(0);
      if (PML == PML_SEP && b == 0) w = txi * tyj * wzk;
      float dSx, dSy, dSz;
      if (PML == PML_CPML && (cc.x || cc.y || cc.z)) {
        // the nine stress derivatives, each through its memory variable
        // where the cell lies in that axis' slab (cells in no slab take the
        // plain form below, bit for bit as the plain kernel)
        float hxx = (sxx - sxx_im) * a.rdx, hxyy = (sxy_jp - sxy) * a.rdy, hxzz = (sxz_kp - sxz) * a.rdz;
        float hyy = (syy - syy_jm) * a.rdy, hxyx = (sxy_ip - sxy) * a.rdx, hyzz = (syz_kp - syz) * a.rdz;
        float hzz = (szz - szz_km) * a.rdz, hxzx = (sxz_ip - sxz) * a.rdx, hyzy = (syz_jp - syz) * a.rdy;
        {
          if (cc.x) {
            float* p = a.psi[0] + po + cc.ox + 3 * a.pvar[0];
            hxx = cpml_v(hxx, px[0], p, qx.bh, qx.ah, qx.kh);
            hxyx = cpml_v(hxyx, px[1], p + a.pvar[0], qx.bi, qx.ai, qx.ki);
            hxzx = cpml_v(hxzx, px[2], p + 2 * a.pvar[0], qx.bi, qx.ai, qx.ki);
          }
          if (cc.y) {
            float* p = a.psi[1] + po + cc.oy + 3 * a.pvar[1];
            hyy = cpml_v(hyy, py[0], p, qy.bh, qy.ah, qy.kh);
            hxyy = cpml_v(hxyy, py[1], p + a.pvar[1], qy.bi, qy.ai, qy.ki);
            hyzy = cpml_v(hyzy, py[2], p + 2 * a.pvar[1], qy.bi, qy.ai, qy.ki);
          }
          if (cc.z) {
            float* p = a.psi[2] + po + cc.oz + 3 * a.pvar[2];
            const float* t = a.ctab[2];
            const int m = nz + 1;
            hzz = cpml_v(hzz, pz[0], p, t[3 * m + k], t[4 * m + k], t[5 * m + k]);
            hxzz = cpml_v(hxzz, pz[1], p + a.pvar[2], t[k], t[m + k], t[2 * m + k]);
            hyzz = cpml_v(hyzz, pz[2], p + 2 * a.pvar[2], t[k], t[m + k], t[2 * m + k]);
          }
        }
        dSx = hxx + hxyy + hxzz;
        dSy = hyy + hxyx + hyzz;
        dSz = hzz + hxzx + hyzy;
      } else {
        dSx = rate3(sxx - sxx_im, a.rdx, sxy_jp - sxy, a.rdy, sxz_kp - sxz, a.rdz);
        dSy = rate3(syy - syy_jm, a.rdy, sxy_ip - sxy, a.rdx, syz_kp - syz, a.rdz);
        dSz = rate3(szz - szz_km, a.rdz, sxz_ip - sxz, a.rdx, syz_jp - syz, a.rdy);
      }
      if (ux) vx = iso_law(MODE) ? iso_vel(wxm, vx, bx, dSx) : ortho_vel(w, vx, bx, dSx);
      if (uy) vy = iso_law(MODE) ? iso_vel(wym, vy, by, dSy) : ortho_vel(w, vy, by, dSy);
      if (uz) vz = iso_law(MODE) ? iso_vel(wzm, vz, bz, dSz) : ortho_vel(w, vz, bz, dSz);
      // every cell is written (unchanged outside the update ranges) with
      // streaming stores, so whole lines leave the CU
      st(Vx + c, vx);
      st(Vy + c, vy);
      st(Vz + c, vz);
      for (int r = r0; r < r1; ++r) {  // this tile's receivers in plane k
        const PointEntry p = a.rec[r];
        const int it = a.it + (a.it_base ? *a.it_base : 0);
        if (p.i == i && p.j == j && p.k == k && p.batch == wb + b && it < a.nt)
          a.traces[(int64_t)p.slot * a.nt + it] = p.comp == 0 ? vx : (p.comp == 1 ? vy : vz);
      }
    }
  };
  plane(k0, std::true_type{});
  for (int k = k0 + 1; k < k1; ++k) plane(k, std::false_type{});
}

// One model per launch: workgroup -> tile of that model.
template <int MODE, int PML, bool FS, int TX, int TY, int NB>
__global__ __launch_bounds__(TX * TY) void k_stress(const StepArgs a) {
  int g, lb;
  split_batch(a, g, lb);
  if (lb >= 0) stress_tile<MODE, PML, FS, TX, TY, NB>(a, lb, g * NB);
}
template <int MODE, int PML, int TX, int TY, int NB>
__global__ __launch_bounds__(TX * TY) void k_velocity(const StepArgs a) {
  int g, lb;
  split_batch(a, g, lb);
  if (lb >= 0) velocity_tile<MODE, PML, TX, TY, NB>(a, lb, g * NB);
}

// Many models per launch (an uneven batch of independent grids): the
// workgroups of all models form one grid; wg[2*lb] is the model of logical
// workgroup lb and wg[2*lb+1] its workgroup within that model (one 8-byte
// scalar load), tab[m] the model's step arguments (read from device memory).
template <int MODE, int PML, bool FS, int TX, int TY>
__global__ __launch_bounds__(TX * TY) void k_stress_group(const StepArgs* __restrict__ tab,
                                                         const int32_t* __restrict__ wg, int32_t order) {
  const int lb = order ? xcd_remap(blockIdx.x, gridDim.x) : (int)blockIdx.x;
  const int2 ml = *reinterpret_cast<const int2*>(wg + 2 * (int64_t)lb);
  stress_tile<MODE, PML, FS, TX, TY, 1>(tab[ml.x], ml.y, 0);
}
template <int MODE, int PML, int TX, int TY>
__global__ __launch_bounds__(TX * TY) void k_velocity_group(const StepArgs* __restrict__ tab,
                                                           const int32_t* __restrict__ wg, int32_t order) {
  const int lb = order ? xcd_remap(blockIdx.x, gridDim.x) : (int)blockIdx.x;
  const int2 ml = *reinterpret_cast<const int2*>(wg + 2 * (int64_t)lb);
  velocity_tile<MODE, PML, TX, TY, 1>(tab[ml.x], ml.y, 0);
}

// ---------------------------------------------------------------------------
// launchers
// ---------------------------------------------------------------------------
template <int MODE, int PML, bool FS, int TX, int NB>
static hipError_t go_stress2(const StepArgs& a, hipStream_t s) {
  const int nblk = launch_blocks(a) * a.nbg;
  if (a.ty_rows == 4) {
    hipLaunchKernelGGL((k_stress<MODE, PML, FS, TX, 4, NB>), dim3(nblk), dim3(TX, 4), 0, s, a);
  } else if (a.ty_rows == 8) {
    hipLaunchKernelGGL((k_stress<MODE, PML, FS, TX, 8, NB>), dim3(nblk), dim3(TX, 8), 0, s, a);
  } else if (a.ty_rows == 2) {
    hipLaunchKernelGGL((k_stress<MODE, PML, FS, TX, 2, NB>), dim3(nblk), dim3(TX, 2), 0, s, a);
  } else {
    return hipErrorInvalidValue;
  }
  return hipGetLastError();
}

template <int MODE, int PML, bool FS, int NB>
static hipError_t go_stress1(const StepArgs& a, hipStream_t s) {
  if (a.tile_x == 64) return go_stress2<MODE, PML, FS, 64, NB>(a, s);
  if (a.tile_x == 128) return go_stress2<MODE, PML, FS, 128, NB>(a, s);
  return hipErrorInvalidValue;
}

template <int MODE, int PML, bool FS>
static hipError_t go_stress(const StepArgs& a, hipStream_t s) {
  switch (a.nb) {
    case 1: return go_stress1<MODE, PML, FS, 1>(a, s);
    case 2: return go_stress1<MODE, PML, FS, 2>(a, s);
    case 3: return go_stress1<MODE, PML, FS, 3>(a, s);
    case 4: return go_stress1<MODE, PML, FS, 4>(a, s);
    case 5: return go_stress1<MODE, PML, FS, 5>(a, s);
    case 6: return go_stress1<MODE, PML, FS, 6>(a, s);
    case 7: return go_stress1<MODE, PML, FS, 7>(a, s);
    case 8: return go_stress1<MODE, PML, FS, 8>(a, s);
    default: return hipErrorInvalidValue;
  }
}

template <int MODE, int PML, int TX, int NB>
static hipError_t go_vel2(const StepArgs& a, hipStream_t s) {
  const int nblk = launch_blocks(a) * a.nbg;
  if (a.ty_rows == 4) {
    hipLaunchKernelGGL((k_velocity<MODE, PML, TX, 4, NB>), dim3(nblk), dim3(TX, 4), 0, s, a);
  } else if (a.ty_rows == 8) {
    hipLaunchKernelGGL((k_velocity<MODE, PML, TX, 8, NB>), dim3(nblk), dim3(TX, 8), 0, s, a);
  } else if (a.ty_rows == 2) {
    hipLaunchKernelGGL((k_velocity<MODE, PML, TX, 2, NB>), dim3(nblk), dim3(TX, 2), 0, s, a);
  } else {
    return hipErrorInvalidValue;
  }
  return hipGetLastError();
}

template <int MODE, int PML, int NB>
static hipError_t go_vel1(const StepArgs& a, hipStream_t s) {
  if (a.tile_x == 64) return go_vel2<MODE, PML, 64, NB>(a, s);
  if (a.tile_x == 128) return go_vel2<MODE, PML, 128, NB>(a, s);
  return hipErrorInvalidValue;
}

template <int MODE, int PML>
static hipError_t go_vel(const StepArgs& a, hipStream_t s) {
  switch (a.nb) {
    case 1: return go_vel1<MODE, PML, 1>(a, s);
    case 2: return go_vel1<MODE, PML, 2>(a, s);
    case 3: return go_vel1<MODE, PML, 3>(a, s);
    case 4: return go_vel1<MODE, PML, 4>(a, s);
    case 5: return go_vel1<MODE, PML, 5>(a, s);
    case 6: return go_vel1<MODE, PML, 6>(a, s);
    case 7: return go_vel1<MODE, PML, 7>(a, s);
    case 8: return go_vel1<MODE, PML, 8>(a, s);
    default: return hipErrorInvalidValue;
  }
}

template <int MODE, int PML, bool FS, int TX>
static hipError_t go_group2(const StepArgs* tab, const int32_t* wg, int nblk, int ty, int order, bool vel,
                            hipStream_t s) {
  if (ty == 2) {
    if (vel) hipLaunchKernelGGL((k_velocity_group<MODE, PML, TX, 2>), dim3(nblk), dim3(TX, 2), 0, s, tab, wg, order);
    else hipLaunchKernelGGL((k_stress_group<MODE, PML, FS, TX, 2>), dim3(nblk), dim3(TX, 2), 0, s, tab, wg, order);
  } else if (ty == 4) {
    if (vel) hipLaunchKernelGGL((k_velocity_group<MODE, PML, TX, 4>), dim3(nblk), dim3(TX, 4), 0, s, tab, wg, order);
    else hipLaunchKernelGGL((k_stress_group<MODE, PML, FS, TX, 4>), dim3(nblk), dim3(TX, 4), 0, s, tab, wg, order);
  } else if (ty == 8) {
    if (vel) hipLaunchKernelGGL((k_velocity_group<MODE, PML, TX, 8>), dim3(nblk), dim3(TX, 8), 0, s, tab, wg, order);
    else hipLaunchKernelGGL((k_stress_group<MODE, PML, FS, TX, 8>), dim3(nblk), dim3(TX, 8), 0, s, tab, wg, order);
  } else {
    return hipErrorInvalidValue;
  }
  return hipGetLastError();
}

template <int MODE, int PML, bool FS>
static hipError_t go_group(const StepArgs* tab, const int32_t* wg, int nblk, int tx, int ty, int order,
                           bool vel, hipStream_t s) {
  if (tx == 64) return go_group2<MODE, PML, FS, 64>(tab, wg, nblk, ty, order, vel, s);
  if (tx == 128) return go_group2<MODE, PML, FS, 128>(tab, wg, nblk, ty, order, vel, s);
  return hipErrorInvalidValue;
}

template <int MODE, int PML>
hipError_t stress_mp(const StepArgs& a, int fs, hipStream_t s) {
  return fs ? go_stress<MODE, PML, true>(a, s) : go_stress<MODE, PML, false>(a, s);
}

template <int MODE, int PML>
hipError_t velocity_mp(const StepArgs& a, hipStream_t s) {
  return go_vel<MODE, PML>(a, s);
}

template <int MODE, int PML>
hipError_t group_mp(const StepArgs* tab, const int32_t* wg, int nblk, int tx, int ty, int order, int fs,
                    bool vel, hipStream_t s) {
  return fs ? go_group<MODE, PML, true>(tab, wg, nblk, tx, ty, order, vel, s)
            : go_group<MODE, PML, false>(tab, wg, nblk, tx, ty, order, vel, s);
}

}  // namespace fwe
