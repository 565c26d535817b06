// fwe_fused.hip -- single-pass gfx950 kernel of one FDTD step.
//
// Same arithmetic as the two-pass kernels (and so as the reference's
// compute_StressPrf/update_Qxft, OK4.cu:285-388, and compute_StressOrtho/
// update_VelocityFromStress, ANISO.cu:287-389), scheduled differently:
// a workgroup owns a 64 x TY column tile and marches its z chunk once per
// step.  At plane m it
//   1. loads V^n of plane m (plus Vz of m+1) and stages it in LDS,
//   2. computes S^{n+1/2} of plane m for its cells, plus the few halo
//      values its velocity update needs from neighbouring tiles
//      (Syy of row j0-1, Sxy/Syz of row j0+TY by two extra waves; Sxx of
//      column i0-1 by lane 0, Sxy/Sxz of column i0+64 by lane 63),
//   3. computes V^{n+1} of plane m-1, whose stress neighbours are now all
//      known: z from registers, y from LDS, x from wavefront shuffles.
// Time level n is read from one buffer set and n+1 written to the other,
// so tiles never race on shared halo cells.  Per cell and step the fields
// are read once and written once (18 streams) and the medium is read once
// (9 moduli + 3 buoyancies): 30 FP32 streams against 39 for two passes.
#include "fwe_internal.h"
#include "fwe_stencil.h"

namespace fwe {

__device__ __forceinline__ int xcd_remap_f(int b, int nb) {
  const int q = nb >> 3, r = nb & 7;
  const int x = b & 7, w = b >> 3;
  return x * q + (x < r ? x : r) + w;
}

// Unconditional load of p[idx] that yields 0 when !use; lanes that do not
// need it read p[0] instead (always in bounds), so no branch is taken.  The
// halo rows and columns are re-read by neighbouring tiles, so these loads go
// through the caches; the outputs leave with streaming stores.
__device__ __forceinline__ float ldz(const float* p, bool use, int64_t idx) {
  const float v = p[use ? idx : 0];
  return use ? v : 0.f;
}

// Workgroup barrier for data exchanged through LDS only.  The fences are
// scoped to LDS ("local"), so the wave waits for its own LDS traffic
// (lgkmcnt) but keeps its global loads in flight across the barrier; the
// compiler still waits for each load before its first use.  A plain
// __syncthreads() would also drain every outstanding global load.
__device__ __forceinline__ void lds_sync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup", "local");
  __builtin_amdgcn_s_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup", "local");
}

// sum of the injections of one stress component at one cell this step
__device__ __forceinline__ float src_at(const FusedArgs& a, int e0, int e1, float amp, int comp,
                                        int i, int j, int k) {
  float s = 0.f;
  for (int e = e0; e < e1; ++e) {
    const PointEntry p = a.inj[e];
    if (p.comp == comp && p.i == i && p.j == j && p.k == k && p.batch == a.batch) s = fmaf(p.coef, amp, s);
  }
  return s;
}

template <int MODE, int PML, int TY>
__global__ __launch_bounds__(64 * (TY + 2)) void k_fused(const FusedArgs a) {
  __shared__ float LV[2][3][TY + 2][64];  // Vx Vy Vz of plane m, rows j0-1..j0+TY
  __shared__ float LS[2][3][TY + 2][64];  // Syy Sxy Syz of plane m
  const int lane = threadIdx.x, w = threadIdx.y, r = w - 1;
  int lb = xcd_remap_f(blockIdx.x, gridDim.x);
  const int tx = lb % a.ntx;
  lb /= a.ntx;
  const int tyi = lb % a.nty;
  const int kc = lb / a.nty;
  const int nx = a.nx, ny = a.ny, nz = a.nz;
  const int i0 = tx * TILE_X, j0 = tyi * TY;
  const int i = i0 + lane, j = j0 + r;
  const bool interior = r >= 0 && r < TY;
  const bool halo_lo = r == -1, halo_hi = r == TY;
  const bool jin = j >= 0 && j < ny;
  const bool ok = i < nx && jin;              // the thread owns (or halo-computes) cell (i,j)
  const bool lx = i <= nx && jin;             // Vx exists at (i,j)
  const bool ly = i < nx && j >= 0 && j <= ny;  // Vy exists at (i,j)
  const int jc = j < 0 ? 0 : (j > ny ? ny : j);
  const int64_t sy = a.sy, sz = a.sz;
  const int k0 = kc * a.kchunk;
  const int k1 = min(k0 + a.kchunk, nz);
  const int tile = tx + tyi * a.ntx;
  const int e0 = a.inj_off[tile], e1 = a.inj_off[tile + 1];
  const int q0 = a.rec_off[tile], q1 = a.rec_off[tile + 1];
  const int it = a.it + (a.it_base ? *a.it_base : 0);
  const float amp = (e1 > e0 && it < a.nstf) ? a.stf[it] : 0.f;
  const bool have_src = e1 > e0;
  const bool im = i >= 1, jm = j >= 1;
  // x-halo duties (interior rows only)
  const bool xl = interior && ok && lane == 0 && i0 >= 1;                 // Sxx(i0-1)
  const bool xr = interior && ok && lane == 63 && i0 + TILE_X <= nx - 1;  // Sxy,Sxz(i0+64)
  // velocity update masks (OK4.cu:357,364,371)
  const bool ux_ij = interior && ok && im && jm && j <= ny - 2;
  const bool uy_ij = interior && ok && im && i <= nx - 2 && jm;
  const bool uz_ij = interior && ok && im && i <= nx - 2 && jm && j <= ny - 2;

  const float* Vx = a.Vi[0];
  const float* Vy = a.Vi[1];
  const float* Vz = a.Vi[2];

  // per-column damping factors
  float fxh = 1.f, fyh = 1.f, wxi = 1.f, wyj = 1.f, wxm = 1.f, wym = 1.f;
  if (PML == PML_PROFILE && ok) {
    fxh = a.whx[i]; fyh = a.why[j]; wxi = a.wx[i]; wyj = a.wy[j];
    wxm = im ? a.wx[i - 1] : 1.f;
    wym = jm ? a.wy[j - 1] : 1.f;
  }
  if (PML == PML_SEP && ok) { wxi = a.wx[i]; wyj = a.wy[j]; }

  const int mstart = k0 >= 1 ? k0 - 1 : 0;
  // carried state: V^n of plane m-1 and Vz of plane m; S^{n+1/2} of m-1, Szz of m-2
  float vx_p = 0.f, vy_p = 0.f, vz_p = 0.f, vz_c = 0.f;
  float sp0 = 0.f, sp1 = 0.f, sp2 = 0.f, sp3 = 0.f, sp4 = 0.f, sp5 = 0.f, szz_pp = 0.f;
  float sxxL_p = 0.f, sxyR_p = 0.f, sxzR_p = 0.f, w_p = 1.f;
  {
    const int64_t c = (int64_t)i + jc * sy + (int64_t)mstart * sz;
    if (mstart >= 1) {
      if (lx) vx_p = Vx[c - sz];
      if (ly) vy_p = Vy[c - sz];
    }
    if (ok) vz_c = Vz[c];
  }

  for (int m = mstart; m <= k1; ++m) {
    const int buf = m & 1;
    const bool sm = m < nz;  // stress plane exists
    const int64_t c = (int64_t)i + jc * sy + (int64_t)m * sz;
    const bool km = m >= 1;
    const int kv = m - 1;
    const bool vel = interior && ok && kv >= k0 && kv < k1;
    const int64_t cv = c - sz;
    const bool ux = vel && ux_ij && kv >= 1 && kv <= nz - 2;
    const bool uy = vel && uy_ij && kv >= 1 && kv <= nz - 2;
    const bool uz = vel && uz_ij && kv >= 1;

    // ---- phase 1: issue every global load of this plane ----
    // Branch-free: a load a lane does not need reads element 0 (one cached
    // line) and is zeroed, so the loop body stays one basic block and the
    // compiler can keep all of these loads in flight together.
    const bool st = sm && ok;
    const float vx_c = ldz(Vx, sm && lx, c);
    const float vy_c = ldz(Vy, sm && ly, c);
    const float vz_n = ldz(Vz, st, c + sz);
    const float vx_r = ldz(Vx, lane == 63 && sm && i + 1 <= nx && jin, c + 1);  // Vx(i0+64)
    const float vy_l = ldz(Vy, lane == 0 && sm && im && jin, c - 1);            // Vy(i0-1)
    const float vz_l = ldz(Vz, lane == 0 && sm && im && jin, c - 1);            // Vz(i0-1)
    // old stress and moduli of the cells this thread updates (by role)
    const bool ri = st && interior, rlo = st && halo_lo, rhi = st && halo_hi;
    const float so0 = ldz(a.Si[0], ri, c);
    const float so1 = ldz(a.Si[1], ri || rlo, c);
    const float so2 = ldz(a.Si[2], ri, c);
    const float so3 = ldz(a.Si[3], ((ri && jm) || rhi) && im, c);
    const float so4 = ldz(a.Si[4], ri && im && km, c);
    const float so5 = ldz(a.Si[5], ((ri && jm) || rhi) && km, c);
    float m0, m1, m2 = 0.f, m3 = 0.f, m4 = 0.f, m5 = 0.f, m6 = 0.f, m7 = 0.f, m8 = 0.f;
    if (MODE == MODE_ISO) {
      m0 = ldz(a.C[0], ri || rlo, c);  // c11
      m1 = ldz(a.C[1], ri || rlo, c);  // c12
    } else {
      m0 = ldz(a.C[0], ri, c);         // C11
      m1 = ldz(a.C[1], ri || rlo, c);  // C22
      m2 = ldz(a.C[2], ri, c);         // C33
      m3 = ldz(a.C[3], ri || rlo, c);  // C12
      m4 = ldz(a.C[4], ri, c);         // C13
      m5 = ldz(a.C[5], ri || rlo, c);  // C23
      m6 = ldz(a.C[6], ri || rhi, c);  // C44
      m7 = ldz(a.C[7], ri, c);         // C55
      m8 = ldz(a.C[8], ri || rhi, c);  // C66
    }
    float w_c = 1.f;
    if (PML == PML_W3D) w_c = ldz(a.w3, st, c);
    // x-halo inputs: Sxx(i0-1) on lane 0, Sxy/Sxz(i0+64) on lane 63
    const bool gl = st && xl, gr = st && xr;
    const int64_t cl = c - 1, cr = c + 1;
    const float xa0 = ldz(Vx, gl, cl), xa1 = ldz(Vy, gl, cl), xa2 = ldz(Vy, gl, cl + sy);
    const float xa3 = ldz(Vz, gl, cl), xa4 = ldz(Vz, gl, cl + sz);
    const float xa5 = ldz(a.Si[0], gl, cl);
    const float xa6 = ldz(a.C[0], gl, cl);                                // c11 / C11
    const float xa7 = ldz(a.C[MODE == MODE_ISO ? 1 : 3], gl, cl);         // c12 / C12
    const float xw = MODE == MODE_ISO ? 0.f : ldz(a.C[4], gl, cl);        // C13
    const float xlw = PML == PML_W3D ? ldz(a.w3, gl, cl) : 1.f;
    const float xb0 = ldz(Vy, gr, cr), xb1 = ldz(Vx, gr && jm, cr - sy);
    const float xb2 = ldz(Vz, gr, cr), xb3 = ldz(Vx, gr && km, cr - sz);
    const float xb4 = ldz(a.Si[3], gr && jm, cr), xb5 = ldz(a.Si[4], gr && km, cr);
    const float xb6 = MODE == MODE_ISO ? 0.f : ldz(a.C[8], gr, cr);       // C66
    const float xb7 = MODE == MODE_ISO ? 0.f : ldz(a.C[7], gr, cr);       // C55
    const float xbw = PML == PML_W3D ? ldz(a.w3, gr, cr) : 1.f;
    // buoyancy of the velocity plane m-1
    float bx = a.bxdt, by = a.bydt, bz = a.bzdt;
    if (MODE == MODE_ORTHO) {
      bx = ldz(a.B[0], ux, cv);
      by = ldz(a.B[1], uy, cv);
      bz = ldz(a.B[2], uz, cv);
    }

    // ---- phase 2: stage V^n of plane m ----
    LV[buf][0][w][lane] = vx_c;
    LV[buf][1][w][lane] = vy_c;
    LV[buf][2][w][lane] = vz_c;
    lds_sync();

    float vx_ip = __shfl_down(vx_c, 1);
    float vy_im = __shfl_up(vy_c, 1);
    float vz_im = __shfl_up(vz_c, 1);
    if (lane == 63) vx_ip = vx_r;
    if (lane == 0) { vy_im = vy_l; vz_im = vz_l; }
    const float vy_jp = (w <= TY) ? LV[buf][1][w + 1][lane] : 0.f;
    const float vx_jm = (w >= 1) ? LV[buf][0][w - 1][lane] : 0.f;
    const float vz_jm = (w >= 1) ? LV[buf][2][w - 1][lane] : 0.f;

    // ---- phase 3: S^{n+1/2} of plane m ----
    float s0 = 0.f, s1 = 0.f, s2 = 0.f, s3 = 0.f, s4 = 0.f, s5 = 0.f;
    float sxxL = 0.f, sxyR = 0.f, sxzR = 0.f;
    if (st) {
      const float ex = (vx_ip - vx_c) * a.rdx, ey = (vy_jp - vy_c) * a.rdy, ez = (vz_n - vz_c) * a.rdz;
      const float dxy = rate2(vy_c - vy_im, a.rdx, vx_c - vx_jm, a.rdy);  // dVy/dx + dVx/dy
      const float dxz = rate2(vz_c - vz_im, a.rdx, vx_c - vx_p, a.rdz);   // dVz/dx + dVx/dz
      const float dyz = rate2(vy_c - vy_p, a.rdz, vz_c - vz_jm, a.rdy);   // dVy/dz + dVz/dy
      float wzk = 1.f, fzh = 1.f;
      if (PML == PML_PROFILE) { wzk = a.wz[m]; fzh = a.whz[m]; }
      if (PML == PML_SEP) { wzk = a.wz[m]; w_c = wxi * wyj * wzk; }
      if (interior) {
        // own cell, all six components (OK4.cu:307-329 / ANISO.cu:306-352)
        if (MODE == MODE_ISO) {
          s0 = iso_normal(fxh, so0, m0, ex, m1, ey, m1, ez);
          s1 = iso_normal(fyh, so1, m1, ex, m0, ey, m1, ez);
          s2 = iso_normal(fzh, so2, m1, ex, m1, ey, m0, ez);
          s3 = iso_shear(wzk * wyj * wxi, so3, a.c66dt, dxy);
          s4 = iso_shear(wyj * wzk * wxi, so4, a.c66dt, dxz);
          s5 = iso_shear(wxi * wyj * wzk, so5, a.c66dt, dyz);
        } else {
          s0 = ortho_normal(w_c, so0, m0, ex, m3, ey, m4, ez);
          s1 = ortho_normal(w_c, so1, m3, ex, m1, ey, m5, ez);
          s2 = ortho_normal(w_c, so2, m4, ex, m5, ey, m2, ez);
          s3 = ortho_shear(w_c, so3, m8, dxy);
          s4 = ortho_shear(w_c, so4, m7, dxz);
          s5 = ortho_shear(w_c, so5, m6, dyz);
        }
        if (!(im && jm)) s3 = 0.f;
        if (!(im && km)) s4 = 0.f;
        if (!(jm && km)) s5 = 0.f;
        if (have_src) {
          s0 += src_at(a, e0, e1, amp, 0, i, j, m);
          s1 += src_at(a, e0, e1, amp, 1, i, j, m);
          s2 += src_at(a, e0, e1, amp, 2, i, j, m);
          s3 += src_at(a, e0, e1, amp, 3, i, j, m);
          s4 += src_at(a, e0, e1, amp, 4, i, j, m);
          s5 += src_at(a, e0, e1, amp, 5, i, j, m);
        }
        if (m >= k0 && m < k1) {
          __builtin_nontemporal_store(s0, a.So[0] + c); __builtin_nontemporal_store(s1, a.So[1] + c); __builtin_nontemporal_store(s2, a.So[2] + c);
          if (im && jm) __builtin_nontemporal_store(s3, a.So[3] + c);
          if (im && km) __builtin_nontemporal_store(s4, a.So[4] + c);
          if (jm && km) __builtin_nontemporal_store(s5, a.So[5] + c);
        }
      } else if (halo_lo) {
        // Syy of row j0-1, needed by the Vy update of row j0
        if (MODE == MODE_ISO) s1 = iso_normal(fyh, so1, m1, ex, m0, ey, m1, ez);
        else s1 = ortho_normal(w_c, so1, m3, ex, m1, ey, m5, ez);
        if (have_src) s1 += src_at(a, e0, e1, amp, 1, i, j, m);
      } else {
        // Sxy and Syz of row j0+TY, needed by the Vx and Vz updates of row j0+TY-1
        if (MODE == MODE_ISO) {
          s3 = iso_shear(wzk * wyj * wxi, so3, a.c66dt, dxy);
          s5 = iso_shear(wxi * wyj * wzk, so5, a.c66dt, dyz);
        } else {
          s3 = ortho_shear(w_c, so3, m8, dxy);
          s5 = ortho_shear(w_c, so5, m6, dyz);
        }
        if (!im) s3 = 0.f;
        if (!km) s5 = 0.f;
        if (have_src) {
          s3 += src_at(a, e0, e1, amp, 3, i, j, m);
          s5 += src_at(a, e0, e1, amp, 5, i, j, m);
        }
      }
      if (xl) {
        // Sxx of column i0-1 (Vx update of lane 0)
        const float exl = (vx_c - xa0) * a.rdx, eyl = (xa2 - xa1) * a.rdy, ezl = (xa4 - xa3) * a.rdz;
        if (MODE == MODE_ISO) {
          sxxL = iso_normal(a.whx[i - 1], xa5, xa6, exl, xa7, eyl, xa7, ezl);
        } else {
          float wl = 1.f;
          if (PML == PML_W3D) wl = xlw;
          if (PML == PML_SEP) wl = a.wx[i - 1] * wyj * wzk;
          sxxL = ortho_normal(wl, xa5, xa6, exl, xa7, eyl, xw, ezl);
        }
        if (have_src) sxxL += src_at(a, e0, e1, amp, 0, i - 1, j, m);
      }
      if (xr) {
        // Sxy and Sxz of column i0+64 (Vy and Vz updates of lane 63)
        const float dxyr = rate2(xb0 - vy_c, a.rdx, vx_r - xb1, a.rdy);
        const float dxzr = rate2(xb2 - vz_c, a.rdx, vx_r - xb3, a.rdz);
        if (MODE == MODE_ISO) {
          const float wxr = a.wx[i + 1];
          sxyR = iso_shear(wzk * wyj * wxr, xb4, a.c66dt, dxyr);
          sxzR = iso_shear(wyj * wzk * wxr, xb5, a.c66dt, dxzr);
        } else {
          float wr = 1.f;
          if (PML == PML_W3D) wr = xbw;
          if (PML == PML_SEP) wr = a.wx[i + 1] * wyj * wzk;
          sxyR = ortho_shear(wr, xb4, xb6, dxyr);
          sxzR = ortho_shear(wr, xb5, xb7, dxzr);
        }
        if (!jm) sxyR = 0.f;
        if (!km) sxzR = 0.f;
        if (have_src) {
          sxyR += src_at(a, e0, e1, amp, 3, i + 1, j, m);
          sxzR += src_at(a, e0, e1, amp, 4, i + 1, j, m);
        }
      }
    }
    LS[buf][0][w][lane] = s1;
    LS[buf][1][w][lane] = s3;
    LS[buf][2][w][lane] = s5;

    // ---- phase 4: V^{n+1} of plane m-1 (OK4.cu:351-378 / ANISO.cu:374-386) ----
    float sxx_im = __shfl_up(sp0, 1);
    float sxy_ip = __shfl_down(sp3, 1);
    float sxz_ip = __shfl_down(sp4, 1);
    if (lane == 0) sxx_im = sxxL_p;
    if (lane == 63) { sxy_ip = sxyR_p; sxz_ip = sxzR_p; }
    if (vel) {
      const int pb = buf ^ 1;
      const float syy_jm = LS[pb][0][w - 1][lane];
      const float sxy_jp = LS[pb][1][w + 1][lane];
      const float syz_jp = LS[pb][2][w + 1][lane];
      float vx = vx_p, vy = vy_p, vz = vz_p;
      float wzm = 1.f;
      if (PML == PML_PROFILE) wzm = kv >= 1 ? a.wz[kv - 1] : 1.f;
      if (ux) {
        const float dS = rate3(sp0 - sxx_im, a.rdx, sxy_jp - sp3, a.rdy, s4 - sp4, a.rdz);
        vx = (MODE == MODE_ISO) ? iso_vel(wxm, vx, bx, dS) : ortho_vel(w_p, vx, bx, dS);
        __builtin_nontemporal_store(vx, a.Vo[0] + cv);
      }
      if (uy) {
        const float dS = rate3(sp1 - syy_jm, a.rdy, sxy_ip - sp3, a.rdx, s5 - sp5, a.rdz);
        vy = (MODE == MODE_ISO) ? iso_vel(wym, vy, by, dS) : ortho_vel(w_p, vy, by, dS);
        __builtin_nontemporal_store(vy, a.Vo[1] + cv);
      }
      if (uz) {
        const float dS = rate3(sp2 - szz_pp, a.rdz, sxz_ip - sp4, a.rdx, syz_jp - sp5, a.rdy);
        vz = (MODE == MODE_ISO) ? iso_vel(wzm, vz, bz, dS) : ortho_vel(w_p, vz, bz, dS);
        __builtin_nontemporal_store(vz, a.Vo[2] + cv);
      }
      for (int q = q0; q < q1; ++q) {
        const PointEntry p = a.rec[q];
        if (p.i == i && p.j == j && p.k == kv && p.batch == a.batch && it < a.nt)
          a.traces[(int64_t)p.slot * a.nt + it] = p.comp == 0 ? vx : (p.comp == 1 ? vy : vz);
      }
    }
    // advance the carried planes
    szz_pp = sp2;
    sp0 = s0; sp1 = s1; sp2 = s2; sp3 = s3; sp4 = s4; sp5 = s5;
    sxxL_p = sxxL; sxyR_p = sxyR; sxzR_p = sxzR;
    vx_p = vx_c; vy_p = vy_c; vz_p = vz_c; vz_c = vz_n;
    w_p = w_c;
  }
}

template <int MODE, int PML>
static hipError_t go_fused(const FusedArgs& a, hipStream_t s) {
  const int nblk = a.ntx * a.nty * a.nkc;
  if (a.ty_rows == 4) {
    hipLaunchKernelGGL((k_fused<MODE, PML, 4>), dim3(nblk), dim3(64, 6), 0, s, a);
  } else if (a.ty_rows == 8) {
    hipLaunchKernelGGL((k_fused<MODE, PML, 8>), dim3(nblk), dim3(64, 10), 0, s, a);
  } else if (a.ty_rows == 2) {
    hipLaunchKernelGGL((k_fused<MODE, PML, 2>), dim3(nblk), dim3(64, 4), 0, s, a);
  } else {
    return hipErrorInvalidValue;
  }
  return hipGetLastError();
}

hipError_t launch_fused(const FusedArgs& a, int mode, int pml, hipStream_t s) {
  if (mode == MODE_ISO) {
    if (pml == PML_PROFILE) return go_fused<MODE_ISO, PML_PROFILE>(a, s);
    if (pml == PML_NONE) return go_fused<MODE_ISO, PML_NONE>(a, s);
    return hipErrorInvalidValue;
  }
  switch (pml) {
    case PML_NONE: return go_fused<MODE_ORTHO, PML_NONE>(a, s);
    case PML_W3D: return go_fused<MODE_ORTHO, PML_W3D>(a, s);
    case PML_SEP: return go_fused<MODE_ORTHO, PML_SEP>(a, s);
    default: return hipErrorInvalidValue;
  }
}

}  // namespace fwe
