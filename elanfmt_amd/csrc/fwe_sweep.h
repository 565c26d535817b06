// fwe_sweep.h -- single-pass gfx950 kernel of one FDTD step ("sweep").
//
// Same cell formulas as the two-pass kernels (fwe_stencil.h; the reference's
// compute_StressPrf/update_Qxft, OK4.cu:285-388, and compute_StressOrtho/
// update_VelocityFromStress, ANISO.cu:287-389, with BC1 OK4.cu:271-283 and
// the injections/receivers of OK4.cu:336-342,383-386), so the same bits,
// scheduled as one z-march per step:
//  * every wavefront is an independent pipeline (no LDS, no barrier): its
//    64 lanes cover columns x0-1 .. x0+62 of a 62-column tile, so lanes
//    1..62 produce the tile and the two edge lanes supply the x-halo
//    stresses (Sxx left, Sxy/Sxz right); x-neighbours move between lanes by
//    DPP wave shifts;
//  * every lane holds R rows of the tile plus one halo row on each side
//    (Syy below, Sxy/Syz above) in registers, so y-neighbours are register
//    reads;
//  * at plane m it computes S^{n+1/2} of plane m from V^n of planes m-1..m+1
//    and then V^{n+1} of plane m-1 from S^{n+1/2} of planes m-2..m, all
//    carried in registers;
//  * time level n is read from one buffer set and n+1 written to the
//    other, so concurrent tiles never race on shared halo cells.
// Per cell and step the nine fields are read once and written once, and
// the medium is read once: 73 bytes with the material table against 110
// for the two passes.  A z chunk recomputes the stresses of one plane on
// each side of it (Szz below, Sxz/Syz above).
#pragma once
#include "fwe_internal.h"
#include "fwe_stencil.h"

namespace fwe {

constexpr int SWEEP_WPG = 4;   // independent wavefronts per workgroup


// lane l receives lane l-1's value (lane 0 receives 0)
__device__ __forceinline__ float from_left(float v) {
  return __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(v), 0x138, 0xf, 0xf, true));
}
// lane l receives lane l+1's value (lane 63 receives 0)
__device__ __forceinline__ float from_right(float v) {
  return __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(v), 0x130, 0xf, 0xf, true));
}

// element at byte offset `b` of array p (offsets are 32-bit: the host
// checks that one array spans less than 4 GiB)
__device__ __forceinline__ float ldb(const float* p, uint32_t b) {
  return *reinterpret_cast<const float*>(reinterpret_cast<const char*>(p) + b);
}
// Plain (write-back) stores: measured faster here than streaming stores,
// whose completion the next plane's loads then wait behind.
__device__ __forceinline__ void stb(float* p, uint32_t b, float v) {
  *reinterpret_cast<float*>(reinterpret_cast<char*>(p) + b) = v;
}

// the nine moduli (times dt) of a cell, in the two-pass kernels' naming:
// m0..m5 = C11 C22 C33 C12 C13 C23, m6 = C44yz, m7 = C55xz, m8 = C66xy;
// iso law: m0 = c11, m1 = c12
struct Mod {
  float m0, m1, m2, m3, m4, m5, m6, m7, m8;
};

template <int MODE>
__device__ __forceinline__ uint32_t sweep_id(const SweepArgs& a, uint32_t eb) {
  if (MODE == MODE_ORTHO_TAB || MODE == MODE_ISO_TAB) return (uint32_t)a.mid[eb >> 2];
  return 0u;
}

// moduli at element byte offset eb (table modes: entry id of the table
// `tab`, a copy of a.mtab in LDS)
template <int MODE>
__device__ __forceinline__ Mod sweep_mod(const SweepArgs& a, const float* tab, uint32_t eb, uint32_t id) {
  Mod d{0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f};
  if (MODE == MODE_ORTHO_TAB) {
    const float* t = tab + 12u * id;
    const float4 q0 = *reinterpret_cast<const float4*>(t);
    const float4 q1 = *reinterpret_cast<const float4*>(t + 4);
    d.m0 = q0.x; d.m1 = q0.y; d.m2 = q0.z; d.m3 = q0.w;
    d.m4 = q1.x; d.m5 = q1.y; d.m6 = q1.z; d.m7 = q1.w;
    d.m8 = t[8];
  } else if (MODE == MODE_ISO_TAB) {
    const float2 q = *reinterpret_cast<const float2*>(tab + 2u * id);
    d.m0 = q.x; d.m1 = q.y;
  } else if (MODE == MODE_ORTHO_ISO) {
    d.m0 = ldb(a.C[0], eb); d.m3 = ldb(a.C[3], eb);
    d.m1 = d.m0; d.m2 = d.m0; d.m4 = d.m3; d.m5 = d.m3;
    d.m6 = ldb(a.C[6], eb); d.m7 = ldb(a.C[7], eb); d.m8 = ldb(a.C[8], eb);
  } else if (MODE == MODE_ORTHO) {
    d.m0 = ldb(a.C[0], eb); d.m1 = ldb(a.C[1], eb); d.m2 = ldb(a.C[2], eb);
    d.m3 = ldb(a.C[3], eb); d.m4 = ldb(a.C[4], eb); d.m5 = ldb(a.C[5], eb);
    d.m6 = ldb(a.C[6], eb); d.m7 = ldb(a.C[7], eb); d.m8 = ldb(a.C[8], eb);
  } else {  // MODE_ISO
    d.m0 = ldb(a.C[0], eb); d.m1 = ldb(a.C[1], eb);
  }
  return d;
}

// staggered buoyancy (times dt) at element byte offset eb
template <int MODE>
__device__ __forceinline__ void sweep_buoy(const SweepArgs& a, const float* tab, uint32_t eb, uint32_t id,
                                           float& bx, float& by, float& bz) {
  if (MODE == MODE_ORTHO_TAB) {
    const float4 q = *reinterpret_cast<const float4*>(tab + 12u * id + 8);
    bx = q.y; by = q.z; bz = q.w;
  } else if (MODE == MODE_ORTHO || MODE == MODE_ORTHO_ISO) {
    bx = ldb(a.B[0], eb); by = ldb(a.B[1], eb); bz = ldb(a.B[2], eb);
  } else {
    bx = a.bxdt; by = a.bydt; bz = a.bzdt;
  }
}

// Inputs of the stress pass at one plane, loaded ahead of use (see issue()
// in k_sweep): S^n of the plane, V^n of the plane above, medium ids, 3-D
// weight, z damping and the receiver list bounds of the plane below.
template <int NR>
struct SweepPlane {
  float xx[NR], yy[NR], zz[NR], xy[NR], xz[NR], yz[NR], w3[NR], vx[NR], vy[NR], vz[NR];
  uint32_t id[NR];
  float wz, whz;
  int rk0, rk1;
};

// ---------------------------------------------------------------------------
// the kernel.  R = output rows per lane; FS = BC1 top (the vacuum-layer top
// is the runtime plane a.vac_k).
// ---------------------------------------------------------------------------
template <int MODE, int PML, bool FS, int R>
__global__ __launch_bounds__(64 * SWEEP_WPG, R == 2 ? 2 : 1) void k_sweep(const SweepArgs a) {
  constexpr int NR = R + 2;  // row 0 = y0-1 (halo), 1..R = tile rows, R+1 = y0+R (halo)
  constexpr bool ISO = iso_law(MODE);
  constexpr bool TAB = MODE == MODE_ORTHO_TAB || MODE == MODE_ISO_TAB;
  // The material table lives in LDS: its lookups then count on the LDS
  // counter, not on the vector-memory counter that the HBM loads of the
  // next plane hold, so waiting for a lookup never waits for those loads.
  // Filled by the whole workgroup before any wavefront leaves.
  __shared__ __align__(16) float tab[TAB ? (MODE == MODE_ORTHO_TAB ? 12 : 2) * FWE_MEDIUM_TABLE_MAX : 4];
  if (TAB) {
    const int nf = (MODE == MODE_ORTHO_TAB ? 12 : 2) * a.ntab;
    for (int q = threadIdx.y * 64 + threadIdx.x; q < nf; q += 64 * SWEEP_WPG) tab[q] = a.mtab[q];
    __syncthreads();
  }
  const int lane = threadIdx.x;
  // Workgroup -> tiles, XCD-aware: workgroup w runs on XCD w mod 8, and
  // XCD x takes the band of tile rows [x*nyb, (x+1)*nyb) of every chunk
  // and wavefield, so the halo rows and overlap columns two tiles share
  // are fetched into one L2.  Within a band, tiles go x fastest, then y,
  // then chunk.  The tile index is uniform over the wavefront: it is read
  // into a scalar register, so the chunk and plane logic below branches on
  // scalars and every array base stays in scalar registers.
  const int xcd = (int)blockIdx.x & 7;
  int t = __builtin_amdgcn_readfirstlane(((int)blockIdx.x >> 3) * SWEEP_WPG + (int)threadIdx.y);
  const int nyb = a.nyb;
  const int per = a.ntx * nyb * a.nkc;
  const int wf = t / per;
  t -= wf * per;
  const int tx = t % a.ntx;
  const int tyi = xcd * nyb + (t / a.ntx) % nyb;
  const int kc = t / (a.ntx * nyb);
  if (wf >= a.nbatch || tyi >= a.nty) return;  // padding wavefronts
  const int nx = a.nx, ny = a.ny, nz = a.nz;
  const int x0 = tx * SWEEP_X, y0 = tyi * R;
  const int i = x0 - 1 + lane;
  const int k0 = kc * a.kchunk, k1 = min(k0 + a.kchunk, nz);
  const uint32_t bsz = 4u * (uint32_t)a.sz;
  const int64_t fo = (int64_t)wf * a.fstride;
  const float* Vxi = a.Vi[0] + fo; const float* Vyi = a.Vi[1] + fo; const float* Vzi = a.Vi[2] + fo;
  const float* Sxxi = a.Si[0] + fo; const float* Syyi = a.Si[1] + fo; const float* Szzi = a.Si[2] + fo;
  const float* Sxyi = a.Si[3] + fo; const float* Sxzi = a.Si[4] + fo; const float* Syzi = a.Si[5] + fo;
  float* Vxo = a.Vo[0] + fo; float* Vyo = a.Vo[1] + fo; float* Vzo = a.Vo[2] + fo;
  float* Sxxo = a.So[0] + fo; float* Syyo = a.So[1] + fo; float* Szzo = a.So[2] + fo;
  float* Sxyo = a.So[3] + fo; float* Sxzo = a.So[4] + fo; float* Syzo = a.So[5] + fo;

  // Lane and row geometry.  Loads at positions outside the padded arrays
  // (column -1 or past nx, row -1 or past ny, plane -1) read element
  // (0, 0, k) instead and are never zeroed: such values only reach stresses
  // and velocities that no tile cell consumes (shear cells off their update
  // range are set to 0, velocities off their range keep the old value), so
  // the loads can stay in flight until first use.
  const bool lx = i >= 0 && i <= nx;            // element i exists in the padded rows
  const bool ox = lane >= 1 && lane <= SWEEP_X && i < nx;  // this lane writes its column
  const bool im = i >= 1;
  int jr[NR];
  bool ld[NR], orow[NR];
  uint32_t rb[NR];  // byte offset of (i, j_r, 0), 0 where not loadable
#pragma unroll
  for (int r = 0; r < NR; ++r) {
    jr[r] = y0 - 1 + r;
    ld[r] = lx && jr[r] >= 0 && jr[r] <= ny;
    orow[r] = r >= 1 && r <= R && jr[r] < ny;
    rb[r] = ld[r] ? 4u * (uint32_t)(i + jr[r] * a.sy) : 0u;
  }
  const int tile = tx + tyi * a.ntx;
  const int e0 = a.inj_off[tile], e1 = a.inj_off[tile + 1];
  const int itn = a.it + (a.it_base ? *a.it_base : 0);
  const float amp = (e1 > e0 && itn < a.nstf) ? a.stf[(int64_t)wf * a.nstf + itn] : 0.f;
  const float third = 1.0f / 3.0f;

  // per-column damping factors
  float wxi = 1.f, fxh = 1.f, wxm = 1.f;
  if (PML == PML_SEP || PML == PML_PROFILE) wxi = lx && i < nx ? a.wx[i] : 1.f;
  if (PML == PML_PROFILE) {
    fxh = lx && i < nx ? a.whx[i] : 1.f;
    wxm = im && i <= nx ? a.wx[i - 1] : 1.f;
  }
  float wyj[NR], fyh[NR], wym[NR];
#pragma unroll
  for (int r = 0; r < NR; ++r) {
    const bool jv = jr[r] >= 0 && jr[r] < ny;
    wyj[r] = 1.f; fyh[r] = 1.f; wym[r] = 1.f;
    if (PML == PML_SEP || PML == PML_PROFILE) wyj[r] = jv ? a.wy[jr[r]] : 1.f;
    if (PML == PML_PROFILE) {
      fyh[r] = jv ? a.why[jr[r]] : 1.f;
      wym[r] = jr[r] >= 1 && jr[r] <= ny ? a.wy[jr[r] - 1] : 1.f;
    }
  }

  // V^n of planes m-1 (p), m (c), m+1 (n); S^{n+1/2} of plane m-1 (q) and
  // Szz of plane m-2
  float vxp[NR], vyp[NR], vzp[NR], vxc[NR], vyc[NR], vzc[NR];
  float qxx[NR], qyy[NR], qzz[NR], qxy[NR], qxz[NR], qyz[NR], qzz2[NR];
  float wq[NR];      // 3-D weight of plane m-1 (rows 1..R)
  float wzq = 1.f, wzq2 = 1.f;  // wz of planes m-1 and m-2
  // V^{n+1} of plane m-2, stored one iteration late (see the loop)
  float pvxs[NR], pvys[NR], pvzs[NR];
#pragma unroll
  for (int r = 0; r < NR; ++r) pvxs[r] = pvys[r] = pvzs[r] = 0.f;
  uint32_t idq[NR];  // material id of plane m-1 (rows 1..R)
#pragma unroll
  for (int r = 0; r < NR; ++r) {
    qxx[r] = qyy[r] = qzz[r] = qxy[r] = qxz[r] = qyz[r] = qzz2[r] = 0.f;
    wq[r] = 1.f;
    idq[r] = 0u;
  }
  auto load_v = [&](int k, float* vx, float* vy, float* vz) __attribute__((always_inline)) {
    const uint32_t kb = k >= 0 && k <= nz ? (uint32_t)k * bsz : 0u;
#pragma unroll
    for (int r = 0; r < NR; ++r) {
      const uint32_t b = rb[r] + kb;
      vx[r] = ldb(Vxi, b); vy[r] = ldb(Vyi, b); vz[r] = ldb(Vzi, b);
    }
  };
  // Inputs of the stress pass at plane k, loaded one plane ahead (software
  // pipeline: plane m+1 is in flight while plane m computes; two planes
  // ahead measured no faster, as a wavefront's 63 outstanding memory
  // operations cap the second plane): medium ids and 3-D weight first, then S^n of plane k and V^n
  // of plane k+1, the z damping of plane k (wz, and whz for the profile
  // form) and the receiver list bounds of plane k-1 (its velocity is
  // updated at iteration k).  Every load is unconditional (a loaded value
  // merged at a branch join would make the wavefront wait for all its
  // loads there); a component the plane does not need, or a plane past the
  // chunk, reads plane 0 instead (a line the caches keep) and the value is
  // unused.  Uniform values are loaded by the vector unit with the rest (a
  // scalar load of memory the kernel may write is not allowed).
  auto issue = [&](int k, SweepPlane<NR>& P) __attribute__((always_inline)) {
    const bool here = k <= k1 && k < nz;
    const bool kfull = k >= k0 && k < k1;
    const uint32_t kb = here ? (uint32_t)k * bsz : 0u;
    const uint32_t kf = kfull ? kb : 0u;
    const uint32_t kzz = (kfull || k == k0 - 1) ? kb : 0u;  // Szz below the chunk
    const uint32_t ksh = (kfull || k == k1) ? kb : 0u;      // Sxz/Syz above it
#pragma unroll
    for (int r = 0; r < NR; ++r) {
      P.id[r] = sweep_id<MODE>(a, rb[r] + kb);
      P.w3[r] = PML == PML_W3D ? ldb(a.w3, rb[r] + kb) : 1.f;
    }
#pragma unroll
    for (int r = 0; r < NR; ++r) {
      // rows 0 (Syy), 1..R (all), R+1 (Sxy, Syz)
      const bool mid_row = r >= 1 && r <= R;
      P.xx[r] = mid_row ? ldb(Sxxi, rb[r] + kf) : 0.f;
      P.yy[r] = r <= R ? ldb(Syyi, rb[r] + kf) : 0.f;
      P.zz[r] = mid_row ? ldb(Szzi, rb[r] + kzz) : 0.f;
      P.xy[r] = r >= 1 ? ldb(Sxyi, rb[r] + kf) : 0.f;
      P.xz[r] = mid_row ? ldb(Sxzi, rb[r] + ksh) : 0.f;
      P.yz[r] = r >= 1 ? ldb(Syzi, rb[r] + ksh) : 0.f;
    }
    load_v(k < k1 ? k + 1 : -1, P.vx, P.vy, P.vz);
    const int kz = here ? k : 0;
    P.wz = (PML == PML_PROFILE || PML == PML_SEP) ? a.wz[kz] : 1.f;
    P.whz = PML == PML_PROFILE ? a.whz[kz] : 1.f;
    const int kr = k >= 1 && k - 1 < nz ? tile * nz + k - 1 : 0;
    P.rk0 = a.rec_koff[kr];
    P.rk1 = a.rec_koff[k >= 1 && k - 1 < nz ? kr + 1 : 0];
  };
  const int mstart = k0 >= 1 ? k0 - 1 : 0;
  load_v(mstart - 1, vxp, vyp, vzp);
  load_v(mstart, vxc, vyc, vzc);
  SweepPlane<NR> P1;
  issue(mstart, P1);

  for (int m = mstart; m <= k1; ++m) {
    const bool stress_here = m < nz;               // no stress plane nz
    // ---- loads: medium ids and weight of plane m first (their table
    // entries are gathered while the field loads are in flight), then S^n
    // of plane m and V^n of plane m+1.  Every load is unconditional (a
    // loaded value merged at a branch join would make the wavefront wait
    // for all its loads there); a component the plane does not need reads
    // plane 0 instead (a line the caches keep), and its value is unused.
    const uint32_t kb = stress_here ? (uint32_t)m * bsz : 0u;
    // plane m's inputs, loaded during the previous plane
    float oxx[NR], oyy[NR], ozz[NR], oxy[NR], oxz[NR], oyz[NR], w3[NR], vxn[NR], vyn[NR], vzn[NR];
    uint32_t idc[NR];
#pragma unroll
    for (int r = 0; r < NR; ++r) {
      oxx[r] = P1.xx[r]; oyy[r] = P1.yy[r]; ozz[r] = P1.zz[r]; oxy[r] = P1.xy[r]; oxz[r] = P1.xz[r];
      oyz[r] = P1.yz[r]; w3[r] = P1.w3[r]; vxn[r] = P1.vx[r]; vyn[r] = P1.vy[r]; vzn[r] = P1.vz[r];
      idc[r] = P1.id[r];
    }
    const float wzc = P1.wz, whzc = P1.whz;
    const int rq0 = __builtin_amdgcn_readfirstlane(P1.rk0), rq1 = __builtin_amdgcn_readfirstlane(P1.rk1);
    // the medium of plane m (stress) and the buoyancy of plane m-1
    // (velocity) -- table gathers that hit the caches -- issued before the
    // next plane's loads, so waiting for them does not wait for those
    Mod md[NR];
    float bxq[NR], byq[NR], bzq[NR];
#pragma unroll
    for (int r = 0; r < NR; ++r) {
      md[r] = sweep_mod<MODE>(a, tab, rb[r] + kb, idc[r]);
      bxq[r] = byq[r] = bzq[r] = 0.f;
      if (r >= 1 && r <= R)
        sweep_buoy<MODE>(a, tab, rb[r] + (m >= 1 ? (uint32_t)(m - 1) * bsz : 0u), idq[r], bxq[r], byq[r], bzq[r]);
    }
    issue(m + 1, P1);
    // Stores run one iteration late: S^{n+1/2} of plane m-1 (the carried
    // q arrays) and V^{n+1} of plane m-2, issued after the loads above.
    // Stores and loads share one in-order completion counter, so waiting
    // for a plane's loads waits for every store issued before them; this
    // order leaves each store two iterations to complete instead of one.
    // (With BC1, plane 1's Sxz/Syz are stored at plane 2, once set.)
    if (m - 1 >= k0 && m - 1 < k1) {
      const uint32_t sb = (uint32_t)(m - 1) * bsz;
#pragma unroll
      for (int r = 1; r <= R; ++r) {
        if (!(ox && orow[r])) continue;
        const uint32_t b = rb[r] + sb;
        stb(Sxxo, b, qxx[r]); stb(Syyo, b, qyy[r]); stb(Szzo, b, qzz[r]); stb(Sxyo, b, qxy[r]);
        if (!(FS && m - 1 == 1)) { stb(Sxzo, b, qxz[r]); stb(Syzo, b, qyz[r]); }
      }
    }
    if (m - 2 >= k0) {
      const uint32_t vb = (uint32_t)(m - 2) * bsz;
#pragma unroll
      for (int r = 1; r <= R; ++r) {
        if (!(ox && orow[r])) continue;
        const uint32_t b = rb[r] + vb;
        stb(Vxo, b, pvxs[r]); stb(Vyo, b, pvys[r]); stb(Vzo, b, pvzs[r]);
      }
    }
    // ---- stress of plane m ----------------------------------------------
    float sxx[NR], syy[NR], szz[NR], sxy[NR], sxz[NR], syz[NR];
#pragma unroll
    for (int r = 0; r < NR; ++r) sxx[r] = syy[r] = szz[r] = sxy[r] = sxz[r] = syz[r] = 0.f;
    if (stress_here) {
      const bool km = m >= 1;
      float wzk = 1.f, fzh = 1.f;
      if (PML == PML_PROFILE) { wzk = wzc; fzh = whzc; }
      if (PML == PML_SEP) wzk = wzc;
      const bool vac = m == a.vac_k;
      const float c66k = vac ? a.vac_cdt : a.c66dt;
      // point injections into this plane (and plane 1, for BC1 at plane 2)
      float src[NR][6], src1xz[NR], src1yz[NR];
#pragma unroll
      for (int r = 0; r < NR; ++r) {
        for (int q = 0; q < 6; ++q) src[r][q] = 0.f;
        src1xz[r] = src1yz[r] = 0.f;
      }
      for (int e = e0; e < e1; ++e) {
        const PointEntry p = a.inj[e];
        if (p.batch != wf || p.i != i) continue;
#pragma unroll
        for (int r = 0; r < NR; ++r) {
          if (p.j != jr[r]) continue;
          // component by unrolled selects (a run-time index would put the
          // array in scratch memory)
#pragma unroll
          for (int q = 0; q < 6; ++q)
            if (p.k == m && p.comp == q) src[r][q] = fmaf(p.coef, amp, src[r][q]);
          if (FS && m == 2 && p.k == 1) {
            if (p.comp == 4) src1xz[r] = fmaf(p.coef, amp, src1xz[r]);
            if (p.comp == 5) src1yz[r] = fmaf(p.coef, amp, src1yz[r]);
          }
        }
      }
#pragma unroll
      for (int r = 0; r < NR; ++r) {
        const bool jm = jr[r] >= 1;
        const Mod& d = md[r];
        float m6 = d.m6, m7 = d.m7, m8 = d.m8;
        if (vac) { m6 = a.vac_cdt; m7 = a.vac_cdt; m8 = a.vac_cdt; }
        float w = 1.f;
        if (PML == PML_W3D) w = w3[r];
        if (PML == PML_SEP) w = wxi * wyj[r] * wzk;
        const float vx = vxc[r], vy = vyc[r], vz = vzc[r];
        const float vx_ip = from_right(vx);
        const float vy_im = from_left(vy);
        const float vz_im = from_left(vz);
        const float vy_jp = r + 1 < NR ? vyc[r + 1 < NR ? r + 1 : r] : 0.f;
        const float vx_jm = r >= 1 ? vxc[r >= 1 ? r - 1 : 0] : 0.f;
        const float vz_jm = r >= 1 ? vzc[r >= 1 ? r - 1 : 0] : 0.f;
        const float vz_kp = vzn[r];
        const float vx_km = vxp[r], vy_km = vyp[r];
        const float ex = (vx_ip - vx) * a.rdx, ey = (vy_jp - vy) * a.rdy, ez = (vz_kp - vz) * a.rdz;
        const float dxy = rate2(vy - vy_im, a.rdx, vx - (jm ? vx_jm : 0.f), a.rdy);
        const float dxz = rate2(vz - vz_im, a.rdx, vx - vx_km, a.rdz);
        const float dyz = rate2(vy - vy_km, a.rdz, vz - (jm ? vz_jm : 0.f), a.rdy);
        float nxx, nyy, nzz, nxy, nxz, nyz;
        if (ISO) {
          nxx = iso_normal(fxh, oxx[r], d.m0, ex, d.m1, ey, d.m1, ez);
          nyy = iso_normal(fyh[r], oyy[r], d.m1, ex, d.m0, ey, d.m1, ez);
          nzz = iso_normal(fzh, ozz[r], d.m1, ex, d.m1, ey, d.m0, ez);
          nxy = iso_shear(wzk * wyj[r] * wxi, oxy[r], c66k, dxy);
          nxz = iso_shear(wyj[r] * wzk * wxi, oxz[r], c66k, dxz);
          nyz = iso_shear(wxi * wyj[r] * wzk, oyz[r], c66k, dyz);
        } else {
          nxx = ortho_normal(w, oxx[r], d.m0, ex, d.m3, ey, d.m4, ez);
          nyy = ortho_normal(w, oyy[r], d.m3, ex, d.m1, ey, d.m5, ez);
          nzz = ortho_normal(w, ozz[r], d.m4, ex, d.m5, ey, d.m2, ez);
          nxy = ortho_shear(w, oxy[r], m8, dxy);
          nxz = ortho_shear(w, oxz[r], m7, dxz);
          nyz = ortho_shear(w, oyz[r], m6, dyz);
        }
        // shear cells outside the reference update ranges keep their value
        // (zero unless a point injection adds to them)
        if (!(im && jm)) nxy = oxy[r];
        if (!(im && km)) nxz = oxz[r];
        if (!(jm && km)) nyz = oyz[r];
        if (!a.src_after_bc) {
          nxx += src[r][0]; nyy += src[r][1]; nzz += src[r][2];
          nxy += src[r][3]; nxz += src[r][4]; nyz += src[r][5];
        }
        if (FS) {
          // BC1 (OK4.cu:279-282): Szz(k=0) = 0, Sxz/Syz(k=1) = Sxz/Syz(k=2)/3
          if (m == 0) nzz = 0.f;
          if (m == 2) {
            if (im) qxz[r] = fmaf(third, nxz, a.src_after_bc ? src1xz[r] : 0.f);
            if (jm) qyz[r] = fmaf(third, nyz, a.src_after_bc ? src1yz[r] : 0.f);
          }
        }
        if (a.src_after_bc) {
          nxx += src[r][0]; nyy += src[r][1]; nzz += src[r][2];
          nxy += src[r][3]; nxz += src[r][4]; nyz += src[r][5];
        }
        sxx[r] = nxx; syy[r] = nyy; szz[r] = nzz; sxy[r] = nxy; sxz[r] = nxz; syz[r] = nyz;
      }
      // BC1: plane 1's Sxz/Syz, set from plane 2 above
      if (FS && m == 2 && k0 <= 1) {
#pragma unroll
        for (int r = 1; r <= R; ++r) {
          if (!(ox && orow[r])) continue;
          const uint32_t b = rb[r] + bsz;
          stb(Sxzo, b, qxz[r]); stb(Syzo, b, qyz[r]);
        }
      }
    }
    // ---- velocity of plane m-1 (OK4.cu:351-378, ANISO.cu:374-386) ----------
    const int p = m - 1;
    if (p >= k0) {
      float wzp = 1.f, wzm = 1.f;
      if (PML == PML_SEP) wzp = wzq;
      if (PML == PML_PROFILE) wzm = p >= 1 ? wzq2 : 1.f;
      const int q0 = rq0, q1 = rq1;
#pragma unroll
      for (int r = 1; r <= R; ++r) {
        const int j = jr[r];
        const bool jm = j >= 1;
        const bool xin = im && jm && p >= 1;
        const bool ux = xin && j <= ny - 2 && p <= nz - 2;
        const bool uy = xin && i <= nx - 2 && p <= nz - 2;
        const bool uz = xin && i <= nx - 2 && j <= ny - 2;
        const float bx = bxq[r], by = byq[r], bz = bzq[r];
        float w = 1.f;
        if (PML == PML_W3D) w = wq[r];
        if (PML == PML_SEP) w = wxi * wyj[r] * wzp;
        const float sxx_ = qxx[r], sxx_im = from_left(qxx[r]);
        const float syy_ = qyy[r], syy_jm = qyy[r - 1];
        const float szz_ = qzz[r], szz_km = qzz2[r];
        const float sxy_ = qxy[r], sxy_jp = qxy[r + 1], sxy_ip = from_right(qxy[r]);
        const float sxz_ = qxz[r], sxz_kp = sxz[r], sxz_ip = from_right(qxz[r]);
        const float syz_ = qyz[r], syz_kp = syz[r], syz_jp = qyz[r + 1];
        const float dSx = rate3(sxx_ - (im ? sxx_im : 0.f), a.rdx, sxy_jp - sxy_, a.rdy, sxz_kp - sxz_, a.rdz);
        const float dSy = rate3(syy_ - (jm ? syy_jm : 0.f), a.rdy, sxy_ip - sxy_, a.rdx, syz_kp - syz_, a.rdz);
        const float dSz = rate3(szz_ - (p >= 1 ? szz_km : 0.f), a.rdz, sxz_ip - sxz_, a.rdx, syz_jp - syz_, a.rdy);
        float vx = vxp[r], vy = vyp[r], vz = vzp[r];
        if (ux) vx = ISO ? iso_vel(wxm, vx, bx, dSx) : ortho_vel(w, vx, bx, dSx);
        if (uy) vy = ISO ? iso_vel(wym[r], vy, by, dSy) : ortho_vel(w, vy, by, dSy);
        if (uz) vz = ISO ? iso_vel(wzm, vz, bz, dSz) : ortho_vel(w, vz, bz, dSz);
        pvxs[r] = vx; pvys[r] = vy; pvzs[r] = vz;  // stored at the next iteration
        if (ox && orow[r]) {
          for (int e = q0; e < q1; ++e) {  // this tile's receivers in plane p
            const PointEntry rc = a.rec[e];
            if (rc.i == i && rc.j == j && rc.batch == wf && itn < a.nt)
              a.traces[(int64_t)rc.slot * a.nt + itn] = rc.comp == 0 ? vx : (rc.comp == 1 ? vy : vz);
          }
        }
      }
    }
    // ---- advance the carried planes -------------------------------------
#pragma unroll
    for (int r = 0; r < NR; ++r) {
      qzz2[r] = qzz[r];
      qxx[r] = sxx[r]; qyy[r] = syy[r]; qzz[r] = szz[r]; qxy[r] = sxy[r]; qxz[r] = sxz[r]; qyz[r] = syz[r];
      vxp[r] = vxc[r]; vyp[r] = vyc[r]; vzp[r] = vzc[r];
      vxc[r] = vxn[r]; vyc[r] = vyn[r]; vzc[r] = vzn[r];
      wq[r] = w3[r];
      idq[r] = idc[r];
    }
    wzq2 = wzq;
    wzq = wzc;
  }
  // the last velocity plane of the chunk, computed in the last iteration
  if (k1 - 1 >= k0) {
    const uint32_t vb = (uint32_t)(k1 - 1) * bsz;
#pragma unroll
    for (int r = 1; r <= R; ++r) {
      if (!(ox && orow[r])) continue;
      const uint32_t b = rb[r] + vb;
      stb(Vxo, b, pvxs[r]); stb(Vyo, b, pvys[r]); stb(Vzo, b, pvzs[r]);
    }
  }
}

template <int MODE, int PML, bool FS>
static hipError_t go_sweep_fs(const SweepArgs& a, hipStream_t s) {
  // eight equal runs of workgroups, run x on XCD x (see k_sweep)
  const int64_t waves = (int64_t)a.ntx * a.nyb * a.nkc * a.nbatch;
  const int64_t blocks = 8 * ((waves + SWEEP_WPG - 1) / SWEEP_WPG);
  if (blocks <= 0 || blocks > 0x7fffffff) return hipErrorInvalidValue;
  const dim3 g((unsigned)blocks), b(64, SWEEP_WPG);
  if (a.rows == 2) {
    hipLaunchKernelGGL((k_sweep<MODE, PML, FS, 2>), g, b, 0, s, a);
  } else if (a.rows == 4) {
    hipLaunchKernelGGL((k_sweep<MODE, PML, FS, 4>), g, b, 0, s, a);
  } else {
    return hipErrorInvalidValue;
  }
  return hipGetLastError();
}

template <int MODE, int PML>
hipError_t sweep_mp(const SweepArgs& a, int fs, hipStream_t s) {
  return fs ? go_sweep_fs<MODE, PML, true>(a, s) : go_sweep_fs<MODE, PML, false>(a, s);
}

}  // namespace fwe
