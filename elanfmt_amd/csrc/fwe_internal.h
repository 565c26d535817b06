// fwe_internal.h -- shared definitions between the C-ABI host code and the
// gfx950 kernels of libfwe.so.  Not part of the public ABI.
#pragma once
#include <stdint.h>

#include <hip/hip_runtime.h>

#include "../../include/fwe.h"

namespace fwe {

// Columns of a single-pass (sweep) tile: a wavefront's 64 lanes cover the
// tile plus one halo column on each side.
constexpr int SWEEP_X = 62;

// MODE_ORTHO_TAB: the ortho law with the medium read through a material
// table (one byte per cell, see StepArgs::mid).  MODE_ORTHO_ISO: the ortho
// law on an isotropic medium (C11 = C22 = C33 and C12 = C13 = C23 in every
// cell), which reads C[0] and C[3] for the six normal moduli.  MODE_ISO_TAB:
// the iso law with c11, c12 read through a material table (two floats per
// entry).
enum Mode { MODE_ISO = 0, MODE_ORTHO = 1, MODE_ORTHO_TAB = 2, MODE_ORTHO_ISO = 3, MODE_ISO_TAB = 4 };
// the iso law's update formulas (OK4.cu) rather than the ortho law's
__host__ __device__ constexpr bool iso_law(int mode) { return mode == MODE_ISO || mode == MODE_ISO_TAB; }
// PML_CPML: convolutional PML (memory variables in slabs along each face;
// StepArgs::cL and below), the two-pass kernels only
enum Pml { PML_NONE = 0, PML_PROFILE = 1, PML_W3D = 2, PML_SEP = 3, PML_CPML = 4 };

// One injection or receiver, sorted by xy tile on the host.
struct PointEntry {
  int32_t i, j, k;
  int32_t comp;
  int32_t batch;
  float coef;  // injection coefficient (unused for receivers)
  int32_t slot;  // receiver index (trace row) or -1
  int32_t pad;
};

// Everything one step needs.  Fields of wavefield b live at base + b*fstride.
// All 3-D arrays share one padded layout: element (i,j,k) at
// i + j*sy + k*sz with sy >= nx+1 (rounded to 32 floats) and sz = sy*(ny+1).
struct StepArgs {
  int32_t nx, ny, nz;
  // wavefields: each workgroup advances nb of them together (the template
  // NB: one medium load per cell for all nb); a launch covers nbg such
  // groups side by side (nb * nbg wavefields; the group kernels use 1, 1)
  int32_t nb, nbg;
  int64_t sy, sz, fstride;
  float* V[3];        // Vx Vy Vz (batch 0)
  float* S[6];        // Sxx Syy Szz Sxy Sxz Syz (batch 0)
  // ortho medium, pre-multiplied by dt: C11 C22 C33 C12 C13 C23 C44yz C55xz C66xy
  const float* C[9];
  // ortho buoyancy pre-multiplied by dt: bx by bz (staggered)
  const float* B[3];
  // the same values as a table: cell c uses entry mid[c], twelve floats at
  // mtab + 12*mid[c] in the order C[0..8], B[0..2] (MODE_ORTHO_TAB), or two
  // at mtab + 2*mid[c], c11 and c12 (MODE_ISO_TAB)
  const uint8_t* mid;
  const float* mtab;
  // iso medium: c11*dt, c12*dt arrays, scalars c66*dt and b*dt
  float c66dt, bxdt, bydt, bzdt;
  // vacuum-layer top: on plane vac_k every shear modulus*dt is vac_cdt
  // (vac_k = -1: off)
  int32_t vac_k;
  float vac_cdt;
  // damping: PROFILE -> w (n+1) and mean-of-neighbours wh (n) per axis;
  // SEP -> tapers (n); W3D -> w3 in the padded layout.
  const float* wx; const float* wy; const float* wz;
  const float* whx; const float* why; const float* whz;
  const float* w3;
  // convolutional PML (PML_CPML): every derivative across a face's slab
  // (the cL cells at either end of an axis) becomes g/kappa + psi with
  // psi <- b psi + a g.  ctab[d]: 6 rows of n_d+1 floats (b, a, 1/kappa at
  // integer nodes, then at half nodes) for axis d.  psi[d]: the memory
  // variables of axis d's two slabs for wavefield 0, six variables of
  // pvar[d] elements each (0..2 read by the stress pass, 3..5 by the
  // velocity pass); slab element of cell (i,j,k): x: xi + 2cL (j + ny k),
  // y: i + nx (yj + 2cL k), z: i + nx (j + ny zk), where xi = i (i < cL) or
  // i - nx + 2cL (i >= nx - cL), likewise yj, zk.  Wavefield w at + w*pstride.
  int32_t cL;
  const float* ctab[3];
  float* psi[3];
  int64_t pvar[3];
  int64_t pstride;
  float rdx, rdy, rdz;   // 1/dx ...
  float dtdx, dtdy, dtdz; // dt/dx ... (iso stress)
  // tiling
  int32_t tile_x;       // columns per tile (64 or 128)
  int32_t xcd_order;    // workgroup order: 0 linear, 1 grouped per XCD (xcd_remap),
                        // 2 plane split (plane_split)
  int32_t split;        // plane split: tiles per XCD part of a chunk, ceil(ntx*nty/8)
  int32_t ty_rows;      // TILE_Y
  int32_t ntx, nty, nkc, kchunk;
  // the tiles one launch covers (tile = tx + ty*ntx + kc*ntx*nty): a box of
  // bnx x bny x bnk tiles from (bx0, by0, bk0), the whole grid by default,
  // or, when `tiles` is set, the ntiles tiles listed there.  The
  // convolutional PML runs the tiles that touch no slab as a box through
  // the plain kernel and lists the others for its own kernel.  Plane-split
  // order always covers the whole grid.
  int32_t bx0, bnx, by0, bny, bk0, bnk;
  const int32_t* tiles;
  int32_t ntiles;
  // sources
  const int32_t* inj_off;   // ntx*nty+1
  const PointEntry* inj;
  const float* stf;         // [nb][nstf]
  int32_t nstf;
  int32_t src_after_bc;
  // receivers, sorted by (xy tile, plane): rec_off per tile, rec_koff per
  // (tile, plane) pair, index tile*nz + k (the two-pass velocity kernel
  // scans only the entries of its own plane)
  const int32_t* rec_off;
  const int32_t* rec_koff;
  const PointEntry* rec;
  float* traces;            // [nrec][nt]
  int32_t nt;
  // step index = it + (it_base ? *it_base : 0); a captured graph of G steps
  // bakes it = 0..G-1 and advances *it_base by G after each replay
  int32_t it;
  const int32_t* it_base;
};

// Single-pass step of the sweep kernel (fwe_sweep.h): every wavefront owns
// a tile of 62 columns x `rows` rows x `kchunk` planes of one wavefield and
// reads time level n (Vi, Si), writes n+1 (Vo, So).  The medium fields are
// those of StepArgs.  Offsets inside one array are 32-bit byte offsets.
struct SweepArgs {
  int32_t nx, ny, nz;
  int32_t sy, sz;           // padded pitches in elements (sz*(nz+1)*4 < 4 GiB)
  int64_t fstride;          // elements between wavefields
  int32_t nbatch;
  const float* Vi[3];
  const float* Si[6];
  float* Vo[3];
  float* So[6];
  const float* C[9];
  const float* B[3];
  const uint8_t* mid;
  const float* mtab;
  int32_t ntab;             // table entries (copied to LDS by each workgroup)
  float c66dt, bxdt, bydt, bzdt;
  int32_t vac_k;
  float vac_cdt;
  const float* wx; const float* wy; const float* wz;
  const float* whx; const float* why; const float* whz;
  const float* w3;
  float rdx, rdy, rdz;
  int32_t rows;             // output rows per lane (2 or 4)
  int32_t ntx, nty, nkc, kchunk;
  int32_t nyb;              // tile rows per XCD band, ceil(nty / 8)
  // injections: per xy tile, every entry whose cell lies in the tile grown
  // by one cell in x and y (the halo stresses include them)
  const int32_t* inj_off;
  const PointEntry* inj;
  const float* stf;         // [nbatch][nstf]
  int32_t nstf;
  int32_t src_after_bc;
  // receivers per (xy tile, plane): rec_koff[tile*nz + k]
  const int32_t* rec_koff;
  const PointEntry* rec;
  float* traces;
  int32_t nt;
  int32_t it;
  const int32_t* it_base;
};

template <int MODE, int PML>
hipError_t sweep_mp(const SweepArgs& a, int fs, hipStream_t s);
hipError_t launch_sweep(const SweepArgs& a, int mode, int pml, int fs, hipStream_t s);

// per (law, damping form) launchers, instantiated in fwe_step.hip
template <int MODE, int PML>
hipError_t stress_mp(const StepArgs& a, int fs, hipStream_t s);
template <int MODE, int PML>
hipError_t velocity_mp(const StepArgs& a, hipStream_t s);
template <int MODE, int PML>
hipError_t group_mp(const StepArgs* tab, const int32_t* wg, int nblk, int tx, int ty, int order, int fs, bool vel,
                    hipStream_t s);

hipError_t launch_stress(const StepArgs& a, int mode, int pml, int fs, hipStream_t s);
hipError_t launch_velocity(const StepArgs& a, int mode, int pml, hipStream_t s);
// one stress (vel=false) or velocity (vel=true) launch over the tiles of a
// group of single-wavefield models: nblk workgroups, tab (one entry per
// model) and wg (per workgroup: model, workgroup within it) in device memory
hipError_t launch_group(const StepArgs* tab, const int32_t* wg, int nblk, int tx, int ty, int order, int mode, int pml,
                        int fs, bool vel, hipStream_t s);
// *it_base = g (set) or *it_base += g
hipError_t launch_advance(int32_t* it_base, int32_t g, int32_t set, hipStream_t s);
hipError_t launch_copy(const float* src, float* dst, int64_t n, hipStream_t s);
// out[m][q] = sum_e coef[m][e] * tr[e][q] for m < mc, q < per; all device
hipError_t launch_combine(const float* tr, int nb, int64_t per, const float* coef, int mc, float* out,
                          hipStream_t s);

// ---------------------------------------------------------------------------
// learned thin boundary (fwe_lb.hip)
// ---------------------------------------------------------------------------
constexpr int LB_IN = 81, LB_H = 64, LB_OUT = 9, LB_MAXL = 16;
// parameters of one face, FP32: W1 [81][64], b1 [L][64], W2 [64][64], b2 [64], W3 [64][9], b3 [9]
__host__ __device__ constexpr int lb_off_w1() { return 0; }
__host__ __device__ constexpr int lb_off_b1() { return LB_IN * LB_H; }
__host__ __device__ constexpr int lb_off_w2(int L) { return LB_IN * LB_H + L * LB_H; }
__host__ __device__ constexpr int lb_off_b2(int L) { return lb_off_w2(L) + LB_H * LB_H; }
__host__ __device__ constexpr int lb_off_w3(int L) { return lb_off_b2(L) + LB_H; }
__host__ __device__ constexpr int lb_off_b3(int L) { return lb_off_w3(L) + LB_H * LB_OUT; }
__host__ __device__ constexpr int lb_nparams(int L) { return lb_off_b3(L) + LB_OUT; }

// a cell of the layer: fd = face << 8 | depth (0 = outermost); faces
// 0..5 = x-, x+, y-, y+, z-, z+
struct LbCell {
  int32_t i, j, k, fd;
};

// one wavefield's field set in the padded layout
struct LbGeom {
  const float* F[9];  // field f at F[0] + f * fstride
  int64_t fstride;
  int64_t sy, sz;
  int32_t nx, ny, nz;
  float zinv;
};

hipError_t lb_launch_mlp(const float* P, int L, int n, const float* X, const int32_t* depth, const float* T,
                         const float* M, float inv_count, float* Y, float* H1, float* H2, float* D1, float* D2,
                         float* G, float* loss_part, int nblk, hipStream_t s);
// out[p][q] = sum_s a(s,p) B[s][q]; a = A[s][p], onehot(depth[s]) or 1;
// part holds ceil(n/chunk) * P * Q floats
hipError_t lb_launch_atb(const float* A, int lda, const int32_t* onehot, int P, const float* B, int ldb, int Q,
                         int n, int chunk, float* part, float* out, hipStream_t s);
hipError_t lb_launch_adam(float* P, float* m, float* v, const float* g, int n, float lr, float b1, float b2,
                          float eps, float c1, float c2, hipStream_t s);
hipError_t lb_launch_gather(const LbGeom& g, const LbCell* cells, int n, float* X, float* R, int32_t* depth,
                            hipStream_t s);
hipError_t lb_launch_targets(const LbGeom& thin, const LbGeom& thick, int off, const LbCell* cells, int n,
                             const float* R, float rmin, float wscale, float* T, float* M, hipStream_t s);

// Inference weights of one face in the operand order of k_lb_infer,
// rewritten by lb_launch_pack whenever the FP32 parameters change: bfloat16
// W1 [4 tiles][3 k-steps][64 lanes][8], W2 [4][2][64][8], W3 [2][64][8],
// then FP32 b1 [LB_MAXL][64] (rows >= L zero), b2 [64], b3 [16].
constexpr int LB_PK_W1 = 0;
constexpr int LB_PK_W2 = LB_PK_W1 + 4 * 3 * 64 * 8 * 2;
constexpr int LB_PK_W3 = LB_PK_W2 + 4 * 2 * 64 * 8 * 2;
constexpr int LB_PK_B1 = LB_PK_W3 + 2 * 64 * 8 * 2;
constexpr int LB_PK_B2 = LB_PK_B1 + LB_MAXL * LB_H * 4;
constexpr int LB_PK_B3 = LB_PK_B2 + LB_H * 4;
constexpr int LB_PK_BYTES = LB_PK_B3 + 16 * 4;

// One inference launch: every layer cell of every wavefield.
struct LbInfer {
  const uint8_t* pk;      // 6 faces x LB_PK_BYTES
  const LbCell* cells;    // layer cells grouped by face
  int32_t face_off[7];    // cells of face f: [face_off[f], face_off[f+1])
  int32_t blk_off[7];     // workgroups of face f (set by lb_launch_infer)
  float* F;               // wavefield 0: field q at F + q * fstride
  int64_t wstride;        // elements between wavefields
  uint32_t fstride;       // elements between the fields of a wavefield
  int32_t sy, sz, nx, ny, nz;
  int32_t nbatch;
  float zinv;
  float* D;               // corrections [nbatch][9][face_off[6]]
  int32_t free_form;      // 1: fields + D; 0: their dissipative part
  float* trash;           // 64 floats, written by lanes that have no output
};
hipError_t lb_launch_pack(const float* P, int np, int L, uint8_t* pk, hipStream_t s);
// corrections D from the fields before the step (inputs and hidden
// activations rounded to bfloat16, sums in FP32)
hipError_t lb_launch_infer(LbInfer& a, hipStream_t s);
// fields <- shrink(fields, D) after the step
hipError_t lb_launch_add(const LbInfer& a, hipStream_t s);
hipError_t lb_launch_apply_y(const LbGeom& g, const LbCell* cells, int n, const float* Y, const float* R,
                             int free_form, hipStream_t s);
hipError_t lb_launch_crop(const LbGeom& thick, const LbGeom& thin, int off, hipStream_t s);

// the convolutional PML memory variables of one wavefield (StepArgs::psi
// layout: x, y, z slabs of six variables, pvar[d] elements each)
struct PsiGeom {
  float* psi;
  int32_t nx, ny, nz, L;
  int64_t pvar[3];
};
// thin slabs <- the same cells of the thick run's (wider) slabs, which must
// hold them: thick L >= thin L + off
hipError_t lb_launch_crop_psi(const PsiGeom& thick, const PsiGeom& thin, int off, hipStream_t s);

}  // namespace fwe
