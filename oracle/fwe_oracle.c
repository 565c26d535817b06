/*
 * fwe_oracle.c -- CPU restatement of the reference FDTD step (test oracle).
 *
 * TEST INFRASTRUCTURE ONLY: used by tests/, __graft_entry__.smoke() and the
 * cpu_baseline leg of bench.py as the checker.  Never part of the product.
 *
 * Built twice by oracle/Makefile: liboracle.so (REAL = double, the parity
 * oracle) and liboracle32.so (REAL = float, an FP32 CPU baseline).
 * Citations: OK4.cu / ANISO.cu as in fwe_oracle.h.
 */
#include "fwe_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#ifdef ORACLE_FP32
typedef float REAL;
#else
typedef double REAL;
#endif

enum { F_VX, F_VY, F_VZ, F_SXX, F_SYY, F_SZZ, F_SXY, F_SXZ, F_SYZ, F_N };
enum { C_XX, C_YY, C_ZZ, C_XY, C_XZ, C_YZ };

struct oracle {
  int mode, nx, ny, nz, pml, fs, src_after_bc, nt, it;
  REAL dx, dy, dz, dt;
  /* fields, reference shapes */
  REAL* f[F_N];
  int64_t fdim[F_N][3];
  /* iso medium (OK4.cu:415-418, pa2[7], pa3[0..2]) */
  REAL *c11, *c12, c66, bx, by, bz;
  /* ortho medium (ANISO.cu:411-454) */
  REAL *C[9], *B[3];
  /* damping */
  REAL *wx, *wy, *wz, *w3;
  /* source */
  REAL* stf;
  int nstf, ninj;
  int32_t *inj_comp, *inj_ijk;
  REAL* inj_coef;
  /* receivers */
  int nrec;
  int32_t *rec_comp, *rec_ijk;
  double* traces;
  /* convolutional PML (pml = 4): layer cells per face, coefficient tables
   * per axis (6 rows of n+1: b, a, 1/kappa at integer then half nodes) and
   * 18 memory variables on the cell grid: axis d owns psi[6d .. 6d+5]
   * (0..2 stress pass, 3..5 velocity pass) */
  int cL;
  float* ct[3];
  REAL* psi[18];
};

/* Reference flat indices (OK4.cu:290-292,323-329,351-353). */
#define IC(i, j, k) ((int64_t)(i) + (int64_t)(j) * nx + (int64_t)(k) * nx * ny)
#define IVX(i, j, k) ((int64_t)(i) + (int64_t)(j) * (nx + 1) + (int64_t)(k) * (nx + 1) * ny)
#define IVY(i, j, k) ((int64_t)(i) + (int64_t)(j) * nx + (int64_t)(k) * nx * (ny + 1))
#define IVZ(i, j, k) IC(i, j, k)
#define IXY(i, j, k) ((int64_t)(i) + (int64_t)(j) * (nx + 1) + (int64_t)(k) * (nx + 1) * (ny + 1))
#define IXZ(i, j, k) IVX(i, j, k)
#define IYZ(i, j, k) IVY(i, j, k)

static void* zalloc(int64_t n) { return calloc((size_t)(n > 0 ? n : 1), sizeof(REAL)); }

oracle_t* oracle_create(int mode, int nx, int ny, int nz, double dx, double dy, double dz,
                        double dt, int pml, int free_surface, int src_after_bc, int nt) {
  if (nx < 3 || ny < 3 || nz < 3 || nt < 0) return NULL;
  oracle_t* o = (oracle_t*)calloc(1, sizeof(oracle_t));
  if (!o) return NULL;
  o->mode = mode; o->nx = nx; o->ny = ny; o->nz = nz;
  o->dx = (REAL)dx; o->dy = (REAL)dy; o->dz = (REAL)dz; o->dt = (REAL)dt;
  o->pml = pml; o->fs = free_surface; o->src_after_bc = src_after_bc; o->nt = nt;
  const int64_t d[F_N][3] = {
      {nx + 1, ny, nz}, {nx, ny + 1, nz}, {nx, ny, nz + 1},
      {nx, ny, nz},     {nx, ny, nz},     {nx, ny, nz},
      {nx + 1, ny + 1, nz}, {nx + 1, ny, nz + 1}, {nx, ny + 1, nz + 1}};
  for (int f = 0; f < F_N; ++f) {
    memcpy(o->fdim[f], d[f], sizeof(d[f]));
    o->f[f] = (REAL*)zalloc(d[f][0] * d[f][1] * d[f][2]);
  }
  o->traces = NULL;
  return o;
}

void oracle_destroy(oracle_t* o) {
  if (!o) return;
  for (int f = 0; f < F_N; ++f) free(o->f[f]);
  free(o->c11); free(o->c12);
  for (int c = 0; c < 9; ++c) free(o->C[c]);
  for (int c = 0; c < 3; ++c) free(o->B[c]);
  free(o->wx); free(o->wy); free(o->wz); free(o->w3);
  free(o->stf); free(o->inj_comp); free(o->inj_ijk); free(o->inj_coef);
  free(o->rec_comp); free(o->rec_ijk); free(o->traces);
  for (int d = 0; d < 3; ++d) free(o->ct[d]);
  for (int v = 0; v < 18; ++v) free(o->psi[v]);
  free(o);
}

static REAL* copy_in(const float* src, int64_t n) {
  REAL* a = (REAL*)zalloc(n);
  if (a && src)
    for (int64_t i = 0; i < n; ++i) a[i] = (REAL)src[i];
  return a;
}

int64_t oracle_field_size(const oracle_t* o, int f) {
  if (f < 0 || f >= F_N) return -1;
  return o->fdim[f][0] * o->fdim[f][1] * o->fdim[f][2];
}

int oracle_set_iso(oracle_t* o, const float* c11, const float* c12, double c66, double bx,
                   double by, double bz) {
  const int64_t n = (int64_t)o->nx * o->ny * o->nz;
  free(o->c11); free(o->c12);
  o->c11 = copy_in(c11, n); o->c12 = copy_in(c12, n);
  o->c66 = (REAL)c66; o->bx = (REAL)bx; o->by = (REAL)by; o->bz = (REAL)bz;
  return 0;
}

int oracle_set_ortho(oracle_t* o, const float* const cij[9], const float* const b[3]) {
  const int nx = o->nx, ny = o->ny, nz = o->nz;
  const int64_t nc = (int64_t)nx * ny * nz;
  const int64_t sz[9] = {nc, nc, nc, nc, nc, nc, (int64_t)nx * (ny + 1) * (nz + 1),
                         (int64_t)(nx + 1) * ny * (nz + 1), (int64_t)(nx + 1) * (ny + 1) * nz};
  const int64_t bs[3] = {(int64_t)(nx + 1) * ny * nz, (int64_t)nx * (ny + 1) * nz,
                         (int64_t)nx * ny * (nz + 1)};
  for (int c = 0; c < 9; ++c) { free(o->C[c]); o->C[c] = copy_in(cij[c], sz[c]); }
  for (int c = 0; c < 3; ++c) { free(o->B[c]); o->B[c] = copy_in(b[c], bs[c]); }
  return 0;
}

int oracle_set_profiles(oracle_t* o, const float* wx, const float* wy, const float* wz) {
  const int add = (o->pml == 1) ? 1 : 0; /* PROFILE: n+1 (OK4.cu:419-424); SEPARABLE: n */
  free(o->wx); free(o->wy); free(o->wz);
  o->wx = copy_in(wx, o->nx + add);
  o->wy = copy_in(wy, o->ny + add);
  o->wz = copy_in(wz, o->nz + add);
  return 0;
}

int oracle_set_weights3d(oracle_t* o, const float* w) {
  free(o->w3);
  o->w3 = copy_in(w, (int64_t)o->nx * o->ny * o->nz);
  return 0;
}

/* Convolutional PML of the HIP library (its own addition: the reference
 * damps with multiplicative profiles only).  The tables are the float
 * coefficients the device receives, so both sides use the same numbers. */
int oracle_set_cpml(oracle_t* o, int L, const float* tx, const float* ty, const float* tz) {
  const int n[3] = {o->nx, o->ny, o->nz};
  const float* t[3] = {tx, ty, tz};
  if (o->pml != 4 || L < 1) return -1;
  for (int d = 0; d < 3; ++d)
    if (2 * L > n[d] || !t[d]) return -1;
  o->cL = L;
  for (int d = 0; d < 3; ++d) {
    free(o->ct[d]);
    o->ct[d] = (float*)malloc(sizeof(float) * 6 * (size_t)(n[d] + 1));
    if (!o->ct[d]) return -1;
    memcpy(o->ct[d], t[d], sizeof(float) * 6 * (size_t)(n[d] + 1));
  }
  for (int v = 0; v < 18; ++v) {
    free(o->psi[v]);
    o->psi[v] = (REAL*)zalloc((int64_t)o->nx * o->ny * o->nz);
    if (!o->psi[v]) return -1;
  }
  return 0;
}

/* Derivative g along axis d at a node with axis index q (h = 0 integer
 * node, 1 half node) of cell c, through memory variable v of that axis:
 * inside the axis' slab (the cL cells at either end) psi <- b psi + a g and
 * the result is g/kappa + psi; elsewhere g. */
static inline REAL cpml(oracle_t* o, int d, int v, int h, int q, int64_t c, REAL g) {
  const int n = d == 0 ? o->nx : (d == 1 ? o->ny : o->nz);
  if (o->pml != 4 || (q >= o->cL && q < n - o->cL)) return g;
  const int64_t m = (int64_t)n + 1;
  const float* t = o->ct[d] + 3 * h * m;
  REAL* p = o->psi[6 * d + v] + c;
  *p = (REAL)t[q] * *p + (REAL)t[m + q] * g;
  return g * (REAL)t[2 * m + q] + *p;
}

int oracle_set_stf(oracle_t* o, const float* stf, int n) {
  free(o->stf);
  o->stf = copy_in(stf, n);
  o->nstf = n;
  return 0;
}

int oracle_set_injections(oracle_t* o, int n, const int32_t* comp, const int32_t* ijk,
                          const float* coef) {
  free(o->inj_comp); free(o->inj_ijk); free(o->inj_coef);
  o->ninj = n;
  o->inj_comp = (int32_t*)malloc(sizeof(int32_t) * (n > 0 ? n : 1));
  o->inj_ijk = (int32_t*)malloc(sizeof(int32_t) * 3 * (n > 0 ? n : 1));
  o->inj_coef = (REAL*)zalloc(n);
  for (int s = 0; s < n; ++s) {
    o->inj_comp[s] = comp[s];
    memcpy(&o->inj_ijk[3 * s], &ijk[3 * s], 3 * sizeof(int32_t));
    o->inj_coef[s] = (REAL)coef[s];
  }
  return 0;
}

int oracle_set_receivers(oracle_t* o, int n, const int32_t* comp, const int32_t* ijk) {
  free(o->rec_comp); free(o->rec_ijk); free(o->traces);
  o->nrec = n;
  o->rec_comp = (int32_t*)malloc(sizeof(int32_t) * (n > 0 ? n : 1));
  o->rec_ijk = (int32_t*)malloc(sizeof(int32_t) * 3 * (n > 0 ? n : 1));
  memcpy(o->rec_comp, comp, sizeof(int32_t) * n);
  memcpy(o->rec_ijk, ijk, sizeof(int32_t) * 3 * n);
  o->traces = (double*)calloc((size_t)(n > 0 ? n : 1) * (o->nt > 0 ? o->nt : 1), sizeof(double));
  return 0;
}

void oracle_reset(oracle_t* o) {
  for (int f = 0; f < F_N; ++f)
    memset(o->f[f], 0, sizeof(REAL) * (size_t)oracle_field_size(o, f));
  for (int v = 0; v < 18; ++v)
    if (o->psi[v]) memset(o->psi[v], 0, sizeof(REAL) * (size_t)o->nx * o->ny * o->nz);
  if (o->traces) memset(o->traces, 0, sizeof(double) * (size_t)o->nrec * o->nt);
  o->it = 0;
}

/* Damping factor lookups.  PROFILE (OK4): normal stress uses the mean of
 * the two profile entries bracketing the cell (OK4.cu:307-309), shear stress
 * uses the product wx[i]*wy[j]*wz[k] in the reference's order
 * (OK4.cu:323,326,329), velocity uses w[i-1] along its own axis
 * (OK4.cu:361,368,375).  WEIGHTS3D/SEPARABLE (ANISO): the cell-centred
 * weight w(i,j,k) multiplies the old value (ANISO.cu:306-352,374-386). */
static inline REAL w3(const oracle_t* o, int i, int j, int k) {
  const int nx = o->nx, ny = o->ny;
  if (o->pml == 2) return o->w3[IC(i, j, k)];
  if (o->pml == 3) return o->wx[i] * o->wy[j] * o->wz[k];
  return (REAL)1;
}

/* Vacuum-layer top (free_surface = 2), the BC2 variant of the iso kernel
 * (diff_BC2_OK3/FastWaveED_GPU3D_v1.cu:309-314): every shear stress on the
 * plane k = 1 uses the modulus 1e-5 instead of the medium's, and BC1 is not
 * applied.  The caller empties the top cell layer (zero moduli at k = 0), as
 * that variant's driver does (vizme_FastBiotED_GPU3D_OK_2024.m:238,243). */
static inline REAL shear_modulus(const oracle_t* o, REAL c, int k) {
  return (o->fs == 2 && k == 1) ? (REAL)1e-5 : c;
}

/* Stress pass, iso law: OK4.cu:285-330 (compute_StressPrf).  Shear cells at
 * i=0, j=0 or k=0 read outside their neighbourhood in the reference and are
 * never consumed by the velocity pass (OK4.cu:351-378); they stay 0 here. */
static void stress_iso(oracle_t* o) {
  const int nx = o->nx, ny = o->ny, nz = o->nz;
  REAL *Vx = o->f[F_VX], *Vy = o->f[F_VY], *Vz = o->f[F_VZ];
  REAL *sxx = o->f[F_SXX], *syy = o->f[F_SYY], *szz = o->f[F_SZZ];
  REAL *sxy = o->f[F_SXY], *sxz = o->f[F_SXZ], *syz = o->f[F_SYZ];
  const REAL dt = o->dt, rdx = (REAL)1 / o->dx, rdy = (REAL)1 / o->dy, rdz = (REAL)1 / o->dz;
  const int prof = (o->pml == 1);
  const REAL *wx = o->wx, *wy = o->wy, *wz = o->wz;
#pragma omp parallel for schedule(static)
  for (int k = 0; k < nz; ++k)
    for (int j = 0; j < ny; ++j)
      for (int i = 0; i < nx; ++i) {
        const int64_t c = IC(i, j, k);
        const REAL c66 = shear_modulus(o, o->c66, k);
        const REAL exx = cpml(o, 0, 0, 0, i, c, (Vx[IVX(i + 1, j, k)] - Vx[IVX(i, j, k)]) * rdx);
        const REAL eyy = cpml(o, 1, 0, 0, j, c, (Vy[IVY(i, j + 1, k)] - Vy[IVY(i, j, k)]) * rdy);
        const REAL ezz = cpml(o, 2, 0, 0, k, c, (Vz[IVZ(i, j, k + 1)] - Vz[IVZ(i, j, k)]) * rdz);
        const REAL a = o->c11[c], b = o->c12[c];
        const REAL fx = prof ? (REAL)0.5 * (wx[i + 1] + wx[i]) : (REAL)1;
        const REAL fy = prof ? (REAL)0.5 * (wy[j + 1] + wy[j]) : (REAL)1;
        const REAL fz = prof ? (REAL)0.5 * (wz[k + 1] + wz[k]) : (REAL)1;
        sxx[c] = fx * (sxx[c] + dt * (a * exx + b * eyy + b * ezz));
        syy[c] = fy * (syy[c] + dt * (b * exx + a * eyy + b * ezz));
        szz[c] = fz * (szz[c] + dt * (b * exx + b * eyy + a * ezz));
        if (i >= 1 && j >= 1) {
          const int64_t e = IXY(i, j, k);
          const REAL w = prof ? wz[k] * wy[j] * wx[i] : (REAL)1;
          sxy[e] = w * (sxy[e] + c66 * dt * (cpml(o, 0, 1, 1, i, c, (Vy[IVY(i, j, k)] - Vy[IVY(i - 1, j, k)]) * rdx) +
                                             cpml(o, 1, 1, 1, j, c, (Vx[IVX(i, j, k)] - Vx[IVX(i, j - 1, k)]) * rdy)));
        }
        if (i >= 1 && k >= 1) {
          const int64_t e = IXZ(i, j, k);
          const REAL w = prof ? wy[j] * wz[k] * wx[i] : (REAL)1;
          sxz[e] = w * (sxz[e] + c66 * dt * (cpml(o, 0, 2, 1, i, c, (Vz[IVZ(i, j, k)] - Vz[IVZ(i - 1, j, k)]) * rdx) +
                                             cpml(o, 2, 1, 1, k, c, (Vx[IVX(i, j, k)] - Vx[IVX(i, j, k - 1)]) * rdz)));
        }
        if (j >= 1 && k >= 1) {
          const int64_t e = IYZ(i, j, k);
          const REAL w = prof ? wx[i] * wy[j] * wz[k] : (REAL)1;
          syz[e] = w * (syz[e] + c66 * dt * (cpml(o, 2, 2, 1, k, c, (Vy[IVY(i, j, k)] - Vy[IVY(i, j, k - 1)]) * rdz) +
                                             cpml(o, 1, 2, 1, j, c, (Vz[IVZ(i, j, k)] - Vz[IVZ(i, j - 1, k)]) * rdy)));
        }
      }
}

/* Stress pass, orthorhombic law: ANISO.cu:287-356 (compute_StressOrtho). */
static void stress_ortho(oracle_t* o) {
  const int nx = o->nx, ny = o->ny, nz = o->nz;
  REAL *Vx = o->f[F_VX], *Vy = o->f[F_VY], *Vz = o->f[F_VZ];
  REAL *sxx = o->f[F_SXX], *syy = o->f[F_SYY], *szz = o->f[F_SZZ];
  REAL *sxy = o->f[F_SXY], *sxz = o->f[F_SXZ], *syz = o->f[F_SYZ];
  const REAL dt = o->dt, dx = o->dx, dy = o->dy, dz = o->dz;
  const REAL rdx = (REAL)1 / dx, rdy = (REAL)1 / dy, rdz = (REAL)1 / dz;
  REAL *C11 = o->C[0], *C22 = o->C[1], *C33 = o->C[2], *C12 = o->C[3], *C13 = o->C[4],
       *C23 = o->C[5], *C44 = o->C[6], *C55 = o->C[7], *C66 = o->C[8];
#pragma omp parallel for schedule(static)
  for (int k = 0; k < nz; ++k)
    for (int j = 0; j < ny; ++j)
      for (int i = 0; i < nx; ++i) {
        const int64_t c = IC(i, j, k);
        const REAL w = w3(o, i, j, k);
        const REAL exx = cpml(o, 0, 0, 0, i, c, (Vx[IVX(i + 1, j, k)] - Vx[IVX(i, j, k)]) * rdx);
        const REAL eyy = cpml(o, 1, 0, 0, j, c, (Vy[IVY(i, j + 1, k)] - Vy[IVY(i, j, k)]) * rdy);
        const REAL ezz = cpml(o, 2, 0, 0, k, c, (Vz[IVZ(i, j, k + 1)] - Vz[IVZ(i, j, k)]) * rdz);
        sxx[c] = sxx[c] * w + dt * (C11[c] * exx + C12[c] * eyy + C13[c] * ezz);
        syy[c] = syy[c] * w + dt * (C12[c] * exx + C22[c] * eyy + C23[c] * ezz);
        szz[c] = szz[c] * w + dt * (C13[c] * exx + C23[c] * eyy + C33[c] * ezz);
        if (i >= 1 && j >= 1) { /* ANISO.cu:327-334 */
          const int64_t e = IXY(i, j, k);
          const REAL dVx_dy = cpml(o, 1, 1, 1, j, c, (Vx[IVX(i, j, k)] - Vx[IVX(i, j - 1, k)]) / dy);
          const REAL dVy_dx = cpml(o, 0, 1, 1, i, c, (Vy[IVY(i, j, k)] - Vy[IVY(i - 1, j, k)]) / dx);
          sxy[e] = sxy[e] * w + dt * shear_modulus(o, C66[e], k) * (dVy_dx + dVx_dy);
        }
        if (i >= 1 && k >= 1) { /* ANISO.cu:336-343 */
          const int64_t e = IXZ(i, j, k);
          const REAL dVx_dz = cpml(o, 2, 1, 1, k, c, (Vx[IVX(i, j, k)] - Vx[IVX(i, j, k - 1)]) / dz);
          const REAL dVz_dx = cpml(o, 0, 2, 1, i, c, (Vz[IVZ(i, j, k)] - Vz[IVZ(i - 1, j, k)]) / dx);
          sxz[e] = sxz[e] * w + dt * shear_modulus(o, C55[e], k) * (dVz_dx + dVx_dz);
        }
        if (j >= 1 && k >= 1) { /* ANISO.cu:345-352 */
          const int64_t e = IYZ(i, j, k);
          const REAL dVy_dz = cpml(o, 2, 2, 1, k, c, (Vy[IVY(i, j, k)] - Vy[IVY(i, j, k - 1)]) / dz);
          const REAL dVz_dy = cpml(o, 1, 2, 1, j, c, (Vz[IVZ(i, j, k)] - Vz[IVZ(i, j - 1, k)]) / dy);
          syz[e] = syz[e] * w + dt * shear_modulus(o, C44[e], k) * (dVy_dz + dVz_dy);
        }
      }
}

/* BC1 free-surface approximation: OK4.cu:271-283 (same body ANISO.cu:270-285).
 * The reference guards the y range with `iy<nx`; this restatement uses the
 * array's own extent (j<ny), identical whenever the launch has no threads
 * beyond ny or nx <= ny. */
static void bc1(oracle_t* o) {
  const int nx = o->nx, ny = o->ny;
  REAL *szz = o->f[F_SZZ], *sxz = o->f[F_SXZ], *syz = o->f[F_SYZ];
  const REAL third = (REAL)1 / (REAL)3;
  for (int j = 0; j < ny; ++j)
    for (int i = 0; i < nx; ++i) szz[IC(i, j, 0)] = (REAL)0;
  for (int j = 0; j < ny; ++j)
    for (int i = 1; i < nx; ++i) sxz[IXZ(i, j, 1)] = third * sxz[IXZ(i, j, 2)];
  for (int j = 1; j < ny; ++j)
    for (int i = 0; i < nx; ++i) syz[IYZ(i, j, 1)] = third * syz[IYZ(i, j, 2)];
}

/* Point injections: S_comp(cell) += coef*Src[it] (OK4.cu:336-342, ANISO.cu:228-252). */
static void inject(oracle_t* o) {
  if (!o->stf || o->it >= o->nstf) return;
  const REAL s = o->stf[o->it];
  const int nx = o->nx, ny = o->ny;
  for (int q = 0; q < o->ninj; ++q) {
    const int i = o->inj_ijk[3 * q], j = o->inj_ijk[3 * q + 1], k = o->inj_ijk[3 * q + 2];
    int64_t e;
    switch (o->inj_comp[q]) {
      case C_XX: case C_YY: case C_ZZ: e = IC(i, j, k); break;
      case C_XY: e = IXY(i, j, k); break;
      case C_XZ: e = IXZ(i, j, k); break;
      default: e = IYZ(i, j, k); break;
    }
    REAL* a = o->f[F_SXX + o->inj_comp[q]];
    a[e] = a[e] + o->inj_coef[q] * s;
  }
}

/* Velocity pass: OK4.cu:346-381 (update_Qxft, iso: scalar buoyancy, 1-D
 * profile w[i-1]) and ANISO.cu:358-389 (update_VelocityFromStress: staggered
 * buoyancy arrays, cell weight on the old value).  Index ranges:
 * Vx 1<=i<=nx-1, 1<=j<=ny-2, 1<=k<=nz-2 ; Vy 1<=i<=nx-2, 1<=j<=ny-1,
 * 1<=k<=nz-2 ; Vz 1<=i<=nx-2, 1<=j<=ny-2, 1<=k<=nz-1. */
static void velocity(oracle_t* o) {
  const int nx = o->nx, ny = o->ny, nz = o->nz;
  REAL *Vx = o->f[F_VX], *Vy = o->f[F_VY], *Vz = o->f[F_VZ];
  REAL *sxx = o->f[F_SXX], *syy = o->f[F_SYY], *szz = o->f[F_SZZ];
  REAL *sxy = o->f[F_SXY], *sxz = o->f[F_SXZ], *syz = o->f[F_SYZ];
  const REAL dt = o->dt, rdt = (REAL)1 / o->dt;
  const REAL rdx = (REAL)1 / o->dx, rdy = (REAL)1 / o->dy, rdz = (REAL)1 / o->dz;
  const int iso = (o->mode == 0), prof = (o->pml == 1);
#pragma omp parallel for schedule(static)
  for (int k = 1; k < nz; ++k)
    for (int j = 1; j < ny; ++j)
      for (int i = 1; i < nx; ++i) {
        const int64_t c = IC(i, j, k);
        if (k <= nz - 2 && j <= ny - 2) {
          const int64_t e = IVX(i, j, k);
          const REAL dS = cpml(o, 0, 3, 1, i, c, (sxx[IC(i, j, k)] - sxx[IC(i - 1, j, k)]) * rdx) +
                          cpml(o, 1, 4, 0, j, c, (sxy[IXY(i, j + 1, k)] - sxy[IXY(i, j, k)]) * rdy) +
                          cpml(o, 2, 4, 0, k, c, (sxz[IXZ(i, j, k + 1)] - sxz[IXZ(i, j, k)]) * rdz);
          if (iso) {
            const REAL w = prof ? o->wx[i - 1] : (REAL)1;
            Vx[e] = w * (Vx[e] * rdt + o->bx * dS) * dt;
          } else {
            Vx[e] = (Vx[e] * rdt * w3(o, i, j, k) + o->B[0][e] * dS) * dt;
          }
        }
        if (k <= nz - 2 && i <= nx - 2) {
          const int64_t e = IVY(i, j, k);
          const REAL dS = cpml(o, 1, 3, 1, j, c, (syy[IC(i, j, k)] - syy[IC(i, j - 1, k)]) * rdy) +
                          cpml(o, 0, 4, 0, i, c, (sxy[IXY(i + 1, j, k)] - sxy[IXY(i, j, k)]) * rdx) +
                          cpml(o, 2, 5, 0, k, c, (syz[IYZ(i, j, k + 1)] - syz[IYZ(i, j, k)]) * rdz);
          if (iso) {
            const REAL w = prof ? o->wy[j - 1] : (REAL)1;
            Vy[e] = w * (Vy[e] * rdt + o->by * dS) * dt;
          } else {
            Vy[e] = (Vy[e] * rdt * w3(o, i, j, k) + o->B[1][e] * dS) * dt;
          }
        }
        if (j <= ny - 2 && i <= nx - 2) {
          const int64_t e = IVZ(i, j, k);
          const REAL dS = cpml(o, 2, 3, 1, k, c, (szz[IC(i, j, k)] - szz[IC(i, j, k - 1)]) * rdz) +
                          cpml(o, 0, 5, 0, i, c, (sxz[IXZ(i + 1, j, k)] - sxz[IXZ(i, j, k)]) * rdx) +
                          cpml(o, 1, 5, 0, j, c, (syz[IYZ(i, j + 1, k)] - syz[IYZ(i, j, k)]) * rdy);
          if (iso) {
            const REAL w = prof ? o->wz[k - 1] : (REAL)1;
            Vz[e] = w * (Vz[e] * rdt + o->bz * dS) * dt;
          } else {
            Vz[e] = (Vz[e] * rdt * w3(o, i, j, k) + o->B[2][e] * dS) * dt;
          }
        }
      }
}

/* Receivers sample after the velocity update of step it (OK4.cu:383-386). */
static void record(oracle_t* o) {
  if (o->it >= o->nt || !o->traces) return;
  const int nx = o->nx, ny = o->ny;
  for (int r = 0; r < o->nrec; ++r) {
    const int i = o->rec_ijk[3 * r], j = o->rec_ijk[3 * r + 1], k = o->rec_ijk[3 * r + 2];
    int64_t e;
    switch (o->rec_comp[r]) {
      case 0: e = IVX(i, j, k); break;
      case 1: e = IVY(i, j, k); break;
      default: e = IVZ(i, j, k); break;
    }
    o->traces[(int64_t)r * o->nt + o->it] = (double)o->f[o->rec_comp[r]][e];
  }
}

/* One step in the reference order: OK4.cu:485-498 (stress+source, BC1,
 * velocity+record) or ANISO.cu:521-550 (stress, BC1, source, velocity,
 * record). */
void oracle_run(oracle_t* o, int nsteps, int nthreads) {
#ifdef _OPENMP
  if (nthreads > 0) omp_set_num_threads(nthreads);
#else
  (void)nthreads;
#endif
  if (o->pml == 4 && !o->ct[0]) return; /* convolutional PML without its tables */
  for (int s = 0; s < nsteps; ++s) {
    if (o->mode == 0) stress_iso(o); else stress_ortho(o);
    if (!o->src_after_bc) inject(o);
    if (o->fs == 1) bc1(o);
    if (o->src_after_bc) inject(o);
    velocity(o);
    record(o);
    o->it++;
  }
}

int oracle_step_index(const oracle_t* o) { return o->it; }

void oracle_get_traces(const oracle_t* o, double* out) {
  if (o->traces) memcpy(out, o->traces, sizeof(double) * (size_t)o->nrec * o->nt);
}

int oracle_get_field(const oracle_t* o, int f, double* out) {
  const int64_t n = oracle_field_size(o, f);
  if (n < 0) return -1;
  for (int64_t i = 0; i < n; ++i) out[i] = (double)o->f[f][i];
  return 0;
}

int oracle_set_field(oracle_t* o, int f, const double* in) {
  const int64_t n = oracle_field_size(o, f);
  if (n < 0) return -1;
  for (int64_t i = 0; i < n; ++i) o->f[f][i] = (REAL)in[i];
  return 0;
}

double oracle_sum_sq_velocity(const oracle_t* o) {
  double s = 0.0;
  for (int f = F_VX; f <= F_VZ; ++f) {
    const int64_t n = oracle_field_size(o, f);
    for (int64_t i = 0; i < n; ++i) s += (double)o->f[f][i] * (double)o->f[f][i];
  }
  return s;
}

int oracle_real_bytes(void) { return (int)sizeof(REAL); }
