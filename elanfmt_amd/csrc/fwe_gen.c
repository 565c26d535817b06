/*
 * fwe_gen.c -- host generators of the reference's input arrays
 * (see include/fwe_gen.h for the formulas and their sources).
 */
#include "../../include/fwe_gen.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>

#ifndef M_PI
#define M_PI 3.14159265358979323846
#endif
#define DEG (M_PI / 180.0)

/* ---- random numbers (splitmix64) ---------------------------------------- */
void fwe_rng_seed(fwe_rng* r, uint64_t seed) { r->s = seed ^ 0x5DEECE66DULL; }

uint64_t fwe_rng_next(fwe_rng* r) {
  uint64_t z = (r->s += 0x9E3779B97F4A7C15ULL);
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
  return z ^ (z >> 31);
}

double fwe_rng_uniform(fwe_rng* r) { return (double)(fwe_rng_next(r) >> 11) * (1.0 / 9007199254740992.0); }

double fwe_rng_normal(fwe_rng* r) {
  const double u1 = 1.0 - fwe_rng_uniform(r); /* (0,1] */
  const double u2 = fwe_rng_uniform(r);
  return sqrt(-2.0 * log(u1)) * cos(2.0 * M_PI * u2);
}

/* ---- source time function ------------------------------------------------ */
void fwe_gen_ricker(int32_t nt, double dt, double f0, double t0, double amp, float* out) {
  for (int32_t it = 0; it < nt; ++it) {
    const double a = M_PI * f0 * ((double)it * dt - t0);
    out[it] = (float)(amp * (1.0 - 2.0 * a * a) * exp(-a * a));
  }
}

/* ---- moment tensors ------------------------------------------------------ */
void fwe_gen_mt_sdr(double strike_deg, double dip_deg, double rake_deg, double m0, double m[6]) {
  const double f = strike_deg * DEG, d = dip_deg * DEG, l = rake_deg * DEG;
  const double sd = sin(d), cd = cos(d), s2d = sin(2 * d), c2d = cos(2 * d);
  const double sl = sin(l), cl = cos(l), sf = sin(f), cf = cos(f), s2f = sin(2 * f), c2f = cos(2 * f);
  m[0] = -m0 * (sd * cl * s2f + s2d * sl * sf * sf);   /* Mxx */
  m[1] = m0 * (sd * cl * s2f - s2d * sl * cf * cf);    /* Myy */
  m[2] = m0 * s2d * sl;                                /* Mzz */
  m[3] = m0 * (sd * cl * c2f + 0.5 * s2d * sl * s2f);  /* Mxy */
  m[4] = -m0 * (cd * cl * cf + c2d * sl * sf);         /* Mxz */
  m[5] = -m0 * (cd * cl * sf - c2d * sl * cf);         /* Myz */
}

void fwe_gen_lune_eigs(double gamma_deg, double delta_deg, double lam[3]) {
  const double g = gamma_deg * DEG, b = (90.0 - delta_deg) * DEG;
  const double r6 = 1.0 / sqrt(6.0), s3 = sqrt(3.0), s2 = sqrt(2.0);
  const double cg = cos(g), sg = sin(g), sb = sin(b), cb = cos(b);
  lam[0] = r6 * (s3 * cg * sb - sg * sb + s2 * cb);
  lam[1] = r6 * (2.0 * sg * sb + s2 * cb);
  lam[2] = r6 * (-s3 * cg * sb - sg * sb + s2 * cb);
}

void fwe_gen_mt_from_eigs(const double lam[3], const double u[9], double m0, double m[6]) {
  double L[3], M[3][3];
  for (int c = 0; c < 3; ++c) L[c] = lam[c] * sqrt(2.0) * m0;
  for (int a = 0; a < 3; ++a)
    for (int b = 0; b < 3; ++b) {
      double s = 0.0;
      for (int c = 0; c < 3; ++c) s += u[3 * a + c] * L[c] * u[3 * b + c];
      M[a][b] = s;
    }
  m[0] = M[0][0]; m[1] = M[1][1]; m[2] = M[2][2];
  m[3] = M[0][1]; m[4] = M[0][2]; m[5] = M[1][2];
}

void fwe_gen_random_rotation(fwe_rng* r, double u[9]) {
  double q[4], n = 0.0;
  do {
    n = 0.0;
    for (int c = 0; c < 4; ++c) { q[c] = fwe_rng_normal(r); n += q[c] * q[c]; }
  } while (n < 1e-12);
  n = sqrt(n);
  const double w = q[0] / n, x = q[1] / n, y = q[2] / n, z = q[3] / n;
  u[0] = 1 - 2 * (y * y + z * z); u[1] = 2 * (x * y - z * w);     u[2] = 2 * (x * z + y * w);
  u[3] = 2 * (x * y + z * w);     u[4] = 1 - 2 * (x * x + z * z); u[5] = 2 * (y * z - x * w);
  u[6] = 2 * (x * z - y * w);     u[7] = 2 * (y * z + x * w);     u[8] = 1 - 2 * (x * x + y * y);
}

static double lune_u(double b) { return 0.75 * b - 0.5 * sin(2 * b) + sin(4 * b) / 16.0; }

void fwe_gen_mt_lune_random(fwe_rng* r, double m0, double m[6], double* gamma_deg, double* delta_deg) {
  const double v = (fwe_rng_uniform(r) * 2.0 - 1.0) / 3.0;
  const double uu = fwe_rng_uniform(r) * 0.75 * M_PI;
  const double g = asin(3.0 * v) / 3.0;
  double lo = 0.0, hi = M_PI;  /* u(beta) is increasing: u' = 2 sin^4(beta) */
  for (int it = 0; it < 200; ++it) {
    const double mid = 0.5 * (lo + hi);
    if (lune_u(mid) < uu) lo = mid; else hi = mid;
  }
  const double beta = 0.5 * (lo + hi);
  const double gd = g / DEG, dd = 90.0 - beta / DEG;
  double lam[3], rot[9];
  fwe_gen_lune_eigs(gd, dd, lam);
  fwe_gen_random_rotation(r, rot);
  fwe_gen_mt_from_eigs(lam, rot, m0, m);
  if (gamma_deg) *gamma_deg = gd;
  if (delta_deg) *delta_deg = dd;
}

/* ---- damping ------------------------------------------------------------- */
void fwe_gen_pml_profile(int32_t n, int32_t L, double R, double dt, int32_t lo, int32_t hi,
                         int32_t top_undamped, float* w) {
  for (int32_t q = 0; q <= n; ++q) w[q] = 1.0f;
  for (int32_t p = 1; p <= L; ++p) {
    const double r = (double)(L - p + 1) / (double)L;
    const double d = 3.0 * 0.5 / ((double)L / (double)n) * log10(1.0 / R) * r * r;
    const float v = (float)exp(-d * dt);
    if (lo && p - 1 <= n) w[p - 1] = v;
    if (hi && n - p + 1 >= 0) w[n - p + 1] = v;
  }
  if (top_undamped)
    for (int32_t q = 0; q < n / 2; ++q) w[q] = 1.0f;
}

void fwe_gen_cell_profile(int32_t n, const float* w, float* t) {
  for (int32_t q = 0; q < n; ++q) t[q] = 0.5f * (w[q + 1] + w[q]);
}

static int margin_depth(int32_t i, int32_t n, int32_t m_lo, int32_t m_hi) {
  int p = 0;
  if (i < m_lo) p = m_lo - i;            /* ANIS.m: ix < lmargin+1 -> lmargin+1-ix */
  if (i >= n - m_hi) p = i + 1 - n + m_hi; /* ANIS.m: nx-rmargin < ix -> ix-nx+rmargin */
  return p;
}

void fwe_gen_taper(int32_t n, int32_t m_lo, int32_t m_hi, double rate, float* t) {
  for (int32_t i = 0; i < n; ++i) {
    const double p = margin_depth(i, n, m_lo, m_hi);
    t[i] = (float)exp(-rate * rate * p * p);
  }
}

/* normalised depth of a node at position x into the layer of one axis end */
static double cpml_depth(double x, int32_t n, int32_t layer, int32_t lo, int32_t hi) {
  double q = 0.0;
  if (lo) q = fmax(q, ((double)layer - 0.5 - x) / layer);
  if (hi) q = fmax(q, (x - ((double)n - 0.5 - layer)) / layer);
  return fmin(q, 1.0);
}

int32_t fwe_gen_cpml_table(int32_t n, int32_t layer, double h, double dt, double vmax, double R,
                           double kappa_max, double alpha_max, double power, int32_t lo, int32_t hi,
                           float* out) {
  if (!out || n < 1 || layer < 1 || 2 * layer > n || !(h > 0) || !(dt > 0) || !(vmax > 0) || !(R > 0 && R < 1) ||
      !(kappa_max >= 1) || !(alpha_max >= 0) || !(power >= 0))
    return -1;
  const int64_t m = (int64_t)n + 1;
  const double d0 = -(power + 1.0) * vmax * log(R) / (2.0 * layer * h);
  for (int32_t s = 0; s < 2; ++s)  /* 0: integer nodes x = i, 1: half nodes x = i - 1/2 */
    for (int64_t i = 0; i < m; ++i) {
      const double q = cpml_depth((double)i - 0.5 * s, n, layer, lo, hi);
      const double qp = q > 0 ? pow(q, power) : 0.0;
      const double d = d0 * qp, kappa = 1.0 + (kappa_max - 1.0) * qp, alpha = alpha_max * (1.0 - q);
      const double b = exp(-(d / kappa + alpha) * dt);
      const double a = d > 0 ? d * (b - 1.0) / (kappa * (d + kappa * alpha)) : 0.0;
      out[(3 * s + 0) * m + i] = (float)b;
      out[(3 * s + 1) * m + i] = (float)a;
      out[(3 * s + 2) * m + i] = (float)(1.0 / kappa);
    }
  return 0;
}

void fwe_gen_weights3d(int32_t nx, int32_t ny, int32_t nz, int32_t mx_lo, int32_t mx_hi,
                       int32_t my_lo, int32_t my_hi, int32_t mz_lo, int32_t mz_hi, double rate,
                       float* w) {
  for (int32_t k = 0; k < nz; ++k)
    for (int32_t j = 0; j < ny; ++j)
      for (int32_t i = 0; i < nx; ++i) {
        const double p = margin_depth(i, nx, mx_lo, mx_hi);
        const double q = margin_depth(j, ny, my_lo, my_hi);
        const double r = margin_depth(k, nz, mz_lo, mz_hi);
        w[(int64_t)i + (int64_t)j * nx + (int64_t)k * nx * ny] = (float)exp(-rate * rate * (p * p + q * q + r * r));
      }
}

/* ---- staggered averaging --------------------------------------------------- */
void fwe_gen_stagger_buoyancy(int32_t nx, int32_t ny, int32_t nz, const float* b, float* bx,
                              float* by, float* bz) {
#define B(i, j, k) ((double)b[(int64_t)(i) + (int64_t)(j) * nx + (int64_t)(k) * nx * ny])
  for (int32_t k = 0; k < nz; ++k)
    for (int32_t j = 0; j < ny; ++j)
      for (int32_t i = 0; i <= nx; ++i)
        bx[(int64_t)i + (int64_t)j * (nx + 1) + (int64_t)k * (nx + 1) * ny] =
            (i >= 1 && i <= nx - 1) ? (float)((B(i - 1, j, k) + B(i, j, k)) / 2.0) : 0.0f;
  for (int32_t k = 0; k < nz; ++k)
    for (int32_t j = 0; j <= ny; ++j)
      for (int32_t i = 0; i < nx; ++i)
        by[(int64_t)i + (int64_t)j * nx + (int64_t)k * nx * (ny + 1)] =
            (j >= 1 && j <= ny - 1) ? (float)((B(i, j - 1, k) + B(i, j, k)) / 2.0) : 0.0f;
  for (int32_t k = 0; k <= nz; ++k)
    for (int32_t j = 0; j < ny; ++j)
      for (int32_t i = 0; i < nx; ++i)
        bz[(int64_t)i + (int64_t)j * nx + (int64_t)k * nx * ny] =
            (k >= 1 && k <= nz - 1) ? (float)((B(i, j, k - 1) + B(i, j, k)) / 2.0) : 0.0f;
#undef B
}

void fwe_gen_harmonic_edges(int32_t nx, int32_t ny, int32_t nz, const float* c44, const float* c55,
                            const float* c66, float* gyz, float* gxz, float* gxy) {
#define C(a, i, j, k) (1.0 / (double)(a)[(int64_t)(i) + (int64_t)(j) * nx + (int64_t)(k) * nx * ny])
  for (int32_t k = 0; k <= nz; ++k)
    for (int32_t j = 0; j <= ny; ++j)
      for (int32_t i = 0; i < nx; ++i) {
        const int in = j >= 1 && j <= ny - 1 && k >= 1 && k <= nz - 1;
        gyz[(int64_t)i + (int64_t)j * nx + (int64_t)k * nx * (ny + 1)] =
            in ? (float)(4.0 / (C(c44, i, j - 1, k - 1) + C(c44, i, j, k - 1) + C(c44, i, j - 1, k) +
                                C(c44, i, j, k)))
               : 0.0f;
      }
  for (int32_t k = 0; k <= nz; ++k)
    for (int32_t j = 0; j < ny; ++j)
      for (int32_t i = 0; i <= nx; ++i) {
        const int in = i >= 1 && i <= nx - 1 && k >= 1 && k <= nz - 1;
        gxz[(int64_t)i + (int64_t)j * (nx + 1) + (int64_t)k * (nx + 1) * ny] =
            in ? (float)(4.0 / (C(c55, i - 1, j, k - 1) + C(c55, i, j, k - 1) + C(c55, i - 1, j, k) +
                                C(c55, i, j, k)))
               : 0.0f;
      }
  for (int32_t k = 0; k < nz; ++k)
    for (int32_t j = 0; j <= ny; ++j)
      for (int32_t i = 0; i <= nx; ++i) {
        const int in = i >= 1 && i <= nx - 1 && j >= 1 && j <= ny - 1;
        gxy[(int64_t)i + (int64_t)j * (nx + 1) + (int64_t)k * (nx + 1) * (ny + 1)] =
            in ? (float)(4.0 / (C(c66, i - 1, j - 1, k) + C(c66, i, j - 1, k) + C(c66, i - 1, j, k) +
                                C(c66, i, j, k)))
               : 0.0f;
      }
#undef C
}

/* ---- media --------------------------------------------------------------- */
int fwe_gen_ortho_stiffness(double vp, double vs, double rho, const double p[7], double c[9]) {
  const double e1 = p[0], e2 = p[1], d1 = p[2], d2 = p[3], d3 = p[4], g1 = p[5], g2 = p[6];
  const double C33 = rho * vp * vp, C55 = rho * vs * vs;
  const double C11 = C33 * (1 + 2 * e2), C22 = C33 * (1 + 2 * e1);
  const double C66 = C55 * (1 + 2 * g1), C44 = C66 / (1 + 2 * g2);
  const double a13 = 2 * d2 * C33 * (C33 - C55) + (C33 - C55) * (C33 - C55);
  const double a23 = 2 * d1 * C33 * (C33 - C44) + (C33 - C44) * (C33 - C44);
  const double a12 = 2 * d3 * C11 * (C11 - C66) + (C11 - C66) * (C11 - C66);
  if (a13 < 0 || a23 < 0 || a12 < 0) return -1;
  const double C13 = sqrt(a13) - C55, C23 = sqrt(a23) - C44, C12 = sqrt(a12) - C66;
  c[0] = C11; c[1] = C22; c[2] = C33; c[3] = C12; c[4] = C13; c[5] = C23;
  c[6] = C44; c[7] = C55; c[8] = C66;
  /* positive definiteness: shear moduli and leading minors of the normal block */
  if (C44 <= 0 || C55 <= 0 || C66 <= 0) return -1;
  const double m1 = C11, m2 = C11 * C22 - C12 * C12;
  const double m3 = C11 * (C22 * C33 - C23 * C23) - C12 * (C12 * C33 - C23 * C13) + C13 * (C12 * C23 - C22 * C13);
  return (m1 > 0 && m2 > 0 && m3 > 0) ? 0 : -1;
}

int fwe_gen_layered_ortho(int32_t nx, int32_t ny, int32_t nz, int32_t nl, uint64_t seed, double vp_lo,
                          double vp_hi, double r_lo, double r_hi, double rho_lo, double rho_hi,
                          double aniso_hi, float* const cij[9], float* bcell, int32_t* layer_top,
                          double* vmax) {
  if (nl < 1 || nl > nz) return -1;
  fwe_rng r;
  fwe_rng_seed(&r, seed);
  /* thicknesses: weights U[0.5,1.5], rounded, at least one plane each */
  double wsum = 0.0, wts[256];
  if (nl > 256) return -1;
  for (int l = 0; l < nl; ++l) { wts[l] = 0.5 + fwe_rng_uniform(&r); wsum += wts[l]; }
  int top = 0;
  double acc = 0.0;
  for (int l = 0; l < nl; ++l) {
    layer_top[l] = top;
    acc += wts[l];
    int next = (l == nl - 1) ? nz : (int)floor(nz * acc / wsum + 0.5);
    if (next < top + 1) next = top + 1;
    if (next > nz - (nl - 1 - l)) next = nz - (nl - 1 - l);
    top = next;
  }
  double vm = 0.0;
  const int64_t plane = (int64_t)nx * ny;
  for (int l = 0; l < nl; ++l) {
    const double vp = vp_lo + (vp_hi - vp_lo) * fwe_rng_uniform(&r);
    const double ratio = r_lo + (r_hi - r_lo) * fwe_rng_uniform(&r);
    const double rho = rho_lo + (rho_hi - rho_lo) * fwe_rng_uniform(&r);
    double c[9], p[7];
    int tries = 0;
    do {
      for (int q = 0; q < 7; ++q) p[q] = aniso_hi * fwe_rng_uniform(&r);
      if (++tries > 1000) return -1;
    } while (fwe_gen_ortho_stiffness(vp, vp / ratio, rho, p, c) != 0);
    double cmax = c[0];
    if (c[1] > cmax) cmax = c[1];
    if (c[2] > cmax) cmax = c[2];
    const double v = sqrt(cmax / rho);
    if (v > vm) vm = v;
    const int k1 = (l == nl - 1) ? nz : layer_top[l + 1];
    for (int k = layer_top[l]; k < k1; ++k)
      for (int64_t q = 0; q < plane; ++q) {
        const int64_t e = q + (int64_t)k * plane;
        for (int s = 0; s < 9; ++s) cij[s][e] = (float)c[s];
        bcell[e] = (float)(1.0 / rho);
      }
  }
  if (vmax) *vmax = vm;
  return 0;
}

static void smooth_axis(float* a, int32_t nx, int32_t ny, int32_t nz, int axis, const double* ker, int rad) {
  const int32_t n[3] = {nx, ny, nz};
  const int64_t st[3] = {1, nx, (int64_t)nx * ny};
  const int32_t len = n[axis];
  double* line = (double*)malloc(sizeof(double) * (size_t)len);
  double* outl = (double*)malloc(sizeof(double) * (size_t)len);
  const int a1 = axis == 0 ? 1 : 0, a2 = axis == 2 ? 1 : 2;
  for (int32_t u = 0; u < n[a1]; ++u)
    for (int32_t v = 0; v < n[a2]; ++v) {
      const int64_t base = (int64_t)u * st[a1] + (int64_t)v * st[a2];
      for (int32_t q = 0; q < len; ++q) line[q] = a[base + (int64_t)q * st[axis]];
      for (int32_t q = 0; q < len; ++q) {
        double s = 0.0;
        for (int t = -rad; t <= rad; ++t) {
          int32_t p = q + t;
          while (p < 0 || p >= len) p = p < 0 ? -p - 1 : 2 * len - p - 1; /* reflect */
          s += ker[t + rad] * line[p];
        }
        outl[q] = s;
      }
      for (int32_t q = 0; q < len; ++q) a[base + (int64_t)q * st[axis]] = (float)outl[q];
    }
  free(line);
  free(outl);
}

void fwe_gen_random_field(int32_t nx, int32_t ny, int32_t nz, double corr, uint64_t seed, float* out) {
  fwe_rng r;
  fwe_rng_seed(&r, seed);
  const int64_t n = (int64_t)nx * ny * nz;
  for (int64_t q = 0; q < n; ++q) out[q] = (float)fwe_rng_normal(&r);
  const double sig = corr / 2.0;
  if (sig > 0.0) {
    int rad = (int)ceil(4.0 * sig);
    double* ker = (double*)malloc(sizeof(double) * (size_t)(2 * rad + 1));
    double ks = 0.0;
    for (int t = -rad; t <= rad; ++t) { ker[t + rad] = exp(-0.5 * t * t / (sig * sig)); ks += ker[t + rad]; }
    for (int t = 0; t < 2 * rad + 1; ++t) ker[t] /= ks;
    for (int ax = 0; ax < 3; ++ax) smooth_axis(out, nx, ny, nz, ax, ker, rad);
    free(ker);
  }
  double m = 0.0, s2 = 0.0;
  for (int64_t q = 0; q < n; ++q) m += out[q];
  m /= (double)n;
  for (int64_t q = 0; q < n; ++q) s2 += (out[q] - m) * (out[q] - m);
  const double sd = sqrt(s2 / (double)n);
  for (int64_t q = 0; q < n; ++q) out[q] = (float)((out[q] - m) / (sd > 0 ? sd : 1.0));
}

double fwe_gen_cfl_dt(double vmax, double dx, double dy, double dz, double cfl) {
  return cfl / (vmax * sqrt(1.0 / (dx * dx) + 1.0 / (dy * dy) + 1.0 / (dz * dz)));
}

int32_t fwe_gen_material_table(int64_t n, int32_t m, const float* const* arrays, int32_t max_entries, uint8_t* idx,
                               float* tab) {
  enum { NSLOT = 1024 };  /* open addressing, more than twice the largest table */
  if (n < 0 || m < 1 || m > 16 || max_entries < 1 || max_entries > 256 || !arrays || (n > 0 && (!idx || !tab)))
    return -1;
  for (int q = 0; q < m; ++q)
    if (!arrays[q] && n > 0) return -1;
  int slot[NSLOT];
  for (int q = 0; q < NSLOT; ++q) slot[q] = -1;
  uint32_t t[16], prev[16];
  int ntab = 0, pe = -1;
  const size_t bytes = sizeof(uint32_t) * (size_t)m;
  for (int64_t e = 0; e < n; ++e) {
    for (int q = 0; q < m; ++q) memcpy(&t[q], &arrays[q][e], 4);
    if (pe >= 0 && memcmp(t, prev, bytes) == 0) {  /* runs of one material are the common case */
      idx[e] = (uint8_t)pe;
      continue;
    }
    uint64_t h = 1469598103934665603ull;  /* FNV-1a over the 32-bit words */
    for (int q = 0; q < m; ++q) h = (h ^ t[q]) * 1099511628211ull;
    int sl = (int)(h & (NSLOT - 1)), id;
    for (;;) {
      id = slot[sl];
      if (id < 0) {
        if (ntab >= max_entries) return -1;
        id = ntab++;
        memcpy(tab + (size_t)id * m, t, bytes);
        slot[sl] = id;
        break;
      }
      if (memcmp(tab + (size_t)id * m, t, bytes) == 0) break;
      sl = (sl + 1) & (NSLOT - 1);
    }
    idx[e] = (uint8_t)id;
    memcpy(prev, t, bytes);
    pe = id;
  }
  return ntab;
}
