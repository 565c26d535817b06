/*
 * fwe_gen.h -- host-side generators of the reference's input files
 * (exported by libfwe.so; pure CPU code, no GPU needed).
 *
 * The reference builds these inputs inside its MATLAB drivers and writes
 * them as raw float32 x-fastest .dat files; these functions produce the same
 * arrays so the MATLAB driver becomes optional.  Citations:
 *   OK4.m  = src/FastWaveED_GPU3D_v1_DC_Earthquake_FULL_OK4_final/vizme_FastBiotED_GPU3D_OK_2024_final.m
 *   ANIS.m = src/FastWaveED_GPU3D_v1_9km_station_ANISO/vizme_FastWaveED_GPU3D_9kmst_ANIS.m
 *   9km.m  = src/FastWaveED_GPU3D_v1_9km_station/vizme_FastWaveED_GPU3D_9kmst.m
 *   OK4.cu = src/FastWaveED_GPU3D_v1_DC_Earthquake_FULL_OK4_final/FastWaveED_GPU3D_v1.cu
 * Coordinates: x = north, y = east, z = down (k = 0 is the top plane).
 * Moment tensors are ordered (Mxx, Myy, Mzz, Mxy, Mxz, Myz) as MTS.dat.
 */
#ifndef FWE_GEN_H
#define FWE_GEN_H
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Seeded generator (splitmix64 state); identical streams on every host. */
typedef struct fwe_rng { uint64_t s; } fwe_rng;
void fwe_rng_seed(fwe_rng* r, uint64_t seed);
uint64_t fwe_rng_next(fwe_rng* r);
double fwe_rng_uniform(fwe_rng* r);            /* [0,1) */
double fwe_rng_normal(fwe_rng* r);             /* N(0,1), Box-Muller */

/* Ricker wavelet of the reference's pert() (OK4.cu:210-218):
 * out[it] = amp*(1-2a^2)*exp(-a^2), a = pi*f0*(it*dt - t0). */
void fwe_gen_ricker(int32_t nt, double dt, double f0, double t0, double amp, float* out);

/* Double couple (Aki & Richards 2002, Box 4.4) of scalar moment m0. */
void fwe_gen_mt_sdr(double strike_deg, double dip_deg, double rake_deg, double m0, double m[6]);
/* Lune eigenvalues (Tape & Tape 2012): unit-norm lambda1 >= lambda2 >= lambda3
 * for lune longitude gamma in [-30,30] deg and latitude delta in [-90,90]. */
void fwe_gen_lune_eigs(double gamma_deg, double delta_deg, double lam[3]);
/* M = U diag(lam*sqrt(2)*m0) U^T with U a row-major 3x3 rotation. */
void fwe_gen_mt_from_eigs(const double lam[3], const double u[9], double m0, double m[6]);
/* Uniform random rotation (unit quaternion from 4 normals). */
void fwe_gen_random_rotation(fwe_rng* r, double u[9]);
/* Moment tensor drawn uniformly on the lune (Tape & Tape 2015: v = sin(3g)/3
 * and u(beta) = 3b/4 - sin(2b)/2 + sin(4b)/16 uniform) with a uniform
 * orientation; returns gamma/delta (deg) if the pointers are non-NULL. */
void fwe_gen_mt_lune_random(fwe_rng* r, double m0, double m[6], double* gamma_deg, double* delta_deg);

/* OK4 1-D damping profile of length n+1 (OK4.m:297-338):
 * d(p) = 1.5*(n/L)*log10(1/R)*((L-p+1)/L)^2 for p = 1..L,
 * w[p-1] = w[n-p+1] = exp(-d(p)*dt), all other entries 1.
 * top_undamped sets w[0..n/2-1] = 1 (OK4.m:338, z axis). lo/hi enable the
 * low/high end of the axis.  Requires n >= 2L when both ends are on. */
void fwe_gen_pml_profile(int32_t n, int32_t L, double R, double dt, int32_t lo, int32_t hi,
                         int32_t top_undamped, float* w);
/* Cell-centred profile (length n) from an OK4 profile (length n+1):
 * t[i] = 0.5*(w[i] + w[i+1]) -- the factor OK4 applies to normal stress. */
void fwe_gen_cell_profile(int32_t n, const float* w, float* t);
/* ANISO taper factor of one axis (ANIS.m:466-506): depth p into a margin of
 * width m_lo (low end) / m_hi (high end); t[i] = exp(-(rate*p)^2). */
void fwe_gen_taper(int32_t n, int32_t m_lo, int32_t m_hi, double rate, float* t);

/* Coefficients of a convolutional PML (no reference counterpart: the
 * reference damps with the multiplicative profiles above).  Recursive
 * convolution of Roden & Gedney (2000) with the complex-frequency-shifted
 * stretching s = kappa + d/(alpha + i w) of Kuzuoglu & Mittra (1996), in the
 * form of Komatitsch & Martin (2007): every derivative g across the layer is
 * replaced by g/kappa + psi with psi <- b psi + a g.  On one axis of n
 * cells with `layer` cells at the low (lo != 0) and/or high (hi != 0) end,
 * a node at normalised depth q in [0, 1] gets
 *   d = d0 q^power, d0 = -(power+1) vmax ln(R) / (2 layer h),
 *   kappa = 1 + (kappa_max - 1) q^power, alpha = alpha_max (1 - q),
 *   b = exp(-(d/kappa + alpha) dt), a = d (b - 1) / (kappa (d + kappa alpha)).
 * Integer nodes (index i at x = i: normal stresses, and the velocities
 * across the axis) and half nodes (x = i - 1/2: the velocity along the
 * axis, the shear stresses with that axis) get q = (layer - 1/2 - x)/layer
 * at the low end and (x - (n - 1/2 - layer))/layer at the high end, clipped
 * to [0, 1] (0 on a disabled end).  out holds 6 rows of n+1 entries:
 * b, a, 1/kappa at the integer nodes, then the same at the half nodes.
 * Returns 0, or -1 for bad arguments (layer < 1, 2 layer > n, R not in
 * (0,1), kappa_max < 1, alpha_max < 0, power < 0). */
int32_t fwe_gen_cpml_table(int32_t n, int32_t layer, double h, double dt, double vmax, double R,
                           double kappa_max, double alpha_max, double power, int32_t lo, int32_t hi,
                           float* out);
/* ANISO 3-D weights.dat: w = exp(-rate^2 (p^2 + q^2 + r^2)) computed in
 * double and rounded once (ANIS.m:466-512). */
void fwe_gen_weights3d(int32_t nx, int32_t ny, int32_t nz, int32_t mx_lo, int32_t mx_hi,
                       int32_t my_lo, int32_t my_hi, int32_t mz_lo, int32_t mz_hi, double rate,
                       float* w);

/* Staggered buoyancy from a cell array b (nx*ny*nz): two-point arithmetic
 * means on interior faces, zero on the outer faces (ANIS.m:386-398). */
void fwe_gen_stagger_buoyancy(int32_t nx, int32_t ny, int32_t nz, const float* b, float* bx,
                              float* by, float* bz);
/* Edge moduli from cell moduli by 4-point harmonic means, zero on the outer
 * edges (ANIS.m:438-454, 9km.m:374-396): gyz from c44, gxz from c55, gxy
 * from c66 (each may alias the same cell array). */
void fwe_gen_harmonic_edges(int32_t nx, int32_t ny, int32_t nz, const float* c44, const float* c55,
                            const float* c66, float* gyz, float* gxz, float* gxy);

/* Orthorhombic stiffness (Pa) from vertical vp, vs (m/s), rho and Tsvankin's
 * (1997) parameters p = {eps1, eps2, delta1, delta2, delta3, gamma1, gamma2}
 * (Thomsen 1986 in each symmetry plane).  c = {C11 C22 C33 C12 C13 C23 C44
 * C55 C66}.  Returns 0 if the tensor is positive definite, -1 otherwise. */
int fwe_gen_ortho_stiffness(double vp, double vs, double rho, const double p[7], double c[9]);

/* Layered orthorhombic medium (a benchmark workload, not a reference
 * input): nl horizontal layers with seeded random thicknesses;
 * per layer vp~U[vp_lo,vp_hi], vp/vs~U[r_lo,r_hi], rho~U[rho_lo,rho_hi],
 * Tsvankin parameters ~U[0,aniso_hi].  Writes the 9 cell Cij arrays
 * (C11 C22 C33 C12 C13 C23 C44 C55 C66, cell-centred) and cell buoyancy.
 * layer_top[l] receives the first k of layer l; vmax the largest
 * quasi-P speed bound sqrt(max Cii / rho). Returns 0 on success. */
int fwe_gen_layered_ortho(int32_t nx, int32_t ny, int32_t nz, int32_t nl, uint64_t seed,
                          double vp_lo, double vp_hi, double r_lo, double r_hi, double rho_lo,
                          double rho_hi, double aniso_hi, float* const cij[9], float* bcell,
                          int32_t* layer_top, double* vmax);

/* Gaussian random field of unit standard deviation: white noise smoothed by
 * a separable Gaussian of std corr/2 cells (covariance ~ exp(-r^2/corr^2)),
 * reflecting boundaries, renormalised to sample std 1. */
void fwe_gen_random_field(int32_t nx, int32_t ny, int32_t nz, double corr, uint64_t seed, float* out);

/* dt = cfl / (vmax * sqrt(1/dx^2 + 1/dy^2 + 1/dz^2)) (OK4.m:165-169). */
double fwe_gen_cfl_dt(double vmax, double dx, double dy, double dz, double cfl);

/* Material table: the distinct m-tuples (arrays[0][e], ..., arrays[m-1][e])
 * over n elements, compared bit for bit (so -0.0 and 0.0 differ), numbered
 * in order of first appearance.  idx[e] receives the number of element e's
 * tuple and tab[t*m + q] the tuples.  Returns the number of distinct tuples,
 * or -1 when there are more than max_entries (idx and tab then hold partial
 * results) or an argument is out of range (1 <= m <= 16,
 * 1 <= max_entries <= 256).  libfwe uses it to read layered media as one
 * byte per cell (fwe_set_medium_encoding); no counterpart in the reference,
 * which streams every array. */
int32_t fwe_gen_material_table(int64_t n, int32_t m, const float* const* arrays, int32_t max_entries, uint8_t* idx,
                               float* tab);

#ifdef __cplusplus
}
#endif
#endif
