/*
 * fwe_oracle.h -- CPU restatement of the reference's FDTD time loop.
 *
 * TEST INFRASTRUCTURE ONLY.  This is the checker the HIP product is compared
 * against (tests/, __graft_entry__.smoke(), bench.py's cpu_baseline leg).
 * Nothing in the product library links, loads or calls it.
 *
 * It restates, in double precision (REAL = double; the same source builds an
 * FP32 variant with -DORACLE_FP32 for the CPU baseline), the update order,
 * index ranges, damping indexing, injection and recording semantics of
 *   OK4.cu   = src/FastWaveED_GPU3D_v1_DC_Earthquake_FULL_OK4_final/FastWaveED_GPU3D_v1.cu
 *   ANISO.cu = src/FastWaveED_GPU3D_v1_9km_station_ANISO/FastWaveED_GPU3D_v1_9kmst_ANISO.cu
 * on the reference's own staggered array shapes (x fastest):
 *   Vx (nx+1)ny nz, Vy nx(ny+1)nz, Vz nx ny(nz+1), Sxx/Syy/Szz nx ny nz,
 *   Sxy (nx+1)(ny+1)nz, Sxz (nx+1)ny(nz+1), Syz nx(ny+1)(nz+1).
 * Each function in fwe_oracle.c cites the reference lines it follows.
 */
#ifndef FWE_ORACLE_H
#define FWE_ORACLE_H
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct oracle oracle_t;

/* mode 0 = iso (OK4), 1 = ortho (ANISO); pml 0 none, 1 profile (iso),
 * 2 weights3d (ortho), 3 separable (ortho), 4 convolutional PML (either law;
 * the HIP library's own thin-layer damping, no reference counterpart: see
 * oracle_set_cpml); free_surface 0 none, 1 BC1,
 * 2 vacuum layer (shear modulus 1e-5 on plane k = 1, no BC1); src_after_bc
 * as in fwe_config. */
oracle_t* oracle_create(int mode, int nx, int ny, int nz, double dx, double dy,
                        double dz, double dt, int pml, int free_surface,
                        int src_after_bc, int nt);
void oracle_destroy(oracle_t* o);
int oracle_set_iso(oracle_t* o, const float* c11, const float* c12, double c66,
                   double bx, double by, double bz);
/* cij: C11 C22 C33 C12 C13 C23 C44yz C55xz C66xy; b: bx by bz (staggered) */
int oracle_set_ortho(oracle_t* o, const float* const cij[9], const float* const b[3]);
int oracle_set_profiles(oracle_t* o, const float* wx, const float* wy, const float* wz);
int oracle_set_weights3d(oracle_t* o, const float* w);
/* pml = 4: memory variables in the L cells at both ends of every axis; per
 * axis 6 rows of n+1 floats (b, a, 1/kappa at integer nodes x = i, then at
 * half nodes x = i - 1/2), as fwe_gen_cpml_table writes them.  Every
 * derivative g across a slab becomes g/kappa + psi with psi <- b psi + a g
 * (no reference counterpart; restated from Komatitsch & Martin 2007). */
int oracle_set_cpml(oracle_t* o, int L, const float* tx, const float* ty, const float* tz);
int oracle_set_stf(oracle_t* o, const float* stf, int n);
/* comp/ijk/coef arrays of length n (comp as FWE_SXX..FWE_SYZ) */
int oracle_set_injections(oracle_t* o, int n, const int32_t* comp, const int32_t* ijk,
                          const float* coef);
/* comp 0 Vx, 1 Vy, 2 Vz */
int oracle_set_receivers(oracle_t* o, int n, const int32_t* comp, const int32_t* ijk);
void oracle_reset(oracle_t* o);
/* nthreads <= 0: OpenMP default */
void oracle_run(oracle_t* o, int nsteps, int nthreads);
int oracle_step_index(const oracle_t* o);
void oracle_get_traces(const oracle_t* o, double* out);
/* field ids as FWE_F_VX..FWE_F_SYZ; reference shape */
int oracle_get_field(const oracle_t* o, int field, double* out);
int oracle_set_field(oracle_t* o, int field, const double* in);
int64_t oracle_field_size(const oracle_t* o, int field);
/* Sum over all fields of 0.5*rho*v^2-like quadratic form: sum V^2 (for the
 * energy/stability checks). */
double oracle_sum_sq_velocity(const oracle_t* o);
int oracle_real_bytes(void);

#ifdef __cplusplus
}
#endif
#endif
