/*
 * fwe.h -- C ABI of the MI355X elastodynamic FDTD time loop (libfwe.so).
 *
 * The reference (FastWaveED_GPU3D, celsoa/ELANFMT) has no library API: its
 * boundary is a process + file-system contract.  A MATLAB driver writes raw
 * float32 x-fastest inputs (pa1/pa2/pa3/MTS/Src + material + PML weights),
 * compiles the solver with -DNBX.. macros, runs `a.exe` in its working
 * directory and reads back `0_infos.inf` and `0_0_<name>.res`
 * (src/FastWaveED_GPU3D_v1_DC_Earthquake_FULL_OK4_final/
 *  vizme_FastBiotED_GPU3D_OK_2024_final.m:348-467).
 * Every entry point below replaces one step of that contract or one piece of
 * the solver's main(); the comment on each names the reference code it
 * stands in for.  Aliases used in citations:
 *   OK4.cu   = src/FastWaveED_GPU3D_v1_DC_Earthquake_FULL_OK4_final/FastWaveED_GPU3D_v1.cu
 *   ANISO.cu = src/FastWaveED_GPU3D_v1_9km_station_ANISO/FastWaveED_GPU3D_v1_9kmst_ANISO.cu
 *   ANIS.m   = src/FastWaveED_GPU3D_v1_9km_station_ANISO/vizme_FastWaveED_GPU3D_9kmst_ANIS.m
 *   9km.cu   = src/FastWaveED_GPU3D_v1_9km_station/FastWaveED_GPU3D_v1_9kmst.cu
 *   OK4.m    = src/FastWaveED_GPU3D_v1_DC_Earthquake_FULL_OK4_final/vizme_FastBiotED_GPU3D_OK_2024_final.m
 *
 * Conventions
 *  - Every call returns fwe_status; nothing in the library calls exit().
 *    fwe_last_error() returns a thread-local message for the last failure.
 *  - The caller owns every host array it passes; a context owns its device
 *    memory.  Host arrays use the reference's staggered shapes, x fastest:
 *      cell      nx*ny*nz            Vx, bx     (nx+1)*ny*nz
 *      Vy, by    nx*(ny+1)*nz        Vz, bz     nx*ny*(nz+1)
 *      Sxy,C66xy (nx+1)*(ny+1)*nz    Sxz,C55xz  (nx+1)*ny*(nz+1)
 *      Syz,C44yz nx*(ny+1)*(nz+1)
 *  - k = 0 is the top plane (the free surface when enabled).
 *  - One context per host thread; calls on one context are serialised by
 *    the caller.  All compute runs on one HIP stream owned by the context.
 *  - There is no CPU fallback: when no HIP device is usable, fwe_create
 *    returns FWE_ERR_HIP.
 */
#ifndef FWE_H
#define FWE_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define FWE_API_VERSION 1
#define FWE_MAX_BATCH 8

typedef enum {
  FWE_OK = 0,
  FWE_ERR_ARG = 1,         /* invalid argument or size */
  FWE_ERR_IO = 2,          /* missing/short input file, write failure */
  FWE_ERR_HIP = 3,         /* HIP runtime error or no usable device */
  FWE_ERR_STATE = 4,       /* call out of order (e.g. run before material) */
  FWE_ERR_NOMEM = 5,       /* host or device allocation failed */
  FWE_ERR_UNSUPPORTED = 6, /* combination not provided (e.g. iso + 3-D weights) */
  FWE_ERR_NONFINITE = 7    /* NaN/Inf detected in the receiver traces */
} fwe_status;

/* Stress law.  ISO = OK4.cu:285-344 (cell-centred c11,c12 arrays, scalar c66,
 * scalar buoyancy pa3[0..2]); ORTHO = ANISO.cu:287-389 (9 heterogeneous Cij,
 * edge-staggered C44yz/C55xz/C66xy, staggered buoyancy arrays). */
typedef enum { FWE_MODE_ISO = 0, FWE_MODE_ORTHO = 1 } fwe_mode;

/* Damping.  PROFILE = OK4's three 1-D profiles of length n+1, applied as in
 * OK4.cu:307-329,361-375 (iso only).  WEIGHTS3D = ANISO's cell-centred 3-D
 * weights.dat multiplying the old value (ANISO.cu:306-352,374-386; ortho
 * only).  SEPARABLE = the same ANISO law with w(i,j,k)=tx[i]*ty[j]*tz[k]
 * rebuilt in the kernel from three 1-D cell-centred profiles of length n
 * (ortho only; the ANISO taper of ANIS.m:466-512 is separable).
 * CPML = a convolutional PML in a thin layer on every face (either law; no
 * reference counterpart, see fwe_set_cpml): it replaces the multiplicative
 * damping of ANISO.cu:306-352,374-386 / OK4.cu:307-329,361-375 and absorbs
 * far better per cell; the two-pass kernels only. */
typedef enum {
  FWE_PML_NONE = 0,
  FWE_PML_PROFILE = 1,
  FWE_PML_WEIGHTS3D = 2,
  FWE_PML_SEPARABLE = 3,
  FWE_PML_CPML = 4
} fwe_pml;

/* Stress components (injection targets). */
enum { FWE_SXX = 0, FWE_SYY = 1, FWE_SZZ = 2, FWE_SXY = 3, FWE_SXZ = 4, FWE_SYZ = 5 };
/* Field ids for fwe_get_field / fwe_set_field. */
enum {
  FWE_F_VX = 0, FWE_F_VY = 1, FWE_F_VZ = 2,
  FWE_F_SXX = 3, FWE_F_SYY = 4, FWE_F_SZZ = 5,
  FWE_F_SXY = 6, FWE_F_SXZ = 7, FWE_F_SYZ = 8
};
/* Receiver components. */
enum { FWE_REC_VX = 0, FWE_REC_VY = 1, FWE_REC_VZ = 2 };
/* Injection offset layouts of the reference's moment-tensor source. */
enum { FWE_SRC_LAYOUT_OK4 = 0, FWE_SRC_LAYOUT_ANISO = 1 };
/* Kernel schedule: TWO_PASS = a stress launch and a velocity launch per step
 * (in place, the reference's two-kernel structure); FUSED = the single-pass
 * sweep: one z-marching launch per step (stress of a plane, then velocity of
 * the plane behind it) over two time-level buffers, every law, damping form
 * and top boundary, the same results bit for bit; AUTO = TWO_PASS until
 * fwe_autotune has timed both, then the faster one. */
enum { FWE_KERNEL_AUTO = 0, FWE_KERNEL_TWO_PASS = 1, FWE_KERNEL_FUSED = 2 };

/* Top boundary.  VACUUM: on the shear plane k = 1 every shear stress uses
 * the modulus FWE_VACUUM_SHEAR_MODULUS (Pa) instead of the medium's, and BC1
 * is off; the caller empties the top cell layer (zero moduli at k = 0), as
 * the BC2 driver does (vizme_FastBiotED_GPU3D_OK_2024.m:238,243). */
enum { FWE_FS_NONE = 0, FWE_FS_BC1 = 1, FWE_FS_VACUUM = 2 };
#define FWE_VACUUM_SHEAR_MODULUS 1e-5

typedef struct fwe_config {
  int32_t nx, ny, nz;     /* cells: OK4.cu:39-41 (nx = 64*NBX - OVERX ...) */
  int32_t nt;             /* trace length in steps (OK4.cu:43) */
  int32_t mode;           /* fwe_mode */
  int32_t pml;            /* fwe_pml */
  int32_t free_surface;   /* fwe_free_surface: FWE_FS_BC1 applies BC1 every
                             step (OK4.cu:271-283,490); FWE_FS_VACUUM is the
                             vacuum-layer top of the BC2 variant
                             (diff_BC2_OK3/FastWaveED_GPU3D_v1.cu:309-314) */
  int32_t src_after_bc;   /* 0: inject before BC1 (OK4 order, OK4.cu:336-342);
                             1: after BC1 (ANISO order, ANISO.cu:536-538) */
  int32_t nbatch;         /* wavefields sharing one medium, 1..FWE_MAX_BATCH */
  int32_t device;         /* HIP device ordinal (OK4.cu:14 GPU_ID) */
  double dx, dy, dz, dt;  /* pa1.dat[0..3] (OK4.cu:404) */
} fwe_config;

/* One point injection: S_comp(ix,iy,iz) of wavefield `batch` gains
 * coef * stf_batch[it] after the stress update of step it.  The reference's
 * moment-tensor source is six of these (OK4.cu:336-342, ANISO.cu:228-252). */
typedef struct fwe_injection {
  int32_t comp;           /* FWE_SXX..FWE_SYZ */
  int32_t ix, iy, iz;     /* index in the component's own staggered array */
  int32_t batch;
  float coef;
} fwe_injection;

/* One receiver: trace[r][it] = V_comp(ix,iy,iz) of wavefield `batch` after
 * the velocity update of step it (OK4.cu:383-386, ANISO.cu:254-268). */
typedef struct fwe_receiver {
  int32_t comp;           /* FWE_REC_VX..FWE_REC_VZ */
  int32_t ix, iy, iz;
  int32_t batch;
} fwe_receiver;

/* Scalars of the legacy parameter packs pa1/pa2/pa3 (OK4.cu:401-408). */
typedef struct fwe_legacy_params {
  double pa1[10];         /* dx dy dz dt Lx Ly Lz lamx lamy lamz */
  double pa2[13];         /* c11u c33u c13u c12u c23u c44u c55u c66u a1 a2 a3 M1 c22u */
  double pa3[12];         /* vrho_x_11 vrho_y_11 vrho_z_11 ... (only [0..2] used) */
  double mts[6];          /* MTS.dat: Mxx Myy Mzz Mxy Mxz Myz (OK4.cu:410-411) */
  int32_t n_pa1, n_pa2, n_pa3, has_mts;
} fwe_legacy_params;

typedef struct fwe_ctx fwe_ctx;

/* ---- library ------------------------------------------------------------ */
/* "fwe <api> <arch>". */
const char* fwe_version(void);
/* Message of the last failing call on this host thread ("" if none). */
const char* fwe_last_error(void);
/* Usable HIP devices; replaces the fixed device pick of set_up_gpu
 * (OK4.cu:87-94).  Returns FWE_OK with 0 on a host without a GPU. */
fwe_status fwe_device_count(int32_t* count);

/* ---- context ------------------------------------------------------------ */
/* Reference defaults for `mode`: free surface on (BC1 runs every step in both
 * OK4.cu:490 and ANISO.cu:536), OK4 or ANISO injection order, PROFILE or
 * WEIGHTS3D damping, nbatch 1, device 0. Sizes and spacings are zeroed. */
fwe_status fwe_config_default(fwe_config* cfg, int32_t mode);
/* Allocate zeroed device fields for cfg (OK4.cu:428-439) on cfg->device and
 * a stream; replaces set_up_gpu and the zeros()/cudaMalloc blocks of main. */
fwe_status fwe_create(const fwe_config* cfg, fwe_ctx** out);
/* Free everything the context owns (OK4.cu:538-557). NULL is accepted. */
fwe_status fwe_destroy(fwe_ctx* ctx);
/* The configuration the context was created with. */
fwe_status fwe_get_config(const fwe_ctx* ctx, fwe_config* out);

/* ---- medium ------------------------------------------------------------- */
/* Iso medium of OK4.cu:415-418 (c11u_het.dat, c12u_het.dat: cell arrays),
 * with the scalar shear modulus pa2[7] (c66u, OK4.cu:323-329) and the scalar
 * buoyancies pa3[0..2] (vrho_x_11.., OK4.cu:361-375). */
fwe_status fwe_set_material_iso(fwe_ctx* ctx, const float* c11, const float* c12,
                                float c66, float bx, float by, float bz);
/* Ortho medium of ANISO.cu:411-457: cij[0..5] = C11 C22 C33 C12 C13 C23
 * (cell), cij[6] = C44yz, cij[7] = C55xz, cij[8] = C66xy (edges);
 * buoy[0..2] = vrho_{x,y,z}_11_hc (staggered with Vx, Vy, Vz). */
fwe_status fwe_set_material_ortho(fwe_ctx* ctx, const float* const cij[9],
                                  const float* const buoy[3]);
/* sqrt(largest normal modulus x largest buoyancy) of the medium set last:
 * a bound of its fastest (P) wave speed, for sizing damping profiles
 * (fwe_gen_cpml_table's vmax).  No reference counterpart (the reference's
 * driver computes its profiles from its own velocity model, OK4.m:297-341). */
fwe_status fwe_get_medium_vmax(const fwe_ctx* ctx, double* vmax);
/* PROFILE: weights_{x,y,z}.dat, lengths nx+1, ny+1, nz+1 (OK4.cu:419-424).
 * SEPARABLE: cell-centred tapers of lengths nx, ny, nz. */
fwe_status fwe_set_pml_profiles(fwe_ctx* ctx, const float* wx, const float* wy,
                                const float* wz);
/* WEIGHTS3D: weights.dat, nx*ny*nz (ANISO.cu:456-457). */
fwe_status fwe_set_pml_weights(fwe_ctx* ctx, const float* w);
/* CPML: `layer` cells at both ends of every axis (1 <= layer, 2 layer <= n)
 * and per axis the coefficient table of fwe_gen_cpml_table (6 rows of n+1
 * floats: b, a, 1/kappa at integer nodes, then at half nodes).  Every
 * derivative across a slab becomes g/kappa + psi with psi <- b psi + a g:
 * in the stress pass dVx/dx, dVy/dx, dVz/dx (x slabs; likewise y, z), in the
 * velocity pass dSxx/dx, dSxy/dx, dSxz/dx (likewise).  Resets the memory
 * variables to zero.  A face without damping has zero entries (its table's
 * end disabled in the generator); with BC1 the top end of z is normally
 * disabled.  If the call fails, runs are refused (FWE_ERR_STATE) until a
 * later call succeeds.  No reference counterpart (the reference damps with
 * the profiles above); replaces them for thin layers. */
fwe_status fwe_set_cpml(fwe_ctx* ctx, int32_t layer, const float* tx, const float* ty,
                        const float* tz);
/* Memory variables of one wavefield (floats per wavefield in *n): saved and
 * restored with the fields for an exact restart of a CPML run. */
fwe_status fwe_cpml_state_size(const fwe_ctx* ctx, int64_t* n);
fwe_status fwe_get_cpml_state(fwe_ctx* ctx, int32_t batch, float* out);
fwe_status fwe_set_cpml_state(fwe_ctx* ctx, int32_t batch, const float* in);

/* ---- source and receivers ------------------------------------------------ */
/* Source time function of wavefield `batch` (Src.dat, OK4.cu:454); steps at
 * or beyond n inject nothing. */
fwe_status fwe_set_stf(fwe_ctx* ctx, int32_t batch, const float* stf, int32_t n);
/* Replace the injection list (the hard-coded guards of OK4.cu:336-342 and
 * ANISO.cu:233-247 become data).  Indices are checked against the shapes. */
fwe_status fwe_set_injections(fwe_ctx* ctx, int32_t n, const fwe_injection* inj);
/* Replace the receiver list (the hard-coded guards of OK4.cu:383-386). */
fwe_status fwe_set_receivers(fwe_ctx* ctx, int32_t n, const fwe_receiver* rec);
/* The six injections of a moment tensor m = (Mxx Myy Mzz Mxy Mxz Myz) at base
 * cell (xs,ys,zs) with the offsets of OK4.cu:336-342 (layout OK4) or
 * ANISO.cu:233-247 (layout ANISO).  Pure host helper; out must hold 6. */
fwe_status fwe_mt_injections(int32_t layout, int32_t xs, int32_t ys, int32_t zs,
                             const float m[6], int32_t batch, fwe_injection out[6]);

/* ---- time loop ---------------------------------------------------------- */
/* Zero fields and traces and rewind the step counter (OK4.cu:428-439). */
fwe_status fwe_reset(fwe_ctx* ctx);
/* Advance nsteps steps of the hot loop (OK4.cu:485-505, ANISO.cu:519-555):
 * stress + damping [+ BC1] [+ injection], velocity + damping + receivers.
 * Asynchronous on the context stream; errors of earlier launches surface
 * here or at the next synchronising call. */
fwe_status fwe_run(fwe_ctx* ctx, int32_t nsteps);
/* As fwe_run, and brackets every stress and every velocity launch with HIP
 * events on the context stream; returns the summed kernel milliseconds and
 * the wall milliseconds of the whole call (the cudaEvent pair of
 * OK4.cu:477-518, per kernel).  Any pointer may be NULL. */
fwe_status fwe_run_timed(fwe_ctx* ctx, int32_t nsteps, double* ms_stress,
                         double* ms_velocity, double* ms_wall);
/* Block until the context stream is idle. */
fwe_status fwe_synchronize(fwe_ctx* ctx);
/* Capture g steps of the loop (2g kernel launches + a step-counter advance)
 * in one HIP graph that fwe_run replays; 0 = launch every kernel directly.
 * Replaces the per-kernel launch + cudaDeviceSynchronize of OK4.cu:485-505.
 * fwe_run_timed always launches directly. */
fwe_status fwe_set_graph_steps(fwe_ctx* ctx, int32_t g);
/* Select the kernel schedule (FWE_KERNEL_*); results agree to FP32
 * round-off.  get returns the schedule fwe_run will use. */
fwe_status fwe_set_kernel(fwe_ctx* ctx, int32_t kind);
fwe_status fwe_get_kernel(const fwe_ctx* ctx, int32_t* kind);
/* Workgroup tile: 64 x tile_rows cells (tile_rows 2, 4 or 8), marching
 * zchunk planes (3..nz).  Replaces the fixed 32x2x8 block of OK4.cu:26-31. */
fwe_status fwe_set_tiling(fwe_ctx* ctx, int32_t tile_rows, int32_t zchunk);
fwe_status fwe_get_tiling(const fwe_ctx* ctx, int32_t* tile_rows, int32_t* zchunk);
/* Time trial_steps steps for each tiling candidate, keep the fastest, and
 * reset the context (fields, traces, step counter). */
fwe_status fwe_autotune(fwe_ctx* ctx, int32_t trial_steps, double* best_ms_per_step);
/* Tiling of the single-pass sweep kernel: every wavefront owns 62 columns x
 * `rows` rows (2 or 4) and marches z chunks of `zchunk` planes (1..nz; 3..nz
 * with BC1), recomputing one plane of stresses on each side of a chunk.
 * fwe_autotune tries both row counts and several chunk heights.  Replaces
 * the fixed 32x2x8 block of OK4.cu:26-31 for this schedule. */
fwe_status fwe_set_sweep_tiling(fwe_ctx* ctx, int32_t rows, int32_t zchunk);
fwe_status fwe_get_sweep_tiling(const fwe_ctx* ctx, int32_t* rows, int32_t* zchunk);
/* Columns per tile of the two-pass kernels: 64 (one wavefront per tile row)
 * or 128 (two; halves the x-halo lines per tile).  Groups take 64 or 128 as
 * well (fwe_group_set_tile_cols, fwe_group_autotune).  fwe_autotune tries
 * both. */
fwe_status fwe_set_tile_cols(fwe_ctx* ctx, int32_t cols);
fwe_status fwe_get_tile_cols(const fwe_ctx* ctx, int32_t* cols);
/* Order in which the two-pass kernels map workgroup ids to tiles: LINEAR
 * (tile b = workgroup b, so the workgroups in flight cover one contiguous
 * stretch of every array), XCD (each of the eight XCDs gets a contiguous
 * range of tiles, so x/y neighbours share its L2) or PLANE_SPLIT (each
 * chunk of planes is cut into eight contiguous runs of tiles, one per XCD,
 * and the XCDs walk the chunks together: neighbours share an L2 and the
 * workgroups in flight still cover a few contiguous planes).  Results do
 * not depend on it; fwe_autotune tries all three, fwe_group_autotune the
 * first two (groups take LINEAR or XCD). */
enum { FWE_ORDER_LINEAR = 0, FWE_ORDER_XCD = 1, FWE_ORDER_PLANE_SPLIT = 2 };
fwe_status fwe_set_wg_order(fwe_ctx* ctx, int32_t order);
fwe_status fwe_get_wg_order(const fwe_ctx* ctx, int32_t* order);
/* How the two-pass kernels read the ortho medium (9 Cij + 3 buoyancy
 * values per cell; the reference streams all twelve arrays,
 * ANISO.cu:306-352,374-386).  TABLE: one byte per cell indexes a table of
 * the distinct 12-value tuples, read through the caches, so a cell update
 * streams 1 byte of medium instead of 48 (the same values, so the same
 * results bit for bit).  fwe_set_material_ortho builds the table when the
 * medium has at most FWE_MEDIUM_TABLE_MAX distinct tuples (layered media:
 * a few tens).  The iso law's (c11, c12) pairs get the same table (8 B per
 * entry; fwe_set_material_iso builds it), so its stress pass streams 1
 * byte of medium instead of 8.  ISOTROPIC (ortho law only): when C11 = C22 = C33 and C12 = C13 = C23 in
 * every cell (the station solver's c11/c12 file set,
 * FastWaveED_GPU3D_v1_9kmst.cu:381-400, or any isotropic medium), the
 * stress pass reads C11 and C12 for the six, eight arrays in all.  AUTO
 * (default) uses the table when it was built, else the isotropic arrays
 * when they apply, else all the arrays; ARRAYS always streams them all.
 * The single-pass kernel reads the same encoding; groups use the arrays (ortho
 * groups the isotropic arrays when every member may, else the twelve).  get returns
 * the encoding fwe_run will use and the number of table entries (0
 * without a table). */
enum { FWE_MEDIUM_AUTO = 0, FWE_MEDIUM_ARRAYS = 1, FWE_MEDIUM_TABLE = 2, FWE_MEDIUM_ISOTROPIC = 3 };
/* Wavefields of a batch that one workgroup of the two-pass kernels advances
 * together (a divisor of nbatch).  nbatch: every cell's medium is loaded
 * once for the whole batch; 1: the wavefields lie side by side in one grid
 * of single-wavefield workgroups, each streaming like a plain run.  0 (the
 * default) = nbatch with the medium arrays, 1 with the material table.  One
 * launch per half-step either way, and the same results bit for bit.
 * fwe_autotune tries nbatch and 1.  No counterpart in the reference, which
 * runs one wavefield per process. */
fwe_status fwe_set_batch_tile(fwe_ctx* ctx, int32_t wavefields);
fwe_status fwe_get_batch_tile(const fwe_ctx* ctx, int32_t* wavefields);
#define FWE_MEDIUM_TABLE_MAX 256
fwe_status fwe_set_medium_encoding(fwe_ctx* ctx, int32_t encoding);
fwe_status fwe_get_medium_encoding(const fwe_ctx* ctx, int32_t* encoding, int32_t* n_materials);
/* Steps advanced since create/reset. */
fwe_status fwe_get_step(const fwe_ctx* ctx, int32_t* it);
/* Set the step counter (restart from saved fields: the reference's unused
 * read_data hook, OK4.cu:166-175, only restores fields). */
fwe_status fwe_set_step(fwe_ctx* ctx, int32_t it);

/* ---- results ------------------------------------------------------------ */
/* Receiver traces, out[r*nt + it] (Rec..Rec_y2 of OK4.cu:454-459,525-529).
 * Returns FWE_ERR_NONFINITE (after copying) if any sample is NaN/Inf. */
fwe_status fwe_get_traces(fwe_ctx* ctx, float* out);
/* Seismograms of n moment tensors by linear combination of the traces of the
 * nbatch elementary-source wavefields: out[(m*nr + r)*nt + it] =
 * sum_b coef[m*nbatch + b] * trace[(b*nr + r)*nt + it], where the receiver
 * list is nbatch repeats of the same nr receivers, repeat b on wavefield b.
 * No counterpart in the reference, which runs one source per process
 * (OK4.cu:336-342 injects the single MTS.dat tensor); this is the batched
 * elementary-tensor form of that injection. */
fwe_status fwe_combine_traces(fwe_ctx* ctx, int32_t n, const float* coef, float* out);
/* As fwe_combine_traces (same arguments, same results bit for bit), but it
 * allocates nothing: the coefficients and the combined traces are staged
 * through a static device buffer of FWE_NOALLOC_SCRATCH_BYTES that the
 * library holds on every device it runs on until the process ends, instead
 * of device memory allocated per call.  One combined source
 * (nbatch + receivers-per-wavefield * nt floats) must fit in it, else
 * FWE_ERR_ARG.  Every call on one device shares that buffer: calls that
 * overlap on one device (from several host threads) race, and callers must
 * serialise them.  No counterpart in the reference; it is for callers that
 * cannot allocate device memory (serving loops, code around captured
 * graphs, allocation-free harnesses). */
#define FWE_NOALLOC_SCRATCH_BYTES (32u << 20)
fwe_status fwe_combine_traces_noalloc(fwe_ctx* ctx, int32_t n, const float* coef, float* out);
/* Copy one field of wavefield `batch` to/from a host array of its reference
 * shape (gather/SaveArray, OK4.cu:75,164,530-532; read_data, OK4.cu:166-175). */
fwe_status fwe_get_field(fwe_ctx* ctx, int32_t batch, int32_t field, float* out);
fwe_status fwe_set_field(fwe_ctx* ctx, int32_t batch, int32_t field, const float* in);

/* ---- legacy file contract ----------------------------------------------- */
/* Read pa1.dat, pa2.dat, pa3.dat and (if present) MTS.dat from dir
 * (load(pa1..), OK4.cu:401-411). */
fwe_status fwe_read_legacy_params(const char* dir, fwe_legacy_params* out);
/* Read the medium and damping files of the context's mode from dir:
 * iso: c11u_het.dat c12u_het.dat weights_{x,y,z}.dat (OK4.cu:415-425);
 * ortho: C11m..C23m, C44myz, C55mxz, C66mxy, vrho_{x,y,z}_11_hc.dat and
 * weights.dat (ANISO.cu:411-457), or, when C11m.dat is absent and
 * c44u_hetxy.dat present, the iso-het set of the station solver
 * (9km.cu:381-400: c11u_het, c12u_het, c44u_het{xy,xz,yz}, the same
 * buoyancy and damping files), run by the ortho law with C11=C22=C33 and
 * C12=C13=C23; plus Src.dat (nt values, OK4.cu:454).
 * Scalars come from p (c66 = pa2[7], buoyancy = pa3[0..2]). */
fwe_status fwe_load_legacy_inputs(fwe_ctx* ctx, const char* dir,
                                  const fwe_legacy_params* p);
/* Write 0_infos.inf ("4 nx ny nz 1 ceil(10/200)", OK4.cu:153-157) and
 * 0_0_<name>.res for Src, each trace name in rec_names (one per receiver,
 * in receiver order) and Vx, Vy, Vz of wavefield 0 (OK4.cu:521-532). */
fwe_status fwe_write_legacy_outputs(fwe_ctx* ctx, const char* dir,
                                    const char* const* rec_names, int32_t n_names);

/* ---- groups of independent models ---------------------------------------- */
/* The reference runs one model per process; a group advances n fully set-up
 * single-wavefield contexts (same device, law, damping kind and free-surface
 * flag; grids, media, sources and receivers free) with ONE stress launch and
 * ONE velocity launch per step covering every model's tiles, so an uneven
 * batch of small grids fills the GPU.  Members keep their own fields and
 * traces (read them with fwe_get_traces / fwe_get_field); their step
 * counters advance together.  The group uses the two-pass kernels and one
 * tile shape for all members.  No counterpart in the reference. */
typedef struct fwe_group fwe_group;
fwe_status fwe_group_create(int32_t n, fwe_ctx* const* members, fwe_group** out);
fwe_status fwe_group_destroy(fwe_group* g);
fwe_status fwe_group_run(fwe_group* g, int32_t nsteps);
/* as fwe_group_run; *ms = GPU milliseconds of the nsteps steps */
fwe_status fwe_group_run_timed(fwe_group* g, int32_t nsteps, double* ms);
fwe_status fwe_group_set_tiling(fwe_group* g, int32_t tile_rows, int32_t zchunk);
/* Columns per tile of every member, 64 or 128 (as fwe_set_tile_cols). */
fwe_status fwe_group_set_tile_cols(fwe_group* g, int32_t cols);
/* time each tile shape, keep the fastest, reset every member */
fwe_status fwe_group_autotune(fwe_group* g, int32_t trial_steps, double* best_ms_per_step);

/* ---- learned thin boundary ---------------------------------------------- */
/* A thin absorbing layer of `layer` cells on all six faces whose cells are
 * corrected after every step by a small per-face network: the nine fields
 * at a 9-cell stencil (the cell, its six neighbours, two more cells along
 * the face normal; stresses divided by an impedance Z; the vector divided
 * by its RMS r) -> 64 -> 64 -> 9 (SiLU), corrections r*y for the velocities
 * and Z*r*y for the stresses, applied after the step in their dissipative
 * part only (a value may move toward zero by at most its magnitude and
 * never grow), so the layer is an adaptive sponge.  W1 [81][64], b1 [layer][64] (one first-layer bias
 * per depth), W2 [64][64], b2 [64], W3 [64][9], b3 [9], stored FP32 and used
 * rounded to bfloat16, FP32 accumulation.  Faces 0..5 = x-, x+, y-, y+, z-,
 * z+.  Replaces the reference's PML weights (OK4.m:297-341, ANIS.m:466-512)
 * over most of their thickness; there is no network in the reference. */
typedef struct fwe_lb fwe_lb;
fwe_status fwe_lb_param_count(int32_t layer, int32_t* n_per_face);
fwe_status fwe_lb_create(int32_t device, int32_t layer, fwe_lb** out);
fwe_status fwe_lb_destroy(fwe_lb* lb);
/* He-normal hidden layers, zero output layer (no correction); resets Adam */
fwe_status fwe_lb_init(fwe_lb* lb, uint64_t seed);
/* How corrections reach the fields after each step: DISSIPATIVE (the
 * default) keeps only the part that moves a value toward zero, by at most
 * its own magnitude; FREE adds them as predicted (for a layer whose own
 * damping is already absorbing, the convolutional PML).  Waits for the steps
 * queued on attached contexts. */
enum { FWE_LB_DISSIPATIVE = 0, FWE_LB_FREE = 1 };
fwe_status fwe_lb_set_correction_form(fwe_lb* lb, int32_t form);
fwe_status fwe_lb_get_correction_form(const fwe_lb* lb, int32_t* form);
fwe_status fwe_lb_set_params(fwe_lb* lb, int32_t face, const float* params);
fwe_status fwe_lb_get_params(fwe_lb* lb, int32_t face, float* params);
/* Network of one face on n explicit samples (host arrays): x [n][81],
 * depth [n].  y [n][9] (optional).  With target and mask [n][9]: loss =
 * sum mask*(y-target)^2 / (9n) and its gradient [n_per_face] w.r.t. the
 * parameters (the used bfloat16 values). */
fwe_status fwe_lb_eval(fwe_lb* lb, int32_t face, int32_t n, const float* x, const int32_t* depth,
                       const float* target, const float* mask, float* y, float* grad, double* loss);
/* One training step.  `thick` is a run whose grid is the thin grid plus
 * `offset` cells on every face (thick damping).  With resync (or when the
 * step counters differ) its fields at step n are first copied into `thin`
 * (same medium, `layer`-cell plain damping or convolutional PML); otherwise
 * `thin` continues from its own state.  With the convolutional PML both runs
 * use it, the thick run's layer covers the thin one plus the offset, and the
 * resync also copies the thin slabs' memory variables.  Both advance one step; each face's network is fitted (one
 * Adam step, rate lr) to the difference between the thick step and the
 * thin plain step in the layer cells, and `thin` then receives the
 * corrections predicted before the update, as with the boundary attached.  Each cell's error is weighted by (r / r_ref)^2, r its
 * stencil RMS and r_ref the RMS over the layer, so the loss is the squared
 * error in physical units; cells with r < 1e-4 r_ref are left out.
 * `thick` ends one step further; *loss is the weighted mean squared error
 * before the update (0 while the layer holds no wave). */
fwe_status fwe_lb_train_step(fwe_lb* lb, fwe_ctx* thick, fwe_ctx* thin, int32_t offset, float impedance, float lr,
                             int32_t resync, double* loss);
/* Attach (lb != NULL) or detach the learned boundary: every later step
 * applies the corrections to the layer cells of every wavefield.  Two-pass
 * kernels, no free surface. */
fwe_status fwe_set_learned_boundary(fwe_ctx* ctx, fwe_lb* lb, float impedance);

/* ---- measurement -------------------------------------------------------- */
/* Stream-copy probe: copies `bytes` bytes device->device `reps` times with a
 * plain vector copy kernel and returns the mean ms per copy (the memory-copy
 * figure of Performance/PerfMC.m:1-15). */
fwe_status fwe_copy_probe(int32_t device, int64_t bytes, int32_t reps, double* ms_per_copy);

#ifdef __cplusplus
}
#endif
#endif
