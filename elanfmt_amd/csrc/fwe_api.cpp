// fwe_api.cpp -- host implementation of the C ABI in include/fwe.h.
//
// Replaces the host half of the reference solver's main()
// (OK4.cu:390-559, ANISO.cu:392-608): device set-up, input loading, the
// time loop, timing and output writing.  The reference's file formats are
// kept byte for byte; the device layout is this library's own (one padded
// x-fastest layout shared by all staggered arrays).
#include <math.h>
#include <stdarg.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include <algorithm>
#include <mutex>
#include <string>
#include <vector>

#include "../../include/fwe.h"
#include "../../include/fwe_gen.h"
#include "fwe_internal.h"

using fwe::PointEntry;
using fwe::StepArgs;

// ---------------------------------------------------------------------------
// errors
// ---------------------------------------------------------------------------
static thread_local char g_err[512];

static fwe_status fail(fwe_status st, const char* fmt, ...) {
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(g_err, sizeof(g_err), fmt, ap);
  va_end(ap);
  return st;
}

#define HIPCHK(expr)                                                               \
  do {                                                                             \
    hipError_t e_ = (expr);                                                        \
    if (e_ != hipSuccess)                                                          \
      return fail(FWE_ERR_HIP, "%s failed: %s (%s:%d)", #expr, hipGetErrorString(e_), \
                  __FILE__, __LINE__);                                             \
  } while (0)

extern "C" const char* fwe_last_error(void) { return g_err; }

extern "C" const char* fwe_version(void) {
  static char v[64];
  snprintf(v, sizeof(v), "fwe %d gfx950", FWE_API_VERSION);
  return v;
}

extern "C" fwe_status fwe_device_count(int32_t* count) {
  if (!count) return fail(FWE_ERR_ARG, "count is NULL");
  int n = 0;
  if (hipGetDeviceCount(&n) != hipSuccess) n = 0;
  *count = n;
  return FWE_OK;
}

// ---------------------------------------------------------------------------
// context
// ---------------------------------------------------------------------------
struct fwe_ctx {
  fwe_config cfg;
  int64_t sy = 0, sz = 0, nelem = 0;
  // fields (nb * 9 * nelem): fbuf[cur] holds the current time level; the
  // single-pass kernel writes the next level into fbuf[cur^1]
  float* fbuf[2] = {nullptr, nullptr};
  int cur = 0;
  // requested schedule (FWE_KERNEL_*); under AUTO, auto_sweep says which
  // one fwe_autotune measured faster
  int kernel = FWE_KERNEL_AUTO;
  bool auto_sweep = false;
  // single-pass (sweep) kernel: output rows per lane, z chunk, tile counts
  // (62-column tiles) and its point lists -- injections per tile grown by
  // one cell in x and y, receivers per (tile, plane) -- rebuilt before the
  // next sweep launch when points or the sweep tiling change
  int sw_rows = 2, sw_kchunk = 32, sw_ntx = 0, sw_nty = 0, sw_nkc = 0;
  bool sw_dirty = true;
  int32_t* sw_inj_off = nullptr;
  PointEntry* sw_inj = nullptr;
  int32_t* sw_rec_koff = nullptr;
  PointEntry* sw_rec = nullptr;
  float* medium = nullptr;  // 12 * nelem (ortho) or 2 * nelem (iso)
  // ortho material table: mid[c] selects twelve floats at mtab + 12*mid[c]
  // (ntab entries; 0 = the medium has too many distinct tuples)
  uint8_t* mid = nullptr;
  float* mtab = nullptr;
  int ntab = 0;
  int medium_enc = FWE_MEDIUM_AUTO;
  // ortho medium with C11 = C22 = C33 and C12 = C13 = C23 everywhere
  bool iso_medium = false;
  // wavefields per workgroup of a batch (a divisor of nbatch; 0 = default)
  int batch_tile = 0;
  float* w3 = nullptr;      // nelem (W3D)
  // convolutional PML: layer cells per face, the three axis tables (6 rows
  // of n+1 each, x then y then z) and the memory variables (pstride floats
  // per wavefield: x slab, y slab, z slab, six variables each)
  int cpml_L = 0;
  double vmax = 0;  // sqrt(largest normal modulus x largest buoyancy) of the medium
  float* ctab = nullptr;
  float* psi = nullptr;
  int64_t pvar[3] = {0, 0, 0};
  int64_t pstride = 0;
  // the tiles that touch no slab (a box, StepArgs::bx0..) run through the
  // plain kernel and the others (listed in cpml_tiles) through the CPML one;
  // rebuilt when the tiling or the layer changes (cpml_box[1] = 0: no box)
  int32_t cpml_box[6] = {0, 0, 0, 0, 0, 0};
  int32_t* cpml_tiles = nullptr;
  int32_t cpml_ntiles = 0;
  float* prof = nullptr;    // wx wy wz whx why whz (profile / taper tables)
  float* stf = nullptr;     // nb * nstf
  float* traces = nullptr;  // nrec * nt
  int32_t* inj_off = nullptr;
  PointEntry* inj = nullptr;
  int32_t* rec_off = nullptr;
  int32_t* rec_koff = nullptr;
  PointEntry* rec = nullptr;
  int32_t nstf = 0, ninj = 0, nrec = 0;
  std::vector<float> stf_host;  // batch 0, for the legacy Src output
  bool have_medium = false, have_damp = false;
  float c66dt = 0, bxdt = 0, bydt = 0, bzdt = 0;
  int tx = 64, ty = 4, kchunk = 8, ntx = 0, nty = 0, nkc = 0;
  int wg_order = FWE_ORDER_LINEAR;
  int it = 0;
  hipStream_t stream = nullptr;
  std::vector<hipEvent_t> events;
  // host copies of the point lists (re-sorted when the tiling changes)
  std::vector<PointEntry> inj_host, rec_host;
  // captured time loop: graph_steps steps + counter advance, replayed
  int32_t* it_dev = nullptr;
  hipGraphExec_t gexec = nullptr;
  int graph_steps = 0;
  int graph_from = 0;  // time-level buffer the captured graph starts from
  // learned boundary attached with fwe_set_learned_boundary
  fwe_lb* lb = nullptr;
  float lb_zinv = 1.f;
  fwe::LbCell* lb_cells = nullptr;  // layer cells grouped by face
  int64_t lb_face_off[7] = {0, 0, 0, 0, 0, 0, 0};
  float* lb_D = nullptr;            // corrections [cells][9]
};

static fwe_status cpml_split_update(fwe_ctx* x);

static void drop_graph(fwe_ctx* x) {
  if (x->gexec) hipGraphExecDestroy(x->gexec);
  x->gexec = nullptr;
}

static float* fields(const fwe_ctx* x) { return x->fbuf[x->cur]; }
static size_t field_bytes(const fwe_ctx* x) { return sizeof(float) * (size_t)x->nelem * 9 * x->cfg.nbatch; }

static int64_t round_up(int64_t a, int64_t m) { return (a + m - 1) / m * m; }

static void field_shape(const fwe_config& c, int f, int64_t d[3]) {
  const int64_t nx = c.nx, ny = c.ny, nz = c.nz;
  switch (f) {
    case FWE_F_VX: d[0] = nx + 1; d[1] = ny; d[2] = nz; break;
    case FWE_F_VY: d[0] = nx; d[1] = ny + 1; d[2] = nz; break;
    case FWE_F_VZ: d[0] = nx; d[1] = ny; d[2] = nz + 1; break;
    case FWE_F_SXY: d[0] = nx + 1; d[1] = ny + 1; d[2] = nz; break;
    case FWE_F_SXZ: d[0] = nx + 1; d[1] = ny; d[2] = nz + 1; break;
    case FWE_F_SYZ: d[0] = nx; d[1] = ny + 1; d[2] = nz + 1; break;
    default: d[0] = nx; d[1] = ny; d[2] = nz; break;
  }
}

// reference shape (ex,ey,ez) -> padded layout, scaled
static void pack(const fwe_ctx* x, const float* src, const int64_t d[3], double scale,
                 std::vector<float>& dst) {
  dst.assign((size_t)x->nelem, 0.f);
  for (int64_t k = 0; k < d[2]; ++k)
    for (int64_t j = 0; j < d[1]; ++j) {
      const float* s = src + j * d[0] + k * d[0] * d[1];
      float* o = dst.data() + j * x->sy + k * x->sz;
      if (scale == 1.0)
        memcpy(o, s, sizeof(float) * (size_t)d[0]);
      else
        for (int64_t i = 0; i < d[0]; ++i) o[i] = (float)((double)s[i] * scale);
    }
}

static void unpack(const fwe_ctx* x, const std::vector<float>& src, const int64_t d[3], float* dst) {
  for (int64_t k = 0; k < d[2]; ++k)
    for (int64_t j = 0; j < d[1]; ++j)
      memcpy(dst + j * d[0] + k * d[0] * d[1], src.data() + j * x->sy + k * x->sz,
             sizeof(float) * (size_t)d[0]);
}

static void choose_tiling(fwe_ctx* x) {
  const fwe_config& c = x->cfg;
  const char* ety = getenv("FWE_TILE_Y");
  const char* ekc = getenv("FWE_KCHUNK");
  int ty = ety ? atoi(ety) : 4;
  if (ty != 2 && ty != 4 && ty != 8) ty = 4;
  x->ty = ty;
  const char* etx = getenv("FWE_TILE_X");
  x->tx = (etx && atoi(etx) == 128) ? 128 : 64;
  const char* eo = getenv("FWE_WG_ORDER");
  x->wg_order = eo ? atoi(eo) : FWE_ORDER_LINEAR;
  if (x->wg_order < FWE_ORDER_LINEAR || x->wg_order > FWE_ORDER_PLANE_SPLIT) x->wg_order = FWE_ORDER_LINEAR;
  const char* eb = getenv("FWE_BATCH_TILE");
  const int bt = eb ? atoi(eb) : 0;
  x->batch_tile = (bt > 0 && bt <= c.nbatch && c.nbatch % bt == 0) ? bt : 0;
  x->ntx = (c.nx + x->tx - 1) / x->tx;
  x->nty = (c.ny + ty - 1) / ty;
  // one plane per workgroup in linear order streams best on the measured
  // cases (the autotuner confirms or replaces this)
  int kc = ekc ? atoi(ekc) : 1;
  const int kmin = c.free_surface == FWE_FS_BC1 ? 3 : 1;  // BC1 needs planes 1 and 2 in one chunk
  if (kc < kmin) kc = kmin;
  if (kc > c.nz) kc = c.nz;
  x->kchunk = kc;
  x->nkc = (c.nz + kc - 1) / kc;
  // sweep kernel: 2 rows per lane, 12-plane chunks (3 at least with BC1)
  const char* esr = getenv("FWE_SWEEP_ROWS");
  const char* esk = getenv("FWE_SWEEP_KCHUNK");
  x->sw_rows = (esr && atoi(esr) == 4) ? 4 : 2;
  int skc = esk ? atoi(esk) : 12;
  if (skc < kmin) skc = kmin;
  if (skc > c.nz) skc = c.nz;
  x->sw_kchunk = skc;
  x->sw_ntx = (c.nx + fwe::SWEEP_X - 1) / fwe::SWEEP_X;
  x->sw_nty = (c.ny + x->sw_rows - 1) / x->sw_rows;
  x->sw_nkc = (c.nz + skc - 1) / skc;
  x->sw_dirty = true;
}

extern "C" fwe_status fwe_config_default(fwe_config* cfg, int32_t mode) {
  if (!cfg) return fail(FWE_ERR_ARG, "cfg is NULL");
  if (mode != FWE_MODE_ISO && mode != FWE_MODE_ORTHO) return fail(FWE_ERR_ARG, "bad mode %d", mode);
  memset(cfg, 0, sizeof(*cfg));
  cfg->mode = mode;
  cfg->pml = mode == FWE_MODE_ISO ? FWE_PML_PROFILE : FWE_PML_WEIGHTS3D;
  cfg->free_surface = FWE_FS_BC1;
  cfg->src_after_bc = mode == FWE_MODE_ISO ? 0 : 1;
  cfg->nbatch = 1;
  return FWE_OK;
}

static fwe_status check_cfg(const fwe_config* c) {
  if (c->nx < 3 || c->ny < 3 || c->nz < 3) return fail(FWE_ERR_ARG, "grid %dx%dx%d too small", c->nx, c->ny, c->nz);
  if ((int64_t)c->nx * c->ny * c->nz > (int64_t)1 << 33) return fail(FWE_ERR_ARG, "grid too large");
  if (c->nt < 0) return fail(FWE_ERR_ARG, "nt < 0");
  if (c->mode != FWE_MODE_ISO && c->mode != FWE_MODE_ORTHO) return fail(FWE_ERR_ARG, "bad mode");
  if (c->nbatch < 1 || c->nbatch > FWE_MAX_BATCH) return fail(FWE_ERR_ARG, "nbatch %d out of 1..%d", c->nbatch, FWE_MAX_BATCH);
  if (!(c->dx > 0 && c->dy > 0 && c->dz > 0 && c->dt > 0)) return fail(FWE_ERR_ARG, "dx,dy,dz,dt must be > 0");
  if (c->mode == FWE_MODE_ISO && c->pml != FWE_PML_NONE && c->pml != FWE_PML_PROFILE && c->pml != FWE_PML_CPML)
    return fail(FWE_ERR_UNSUPPORTED, "iso law takes PROFILE, CPML or NONE damping");
  if (c->mode == FWE_MODE_ORTHO && c->pml == FWE_PML_PROFILE)
    return fail(FWE_ERR_UNSUPPORTED, "ortho law takes WEIGHTS3D, SEPARABLE, CPML or NONE damping");
  if (c->pml < 0 || c->pml > 4) return fail(FWE_ERR_ARG, "bad pml");
  if (c->free_surface < FWE_FS_NONE || c->free_surface > FWE_FS_VACUUM)
    return fail(FWE_ERR_ARG, "free_surface must be 0 (none), 1 (BC1) or 2 (vacuum layer)");
  return FWE_OK;
}

extern "C" fwe_status fwe_create(const fwe_config* cfg, fwe_ctx** out) {
  if (!cfg || !out) return fail(FWE_ERR_ARG, "NULL argument");
  *out = nullptr;
  fwe_status st = check_cfg(cfg);
  if (st != FWE_OK) return st;
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0)
    return fail(FWE_ERR_HIP, "no HIP device available (libfwe has no CPU path)");
  if (cfg->device < 0 || cfg->device >= ndev) return fail(FWE_ERR_ARG, "device %d of %d", cfg->device, ndev);
  HIPCHK(hipSetDevice(cfg->device));
  fwe_ctx* x = new (std::nothrow) fwe_ctx();
  if (!x) return fail(FWE_ERR_NOMEM, "host allocation");
  x->cfg = *cfg;
  x->sy = round_up(cfg->nx + 1, 32);
  x->sz = x->sy * (cfg->ny + 1);
  x->nelem = x->sz * (cfg->nz + 1);
  choose_tiling(x);
  const size_t fbytes = sizeof(float) * (size_t)x->nelem * 9 * cfg->nbatch;
  const int nmed = cfg->mode == FWE_MODE_ORTHO ? 12 : 2;
  hipError_t e = hipMalloc(&x->fbuf[0], fbytes);
  if (e == hipSuccess) e = hipMalloc(&x->medium, sizeof(float) * (size_t)x->nelem * nmed);
  if (e == hipSuccess && cfg->pml == FWE_PML_WEIGHTS3D) e = hipMalloc(&x->w3, sizeof(float) * (size_t)x->nelem);
  if (e == hipSuccess) {
    const size_t np = 2 * ((size_t)cfg->nx + cfg->ny + cfg->nz + 3);
    e = hipMalloc(&x->prof, sizeof(float) * np);
  }
  if (e == hipSuccess) e = hipMalloc(&x->it_dev, sizeof(int32_t));
  if (e == hipSuccess) e = hipStreamCreateWithFlags(&x->stream, hipStreamNonBlocking);
  if (e == hipSuccess) e = hipMemsetAsync(x->fbuf[0], 0, fbytes, x->stream);
  if (e == hipSuccess) e = hipMemsetAsync(x->medium, 0, sizeof(float) * (size_t)x->nelem * nmed, x->stream);
  if (e == hipSuccess) e = hipStreamSynchronize(x->stream);
  if (e != hipSuccess) {
    fwe_destroy(x);
    return fail(e == hipErrorOutOfMemory ? FWE_ERR_NOMEM : FWE_ERR_HIP, "create: %s", hipGetErrorString(e));
  }
  // default: no injections, no receivers (empty tile lists)
  fwe_set_injections(x, 0, nullptr);
  fwe_set_receivers(x, 0, nullptr);
  *out = x;
  g_err[0] = 0;
  return FWE_OK;
}

static void lb_detach(fwe_ctx* x);

extern "C" fwe_status fwe_destroy(fwe_ctx* x) {
  if (!x) return FWE_OK;
  hipSetDevice(x->cfg.device);
  if (x->stream) hipStreamSynchronize(x->stream);
  lb_detach(x);
  drop_graph(x);
  hipFree(x->it_dev);
  for (auto ev : x->events) hipEventDestroy(ev);
  hipFree(x->fbuf[0]); hipFree(x->fbuf[1]); hipFree(x->medium); hipFree(x->w3); hipFree(x->prof);
  hipFree(x->ctab); hipFree(x->psi); hipFree(x->cpml_tiles);
  hipFree(x->mid); hipFree(x->mtab);
  hipFree(x->lb_cells); hipFree(x->lb_D);
  hipFree(x->stf); hipFree(x->traces); hipFree(x->inj_off); hipFree(x->inj);
  hipFree(x->sw_inj_off); hipFree(x->sw_inj); hipFree(x->sw_rec_koff); hipFree(x->sw_rec);
  hipFree(x->rec_off); hipFree(x->rec_koff); hipFree(x->rec);
  if (x->stream) hipStreamDestroy(x->stream);
  delete x;
  return FWE_OK;
}

extern "C" fwe_status fwe_get_config(const fwe_ctx* x, fwe_config* out) {
  if (!x || !out) return fail(FWE_ERR_ARG, "NULL argument");
  *out = x->cfg;
  return FWE_OK;
}

static fwe_status upload(fwe_ctx* x, float* dst, const std::vector<float>& h) {
  HIPCHK(hipSetDevice(x->cfg.device));
  HIPCHK(hipMemcpyAsync(dst, h.data(), sizeof(float) * h.size(), hipMemcpyHostToDevice, x->stream));
  HIPCHK(hipStreamSynchronize(x->stream));
  return FWE_OK;
}

// ---------------------------------------------------------------------------
// medium and damping
// ---------------------------------------------------------------------------
// Material table of a medium: the distinct (bitwise) tuples over every
// element of the padded layout, padding included, in order of first
// appearance, and one byte per element naming its tuple.  A tuple is
// (C[0..8], B[0..2]) for the ortho law and (c11, c12) for the iso law.  Built when there are at most FWE_MEDIUM_TABLE_MAX tuples;
// otherwise ntab = 0 and the kernels stream the arrays.
static fwe_status build_material_table(fwe_ctx* x, const std::vector<std::vector<float>>& h) {
  const int64_t n = x->nelem;
  const int m = (int)h.size();  // 12 (ortho) or 2 (iso: c11, c12)
  std::vector<uint8_t> idx((size_t)n);
  std::vector<float> tab((size_t)12 * FWE_MEDIUM_TABLE_MAX);
  const float* arrays[12];
  for (int q = 0; q < m; ++q) arrays[q] = h[q].data();
  const int32_t nt = fwe_gen_material_table(n, m, arrays, FWE_MEDIUM_TABLE_MAX, idx.data(), tab.data());
  x->ntab = 0;
  if (nt < 1) return FWE_OK;
  HIPCHK(hipSetDevice(x->cfg.device));
  if (!x->mid) HIPCHK(hipMalloc(&x->mid, (size_t)n));
  if (!x->mtab) HIPCHK(hipMalloc(&x->mtab, sizeof(float) * 12 * FWE_MEDIUM_TABLE_MAX));
  HIPCHK(hipMemcpyAsync(x->mid, idx.data(), (size_t)n, hipMemcpyHostToDevice, x->stream));
  HIPCHK(hipMemcpyAsync(x->mtab, tab.data(), sizeof(float) * m * (size_t)nt, hipMemcpyHostToDevice, x->stream));
  HIPCHK(hipStreamSynchronize(x->stream));
  x->ntab = nt;
  return FWE_OK;
}

extern "C" fwe_status fwe_set_material_iso(fwe_ctx* x, const float* c11, const float* c12, float c66,
                                           float bx, float by, float bz) {
  if (!x || !c11 || !c12) return fail(FWE_ERR_ARG, "NULL argument");
  if (x->cfg.mode != FWE_MODE_ISO) return fail(FWE_ERR_STATE, "context is not iso");
  const double dt = x->cfg.dt;
  int64_t d[3];
  field_shape(x->cfg, FWE_F_SXX, d);
  drop_graph(x);
  std::vector<std::vector<float>> h(2);
  pack(x, c11, d, dt, h[0]);
  fwe_status st = upload(x, x->medium, h[0]);
  if (st) return st;
  pack(x, c12, d, dt, h[1]);
  st = upload(x, x->medium + x->nelem, h[1]);
  if (st) return st;
  x->c66dt = (float)((double)c66 * dt);
  x->bxdt = (float)((double)bx * dt);
  x->bydt = (float)((double)by * dt);
  x->bzdt = (float)((double)bz * dt);
  {
    double cmax = 0;
    for (int64_t q = 0; q < d[0] * d[1] * d[2]; ++q) cmax = std::max(cmax, (double)c11[q]);
    x->vmax = sqrt(cmax * std::max({(double)bx, (double)by, (double)bz}));
  }
  x->have_medium = true;
  return build_material_table(x, h);
}

extern "C" fwe_status fwe_get_medium_vmax(const fwe_ctx* x, double* vmax) {
  if (!x || !vmax) return fail(FWE_ERR_ARG, "NULL argument");
  if (!x->have_medium) return fail(FWE_ERR_STATE, "medium not set");
  *vmax = x->vmax;
  return FWE_OK;
}

extern "C" fwe_status fwe_set_material_ortho(fwe_ctx* x, const float* const cij[9],
                                             const float* const buoy[3]) {
  if (!x || !cij || !buoy) return fail(FWE_ERR_ARG, "NULL argument");
  if (x->cfg.mode != FWE_MODE_ORTHO) return fail(FWE_ERR_STATE, "context is not ortho");
  for (int c = 0; c < 9; ++c) if (!cij[c]) return fail(FWE_ERR_ARG, "cij[%d] NULL", c);
  for (int c = 0; c < 3; ++c) if (!buoy[c]) return fail(FWE_ERR_ARG, "buoy[%d] NULL", c);
  const double dt = x->cfg.dt;
  static const int shape_of[12] = {FWE_F_SXX, FWE_F_SXX, FWE_F_SXX, FWE_F_SXX, FWE_F_SXX, FWE_F_SXX,
                                   FWE_F_SYZ, FWE_F_SXZ, FWE_F_SXY, FWE_F_VX,  FWE_F_VY,  FWE_F_VZ};
  drop_graph(x);
  {
    double cmax = 0, bmax = 0;
    const size_t nc = (size_t)x->cfg.nx * x->cfg.ny * x->cfg.nz;
    for (int c = 0; c < 3; ++c)
      for (size_t q = 0; q < nc; ++q) cmax = std::max(cmax, (double)cij[c][q]);
    for (int c = 0; c < 3; ++c) {
      int64_t d[3];
      field_shape(x->cfg, c, d);
      for (int64_t q = 0; q < d[0] * d[1] * d[2]; ++q) bmax = std::max(bmax, (double)buoy[c][q]);
    }
    x->vmax = sqrt(cmax * bmax);
  }
  std::vector<std::vector<float>> h(12);
  for (int c = 0; c < 12; ++c) {
    int64_t d[3];
    field_shape(x->cfg, shape_of[c], d);
    pack(x, c < 9 ? cij[c] : buoy[c - 9], d, dt, h[c]);
    fwe_status st = upload(x, x->medium + (int64_t)c * x->nelem, h[c]);
    if (st) return st;
  }
  x->have_medium = true;
  // an isotropic medium run through the ortho law (the station solver's
  // file set, random-field cubes): the kernels can read two of the six
  // normal-moduli arrays
  const size_t nb = sizeof(float) * (size_t)x->nelem;
  x->iso_medium = !memcmp(h[0].data(), h[1].data(), nb) && !memcmp(h[0].data(), h[2].data(), nb) &&
                  !memcmp(h[3].data(), h[4].data(), nb) && !memcmp(h[3].data(), h[5].data(), nb);
  return build_material_table(x, h);
}

extern "C" fwe_status fwe_set_pml_profiles(fwe_ctx* x, const float* wx, const float* wy, const float* wz) {
  if (!x || !wx || !wy || !wz) return fail(FWE_ERR_ARG, "NULL argument");
  const fwe_config& c = x->cfg;
  if (c.pml != FWE_PML_PROFILE && c.pml != FWE_PML_SEPARABLE)
    return fail(FWE_ERR_STATE, "context damping is not PROFILE or SEPARABLE");
  const int add = c.pml == FWE_PML_PROFILE ? 1 : 0;
  const int n[3] = {c.nx, c.ny, c.nz};
  const float* w[3] = {wx, wy, wz};
  // device table: [wx (nx+1)][wy (ny+1)][wz (nz+1)][whx (nx+1)][why][whz]
  std::vector<float> h(2 * ((size_t)c.nx + c.ny + c.nz + 3), 1.f);
  size_t o = 0;
  for (int a = 0; a < 3; ++a) {
    for (int q = 0; q < n[a] + add; ++q) h[o + q] = w[a][q];
    o += n[a] + 1;
  }
  for (int a = 0; a < 3; ++a) {
    if (add)  // OK4.cu:307-309: (DAT)0.5*(w[i+1]+w[i])
      for (int q = 0; q < n[a]; ++q) h[o + q] = 0.5f * (w[a][q + 1] + w[a][q]);
    o += n[a] + 1;
  }
  // on the context stream, after the steps already queued there
  fwe_status st = upload(x, x->prof, h);
  if (st) return st;
  x->have_damp = true;
  return FWE_OK;
}

extern "C" fwe_status fwe_set_pml_weights(fwe_ctx* x, const float* w) {
  if (!x || !w) return fail(FWE_ERR_ARG, "NULL argument");
  if (x->cfg.pml != FWE_PML_WEIGHTS3D) return fail(FWE_ERR_STATE, "context damping is not WEIGHTS3D");
  int64_t d[3];
  field_shape(x->cfg, FWE_F_SXX, d);
  std::vector<float> h;
  pack(x, w, d, 1.0, h);
  fwe_status st = upload(x, x->w3, h);
  if (st) return st;
  x->have_damp = true;
  return FWE_OK;
}

// ---------------------------------------------------------------------------
// source and receivers
// ---------------------------------------------------------------------------
extern "C" fwe_status fwe_set_cpml(fwe_ctx* x, int32_t layer, const float* tx, const float* ty, const float* tz) {
  if (!x || !tx || !ty || !tz) return fail(FWE_ERR_ARG, "NULL argument");
  const fwe_config& c = x->cfg;
  if (c.pml != FWE_PML_CPML) return fail(FWE_ERR_STATE, "context damping is not CPML");
  if (layer < 1 || 2 * layer > c.nx || 2 * layer > c.ny || 2 * layer > c.nz)
    return fail(FWE_ERR_ARG, "CPML layer %d: need 1 <= layer and 2*layer <= nx, ny, nz", layer);
  HIPCHK(hipSetDevice(c.device));
  HIPCHK(hipStreamSynchronize(x->stream));  // steps still queued read the old state
  drop_graph(x);
  const int64_t L2 = 2 * (int64_t)layer;
  const int64_t pv[3] = {L2 * c.ny * c.nz, (int64_t)c.nx * L2 * c.nz, (int64_t)c.nx * c.ny * L2};
  const int64_t ps = 6 * (pv[0] + pv[1] + pv[2]);
  // runs are refused until the new layer is complete, so a failure below
  // never leaves a context that steps with a missing or mis-sized state
  x->have_damp = false;
  if (layer != x->cpml_L || !x->psi) {
    hipFree(x->psi);
    x->psi = nullptr;
    x->cpml_L = 0;
    x->pstride = 0;
    hipError_t e = hipMalloc(&x->psi, sizeof(float) * (size_t)ps * c.nbatch);
    if (e != hipSuccess) {
      x->psi = nullptr;
      return fail(e == hipErrorOutOfMemory ? FWE_ERR_NOMEM : FWE_ERR_HIP, "CPML memory: %s", hipGetErrorString(e));
    }
  }
  // the sizes always describe the allocation (reset and the state calls use them)
  x->cpml_L = layer;
  for (int d = 0; d < 3; ++d) x->pvar[d] = pv[d];
  x->pstride = ps;
  if (!x->ctab) HIPCHK(hipMalloc(&x->ctab, sizeof(float) * 6 * ((size_t)c.nx + c.ny + c.nz + 3)));
  std::vector<float> h;
  h.reserve(6 * ((size_t)c.nx + c.ny + c.nz + 3));
  h.insert(h.end(), tx, tx + 6 * ((size_t)c.nx + 1));
  h.insert(h.end(), ty, ty + 6 * ((size_t)c.ny + 1));
  h.insert(h.end(), tz, tz + 6 * ((size_t)c.nz + 1));
  fwe_status st = upload(x, x->ctab, h);
  if (st) return st;
  // the memory variables start from zero with the new coefficients
  HIPCHK(hipMemsetAsync(x->psi, 0, sizeof(float) * (size_t)ps * c.nbatch, x->stream));
  HIPCHK(hipStreamSynchronize(x->stream));
  if ((st = cpml_split_update(x))) return st;
  x->have_damp = true;
  return FWE_OK;
}

extern "C" fwe_status fwe_cpml_state_size(const fwe_ctx* x, int64_t* n) {
  if (!x || !n) return fail(FWE_ERR_ARG, "NULL argument");
  if (!x->psi) return fail(FWE_ERR_STATE, "no CPML set");
  *n = x->pstride;
  return FWE_OK;
}

extern "C" fwe_status fwe_get_cpml_state(fwe_ctx* x, int32_t batch, float* out) {
  if (!x || !out) return fail(FWE_ERR_ARG, "NULL argument");
  if (!x->psi) return fail(FWE_ERR_STATE, "no CPML set");
  if (batch < 0 || batch >= x->cfg.nbatch) return fail(FWE_ERR_ARG, "bad batch");
  HIPCHK(hipSetDevice(x->cfg.device));
  HIPCHK(hipStreamSynchronize(x->stream));
  HIPCHK(hipMemcpy(out, x->psi + (int64_t)batch * x->pstride, sizeof(float) * (size_t)x->pstride,
                   hipMemcpyDeviceToHost));
  return FWE_OK;
}

extern "C" fwe_status fwe_set_cpml_state(fwe_ctx* x, int32_t batch, const float* in) {
  if (!x || !in) return fail(FWE_ERR_ARG, "NULL argument");
  if (!x->psi) return fail(FWE_ERR_STATE, "no CPML set");
  if (batch < 0 || batch >= x->cfg.nbatch) return fail(FWE_ERR_ARG, "bad batch");
  std::vector<float> h(in, in + x->pstride);
  return upload(x, x->psi + (int64_t)batch * x->pstride, h);
}

extern "C" fwe_status fwe_set_stf(fwe_ctx* x, int32_t batch, const float* stf, int32_t n) {
  if (!x || (!stf && n > 0) || n < 0) return fail(FWE_ERR_ARG, "bad stf arguments");
  if (batch < 0 || batch >= x->cfg.nbatch) return fail(FWE_ERR_ARG, "batch %d out of range", batch);
  HIPCHK(hipSetDevice(x->cfg.device));
  HIPCHK(hipStreamSynchronize(x->stream));
  if (n > x->nstf) {
    // grow, keeping the other batches' functions
    float* nb = nullptr;
    HIPCHK(hipMalloc(&nb, sizeof(float) * (size_t)n * x->cfg.nbatch));
    HIPCHK(hipMemset(nb, 0, sizeof(float) * (size_t)n * x->cfg.nbatch));
    for (int b = 0; b < x->cfg.nbatch && x->stf; ++b)
      HIPCHK(hipMemcpy(nb + (size_t)b * n, x->stf + (size_t)b * x->nstf, sizeof(float) * x->nstf,
                       hipMemcpyDeviceToDevice));
    hipFree(x->stf);
    x->stf = nb;
    x->nstf = n;
    drop_graph(x);
  }
  std::vector<float> h((size_t)x->nstf, 0.f);
  if (n) memcpy(h.data(), stf, sizeof(float) * n);
  HIPCHK(hipMemcpy(x->stf + (size_t)batch * x->nstf, h.data(), sizeof(float) * h.size(),
                   hipMemcpyHostToDevice));
  if (batch == 0) x->stf_host.assign(stf, stf + n);
  return FWE_OK;
}

// sort point entries by xy tile and upload with their offsets
// (tile, entry) pairs -> per-tile CSR lists on the device
static fwe_status upload_tiled(fwe_ctx* x, std::vector<std::pair<int, PointEntry>> tv, int32_t** off,
                               PointEntry** dev) {
  drop_graph(x);
  const int ntile = x->ntx * x->nty;
  std::vector<int32_t> counts(ntile + 1, 0);
  std::stable_sort(tv.begin(), tv.end(),
                   [](const std::pair<int, PointEntry>& a, const std::pair<int, PointEntry>& b) {
                     return a.first < b.first;
                   });
  std::vector<PointEntry> v;
  for (auto& p : tv) {
    counts[p.first + 1]++;
    v.push_back(p.second);
  }
  for (int t = 0; t < ntile; ++t) counts[t + 1] += counts[t];
  HIPCHK(hipSetDevice(x->cfg.device));
  HIPCHK(hipStreamSynchronize(x->stream));
  hipFree(*off);
  hipFree(*dev);
  *off = nullptr;
  *dev = nullptr;
  HIPCHK(hipMalloc(off, sizeof(int32_t) * counts.size()));
  HIPCHK(hipMemcpy(*off, counts.data(), sizeof(int32_t) * counts.size(), hipMemcpyHostToDevice));
  HIPCHK(hipMalloc(dev, sizeof(PointEntry) * std::max<size_t>(1, v.size())));
  if (!v.empty())
    HIPCHK(hipMemcpy(*dev, v.data(), sizeof(PointEntry) * v.size(), hipMemcpyHostToDevice));
  return FWE_OK;
}

// each entry in the list of the tile that owns its cell
static fwe_status upload_points(fwe_ctx* x, const std::vector<PointEntry>& v, int32_t** off, PointEntry** dev) {
  std::vector<std::pair<int, PointEntry>> tv;
  for (auto& p : v) tv.push_back({p.i / x->tx + (p.j / x->ty) * x->ntx, p});
  return upload_tiled(x, tv, off, dev);
}

// receivers sorted by (xy tile, plane), with per-tile offsets (single-pass
// kernel) and per-(tile, plane) offsets (two-pass velocity kernel: a
// workgroup then scans only its own plane's receivers, not the whole
// column's at every plane)
static fwe_status upload_receivers(fwe_ctx* x) {
  drop_graph(x);
  const int64_t ntile = (int64_t)x->ntx * x->nty, nz = x->cfg.nz;
  std::vector<std::pair<int64_t, PointEntry>> tv;
  for (auto& p : x->rec_host) tv.push_back({(int64_t)(p.i / x->tx + (p.j / x->ty) * x->ntx) * nz + p.k, p});
  std::stable_sort(tv.begin(), tv.end(),
                   [](const std::pair<int64_t, PointEntry>& a, const std::pair<int64_t, PointEntry>& b) {
                     return a.first < b.first;
                   });
  std::vector<int32_t> fine((size_t)(ntile * nz + 1), 0), coarse((size_t)ntile + 1, 0);
  std::vector<PointEntry> v;
  for (auto& p : tv) {
    fine[(size_t)p.first + 1]++;
    coarse[(size_t)(p.first / nz) + 1]++;
    v.push_back(p.second);
  }
  for (size_t t = 1; t < fine.size(); ++t) fine[t] += fine[t - 1];
  for (size_t t = 1; t < coarse.size(); ++t) coarse[t] += coarse[t - 1];
  HIPCHK(hipSetDevice(x->cfg.device));
  HIPCHK(hipStreamSynchronize(x->stream));
  hipFree(x->rec_off);
  hipFree(x->rec_koff);
  hipFree(x->rec);
  x->rec_off = x->rec_koff = nullptr;
  x->rec = nullptr;
  HIPCHK(hipMalloc(&x->rec_off, sizeof(int32_t) * coarse.size()));
  HIPCHK(hipMemcpy(x->rec_off, coarse.data(), sizeof(int32_t) * coarse.size(), hipMemcpyHostToDevice));
  HIPCHK(hipMalloc(&x->rec_koff, sizeof(int32_t) * fine.size()));
  HIPCHK(hipMemcpy(x->rec_koff, fine.data(), sizeof(int32_t) * fine.size(), hipMemcpyHostToDevice));
  HIPCHK(hipMalloc(&x->rec, sizeof(PointEntry) * std::max<size_t>(1, v.size())));
  if (!v.empty()) HIPCHK(hipMemcpy(x->rec, v.data(), sizeof(PointEntry) * v.size(), hipMemcpyHostToDevice));
  return FWE_OK;
}

extern "C" fwe_status fwe_set_injections(fwe_ctx* x, int32_t n, const fwe_injection* inj) {
  if (!x || n < 0 || (n > 0 && !inj)) return fail(FWE_ERR_ARG, "bad injection arguments");
  const fwe_config& c = x->cfg;
  std::vector<PointEntry> v;
  for (int q = 0; q < n; ++q) {
    const fwe_injection& s = inj[q];
    if (s.comp < FWE_SXX || s.comp > FWE_SYZ) return fail(FWE_ERR_ARG, "injection %d: bad component %d", q, s.comp);
    if (s.batch < 0 || s.batch >= c.nbatch) return fail(FWE_ERR_ARG, "injection %d: bad batch", q);
    // the stress kernel covers cells i<nx, j<ny, k<nz of every component
    if (s.ix < 0 || s.iy < 0 || s.iz < 0 || s.ix >= c.nx || s.iy >= c.ny || s.iz >= c.nz)
      return fail(FWE_ERR_ARG, "injection %d at (%d,%d,%d) outside %dx%dx%d", q, s.ix, s.iy, s.iz, c.nx, c.ny, c.nz);
    v.push_back(PointEntry{s.ix, s.iy, s.iz, s.comp, s.batch, s.coef, -1, 0});
  }
  fwe_status st = upload_points(x, v, &x->inj_off, &x->inj);
  if (st) return st;
  x->inj_host = v;
  x->ninj = n;
  x->sw_dirty = true;
  return FWE_OK;
}

extern "C" fwe_status fwe_set_receivers(fwe_ctx* x, int32_t n, const fwe_receiver* rec) {
  if (!x || n < 0 || (n > 0 && !rec)) return fail(FWE_ERR_ARG, "bad receiver arguments");
  const fwe_config& c = x->cfg;
  std::vector<PointEntry> v;
  for (int r = 0; r < n; ++r) {
    const fwe_receiver& s = rec[r];
    if (s.comp < 0 || s.comp > 2) return fail(FWE_ERR_ARG, "receiver %d: bad component", r);
    if (s.batch < 0 || s.batch >= c.nbatch) return fail(FWE_ERR_ARG, "receiver %d: bad batch", r);
    if (s.ix < 0 || s.iy < 0 || s.iz < 0 || s.ix >= c.nx || s.iy >= c.ny || s.iz >= c.nz)
      return fail(FWE_ERR_ARG, "receiver %d at (%d,%d,%d) outside %dx%dx%d", r, s.ix, s.iy, s.iz, c.nx, c.ny, c.nz);
    v.push_back(PointEntry{s.ix, s.iy, s.iz, s.comp, s.batch, 0.f, r, 0});
  }
  const std::vector<PointEntry> old = x->rec_host;
  x->rec_host = v;
  x->sw_dirty = true;
  fwe_status st = upload_receivers(x);
  if (st) {
    x->rec_host = old;
    return st;
  }
  hipFree(x->traces);
  x->traces = nullptr;
  x->nrec = n;
  const size_t tb = sizeof(float) * std::max<size_t>(1, (size_t)n * std::max(1, c.nt));
  HIPCHK(hipMalloc(&x->traces, tb));
  HIPCHK(hipMemset(x->traces, 0, tb));
  return FWE_OK;
}

extern "C" fwe_status fwe_mt_injections(int32_t layout, int32_t xs, int32_t ys, int32_t zs,
                                        const float m[6], int32_t batch, fwe_injection out[6]) {
  if (!m || !out) return fail(FWE_ERR_ARG, "NULL argument");
  // offsets of Sxx Syy Szz Sxy Sxz Syz relative to (xs,ys,zs)
  static const int ok4[6][3] = {{1, 1, 1}, {1, 1, 1}, {1, 1, 1}, {3, 3, 1}, {3, 3, 1}, {2, 2, 1}};
  static const int an[6][3] = {{1, 1, 1}, {1, 1, 1}, {1, 1, 1}, {2, 2, 1}, {-1, 1, 1}, {3, 2, 1}};
  if (layout != FWE_SRC_LAYOUT_OK4 && layout != FWE_SRC_LAYOUT_ANISO) return fail(FWE_ERR_ARG, "bad layout");
  const int(*o)[3] = layout == FWE_SRC_LAYOUT_OK4 ? ok4 : an;
  for (int c = 0; c < 6; ++c) {
    out[c].comp = c;
    out[c].ix = xs + o[c][0];
    out[c].iy = ys + o[c][1];
    out[c].iz = zs + o[c][2];
    out[c].batch = batch;
    out[c].coef = m[c];
  }
  return FWE_OK;
}

// ---------------------------------------------------------------------------
// time loop
// ---------------------------------------------------------------------------
static bool use_table(const fwe_ctx* x);

// Wavefields each workgroup advances together.  Default: all of them with
// the medium arrays (one medium load per cell serves the batch), one with
// the material table (its medium costs a byte per cell, and single-field
// workgroups stream at the plain kernel's rate).
static int batch_tile(const fwe_ctx* x) {
  if (x->batch_tile > 0) return x->batch_tile;
  return use_table(x) ? 1 : x->cfg.nbatch;
}

static StepArgs make_args(fwe_ctx* x, int from) {
  const fwe_config& c = x->cfg;
  StepArgs a;
  memset(&a, 0, sizeof(a));
  a.nx = c.nx; a.ny = c.ny; a.nz = c.nz;
  a.nb = batch_tile(x);
  a.nbg = c.nbatch / a.nb;
  a.sy = x->sy; a.sz = x->sz; a.fstride = 9 * x->nelem;
  a.tile_x = x->tx;
  a.xcd_order = x->wg_order;
  a.split = (x->ntx * x->nty + 7) / 8;
  for (int q = 0; q < 3; ++q) a.V[q] = x->fbuf[from] + (int64_t)q * x->nelem;
  for (int q = 0; q < 6; ++q) a.S[q] = x->fbuf[from] + (int64_t)(3 + q) * x->nelem;
  if (c.mode == FWE_MODE_ORTHO) {
    for (int q = 0; q < 9; ++q) a.C[q] = x->medium + (int64_t)q * x->nelem;
    for (int q = 0; q < 3; ++q) a.B[q] = x->medium + (int64_t)(9 + q) * x->nelem;
  } else {
    a.C[0] = x->medium;
    a.C[1] = x->medium + x->nelem;
  }
  if (x->ntab > 0) { a.mid = x->mid; a.mtab = x->mtab; }
  a.c66dt = x->c66dt; a.bxdt = x->bxdt; a.bydt = x->bydt; a.bzdt = x->bzdt;
  a.vac_k = c.free_surface == FWE_FS_VACUUM ? 1 : -1;
  a.vac_cdt = (float)(FWE_VACUUM_SHEAR_MODULUS * c.dt);
  const size_t ox = 0, oy = c.nx + 1, oz = oy + c.ny + 1, oh = oz + c.nz + 1;
  a.wx = x->prof + ox; a.wy = x->prof + oy; a.wz = x->prof + oz;
  a.whx = x->prof + oh + ox; a.why = x->prof + oh + oy; a.whz = x->prof + oh + oz;
  a.w3 = x->w3;
  if (c.pml == FWE_PML_CPML) {
    a.cL = x->cpml_L;
    a.ctab[0] = x->ctab;
    a.ctab[1] = x->ctab + 6 * ((int64_t)c.nx + 1);
    a.ctab[2] = a.ctab[1] + 6 * ((int64_t)c.ny + 1);
    a.psi[0] = x->psi;
    a.psi[1] = x->psi + 6 * x->pvar[0];
    a.psi[2] = a.psi[1] + 6 * x->pvar[1];
    for (int d = 0; d < 3; ++d) a.pvar[d] = x->pvar[d];
    a.pstride = x->pstride;
  }
  a.rdx = (float)(1.0 / c.dx); a.rdy = (float)(1.0 / c.dy); a.rdz = (float)(1.0 / c.dz);
  a.dtdx = (float)(c.dt / c.dx); a.dtdy = (float)(c.dt / c.dy); a.dtdz = (float)(c.dt / c.dz);
  a.ty_rows = x->ty; a.ntx = x->ntx; a.nty = x->nty; a.nkc = x->nkc; a.kchunk = x->kchunk;
  a.bx0 = 0; a.bnx = x->ntx; a.by0 = 0; a.bny = x->nty; a.bk0 = 0; a.bnk = x->nkc;
  a.tiles = nullptr; a.ntiles = 0;
  a.inj_off = x->inj_off; a.inj = x->inj; a.stf = x->stf; a.nstf = x->nstf;
  a.src_after_bc = c.src_after_bc;
  a.rec_off = x->rec_off; a.rec_koff = x->rec_koff; a.rec = x->rec; a.traces = x->traces; a.nt = c.nt;
  return a;
}

// Point lists of the sweep kernel's tiles (62 columns x sw_rows rows): an
// injection goes to every tile whose footprint grown by one cell in x and y
// holds it (those tiles recompute its halo stresses), a receiver to the
// (tile, plane) that owns its cell.  Both keep the list order of the host
// copies, so sums of several injections into one cell round as in the
// two-pass kernels.
static fwe_status sweep_upload(fwe_ctx* x) {
  if (!x->sw_dirty) return FWE_OK;
  drop_graph(x);
  const int W = fwe::SWEEP_X, R = x->sw_rows, nz = x->cfg.nz;
  const int64_t ntile = (int64_t)x->sw_ntx * x->sw_nty;
  std::vector<std::vector<PointEntry>> per((size_t)ntile);
  for (const PointEntry& p : x->inj_host)
    for (int tx = std::max(0, (p.i - W) / W); tx < x->sw_ntx; ++tx) {
      const int lo = tx * W - 1, hi = tx * W + W;
      if (p.i < lo) break;
      if (p.i > hi) continue;
      for (int ty = std::max(0, (p.j - R) / R); ty < x->sw_nty; ++ty) {
        const int jl = ty * R - 1, jh = ty * R + R;
        if (p.j < jl) break;
        if (p.j <= jh) per[(size_t)(tx + (int64_t)ty * x->sw_ntx)].push_back(p);
      }
    }
  std::vector<int32_t> ioff((size_t)ntile + 1, 0);
  std::vector<PointEntry> iv;
  for (int64_t t = 0; t < ntile; ++t) {
    iv.insert(iv.end(), per[(size_t)t].begin(), per[(size_t)t].end());
    ioff[(size_t)t + 1] = (int32_t)iv.size();
  }
  std::vector<std::pair<int64_t, PointEntry>> rv;
  for (const PointEntry& p : x->rec_host) rv.push_back({(int64_t)(p.i / W + (p.j / R) * x->sw_ntx) * nz + p.k, p});
  std::stable_sort(rv.begin(), rv.end(), [](const std::pair<int64_t, PointEntry>& a,
                                            const std::pair<int64_t, PointEntry>& b) { return a.first < b.first; });
  std::vector<int32_t> roff((size_t)(ntile * nz + 1), 0);
  std::vector<PointEntry> rl;
  for (auto& q : rv) {
    roff[(size_t)q.first + 1]++;
    rl.push_back(q.second);
  }
  for (size_t t = 1; t < roff.size(); ++t) roff[t] += roff[t - 1];
  HIPCHK(hipSetDevice(x->cfg.device));
  HIPCHK(hipStreamSynchronize(x->stream));
  hipFree(x->sw_inj_off); hipFree(x->sw_inj); hipFree(x->sw_rec_koff); hipFree(x->sw_rec);
  x->sw_inj_off = x->sw_rec_koff = nullptr;
  x->sw_inj = x->sw_rec = nullptr;
  HIPCHK(hipMalloc(&x->sw_inj_off, sizeof(int32_t) * ioff.size()));
  HIPCHK(hipMemcpy(x->sw_inj_off, ioff.data(), sizeof(int32_t) * ioff.size(), hipMemcpyHostToDevice));
  HIPCHK(hipMalloc(&x->sw_inj, sizeof(PointEntry) * std::max<size_t>(1, iv.size())));
  if (!iv.empty()) HIPCHK(hipMemcpy(x->sw_inj, iv.data(), sizeof(PointEntry) * iv.size(), hipMemcpyHostToDevice));
  HIPCHK(hipMalloc(&x->sw_rec_koff, sizeof(int32_t) * roff.size()));
  HIPCHK(hipMemcpy(x->sw_rec_koff, roff.data(), sizeof(int32_t) * roff.size(), hipMemcpyHostToDevice));
  HIPCHK(hipMalloc(&x->sw_rec, sizeof(PointEntry) * std::max<size_t>(1, rl.size())));
  if (!rl.empty()) HIPCHK(hipMemcpy(x->sw_rec, rl.data(), sizeof(PointEntry) * rl.size(), hipMemcpyHostToDevice));
  x->sw_dirty = false;
  return FWE_OK;
}

// the sweep kernel addresses one array with 32-bit byte offsets
static bool sweep_fits(const fwe_ctx* x) { return (uint64_t)x->nelem * sizeof(float) < ((uint64_t)1 << 32); }

// sweep arguments: all wavefields, time level fbuf[from] -> fbuf[from ^ 1]
static fwe::SweepArgs make_sweep(fwe_ctx* x, int from) {
  const StepArgs s = make_args(x, from);
  fwe::SweepArgs f;
  memset(&f, 0, sizeof(f));
  f.nx = s.nx; f.ny = s.ny; f.nz = s.nz;
  f.sy = (int32_t)s.sy; f.sz = (int32_t)s.sz;
  f.fstride = 9 * x->nelem;
  f.nbatch = x->cfg.nbatch;
  for (int q = 0; q < 3; ++q) {
    f.Vi[q] = x->fbuf[from] + (int64_t)q * x->nelem;
    f.Vo[q] = x->fbuf[from ^ 1] + (int64_t)q * x->nelem;
  }
  for (int q = 0; q < 6; ++q) {
    f.Si[q] = x->fbuf[from] + (int64_t)(3 + q) * x->nelem;
    f.So[q] = x->fbuf[from ^ 1] + (int64_t)(3 + q) * x->nelem;
  }
  for (int q = 0; q < 9; ++q) f.C[q] = s.C[q];
  for (int q = 0; q < 3; ++q) f.B[q] = s.B[q];
  f.mid = s.mid; f.mtab = s.mtab;
  f.ntab = s.mtab ? x->ntab : 0;
  f.c66dt = s.c66dt; f.bxdt = s.bxdt; f.bydt = s.bydt; f.bzdt = s.bzdt;
  f.vac_k = s.vac_k; f.vac_cdt = s.vac_cdt;
  f.wx = s.wx; f.wy = s.wy; f.wz = s.wz; f.whx = s.whx; f.why = s.why; f.whz = s.whz; f.w3 = s.w3;
  f.rdx = s.rdx; f.rdy = s.rdy; f.rdz = s.rdz;
  f.rows = x->sw_rows;
  f.ntx = x->sw_ntx; f.nty = x->sw_nty; f.nkc = x->sw_nkc; f.kchunk = x->sw_kchunk;
  f.nyb = (x->sw_nty + 7) / 8;
  f.inj_off = x->sw_inj_off; f.inj = x->sw_inj;
  f.stf = x->stf; f.nstf = x->nstf;
  f.src_after_bc = s.src_after_bc;
  f.rec_koff = x->sw_rec_koff; f.rec = x->sw_rec; f.traces = x->traces; f.nt = s.nt;
  return f;
}

// the single-pass sweep kernel runs when asked for, or under AUTO when
// fwe_autotune measured it faster than the two passes
static bool use_fused(const fwe_ctx* x) {
  if (x->cfg.pml == FWE_PML_CPML) return false;  // the two-pass kernels only
  return x->kernel == FWE_KERNEL_FUSED || (x->kernel == FWE_KERNEL_AUTO && x->auto_sweep);
}

// one step from time level fbuf[from]; step index it (+ *it_base); `mid`
// (optional) is recorded between the stress and velocity passes
static hipError_t lb_before(fwe_ctx* x, int from);
static hipError_t lb_after(fwe_ctx* x, int from);

// the two-pass kernels read the medium through the material table when one
// was built, unless the arrays were asked for
static bool use_table(const fwe_ctx* x) {
  return x->ntab > 0 && (x->medium_enc == FWE_MEDIUM_AUTO || x->medium_enc == FWE_MEDIUM_TABLE);
}
// otherwise an isotropic medium reads eight of the twelve arrays
static bool use_iso(const fwe_ctx* x) {
  return x->cfg.mode == FWE_MODE_ORTHO && x->iso_medium && !use_table(x) &&
         (x->medium_enc == FWE_MEDIUM_AUTO || x->medium_enc == FWE_MEDIUM_ISOTROPIC);
}
static int kernel_mode(const fwe_ctx* x) {
  if (use_table(x)) return x->cfg.mode == FWE_MODE_ISO ? fwe::MODE_ISO_TAB : fwe::MODE_ORTHO_TAB;
  if (use_iso(x)) return fwe::MODE_ORTHO_ISO;
  return x->cfg.mode;
}

// The stress and velocity passes of one step, `mid` (optional) recorded
// between them.  With the convolutional PML and a box of slab-free tiles
// (not in plane-split order, which covers the whole grid), each pass is two
// launches: the box through the plain kernel, the listed tiles through the
// CPML one; cells in no slab compute the same bits in either.
static hipError_t two_pass(fwe_ctx* x, const StepArgs& a, hipEvent_t mid) {
  const fwe_config& c = x->cfg;
  const int km = kernel_mode(x);
  const int fs = c.free_surface == FWE_FS_BC1;
  hipError_t e = hipSuccess;
  if (c.pml == FWE_PML_CPML && x->cpml_box[1] > 0 && x->wg_order != FWE_ORDER_PLANE_SPLIT) {
    StepArgs in = a, sh = a;
    in.bx0 = x->cpml_box[0]; in.bnx = x->cpml_box[1];
    in.by0 = x->cpml_box[2]; in.bny = x->cpml_box[3];
    in.bk0 = x->cpml_box[4]; in.bnk = x->cpml_box[5];
    sh.tiles = x->cpml_tiles;
    sh.ntiles = x->cpml_ntiles;
    e = fwe::launch_stress(in, km, FWE_PML_NONE, fs, x->stream);
    if (e == hipSuccess && sh.ntiles > 0) e = fwe::launch_stress(sh, km, FWE_PML_CPML, fs, x->stream);
    if (e == hipSuccess && mid) e = hipEventRecord(mid, x->stream);
    if (e == hipSuccess) e = fwe::launch_velocity(in, km, FWE_PML_NONE, x->stream);
    if (e == hipSuccess && sh.ntiles > 0) e = fwe::launch_velocity(sh, km, FWE_PML_CPML, x->stream);
    return e;
  }
  e = fwe::launch_stress(a, km, c.pml, fs, x->stream);
  if (e == hipSuccess && mid) e = hipEventRecord(mid, x->stream);
  if (e == hipSuccess) e = fwe::launch_velocity(a, km, c.pml, x->stream);
  return e;
}

static hipError_t step_launch(fwe_ctx* x, int from, int it, const int32_t* it_base, hipEvent_t mid) {
  const fwe_config& c = x->cfg;
  const int km = kernel_mode(x);
  hipError_t e = hipSuccess;
  if (x->lb) {
    // learned boundary: corrections from the fields before the step, added
    // after it (two-pass kernels only, checked when attached)
    e = lb_before(x, from);
    if (e == hipSuccess) {
      StepArgs a = make_args(x, from);
      a.it = it;
      a.it_base = it_base;
      e = two_pass(x, a, mid);
    }
    if (e == hipSuccess) e = lb_after(x, from);
    return e;
  }
  if (use_fused(x)) {
    // one launch for the whole step and every wavefield; `mid` (the event
    // between the passes of the two-pass schedule) follows it
    fwe::SweepArgs f = make_sweep(x, from);
    f.it = it;
    f.it_base = it_base;
    e = fwe::launch_sweep(f, km, c.pml, c.free_surface == FWE_FS_BC1, x->stream);
    if (e == hipSuccess && mid) e = hipEventRecord(mid, x->stream);
  } else {
    StepArgs a = make_args(x, from);
    a.it = it;
    a.it_base = it_base;
    e = two_pass(x, a, mid);
  }
  return e;
}

static fwe_status ready(fwe_ctx* x) {
  if (!x->have_medium) return fail(FWE_ERR_STATE, "medium not set");
  if (x->medium_enc == FWE_MEDIUM_TABLE && !use_table(x))
    return fail(FWE_ERR_UNSUPPORTED, "material table asked for, but the medium has more than %d distinct tuples",
                FWE_MEDIUM_TABLE_MAX);
  if (x->medium_enc == FWE_MEDIUM_ISOTROPIC && !use_iso(x))
    return fail(FWE_ERR_UNSUPPORTED, "isotropic medium arrays asked for, but the medium is not isotropic");
  if (x->cfg.pml != FWE_PML_NONE && !x->have_damp) return fail(FWE_ERR_STATE, "damping not set");
  return FWE_OK;
}

extern "C" fwe_status fwe_reset(fwe_ctx* x) {
  if (!x) return fail(FWE_ERR_ARG, "NULL ctx");
  HIPCHK(hipSetDevice(x->cfg.device));
  for (int b = 0; b < 2; ++b)
    if (x->fbuf[b]) HIPCHK(hipMemsetAsync(x->fbuf[b], 0, field_bytes(x), x->stream));
  x->cur = 0;
  if (x->traces)
    HIPCHK(hipMemsetAsync(x->traces, 0, sizeof(float) * std::max<size_t>(1, (size_t)x->nrec * std::max(1, x->cfg.nt)), x->stream));
  if (x->psi) HIPCHK(hipMemsetAsync(x->psi, 0, sizeof(float) * (size_t)x->pstride * x->cfg.nbatch, x->stream));
  HIPCHK(hipStreamSynchronize(x->stream));
  x->it = 0;
  return FWE_OK;
}

static fwe_status run_impl(fwe_ctx* x, int32_t nsteps, double* ms_s, double* ms_v, double* ms_w) {
  if (!x || nsteps < 0) return fail(FWE_ERR_ARG, "bad run arguments");
  fwe_status st = ready(x);
  if (st) return st;
  HIPCHK(hipSetDevice(x->cfg.device));
  const bool fused = use_fused(x);
  if (fused && !sweep_fits(x))
    return fail(FWE_ERR_UNSUPPORTED, "the single-pass kernel needs arrays under 4 GiB; use the two-pass kernels");
  if (fused && (st = sweep_upload(x))) return st;
  if (fused && !x->fbuf[1]) {
    // second time level, starting as a copy so cells no kernel updates agree
    HIPCHK(hipMalloc(&x->fbuf[1], field_bytes(x)));
    HIPCHK(hipMemcpyAsync(x->fbuf[1], x->fbuf[x->cur], field_bytes(x), hipMemcpyDeviceToDevice, x->stream));
  }
  const bool timed = ms_s || ms_v || ms_w;
  if (timed && (int)x->events.size() < 2 * nsteps + 1) {
    while ((int)x->events.size() < 2 * nsteps + 1) {
      hipEvent_t ev;
      HIPCHK(hipEventCreate(&ev));
      x->events.push_back(ev);
    }
  }
  const int flip = fused ? 1 : 0;
  int g = x->graph_steps;
  if (fused && (g & 1)) ++g;  // a replay must return to the buffer it started from
  if (!timed && g > 0 && nsteps >= g) {
    // replay a captured G-step graph; the kernels read the step index from
    // it_dev, which each replay advances by G
    if (x->gexec && x->graph_from != x->cur) drop_graph(x);
    if (!x->gexec) {
      hipGraph_t graph = nullptr;
      HIPCHK(hipStreamBeginCapture(x->stream, hipStreamCaptureModeThreadLocal));
      hipError_t e = hipSuccess;
      int from = x->cur;
      for (int s = 0; s < g && e == hipSuccess; ++s) {
        e = step_launch(x, from, s, x->it_dev, nullptr);
        from ^= flip;
      }
      if (e == hipSuccess) e = fwe::launch_advance(x->it_dev, g, 0, x->stream);
      hipError_t e2 = hipStreamEndCapture(x->stream, &graph);
      if (e == hipSuccess) e = e2;
      if (e == hipSuccess) e = hipGraphInstantiate(&x->gexec, graph, nullptr, nullptr, 0);
      if (graph) hipGraphDestroy(graph);
      if (e != hipSuccess) {
        x->gexec = nullptr;
        return fail(FWE_ERR_HIP, "time-loop graph capture: %s", hipGetErrorString(e));
      }
      x->graph_from = x->cur;
    }
    HIPCHK(fwe::launch_advance(x->it_dev, x->it, 1, x->stream));
    while (nsteps >= g) {
      HIPCHK(hipGraphLaunch(x->gexec, x->stream));
      x->it += g;
      nsteps -= g;
    }
  }
  if (timed) HIPCHK(hipEventRecord(x->events[0], x->stream));
  for (int s = 0; s < nsteps; ++s) {
    HIPCHK(step_launch(x, x->cur, x->it, nullptr, timed ? x->events[2 * s + 1] : nullptr));
    if (timed) HIPCHK(hipEventRecord(x->events[2 * s + 2], x->stream));
    x->cur ^= flip;
    x->it++;
  }
  if (timed) {
    HIPCHK(hipEventSynchronize(x->events[2 * nsteps]));
    double ts = 0, tv = 0;
    for (int s = 0; s < nsteps; ++s) {
      float a1 = 0, a2 = 0;
      HIPCHK(hipEventElapsedTime(&a1, x->events[2 * s], x->events[2 * s + 1]));
      HIPCHK(hipEventElapsedTime(&a2, x->events[2 * s + 1], x->events[2 * s + 2]));
      ts += a1;
      tv += a2;
    }
    float wall = 0;
    HIPCHK(hipEventElapsedTime(&wall, x->events[0], x->events[2 * nsteps]));
    if (ms_s) *ms_s = ts;
    if (ms_v) *ms_v = tv;
    if (ms_w) *ms_w = wall;
  }
  return FWE_OK;
}

extern "C" fwe_status fwe_run(fwe_ctx* x, int32_t nsteps) { return run_impl(x, nsteps, nullptr, nullptr, nullptr); }

extern "C" fwe_status fwe_set_graph_steps(fwe_ctx* x, int32_t g) {
  if (!x || g < 0 || g > 4096) return fail(FWE_ERR_ARG, "graph steps must be in 0..4096");
  drop_graph(x);
  x->graph_steps = g;
  return FWE_OK;
}

// The convolutional PML's tile split for the current tiling and layer: the
// box of tiles whose cells all lie outside every slab, and the list of the
// others (x fastest, then y, then chunk).  No box when the layer or the
// tiling leaves no such tile.
static fwe_status cpml_split_update(fwe_ctx* x) {
  for (int q = 0; q < 6; ++q) x->cpml_box[q] = 0;
  x->cpml_ntiles = 0;
  const fwe_config& c = x->cfg;
  const int L = x->cpml_L;
  if (c.pml != FWE_PML_CPML || L <= 0 || x->ntx <= 0) return FWE_OK;
  auto range = [](int n, int t, int L, int& lo, int& cnt) {
    lo = (L + t - 1) / t;           // first tile with every cell at index >= L
    const int hi = (n - L) / t;     // tiles below hi have every cell below n - L
    cnt = std::max(0, hi - lo);
  };
  int b[6];
  range(c.nx, x->tx, L, b[0], b[1]);
  range(c.ny, x->ty, L, b[2], b[3]);
  range(c.nz, x->kchunk, L, b[4], b[5]);
  if (b[1] <= 0 || b[3] <= 0 || b[5] <= 0) return FWE_OK;
  std::vector<int32_t> list;
  for (int kc = 0; kc < x->nkc; ++kc)
    for (int ty = 0; ty < x->nty; ++ty)
      for (int tx = 0; tx < x->ntx; ++tx) {
        const bool in = tx >= b[0] && tx < b[0] + b[1] && ty >= b[2] && ty < b[2] + b[3] && kc >= b[4] &&
                        kc < b[4] + b[5];
        if (!in) list.push_back(tx + (ty + kc * x->nty) * x->ntx);
      }
  HIPCHK(hipSetDevice(c.device));
  HIPCHK(hipStreamSynchronize(x->stream));  // queued steps read the old list
  hipFree(x->cpml_tiles);
  x->cpml_tiles = nullptr;
  HIPCHK(hipMalloc(&x->cpml_tiles, sizeof(int32_t) * std::max<size_t>(1, list.size())));
  HIPCHK(hipMemcpy(x->cpml_tiles, list.data(), sizeof(int32_t) * list.size(), hipMemcpyHostToDevice));
  x->cpml_ntiles = (int32_t)list.size();
  for (int q = 0; q < 6; ++q) x->cpml_box[q] = b[q];
  return FWE_OK;
}

static fwe_status apply_tiling(fwe_ctx* x, int tx, int ty, int kchunk) {
  const fwe_config& c = x->cfg;
  if (ty != 2 && ty != 4 && ty != 8) return fail(FWE_ERR_ARG, "tile rows must be 2, 4 or 8");
  if (tx != 64 && tx != 128) return fail(FWE_ERR_ARG, "tile columns must be 64 or 128");
  // the free surface (BC1 of plane 1 written after plane 2) needs planes
  // 0..2 in one chunk
  const int kmin = c.free_surface == FWE_FS_BC1 ? 3 : 1;
  if (kchunk < kmin || kchunk > c.nz) return fail(FWE_ERR_ARG, "z chunk must be in %d..nz", kmin);
  x->tx = tx;
  x->ty = ty;
  x->ntx = (c.nx + tx - 1) / tx;
  x->nty = (c.ny + ty - 1) / ty;
  x->kchunk = kchunk;
  x->nkc = (c.nz + kchunk - 1) / kchunk;
  fwe_status st = upload_points(x, x->inj_host, &x->inj_off, &x->inj);
  if (st) return st;
  if ((st = cpml_split_update(x))) return st;
  return upload_receivers(x);
}

extern "C" fwe_status fwe_set_kernel(fwe_ctx* x, int32_t kind) {
  if (!x) return fail(FWE_ERR_ARG, "NULL ctx");
  if (kind != FWE_KERNEL_AUTO && kind != FWE_KERNEL_TWO_PASS && kind != FWE_KERNEL_FUSED)
    return fail(FWE_ERR_ARG, "bad kernel kind %d", kind);
  if (kind == FWE_KERNEL_FUSED && x->lb)
    return fail(FWE_ERR_UNSUPPORTED, "the single-pass kernel has no learned boundary");
  if (kind == FWE_KERNEL_FUSED && x->cfg.pml == FWE_PML_CPML)
    return fail(FWE_ERR_UNSUPPORTED, "the single-pass kernel has no convolutional PML");
  if (kind == FWE_KERNEL_FUSED && !sweep_fits(x))
    return fail(FWE_ERR_UNSUPPORTED, "the single-pass kernel needs arrays under 4 GiB");
  drop_graph(x);
  x->kernel = kind;
  return FWE_OK;
}

extern "C" fwe_status fwe_get_kernel(const fwe_ctx* x, int32_t* kind) {
  if (!x || !kind) return fail(FWE_ERR_ARG, "NULL argument");
  *kind = use_fused(x) ? FWE_KERNEL_FUSED : FWE_KERNEL_TWO_PASS;
  return FWE_OK;
}

static fwe_status apply_sweep_tiling(fwe_ctx* x, int rows, int kchunk) {
  const fwe_config& c = x->cfg;
  if (rows != 2 && rows != 4) return fail(FWE_ERR_ARG, "sweep rows per lane must be 2 or 4");
  const int kmin = c.free_surface == FWE_FS_BC1 ? 3 : 1;
  if (kchunk < std::min(kmin, c.nz) || kchunk > c.nz)
    return fail(FWE_ERR_ARG, "sweep z chunk must be in %d..nz", kmin);
  drop_graph(x);
  x->sw_rows = rows;
  x->sw_kchunk = kchunk;
  x->sw_ntx = (c.nx + fwe::SWEEP_X - 1) / fwe::SWEEP_X;
  x->sw_nty = (c.ny + rows - 1) / rows;
  x->sw_nkc = (c.nz + kchunk - 1) / kchunk;
  x->sw_dirty = true;
  return FWE_OK;
}

extern "C" fwe_status fwe_set_sweep_tiling(fwe_ctx* x, int32_t rows, int32_t zchunk) {
  if (!x) return fail(FWE_ERR_ARG, "NULL ctx");
  return apply_sweep_tiling(x, rows, zchunk);
}

extern "C" fwe_status fwe_get_sweep_tiling(const fwe_ctx* x, int32_t* rows, int32_t* zchunk) {
  if (!x || !rows || !zchunk) return fail(FWE_ERR_ARG, "NULL argument");
  *rows = x->sw_rows;
  *zchunk = x->sw_kchunk;
  return FWE_OK;
}

extern "C" fwe_status fwe_set_tiling(fwe_ctx* x, int32_t tile_rows, int32_t zchunk) {
  if (!x) return fail(FWE_ERR_ARG, "NULL ctx");
  return apply_tiling(x, x->tx, tile_rows, zchunk);
}

extern "C" fwe_status fwe_set_tile_cols(fwe_ctx* x, int32_t cols) {
  if (!x) return fail(FWE_ERR_ARG, "NULL ctx");
  return apply_tiling(x, cols, x->ty, x->kchunk);
}

extern "C" fwe_status fwe_set_wg_order(fwe_ctx* x, int32_t order) {
  if (!x) return fail(FWE_ERR_ARG, "NULL ctx");
  if (order < FWE_ORDER_LINEAR || order > FWE_ORDER_PLANE_SPLIT) return fail(FWE_ERR_ARG, "bad workgroup order %d", order);
  drop_graph(x);
  x->wg_order = order;
  return FWE_OK;
}

extern "C" fwe_status fwe_get_wg_order(const fwe_ctx* x, int32_t* order) {
  if (!x || !order) return fail(FWE_ERR_ARG, "NULL argument");
  *order = x->wg_order;
  return FWE_OK;
}

extern "C" fwe_status fwe_set_batch_tile(fwe_ctx* x, int32_t wavefields) {
  if (!x) return fail(FWE_ERR_ARG, "NULL ctx");
  if (wavefields < 0 || wavefields > x->cfg.nbatch || (wavefields > 0 && x->cfg.nbatch % wavefields))
    return fail(FWE_ERR_ARG, "wavefields per workgroup must divide the batch of %d (0: default)", x->cfg.nbatch);
  drop_graph(x);
  x->batch_tile = wavefields;
  return FWE_OK;
}

extern "C" fwe_status fwe_get_batch_tile(const fwe_ctx* x, int32_t* wavefields) {
  if (!x || !wavefields) return fail(FWE_ERR_ARG, "NULL argument");
  *wavefields = batch_tile(x);
  return FWE_OK;
}

extern "C" fwe_status fwe_set_medium_encoding(fwe_ctx* x, int32_t encoding) {
  if (!x) return fail(FWE_ERR_ARG, "NULL ctx");
  if (encoding < FWE_MEDIUM_AUTO || encoding > FWE_MEDIUM_ISOTROPIC)
    return fail(FWE_ERR_ARG, "bad medium encoding %d", encoding);
  if (encoding == FWE_MEDIUM_ISOTROPIC && x->cfg.mode != FWE_MODE_ORTHO)
    return fail(FWE_ERR_UNSUPPORTED, "the isotropic arrays are a form of the ortho law's medium");
  if (encoding == FWE_MEDIUM_TABLE && x->have_medium && x->ntab == 0)
    return fail(FWE_ERR_UNSUPPORTED, "the medium has more than %d distinct tuples", FWE_MEDIUM_TABLE_MAX);
  if (encoding == FWE_MEDIUM_ISOTROPIC && x->have_medium && !x->iso_medium)
    return fail(FWE_ERR_UNSUPPORTED, "the medium is not isotropic (C11 = C22 = C33, C12 = C13 = C23)");
  drop_graph(x);
  x->medium_enc = encoding;
  return FWE_OK;
}

extern "C" fwe_status fwe_get_medium_encoding(const fwe_ctx* x, int32_t* encoding, int32_t* n_materials) {
  if (!x || !encoding) return fail(FWE_ERR_ARG, "NULL argument");
  *encoding = use_table(x) ? FWE_MEDIUM_TABLE
              : use_iso(x)   ? FWE_MEDIUM_ISOTROPIC
                             : FWE_MEDIUM_ARRAYS;
  if (n_materials) *n_materials = x->ntab;
  return FWE_OK;
}

extern "C" fwe_status fwe_get_tile_cols(const fwe_ctx* x, int32_t* cols) {
  if (!x || !cols) return fail(FWE_ERR_ARG, "NULL argument");
  *cols = x->tx;
  return FWE_OK;
}

extern "C" fwe_status fwe_get_tiling(const fwe_ctx* x, int32_t* tile_rows, int32_t* zchunk) {
  if (!x || !tile_rows || !zchunk) return fail(FWE_ERR_ARG, "NULL argument");
  *tile_rows = x->ty;
  *zchunk = x->kchunk;
  return FWE_OK;
}

extern "C" fwe_status fwe_autotune(fwe_ctx* x, int32_t trial_steps, double* best_ms_per_step) {
  if (!x || trial_steps < 1) return fail(FWE_ERR_ARG, "bad autotune arguments");
  fwe_status st = ready(x);
  if (st) return st;
  const int nz = x->cfg.nz;
  // a candidate: the two-pass schedule (sweep = false: batch layout,
  // workgroup order, tile and z chunk) or the single-pass sweep (rows per
  // lane in ty, z chunk)
  struct Cand {
    double ms;
    bool sweep;
    int bt, order, tx, ty, kc;
  };
  std::vector<Cand> cands;
  const bool try_two = x->kernel != FWE_KERNEL_FUSED;
  const bool try_sweep = x->kernel != FWE_KERNEL_TWO_PASS && !x->lb && sweep_fits(x) && x->cfg.pml != FWE_PML_CPML;
  // batch layouts: all wavefields per workgroup, or one
  std::vector<int> bts(1, x->cfg.nbatch);
  if (x->cfg.nbatch > 1) bts.push_back(1);
  static const int txs[2] = {64, 128};
  static const int tys[3] = {2, 4, 8};
  static const int kcs[10] = {1, 2, 4, 8, 12, 16, 24, 32, 48, 64};
  static const int skcs[7] = {4, 8, 12, 16, 24, 32, 64};
  const int kmin = x->cfg.free_surface == FWE_FS_BC1 ? 3 : 1;
  drop_graph(x);
  const int kernel0 = x->kernel;
  // time one candidate: a warm-up step, then `steps` timed steps; ms per step
  // is the sum of the kernel times (launch gaps do not depend on the tiling)
  auto trial = [&](const Cand& c, int steps, double* ms) -> fwe_status {
    fwe_status s2;
    if (c.sweep) {
      x->kernel = FWE_KERNEL_FUSED;
      s2 = apply_sweep_tiling(x, c.ty, c.kc);
    } else {
      x->kernel = FWE_KERNEL_TWO_PASS;
      x->wg_order = c.order;
      x->batch_tile = c.bt;
      s2 = apply_tiling(x, c.tx, c.ty, c.kc);
    }
    if (!s2) s2 = run_impl(x, 1, nullptr, nullptr, nullptr);
    double a = 0, b = 0, w = 0;
    if (!s2) s2 = run_impl(x, steps, &a, &b, &w);
    x->kernel = kernel0;
    if (s2) return s2;
    *ms = (a + b) / steps;
    return FWE_OK;
  };
  if (try_two)
    for (int bt : bts)
      for (int order = FWE_ORDER_LINEAR; order <= FWE_ORDER_PLANE_SPLIT; ++order)
        for (int tx : txs)
          for (int ty : tys)
            for (int q = 0; q <= 10; ++q) {
              // the listed chunk heights, then the whole column (so a short
              // grid always has a candidate)
              const int kc = q < 10 ? kcs[q] : nz;
              if (kc > nz || kc < kmin || (q == 10 && std::find(kcs, kcs + 10, nz) != kcs + 10)) continue;
              Cand c{0.0, false, bt, order, tx, ty, kc};
              if ((st = trial(c, trial_steps, &c.ms))) return st;
              cands.push_back(c);
            }
  if (try_sweep)
    for (int rows = 2; rows <= 4; rows += 2)
      for (int q = 0; q <= 7; ++q) {
        const int kc = q < 7 ? skcs[q] : nz;
        if (kc > nz || kc < kmin || (q == 7 && std::find(skcs, skcs + 7, nz) != skcs + 7)) continue;
        Cand c{0.0, true, 0, 0, 64, rows, kc};
        if ((st = trial(c, trial_steps, &c.ms))) return st;
        cands.push_back(c);
      }
  if (cands.empty()) return fail(FWE_ERR_UNSUPPORTED, "no kernel schedule fits this context");
  // short trials rank candidates within a few percent of each other by
  // noise: time the best few again, longer, and keep the fastest of those
  std::sort(cands.begin(), cands.end(), [](const Cand& p, const Cand& q) { return p.ms < q.ms; });
  const size_t nfinal = std::min<size_t>(4, cands.size());
  Cand best = cands[0];
  double best_ms = 1e30;
  for (size_t q = 0; q < nfinal; ++q) {
    double ms = 0;
    if ((st = trial(cands[q], 4 * trial_steps, &ms))) return st;
    if (ms < best_ms) { best_ms = ms; best = cands[q]; }
  }
  if (best.sweep) {
    if ((st = apply_sweep_tiling(x, best.ty, best.kc))) return st;
  } else {
    x->wg_order = best.order;
    x->batch_tile = best.bt;
    if ((st = apply_tiling(x, best.tx, best.ty, best.kc))) return st;
  }
  if (x->kernel == FWE_KERNEL_AUTO) x->auto_sweep = best.sweep;
  if (best_ms_per_step) *best_ms_per_step = best_ms;
  return fwe_reset(x);
}

extern "C" fwe_status fwe_run_timed(fwe_ctx* x, int32_t nsteps, double* ms_stress, double* ms_velocity,
                                    double* ms_wall) {
  double a = 0, b = 0, w = 0;
  fwe_status st = run_impl(x, nsteps, &a, &b, &w);
  if (st) return st;
  if (ms_stress) *ms_stress = a;
  if (ms_velocity) *ms_velocity = b;
  if (ms_wall) *ms_wall = w;
  return FWE_OK;
}

extern "C" fwe_status fwe_synchronize(fwe_ctx* x) {
  if (!x) return fail(FWE_ERR_ARG, "NULL ctx");
  HIPCHK(hipSetDevice(x->cfg.device));
  HIPCHK(hipStreamSynchronize(x->stream));
  return FWE_OK;
}

extern "C" fwe_status fwe_get_step(const fwe_ctx* x, int32_t* it) {
  if (!x || !it) return fail(FWE_ERR_ARG, "NULL argument");
  *it = x->it;
  return FWE_OK;
}

extern "C" fwe_status fwe_set_step(fwe_ctx* x, int32_t it) {
  if (!x || it < 0) return fail(FWE_ERR_ARG, "bad step");
  x->it = it;
  return FWE_OK;
}

// ---------------------------------------------------------------------------
// results
// ---------------------------------------------------------------------------
extern "C" fwe_status fwe_get_traces(fwe_ctx* x, float* out) {
  if (!x || (!out && x->nrec > 0)) return fail(FWE_ERR_ARG, "NULL argument");
  if (x->nrec == 0 || x->cfg.nt == 0) return FWE_OK;
  HIPCHK(hipSetDevice(x->cfg.device));
  HIPCHK(hipStreamSynchronize(x->stream));
  const size_t n = (size_t)x->nrec * x->cfg.nt;
  HIPCHK(hipMemcpy(out, x->traces, sizeof(float) * n, hipMemcpyDeviceToHost));
  for (size_t q = 0; q < n; ++q)
    if (!isfinite(out[q])) return fail(FWE_ERR_NONFINITE, "trace sample %zu is not finite", q);
  return FWE_OK;
}

// Arguments and receiver layout of a trace combination: the receiver list
// must be nbatch repeats of the same nr receivers.  *nr_out = nr.
static fwe_status combine_check(const fwe_ctx* x, int32_t n, const float* coef, const float* out, int* nr_out) {
  if (!x || n < 0 || (n > 0 && (!coef || !out))) return fail(FWE_ERR_ARG, "bad combine arguments");
  const int nb = x->cfg.nbatch;
  if (x->nrec == 0 || x->nrec % nb) return fail(FWE_ERR_STATE, "receivers must be %d equal per-wavefield sets", nb);
  const int nr = x->nrec / nb;
  for (int b = 0; b < nb; ++b)
    for (int r = 0; r < nr; ++r) {
      const PointEntry& p = x->rec_host[(size_t)b * nr + r];
      const PointEntry& p0 = x->rec_host[r];
      if (p.batch != b || p.comp != p0.comp || p.i != p0.i || p.j != p0.j || p.k != p0.k)
        return fail(FWE_ERR_STATE, "receiver %d of wavefield %d does not repeat receiver %d of wavefield 0",
                    b * nr + r, b, r);
    }
  *nr_out = nr;
  return FWE_OK;
}

extern "C" fwe_status fwe_combine_traces(fwe_ctx* x, int32_t n, const float* coef, float* out) {
  int nr = 0;
  const fwe_status cs = combine_check(x, n, coef, out, &nr);
  if (cs != FWE_OK) return cs;
  const int nb = x->cfg.nbatch, nt = x->cfg.nt;
  if (n == 0 || nt == 0) return FWE_OK;
  HIPCHK(hipSetDevice(x->cfg.device));
  const int64_t per = (int64_t)nr * nt;
  // output staged through a bounded device buffer (<= 256 MiB or one tensor)
  const int mc = (int)std::max<int64_t>(1, std::min<int64_t>(n, (int64_t(256) << 20) / (4 * per)));
  float *dc = nullptr, *dout = nullptr;
  fwe_status st = FWE_OK;
  hipError_t e = hipMalloc(&dc, sizeof(float) * (size_t)mc * nb);
  if (e == hipSuccess) e = hipMalloc(&dout, sizeof(float) * (size_t)mc * per);
  for (int m0 = 0; m0 < n && e == hipSuccess; m0 += mc) {
    const int m = std::min(mc, n - m0);
    e = hipMemcpyAsync(dc, coef + (size_t)m0 * nb, sizeof(float) * (size_t)m * nb, hipMemcpyHostToDevice,
                       x->stream);
    if (e == hipSuccess) e = fwe::launch_combine(x->traces, nb, per, dc, m, dout, x->stream);
    if (e == hipSuccess)
      e = hipMemcpyAsync(out + (size_t)m0 * per, dout, sizeof(float) * (size_t)m * per, hipMemcpyDeviceToHost,
                         x->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(x->stream);
  }
  if (e != hipSuccess) st = fail(FWE_ERR_HIP, "combine: %s", hipGetErrorString(e));
  hipFree(dc);
  hipFree(dout);
  return st;
}

// ---------------------------------------------------------------------------
// trace combination without allocation
// ---------------------------------------------------------------------------
// Staging buffer of fwe_combine_traces_noalloc: the coefficients of a chunk
// of sources, then (from a 256-byte boundary) their combined traces.  Each
// device holds its own copy at its own address; host code never uses the
// array's name as a device address, only the address hipGetSymbolAddress
// returned for the current device (noalloc_scratch below).
static constexpr int64_t NOALLOC_FLOATS = FWE_NOALLOC_SCRATCH_BYTES / sizeof(float);
static __device__ __align__(256) float scratch[NOALLOC_FLOATS];

static constexpr int NOALLOC_MAX_DEVICES = 64;
static std::mutex g_scratch_mu;
static float* g_scratch_ptr[NOALLOC_MAX_DEVICES];

// Device address of `scratch` on `dev`, looked up once per device with
// that device current (the caller has made it current).
static fwe_status noalloc_scratch(int dev, float** ptr) {
  if (dev < 0 || dev >= NOALLOC_MAX_DEVICES) return fail(FWE_ERR_ARG, "device %d out of range", dev);
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (!g_scratch_ptr[dev]) {
    void* p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p)
      return fail(FWE_ERR_HIP, "static staging buffer on device %d: %s", dev, hipGetErrorString(e));
    g_scratch_ptr[dev] = static_cast<float*>(p);
  }
  *ptr = g_scratch_ptr[dev];
  return FWE_OK;
}

extern "C" fwe_status fwe_combine_traces_noalloc(fwe_ctx* x, int32_t n, const float* coef, float* out) {
  int nr = 0;
  const fwe_status cs = combine_check(x, n, coef, out, &nr);
  if (cs != FWE_OK) return cs;
  const int nb = x->cfg.nbatch, nt = x->cfg.nt;
  if (n == 0 || nt == 0) return FWE_OK;
  const int64_t per = (int64_t)nr * nt;
  // a chunk of mc sources: coefficients at [0, cb), traces at [cb, cb + mc*per)
  auto need = [&](int64_t mc) { return round_up(mc * nb, 64) + mc * per; };
  if (need(1) > NOALLOC_FLOATS)
    return fail(FWE_ERR_ARG, "one combined source needs %lld floats; the static staging buffer holds %lld",
                (long long)need(1), (long long)NOALLOC_FLOATS);
  int64_t mc = std::min<int64_t>(n, NOALLOC_FLOATS / (nb + per));
  while (mc > 1 && need(mc) > NOALLOC_FLOATS) --mc;
  HIPCHK(hipSetDevice(x->cfg.device));
  float* base = nullptr;
  const fwe_status ss = noalloc_scratch(x->cfg.device, &base);
  if (ss != FWE_OK) return ss;
  float* dc = base;
  float* dout = base + round_up(mc * nb, 64);
  hipError_t e = hipSuccess;
  for (int m0 = 0; m0 < n && e == hipSuccess; m0 += (int)mc) {
    const int m = (int)std::min<int64_t>(mc, n - m0);
    e = hipMemcpyAsync(dc, coef + (size_t)m0 * nb, sizeof(float) * (size_t)m * nb, hipMemcpyHostToDevice,
                       x->stream);
    if (e == hipSuccess) e = fwe::launch_combine(x->traces, nb, per, dc, m, dout, x->stream);
    if (e == hipSuccess)
      e = hipMemcpyAsync(out + (size_t)m0 * per, dout, sizeof(float) * (size_t)m * per, hipMemcpyDeviceToHost,
                         x->stream);
    // the next chunk overwrites the staging buffer
    if (e == hipSuccess) e = hipStreamSynchronize(x->stream);
  }
  if (e != hipSuccess) return fail(FWE_ERR_HIP, "combine (no allocation): %s", hipGetErrorString(e));
  return FWE_OK;
}

extern "C" fwe_status fwe_get_field(fwe_ctx* x, int32_t batch, int32_t field, float* out) {
  if (!x || !out) return fail(FWE_ERR_ARG, "NULL argument");
  if (batch < 0 || batch >= x->cfg.nbatch || field < 0 || field > 8) return fail(FWE_ERR_ARG, "bad batch/field");
  HIPCHK(hipSetDevice(x->cfg.device));
  HIPCHK(hipStreamSynchronize(x->stream));
  std::vector<float> h((size_t)x->nelem);
  HIPCHK(hipMemcpy(h.data(), fields(x) + ((int64_t)batch * 9 + field) * x->nelem, sizeof(float) * h.size(),
                   hipMemcpyDeviceToHost));
  int64_t d[3];
  field_shape(x->cfg, field, d);
  unpack(x, h, d, out);
  return FWE_OK;
}

extern "C" fwe_status fwe_set_field(fwe_ctx* x, int32_t batch, int32_t field, const float* in) {
  if (!x || !in) return fail(FWE_ERR_ARG, "NULL argument");
  if (batch < 0 || batch >= x->cfg.nbatch || field < 0 || field > 8) return fail(FWE_ERR_ARG, "bad batch/field");
  int64_t d[3];
  field_shape(x->cfg, field, d);
  std::vector<float> h;
  pack(x, in, d, 1.0, h);
  // both time-level buffers: cells no kernel updates must agree in both
  for (int b = 0; b < 2; ++b)
    if (x->fbuf[b]) {
      fwe_status st = upload(x, x->fbuf[b] + ((int64_t)batch * 9 + field) * x->nelem, h);
      if (st) return st;
    }
  return FWE_OK;
}

// ---------------------------------------------------------------------------
// legacy file contract (OK4.cu:153-199, 401-457, 521-532)
// ---------------------------------------------------------------------------
static std::string path_of(const char* dir, const char* name) {
  std::string p = (dir && dir[0]) ? dir : ".";
  if (p.back() != '/') p += '/';
  return p + name;
}

// read exactly n float32 values; FWE_ERR_IO on a missing or short file
// (the reference would fread(NULL) and crash, OK4.cu:190)
static bool file_exists(const char* dir, const char* name) {
  FILE* f = fopen(path_of(dir, name).c_str(), "rb");
  if (f) fclose(f);
  return f != nullptr;
}

static fwe_status read_f32(const char* dir, const char* name, size_t n, std::vector<float>& out, size_t* got) {
  const std::string p = path_of(dir, name);
  FILE* f = fopen(p.c_str(), "rb");
  if (!f) return fail(FWE_ERR_IO, "cannot open %s", p.c_str());
  out.assign(n, 0.f);
  const size_t r = n ? fread(out.data(), sizeof(float), n, f) : 0;
  fclose(f);
  if (got) *got = r;
  if (!got && r != n) return fail(FWE_ERR_IO, "%s: read %zu of %zu float32 values", p.c_str(), r, n);
  return FWE_OK;
}

extern "C" fwe_status fwe_read_legacy_params(const char* dir, fwe_legacy_params* out) {
  if (!out) return fail(FWE_ERR_ARG, "NULL argument");
  memset(out, 0, sizeof(*out));
  std::vector<float> v;
  size_t got = 0;
  fwe_status st = read_f32(dir, "pa1.dat", 10, v, &got);
  if (st) return st;
  if (got < 4) return fail(FWE_ERR_IO, "pa1.dat holds %zu values, need >= 4", got);
  out->n_pa1 = (int)got;
  for (size_t q = 0; q < got; ++q) out->pa1[q] = v[q];
  st = read_f32(dir, "pa2.dat", 13, v, &got);
  if (st) return st;
  out->n_pa2 = (int)got;
  for (size_t q = 0; q < got; ++q) out->pa2[q] = v[q];
  st = read_f32(dir, "pa3.dat", 12, v, &got);
  if (st) return st;
  if (got < 3) return fail(FWE_ERR_IO, "pa3.dat holds %zu values, need >= 3", got);
  out->n_pa3 = (int)got;
  for (size_t q = 0; q < got; ++q) out->pa3[q] = v[q];
  if (read_f32(dir, "MTS.dat", 6, v, &got) == FWE_OK && got == 6) {
    out->has_mts = 1;
    for (int q = 0; q < 6; ++q) out->mts[q] = v[q];
  }
  g_err[0] = 0;
  return FWE_OK;
}

extern "C" fwe_status fwe_load_legacy_inputs(fwe_ctx* x, const char* dir, const fwe_legacy_params* p) {
  if (!x || !p) return fail(FWE_ERR_ARG, "NULL argument");
  const fwe_config& c = x->cfg;
  const size_t nc = (size_t)c.nx * c.ny * c.nz;
  std::vector<float> a, b;
  fwe_status st;
  if (c.mode == FWE_MODE_ISO) {
    if ((st = read_f32(dir, "c11u_het.dat", nc, a, nullptr))) return st;
    if ((st = read_f32(dir, "c12u_het.dat", nc, b, nullptr))) return st;
    if (p->n_pa2 < 8) return fail(FWE_ERR_IO, "pa2.dat lacks c66u (index 7)");
    st = fwe_set_material_iso(x, a.data(), b.data(), (float)p->pa2[7], (float)p->pa3[0], (float)p->pa3[1],
                              (float)p->pa3[2]);
    if (st) return st;
    if (c.pml == FWE_PML_PROFILE) {
      std::vector<float> wx, wy, wz;
      if ((st = read_f32(dir, "weights_x.dat", c.nx + 1, wx, nullptr))) return st;
      if ((st = read_f32(dir, "weights_y.dat", c.ny + 1, wy, nullptr))) return st;
      if ((st = read_f32(dir, "weights_z.dat", c.nz + 1, wz, nullptr))) return st;
      if ((st = fwe_set_pml_profiles(x, wx.data(), wy.data(), wz.data()))) return st;
    }
  } else {
    static const char* cn[9] = {"C11m.dat", "C22m.dat", "C33m.dat", "C12m.dat", "C13m.dat",
                                "C23m.dat", "C44myz.dat", "C55mxz.dat", "C66mxy.dat"};
    // the iso-het file set of the station solver (9km.cu:381-400): the same
    // update with C11=C22=C33 = c11u_het, C12=C13=C23 = c12u_het and the
    // shear moduli on their staggered edges
    static const char* hn[9] = {"c11u_het.dat", "c11u_het.dat", "c11u_het.dat", "c12u_het.dat", "c12u_het.dat",
                                "c12u_het.dat", "c44u_hetyz.dat", "c44u_hetxz.dat", "c44u_hetxy.dat"};
    static const int shp[9] = {FWE_F_SXX, FWE_F_SXX, FWE_F_SXX, FWE_F_SXX, FWE_F_SXX,
                               FWE_F_SXX, FWE_F_SYZ, FWE_F_SXZ, FWE_F_SXY};
    static const char* bn[3] = {"vrho_x_11_hc.dat", "vrho_y_11_hc.dat", "vrho_z_11_hc.dat"};
    const bool isohet = !file_exists(dir, cn[0]) && file_exists(dir, hn[8]);
    const char* const* names = isohet ? hn : cn;
    std::vector<std::vector<float>> cv(9), bv(3);
    const float* cp[9];
    const float* bp[3];
    for (int q = 0; q < 9; ++q) {
      int64_t d[3];
      field_shape(c, shp[q], d);
      if ((st = read_f32(dir, names[q], (size_t)(d[0] * d[1] * d[2]), cv[q], nullptr))) return st;
      cp[q] = cv[q].data();
    }
    for (int q = 0; q < 3; ++q) {
      int64_t d[3];
      field_shape(c, q, d);
      if ((st = read_f32(dir, bn[q], (size_t)(d[0] * d[1] * d[2]), bv[q], nullptr))) return st;
      bp[q] = bv[q].data();
    }
    if ((st = fwe_set_material_ortho(x, cp, bp))) return st;
    if (c.pml == FWE_PML_WEIGHTS3D) {
      if ((st = read_f32(dir, "weights.dat", nc, a, nullptr))) return st;
      if ((st = fwe_set_pml_weights(x, a.data()))) return st;
    } else if (c.pml == FWE_PML_SEPARABLE) {
      std::vector<float> wx, wy, wz;
      if ((st = read_f32(dir, "taper_x.dat", c.nx, wx, nullptr))) return st;
      if ((st = read_f32(dir, "taper_y.dat", c.ny, wy, nullptr))) return st;
      if ((st = read_f32(dir, "taper_z.dat", c.nz, wz, nullptr))) return st;
      if ((st = fwe_set_pml_profiles(x, wx.data(), wy.data(), wz.data()))) return st;
    }
  }
  if ((st = read_f32(dir, "Src.dat", (size_t)c.nt, a, nullptr))) return st;
  for (int b2 = 0; b2 < c.nbatch; ++b2)
    if ((st = fwe_set_stf(x, b2, a.data(), c.nt))) return st;
  g_err[0] = 0;
  return FWE_OK;
}

static fwe_status write_f32(const char* dir, const char* name, const float* a, size_t n) {
  // "%d_%d_%s.res" with isave = 0 and rank 0 (OK4.cu:159-164)
  char fn[256];
  snprintf(fn, sizeof(fn), "0_0_%s.res", name);
  const std::string p = path_of(dir, fn);
  FILE* f = fopen(p.c_str(), "wb");
  if (!f) return fail(FWE_ERR_IO, "cannot write %s", p.c_str());
  const size_t w = n ? fwrite(a, sizeof(float), n, f) : 0;
  fclose(f);
  if (w != n) return fail(FWE_ERR_IO, "short write %s", p.c_str());
  return FWE_OK;
}

extern "C" fwe_status fwe_write_legacy_outputs(fwe_ctx* x, const char* dir, const char* const* rec_names,
                                               int32_t n_names) {
  if (!x) return fail(FWE_ERR_ARG, "NULL ctx");
  if (n_names < 0 || (n_names > 0 && !rec_names) || n_names > x->nrec)
    return fail(FWE_ERR_ARG, "rec_names must name at most the %d receivers", x->nrec);
  const fwe_config& c = x->cfg;
  {
    // save_info (OK4.cu:153-157): PRECIS nx ny nz NB_PARAMS ceil(niter/nout)
    const std::string p = path_of(dir, "0_infos.inf");
    FILE* f = fopen(p.c_str(), "w");
    if (!f) return fail(FWE_ERR_IO, "cannot write %s", p.c_str());
    fprintf(f, "%d %d %d %d %d %d", 4, c.nx, c.ny, c.nz, 1, (int)ceil(10.0f / 200.0f));
    fclose(f);
  }
  fwe_status st;
  std::vector<float> src((size_t)c.nt, 0.f);
  for (size_t q = 0; q < x->stf_host.size() && q < src.size(); ++q) src[q] = x->stf_host[q];
  if ((st = write_f32(dir, "Src", src.data(), src.size()))) return st;
  if (x->nrec > 0 && c.nt > 0) {
    std::vector<float> tr((size_t)x->nrec * c.nt);
    st = fwe_get_traces(x, tr.data());
    if (st && st != FWE_ERR_NONFINITE) return st;
    for (int r = 0; r < n_names; ++r)
      if ((st = write_f32(dir, rec_names[r], tr.data() + (size_t)r * c.nt, (size_t)c.nt))) return st;
  }
  static const char* fn[3] = {"Vx", "Vy", "Vz"};
  for (int f = 0; f < 3; ++f) {
    int64_t d[3];
    field_shape(c, f, d);
    std::vector<float> h((size_t)(d[0] * d[1] * d[2]));
    if ((st = fwe_get_field(x, 0, f, h.data()))) return st;
    if ((st = write_f32(dir, fn[f], h.data(), h.size()))) return st;
  }
  return FWE_OK;
}

// ---------------------------------------------------------------------------
// measurement
// ---------------------------------------------------------------------------
extern "C" fwe_status fwe_copy_probe(int32_t device, int64_t bytes, int32_t reps, double* ms_per_copy) {
  if (bytes < 1024 || reps < 1 || !ms_per_copy) return fail(FWE_ERR_ARG, "bad probe arguments");
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || device < 0 || device >= ndev)
    return fail(FWE_ERR_HIP, "no HIP device %d", device);
  HIPCHK(hipSetDevice(device));
  const int64_t n = bytes / 16 * 4;
  float *a = nullptr, *b = nullptr;
  HIPCHK(hipMalloc(&a, sizeof(float) * n));
  HIPCHK(hipMalloc(&b, sizeof(float) * n));
  hipStream_t s;
  HIPCHK(hipStreamCreateWithFlags(&s, hipStreamNonBlocking));
  HIPCHK(hipMemsetAsync(a, 0, sizeof(float) * n, s));
  HIPCHK(fwe::launch_copy(a, b, n, s));
  hipEvent_t e0, e1;
  HIPCHK(hipEventCreate(&e0));
  HIPCHK(hipEventCreate(&e1));
  HIPCHK(hipEventRecord(e0, s));
  for (int r = 0; r < reps; ++r) HIPCHK(fwe::launch_copy(a, b, n, s));
  HIPCHK(hipEventRecord(e1, s));
  HIPCHK(hipEventSynchronize(e1));
  float ms = 0;
  HIPCHK(hipEventElapsedTime(&ms, e0, e1));
  *ms_per_copy = ms / reps;
  hipEventDestroy(e0); hipEventDestroy(e1);
  hipStreamDestroy(s);
  hipFree(a); hipFree(b);
  return FWE_OK;
}

// ---------------------------------------------------------------------------
// groups: several independent single-wavefield models advanced together,
// one stress launch and one velocity launch per step for all of them
// ---------------------------------------------------------------------------
struct fwe_group {
  std::vector<fwe_ctx*> m;
  int device = 0, mode = 0, pml = 0, fs = 0, tx = 64, ty = 4, order = FWE_ORDER_LINEAR;
  hipStream_t stream = nullptr;
  StepArgs* tab = nullptr;
  int32_t* wg = nullptr;  // per workgroup: (model, workgroup within the model)
  int64_t wg_cap = 0;     // workgroups wg has room for
  std::vector<int32_t> wg_start;  // first workgroup of every model, as wg holds them
  int32_t* it_dev = nullptr;
  int nblk = 0;
};

// rebuild the device table of the members' current step arguments
static fwe_status group_table(fwe_group* g) {
  std::vector<StepArgs> tab;
  std::vector<int32_t> start(1, 0);
  for (fwe_ctx* x : g->m) {
    StepArgs a = make_args(x, x->cur);
    a.it = 0;
    a.it_base = g->it_dev;
    tab.push_back(a);
    start.push_back(start.back() + x->ntx * x->nty * x->nkc);
  }
  g->nblk = start.back();
  HIPCHK(hipSetDevice(g->device));
  HIPCHK(hipMemcpy(g->tab, tab.data(), sizeof(StepArgs) * tab.size(), hipMemcpyHostToDevice));
  if (start == g->wg_start) return FWE_OK;  // the tilings did not change
  std::vector<int32_t> wg(2 * (size_t)g->nblk);
  for (size_t m = 0; m + 1 < start.size(); ++m)
    for (int32_t b = start[m]; b < start[m + 1]; ++b) {
      wg[2 * (size_t)b] = (int32_t)m;
      wg[2 * (size_t)b + 1] = b - start[m];
    }
  g->wg_start.clear();
  if (g->nblk > g->wg_cap) {
    HIPCHK(hipFree(g->wg));
    g->wg = nullptr;
    g->wg_cap = 0;
    HIPCHK(hipMalloc(&g->wg, sizeof(int32_t) * wg.size()));
    g->wg_cap = g->nblk;
  }
  HIPCHK(hipMemcpy(g->wg, wg.data(), sizeof(int32_t) * wg.size(), hipMemcpyHostToDevice));
  g->wg_start = start;
  return FWE_OK;
}

extern "C" fwe_status fwe_group_create(int32_t n, fwe_ctx* const* members, fwe_group** out) {
  if (!members || !out || n < 1) return fail(FWE_ERR_ARG, "bad group arguments");
  *out = nullptr;
  const fwe_config& c0 = members[0]->cfg;
  for (int q = 0; q < n; ++q) {
    fwe_ctx* x = members[q];
    if (!x) return fail(FWE_ERR_ARG, "member %d is NULL", q);
    const fwe_config& c = x->cfg;
    if (c.device != c0.device || c.mode != c0.mode || c.pml != c0.pml || c.free_surface != c0.free_surface)
      return fail(FWE_ERR_ARG, "member %d differs in device, law, damping or free surface", q);
    if (c.nbatch != 1) return fail(FWE_ERR_UNSUPPORTED, "member %d has %d wavefields; groups take 1", q, c.nbatch);
    if (x->lb) return fail(FWE_ERR_UNSUPPORTED, "member %d has a learned boundary; groups run plain damping", q);
    if (c.pml == FWE_PML_CPML)
      return fail(FWE_ERR_UNSUPPORTED, "member %d uses the convolutional PML; groups run the profile forms", q);
    if (x->it != members[0]->it) return fail(FWE_ERR_STATE, "member %d is at another step", q);
    fwe_status st = ready(x);
    if (st) return st;
  }
  fwe_group* g = new (std::nothrow) fwe_group();
  if (!g) return fail(FWE_ERR_NOMEM, "host allocation");
  g->m.assign(members, members + n);
  g->device = c0.device; g->mode = c0.mode; g->pml = c0.pml; g->fs = c0.free_surface == FWE_FS_BC1;
  g->ty = members[0]->ty;
  g->tx = members[0]->tx;
  HIPCHK(hipSetDevice(g->device));
  for (fwe_ctx* x : g->m) {
    HIPCHK(hipStreamSynchronize(x->stream));
    if (x->ty != g->ty || x->tx != g->tx) {
      fwe_status st = apply_tiling(x, g->tx, g->ty, std::min(x->kchunk, x->cfg.nz));
      if (st) { delete g; return st; }
    }
  }
  hipError_t e = hipStreamCreateWithFlags(&g->stream, hipStreamNonBlocking);
  if (e == hipSuccess) e = hipMalloc(&g->tab, sizeof(StepArgs) * n);
  if (e == hipSuccess) e = hipMalloc(&g->it_dev, sizeof(int32_t));
  if (e != hipSuccess) {
    fwe_group_destroy(g);
    return fail(FWE_ERR_HIP, "group create: %s", hipGetErrorString(e));
  }
  fwe_status st = group_table(g);
  if (st) { fwe_group_destroy(g); return st; }
  *out = g;
  return FWE_OK;
}

extern "C" fwe_status fwe_group_destroy(fwe_group* g) {
  if (!g) return FWE_OK;
  hipSetDevice(g->device);
  if (g->stream) hipStreamSynchronize(g->stream);
  hipFree(g->tab); hipFree(g->wg); hipFree(g->it_dev);
  if (g->stream) hipStreamDestroy(g->stream);
  delete g;
  return FWE_OK;
}

static fwe_status group_run_impl(fwe_group* g, int32_t nsteps, double* ms) {
  if (!g || nsteps < 0) return fail(FWE_ERR_ARG, "bad group run arguments");
  HIPCHK(hipSetDevice(g->device));
  for (fwe_ctx* x : g->m) {
    if (x->it != g->m[0]->it) return fail(FWE_ERR_STATE, "members were advanced separately");
    if (x->ty != g->ty || x->tx != g->tx) return fail(FWE_ERR_STATE, "a member was re-tiled outside the group");
    if (x->lb) return fail(FWE_ERR_UNSUPPORTED, "a member has a learned boundary attached; groups run plain damping");
    HIPCHK(hipStreamSynchronize(x->stream));
  }
  fwe_status st = group_table(g);
  if (st) return st;
  hipEvent_t e0 = nullptr, e1 = nullptr;
  if (ms) {
    HIPCHK(hipEventCreate(&e0));
    HIPCHK(hipEventCreate(&e1));
  }
  // groups read the medium arrays: the eight of an isotropic medium when
  // every member's would be read that way, else the full set
  int gm = g->mode;
  if (gm == FWE_MODE_ORTHO) {
    bool iso = true;
    for (fwe_ctx* x : g->m) iso = iso && x->iso_medium && (x->medium_enc == FWE_MEDIUM_AUTO || x->medium_enc == FWE_MEDIUM_ISOTROPIC);
    if (iso) gm = fwe::MODE_ORTHO_ISO;
  }
  HIPCHK(fwe::launch_advance(g->it_dev, g->m[0]->it, 1, g->stream));
  if (ms) HIPCHK(hipEventRecord(e0, g->stream));
  for (int s = 0; s < nsteps; ++s) {
    HIPCHK(fwe::launch_group(g->tab, g->wg, g->nblk, g->tx, g->ty, g->order, gm, g->pml, g->fs, false,
                             g->stream));
    HIPCHK(fwe::launch_group(g->tab, g->wg, g->nblk, g->tx, g->ty, g->order, gm, g->pml, g->fs, true,
                             g->stream));
    HIPCHK(fwe::launch_advance(g->it_dev, 1, 0, g->stream));
  }
  if (ms) HIPCHK(hipEventRecord(e1, g->stream));
  HIPCHK(hipStreamSynchronize(g->stream));
  for (fwe_ctx* x : g->m) x->it += nsteps;
  if (ms) {
    float t = 0;
    HIPCHK(hipEventElapsedTime(&t, e0, e1));
    *ms = t;
    hipEventDestroy(e0);
    hipEventDestroy(e1);
  }
  return FWE_OK;
}

extern "C" fwe_status fwe_group_run(fwe_group* g, int32_t nsteps) { return group_run_impl(g, nsteps, nullptr); }

extern "C" fwe_status fwe_group_run_timed(fwe_group* g, int32_t nsteps, double* ms) {
  return group_run_impl(g, nsteps, ms);
}

extern "C" fwe_status fwe_group_set_tiling(fwe_group* g, int32_t tile_rows, int32_t zchunk) {
  if (!g) return fail(FWE_ERR_ARG, "NULL group");
  if (tile_rows != 2 && tile_rows != 4 && tile_rows != 8) return fail(FWE_ERR_ARG, "tile rows must be 2, 4 or 8");
  for (fwe_ctx* x : g->m)
    if (std::min<int>(zchunk, x->cfg.nz) < (x->cfg.free_surface == FWE_FS_BC1 ? 3 : 1))
      return fail(FWE_ERR_ARG, "z chunk too small for member with free surface");
  for (fwe_ctx* x : g->m) {
    fwe_status st = apply_tiling(x, g->tx, tile_rows, std::min<int>(zchunk, x->cfg.nz));
    if (st) return st;
  }
  g->ty = tile_rows;
  return group_table(g);
}

extern "C" fwe_status fwe_group_set_tile_cols(fwe_group* g, int32_t cols) {
  if (!g) return fail(FWE_ERR_ARG, "NULL group");
  if (cols != 64 && cols != 128) return fail(FWE_ERR_ARG, "tile columns must be 64 or 128");
  for (fwe_ctx* x : g->m) {
    fwe_status st = apply_tiling(x, cols, g->ty, x->kchunk);
    if (st) return st;
  }
  g->tx = cols;
  return group_table(g);
}

extern "C" fwe_status fwe_group_autotune(fwe_group* g, int32_t trial_steps, double* best_ms_per_step) {
  if (!g || trial_steps < 1) return fail(FWE_ERR_ARG, "bad autotune arguments");
  static const int tys[3] = {2, 4, 8};
  static const int kcs[8] = {1, 2, 4, 8, 16, 32, 64, 256};
  fwe_status st;
  const int kmin = g->fs ? 3 : 1;
  struct Cand {
    double ms;
    int order, tx, ty, kc;
  };
  auto trial = [&](const Cand& c, int steps, double* ms) -> fwe_status {
    g->order = c.order;
    g->tx = c.tx;
    fwe_status s2 = fwe_group_set_tiling(g, c.ty, c.kc);
    if (s2) return s2;
    double t = 0;
    if ((s2 = group_run_impl(g, 1, &t))) return s2;
    if ((s2 = group_run_impl(g, steps, &t))) return s2;
    *ms = t / steps;
    return FWE_OK;
  };
  std::vector<Cand> cands;
  for (int order = FWE_ORDER_LINEAR; order <= FWE_ORDER_XCD; ++order)
    for (int tx : {64, 128})
      for (int ty : tys)
        for (int kc : kcs) {
          if (kc < kmin) continue;
          Cand c{0.0, order, tx, ty, kc};
          if ((st = trial(c, trial_steps, &c.ms))) return st;
          cands.push_back(c);
        }
  // as fwe_autotune: the best few again with longer trials
  std::sort(cands.begin(), cands.end(), [](const Cand& p, const Cand& q) { return p.ms < q.ms; });
  Cand best = cands[0];
  double best_ms = 1e30;
  for (size_t q = 0; q < std::min<size_t>(4, cands.size()); ++q) {
    double ms = 0;
    if ((st = trial(cands[q], 4 * trial_steps, &ms))) return st;
    if (ms < best_ms) { best_ms = ms; best = cands[q]; }
  }
  g->order = best.order;
  g->tx = best.tx;
  if ((st = fwe_group_set_tiling(g, best.ty, best.kc))) return st;
  for (fwe_ctx* x : g->m)
    if ((st = fwe_reset(x))) return st;
  if (best_ms_per_step) *best_ms_per_step = best_ms;
  return FWE_OK;
}

// ---------------------------------------------------------------------------
// learned thin boundary (kernels in fwe_lb.hip)
// ---------------------------------------------------------------------------
using fwe::LbCell;
using fwe::LbGeom;

struct fwe_lb {
  int device = 0, L = 6, np = 0;
  hipStream_t stream = nullptr;
  float* P = nullptr;     // 6 faces x np, FP32 masters
  uint8_t* pk = nullptr;  // 6 faces x LB_PK_BYTES: P packed for inference (lb_repack after every change)
  float* Am = nullptr;    // Adam first moments
  float* Av = nullptr;    // Adam second moments
  float* grad = nullptr;  // np (one face at a time)
  int adam_t[6] = {0, 0, 0, 0, 0, 0};
  int form = FWE_LB_DISSIPATIVE;  // how corrections are applied
  // cell lists of the last thin geometry trained on
  int gx = -1, gy = -1, gz = -1;
  LbCell* cells = nullptr;
  int64_t face_off[7] = {0, 0, 0, 0, 0, 0, 0};
  // per-sample scratch, capacity cap
  int64_t cap = 0;
  float *X = nullptr, *R = nullptr, *T = nullptr, *M = nullptr, *Y = nullptr;
  float *H1 = nullptr, *H2 = nullptr, *D1 = nullptr, *D2 = nullptr, *G = nullptr;
  int32_t* depth = nullptr;
  float* part = nullptr;
  float* loss = nullptr;
  // contexts the boundary is attached to (their queued steps read P)
  std::vector<fwe_ctx*> attached;
};

// P changed: rewrite the inference copy.  Callers have drained the streams
// of the attached contexts, whose queued steps read it.
static hipError_t lb_repack(fwe_lb* lb) {
  hipError_t e = fwe::lb_launch_pack(lb->P, lb->np, lb->L, lb->pk, lb->stream);
  return e == hipSuccess ? hipStreamSynchronize(lb->stream) : e;
}

// forget x's learned boundary (both directions)
static void lb_detach(fwe_ctx* x) {
  if (!x->lb) return;
  auto& v = x->lb->attached;
  v.erase(std::remove(v.begin(), v.end(), x), v.end());
  x->lb = nullptr;
}

static const int LB_MLP_BLOCKS = 1024;  // resident blocks of the MLP kernels
static const int LB_CHUNK = 512;        // samples per gradient partial sum

static int lb_mlp_blocks(int64_t n) { return (int)std::max<int64_t>(1, std::min<int64_t>(LB_MLP_BLOCKS, (n + 3) / 4)); }

// Cells of an nx x ny x nz grid closer than L cells to a face, grouped by
// face (the first of x-, x+, y-, y+, z-, z+ at the smallest distance),
// x fastest within a face.
static void lb_cells_host(int nx, int ny, int nz, int L, std::vector<LbCell>& out, int64_t off[7]) {
  std::vector<LbCell> per[6];
  auto visit = [&](int i, int j, int k) {
    const int dd[6] = {i, nx - 1 - i, j, ny - 1 - j, k, nz - 1 - k};
    int f = 0;
    for (int q = 1; q < 6; ++q)
      if (dd[q] < dd[f]) f = q;
    if (dd[f] < L) per[f].push_back(LbCell{i, j, k, (f << 8) | dd[f]});
  };
  for (int k = 0; k < nz; ++k)
    for (int j = 0; j < ny; ++j) {
      const bool inner_row = j >= L && j < ny - L && k >= L && k < nz - L;
      if (inner_row) {
        for (int i = 0; i < std::min(L, nx); ++i) visit(i, j, k);
        for (int i = std::max(L, nx - L); i < nx; ++i) visit(i, j, k);
      } else {
        for (int i = 0; i < nx; ++i) visit(i, j, k);
      }
    }
  out.clear();
  off[0] = 0;
  for (int f = 0; f < 6; ++f) {
    out.insert(out.end(), per[f].begin(), per[f].end());
    off[f + 1] = (int64_t)out.size();
  }
}

// the stencil gather of the learned boundary addresses a wavefield's nine
// fields with 32-bit byte offsets
static bool lb_fits(const fwe_ctx* x) { return (uint64_t)x->nelem * 9 * sizeof(float) < ((uint64_t)1 << 32); }

static LbGeom lb_geom(const fwe_ctx* x, int from, float zinv) {
  LbGeom g;
  for (int f = 0; f < 9; ++f) g.F[f] = x->fbuf[from] + (int64_t)f * x->nelem;
  g.fstride = x->nelem;
  g.sy = x->sy;
  g.sz = x->sz;
  g.nx = x->cfg.nx;
  g.ny = x->cfg.ny;
  g.nz = x->cfg.nz;
  g.zinv = zinv;
  return g;
}

// one inference (or add) launch over every layer cell of every wavefield
static fwe::LbInfer lb_infer_args(const fwe_ctx* x, int from) {
  fwe::LbInfer a{};
  a.pk = x->lb->pk;
  a.cells = x->lb_cells;
  for (int f = 0; f < 7; ++f) a.face_off[f] = (int32_t)x->lb_face_off[f];
  a.F = x->fbuf[from];
  a.wstride = (int64_t)9 * x->nelem;
  a.fstride = (uint32_t)x->nelem;
  a.sy = (int32_t)x->sy;
  a.sz = (int32_t)x->sz;
  a.nx = x->cfg.nx;
  a.ny = x->cfg.ny;
  a.nz = x->cfg.nz;
  a.nbatch = x->cfg.nbatch;
  a.zinv = x->lb_zinv;
  a.D = x->lb_D;
  a.free_form = x->lb->form == FWE_LB_FREE;
  a.trash = x->lb_D + (int64_t)x->lb_face_off[6] * fwe::LB_OUT * x->cfg.nbatch;
  return a;
}

static hipError_t lb_before(fwe_ctx* x, int from) {
  fwe::LbInfer a = lb_infer_args(x, from);
  return fwe::lb_launch_infer(a, x->stream);
}

static hipError_t lb_after(fwe_ctx* x, int from) { return fwe::lb_launch_add(lb_infer_args(x, from), x->stream); }

extern "C" fwe_status fwe_lb_param_count(int32_t layer, int32_t* n) {
  if (!n || layer < 1 || layer > fwe::LB_MAXL) return fail(FWE_ERR_ARG, "layer must be in 1..%d", fwe::LB_MAXL);
  *n = fwe::lb_nparams(layer);
  return FWE_OK;
}

static void lb_free_scratch(fwe_lb* lb) {
  float* fs[] = {lb->X, lb->R, lb->T, lb->M, lb->Y, lb->H1, lb->H2, lb->D1, lb->D2, lb->G, lb->part};
  for (float* p : fs) hipFree(p);
  hipFree(lb->depth);
  lb->X = lb->R = lb->T = lb->M = lb->Y = lb->H1 = lb->H2 = lb->D1 = lb->D2 = lb->G = lb->part = nullptr;
  lb->depth = nullptr;
  lb->cap = 0;
}

static fwe_status lb_reserve(fwe_lb* lb, int64_t n) {
  if (n <= lb->cap) return FWE_OK;
  lb_free_scratch(lb);
  const int64_t nchunk = (n + LB_CHUNK - 1) / LB_CHUNK;
  hipError_t e = hipSuccess;
  auto al = [&](float** p, int64_t count) {
    if (e == hipSuccess) e = hipMalloc(p, sizeof(float) * (size_t)std::max<int64_t>(1, count));
  };
  al(&lb->X, n * fwe::LB_IN);
  al(&lb->R, n);
  al(&lb->T, n * fwe::LB_OUT);
  al(&lb->M, n * fwe::LB_OUT);
  al(&lb->Y, n * fwe::LB_OUT);
  al(&lb->H1, n * fwe::LB_H);
  al(&lb->H2, n * fwe::LB_H);
  al(&lb->D1, n * fwe::LB_H);
  al(&lb->D2, n * fwe::LB_H);
  al(&lb->G, n * fwe::LB_OUT);
  al(&lb->part, nchunk * fwe::LB_IN * fwe::LB_H);
  if (e == hipSuccess) e = hipMalloc(&lb->depth, sizeof(int32_t) * (size_t)std::max<int64_t>(1, n));
  if (e != hipSuccess) {
    lb_free_scratch(lb);
    return fail(FWE_ERR_NOMEM, "learned boundary scratch for %lld samples: %s", (long long)n, hipGetErrorString(e));
  }
  lb->cap = n;
  return FWE_OK;
}

extern "C" fwe_status fwe_lb_init(fwe_lb* lb, uint64_t seed) {
  if (!lb) return fail(FWE_ERR_ARG, "NULL learned boundary");
  const int np = lb->np, L = lb->L;
  std::vector<float> h((size_t)np * 6, 0.f);
  fwe_rng r;
  fwe_rng_seed(&r, seed);
  for (int f = 0; f < 6; ++f) {
    float* p = h.data() + (size_t)f * np;
    // He-normal hidden layers; zero output layer, so an initialised network
    // adds no correction and the layer starts as plain damping
    const double s1 = sqrt(2.0 / fwe::LB_IN), s2 = sqrt(2.0 / fwe::LB_H);
    for (int q = 0; q < fwe::LB_IN * fwe::LB_H; ++q) p[fwe::lb_off_w1() + q] = (float)(s1 * fwe_rng_normal(&r));
    for (int q = 0; q < fwe::LB_H * fwe::LB_H; ++q) p[fwe::lb_off_w2(L) + q] = (float)(s2 * fwe_rng_normal(&r));
  }
  HIPCHK(hipSetDevice(lb->device));
  for (fwe_ctx* x : lb->attached) HIPCHK(hipStreamSynchronize(x->stream));
  HIPCHK(hipMemcpy(lb->P, h.data(), sizeof(float) * h.size(), hipMemcpyHostToDevice));
  HIPCHK(hipMemset(lb->Am, 0, sizeof(float) * h.size()));
  HIPCHK(hipMemset(lb->Av, 0, sizeof(float) * h.size()));
  for (int f = 0; f < 6; ++f) lb->adam_t[f] = 0;
  HIPCHK(lb_repack(lb));
  return FWE_OK;
}

extern "C" fwe_status fwe_lb_destroy(fwe_lb* lb) {
  if (!lb) return FWE_OK;
  if (!lb->attached.empty())
    return fail(FWE_ERR_STATE, "learned boundary still attached to %zu context(s): detach it first",
                lb->attached.size());
  hipSetDevice(lb->device);
  if (lb->stream) hipStreamSynchronize(lb->stream);
  lb_free_scratch(lb);
  hipFree(lb->P); hipFree(lb->pk); hipFree(lb->Am); hipFree(lb->Av); hipFree(lb->grad); hipFree(lb->cells); hipFree(lb->loss);
  if (lb->stream) hipStreamDestroy(lb->stream);
  delete lb;
  return FWE_OK;
}

extern "C" fwe_status fwe_lb_create(int32_t device, int32_t layer, fwe_lb** out) {
  if (!out) return fail(FWE_ERR_ARG, "out is NULL");
  *out = nullptr;
  if (layer < 1 || layer > fwe::LB_MAXL) return fail(FWE_ERR_ARG, "layer must be in 1..%d", fwe::LB_MAXL);
  int nd = 0;
  if (hipGetDeviceCount(&nd) != hipSuccess || device < 0 || device >= nd)
    return fail(FWE_ERR_HIP, "no HIP device %d", device);
  fwe_lb* lb = new (std::nothrow) fwe_lb();
  if (!lb) return fail(FWE_ERR_NOMEM, "host allocation");
  lb->device = device;
  lb->L = layer;
  lb->np = fwe::lb_nparams(layer);
  HIPCHK(hipSetDevice(device));
  const size_t pb = sizeof(float) * (size_t)lb->np * 6;
  hipError_t e = hipStreamCreateWithFlags(&lb->stream, hipStreamNonBlocking);
  if (e == hipSuccess) e = hipMalloc(&lb->P, pb);
  if (e == hipSuccess) e = hipMalloc(&lb->pk, (size_t)fwe::LB_PK_BYTES * 6);
  if (e == hipSuccess) e = hipMalloc(&lb->Am, pb);
  if (e == hipSuccess) e = hipMalloc(&lb->Av, pb);
  if (e == hipSuccess) e = hipMalloc(&lb->grad, sizeof(float) * lb->np);
  if (e == hipSuccess) e = hipMalloc(&lb->loss, sizeof(float) * LB_MLP_BLOCKS);
  if (e != hipSuccess) {
    fwe_lb_destroy(lb);
    return fail(FWE_ERR_HIP, "learned boundary create: %s", hipGetErrorString(e));
  }
  fwe_status st = fwe_lb_init(lb, 1);
  if (st) { fwe_lb_destroy(lb); return st; }
  *out = lb;
  return FWE_OK;
}

extern "C" fwe_status fwe_lb_set_correction_form(fwe_lb* lb, int32_t form) {
  if (!lb || (form != FWE_LB_DISSIPATIVE && form != FWE_LB_FREE)) return fail(FWE_ERR_ARG, "bad correction form");
  HIPCHK(hipSetDevice(lb->device));
  // steps queued on attached contexts apply corrections in the old form;
  // their captured time loops hold it as a launch argument
  for (fwe_ctx* x : lb->attached) {
    HIPCHK(hipStreamSynchronize(x->stream));
    drop_graph(x);
  }
  lb->form = form;
  return FWE_OK;
}

extern "C" fwe_status fwe_lb_get_correction_form(const fwe_lb* lb, int32_t* form) {
  if (!lb || !form) return fail(FWE_ERR_ARG, "NULL argument");
  *form = lb->form;
  return FWE_OK;
}

extern "C" fwe_status fwe_lb_set_params(fwe_lb* lb, int32_t face, const float* p) {
  if (!lb || !p || face < 0 || face > 5) return fail(FWE_ERR_ARG, "bad learned-boundary parameter arguments");
  HIPCHK(hipSetDevice(lb->device));
  HIPCHK(hipStreamSynchronize(lb->stream));
  // steps queued on the contexts it is attached to read the parameters
  for (fwe_ctx* x : lb->attached) HIPCHK(hipStreamSynchronize(x->stream));
  HIPCHK(hipMemcpy(lb->P + (int64_t)face * lb->np, p, sizeof(float) * lb->np, hipMemcpyHostToDevice));
  HIPCHK(lb_repack(lb));
  return FWE_OK;
}

extern "C" fwe_status fwe_lb_get_params(fwe_lb* lb, int32_t face, float* p) {
  if (!lb || !p || face < 0 || face > 5) return fail(FWE_ERR_ARG, "bad learned-boundary parameter arguments");
  HIPCHK(hipSetDevice(lb->device));
  HIPCHK(hipStreamSynchronize(lb->stream));
  HIPCHK(hipMemcpy(p, lb->P + (int64_t)face * lb->np, sizeof(float) * lb->np, hipMemcpyDeviceToHost));
  return FWE_OK;
}

// gradient of the loss w.r.t. one face's parameters from the activations of
// samples [o, o+n) into lb->grad
static hipError_t lb_grad(fwe_lb* lb, int64_t o, int n) {
  const int L = lb->L;
  float* g = lb->grad;
  hipStream_t s = lb->stream;
  const float* X = lb->X + o * fwe::LB_IN;
  const float* H1 = lb->H1 + o * fwe::LB_H;
  const float* H2 = lb->H2 + o * fwe::LB_H;
  const float* D1 = lb->D1 + o * fwe::LB_H;
  const float* D2 = lb->D2 + o * fwe::LB_H;
  const float* G = lb->G + o * fwe::LB_OUT;
  const int32_t* dep = lb->depth + o;
  hipError_t e = fwe::lb_launch_atb(X, fwe::LB_IN, nullptr, fwe::LB_IN, D1, fwe::LB_H, fwe::LB_H, n, LB_CHUNK,
                                    lb->part, g + fwe::lb_off_w1(), s);
  if (e == hipSuccess)
    e = fwe::lb_launch_atb(nullptr, 0, dep, L, D1, fwe::LB_H, fwe::LB_H, n, LB_CHUNK, lb->part, g + fwe::lb_off_b1(), s);
  if (e == hipSuccess)
    e = fwe::lb_launch_atb(H1, fwe::LB_H, nullptr, fwe::LB_H, D2, fwe::LB_H, fwe::LB_H, n, LB_CHUNK, lb->part,
                           g + fwe::lb_off_w2(L), s);
  if (e == hipSuccess)
    e = fwe::lb_launch_atb(nullptr, 0, nullptr, 1, D2, fwe::LB_H, fwe::LB_H, n, LB_CHUNK, lb->part, g + fwe::lb_off_b2(L), s);
  if (e == hipSuccess)
    e = fwe::lb_launch_atb(H2, fwe::LB_H, nullptr, fwe::LB_H, G, fwe::LB_OUT, fwe::LB_OUT, n, LB_CHUNK, lb->part,
                           g + fwe::lb_off_w3(L), s);
  if (e == hipSuccess)
    e = fwe::lb_launch_atb(nullptr, 0, nullptr, 1, G, fwe::LB_OUT, fwe::LB_OUT, n, LB_CHUNK, lb->part,
                           g + fwe::lb_off_b3(L), s);
  return e;
}

// forward + backward of samples [o, o+n) of face f (targets in lb->T, M);
// adds the summed squared error to *lsum
static fwe_status lb_fb(fwe_lb* lb, int f, int64_t o, int n, bool want_y, double* lsum) {
  const int nblk = lb_mlp_blocks(n);
  HIPCHK(fwe::lb_launch_mlp(lb->P + (int64_t)f * lb->np, lb->L, n, lb->X + o * fwe::LB_IN, lb->depth + o,
                            lb->T + o * fwe::LB_OUT, lb->M + o * fwe::LB_OUT, 1.f / (9.f * (float)n),
                            want_y ? lb->Y + o * fwe::LB_OUT : nullptr, lb->H1 + o * fwe::LB_H,
                            lb->H2 + o * fwe::LB_H, lb->D1 + o * fwe::LB_H, lb->D2 + o * fwe::LB_H,
                            lb->G + o * fwe::LB_OUT, lb->loss, nblk, lb->stream));
  HIPCHK(lb_grad(lb, o, n));
  std::vector<float> lp(nblk);
  HIPCHK(hipMemcpyAsync(lp.data(), lb->loss, sizeof(float) * nblk, hipMemcpyDeviceToHost, lb->stream));
  HIPCHK(hipStreamSynchronize(lb->stream));
  double t = 0;
  for (float v : lp) t += v;
  *lsum += t;
  return FWE_OK;
}

extern "C" fwe_status fwe_lb_eval(fwe_lb* lb, int32_t face, int32_t n, const float* x, const int32_t* depth,
                                  const float* target, const float* mask, float* y, float* grad, double* loss) {
  if (!lb || face < 0 || face > 5 || n < 1 || !x || !depth || (grad && (!target || !mask)))
    return fail(FWE_ERR_ARG, "bad learned-boundary eval arguments");
  for (int s = 0; s < n; ++s)
    if (depth[s] < 0 || depth[s] >= lb->L) return fail(FWE_ERR_ARG, "sample %d: depth %d outside 0..%d", s, depth[s], lb->L - 1);
  HIPCHK(hipSetDevice(lb->device));
  fwe_status st = lb_reserve(lb, n);
  if (st) return st;
  hipStream_t s = lb->stream;
  HIPCHK(hipMemcpyAsync(lb->X, x, sizeof(float) * (size_t)n * fwe::LB_IN, hipMemcpyHostToDevice, s));
  HIPCHK(hipMemcpyAsync(lb->depth, depth, sizeof(int32_t) * n, hipMemcpyHostToDevice, s));
  const bool train = target && mask;
  if (train) {
    HIPCHK(hipMemcpyAsync(lb->T, target, sizeof(float) * (size_t)n * fwe::LB_OUT, hipMemcpyHostToDevice, s));
    HIPCHK(hipMemcpyAsync(lb->M, mask, sizeof(float) * (size_t)n * fwe::LB_OUT, hipMemcpyHostToDevice, s));
    double l = 0;
    if ((st = lb_fb(lb, face, 0, n, true, &l))) return st;
    if (loss) *loss = l / (9.0 * n);
    if (grad) HIPCHK(hipMemcpyAsync(grad, lb->grad, sizeof(float) * lb->np, hipMemcpyDeviceToHost, s));
  } else {
    HIPCHK(fwe::lb_launch_mlp(lb->P + (int64_t)face * lb->np, lb->L, n, lb->X, lb->depth, nullptr, nullptr, 0.f,
                              lb->Y, nullptr, nullptr, nullptr, nullptr, nullptr, nullptr, lb_mlp_blocks(n), s));
  }
  if (y) HIPCHK(hipMemcpyAsync(y, lb->Y, sizeof(float) * (size_t)n * fwe::LB_OUT, hipMemcpyDeviceToHost, s));
  HIPCHK(hipStreamSynchronize(s));
  return FWE_OK;
}

static fwe_status lb_plain_ctx(const fwe_ctx* x, const char* what) {
  if (x->cfg.nbatch != 1) return fail(FWE_ERR_UNSUPPORTED, "%s: one wavefield only", what);
  if (x->cfg.free_surface) return fail(FWE_ERR_UNSUPPORTED, "%s: the learned boundary replaces damping on all six faces; no free surface", what);
  return FWE_OK;
}

extern "C" fwe_status fwe_lb_train_step(fwe_lb* lb, fwe_ctx* thick, fwe_ctx* thin, int32_t offset, float impedance,
                                        float lr, int32_t resync, double* loss) {
  if (!lb || !thick || !thin || offset < 0 || !(impedance > 0.f) || !(lr >= 0.f))
    return fail(FWE_ERR_ARG, "bad learned-boundary training arguments");
  fwe_status st;
  if ((st = lb_plain_ctx(thick, "thick run")) || (st = lb_plain_ctx(thin, "thin run"))) return st;
  if (thin->lb) return fail(FWE_ERR_STATE, "the thin run must not have a learned boundary attached while training");
  if (thick->cfg.device != lb->device || thin->cfg.device != lb->device)
    return fail(FWE_ERR_ARG, "runs and learned boundary on different devices");
  const fwe_config &a = thick->cfg, &b = thin->cfg;
  if (a.nx != b.nx + 2 * offset || a.ny != b.ny + 2 * offset || a.nz != b.nz + 2 * offset)
    return fail(FWE_ERR_ARG, "thick grid must be the thin grid plus %d cells on every face", offset);
  if ((st = ready(thick)) || (st = ready(thin))) return st;
  const bool cp = thin->cfg.pml == FWE_PML_CPML;
  if (cp != (thick->cfg.pml == FWE_PML_CPML))
    return fail(FWE_ERR_ARG, "thick and thin runs must both use the convolutional PML, or neither");
  if (cp && thick->cpml_L < thin->cpml_L + offset)
    return fail(FWE_ERR_ARG, "the thick run's CPML layer (%d) must cover the thin layer plus the offset (%d)",
                thick->cpml_L, thin->cpml_L + offset);
  if (!lb_fits(thick) || !lb_fits(thin))
    return fail(FWE_ERR_UNSUPPORTED, "learned boundary: the nine fields of a wavefield exceed 4 GiB");
  HIPCHK(hipSetDevice(lb->device));
  if (lb->gx != b.nx || lb->gy != b.ny || lb->gz != b.nz) {
    std::vector<LbCell> h;
    lb_cells_host(b.nx, b.ny, b.nz, lb->L, h, lb->face_off);
    hipFree(lb->cells);
    lb->cells = nullptr;
    lb->gx = -1;
    HIPCHK(hipMalloc(&lb->cells, sizeof(LbCell) * std::max<size_t>(1, h.size())));
    HIPCHK(hipMemcpy(lb->cells, h.data(), sizeof(LbCell) * h.size(), hipMemcpyHostToDevice));
    lb->gx = b.nx; lb->gy = b.ny; lb->gz = b.nz;
  }
  const int64_t ncell = lb->face_off[6];
  if ((st = lb_reserve(lb, ncell))) return st;
  const float zinv = 1.f / impedance;
  HIPCHK(hipStreamSynchronize(thick->stream));
  HIPCHK(hipStreamSynchronize(thin->stream));
  // thin window of the thick run at step n -> thin fields (resync), or the
  // thin run's own state; features
  if (resync || thin->it != thick->it) {
    HIPCHK(fwe::lb_launch_crop(lb_geom(thick, thick->cur, zinv), lb_geom(thin, thin->cur, zinv), offset, lb->stream));
    if (cp) {
      auto pg = [](const fwe_ctx* x) {
        fwe::PsiGeom g{x->psi, x->cfg.nx, x->cfg.ny, x->cfg.nz, x->cpml_L, {x->pvar[0], x->pvar[1], x->pvar[2]}};
        return g;
      };
      HIPCHK(fwe::lb_launch_crop_psi(pg(thick), pg(thin), offset, lb->stream));
    }
  }
  HIPCHK(fwe::lb_launch_gather(lb_geom(thin, thin->cur, zinv), lb->cells, (int)ncell, lb->X, lb->R, lb->depth,
                               lb->stream));
  // RMS level of the batch (weights of the loss)
  std::vector<float> rh((size_t)ncell);
  HIPCHK(hipMemcpyAsync(rh.data(), lb->R, sizeof(float) * rh.size(), hipMemcpyDeviceToHost, lb->stream));
  HIPCHK(hipStreamSynchronize(lb->stream));
  double r2 = 0;
  for (float v : rh) r2 += (double)v * v;
  r2 /= (double)std::max<int64_t>(1, ncell);
  // one plain thin step and one thick step from the same state
  thin->it = thick->it;
  if ((st = run_impl(thin, 1, nullptr, nullptr, nullptr))) return st;
  if ((st = run_impl(thick, 1, nullptr, nullptr, nullptr))) return st;
  HIPCHK(hipStreamSynchronize(thin->stream));
  HIPCHK(hipStreamSynchronize(thick->stream));
  if (!(r2 > 1e-60)) {  // no wave in the layer yet: nothing to fit
    if (loss) *loss = 0.0;
    return FWE_OK;
  }
  HIPCHK(fwe::lb_launch_targets(lb_geom(thin, thin->cur, zinv), lb_geom(thick, thick->cur, zinv), offset, lb->cells,
                                (int)ncell, lb->R, (float)(1e-4 * sqrt(r2)), (float)(1.0 / r2), lb->T, lb->M,
                                lb->stream));
  double lsum = 0;
  // the update rewrites P: steps queued on attached contexts read it
  if (lr > 0.f)
    for (fwe_ctx* x : lb->attached) HIPCHK(hipStreamSynchronize(x->stream));
  for (int f = 0; f < 6; ++f) {
    const int64_t o = lb->face_off[f], n = lb->face_off[f + 1] - o;
    if (n == 0) continue;
    if ((st = lb_fb(lb, f, o, (int)n, true, &lsum))) return st;
    if (lr > 0.f) {
      const int t = ++lb->adam_t[f];
      const float c1 = (float)(1.0 / (1.0 - pow(0.9, t))), c2 = (float)(1.0 / (1.0 - pow(0.999, t)));
      const int64_t po = (int64_t)f * lb->np;
      HIPCHK(fwe::lb_launch_adam(lb->P + po, lb->Am + po, lb->Av + po, lb->grad, lb->np, lr, 0.9f, 0.999f, 1e-8f, c1,
                                 c2, lb->stream));
    }
  }
  // the thin run continues as it would with the boundary attached
  HIPCHK(fwe::lb_launch_apply_y(lb_geom(thin, thin->cur, zinv), lb->cells, (int)ncell, lb->Y, lb->R,
                                 lb->form == FWE_LB_FREE, lb->stream));
  if (lr > 0.f) HIPCHK(fwe::lb_launch_pack(lb->P, lb->np, lb->L, lb->pk, lb->stream));
  HIPCHK(hipStreamSynchronize(lb->stream));
  if (loss) *loss = lsum / (9.0 * (double)std::max<int64_t>(1, ncell));
  return FWE_OK;
}

extern "C" fwe_status fwe_set_learned_boundary(fwe_ctx* x, fwe_lb* lb, float impedance) {
  if (!x) return fail(FWE_ERR_ARG, "NULL ctx");
  HIPCHK(hipSetDevice(x->cfg.device));
  HIPCHK(hipStreamSynchronize(x->stream));
  drop_graph(x);
  hipFree(x->lb_cells);
  hipFree(x->lb_D);
  x->lb_cells = nullptr;
  x->lb_D = nullptr;
  lb_detach(x);
  if (!lb) return FWE_OK;
  if (x->cfg.free_surface)
    return fail(FWE_ERR_UNSUPPORTED, "learned boundary: it replaces damping on all six faces; no free surface");
  if (!(impedance > 0.f)) return fail(FWE_ERR_ARG, "impedance must be positive");
  if (lb->device != x->cfg.device) return fail(FWE_ERR_ARG, "learned boundary on another device");
  if (x->kernel == FWE_KERNEL_FUSED)
    return fail(FWE_ERR_UNSUPPORTED, "the single-pass kernel has no learned boundary");
  x->auto_sweep = false;  // AUTO keeps the two-pass kernels while it is attached
  if (!lb_fits(x)) return fail(FWE_ERR_UNSUPPORTED, "learned boundary: the nine fields of a wavefield exceed 4 GiB");
  std::vector<LbCell> h;
  lb_cells_host(x->cfg.nx, x->cfg.ny, x->cfg.nz, lb->L, h, x->lb_face_off);
  HIPCHK(hipMalloc(&x->lb_cells, sizeof(LbCell) * std::max<size_t>(1, h.size())));
  HIPCHK(hipMemcpy(x->lb_cells, h.data(), sizeof(LbCell) * h.size(), hipMemcpyHostToDevice));
  // corrections, then the inference kernel's 64-float trash slot
  const size_t nd = h.size() * fwe::LB_OUT * x->cfg.nbatch + 64;
  HIPCHK(hipMalloc(&x->lb_D, sizeof(float) * nd));
  HIPCHK(hipMemset(x->lb_D, 0, sizeof(float) * nd));
  x->lb = lb;
  lb->attached.push_back(x);
  x->lb_zinv = 1.f / impedance;
  return FWE_OK;
}
