/*
 * fwe_run -- drop-in replacement for the reference's `a.exe`
 * (main() of OK4.cu:390-559 and ANISO.cu:392-608).
 *
 * Run in the directory that holds the legacy inputs.  The reference is
 * configured at compile time (-DNBX= -DNBY= -DNBZ= -DOVERX= -DOVERY= -DOVERZ=
 * -DNPARS1..3, OK4.m:432-442); this binary takes the same names as KEY=VALUE
 * arguments at run time, so the MATLAB driver's run_cmd only needs its
 * command string changed:
 *
 *   fwe_run NBX=12 NBY=12 NBZ=7 OVERX=0 OVERY=0 OVERZ=0 NPARS1=10 NPARS2=13 NPARS3=12
 *
 * Further keys: nx= ny= nz= (instead of the NB and OVER keys), nt= (default 3000 iso,
 * 4500 ortho as OK4.cu:43 / ANISO.cu:43), mode=iso|ortho|isohet (isohet: the
 * ortho loop on the iso-het file set of 9km.cu:381-400), pml=profile|
 * weights|taper|none, fs=0|1|2 (none, BC1, vacuum layer of the BC2 variant),
 * xsrc= ysrc= zsrc=, rec=C:i:j:k:name[,..]
 * (C in x|y|z), dir=<path>, device=<n>, nout=<n> (progress every n steps),
 * snap=<n> (every n steps, all nine fields as <s>_0_<field>.res, s = 1, 2,
 * ..., and the step in <s>_0_step.inf), restart=<s> (start from snapshot s:
 * the read_data convention "%d_%d_%s.%s" of OK4.cu:166-175, for all nine
 * fields; traces before the restart step stay zero), tune=<n> (autotune the
 * tiling with n-step trials before the run, default 3; 0 keeps the default
 * tiling), graph=<g> (replay g-step HIP graphs instead of launching every
 * kernel, default 0), cpml=<L> (a convolutional PML of L cells on every face
 * but a free-surface top instead of the weight files: no reference
 * counterpart; coefficients from fwe_gen_cpml_table with R = 1e-4, power 2,
 * no kappa or frequency shift, sized by the medium's fastest speed; its
 * memory variables join the snapshots as <s>_0_cpml.res), elementary=1 (the
 * six unit moment tensors Mxx Myy Mzz Mxy Mxz Myz advance as one batch of
 * wavefields in this one process, instead of one source per process as in
 * the reference (OK4.cu:336-342): the usual outputs then hold the MTS.dat
 * tensor synthesised from the six by linearity, and 0_0_<name>_<Mc>.res the
 * traces of each unit tensor, from which any tensor's traces are the same
 * linear combination; takes neither snap= nor restart=).
 * Outputs: 0_infos.inf and 0_0_{Src,Rec,Rec_x,Rec_y,Rec_y2,Vx,Vy,Vz}.res.
 */
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <time.h>

#include "../../include/fwe.h"
#include "../../include/fwe_gen.h"

#define MAXREC 1024

static int usage(void) {
  fprintf(stderr,
          "usage: fwe_run [NBX=.. NBY=.. NBZ=.. OVERX=.. OVERY=.. OVERZ=..] [nx= ny= nz=] [nt=]\n"
          "               [mode=iso|ortho|isohet] [pml=profile|weights|taper|none] [fs=0|1|2]\n"
          "               [xsrc= ysrc= zsrc=] [rec=x:i:j:k:name,...] [dir=.] [device=0] [nout=200]\n"
          "               [snap=<steps>] [restart=<snapshot>] [tune=3] [graph=0] [cpml=<cells>]\n"
          "               [elementary=0|1]\n");
  return 2;
}

static double now_s(void) {
  struct timespec t;
  clock_gettime(CLOCK_MONOTONIC, &t);
  return t.tv_sec + 1e-9 * t.tv_nsec;
}

#define CHECK(call)                                                                     \
  do {                                                                                  \
    fwe_status s_ = (call);                                                             \
    if (s_ != FWE_OK) {                                                                 \
      fprintf(stderr, "fwe_run: %s -> status %d: %s\n", #call, (int)s_, fwe_last_error()); \
      return 1;                                                                         \
    }                                                                                   \
  } while (0)

static const char* const FIELD_NAMES[9] = {"Vx", "Vy", "Vz", "Sxx", "Syy", "Szz", "Sxy", "Sxz", "Syz"};

/* elements of field f in its reference (staggered) shape */
static size_t field_elems(const fwe_config* c, int f) {
  const size_t nx = c->nx, ny = c->ny, nz = c->nz;
  switch (f) {
    case 0: return (nx + 1) * ny * nz;
    case 1: return nx * (ny + 1) * nz;
    case 2: return nx * ny * (nz + 1);
    case 6: return (nx + 1) * (ny + 1) * nz;
    case 7: return (nx + 1) * ny * (nz + 1);
    case 8: return nx * (ny + 1) * (nz + 1);
    default: return nx * ny * nz;
  }
}

/* raw float32 output 0_0_<name>.res (the "%d_%d_%s.res" convention of
 * OK4.cu:159-164 with isave = 0, rank 0) */
static int write_res(const char* dir, const char* name, const float* a, size_t n) {
  char fn[4096];
  snprintf(fn, sizeof(fn), "%s/0_0_%s.res", dir, name);
  FILE* fp = fopen(fn, "wb");
  const int ok = fp && fwrite(a, sizeof(float), n, fp) == n;
  if (fp) fclose(fp);
  if (!ok) fprintf(stderr, "fwe_run: cannot write %s\n", fn);
  return !ok;
}

static char* snap_path(const char* dir, int s, const char* name, const char* ext) {
  static char buf[4096];
  snprintf(buf, sizeof(buf), "%s/%d_0_%s.%s", dir, s, name, ext);
  return buf;
}

/* all nine fields of wavefield 0 as <s>_0_<field>.res, and the step */
static int write_snapshot(fwe_ctx* ctx, const char* dir, int s, const fwe_config* c, int step) {
  for (int f = 0; f < 9; ++f) {
    const size_t n = field_elems(c, f);
    float* h = (float*)malloc(sizeof(float) * n);
    if (!h) { fprintf(stderr, "fwe_run: out of host memory\n"); return 1; }
    if (fwe_get_field(ctx, 0, f, h) != FWE_OK) { fprintf(stderr, "fwe_run: %s\n", fwe_last_error()); free(h); return 1; }
    FILE* fp = fopen(snap_path(dir, s, FIELD_NAMES[f], "res"), "wb");
    const int ok = fp && fwrite(h, sizeof(float), n, fp) == n;
    if (fp) fclose(fp);
    free(h);
    if (!ok) { fprintf(stderr, "fwe_run: cannot write snapshot %d\n", s); return 1; }
  }
  if (c->pml == FWE_PML_CPML) {  /* the memory variables are part of the state */
    int64_t n = 0;
    if (fwe_cpml_state_size(ctx, &n) != FWE_OK) { fprintf(stderr, "fwe_run: %s\n", fwe_last_error()); return 1; }
    float* h = (float*)malloc(sizeof(float) * (size_t)n);
    if (!h) { fprintf(stderr, "fwe_run: out of host memory\n"); return 1; }
    if (fwe_get_cpml_state(ctx, 0, h) != FWE_OK) { fprintf(stderr, "fwe_run: %s\n", fwe_last_error()); free(h); return 1; }
    FILE* fc = fopen(snap_path(dir, s, "cpml", "res"), "wb");
    const int ok = fc && fwrite(h, sizeof(float), (size_t)n, fc) == (size_t)n;
    if (fc) fclose(fc);
    free(h);
    if (!ok) { fprintf(stderr, "fwe_run: cannot write snapshot %d\n", s); return 1; }
  }
  FILE* fp = fopen(snap_path(dir, s, "step", "inf"), "w");
  if (!fp) { fprintf(stderr, "fwe_run: cannot write snapshot %d\n", s); return 1; }
  fprintf(fp, "%d\n", step);
  fclose(fp);
  return 0;
}

static int read_snapshot(fwe_ctx* ctx, const char* dir, int s, const fwe_config* c, int* step) {
  FILE* fp = fopen(snap_path(dir, s, "step", "inf"), "r");
  if (!fp || fscanf(fp, "%d", step) != 1 || *step < 0) {
    fprintf(stderr, "fwe_run: snapshot %d: missing or bad %s\n", s, snap_path(dir, s, "step", "inf"));
    if (fp) fclose(fp);
    return 1;
  }
  fclose(fp);
  for (int f = 0; f < 9; ++f) {
    const size_t n = field_elems(c, f);
    float* h = (float*)malloc(sizeof(float) * n);
    if (!h) { fprintf(stderr, "fwe_run: out of host memory\n"); return 1; }
    fp = fopen(snap_path(dir, s, FIELD_NAMES[f], "res"), "rb");
    const int ok = fp && fread(h, sizeof(float), n, fp) == n;
    if (fp) fclose(fp);
    if (!ok) { fprintf(stderr, "fwe_run: snapshot %d: missing or short %s\n", s, snap_path(dir, s, FIELD_NAMES[f], "res")); free(h); return 1; }
    const fwe_status st = fwe_set_field(ctx, 0, f, h);
    free(h);
    if (st != FWE_OK) { fprintf(stderr, "fwe_run: %s\n", fwe_last_error()); return 1; }
  }
  if (c->pml == FWE_PML_CPML) {
    int64_t n = 0;
    if (fwe_cpml_state_size(ctx, &n) != FWE_OK) { fprintf(stderr, "fwe_run: %s\n", fwe_last_error()); return 1; }
    float* h = (float*)malloc(sizeof(float) * (size_t)n);
    if (!h) { fprintf(stderr, "fwe_run: out of host memory\n"); return 1; }
    fp = fopen(snap_path(dir, s, "cpml", "res"), "rb");
    const int ok = fp && fread(h, sizeof(float), (size_t)n, fp) == (size_t)n;
    if (fp) fclose(fp);
    if (!ok) { fprintf(stderr, "fwe_run: snapshot %d: missing or short %s\n", s, snap_path(dir, s, "cpml", "res")); free(h); return 1; }
    const fwe_status st = fwe_set_cpml_state(ctx, 0, h);
    free(h);
    if (st != FWE_OK) { fprintf(stderr, "fwe_run: %s\n", fwe_last_error()); return 1; }
  }
  if (fwe_set_step(ctx, *step) != FWE_OK) { fprintf(stderr, "fwe_run: %s\n", fwe_last_error()); return 1; }
  return 0;
}

/* Outputs of elementary=1 (see the header): traces of the tensor m by
 * fwe_combine_traces, velocities sum_b m_b V_b (b in order), and the traces
 * of each unit tensor as 0_0_<name>_<Mc>.res. */
static int write_elementary(fwe_ctx* ctx, const char* dir, const fwe_config* c, const char* const* np, int nrec,
                            const float m[6]) {
  static const char* const MC[6] = {"Mxx", "Myy", "Mzz", "Mxy", "Mxz", "Myz"};
  const size_t nt = (size_t)c->nt;
  int bad = 0;
  float* all = (float*)malloc(sizeof(float) * 6 * (size_t)nrec * (nt ? nt : 1));
  float* comb = (float*)malloc(sizeof(float) * (size_t)nrec * (nt ? nt : 1));
  if (!all || !comb) { fprintf(stderr, "fwe_run: out of host memory\n"); bad = 1; }
  if (!bad && nrec > 0 && nt > 0) {
    fwe_status st = fwe_get_traces(ctx, all);
    if (st == FWE_OK || st == FWE_ERR_NONFINITE) st = fwe_combine_traces(ctx, 1, m, comb);
    if (st != FWE_OK && st != FWE_ERR_NONFINITE) { fprintf(stderr, "fwe_run: %s\n", fwe_last_error()); bad = 1; }
    char nm[160];
    for (int r = 0; r < nrec && !bad; ++r) {
      bad |= write_res(dir, np[r], comb + (size_t)r * nt, nt);
      for (int b = 0; b < 6 && !bad; ++b) {
        snprintf(nm, sizeof(nm), "%s_%s", np[r], MC[b]);
        bad |= write_res(dir, nm, all + ((size_t)b * nrec + r) * nt, nt);
      }
    }
  }
  free(all);
  free(comb);
  for (int f = 0; f < 3 && !bad; ++f) {
    const size_t n = field_elems(c, f);
    float* acc = (float*)calloc(n, sizeof(float));
    float* h = (float*)malloc(sizeof(float) * n);
    if (!acc || !h) { fprintf(stderr, "fwe_run: out of host memory\n"); bad = 1; }
    for (int b = 0; b < 6 && !bad; ++b) {
      if (fwe_get_field(ctx, b, f, h) != FWE_OK) { fprintf(stderr, "fwe_run: %s\n", fwe_last_error()); bad = 1; break; }
      for (size_t q = 0; q < n; ++q) acc[q] = fmaf(m[b], h[q], acc[q]);
    }
    if (!bad) bad = write_res(dir, FIELD_NAMES[f], acc, n);
    free(acc);
    free(h);
  }
  return bad;
}

int main(int argc, char** argv) {
  int nbx = -1, nby = -1, nbz = -1, ovx = 0, ovy = 0, ovz = 0, nx = -1, ny = -1, nz = -1;
  int nt = -1, mode = FWE_MODE_ISO, pml = -1, fs = 1, device = 0, nout = 200, snap = 0, restart = -1, cpml = 0;
  int tune = 3, graph = 0, elementary = 0;
  int xsrc = -1, ysrc = -1, zsrc = -1;
  int npars[3] = {-1, -1, -1};
  const char* dir = ".";
  const char* recspec = NULL;
  for (int a = 1; a < argc; ++a) {
    const char* eq = strchr(argv[a], '=');
    if (!eq) return usage();
    const size_t kl = (size_t)(eq - argv[a]);
    const char* v = eq + 1;
#define KEY(s) (kl == strlen(s) && strncmp(argv[a], s, kl) == 0)
    if (KEY("NBX")) nbx = atoi(v);
    else if (KEY("NBY")) nby = atoi(v);
    else if (KEY("NBZ")) nbz = atoi(v);
    else if (KEY("OVERX")) ovx = atoi(v);
    else if (KEY("OVERY")) ovy = atoi(v);
    else if (KEY("OVERZ")) ovz = atoi(v);
    else if (KEY("NPARS1")) npars[0] = atoi(v);
    else if (KEY("NPARS2")) npars[1] = atoi(v);
    else if (KEY("NPARS3")) npars[2] = atoi(v);
    else if (KEY("nx")) nx = atoi(v);
    else if (KEY("ny")) ny = atoi(v);
    else if (KEY("nz")) nz = atoi(v);
    else if (KEY("nt")) nt = atoi(v);
    else if (KEY("mode"))
      mode = (strcmp(v, "ortho") == 0 || strcmp(v, "isohet") == 0) ? FWE_MODE_ORTHO : FWE_MODE_ISO;
    else if (KEY("pml")) pml = strcmp(v, "profile") == 0 ? FWE_PML_PROFILE
                            : strcmp(v, "weights") == 0 ? FWE_PML_WEIGHTS3D
                            : strcmp(v, "taper") == 0   ? FWE_PML_SEPARABLE
                                                        : FWE_PML_NONE;
    else if (KEY("fs")) fs = atoi(v);
    else if (KEY("xsrc")) xsrc = atoi(v);
    else if (KEY("ysrc")) ysrc = atoi(v);
    else if (KEY("zsrc")) zsrc = atoi(v);
    else if (KEY("rec")) recspec = v;
    else if (KEY("dir")) dir = v;
    else if (KEY("device")) device = atoi(v);
    else if (KEY("nout")) nout = atoi(v);
    else if (KEY("snap")) snap = atoi(v);
    else if (KEY("restart")) restart = atoi(v);
    else if (KEY("tune")) tune = atoi(v);
    else if (KEY("graph")) graph = atoi(v);
    else if (KEY("cpml")) cpml = atoi(v);
    else if (KEY("elementary")) elementary = atoi(v) != 0;
    else { fprintf(stderr, "fwe_run: unknown key in '%s'\n", argv[a]); return usage(); }
#undef KEY
  }
  /* OK4.cu:26-41: nx = GRID_X*BLOCKS_X - OVERX = 64*NBX - OVERX */
  if (nx < 0 && nbx > 0) nx = 64 * nbx - ovx;
  if (ny < 0 && nby > 0) ny = 64 * nby - ovy;
  if (nz < 0 && nbz > 0) nz = 64 * nbz - ovz;
  if (nx < 3 || ny < 3 || nz < 3) { fprintf(stderr, "fwe_run: grid size not given\n"); return usage(); }
  if (nt < 0) nt = mode == FWE_MODE_ISO ? 3000 : 4500;
  if (elementary && (snap > 0 || restart >= 0)) {
    fprintf(stderr, "fwe_run: snapshots hold one wavefield; elementary=1 takes neither snap= nor restart=\n");
    return 2;
  }
  const int nb = elementary ? 6 : 1;

  fwe_legacy_params p;
  CHECK(fwe_read_legacy_params(dir, &p));
  if (npars[0] > 0 && npars[0] != p.n_pa1) fprintf(stderr, "fwe_run: note: NPARS1=%d, pa1.dat holds %d\n", npars[0], p.n_pa1);
  if (npars[1] > 0 && npars[1] != p.n_pa2) fprintf(stderr, "fwe_run: note: NPARS2=%d, pa2.dat holds %d\n", npars[1], p.n_pa2);
  if (npars[2] > 0 && npars[2] != p.n_pa3) fprintf(stderr, "fwe_run: note: NPARS3=%d, pa3.dat holds %d\n", npars[2], p.n_pa3);

  fwe_config cfg;
  CHECK(fwe_config_default(&cfg, mode));
  cfg.nx = nx; cfg.ny = ny; cfg.nz = nz; cfg.nt = nt;
  cfg.dx = p.pa1[0]; cfg.dy = p.pa1[1]; cfg.dz = p.pa1[2]; cfg.dt = p.pa1[3];
  cfg.free_surface = fs;
  cfg.device = device;
  cfg.nbatch = nb;
  if (pml >= 0) cfg.pml = pml;
  if (cpml > 0) cfg.pml = FWE_PML_CPML;
  fwe_ctx* ctx = NULL;
  CHECK(fwe_create(&cfg, &ctx));
  CHECK(fwe_load_legacy_inputs(ctx, dir, &p));
  if (cpml > 0) {
    /* thin convolutional PML on every face but a free-surface top */
    double vmax = 0;
    CHECK(fwe_get_medium_vmax(ctx, &vmax));
    const int n[3] = {nx, ny, nz};
    const double h[3] = {cfg.dx, cfg.dy, cfg.dz};
    float* t[3];
    for (int a = 0; a < 3; ++a) {
      t[a] = (float*)malloc(sizeof(float) * 6 * (size_t)(n[a] + 1));
      if (!t[a] || fwe_gen_cpml_table(n[a], cpml, h[a], cfg.dt, vmax, 1e-4, 1.0, 0.0, 2.0,
                                      a == 2 && fs != FWE_FS_NONE ? 0 : 1, 1, t[a]) != 0) {
        fprintf(stderr, "fwe_run: cpml=%d does not fit the grid (2 x cells <= nx, ny, nz)\n", cpml);
        return 1;
      }
    }
    CHECK(fwe_set_cpml(ctx, cpml, t[0], t[1], t[2]));
    for (int a = 0; a < 3; ++a) free(t[a]);
  }

  /* source: OK4.cu:450-452 (234, round(Ly/2/dy), 100) and MTS.dat;
   * ANISO.cu:483-485 (180, round(Ly/2/dy), 200) with the coefficients of
   * ANISO.cu:233-247 unless MTS.dat is present */
  const double Ly = p.n_pa1 > 5 ? p.pa1[5] : ny * cfg.dy;
  if (xsrc < 0) xsrc = mode == FWE_MODE_ISO ? 234 : 180;
  if (ysrc < 0) ysrc = (int)lround((Ly - Ly / 2.0) / cfg.dy);
  if (zsrc < 0) zsrc = mode == FWE_MODE_ISO ? 100 : 200;
  float m[6];
  if (p.has_mts) {
    for (int c = 0; c < 6; ++c) m[c] = (float)p.mts[c];
  } else if (mode == FWE_MODE_ORTHO) {
    const float an[6] = {-0.6282f, 0.9437f, 0.2439f, 0.4455f, -0.2709f, -0.2362f};
    memcpy(m, an, sizeof(m));
  } else {
    fprintf(stderr, "fwe_run: MTS.dat missing (required in iso mode, OK4.cu:410)\n");
    fwe_destroy(ctx);
    return 1;
  }
  /* one tensor, or the six unit tensors on wavefields 0..5 */
  fwe_injection inj[36];
  for (int b = 0; b < nb; ++b) {
    float e[6] = {0.f, 0.f, 0.f, 0.f, 0.f, 0.f};
    e[b] = 1.f;
    CHECK(fwe_mt_injections(mode == FWE_MODE_ISO ? FWE_SRC_LAYOUT_OK4 : FWE_SRC_LAYOUT_ANISO, xsrc, ysrc, zsrc,
                            elementary ? e : m, b, inj + 6 * b));
  }
  CHECK(fwe_set_injections(ctx, 6 * nb, inj));

  /* receivers: OK4.cu:383-386 or ANISO.cu:264-267 unless rec= is given */
  static fwe_receiver rec[6 * MAXREC];
  static char names[MAXREC][64];
  const char* np[MAXREC];
  int nrec = 0;
  if (recspec) {
    const char* s = recspec;
    while (*s && nrec < MAXREC) {
      char c;
      int i, j, k, used = 0;
      if (sscanf(s, "%c:%d:%d:%d:%63[^,]%n", &c, &i, &j, &k, names[nrec], &used) != 5) {
        fprintf(stderr, "fwe_run: bad rec= entry at '%s'\n", s);
        return usage();
      }
      rec[nrec].comp = c == 'x' ? FWE_REC_VX : c == 'y' ? FWE_REC_VY : FWE_REC_VZ;
      rec[nrec].ix = i; rec[nrec].iy = j; rec[nrec].iz = k; rec[nrec].batch = 0;
      ++nrec;
      s += used;
      if (*s == ',') ++s;
    }
  } else if (mode == FWE_MODE_ISO) {
    const int t[4][4] = {{FWE_REC_VZ, 504, 385, 1}, {FWE_REC_VX, 504, 385, 1}, {FWE_REC_VY, 504, 388, 1},
                         {FWE_REC_VY, 504, 387, 1}};
    const char* nm[4] = {"Rec", "Rec_x", "Rec_y", "Rec_y2"};
    for (nrec = 0; nrec < 4; ++nrec) {
      rec[nrec] = (fwe_receiver){t[nrec][0], t[nrec][1], t[nrec][2], t[nrec][3], 0};
      snprintf(names[nrec], 64, "%s", nm[nrec]);
    }
  } else {
    const int t[4] = {FWE_REC_VZ, FWE_REC_VX, FWE_REC_VY, FWE_REC_VY};
    const char* nm[4] = {"Rec", "Rec_x", "Rec_y", "Rec_y2"};
    for (nrec = 0; nrec < 4; ++nrec) {
      rec[nrec] = (fwe_receiver){t[nrec], 631, 257, 1, 0};
      snprintf(names[nrec], 64, "%s", nm[nrec]);
    }
  }
  for (int r = 0; r < nrec; ++r) np[r] = names[r];
  /* the same receivers on every wavefield of the batch */
  for (int b = 1; b < nb; ++b)
    for (int r = 0; r < nrec; ++r) {
      rec[b * nrec + r] = rec[r];
      rec[b * nrec + r].batch = b;
    }
  CHECK(fwe_set_receivers(ctx, nb * nrec, rec));

  printf("fwe_run: %s law, %dx%dx%d cells, %d steps, dt=%.6g, free surface %s (%s)\n",
         mode == FWE_MODE_ISO ? "iso" : "ortho", nx, ny, nz, nt, cfg.dt,
         fs == FWE_FS_BC1 ? "BC1" : (fs == FWE_FS_VACUUM ? "vacuum layer" : "off"), fwe_version());
  if (elementary) printf("fwe_run: the six unit moment tensors as one batch of wavefields\n");
  if (cpml > 0) printf("fwe_run: %d-cell convolutional PML instead of the weight files\n", cpml);
  /* tiling: autotuned (the context is reset afterwards, before any restart) */
  if (tune > 0) CHECK(fwe_autotune(ctx, tune, NULL));
  if (graph > 0) CHECK(fwe_set_graph_steps(ctx, graph));
  {
    int32_t rows = 0, zc = 0, cols = 0, order = 0, enc = 0, ntab = 0;
    CHECK(fwe_get_tiling(ctx, &rows, &zc));
    CHECK(fwe_get_tile_cols(ctx, &cols));
    CHECK(fwe_get_wg_order(ctx, &order));
    CHECK(fwe_get_medium_encoding(ctx, &enc, &ntab));
    static const char* const ENC[4] = {"auto", "arrays", "material table", "isotropic arrays"};
    printf("fwe_run: tile %dx%d, %d-plane chunks, order %d%s, medium %s", cols, rows, zc, order,
           tune > 0 ? " (autotuned)" : "", ENC[enc & 3]);
    if (enc == FWE_MEDIUM_TABLE) printf(" (%d entries)", ntab);
    if (graph > 0) printf(", %d-step graphs", graph);
    printf("\n");
  }
  int done = 0, isnap = 0;
  if (restart >= 0) {
    if (read_snapshot(ctx, dir, restart, &cfg, &done)) { fwe_destroy(ctx); return 1; }
    printf("fwe_run: restart from snapshot %d at step %d\n", restart, done);
    isnap = restart;
  }
  const double t0 = now_s();
  const int first = done;
  while (done < nt) {
    int chunk = nt - done;
    if (nout > 0 && chunk > nout - done % nout) chunk = nout - done % nout;
    if (snap > 0 && chunk > snap - done % snap) chunk = snap - done % snap;
    CHECK(fwe_run(ctx, chunk));
    done += chunk;
    if (snap > 0 && done % snap == 0 && done < nt) {
      if (write_snapshot(ctx, dir, ++isnap, &cfg, done)) { fwe_destroy(ctx); return 1; }
    }
    if (nout > 0 && done % nout == 0 && done < nt) {
      CHECK(fwe_synchronize(ctx));
      printf("it=%05d > \n", done);
      fflush(stdout);
    }
  }
  CHECK(fwe_synchronize(ctx));
  const double el = now_s() - t0;
  const double cells = (double)nx * ny * nz;
  printf("time %.3f s after %d iterations, %.2f it/s, %.3f Gcell-updates/s\n", el, nt - first,
         (nt - first) / el, cells * nb * (nt - first) / el * 1e-9);
  fwe_status st = fwe_write_legacy_outputs(ctx, dir, np, nrec);
  if (st == FWE_ERR_NONFINITE) fprintf(stderr, "fwe_run: warning: %s\n", fwe_last_error());
  else if (st != FWE_OK) { fprintf(stderr, "fwe_run: %s\n", fwe_last_error()); fwe_destroy(ctx); return 1; }
  /* elementary=1: the files above hold wavefield 0 (unit Mxx); replace the
   * traces and velocities by the MTS.dat tensor's, the same linear
   * combination of the six wavefields, and add the six unit tensors' traces */
  if (elementary && write_elementary(ctx, dir, &cfg, np, nrec, m)) { fwe_destroy(ctx); return 1; }
  fwe_destroy(ctx);
  return 0;
}
