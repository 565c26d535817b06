/*
 * fwe_gen_inputs -- writes the legacy input set that the reference's MATLAB
 * drivers produce before running the solver, so that fwe_run (or the
 * reference's a.exe) can be driven without MATLAB.
 *
 *   iso   (OK4.m:160-429):  pa1.dat pa2.dat pa3.dat MTS.dat Src.dat
 *                           c11u_het.dat c12u_het.dat weights_{x,y,z}.dat
 *   ortho (ANIS.m:378-533): pa1.dat pa2.dat pa3.dat MTS.dat Src.dat
 *                           C11m..C66mxy.dat vrho_{x,y,z}_11_hc.dat and
 *                           weights.dat (pml=weights) or taper_{x,y,z}.dat
 *   isohet (9km.cu:381-400): as ortho with the medium as c11u_het.dat,
 *                           c12u_het.dat and c44u_het{xy,xz,yz}.dat (layered
 *                           isotropic: the ortho layers without anisotropy)
 *
 * Raw float32, x fastest, in the reference's staggered shapes.  Keys:
 *   dir= nx= ny= nz= dx=(50) nt= cfl=(0.5) mode=iso|ortho|isohet
 *   iso:   vp=(6000) vs=(3464) rho=(2700) L=(20) R=(1e-8)
 *   fs=(0|1|2): 1 and 2 leave the top undamped; 2 also empties the top cell
 *          layer (zero moduli at k = 0) for the vacuum-layer free surface, as
 *          the BC2 driver does (vizme_FastBiotED_GPU3D_OK_2024.m:238,243)
 *   ortho: seed=(1) layers=(5) L=(24) R=(1e-8) pml=taper|weights
 *   source: sdr=strike:dip:rake (30:60:90) or lune=<seed>; m0=(1e15); f0=(4)
 * The source time function is the Ricker of OK4.cu:210-218 with t0 = 1.2/f0,
 * scaled by dt/(dx dy dz) so that one step adds the moment released in dt.
 */
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "../../include/fwe_gen.h"

static int usage(void) {
  fprintf(stderr,
          "usage: fwe_gen_inputs nx= ny= nz= nt= [dir=.] [dx=50] [cfl=0.5] [mode=iso|ortho|isohet]\n"
          "       iso:   [vp=6000 vs=3464 rho=2700 L=20 R=1e-8]\n"
          "       ortho: [seed=1 layers=5 L=24 R=1e-8 pml=taper|weights]\n"
          "       [fs=0|1|2]\n"
          "       [sdr=30:60:90 | lune=<seed>] [m0=1e15] [f0=4]\n");
  return 2;
}

static int put(const char* dir, const char* name, const float* a, size_t n) {
  char path[4096];
  snprintf(path, sizeof(path), "%s/%s", dir, name);
  FILE* f = fopen(path, "wb");
  if (!f || fwrite(a, sizeof(float), n, f) != n) {
    fprintf(stderr, "fwe_gen_inputs: cannot write %s\n", path);
    if (f) fclose(f);
    return 1;
  }
  fclose(f);
  return 0;
}

static float* alloc(size_t n) {
  float* p = (float*)calloc(n ? n : 1, sizeof(float));
  if (!p) {
    fprintf(stderr, "fwe_gen_inputs: out of host memory\n");
    exit(1);
  }
  return p;
}

/* the ANISO damping of one axis from the OK4 profile: cell-centred factors */
static float* cell_taper(int n, int L, double R, double dt, int top_undamped) {
  float* w = alloc((size_t)n + 1);
  float* t = alloc((size_t)n);
  fwe_gen_pml_profile(n, L, R, dt, 1, 1, top_undamped, w);
  fwe_gen_cell_profile(n, w, t);
  free(w);
  return t;
}

int main(int argc, char** argv) {
  int nx = -1, ny = -1, nz = -1, nt = -1, ortho = 0, isohet = 0, L = -1, fs = 0, layers = 5, weights3d = 0;
  double dx = 50.0, cfl = 0.5, vp = 6000.0, vs = 3464.0, rho = 2700.0, R = 1e-8, m0 = 1e15, f0 = 4.0;
  double sdr[3] = {30.0, 60.0, 90.0};
  long long seed = 1, lune = -1;
  const char* dir = ".";
  for (int a = 1; a < argc; ++a) {
    const char* eq = strchr(argv[a], '=');
    if (!eq) return usage();
    const size_t kl = (size_t)(eq - argv[a]);
    const char* v = eq + 1;
#define KEY(s) (kl == strlen(s) && strncmp(argv[a], s, kl) == 0)
    if (KEY("nx")) nx = atoi(v);
    else if (KEY("ny")) ny = atoi(v);
    else if (KEY("nz")) nz = atoi(v);
    else if (KEY("nt")) nt = atoi(v);
    else if (KEY("dx")) dx = atof(v);
    else if (KEY("cfl")) cfl = atof(v);
    else if (KEY("mode")) {
      isohet = strcmp(v, "isohet") == 0;
      ortho = isohet || strcmp(v, "ortho") == 0;
    }
    else if (KEY("vp")) vp = atof(v);
    else if (KEY("vs")) vs = atof(v);
    else if (KEY("rho")) rho = atof(v);
    else if (KEY("L")) L = atoi(v);
    else if (KEY("R")) R = atof(v);
    else if (KEY("fs")) fs = atoi(v);
    else if (KEY("seed")) seed = atoll(v);
    else if (KEY("layers")) layers = atoi(v);
    else if (KEY("pml")) weights3d = strcmp(v, "weights") == 0;
    else if (KEY("sdr")) {
      if (sscanf(v, "%lf:%lf:%lf", &sdr[0], &sdr[1], &sdr[2]) != 3) return usage();
    } else if (KEY("lune")) lune = atoll(v);
    else if (KEY("m0")) m0 = atof(v);
    else if (KEY("f0")) f0 = atof(v);
    else if (KEY("dir")) dir = v;
    else {
      fprintf(stderr, "fwe_gen_inputs: unknown key in '%s'\n", argv[a]);
      return usage();
    }
#undef KEY
  }
  if (nx < 3 || ny < 3 || nz < 3 || nt < 1 || fs < 0 || fs > 2) return usage();
  if (L < 0) L = ortho ? 24 : 20;
  if (2 * L > nx || 2 * L > ny || 2 * L > nz) {
    fprintf(stderr, "fwe_gen_inputs: damping of %d cells does not fit the grid\n", L);
    return 2;
  }
  const size_t nc = (size_t)nx * ny * nz;
  const double dy = dx, dz = dx;
  float* cij[9];
  float* bcell = NULL;
  double vmax = vp;
  if (ortho) {
    for (int q = 0; q < 9; ++q) cij[q] = alloc(nc);
    bcell = alloc(nc);
    int32_t* tops = (int32_t*)calloc((size_t)layers, sizeof(int32_t));
    if (!tops || fwe_gen_layered_ortho(nx, ny, nz, layers, (uint64_t)seed, 4000.0, 6500.0, 1.65, 1.85, 2400.0,
                                       2900.0, isohet ? 0.0 : 0.15, cij, bcell, tops, &vmax) != 0) {
      fprintf(stderr, "fwe_gen_inputs: layered medium generation failed\n");
      return 1;
    }
    free(tops);
    if (fs == 2)  // vacuum cells at k = 0 (the edge moduli follow by harmonic averaging)
      for (int q = 0; q < 9; ++q) memset(cij[q], 0, sizeof(float) * (size_t)nx * ny);
  }
  const double dt = fwe_gen_cfl_dt(vmax, dx, dy, dz, cfl);

  /* pa1 (OK4.m:160-176): dx dy dz dt Lx Ly Lz lamx lamy lamz */
  const float pa1[10] = {(float)dx, (float)dy, (float)dz, (float)dt, (float)((nx - 1) * dx),
                         (float)((ny - 1) * dy), (float)((nz - 1) * dz), 1.f, 1.f, 1.f};
  float pa2[13] = {0}, pa3[12] = {0};
  if (!ortho) {
    const double c11 = rho * vp * vp, c12 = rho * (vp * vp - 2 * vs * vs), c66 = rho * vs * vs;
    const float p2[13] = {(float)c11, (float)c11, (float)c12, (float)c12, (float)c12, (float)c66,
                          (float)c66, (float)c66, 0.f, 0.f, 0.f, 0.f, (float)c11};
    memcpy(pa2, p2, sizeof(pa2));
    pa3[0] = pa3[1] = pa3[2] = (float)(1.0 / rho);
  }
  int bad = put(dir, "pa1.dat", pa1, 10) || put(dir, "pa2.dat", pa2, 13) || put(dir, "pa3.dat", pa3, 12);

  /* source */
  double m[6];
  if (lune >= 0) {
    fwe_rng r;
    fwe_rng_seed(&r, (uint64_t)lune);
    fwe_gen_mt_lune_random(&r, m0, m, NULL, NULL);
  } else {
    fwe_gen_mt_sdr(sdr[0], sdr[1], sdr[2], m0, m);
  }
  float mts[6];
  for (int c = 0; c < 6; ++c) mts[c] = (float)m[c];
  float* src = alloc((size_t)nt);
  fwe_gen_ricker(nt, dt, f0, 1.2 / f0, dt / (dx * dy * dz), src);
  bad = bad || put(dir, "MTS.dat", mts, 6) || put(dir, "Src.dat", src, (size_t)nt);
  free(src);

  if (!ortho) {
    float* a = alloc(nc);
    const size_t top = fs == 2 ? (size_t)nx * ny : 0;  // vacuum cells at k = 0
    for (size_t q = 0; q < nc; ++q) a[q] = q < top ? 0.f : (float)(rho * vp * vp);
    bad = bad || put(dir, "c11u_het.dat", a, nc);
    for (size_t q = 0; q < nc; ++q) a[q] = q < top ? 0.f : (float)(rho * (vp * vp - 2 * vs * vs));
    bad = bad || put(dir, "c12u_het.dat", a, nc);
    free(a);
    const int ns[3] = {nx, ny, nz};
    const char* wn[3] = {"weights_x.dat", "weights_y.dat", "weights_z.dat"};
    for (int ax = 0; ax < 3 && !bad; ++ax) {
      float* w = alloc((size_t)ns[ax] + 1);
      fwe_gen_pml_profile(ns[ax], L, R, dt, 1, 1, ax == 2 && fs ? 1 : 0, w);
      bad = put(dir, wn[ax], w, (size_t)ns[ax] + 1);
      free(w);
    }
  } else {
    static const char* cn[9] = {"C11m.dat", "C22m.dat", "C33m.dat", "C12m.dat", "C13m.dat",
                                "C23m.dat", "C44myz.dat", "C55mxz.dat", "C66mxy.dat"};
    static const char* hn[9] = {"c11u_het.dat", NULL, NULL, "c12u_het.dat", NULL,
                                NULL, "c44u_hetyz.dat", "c44u_hetxz.dat", "c44u_hetxy.dat"};
    const char* const* names = isohet ? hn : cn;
    const size_t nyz = (size_t)nx * (ny + 1) * (nz + 1), nxz = (size_t)(nx + 1) * ny * (nz + 1),
                 nxy = (size_t)(nx + 1) * (ny + 1) * nz;
    float *gyz = alloc(nyz), *gxz = alloc(nxz), *gxy = alloc(nxy);
    fwe_gen_harmonic_edges(nx, ny, nz, cij[6], cij[7], cij[8], gyz, gxz, gxy);
    for (int q = 0; q < 6 && !bad; ++q)
      if (names[q]) bad = put(dir, names[q], cij[q], nc);
    bad = bad || put(dir, names[6], gyz, nyz) || put(dir, names[7], gxz, nxz) || put(dir, names[8], gxy, nxy);
    free(gyz); free(gxz); free(gxy);
    const size_t nbx = (size_t)(nx + 1) * ny * nz, nby = (size_t)nx * (ny + 1) * nz, nbz = (size_t)nx * ny * (nz + 1);
    float *bx = alloc(nbx), *by = alloc(nby), *bz = alloc(nbz);
    fwe_gen_stagger_buoyancy(nx, ny, nz, bcell, bx, by, bz);
    bad = bad || put(dir, "vrho_x_11_hc.dat", bx, nbx) || put(dir, "vrho_y_11_hc.dat", by, nby) ||
          put(dir, "vrho_z_11_hc.dat", bz, nbz);
    free(bx); free(by); free(bz);
    float *tx = cell_taper(nx, L, R, dt, 0), *ty = cell_taper(ny, L, R, dt, 0), *tz = cell_taper(nz, L, R, dt, fs != 0);
    if (weights3d) {
      float* w = alloc(nc);
      for (int k = 0; k < nz; ++k)
        for (int j = 0; j < ny; ++j)
          for (int i = 0; i < nx; ++i) w[i + (size_t)nx * (j + (size_t)ny * k)] = tx[i] * ty[j] * tz[k];
      bad = bad || put(dir, "weights.dat", w, nc);
      free(w);
    } else {
      bad = bad || put(dir, "taper_x.dat", tx, (size_t)nx) || put(dir, "taper_y.dat", ty, (size_t)ny) ||
            put(dir, "taper_z.dat", tz, (size_t)nz);
    }
    free(tx); free(ty); free(tz);
    for (int q = 0; q < 9; ++q) free(cij[q]);
    free(bcell);
  }
  if (bad) return 1;
  printf("fwe_gen_inputs: %s inputs for %dx%dx%d, dt=%.6g, %d steps, %d damping cells -> %s\n",
         isohet ? "iso-het" : (ortho ? "ortho" : "iso"), nx, ny, nz, dt, nt, L, dir);
  return 0;
}
