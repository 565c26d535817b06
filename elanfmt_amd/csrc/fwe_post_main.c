/*
 * fwe_post -- host utility for the reference's trace post-processing
 * (host code only; include/fwe_post.h).
 *
 *   fwe_post filter in=<traces.res> out=<file> n=<samples per trace>
 *                   dt=<s> flo=<Hz> fhi=<Hz> [order=4] [pad=<samples>]
 *     Band-passes every trace of a raw float32 file (ntr x n, trace-major,
 *     as 0_0_Rec.res or fwe_run's traces) with the Butterworth sections of
 *     butter(order,[flo fhi],'bandpass') applied causally, after zero-padding
 *     each trace to `pad` samples as OK4.m:548-549 does; writes ntr x pad.
 *
 *   fwe_post rotate az=<deg> (mts=<MTS.dat> | m=Mxx,Myy,Mzz,Mxy,Mxz,Myz) [out=<file>]
 *     Rotates a moment tensor by the station azimuth, R = Q M Q^T
 *     (runmerotation.m:203-207); prints R and writes it as MTS.dat if `out`
 *     is given.
 */
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "../../include/fwe_post.h"

static int usage(void) {
  fprintf(stderr,
          "usage: fwe_post filter in= out= n= dt= flo= fhi= [order=4] [pad=n]\n"
          "       fwe_post rotate az= (mts=MTS.dat | m=Mxx,Myy,Mzz,Mxy,Mxz,Myz) [out=]\n");
  return 2;
}

static const char* val(int argc, char** argv, const char* key) {
  const size_t kl = strlen(key);
  for (int a = 2; a < argc; ++a)
    if (strncmp(argv[a], key, kl) == 0 && argv[a][kl] == '=') return argv[a] + kl + 1;
  return NULL;
}

static int filter(int argc, char** argv) {
  const char *in = val(argc, argv, "in"), *out = val(argc, argv, "out"), *sn = val(argc, argv, "n"),
             *sdt = val(argc, argv, "dt"), *slo = val(argc, argv, "flo"), *shi = val(argc, argv, "fhi"),
             *sord = val(argc, argv, "order"), *spad = val(argc, argv, "pad");
  if (!in || !out || !sn || !sdt || !slo || !shi) return usage();
  const long n = atol(sn), pad = spad ? atol(spad) : n;
  const int order = sord ? atoi(sord) : 4;
  const double dt = atof(sdt), flo = atof(slo), fhi = atof(shi);
  if (n < 1 || pad < n || !(dt > 0)) return usage();
  double sos[16 * 6];
  if (fwe_post_butter_bandpass(order, flo, fhi, 1.0 / dt, sos) != 0) {
    fprintf(stderr, "fwe_post: band %g..%g Hz does not fit below the Nyquist rate %g Hz (order 1..16)\n", flo, fhi,
            0.5 / dt);
    return 2;
  }
  FILE* f = fopen(in, "rb");
  if (!f) { fprintf(stderr, "fwe_post: cannot read %s\n", in); return 1; }
  fseek(f, 0, SEEK_END);
  const long bytes = ftell(f);
  fseek(f, 0, SEEK_SET);
  if (bytes % (long)(n * sizeof(float)) != 0) {
    fprintf(stderr, "fwe_post: %s holds %ld bytes, not a multiple of %ld samples\n", in, bytes, n);
    fclose(f);
    return 2;
  }
  const long ntr = bytes / (long)(n * sizeof(float));
  float* x = (float*)calloc((size_t)pad, sizeof(float));
  FILE* g = fopen(out, "wb");
  if (!x || !g) { fprintf(stderr, "fwe_post: cannot write %s\n", out); fclose(f); return 1; }
  int bad = 0;
  for (long t = 0; t < ntr && !bad; ++t) {
    memset(x, 0, sizeof(float) * (size_t)pad);
    bad = fread(x, sizeof(float), (size_t)n, f) != (size_t)n;
    bad = bad || fwe_post_sosfilt(order, sos, x, pad, x) != 0;
    bad = bad || fwrite(x, sizeof(float), (size_t)pad, g) != (size_t)pad;
  }
  fclose(f);
  fclose(g);
  free(x);
  if (bad) { fprintf(stderr, "fwe_post: I/O error\n"); return 1; }
  printf("fwe_post: %ld traces band-passed %g..%g Hz (order %d), %ld samples each -> %s\n", ntr, flo, fhi, order, pad,
         out);
  return 0;
}

static int rotate(int argc, char** argv) {
  const char *saz = val(argc, argv, "az"), *mts = val(argc, argv, "mts"), *sm = val(argc, argv, "m"),
             *out = val(argc, argv, "out");
  if (!saz || (!mts && !sm)) return usage();
  double m[6], r[6];
  if (mts) {
    float v[6];
    FILE* f = fopen(mts, "rb");
    if (!f || fread(v, sizeof(float), 6, f) != 6) {
      fprintf(stderr, "fwe_post: cannot read 6 floats from %s\n", mts);
      if (f) fclose(f);
      return 1;
    }
    fclose(f);
    for (int c = 0; c < 6; ++c) m[c] = v[c];
  } else if (sscanf(sm, "%lf,%lf,%lf,%lf,%lf,%lf", &m[0], &m[1], &m[2], &m[3], &m[4], &m[5]) != 6) {
    return usage();
  }
  fwe_post_rotate_mt(atof(saz), m, r);
  printf("%.9g %.9g %.9g %.9g %.9g %.9g\n", r[0], r[1], r[2], r[3], r[4], r[5]);
  if (out) {
    float v[6];
    for (int c = 0; c < 6; ++c) v[c] = (float)r[c];
    FILE* f = fopen(out, "wb");
    if (!f || fwrite(v, sizeof(float), 6, f) != 6) {
      fprintf(stderr, "fwe_post: cannot write %s\n", out);
      if (f) fclose(f);
      return 1;
    }
    fclose(f);
  }
  return 0;
}

int main(int argc, char** argv) {
  if (argc < 2) return usage();
  if (strcmp(argv[1], "filter") == 0) return filter(argc, argv);
  if (strcmp(argv[1], "rotate") == 0) return rotate(argc, argv);
  return usage();
}
