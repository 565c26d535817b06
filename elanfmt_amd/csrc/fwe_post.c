/*
 * fwe_post.c -- seismogram post-processing on the host (include/fwe_post.h).
 *
 * The band-pass design is the textbook Butterworth construction: analog
 * prototype poles on the unit circle, the low-pass to band-pass
 * substitution s -> (s^2 + w0^2) / (B s) on edges pre-warped with
 * 2 fs tan(pi f / fs), and the bilinear transform z = (2fs + s) / (2fs - s).
 * Each band-pass section takes one conjugate pole pair and the zero pair
 * {+1, -1} (the transformed analog zeros at s = 0 and s = infinity).
 */
#include <complex.h>
#include <math.h>
#include <stdint.h>
#include <string.h>

#include "../../include/fwe_post.h"

static const double kPi = 3.14159265358979323846;

void fwe_post_rotate_mt(double azimuth_deg, const double m[6], double out[6]) {
  const double a = azimuth_deg * (kPi / 180.0), c = cos(a), s = sin(a);
  const double q[3][3] = {{c, s, 0.0}, {-s, c, 0.0}, {0.0, 0.0, 1.0}};
  const double M[3][3] = {{m[0], m[3], m[4]}, {m[3], m[1], m[5]}, {m[4], m[5], m[2]}};
  double t[3][3], r[3][3];
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j) {
      double v = 0.0;
      for (int k = 0; k < 3; ++k) v += q[i][k] * M[k][j];
      t[i][j] = v;
    }
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j) {
      double v = 0.0;
      for (int k = 0; k < 3; ++k) v += t[i][k] * q[j][k];
      r[i][j] = v;
    }
  out[0] = r[0][0]; out[1] = r[1][1]; out[2] = r[2][2];
  out[3] = 0.5 * (r[0][1] + r[1][0]); out[4] = 0.5 * (r[0][2] + r[2][0]); out[5] = 0.5 * (r[1][2] + r[2][1]);
}

int32_t fwe_post_butter_bandpass(int32_t order, double flo, double fhi, double fs, double* sos) {
  if (order < 1 || order > 16 || !sos || !(fs > 0) || !(flo > 0) || !(flo < fhi) || !(fhi < 0.5 * fs)) return -1;
  const int n = order;
  const double fs2 = 2.0 * fs;
  const double wl = fs2 * tan(kPi * flo / fs), wh = fs2 * tan(kPi * fhi / fs);  // pre-warped edges, rad/s
  const double bw = wh - wl, w0sq = wl * wh;
  // band-pass poles: each prototype pole p gives the roots of s^2 - p bw s + w0^2
  double complex pz[32];
  double gain = pow(bw, n);  // analog gain: bw^n (prototype gain 1)
  double complex den = 1.0, num = 1.0;
  int np = 0;
  for (int k = 0; k < n; ++k) {
    // prototype poles on the left half of the unit circle; the real one of
    // an odd order is set exactly, so that its band-pass poles come out
    // either exactly real or exactly conjugate
    const double th = kPi * (2.0 * k + n + 1) / (2.0 * n);
    const double complex p = (2 * k + n + 1 == 2 * n) ? -1.0 : cexp(I * th);
    const double complex h = 0.5 * p * bw, d = csqrt(h * h - w0sq);
    const double complex s1 = h + d, s2 = h - d;
    const double complex ps[2] = {s1, s2};
    for (int q = 0; q < 2; ++q) {
      pz[np++] = (fs2 + ps[q]) / (fs2 - ps[q]);
      den *= (fs2 - ps[q]);
    }
  }
  // the n analog zeros at s = 0 map to z = 1 with factor fs2 each; the n at
  // infinity map to z = -1
  for (int k = 0; k < n; ++k) num *= fs2;
  gain *= creal(num / den);
  // one section per conjugate pole pair (kept as its upper-half-plane pole)
  // and one per pair of real poles (a wide band of an odd order has two),
  // in order of increasing modulus: the section nearest the unit circle
  // comes last
  double sa1[16], sa2[16], mod[16];
  double re[2];
  int ns = 0, nr = 0;
  for (int q = 0; q < np; ++q) {
    const double x = creal(pz[q]), y = cimag(pz[q]);
    if (y > 0.0) {
      if (ns >= 16) return -1;
      sa1[ns] = -2.0 * x; sa2[ns] = x * x + y * y; mod[ns] = cabs(pz[q]); ++ns;
    } else if (y == 0.0) {
      if (nr >= 2) return -1;
      re[nr++] = x;
    }
  }
  if (nr == 2) {
    sa1[ns] = -(re[0] + re[1]); sa2[ns] = re[0] * re[1]; mod[ns] = fmax(fabs(re[0]), fabs(re[1])); ++ns;
  } else if (nr != 0) {
    return -1;
  }
  if (ns != n) return -1;
  for (int a = 1; a < ns; ++a)
    for (int b = a; b > 0 && mod[b] < mod[b - 1]; --b) {
      double t = mod[b]; mod[b] = mod[b - 1]; mod[b - 1] = t;
      t = sa1[b]; sa1[b] = sa1[b - 1]; sa1[b - 1] = t;
      t = sa2[b]; sa2[b] = sa2[b - 1]; sa2[b - 1] = t;
    }
  for (int k = 0; k < n; ++k) {
    double* sec = sos + 6 * k;
    const double g = k == 0 ? gain : 1.0;
    sec[0] = g; sec[1] = 0.0; sec[2] = -g;  // zeros at +1 and -1
    sec[3] = 1.0; sec[4] = sa1[k]; sec[5] = sa2[k];
  }
  return 0;
}

int32_t fwe_post_sosfilt(int32_t nsec, const double* sos, const float* x, int64_t n, float* y) {
  if (nsec < 1 || nsec > 64 || !sos || n < 0 || (n > 0 && (!x || !y))) return -1;
  double z[64][2];
  memset(z, 0, sizeof(z));
  for (int64_t t = 0; t < n; ++t) {
    double v = (double)x[t];
    for (int k = 0; k < nsec; ++k) {
      const double* s = sos + 6 * k;
      const double out = s[0] * v + z[k][0];
      z[k][0] = s[1] * v - s[4] * out + z[k][1];
      z[k][1] = s[2] * v - s[5] * out;
      v = out;
    }
    y[t] = (float)v;
  }
  return 0;
}
