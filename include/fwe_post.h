/*
 * fwe_post.h -- host-side seismogram post-processing (exported by libfwe.so;
 * pure CPU code, no GPU needed), for comparing traces the way the
 * reference's MATLAB drivers do before plotting:
 *   - moment-tensor rotation by station azimuth, R = Q M Q^T
 *     (ROT.m = src/runmerotation.m:203-207);
 *   - Butterworth band-pass of order n as second-order sections, applied
 *     causally in cascade, the butter(4,[flo fhi],'bandpass') / zp2sos /
 *     sosfilt sequence of OK4.m:551-558
 *     (OK4.m = src/FastWaveED_GPU3D_v1_DC_Earthquake_FULL_OK4_final/
 *      vizme_FastBiotED_GPU3D_OK_2024_final.m).
 * Moment tensors are ordered (Mxx, Myy, Mzz, Mxy, Mxz, Myz) as MTS.dat.
 */
#ifndef FWE_POST_H
#define FWE_POST_H
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* R = Q M Q^T with Q = [[cos a, sin a, 0], [-sin a, cos a, 0], [0, 0, 1]],
 * a = azimuth in degrees (ROT.m:203-207).  `out` may alias `m`. */
void fwe_post_rotate_mt(double azimuth_deg, const double m[6], double out[6]);

/* Digital Butterworth band-pass with `order` poles per band edge (2*order
 * in all, MATLAB butter(order,[flo fhi]/(fs/2),'bandpass')): analog
 * prototype, low-pass to band-pass transform on the pre-warped edges,
 * bilinear transform.  Writes `order` sections to sos[order][6] as
 * (b0 b1 b2 a0 a1 a2) with a0 = 1, the overall gain in the first section.
 * Requires 1 <= order <= 16 and 0 < flo < fhi < fs/2; returns 0, or -1 on
 * bad arguments. */
int32_t fwe_post_butter_bandpass(int32_t order, double flo, double fhi, double fs, double* sos);

/* y = the cascade of nsec sections applied to x (n samples) from zero state,
 * direct form II transposed in double precision (MATLAB / scipy sosfilt).
 * `y` may alias `x`.  Returns 0, or -1 on bad arguments. */
int32_t fwe_post_sosfilt(int32_t nsec, const double* sos, const float* x, int64_t n, float* y);

#ifdef __cplusplus
}
#endif
#endif
