"""Trace post-processing of libfwe.so (include/fwe_post.h; host code, no GPU):
the moment-tensor rotation of runmerotation.m:203-207 and the band-pass
butter / zp2sos / sosfilt chain of vizme_FastBiotED_GPU3D_OK_2024_final.m:
551-558.  The filter design is checked against properties a Butterworth
band-pass has by construction and against scipy.signal, which implements
the same published construction."""
import os
import subprocess

import numpy as np
import pytest

from elanfmt_amd import fwe as F

TOOL = os.path.join(os.path.dirname(F.RUN_PATH), "fwe_post")


def _H(sos, f, fs):
    z = np.exp(2j * np.pi * np.asarray(f) / fs)
    h = np.ones_like(z)
    for b0, b1, b2, a0, a1, a2 in sos:
        h *= (b0 + b1 / z + b2 / z ** 2) / (a0 + a1 / z + a2 / z ** 2)
    return h


@pytest.mark.parametrize("order,flo,fhi,fs", [(4, 0.4, 2.0, 415.6), (4, 1 / 25, 1 / 12.5, 20.0), (2, 5.0, 40.0, 200.0),
                                              (6, 1.0, 3.0, 50.0)])
def test_butterworth_bandpass_by_construction(fwe, order, flo, fhi, fs):
    """Zeros at DC and Nyquist, unit gain at the (pre-warped) centre
    frequency, -3 dB at both band edges, monotone decay outside the band,
    poles inside the unit circle."""
    sos = fwe.butter_bandpass(order, flo, fhi, fs)
    assert sos.shape == (order, 6) and np.all(sos[:, 3] == 1)
    assert np.all(np.abs(_H(sos, [0.0, fs / 2], fs)) < 1e-12)
    w = lambda f: np.tan(np.pi * f / fs)
    fc = fs / np.pi * np.arctan(np.sqrt(w(flo) * w(fhi)))
    assert abs(np.abs(_H(sos, fc, fs)) - 1) < 1e-9
    assert np.allclose(np.abs(_H(sos, [flo, fhi], fs)), np.sqrt(0.5), rtol=1e-9, atol=0)
    below = np.abs(_H(sos, np.linspace(1e-6, flo, 50), fs))
    above = np.abs(_H(sos, np.linspace(fhi, fs / 2 * 0.999, 50), fs))
    assert np.all(np.diff(below) > 0) and np.all(np.diff(above) < 0)
    r = np.sqrt(sos[:, 5])
    assert np.all(r < 1)


@pytest.mark.parametrize("order,flo,fhi,fs", [(4, 0.4, 2.0, 415.6), (4, 1 / 25, 1 / 12.5, 20.0), (3, 5.0, 40.0, 200.0)])
def test_bandpass_agrees_with_scipy(fwe, order, flo, fhi, fs):
    from scipy import signal

    ours = fwe.butter_bandpass(order, flo, fhi, fs)
    ref = signal.butter(order, [flo, fhi], btype="bandpass", fs=fs, output="sos")
    f = np.linspace(0, fs / 2, 2001)
    assert np.max(np.abs(_H(ours, f, fs) - _H(ref, f, fs))) < 1e-9
    rng = np.random.default_rng(3)
    x = rng.standard_normal(4000).astype(np.float32)
    y = fwe.sosfilt(ours, x)
    # the cascade itself: same sections through scipy (double) to FP32 output rounding
    y_same = signal.sosfilt(ours, x.astype(np.float64))
    assert np.max(np.abs(y - y_same)) <= 2e-7 * np.max(np.abs(y_same)) + 1e-30
    # the whole chain against scipy's own design
    y_ref = signal.sosfilt(ref, x.astype(np.float64))
    assert np.max(np.abs(y - y_ref)) <= 1e-6 * np.max(np.abs(y_ref))


def test_sosfilt_impulse_is_the_impulse_response(fwe):
    """Filtering a unit impulse gives the inverse z-transform of the sections:
    its DFT matches H at the DFT frequencies once the response has decayed."""
    sos = fwe.butter_bandpass(4, 2.0, 8.0, 100.0)
    x = np.zeros(8192, np.float32)
    x[0] = 1
    h = fwe.sosfilt(sos, x).astype(np.float64)
    f = np.fft.rfftfreq(x.size, 1 / 100.0)
    assert np.max(np.abs(np.fft.rfft(h) - _H(sos, f, 100.0))) < 1e-6
    assert np.array_equal(fwe.sosfilt(sos, np.zeros(10, np.float32)), np.zeros(10, np.float32))


def test_rotate_mt_special_cases(fwe):
    m = np.array([0.4895, -0.1740, 0.2439, 0.8403, 0.3594, 0.0005])  # runmerotation.m:200-202
    assert np.allclose(fwe.rotate_mt(0.0, m), m, rtol=0, atol=1e-15)
    r = fwe.rotate_mt(90.0, m)
    # Q(90) = [[0,1,0],[-1,0,0],[0,0,1]]: x' = y, y' = -x
    exp = np.array([m[1], m[0], m[2], -m[3], m[5], -m[4]])
    assert np.allclose(r, exp, rtol=0, atol=1e-15)
    r = fwe.rotate_mt(139.0, m)
    full = lambda v: np.array([[v[0], v[3], v[4]], [v[3], v[1], v[5]], [v[4], v[5], v[2]]])
    assert np.allclose(np.linalg.eigvalsh(full(r)), np.linalg.eigvalsh(full(m)), atol=1e-14)
    assert r[2] == pytest.approx(m[2], abs=1e-15)  # the vertical axis is kept
    assert np.allclose(fwe.rotate_mt(-139.0, r), m, atol=1e-14)
    assert np.allclose(fwe.rotate_mt(139.0 + 360.0, m), r, atol=1e-13)


def test_bad_band_is_refused(fwe):
    with pytest.raises(fwe.FweError):
        fwe.butter_bandpass(4, 3.0, 2.0, 100.0)
    with pytest.raises(fwe.FweError):
        fwe.butter_bandpass(4, 1.0, 60.0, 100.0)


def test_cli_filter_and_rotate(fwe, tmp_path):
    rng = np.random.default_rng(5)
    n, ntr, dt = 700, 3, 0.0024
    tr = rng.standard_normal((ntr, n)).astype(np.float32)
    src = tmp_path / "0_0_Rec.res"
    tr.tofile(src)
    out = tmp_path / "f.res"
    r = subprocess.run([TOOL, "filter", f"in={src}", f"out={out}", f"n={n}", f"dt={dt}", "flo=0.4", "fhi=2",
                        "pad=2000"], capture_output=True, text=True, timeout=60)
    assert r.returncode == 0, r.stderr
    got = np.fromfile(out, np.float32).reshape(ntr, 2000)
    sos = fwe.butter_bandpass(4, 0.4, 2.0, 1 / dt)
    for t in range(ntr):
        x = np.zeros(2000, np.float32)
        x[:n] = tr[t]
        y = fwe.sosfilt(sos, x)
        assert np.max(np.abs(got[t] - y)) <= 1e-6 * np.max(np.abs(y))
    r = subprocess.run([TOOL, "filter", f"in={src}", f"out={out}", f"n={n}", f"dt={dt}", "flo=1", "fhi=300"],
                       capture_output=True, text=True, timeout=60)
    assert r.returncode == 2
    m = np.array([1e15, -2e14, 3e14, 4e14, -5e14, 6e14], np.float32)
    m.tofile(tmp_path / "MTS.dat")
    r = subprocess.run([TOOL, "rotate", "az=139", f"mts={tmp_path / 'MTS.dat'}", f"out={tmp_path / 'R.dat'}"],
                       capture_output=True, text=True, timeout=60)
    assert r.returncode == 0, r.stderr
    rot = np.fromfile(tmp_path / "R.dat", np.float32)
    assert np.allclose(rot, fwe.rotate_mt(139.0, m.astype(np.float64)), rtol=1e-6)
