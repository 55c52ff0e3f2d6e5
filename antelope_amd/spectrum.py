"""HHG spectrum of the current, as the reference's post-processing defines it.

Restates src/GeneralModel/AlsisPythonShort/AnalysisReader_PyDir.py:
  :163-167  w = linspace(-pi/dt, pi/dt - dw, Nt), dw = 2 pi / (Nt dt)
  :249-253  Blackman window over the whole series
  :323-330  F_a = -i w fftshift(fft(window j_a) dt), per axis, inter and intra
  :354      S = |F_x|^2 + |F_y|^2 + hzero (hzero = 0.98e-25, :118), Dim = 2
  :367-372  output half w[Nt//2 : Nt-1] / w0 (harmonic order)
and the error measure used to compare two solves: the largest |10 log10(S /
S_ref)| over integer harmonics 1..H_cut, where H_cut is the last harmonic whose
reference intensity lies within `window_db` of the spectrum's maximum.
"""
from __future__ import annotations

import numpy as np

HZERO = 0.98e-25


def hhg_spectrum(j_inter: np.ndarray, j_intra: np.ndarray, dt: float, w0: float, dim: int = 2):
    """j_*: [Nt][3] current series as written to the .dat files (x dV).
    Returns (harmonic order, S) over the non-negative half of the axis."""
    j_inter = np.asarray(j_inter, dtype=np.float64)
    j_intra = np.asarray(j_intra, dtype=np.float64)
    nt = j_inter.shape[0]
    wmax = np.pi / dt
    dw = 2.0 * wmax / float(nt)
    w = np.linspace(-wmax, wmax - dw, num=nt)
    win = np.blackman(nt)
    S = np.full(nt, HZERO)
    for a in range(dim):
        fi = -1j * w * np.fft.fftshift(np.fft.fft(win * j_inter[:, a]) * dt)
        fa = -1j * w * np.fft.fftshift(np.fft.fft(win * j_intra[:, a]) * dt)
        S += np.abs(fi + fa) ** 2
    half = nt // 2
    return w[half:nt - 1] / w0, S[half:nt - 1]


def harmonic_intensities(order: np.ndarray, S: np.ndarray, hmax: int, halfwidth: float = 0.25):
    """Peak spectral intensity within +-halfwidth of each harmonic 1..hmax."""
    out = np.zeros(hmax)
    for h in range(1, hmax + 1):
        sel = np.abs(order - h) <= halfwidth
        out[h - 1] = S[sel].max() if sel.any() else 0.0
    return out


def harmonic_error_db(order, S, S_ref, window_db: float = 60.0, hmax: int = 60):
    """(max dB error over harmonics up to the cutoff, cutoff harmonic)."""
    H = harmonic_intensities(order, S, hmax)
    Hr = harmonic_intensities(order, S_ref, hmax)
    floor = Hr.max() * 10 ** (-window_db / 10)
    live = np.nonzero(Hr >= floor)[0]
    hcut = int(live.max()) + 1 if live.size else 0
    if hcut == 0:
        return 0.0, 0
    err = np.abs(10 * np.log10(H[:hcut] / Hr[:hcut]))
    return float(err.max()), hcut
