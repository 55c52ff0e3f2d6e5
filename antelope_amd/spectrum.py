"""HHG spectrum of the current, as the reference's post-processing defines it.

`hhg_spectrum` calls the library's C++ implementation (sbe_hhg_spectrum,
include/antelope/sbe_frontend.h) of
src/GeneralModel/AlsisPythonShort/AnalysisReader_PyDir.py:
  :163-167  w = linspace(-pi/dt, pi/dt - dw, Nt), dw = 2 pi / (Nt dt)
  :249-253  Blackman window over the whole series
  :323-330  F_a = -i w fftshift(fft(window j_a) dt), per axis, inter and intra
  :354      S = |F_x|^2 + |F_y|^2 + hzero (hzero = 0.98e-25, :118), Dim = 2
  :367-372  output half w[Nt//2 : Nt-1] / w0 (harmonic order)
`harmonic_error_db` is the error measure used to compare two solves: the
largest |10 log10(S / S_ref)| over integer harmonics 1..H_cut, where H_cut is
the last harmonic whose reference intensity lies within `window_db` of the
spectrum's maximum.
"""
from __future__ import annotations

import numpy as np

from . import _abi as A
from .solver import load_library

HZERO = 0.98e-25


def _series(j, nt):
    a = np.zeros((nt, 3))
    j = np.asarray(j, dtype=np.float64)
    a[:, :j.shape[1]] = j
    return np.ascontiguousarray(a)


def hhg_spectrum(j_inter: np.ndarray, j_intra: np.ndarray, dt: float, w0: float, dim: int = 2,
                 parts: bool = False):
    """j_*: [Nt][>=dim] current series as written to the .dat files (x dV).
    Returns (harmonic order, S) over the script's half of the axis; with
    parts=True also (F [n][dim] complex, S_inter, S_intra)."""
    lib = load_library()
    nt = int(np.asarray(j_inter).shape[0])
    ji, ja = _series(j_inter, nt), _series(j_intra, nt)
    n = int(lib.sbe_hhg_points(nt))
    order, S, Si, Sa = np.zeros(n), np.zeros(n), np.zeros(n), np.zeros(n)
    F = np.zeros((n, dim, 2))
    p = lambda a: a.ctypes.data_as(A.c_double_p)  # noqa: E731
    rc = lib.sbe_hhg_spectrum(nt, p(ji), p(ja), float(dt), float(w0), int(dim), p(order), p(F), p(S), p(Si), p(Sa))
    if rc != A.SBE_OK:
        raise ValueError(lib.sbe_spectrum_last_error().decode())
    if parts:
        return order, S, F[..., 0] + 1j * F[..., 1], Si, Sa
    return order, S


def write_hhg(path, j_inter, j_intra, dt: float, w0: float, dim: int = 2) -> None:
    """The script's HHG__*.dat file: w/w0, Re F_a, Im F_a per axis, S (%1.16e)."""
    lib = load_library()
    nt = int(np.asarray(j_inter).shape[0])
    ji, ja = _series(j_inter, nt), _series(j_intra, nt)
    rc = lib.sbe_write_hhg(str(path).encode(), nt, ji.ctypes.data_as(A.c_double_p), ja.ctypes.data_as(A.c_double_p),
                           float(dt), float(w0), int(dim))
    if rc != A.SBE_OK:
        raise ValueError(lib.sbe_spectrum_last_error().decode())


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
