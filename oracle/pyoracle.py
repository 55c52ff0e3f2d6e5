"""ctypes binding of the FP64 CPU oracle (oracle/liboracle.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and the
CPU-baseline leg of bench.py, never by the product package.
"""
from __future__ import annotations

import ctypes as C
import subprocess
from pathlib import Path

import numpy as np

from antelope_amd import _abi as A

_HERE = Path(__file__).resolve().parent
_LIB = _HERE / "liboracle.so"
_lib = None


class orc_info(C.Structure):
    _fields_ = [("nt", C.c_int64), ("nk_total", C.c_int64), ("nk_local", C.c_int64),
                ("atmin", C.c_double), ("atmax", C.c_double), ("dt", C.c_double),
                ("dV", C.c_double), ("total_weight", C.c_double), ("seconds", C.c_double)]


_D = A.c_double_p
_SIGS = {
    "orc_laser_eval": ([C.POINTER(A.sbe_config), C.c_int64, _D, _D, _D, _D, _D, C.POINTER(C.c_int64)], C.c_int),
    "orc_kgrid": ([C.POINTER(A.sbe_config), C.POINTER(A.sbe_model), C.c_int64, C.c_int64, _D, _D, _D, _D], C.c_int),
    "orc_hamiltonian": ([C.POINTER(A.sbe_model), _D, _D], C.c_int),
    "orc_jmatrix": ([C.POINTER(A.sbe_model), _D, _D], C.c_int),
    "orc_dipole": ([C.POINTER(A.sbe_model), _D, _D], C.c_int),
    "orc_eigh": ([C.c_int, _D, _D, _D], C.c_int),
    "orc_rhs": ([C.POINTER(A.sbe_config), C.POINTER(A.sbe_model), _D, C.c_double, _D, _D], C.c_int),
    "orc_velocity_setup": ([C.POINTER(A.sbe_config), C.POINTER(A.sbe_model), _D, _D, _D, _D], C.c_int),
    "orc_run": ([C.POINTER(A.sbe_config), C.POINTER(A.sbe_model), C.c_int64, C.c_int64, C.c_int64, C.c_int,
                 _D, _D, _D, _D, _D, C.POINTER(orc_info)], C.c_int),
    "orc_run_adaptive": ([C.POINTER(A.sbe_config), C.POINTER(A.sbe_model), C.c_int64, C.c_int64, C.c_int64,
                          C.c_int64, C.c_int, _D, _D, _D, _D, _D, C.POINTER(C.c_int64), C.POINTER(C.c_int64),
                          C.POINTER(orc_info)], C.c_int),
    "orc_cash_karp": ([_D, _D, _D, _D], C.c_int),
    "orc_observe": ([C.POINTER(A.sbe_config), C.POINTER(A.sbe_model), _D, C.c_double, _D, _D, _D, _D], C.c_int),
    "orc_currents": ([C.POINTER(A.sbe_config), C.POINTER(A.sbe_model), _D, C.c_double, _D, _D, _D, _D], C.c_int),
    "orc_last_error": ([], C.c_char_p),
}


def lib():
    global _lib
    if _lib is None:
        if not _LIB.exists():
            subprocess.run(["make", "-s", "-C", str(_HERE)], check=True)
        _lib = A.bind(C.CDLL(str(_LIB)), _SIGS)
    return _lib


def _p(a):
    return None if a is None else a.ctypes.data_as(_D)


def _chk(rc):
    if rc != 0:
        raise RuntimeError(f"oracle error {rc}: {lib().orc_last_error().decode()}")


def _c(a):
    """complex ndarray -> contiguous float64 [..., 2]"""
    return np.ascontiguousarray(np.stack([np.real(a), np.imag(a)], -1), dtype=np.float64)


def laser_eval(config, t):
    cfg, keep = config.to_c()
    t = np.ascontiguousarray(t, dtype=np.float64)
    Av, Ev = np.zeros((t.size, 3)), np.zeros((t.size, 3))
    atmin, atmax, nt = C.c_double(), C.c_double(), C.c_int64()
    _chk(lib().orc_laser_eval(C.byref(cfg), t.size, _p(t), _p(Av), _p(Ev), C.byref(atmin), C.byref(atmax), C.byref(nt)))
    return Av, Ev, atmin.value, atmax.value, nt.value


def kgrid(config, model, k_begin=-1, k_end=-1):
    cfg, k1 = config.to_c()
    mdl, k2 = model.to_c()
    ntot = int(np.prod(list(config.npoints)))
    nk = ntot if k_begin < 0 else k_end - k_begin
    k = np.zeros((nk, 3))
    w = np.zeros(nk)
    dV, tw = C.c_double(), C.c_double()
    _chk(lib().orc_kgrid(C.byref(cfg), C.byref(mdl), k_begin, k_end, _p(k), _p(w), C.byref(dV), C.byref(tw)))
    return k, w, dV.value, tw.value


def _model_mat(fn, model, k, shape):
    mdl, keep = model.to_c()
    out = np.zeros(shape + (2,))
    kk = np.ascontiguousarray(k, dtype=np.float64)
    _chk(fn(C.byref(mdl), _p(kk), _p(out)))
    return out[..., 0] + 1j * out[..., 1]


def hamiltonian(model, k):
    N = model.nband
    return _model_mat(lib().orc_hamiltonian, model, k, (N, N))


def jmatrix(model, k):
    N = model.nband
    return _model_mat(lib().orc_jmatrix, model, k, (3, N, N))


def dipole(model, k):
    N = model.nband
    return _model_mat(lib().orc_dipole, model, k, (3, N, N))


def eigh(H):
    H = np.asarray(H, dtype=complex)
    n = H.shape[0]
    ev = np.zeros(n)
    U = np.zeros((n, n, 2))
    _chk(lib().orc_eigh(n, _p(_c(H)), _p(ev), _p(U)))
    return ev, U[..., 0] + 1j * U[..., 1]


def rhs(config, model, k, t, rho):
    cfg, k1 = config.to_c()
    mdl, k2 = model.to_c()
    N = model.nband
    out = np.zeros((N, N, 2))
    kk = np.ascontiguousarray(k, dtype=np.float64)
    _chk(lib().orc_rhs(C.byref(cfg), C.byref(mdl), _p(kk), t, _p(_c(rho)), _p(out)))
    return out[..., 0] + 1j * out[..., 1]


def velocity_setup(config, model, k):
    """VelocityHamiltonian per-k constants: E[N], P[3][N][N] = U^dag J U, rho0."""
    cfg, k1 = config.to_c()
    mdl, k2 = model.to_c()
    N = model.nband
    E, P, r0 = np.zeros(N), np.zeros((3, N, N, 2)), np.zeros((N, N, 2))
    kk = np.ascontiguousarray(k, dtype=np.float64)
    _chk(lib().orc_velocity_setup(C.byref(cfg), C.byref(mdl), _p(kk), _p(E), _p(P), _p(r0)))
    return E, P[..., 0] + 1j * P[..., 1], r0[..., 0] + 1j * r0[..., 1]


def run(config, model, k_begin=-1, k_end=-1, nsteps=-1, nthreads=1, want_state=False):
    """Full oracle propagation; returns dict of series (as sbe_get_series) + info."""
    cfg, k1 = config.to_c()
    mdl, k2 = model.to_c()
    N = model.nband
    _, _, _, _, nt = laser_eval(config, np.zeros(1))
    ntot = int(np.prod(list(config.npoints)))
    nk = ntot if k_begin < 0 else k_end - k_begin
    ji, ja = np.zeros((nt, 3)), np.zeros((nt, 3))
    rb, bp = np.zeros((nt, N, N, 2)), np.zeros((nt, N))
    st = np.zeros((nk, N, N, 2)) if want_state else None
    info = orc_info()
    _chk(lib().orc_run(C.byref(cfg), C.byref(mdl), k_begin, k_end, nsteps, nthreads,
                       _p(ji), _p(ja), _p(rb), _p(bp), _p(st), C.byref(info)))
    out = {"j_inter": ji, "j_intra": ja, "rho_bar": rb[..., 0] + 1j * rb[..., 1], "band_pop": bp, "info": info}
    if want_state:
        out["state"] = st[..., 0] + 1j * st[..., 1]
    return out


def run_adaptive(config, model, k_begin=-1, k_end=-1, block=64, window=-1, nthreads=1, want_state=False):
    """Adaptive Cash-Karp restatement (config.rk_order must be the Cash-Karp pair)."""
    cfg, k1 = config.to_c()
    mdl, k2 = model.to_c()
    N = model.nband
    _, _, _, _, nt = laser_eval(config, np.zeros(1))
    ntot = int(np.prod(list(config.npoints)))
    nk = ntot if k_begin < 0 else k_end - k_begin
    nblk = (nk + block - 1) // block
    ji, ja = np.zeros((nt, 3)), np.zeros((nt, 3))
    rb, bp = np.zeros((nt, N, N, 2)), np.zeros((nt, N))
    st = np.zeros((nk, N, N, 2)) if want_state else None
    acc, rej = np.zeros(nblk, np.int64), np.zeros(nblk, np.int64)
    info = orc_info()
    _chk(lib().orc_run_adaptive(C.byref(cfg), C.byref(mdl), k_begin, k_end, block, window, nthreads,
                                _p(ji), _p(ja), _p(rb), _p(bp), _p(st),
                                acc.ctypes.data_as(C.POINTER(C.c_int64)), rej.ctypes.data_as(C.POINTER(C.c_int64)),
                                C.byref(info)))
    out = {"j_inter": ji, "j_intra": ja, "rho_bar": rb[..., 0] + 1j * rb[..., 1], "band_pop": bp, "info": info,
           "accepted": acc, "rejected": rej}
    if want_state:
        out["state"] = st[..., 0] + 1j * st[..., 1]
    return out


def cash_karp():
    """(c, a, b, bhat) exactly as the restatement uses them."""
    c, a, b, bh = np.zeros(6), np.zeros((6, 5)), np.zeros(6), np.zeros(6)
    _chk(lib().orc_cash_karp(_p(c), _p(a), _p(b), _p(bh)))
    return c, a, b, bh


def observe(config, model, k, t, rho, drho):
    """2-band observables (inter x,y, intra x,y, s, s.dhat) and their time derivatives."""
    cfg, k1 = config.to_c()
    mdl, k2 = model.to_c()
    o, od = np.zeros(8), np.zeros(8)
    kk = np.ascontiguousarray(k, dtype=np.float64)
    _chk(lib().orc_observe(C.byref(cfg), C.byref(mdl), _p(kk), t, _p(_c(rho)), _p(_c(drho)), _p(o), _p(od)))
    return o, od


def currents(config, model, k, t, rho):
    """Generic-path currents (inter[3], intra[3]) and band populations."""
    cfg, k1 = config.to_c()
    mdl, k2 = model.to_c()
    inter, intra, pop = np.zeros(3), np.zeros(3), np.zeros(model.nband)
    kk = np.ascontiguousarray(k, dtype=np.float64)
    _chk(lib().orc_currents(C.byref(cfg), C.byref(mdl), _p(kk), t, _p(_c(rho)), _p(inter), _p(intra), _p(pop)))
    return inter, intra, pop
