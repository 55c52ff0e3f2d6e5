"""sbe_run_haldane_noalloc (propagation over caller-owned device buffers, no
device allocation) against sbe_create + sbe_run on the same problems.

Two calls in a row, the second on a smaller k set, then a continuation of
the second from its propagated state. Bar: 1e-10 relative L2 per series,
the FP64 bar of the path against its oracle (the laser rows are evaluated on
the device instead of the host, so the two entry points agree to rounding,
not bit for bit: about 1e-12 on the intra-band current, whose early values
are rounding-sized).
"""
import ctypes as C
import dataclasses

import numpy as np
import pytest

from antelope_amd import _abi as A
from antelope_amd import SBESolver
from antelope_amd.solver import load_library
from tests.cases import haldane_cfg1_model, rel_l2, small_config

TOL = 1e-10


def _rows_from_solver(cfg, model, nsteps):
    """The 8 raw k sums per step, recovered from sbe_get_series of a context."""
    with SBESolver(cfg, model) as s:
        s.run(0, nsteps)
        s.synchronize()
        g = s.series()
    rb = g["rho_bar"]
    r = np.zeros((nsteps, 8))
    r[:, 0:2] = g["j_inter"][:nsteps, :2]
    r[:, 2:4] = g["j_intra"][:nsteps, :2]
    r[:, 4] = 2.0 * rb[:nsteps, 0, 1].real
    r[:, 5] = 2.0 * rb[:nsteps, 1, 0].imag
    r[:, 6] = (rb[:nsteps, 0, 0] - rb[:nsteps, 1, 1]).real
    r[:, 7] = g["band_pop"][:nsteps, 1] - g["band_pop"][:nsteps, 0]
    return r


def _hip():
    """The HIP runtime the library itself links (already loaded with it), for
    the caller's own device buffers."""
    load_library()
    hip = C.CDLL("libamdhip64.so.7")
    hip.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
    hip.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
    hip.hipFree.argtypes = [C.c_void_p]
    return hip


H2D, D2H = 1, 2


class Problem:
    """Caller-owned device buffers (allocated by the caller, outside the call)."""

    def __init__(self, cfg, model):
        from oracle import pyoracle
        self.hip = _hip()
        k, w, _, _ = pyoracle.kgrid(cfg, model)
        self.nk = k.shape[0]
        self.stride = (self.nk + A.SBE_NOALLOC_K_ALIGN - 1) // A.SBE_NOALLOC_K_ALIGN * A.SBE_NOALLOC_K_ALIGN
        host = np.zeros((4, self.stride))
        host[:3, :self.nk] = k.T
        host[3, :self.nk] = w
        self.kpts, self.state = self._alloc(host.nbytes), self._alloc(host.nbytes)
        assert self.hip.hipMemcpy(self.kpts, host.ctypes.data, host.nbytes, H2D) == 0
        self.cfg, self.model = cfg, model

    def _alloc(self, nbytes):
        p = C.c_void_p()
        assert self.hip.hipMalloc(C.byref(p), nbytes) == 0
        return p

    def free(self):
        self.hip.hipFree(self.kpts)
        self.hip.hipFree(self.state)

    def run(self, step_begin, step_end, init):
        lib = load_library()
        out = np.full((step_end - step_begin, 8), np.nan)
        series = self._alloc(out.nbytes)
        assert self.hip.hipMemcpy(series, out.ctypes.data, out.nbytes, H2D) == 0  # NaN-filled
        c, k1 = self.cfg.to_c()
        m, k2 = self.model.to_c()
        rc = lib.sbe_run_haldane_noalloc(C.byref(c), C.byref(m), self.kpts, self.state, self.nk, self.stride,
                                         step_begin, step_end, 1 if init else 0, series, None)
        assert rc == A.SBE_OK, lib.sbe_last_error(None).decode()
        # the null stream: this copy waits for the enqueued kernels
        assert self.hip.hipMemcpy(out.ctypes.data, series, out.nbytes, D2H) == 0
        self.hip.hipFree(series)
        return out


def _check(got, ref, what):
    assert np.isfinite(got).all(), what
    for v in range(8):
        e = rel_l2(got[:, v], ref[:, v])
        assert e < TOL, f"{what}: value {v} off by {e:.3e}"


@pytest.mark.gpu
def test_noalloc_matches_sbe_run_two_calls():
    model = haldane_cfg1_model()
    # first call: 300 x 300 (two k points per lane, two scratch chunks of steps)
    big = small_config(n=300, ncycles=2.0, ellip=0.5)
    ref_big = _rows_from_solver(big, model, 500)
    pb = Problem(big, model)
    print("noalloc: 300x300, steps [0, 500)", flush=True)
    _check(pb.run(0, 500, init=True), ref_big, "300x300")
    pb.free()
    # second call, smaller: 40 x 40 with another pulse, then its continuation
    small = dataclasses.replace(small_config(n=40, ncycles=3.0, rk=5), T2=150.0)
    ref_small = _rows_from_solver(small, model, 400)
    ps = Problem(small, model)
    print("noalloc: 40x40, steps [0, 250) then [250, 400)", flush=True)
    a = ps.run(0, 250, init=True)
    b = ps.run(250, 400, init=False)
    ps.free()
    _check(np.concatenate([a, b]), ref_small, "40x40")


def test_noalloc_rejects_what_it_does_not_run():
    lib = load_library()
    model = haldane_cfg1_model()
    c, k1 = small_config(n=12, precision="fp32").to_c()
    m, k2 = model.to_c()
    dummy = C.c_void_p(16)  # never dereferenced: the checks come first
    rc = lib.sbe_run_haldane_noalloc(C.byref(c), C.byref(m), dummy, dummy, 144, 512, 0, 10, 1, dummy, None)
    assert rc == A.SBE_ERR_UNSUPPORTED
    c, k1 = small_config(n=12).to_c()
    rc = lib.sbe_run_haldane_noalloc(C.byref(c), C.byref(m), dummy, dummy, 144, 500, 0, 10, 1, dummy, None)
    assert rc == A.SBE_ERR_RANGE
    rc = lib.sbe_run_haldane_noalloc(C.byref(c), C.byref(m), dummy, dummy, 144, 512, 0, 10 ** 9, 1, dummy, None)
    assert rc == A.SBE_ERR_RANGE
    # one step's partial rows of 40 M k points exceed the 16 MiB scratch
    n = 40 * 1024 * 1024
    rc = lib.sbe_run_haldane_noalloc(C.byref(c), C.byref(m), dummy, dummy, n, n, 0, 10, 1, dummy, None)
    assert rc == A.SBE_ERR_RANGE
    assert b"scratch" in lib.sbe_last_error(None)
