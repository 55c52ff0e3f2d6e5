"""Materials given as code, sampled into tight-binding blocks
(include/antelope/sbe_c.h: sbe_sample_material).

The reference's materials are classes with GenHamiltonian / GenDipole
(BaseMaterial.h:22-78) evaluated per k point on the CPU. `sample_material`
takes the same two functions as Python callables k -> matrix, samples them on
the (2 nmax_i + 1)-point grid of each reciprocal axis in the C library and
returns the WannierModel the GPU path runs, with the largest deviation of the
sampled H(k) (and D(k)) from the callables at 16 k points off the grid.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import _abi as A
from .frontend import _wannier_from_c
from .solver import SBEError, load_library


def sample_material(gen_h, nband: int, lattice, nmax, nval: int = 1, gen_d=None, tau=None,
                    is_dipole_zero: bool | None = None):
    """gen_h(k) -> [N][N] complex H(k) (Hartree); gen_d(k) -> [3][N][N] complex
    Wannier-basis dipole (bohr), or None for a dipole-free material, whose
    Fourier phases use R + tau_n - tau_m with the offsets tau ([N][3] bohr).
    lattice: rows a1, a2, a3 (bohr); nmax: hopping range along each a_i.
    Returns (WannierModel, max_dev)."""
    lib = load_library()
    N = int(nband)
    dz = gen_d is None if is_dipole_zero is None else bool(is_dipole_zero)
    errors = []

    def wrap(fn, shape):
        def cb(_user, k, out):
            try:
                m = np.asarray(fn(np.array([k[0], k[1], k[2]])), dtype=complex).reshape(shape)
                v = np.ctypeslib.as_array(out, shape=shape + (2,))
                v[..., 0], v[..., 1] = m.real, m.imag
            except Exception as e:  # an exception cannot cross the C frame
                errors.append(e)
                np.ctypeslib.as_array(out, shape=shape + (2,))[...] = np.nan
        return A.sbe_gen_matrix_fn(cb)

    ch = wrap(gen_h, (N, N))
    cd = wrap(gen_d, (3, N, N)) if gen_d is not None else A.sbe_gen_matrix_fn()
    lat = np.ascontiguousarray(lattice, dtype=np.float64).reshape(9)
    tv = None if tau is None else np.ascontiguousarray(tau, dtype=np.float64).reshape(N * 3)
    nm = (C.c_int32 * 3)(*[int(x) for x in nmax])
    m = A.sbe_model()
    dev = C.c_double()
    rc = lib.sbe_sample_material(N, int(nval), int(dz), lat.ctypes.data_as(A.c_double_p),
                                 None if tv is None else tv.ctypes.data_as(A.c_double_p), nm, ch, cd, None,
                                 C.byref(m), C.byref(dev))
    if errors:
        if rc == A.SBE_OK:
            lib.sbe_free_sampled_material(C.byref(m))
        raise errors[0]
    if rc != A.SBE_OK:
        raise SBEError(rc, lib.sbe_last_error(None).decode())
    try:
        model = _wannier_from_c(m)
    finally:
        lib.sbe_free_sampled_material(C.byref(m))
    return model, float(dev.value)
