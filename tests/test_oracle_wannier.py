"""Pins of the oracle's Wannier model sums (Materials/Wannier90.h:297-387), CPU only.

The reference never compiles this model, so there are no vectors to compare
with. The sums are checked against: the Haldane model (pinned on its own in
test_oracle.py) written as N = 2 Wannier blocks in the isDipoleZero phase
convention r = R + tau_n - tau_m (Wannier90.h:178-204); a numpy evaluation
of H, D on a random model; J = -dH/dk by finite differences; periodicity
over reciprocal lattice vectors.
"""
import math

import numpy as np
import pytest

from antelope_amd import WannierModel
from antelope_amd.models import synthetic_wannier
from tests.cases import haldane_cfg1_model, rel_l2, small_config


def haldane_as_wannier(h):
    """Haldane.h:86-170 as hopping blocks: A at tau_0 = 0, B at tau_1 = -a_0;
    H_AB(R) = t1 for R + tau_B - tau_A = -a_i; H_AA(+-b_j) = t2 e^{+-i phi0},
    H_BB(+-b_j) = t2 e^{-+i phi0}; on-site +-M0. Lattice L1 = b_0, L2 = -b_2."""
    a0 = h.a0
    L = np.array([[math.sqrt(3) * a0, 0, 0], [math.sqrt(3) / 2 * a0, 1.5 * a0, 0], [0, 0, 10.0]])
    # b_0 = L1, b_1 = L2 - L1, b_2 = -L2;  a_0 - a_i: 0, L2, L2 - L1
    b = [(1, 0), (-1, 1), (0, -1)]
    nn = [(0, 0), (0, 1), (-1, 1)]
    blocks = {}

    def add(R, a, c, v):
        key = (R[0], R[1], 0)
        blocks.setdefault(key, np.zeros((2, 2), complex))[a, c] += v

    add((0, 0), 0, 0, h.M0)
    add((0, 0), 1, 1, -h.M0)
    for R in nn:
        add(R, 0, 1, h.t1)
        add((-R[0], -R[1]), 1, 0, h.t1)
    for R in b:
        add(R, 0, 0, h.t2 * np.exp(1j * h.phi0))
        add((-R[0], -R[1]), 0, 0, h.t2 * np.exp(-1j * h.phi0))
        add(R, 1, 1, h.t2 * np.exp(-1j * h.phi0))
        add((-R[0], -R[1]), 1, 1, h.t2 * np.exp(1j * h.phi0))
    keys = sorted(blocks)
    pos = np.zeros((len(keys), 3, 2, 2), complex)
    tau_b = -np.array([0.0, a0, 0.0])
    pos[keys.index((0, 0, 0)), :, 1, 1] = tau_b
    return WannierModel(lattice=L, rvec=np.array(keys), rweight=np.ones(len(keys)),
                        ham=np.array([blocks[k] for k in keys]), pos=pos, nval=1, is_dipole_zero=True)


def test_haldane_written_as_wannier_blocks(oracle):
    h = haldane_cfg1_model()
    w = haldane_as_wannier(h)
    rng = np.random.default_rng(4)
    for _ in range(6):
        k = np.array([*rng.uniform(-1.5, 1.5, 2), 0.0])
        np.testing.assert_allclose(oracle.hamiltonian(w, k), oracle.hamiltonian(h, k), atol=1e-15)
        np.testing.assert_allclose(oracle.jmatrix(w, k), oracle.jmatrix(h, k), atol=1e-14)
        np.testing.assert_allclose(oracle.dipole(w, k), 0.0, atol=0)
    # the same solve through both models on the Haldane Brillouin zone (Haldane.h:77-81)
    a0 = h.a0
    ax = [2 * math.pi / math.sqrt(3) / a0, 0, 0, 0, 4 * math.pi / 3 / a0, 0, 0, 0, 0]
    cfg = small_config(n=6, ncycles=1.0, bz_axes=ax, bz_origin=[0, 0, 0])
    r_h = oracle.run(cfg, h, nthreads=4)
    r_w = oracle.run(cfg, w, nthreads=4)
    for key in ("j_inter", "j_intra", "band_pop", "rho_bar"):
        assert rel_l2(r_w[key], r_h[key]) < 1e-12, key


def _numpy_sums(m, k):
    R = m.rvec @ np.asarray(m.lattice)
    ph = m.rweight * np.exp(1j * (R @ k))
    H = np.einsum("r,rab->ab", ph, m.ham)
    D = np.einsum("r,rxab->xab", ph, m.pos)
    J = -1j * np.einsum("r,rx,rab->xab", ph, R, m.ham)
    return H, D, J


@pytest.mark.parametrize("N", [3, 6])
def test_wannier_sums_against_numpy(oracle, N):
    m = synthetic_wannier(N, nmax=(2, 1, 1), seed=90 + N)
    m.rweight = np.linspace(0.5, 1.0, len(m.rweight))  # degeneracy weights enter every sum
    rng = np.random.default_rng(N)
    for _ in range(4):
        k = rng.uniform(-0.7, 0.7, 3)
        H, D, J = _numpy_sums(m, k)
        np.testing.assert_allclose(oracle.hamiltonian(m, k), H, atol=1e-14)
        np.testing.assert_allclose(oracle.dipole(m, k), D, atol=1e-14)
        np.testing.assert_allclose(oracle.jmatrix(m, k), J, atol=1e-13)


def test_wannier_current_is_minus_dH_dk_and_periodic(oracle):
    m = synthetic_wannier(4, nmax=(2, 2, 1), seed=5)
    k = np.array([0.13, -0.29, 0.07])
    J = oracle.jmatrix(m, k)
    h = 1e-5
    for a in range(3):
        e = np.zeros(3)
        e[a] = h
        fd = -(oracle.hamiltonian(m, k + e) - oracle.hamiltonian(m, k - e)) / (2 * h)
        np.testing.assert_allclose(J[a], fd, atol=1e-9)
    G = 2 * np.pi * np.linalg.inv(np.asarray(m.lattice)).T  # rows: reciprocal vectors
    for g in G:
        np.testing.assert_allclose(oracle.hamiltonian(m, k + g), oracle.hamiltonian(m, k), atol=1e-13)
        np.testing.assert_allclose(oracle.dipole(m, k + g), oracle.dipole(m, k), atol=1e-13)
    H = oracle.hamiltonian(m, k)
    np.testing.assert_allclose(H, H.conj().T, atol=1e-15)
