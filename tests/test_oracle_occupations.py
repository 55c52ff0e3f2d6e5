"""Pins of the oracle's initial occupations (SBEs.h:273-274, 318-340).

The reference fills the band-basis initial state with the first Nval bands,
or with Fermi-Dirac occupations at T = 300 K normalised per k point to sum 1
(kT = 300 / 3.15775024804e5 Hartree, SBEs.h:273-274). The checks below derive
kT from the CODATA Boltzmann constant and Hartree energy instead of the
reference's constant, place the Fermi level where the occupations have closed
forms, and read them back through both gauges' initial states.
"""
import math

import numpy as np
import pytest

from antelope_amd import HaldaneModel
from tests.cases import small_config

KB_EV = 8.617333262e-5      # Boltzmann constant, eV/K (CODATA 2018, exact in SI)
EH_EV = 27.211386245988     # Hartree energy, eV (CODATA 2018)
KT = 300.0 * KB_EV / EH_EV  # 9.5004e-4 Hartree


def _fd(E, ef):
    f = 1.0 / (np.exp((np.asarray(E) - ef) / KT) + 1.0)
    return f / f.sum()


def _small_gap_model(x):
    """Haldane model whose gap at the K point k = -(4 pi / (3 sqrt(3) a0), 0)
    is 2 x kT: there d = (0, 0, M0 - 3 sqrt(3) t2 sin(phi0)) (the nearest-
    neighbour sum vanishes), so M0 = 3 sqrt(3) t2 sin(phi0) + x kT."""
    t1, phi0, a0 = 0.075, 0.3 * math.pi, 2.683
    t2 = t1 / 3
    m = HaldaneModel(a0=a0, t1=t1, t2=t2, phi0=phi0, M0=3 * math.sqrt(3) * t2 * math.sin(phi0) + x * KT)
    K = np.array([-4 * math.pi / (3 * math.sqrt(3) * a0), 0.0, 0.0])
    return m, K


def test_kt_constant_matches_codata():
    # the reference's 3.15775024804e5 K per Hartree is E_h / k_B
    assert 300.0 / 3.15775024804e5 == pytest.approx(KT, rel=1e-10)


@pytest.mark.parametrize("x", [0.4, 1.3])
def test_fermi_dirac_closed_forms_velocity_gauge(oracle, x):
    m, K = _small_gap_model(x)
    H = oracle.hamiltonian(m, K)
    d0 = H.trace().real / 2
    dz = abs(H[0, 0].real - d0)
    assert abs(H[0, 1]) < 1e-15 and dz == pytest.approx(x * KT, rel=1e-9)
    # Fermi level at the k point's mid-gap: 1/(e^{-|d|/kT}+1) and 1/(e^{|d|/kT}+1), already summing to 1
    cfg = small_config(n=2, gauge="VelocityHamiltonian", init="fermi", fermi_energy=d0)
    E, _, rho0 = oracle.velocity_setup(cfg, m, K)
    np.testing.assert_allclose(E, [d0 - dz, d0 + dz], rtol=1e-13)
    occ = np.diag(rho0).real
    np.testing.assert_allclose(occ, [1 / (math.exp(-x) + 1), 1 / (math.exp(x) + 1)], rtol=1e-9)
    assert np.abs(rho0 - np.diag(np.diag(rho0))).max() == 0
    # Fermi level 0.3 kT above the upper band: raw occupations sum to more than 1 and are normalised
    ef = d0 + dz + 0.3 * KT
    cfg = small_config(n=2, gauge="VelocityHamiltonian", init="fermi", fermi_energy=ef)
    _, _, rho0 = oracle.velocity_setup(cfg, m, K)
    raw = np.array([1 / (math.exp(-(2 * x + 0.3)) + 1), 1 / (math.exp(-0.3) + 1)])
    assert raw.sum() > 1.2
    np.testing.assert_allclose(np.diag(rho0).real, raw / raw.sum(), rtol=1e-9)
    # far from the gap (|d| >> kT) with the level mid-gap: the valence band full, the conduction band empty
    m1, K1 = _small_gap_model(200.0)
    H1 = oracle.hamiltonian(m1, K1)
    cfg = small_config(n=2, gauge="VelocityHamiltonian", init="fermi", fermi_energy=H1.trace().real / 2)
    _, _, rho0 = oracle.velocity_setup(cfg, m1, K1)
    np.testing.assert_allclose(np.diag(rho0).real, [1.0, 0.0], atol=1e-15)


@pytest.mark.parametrize("gauge", ["LengthWannier", "VelocityHamiltonian"])
def test_fermi_dirac_initial_state_of_a_run(oracle, gauge):
    """A one-point grid placed on the K point (BZorigin): the initial state's
    spectrum (length gauge: U diag(occ) U^dag in the orbital basis) and the
    first band-population sample are the closed-form occupations."""
    x = 0.9
    m, K = _small_gap_model(x)
    H = oracle.hamiltonian(m, K)
    d0 = H.trace().real / 2
    for ef in (d0, d0 - 0.8 * x * KT, d0 + 2.5 * x * KT):
        cfg = small_config(n=1, ncycles=0.05, gauge=gauge, init="fermi", fermi_energy=ef, bz_origin=list(K))
        r = oracle.run(cfg, m, nsteps=0, want_state=True)
        k, w, _, _ = oracle.kgrid(cfg, m)
        np.testing.assert_allclose(k[0], K, atol=1e-15)
        want = _fd([d0 - x * KT, d0 + x * KT], ef)
        np.testing.assert_allclose(np.linalg.eigvalsh(r["state"][0]), np.sort(want), rtol=1e-9, atol=1e-15)
        r1 = oracle.run(cfg, m, nsteps=1)
        np.testing.assert_allclose(r1["band_pop"][0] / w[0], want, rtol=1e-9)


def test_uniform_valence_initial_state(oracle):
    m, K = _small_gap_model(0.9)
    cfg = small_config(n=1, ncycles=0.05, bz_origin=list(K))
    r = oracle.run(cfg, m, nsteps=1, want_state=True)
    _, w, _, _ = oracle.kgrid(cfg, m)
    np.testing.assert_allclose(r["band_pop"][0] / w[0], [1.0, 0.0], atol=1e-15)
