"""Pins of the oracle's analytic model classes (Materials/*.h), CPU only.

Each class's H(k) is checked against symmetries its lattice fixes (C3 for
the honeycomb and triangular classes, C4 for the square ones, time reversal
where the class has it) and its J(k) against finite differences of H with
the relation the class's GenJMatrix implements (-dH/dk, +dH/dk, or the
noted elements). A dropped term or a wrong coefficient breaks the symmetry;
a wrong sign or index breaks the finite-difference relation.
"""
import math
import re
from pathlib import Path

import numpy as np
import pytest

from antelope_amd import AnalyticModel, tmdc
from tests.cases import rel_l2, small_config

REF = Path("/root/reference/src/GeneralModel/Materials")


def models():
    return {
        "wilson": AnalyticModel.from_reference("WilsonMass", t=0.02, delta=0.03, mu=0.01, a0=4.0),
        "wilson_rect": AnalyticModel.from_reference("WilsonMass", t=0.02, delta=0.03, mu=-0.015, a0=4.0, a1=5.5),
        "square": AnalyticModel.from_reference("SquareSBand", t1=0.03, M0=0.05, a0=2.5),
        "kanemele": AnalyticModel.from_reference("KaneMele", t1=0.075, t2=0.025, tR=0.0, Mt2=1.5, phi0=0.9, a0=1.42),
        "kanemele_rashba": AnalyticModel.from_reference("KaneMele", t1=0.075, t2=0.025, tR=0.01, Mt2=1.5, phi0=0.9,
                                                        a0=1.42),
        "xyh": AnalyticModel.from_reference("XYHlattice", tl=0.05, tr=0.03, **{"lambda": 0.01}, a0=2.46),
        "mos2": tmdc("MoS2"),
        "wse2_soc": tmdc("WSe2", soc=True),
        "h2l_2h": AnalyticModel.from_reference("Haldane2L", t1=0.075, t2=0.025, phi0=0.9, a0=1.42, t11=0.01,
                                               polytype="2H", Mt2=1.0),
        "h2l_ab_bottom": AnalyticModel.from_reference("Haldane2L", t1=0.075, t2=0.025, phi0=0.9, a0=1.42, t11=0.012,
                                                      polytype="AB", M0=0.02, t2_bottom=0.02, phi0_bottom=0.3),
        "bi2se3": AnalyticModel.from_reference("Bi2Se3surf"),
        "multiwsm": AnalyticModel.from_reference("MultiWSM", n=2.0, v=0.05, alphan=0.3, phi0=0.0, a0=3.0),
        "wdsm": AnalyticModel.from_reference("WDSM_TaAs", t=0.02, delta=0.5, my=0.8, mz=0.9, a0=3.4),
    }


def _rot(k, ang):
    c, s = math.cos(ang), math.sin(ang)
    return np.array([c * k[0] - s * k[1], s * k[0] + c * k[1], 0.0])


@pytest.mark.parametrize("name", list(models()))
def test_hamiltonian_hermitian_and_symmetric(oracle, name):
    m = models()[name]
    rng = np.random.default_rng(7)
    rot = {"wilson": math.pi / 2, "square": math.pi / 2, "kanemele": 2 * math.pi / 3,
           "kanemele_rashba": 2 * math.pi / 3, "xyh": 2 * math.pi / 3, "mos2": 2 * math.pi / 3,
           "wse2_soc": 2 * math.pi / 3, "h2l_2h": 2 * math.pi / 3, "h2l_ab_bottom": 2 * math.pi / 3,
           "bi2se3": 2 * math.pi / 3, "multiwsm": 0.37}.get(name)
    for _ in range(5):
        k = np.array([*rng.uniform(-0.6, 0.6, 2), 0.0])
        H = oracle.hamiltonian(m, k)
        np.testing.assert_allclose(H, H.conj().T, atol=1e-16)
        e = np.linalg.eigvalsh(H)
        if rot is not None:  # point-group rotation of the lattice leaves the spectrum unchanged
            np.testing.assert_allclose(np.linalg.eigvalsh(oracle.hamiltonian(m, _rot(k, rot))), e, atol=1e-13)
        if name in ("wilson", "wilson_rect", "square", "mos2", "wse2_soc", "kanemele_rashba", "xyh", "bi2se3"):
            # time reversal (spinless: H(-k) = H(k)*; with spin: Kramers pairs) -> E(-k) = E(k)
            np.testing.assert_allclose(np.linalg.eigvalsh(oracle.hamiltonian(m, -k)), e, atol=1e-13)


def _fd(oracle, m, k, h=1e-5):
    out = []
    for a in range(3):
        e = np.zeros(3)
        e[a] = h
        out.append((oracle.hamiltonian(m, k + e) - oracle.hamiltonian(m, k - e)) / (2 * h))
    return np.array(out)


def _fd3(oracle, m, k, h=1e-5):
    return _fd(oracle, m, k, h)


@pytest.mark.parametrize("name", list(models()))
def test_current_matrix_against_finite_differences(oracle, name):
    m = models()[name]
    k = np.array([0.137, -0.211, 0.0])
    J = oracle.jmatrix(m, k)
    dH = _fd(oracle, m, k)
    if name == "multiwsm":
        np.testing.assert_allclose(J[2], np.diag([-m.params[1], m.params[1]]), atol=0)
    else:
        np.testing.assert_allclose(J[2], 0.0, atol=0)
    if name.startswith("wilson"):
        np.testing.assert_allclose(J, -dH, atol=1e-9)            # J = -dH/dk
    elif name in ("xyh", "mos2", "wse2_soc"):
        np.testing.assert_allclose(J, dH, atol=1e-9)             # XYHlattice.h, TMDCs.h write +dH/dk
    elif name == "square":                                       # SquareSBand.h:69-78: only (0,1)
        np.testing.assert_allclose(J[:2, 0, 1], -dH[:2, 0, 1], atol=1e-9)
        assert np.all(J[:2, 1, 0] == 0) and np.all(J[:2, 0, 0] == 0) and np.all(J[:2, 1, 1] == 0)
    elif name in ("h2l_2h", "bi2se3"):
        np.testing.assert_allclose(J, -dH, atol=1e-9)            # Haldane2L.h, Bi2Se3surf.h: -dH/dk
    elif name == "h2l_ab_bottom":                                # bottom H uses the top's B3 (Haldane2L.h:307)
        np.testing.assert_allclose(J[:, :2, :2], -dH[:, :2, :2], atol=1e-9)
        assert np.all(J[:, :2, 2:] == 0) and np.all(J[:, 2:, :2] == 0)
        m2 = AnalyticModel("Haldane2L", [*m.params[4:8], *m.params[4:8], *m.params[8:]])  # bottom as a top layer
        np.testing.assert_allclose(J[:, 2:, 2:], oracle.jmatrix(m2, k)[:, :2, :2], atol=1e-15)
        assert np.abs(J[:, 2:, 2:] + dH[:, 2:, 2:]).max() > 1e-3   # not -dH/dk of the written H
    elif name == "multiwsm":                                     # MultiWSM.h: J_z = -v sz as well
        dHz = _fd(oracle, m, k)
        np.testing.assert_allclose(J[:2], -dHz[:2], atol=1e-9)
        return
    elif name == "wdsm":                                         # WDSM_TaAs.h:126-185
        a0 = m.params[4]
        np.testing.assert_allclose(J[0], -dH[0] / a0, atol=1e-9)
        np.testing.assert_allclose(J[1], -_fd3(oracle, m, k)[2].real / a0, atol=1e-9)
        return
    elif name.startswith("kanemele"):                            # KaneMele.h:98-176
        blk = np.zeros((4, 4), bool)
        blk[:2, :2] = blk[2:, 2:] = True
        np.testing.assert_allclose(J[:2][:, blk], -dH[:2][:, blk], atol=1e-9)   # Haldane blocks: -dH/dk
        for (i, j) in ((0, 3), (1, 2), (2, 1)):
            np.testing.assert_allclose(J[:2, i, j], dH[:2, i, j], atol=1e-9)    # Rashba: +dH/dk
        # (3,0) misses the i tR factor: -sum_i conj(c_i) e^{-i k.a_i} (-i a_i), non-zero even at tR = 0
        a0 = m.params[5]
        va = np.array([[0, a0], [-math.sqrt(3) / 2 * a0, -0.5 * a0], [math.sqrt(3) / 2 * a0, -0.5 * a0]])
        c = np.array([1.0, -0.5 + 0.5j * math.sqrt(3), -0.5 - 0.5j * math.sqrt(3)])  # cx + i cy
        ph = np.exp(-1j * (va @ k[:2]))
        for ax in range(2):
            assert J[ax, 3, 0] == pytest.approx(-np.sum(c * ph * (-1j * va[:, ax])), abs=1e-14)
        if m.params[2] != 0:
            np.testing.assert_allclose(J[:2, 3, 0], dH[:2, 3, 0] / (1j * m.params[2]), atol=1e-7)


def test_tmdc_gamma_doublet_and_gap(oracle):
    """At Gamma the (dxy, dx2-y2) pair is degenerate (E' of D3h); MoS2's
    direct gap sits at K = (4 pi / 3a, 0) and is about 1.6-1.7 eV in this GGA fit."""
    m = tmdc("MoS2")
    e = np.linalg.eigvalsh(oracle.hamiltonian(m, np.zeros(3)))
    assert abs(e[1] - e[2]) < 1e-14 or abs(e[0] - e[1]) < 1e-14
    a0 = m.params[0]
    K = np.array([4 * math.pi / 3 / a0, 0.0, 0.0])
    eK = np.linalg.eigvalsh(oracle.hamiltonian(m, K)) * 27.211386245988
    assert 1.5 < eK[1] - eK[0] < 1.9


@pytest.mark.skipif(not (REF / "TMDCs.h").exists(), reason="reference sources not present")
def test_tmdc_table_matches_reference_header():
    """The library's parameter table equals the numbers TMDCs.h hard-codes."""
    text = (REF / "TMDCs.h").read_text()
    body = text[text.index("gga_params[6][19] = {") + len("gga_params[6][19] = {"):]
    body = body[: body.index("};")]
    vals = np.array([float(x) for x in re.findall(r"-?\d+\.\d+", body)]).reshape(6, 19)
    for i, (name, a, soc) in enumerate([("MoS2", 3.190, 0.073), ("WS2", 3.191, 0.211), ("MoSe2", 3.326, 0.091),
                                        ("WSe2", 3.325, 0.228), ("MoTe2", 3.557, 0.107), ("WTe2", 3.560, 0.237)]):
        m = tmdc(name)
        assert m.params[0] == pytest.approx(a / 0.529177210903, rel=1e-15)
        assert m.params[1] == pytest.approx(soc / 27.211386245988, rel=1e-15)
        np.testing.assert_allclose(np.array(m.params[2:21]) * 27.211386245988, vals[i], rtol=1e-14)
        assert f'a0 = {a:.3f}/au_angstrom' in text and f'socparam = {soc:.3f}/au_eV' in text


@pytest.mark.parametrize("name", ["wilson_rect", "square", "kanemele_rashba", "xyh", "mos2", "wse2_soc", "h2l_2h",
                                  "h2l_ab_bottom", "bi2se3", "multiwsm", "wdsm"])
def test_model_runs_keep_invariants(oracle, name):
    """Short oracle runs of each class: trace of the k-summed rho conserved,
    inter + intra = the InterIntra = false total, lower bands full before the pulse."""
    m = models()[name]
    n = 6 if name == "multiwsm" else 5  # MultiWSM's J divides by |k_perp|^2: keep k = 0 off the grid
    cfg = small_config(n=n, ncycles=1.0, ellip=0.3, T2=150.0)
    r1 = oracle.run(cfg, m, nthreads=4)
    r0 = oracle.run(small_config(n=n, ncycles=1.0, ellip=0.3, T2=150.0, inter_intra=False), m, nthreads=4)
    tw = r1["info"].total_weight
    np.testing.assert_allclose(np.trace(r1["rho_bar"], axis1=1, axis2=2).real, m.nval * tw, rtol=1e-12)
    np.testing.assert_allclose(r1["j_inter"] + r1["j_intra"], r0["j_inter"], atol=1e-13)
    occ = np.zeros(m.nband)
    occ[: m.nval] = tw
    np.testing.assert_allclose(r1["band_pop"][0], occ, atol=1e-12)
    assert rel_l2(r1["j_inter"] + r1["j_intra"], 0 * r1["j_inter"]) > 0
