"""Pins of the oracle's adaptive Cash-Karp restatement (CPU only).

The coefficients are checked against the Runge-Kutta order conditions
(Butcher trees up to order 5); the 2-band observables against the generic
eigen-decomposition currents, and their time derivatives against central
differences; the stepper against the oracle's fixed-step RK4 at a step size
small enough to be converged, which pins the controller, the sample
interpolation and the window handling together.
"""
import math

import numpy as np
import pytest

from antelope_amd import HaldaneModel, Pulse, RunConfig
from oracle import pyoracle


def _order_conditions(A, c, b):
    """(value, exact) pairs of the Butcher conditions up to order 5 for weights b."""
    Ac = A @ c
    c2, c3 = c * c, c * c * c
    return [
        (b.sum(), 1.0),
        (b @ c, 1 / 2),
        (b @ c2, 1 / 3), (b @ Ac, 1 / 6),
        (b @ c3, 1 / 4), (b @ (c * Ac), 1 / 8), (b @ (A @ c2), 1 / 12), (b @ (A @ Ac), 1 / 24),
        (b @ (c2 * c2), 1 / 5), (b @ (c2 * Ac), 1 / 10), (b @ (c * (A @ c2)), 1 / 15),
        (b @ (c * (A @ Ac)), 1 / 30), (b @ (Ac * Ac), 1 / 20), (b @ (A @ c3), 1 / 20),
        (b @ (A @ (c * Ac)), 1 / 40), (b @ (A @ (A @ c2)), 1 / 60), (b @ (A @ (A @ Ac)), 1 / 120),
    ]


def _tableau():
    c, a, b, bh = pyoracle.cash_karp()
    A = np.zeros((6, 6))
    A[:, :5] = a
    return c, A, b, bh


def test_cash_karp_pair_orders():
    c, A, b, bh = _tableau()
    np.testing.assert_allclose(A.sum(1), c, atol=1e-15)
    conds5 = _order_conditions(A, c, b)
    assert all(abs(v - x) < 1e-14 for v, x in conds5)            # b: order 5
    conds4 = _order_conditions(A, c, bh)
    assert all(abs(v - x) < 1e-14 for v, x in conds4[:8])        # bhat: order 4 ...
    assert max(abs(v - x) for v, x in conds4[8:]) > 1e-4          # ... and not 5


def _random_rho(rng):
    """A 2x2 density matrix (Hermitian, trace 1, |s| < 1)."""
    s = rng.normal(size=3)
    s *= 0.9 / np.linalg.norm(s)
    return 0.5 * np.array([[1 + s[2], s[0] - 1j * s[1]], [s[0] + 1j * s[1], 1 - s[2]]])


@pytest.mark.parametrize("ii", [True, False])
def test_observables_match_generic_currents(ii):
    """The Bloch-form observables equal the eigen-decomposition currents and band
    populations of the generic path (SBEs.h:626-760)."""
    cfg = _cfg(45)
    cfg.inter_intra = ii
    rng = np.random.default_rng(4)
    _, _, atmin, _, nt = pyoracle.laser_eval(cfg, np.zeros(1))
    for _ in range(6):
        k = np.append(rng.uniform(-1, 1, 2), 0.0)
        t = atmin + rng.uniform(0, nt * 0.25)
        rho = _random_rho(rng)
        o, _ = pyoracle.observe(cfg, _MODEL, k, t, rho, np.zeros((2, 2)))
        inter, intra, pop = pyoracle.currents(cfg, _MODEL, k, t, rho)
        np.testing.assert_allclose(o[:2], inter[:2], rtol=1e-10, atol=1e-14)
        np.testing.assert_allclose(o[2:4], intra[:2], rtol=1e-10, atol=1e-14)
        assert abs(o[7] - (pop[1] - pop[0])) < 1e-12
        np.testing.assert_allclose(o[4:7], [2 * rho[0, 1].real, -2 * rho[0, 1].imag, (rho[0, 0] - rho[1, 1]).real],
                                   atol=1e-15)


def test_observable_rates_match_finite_differences():
    """d/dt of the observables along the SBE trajectory (state through rho' = RHS,
    k + A(t) through the laser) against central differences."""
    cfg = _cfg(45)
    rng = np.random.default_rng(5)
    _, _, atmin, _, nt = pyoracle.laser_eval(cfg, np.zeros(1))
    for _ in range(5):
        k = np.append(rng.uniform(-1, 1, 2), 0.0)
        t = atmin + rng.uniform(0.3, 0.7) * nt * 0.25      # inside the pulse
        rho = _random_rho(rng)
        f = pyoracle.rhs(cfg, _MODEL, k, t, rho)
        _, od = pyoracle.observe(cfg, _MODEL, k, t, rho, f)
        eps = 1e-4
        op, _ = pyoracle.observe(cfg, _MODEL, k, t + eps, rho + eps * f, f)
        om, _ = pyoracle.observe(cfg, _MODEL, k, t - eps, rho - eps * f, f)
        # rho +- eps f is the trajectory to first order; its O(eps^2) error cancels in the central difference
        np.testing.assert_allclose(od, (op - om) / (2 * eps), rtol=1e-5, atol=1e-9)


def _cfg(rk, dt=0.25, rtol=1e-8, G=10, E0=0.006):
    return RunConfig(npoints=[G, G], pulses=[Pulse(E0=E0, w0=0.014, ncycles=1.0, ellip=0.5, envelope="sin2")],
                     dt=dt, T2=220.0, rk_order=rk, inter_intra=True, rtol=rtol, atol=1e-3 * rtol)


_MODEL = HaldaneModel(a0=2.683, t1=0.075, t2=0.025, phi0=0.3 * math.pi, M0=2.54 * 0.025)


@pytest.fixture(scope="module")
def converged():
    """RK4 at dt/4 (error about 3e-10 here), sampled on the dt grid."""
    r = pyoracle.run(_cfg(4, dt=0.0625), _MODEL, nthreads=8)
    return r["j_inter"][::4] + r["j_intra"][::4], r["band_pop"][::4]


def _rel(x, ref):
    return float(np.linalg.norm(x - ref[: len(x)]) / np.linalg.norm(ref[: len(x)]))


def test_adaptive_converges(converged):
    jref, pref = converged
    errs = {}
    for rtol in (1e-6, 1e-8):
        r = pyoracle.run_adaptive(_cfg(45, rtol=rtol), _MODEL, nthreads=8)
        errs[rtol] = _rel(r["j_inter"] + r["j_intra"], jref)
        assert _rel(r["band_pop"], pref) < 10 * rtol
        assert r["accepted"].min() > 0
    assert errs[1e-8] < 2e-8, errs
    assert errs[1e-6] > 10 * errs[1e-8], errs
    # the fixed RK4 step of the reference (dt = 0.25) is less accurate than rtol 1e-8
    r4 = pyoracle.run(_cfg(4), _MODEL, nthreads=8)
    assert _rel(r4["j_inter"] + r4["j_intra"], jref) > errs[1e-8]


def test_adaptive_blocks_and_windows(converged):
    """Block size and window boundaries change the step sequence, not the accuracy."""
    jref, _ = converged
    base = pyoracle.run_adaptive(_cfg(45), _MODEL, nthreads=8)
    other = pyoracle.run_adaptive(_cfg(45), _MODEL, block=16, window=97, nthreads=8)
    jb, jo = base["j_inter"] + base["j_intra"], other["j_inter"] + other["j_intra"]
    assert not np.array_equal(jb, jo)
    assert _rel(jo, jref) < 2e-8 and _rel(jo, jb) < 2e-8
    assert len(other["accepted"]) == 7  # ceil(100 / 16) blocks


def test_adaptive_step_underflow_is_an_error():
    with pytest.raises(RuntimeError, match="underflow"):
        pyoracle.run_adaptive(_cfg(45, rtol=1e-30), _MODEL, nthreads=2)
