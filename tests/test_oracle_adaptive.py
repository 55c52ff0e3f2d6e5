"""Pins of the oracle's adaptive Cash-Karp restatement (CPU only).

The coefficients are checked against the Runge-Kutta order conditions
(Butcher trees up to order 5), written out here independently of how they
were derived; the stepper is checked against the oracle's fixed-step RK4 at a
step size small enough to be converged, which pins the controller, the
dense output and the window handling together.
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
    c, a, b, bh, beta = pyoracle.cash_karp()
    A = np.zeros((6, 6))
    A[:, :5] = a
    return c, A, b, bh, beta


def test_cash_karp_pair_orders():
    c, A, b, bh, _ = _tableau()
    np.testing.assert_allclose(A.sum(1), c, atol=1e-15)
    conds5 = _order_conditions(A, c, b)
    assert all(abs(v - x) < 1e-14 for v, x in conds5)            # b: order 5
    conds4 = _order_conditions(A, c, bh)
    assert all(abs(v - x) < 1e-14 for v, x in conds4[:8])        # bhat: order 4 ...
    assert max(abs(v - x) for v, x in conds4[8:]) > 1e-4          # ... and not 5


def test_dense_output_weights():
    """b_i(th) over the six stages and f(t+h, y5): order 4 for every th, C1 at both ends."""
    c, A6, b, _, beta = _tableau()
    A = np.zeros((7, 7))
    A[:6, :6] = A6
    A[6, :6] = b          # the seventh stage evaluates f at y5 = y + h sum b_i f_i
    c7 = np.append(c, 1.0)
    bt = lambda th: np.array([sum(beta[i][m] * th ** (m + 1) for m in range(4)) for i in range(7)])
    dbt = lambda th: np.array([sum((m + 1) * beta[i][m] * th ** m for m in range(4)) for i in range(7)])
    for th in (0.0, 0.13, 0.5, 0.77, 1.0):
        w = bt(th)
        Ac = A @ c7
        got = [w.sum(), w @ c7, w @ c7 ** 2, w @ Ac, w @ c7 ** 3, w @ (c7 * Ac), w @ (A @ c7 ** 2), w @ (A @ Ac)]
        want = [th, th ** 2 / 2, th ** 3 / 3, th ** 3 / 6, th ** 4 / 4, th ** 4 / 8, th ** 4 / 12, th ** 4 / 24]
        np.testing.assert_allclose(got, want, atol=1e-14)
    np.testing.assert_allclose(bt(1.0), np.append(b, 0.0), atol=1e-14)
    np.testing.assert_allclose(dbt(0.0), np.eye(7)[0], atol=1e-14)
    np.testing.assert_allclose(dbt(1.0), np.eye(7)[6], atol=1e-14)


def test_dense_output_local_order():
    """On y' = i w y one step of Cash-Karp + the continuous extension has interior
    error O(h^5): halving h divides it by about 32."""
    c, A6, b, _, beta = _tableau()
    w = 1.3

    def interp_err(h, th=0.4):
        f = lambda y: 1j * w * y
        k = []
        for i in range(6):
            k.append(f(1.0 + h * sum(A6[i, j] * k[j] for j in range(i))))
        y5 = 1.0 + h * sum(b[i] * k[i] for i in range(6))
        k.append(f(y5))
        bt = [sum(beta[i][m] * th ** (m + 1) for m in range(4)) for i in range(7)]
        return abs(1.0 + h * sum(bt[i] * k[i] for i in range(7)) - np.exp(1j * w * th * h))

    r = interp_err(0.2) / interp_err(0.1)
    assert 25 < r < 40, r


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
