"""Shared problem definitions for the tests (seeded, small)."""
import math

import numpy as np

from antelope_amd import HaldaneModel, Pulse, RunConfig

T_PERIOD = 2 * math.pi / 0.014


def haldane_cfg1_model():
    """The plain-method reference case: a = 2.683 bohr, t1 = 0.075, t2 = t1/3,
    phi0 = 0.3 pi, M0 = 2.54 t2."""
    t1 = 0.075
    t2 = t1 / 3
    return HaldaneModel(a0=2.683, t1=t1, t2=t2, phi0=0.3 * math.pi, M0=2.54 * t2)


def small_config(n=12, ncycles=2.0, rk=4, T2=220.0, inter_intra=True, env="sin2", ellip=0.0,
                 dt=0.25, precision="fp64", **kw):
    p = Pulse(E0=0.0045, w0=0.014, ncycles=ncycles, ellip=ellip, envelope=env)
    return RunConfig(npoints=[n, n] if np.isscalar(n) else list(n), pulses=[p], dt=dt, T2=T2,
                     rk_order=rk, inter_intra=inter_intra, precision=precision, **kw)


def rel_l2(x, ref):
    x = np.asarray(x)
    ref = np.asarray(ref)
    den = np.linalg.norm(ref)
    num = np.linalg.norm(x - ref)
    return num / den if den > 0 else num
