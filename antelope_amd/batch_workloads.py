"""Seeded batches of Haldane solves (the batched configuration): per solve
E0 log-uniform in [0.001, 0.006], ellipticity uniform in [0, 1], CEP uniform
in [0, 360) degrees, phi0 uniform in [0, pi/2], M0/t2 uniform in [0, 4] and
6..16 cycles, on square grids, from SplitMix64 so any language reproduces
the same batch."""
from __future__ import annotations

import math

from .models import SplitMix64
from .params import HaldaneModel, Pulse, RunConfig


def haldane_batch(n: int = 64, grid: int = 200, seed: int = 7, dt: float = 0.25, rk_order=4,
                  precision: str = "fp64", T2: float = 220.0, cycles=(6, 16), rtol: float = 1e-8,
                  atol: float = 1e-8):
    rng = SplitMix64(seed)
    cfgs, models = [], []
    t1 = 0.075
    for _ in range(n):
        E0 = math.exp(math.log(0.001) + rng.uniform() * (math.log(0.006) - math.log(0.001)))
        ellip = rng.uniform()
        cep = 2 * math.pi * rng.uniform()
        phi0 = 0.5 * math.pi * rng.uniform()
        mt2 = 4.0 * rng.uniform()
        ncyc = cycles[0] + int(rng.uniform() * (cycles[1] - cycles[0] + 1))
        models.append(HaldaneModel(a0=2.683, t1=t1, t2=t1 / 3, phi0=phi0, M0=mt2 * t1 / 3))
        cfgs.append(RunConfig(npoints=[grid, grid], pulses=[Pulse(E0=E0, w0=0.014, ncycles=float(ncyc), ellip=ellip,
                                                                   cep=cep, envelope="sin2")],
                              dt=dt, T2=T2, rk_order=rk_order, inter_intra=True, precision=precision,
                              rtol=rtol, atol=atol))
    return cfgs, models
