"""Seeded synthetic material models for tests, benchmarks and serving.

`synthetic_wannier` builds a tight-binding model in the data layout of the
reference's Wannier90 reader (Materials/Wannier90.h:107-218: lattice, integer
R vectors with degeneracy weights, <0m|H|Rn>, <0m|r|Rn>): complex Gaussian
hoppings scaled by `scale * exp(-|n| / decay)` (|n| = Euclidean norm of the
integer R index), H(-R) = H(R)^dagger so H(k) is Hermitian, and on-site
energies spread evenly over [-onsite, onsite] to open gaps. Position matrices
are generated the same way at `pos_scale`. The random numbers come from
SplitMix64 + Box-Muller so any language can regenerate the same model.
"""
from __future__ import annotations

import math

import numpy as np

from .params import WannierModel

_MASK = (1 << 64) - 1


class SplitMix64:
    """Steele, Lea & Flood's SplitMix64 (2014)."""

    def __init__(self, seed: int):
        self.state = seed & _MASK

    def next_u64(self) -> int:
        self.state = (self.state + 0x9E3779B97F4A7C15) & _MASK
        z = self.state
        z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & _MASK
        z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & _MASK
        return z ^ (z >> 31)

    def uniform(self) -> float:
        """(0, 1) with 53 random bits."""
        return ((self.next_u64() >> 11) + 0.5) * (1.0 / (1 << 53))

    def normal_pair(self):
        u1, u2 = self.uniform(), self.uniform()
        r = math.sqrt(-2.0 * math.log(u1))
        return r * math.cos(2 * math.pi * u2), r * math.sin(2 * math.pi * u2)

    def cnormal(self) -> complex:
        """Complex Gaussian with E|z|^2 = 1."""
        a, b = self.normal_pair()
        return complex(a, b) / math.sqrt(2.0)


def synthetic_wannier(nband: int = 16, nmax=(3, 3, 1), lattice=None, scale: float = 0.05,
                      decay: float = 1.5, onsite: float = 0.3, pos_scale: float = 0.5,
                      nval: int | None = None, seed: int = 2024, with_pos: bool = True,
                      is_dipole_zero: bool = False) -> WannierModel:
    N = int(nband)
    if lattice is None:
        lattice = np.diag([5.0, 5.0, 10.0])
    rng = SplitMix64(seed)
    rv = [(a, b, c) for a in range(-nmax[0], nmax[0] + 1) for b in range(-nmax[1], nmax[1] + 1)
          for c in range(-nmax[2], nmax[2] + 1)]
    index = {r: i for i, r in enumerate(rv)}
    nr = len(rv)
    ham = np.zeros((nr, N, N), complex)
    pos = np.zeros((nr, 3, N, N), complex)
    done = set()
    for r in rv:
        if r in done:
            continue
        neg = (-r[0], -r[1], -r[2])
        i, j = index[r], index[neg]
        amp = scale * math.exp(-math.sqrt(r[0] ** 2 + r[1] ** 2 + r[2] ** 2) / decay)
        pamp = pos_scale * math.exp(-math.sqrt(r[0] ** 2 + r[1] ** 2 + r[2] ** 2) / decay)
        H = np.array([[rng.cnormal() for _ in range(N)] for _ in range(N)]) * amp
        P = np.array([[[rng.cnormal() for _ in range(N)] for _ in range(N)] for _ in range(3)]) * pamp
        if i == j:  # R = 0: Hermitian on-site block with spread diagonal
            H = 0.5 * (H + H.conj().T)
            H[np.diag_indices(N)] = np.linspace(-onsite, onsite, N)
            P = 0.5 * (P + np.conj(np.transpose(P, (0, 2, 1))))
            ham[i], pos[i] = H, P
        else:
            ham[i], ham[j] = H, H.conj().T
            pos[i], pos[j] = P, np.conj(np.transpose(P, (0, 2, 1)))
        done.update({r, neg})
    return WannierModel(lattice=np.asarray(lattice, float), rvec=np.array(rv, np.int32),
                        rweight=np.ones(nr), ham=ham, pos=pos if with_pos else None,
                        nval=N // 2 if nval is None else nval, is_dipole_zero=is_dipole_zero)
