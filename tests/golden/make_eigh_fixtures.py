"""Writes eigh_fixtures.json: seeded Hermitian matrices with their
eigenvalues and spectral projectors from numpy.linalg.eigh (LAPACK zheevd),
used as a library. Projectors are independent of eigenvector phases, so they
pin the oracle's Jacobi solver without depending on zheevr's phase choice.

    python tests/golden/make_eigh_fixtures.py
"""
import json
from pathlib import Path

import numpy as np


def main():
    cases = []
    for n, seed in ((2, 11), (3, 12), (4, 13), (8, 14)):
        rng = np.random.default_rng(seed)
        X = rng.normal(size=(n, n)) + 1j * rng.normal(size=(n, n))
        H = (X + X.conj().T) / 2
        ev, U = np.linalg.eigh(H)
        P = [np.outer(U[:, j], U[:, j].conj()) for j in range(n)]
        cases.append({
            "n": n, "seed": seed,
            "H_re": H.real.tolist(), "H_im": H.imag.tolist(),
            "evals": ev.tolist(),
            "P_re": [p.real.tolist() for p in P], "P_im": [p.imag.tolist() for p in P],
        })
    out = Path(__file__).with_name("eigh_fixtures.json")
    out.write_text(json.dumps({"generator": "numpy.linalg.eigh " + np.__version__, "cases": cases}))
    print(out)


if __name__ == "__main__":
    main()
