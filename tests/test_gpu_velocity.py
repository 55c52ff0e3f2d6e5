"""VelocityHamiltonian gauge on the GPU against the FP64 CPU oracle.

The oracle restates SBEs.h's generic N x N velocity-gauge path (per-k E, U,
pMatrix = U^dag J U built once; band-basis rho; relaxation toward rho0;
j = Tr(rho P) + Nval A). The 2-band kernel propagates the band-basis Bloch
vector in closed form, the N-band kernel the band-basis matrix on FP64 MFMA.
Bar for FP64: 1e-10 relative L2 over the whole series for j (inter, intra,
total), band populations and the band-basis k-summed rho (whose off-diagonal
elements use the eigenvector phase rule both sides share), and for the final
state in both bases.
"""
import numpy as np
import pytest

from antelope_amd import SBESolver
from antelope_amd.models import synthetic_wannier
from tests.cases import haldane_cfg1_model, rel_l2, small_config

pytestmark = pytest.mark.gpu

TOL64 = 1e-10


def _errs(gpu, ref):
    e = {
        "j_total": rel_l2(gpu["j_inter"] + gpu["j_intra"], ref["j_inter"] + ref["j_intra"]),
        "j_inter": rel_l2(gpu["j_inter"], ref["j_inter"]),
        "band_pop": rel_l2(gpu["band_pop"], ref["band_pop"]),
        "rho_bar": rel_l2(gpu["rho_bar"], ref["rho_bar"]),
    }
    if np.linalg.norm(ref["j_intra"]) > 0:
        e["j_intra"] = rel_l2(gpu["j_intra"], ref["j_intra"])
    else:
        e["j_intra"] = float(np.abs(gpu["j_intra"]).max())
    return e


def _run(cfg, model, splits=None, basis=("band", "orbital")):
    with SBESolver(cfg, model) as s:
        b = 0
        for e in (splits or []) + [s.nt]:
            s.run(b, e)
            b = e
        s.synchronize()
        return s.series(), {x: s.state(x) for x in basis}


def _orbital(oracle, cfg, model, state):
    """U rho_H U^dag per k with the oracle's (phase-fixed) eigenvectors."""
    k, *_ = oracle.kgrid(cfg, model)
    out = np.empty_like(state)
    for i in range(len(k)):
        H = oracle.hamiltonian(model, k[i])
        ev, U = oracle.eigh(H)
        N = U.shape[0]
        for j in range(N):
            p = next(r for r in range(N) if abs(U[r, j]) ** 2 * 2 * N >= 1.0)
            U[:, j] *= np.conj(U[p, j]) / abs(U[p, j])
        out[i] = U @ state[i] @ U.conj().T
    return out


@pytest.mark.parametrize(
    "label,kw",
    [
        ("rk4_t2", dict()),
        ("rk5_fehlberg", dict(rk=5)),
        ("no_dephasing", dict(T2=-1.0)),
        ("t1_and_t2", dict(T1=150.0, T2=90.0)),
        ("total_current_only", dict(inter_intra=False)),
        ("elliptic_cos4", dict(env="cos4", ellip=0.5)),
        ("fermi_dirac", dict(init="fermi", fermi_energy=0.01)),
        ("ragged_mp", dict(n=(7, 5), grid="mp")),
    ],
)
def test_haldane_velocity_parity(oracle, label, kw):
    cfg = small_config(gauge="VelocityHamiltonian", **kw)
    model = haldane_cfg1_model()
    print(label, flush=True)
    ref = oracle.run(cfg, model, nthreads=8, want_state=True)
    gpu, st = _run(cfg, model)
    e = _errs(gpu, ref)
    e["state_band"] = rel_l2(st["band"], ref["state"])
    e["state_orbital"] = rel_l2(st["orbital"], _orbital(oracle, cfg, model, ref["state"]))
    print(label, {k: f"{v:.2e}" for k, v in e.items()}, flush=True)
    assert all(v <= TOL64 for v in e.values()), e


def test_haldane_velocity_sliced_runs_same_bits():
    cfg = small_config(gauge="VelocityHamiltonian", n=9)
    model = haldane_cfg1_model()
    a, _ = _run(cfg, model, basis=())
    b, _ = _run(cfg, model, splits=[1, 37, 500], basis=())
    for key in ("j_inter", "j_intra", "rho_bar", "band_pop"):
        assert np.array_equal(a[key], b[key]), key


def test_haldane_velocity_fp32(oracle):
    """FP32 band-basis state with Kahan-compensated updates and FP64 k sums:
    the band basis has no large cancelling terms, so the error stays near
    FP32 rounding of the per-k constants."""
    cfg = small_config(gauge="VelocityHamiltonian", precision="fp32")
    model = haldane_cfg1_model()
    ref = oracle.run(small_config(gauge="VelocityHamiltonian"), model, nthreads=8)
    gpu, _ = _run(cfg, model, basis=())
    e = _errs(gpu, ref)
    print("fp32", {k: f"{v:.2e}" for k, v in e.items()}, flush=True)
    assert e["j_total"] < 1e-5 and e["band_pop"] < 1e-6 and e["rho_bar"] < 1e-5


@pytest.mark.parametrize(
    "label,N,kw",
    [
        ("n4_rk4", 4, dict()),
        ("n4_rk5_t1", 4, dict(rk=5, T1=150.0, T2=90.0)),
        ("n4_total_only", 4, dict(inter_intra=False, T2=-1.0)),
        ("n8_fermi", 8, dict(init="fermi", fermi_energy=0.0)),
        ("n16", 16, dict()),
        ("n32", 32, dict()),
    ],
)
def test_wannier_velocity_parity(oracle, label, N, kw):
    model = synthetic_wannier(N, nmax=(1, 1, 0), seed=21 + N)
    model.nval = N // 2
    npts = (4, 3) if N >= 16 else (5, 4)
    cfg = small_config(gauge="VelocityHamiltonian", n=npts, ncycles=1.0, dt=0.5, ellip=0.3, **kw)
    print(label, flush=True)
    ref = oracle.run(cfg, model, nthreads=8, want_state=True)
    gpu, st = _run(cfg, model, splits=[3])
    e = _errs(gpu, ref)
    e["state_band"] = rel_l2(st["band"], ref["state"])
    e["state_orbital"] = rel_l2(st["orbital"], _orbital(oracle, cfg, model, ref["state"]))
    print(label, {k: f"{v:.2e}" for k, v in e.items()}, flush=True)
    assert all(v <= TOL64 for v in e.values()), e
