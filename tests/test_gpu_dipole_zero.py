"""isDipoleZero Wannier models on the GPU (Wannier90.h:180-194,297-387): the
Fourier sums take the phase of R + tau_n - tau_m (tau_n = Re <0n|r|0n> of the
R = 0 block) and there is no E.D term.

- The Haldane model written as N = 2 Wannier blocks with tau_B = -a_0 (the
  identity tests/test_oracle_wannier.py pins on the CPU to 1e-12) gives the
  Haldane path's series on the GPU.
- Synthetic models with orbital offsets, in both gauges and both forms of the
  Fourier sums (separable passes on a product grid, dense GEMM), and a model
  without position blocks (tau = 0, as a _tb.dat without them reads), against
  the FP64 oracle. Bar: 1e-10 relative L2.
"""
import dataclasses
import math

import numpy as np
import pytest

from antelope_amd import HaldaneModel, SBESolver
from antelope_amd.models import synthetic_wannier
from tests.cases import rel_l2, small_config
from tests.test_oracle_wannier import haldane_as_wannier

pytestmark = pytest.mark.gpu
TOL = 1e-10


def _gpu(cfg, model):
    with SBESolver(cfg, model) as s:
        s.run()
        s.synchronize()
        return s.series()


def _check(got, ref, label):
    errs = {"j": rel_l2(got["j_inter"] + got["j_intra"], ref["j_inter"] + ref["j_intra"]),
            "j_inter": rel_l2(got["j_inter"], ref["j_inter"]),
            "band_pop": rel_l2(got["band_pop"], ref["band_pop"]),
            "rho_bar": rel_l2(got["rho_bar"], ref["rho_bar"])}
    print(label, {k: f"{v:.2e}" for k, v in errs.items()}, flush=True)
    assert all(v < TOL for v in errs.values()), (label, errs)


def test_haldane_as_wannier_blocks_on_gpu(oracle):
    h = HaldaneModel(a0=2.683, t1=0.075, t2=0.025, phi0=0.3 * math.pi, M0=2.54 * 0.025)
    w = haldane_as_wannier(h)
    a0 = h.a0
    ax = [2 * math.pi / math.sqrt(3) / a0, 0, 0, 0, 4 * math.pi / 3 / a0, 0, 0, 0, 0]
    cfg = small_config(n=12, ncycles=1.0, bz_axes=ax, bz_origin=[0, 0, 0])
    print("case: Haldane as Wannier blocks, 12x12 Haldane zone, 1 cycle", flush=True)
    g_w, g_h = _gpu(cfg, w), _gpu(cfg, h)
    _check(g_w, g_h, "wannier form vs Haldane path")
    _check(g_w, oracle.run(cfg, w, nthreads=16), "wannier form vs oracle")


@pytest.mark.parametrize("gauge", ["LengthWannier", "VelocityHamiltonian"])
@pytest.mark.parametrize("dense", [False, True])
def test_offset_orbitals_against_oracle(oracle, monkeypatch, gauge, dense):
    model = synthetic_wannier(6, nmax=(1, 1, 1), seed=31, is_dipole_zero=True)
    tau = np.real(model.pos[len(model.rvec) // 2, :, range(6), range(6)])
    assert np.abs(tau).max() > 0.1  # the offsets matter
    if dense:
        monkeypatch.setenv("ANTELOPE_NB_DENSE", "1")
    cfg = dataclasses.replace(small_config(n=(3, 3, 2), ncycles=0.5, dt=0.5), gauge=gauge)
    cfg.pulses[0].E0, cfg.pulses[0].w0 = 0.01, 0.05
    print(f"case: N=6 offsets, {gauge}, {'dense' if dense else 'separable'} sums", flush=True)
    _check(_gpu(cfg, model), oracle.run(cfg, model, nthreads=16), f"{gauge} {'dense' if dense else 'box'}")


def test_model_without_positions(oracle):
    model = dataclasses.replace(synthetic_wannier(4, nmax=(1, 1, 0), seed=5), pos=None, is_dipole_zero=True)
    cfg = small_config(n=(4, 4, 1), ncycles=0.5, dt=0.5)
    cfg.pulses[0].E0, cfg.pulses[0].w0 = 0.01, 0.05
    print("case: N=4 without position blocks", flush=True)
    _check(_gpu(cfg, model), oracle.run(cfg, model, nthreads=16), "no positions")


def test_sampled_haldane_class_on_gpu(oracle):
    """The Haldane class's H(k) (Haldane.h:86-92) sampled by sbe_sample_material
    with its orbital offsets runs on the GPU as the Haldane path does."""
    from antelope_amd.material import sample_material
    h = HaldaneModel(a0=2.683, t1=0.075, t2=0.025, phi0=0.3 * math.pi, M0=2.54 * 0.025)
    a0 = h.a0
    L = np.array([[math.sqrt(3) * a0, 0, 0], [math.sqrt(3) / 2 * a0, 1.5 * a0, 0], [0, 0, 10.0]])
    w, dev = sample_material(lambda k: oracle.hamiltonian(h, k), 2, L, (1, 1, 0), tau=[[0, 0, 0], [0, -a0, 0]])
    assert dev < 1e-15
    ax = [2 * math.pi / math.sqrt(3) / a0, 0, 0, 0, 4 * math.pi / 3 / a0, 0, 0, 0, 0]
    cfg = small_config(n=12, ncycles=1.0, bz_axes=ax, bz_origin=[0, 0, 0])
    print("case: sampled Haldane class, 12x12 Haldane zone, 1 cycle", flush=True)
    _check(_gpu(cfg, w), _gpu(cfg, h), "sampled class vs Haldane path")
