"""The launch-plan autotuner on the GPU: it times its candidates and caches the
fastest; contexts created afterwards follow the cached plan. A plan that only
changes steps per launch gives the same bits as the default plan; one that
changes k points per lane or the N-band kernels agrees to rounding."""
import numpy as np
import pytest

from antelope_amd import SBEBatch, SBESolver
from antelope_amd.batch_workloads import haldane_batch
from antelope_amd.models import synthetic_wannier
from antelope_amd.solver import autotune, batch_autotune, tuning_clear, tuning_get, tuning_set
from tests.cases import haldane_cfg1_model, rel_l2, small_config

pytestmark = pytest.mark.gpu


def _series(cfg, model):
    with SBESolver(cfg, model) as s:
        s.run()
        s.synchronize()
        return s.series()


def _close(a, b, tol):
    for k in a:
        den = np.linalg.norm(b[k])
        err = np.linalg.norm(a[k] - b[k]) / den if den > 0 else np.abs(a[k]).max()
        assert err < tol, (k, err)


def test_autotune_haldane_and_same_bits_for_launch_shape():
    tuning_clear()
    m = haldane_cfg1_model()
    cfg = small_config(n=300, ncycles=1.0)
    default = _series(cfg, m)
    print("case: autotune 300x300 fixed-step", flush=True)
    t = autotune(cfg, m, trial_steps=128)
    assert t["tuned"] == 1 and t["candidates"] == 4 and t["ms_per_step"] > 0 and t["kpt"] in (1, 2)
    assert tuning_get(cfg, m) == t
    _close(_series(cfg, m), default, 1e-12)       # the tuned plan: rounding at most
    kpt0 = 2                                        # the default rule for 90,000 FP64 k points
    tuning_set(cfg, m, kpt=kpt0, chunk_steps=37)   # launch shape only: the same bits
    got = _series(cfg, m)
    for k in default:
        assert np.array_equal(got[k], default[k]), k
    tuning_set(cfg, m, kpt=1)                       # another grouping of the wave sums
    _close(_series(cfg, m), default, 1e-12)
    tuning_clear()


def test_autotune_nband_kernels_and_sums():
    tuning_clear()
    model = synthetic_wannier(8, nmax=(1, 1, 1), seed=3)
    cfg = small_config(n=(4, 4, 2), ncycles=0.25, dt=0.5)
    cfg.pulses[0].E0, cfg.pulses[0].w0 = 0.01, 0.05
    default = _series(cfg, model)
    print("case: autotune N=8 Wannier", flush=True)
    t = autotune(cfg, model, trial_steps=16)
    assert t["tuned"] == 1 and t["candidates"] == 4 and t["nb_kernels"] in (1, 2) and t["nb_sums"] in (1, 2)
    _close(_series(cfg, model), default, 1e-12)
    tuning_clear()


def test_batch_autotune_adaptive():
    tuning_clear()
    cfgs, models = haldane_batch(n=4, grid=40, cycles=(1, 1), rk_order=45, rtol=1e-8, atol=1e-9)
    print("case: batch autotune, 4 adaptive solves", flush=True)
    t = batch_autotune(cfgs, models, trial_steps=256)
    assert t["tuned"] == 1 and t["candidates"] == 2 and t["kpt"] in (1, 2)
    assert tuning_get(cfgs[0], models[0])["kpt"] == t["kpt"]
    with SBEBatch(cfgs, models) as b:
        b.run()
        st = b.step_stats(0)
        assert st.k_per_block == 64 * t["kpt"]
    tuning_clear()
