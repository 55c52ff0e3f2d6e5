"""Adaptive Cash-Karp stepper on the GPU against the oracle's restatement of
the same method (FP64: same step sequence, so agreement to rounding), the
FP32-state form's error, batches and the error path."""
import math

import numpy as np
import pytest

from antelope_amd import HaldaneModel, Pulse, RunConfig, SBEBatch, SBESolver
from antelope_amd import _abi as A
from antelope_amd.solver import SBEError
from oracle import pyoracle

pytestmark = pytest.mark.gpu

MODEL = HaldaneModel(a0=2.683, t1=0.075, t2=0.025, phi0=0.3 * math.pi, M0=2.54 * 0.025)


def _cfg(precision="fp64", G=16, rtol=1e-8, ncycles=1.0, E0=0.006, ellip=0.5, cep=0.0, G2=None):
    return RunConfig(npoints=[G, G2 or G], pulses=[Pulse(E0=E0, w0=0.014, ncycles=ncycles, ellip=ellip, cep=cep,
                                                   envelope="sin2")],
                     dt=0.25, T2=220.0, rk_order=45, inter_intra=True, precision=precision,
                     rtol=rtol, atol=1e-3 * rtol)


def _rel(x, ref):
    return float(np.linalg.norm(x - ref) / np.linalg.norm(ref))


def _gpu(cfg, model=MODEL, slices=None):
    with SBESolver(cfg, model) as s:
        if slices:
            for b in range(0, s.nt, slices):
                s.run(b, min(s.nt, b + slices))
        else:
            s.run()
        return s.series(), s.step_stats()


def test_fp64_matches_oracle_restatement():
    cfg = _cfg(G=32, G2=16)   # 512 k points: whole step-control blocks, no padding
    print("case: 32x16 grid, 1 cycle, fp64, rtol 1e-8", flush=True)
    g, st = _gpu(cfg)
    r = pyoracle.run_adaptive(cfg, MODEL, block=st.k_per_block, nthreads=8)
    for key in ("j_inter", "j_intra", "band_pop"):
        assert _rel(g[key], r[key]) < 1e-10, key
    assert _rel(g["rho_bar"], r["rho_bar"]) < 1e-10
    assert st.blocks * st.k_per_block == 512 and st.blocks == len(r["accepted"])
    assert st.accepted == r["accepted"].sum() and st.rejected == r["rejected"].sum()
    assert st.max_accepted == r["accepted"].max() and st.min_accepted == r["accepted"].min()


def test_sliced_runs_match_oracle_windows():
    cfg = _cfg(ncycles=1.5, cep=1.0)
    print("case: sbe_run in slices of 97 samples", flush=True)
    g, st = _gpu(cfg, slices=97)
    r = pyoracle.run_adaptive(cfg, MODEL, block=st.k_per_block, window=97, nthreads=8)
    assert _rel(g["j_inter"] + g["j_intra"], r["j_inter"] + r["j_intra"]) < 1e-10
    assert _rel(g["band_pop"], r["band_pop"]) < 1e-10


def test_fp32_state_error_vs_fp64_oracle():
    """FP32 state (Kahan-compensated) with FP64 error control and k sums."""
    print("case: fp32 state, 24x24 grid", flush=True)
    g, st = _gpu(_cfg("fp32", G=24))
    r = pyoracle.run_adaptive(_cfg(G=24), MODEL, block=st.k_per_block, nthreads=8)
    ej = _rel(g["j_inter"] + g["j_intra"], r["j_inter"] + r["j_intra"])
    ep = _rel(g["band_pop"], r["band_pop"])
    print(f"fp32 adaptive: j {ej:.2e}, populations {ep:.2e}, accepted/block {st.accepted / st.blocks:.0f}")
    assert ej < 1e-6 and ep < 1e-7


def test_batch_equals_single_solves():
    cfgs = [_cfg(E0=0.002, ncycles=1.0), _cfg(E0=0.005, ncycles=1.5, ellip=0.2, G=12),
            _cfg(E0=0.006, ncycles=2.0, ellip=1.0, cep=2.0, G=20)]
    models = [MODEL, HaldaneModel(a0=2.683, t1=0.075, t2=0.025, phi0=0.1, M0=0.02),
              HaldaneModel(a0=2.683, t1=0.075, t2=0.025, phi0=1.2, M0=0.09)]
    for prec in ("fp64", "fp32"):
        print(f"case: batch of 3 adaptive solves, {prec}", flush=True)
        cs = [_cfg(prec, E0=c.pulses[0].E0, ncycles=c.pulses[0].ncycles, ellip=c.pulses[0].ellip,
                   cep=c.pulses[0].cep, G=c.npoints[0]) for c in cfgs]
        with SBEBatch(cs, models) as b:
            b.run()
            got = [(b.series(i), b.step_stats(i)) for i in range(3)]
        for (gs, gst), c, m in zip(got, cs, models):
            s, st = _gpu(c, m)
            for key in s:
                assert np.array_equal(gs[key], s[key]), key
            assert (gst.accepted, gst.rejected) == (st.accepted, st.rejected)


def test_batch_groups_partial_runs_and_rerun():
    """Five solves of one to six cycles (one to five launch windows) in the batch's two
    solve groups, advanced in parts that end on window boundaries (2048 samples: the
    windows are part of the method, so a part ending elsewhere defines other windows),
    then reset and run again: every solve's series and step counts equal its solve
    alone, both times."""
    e0 = [0.006, 0.001, 0.004, 0.0025, 0.005]
    cyc = [6.0, 1.0, 4.0, 2.5, 5.0]
    grids = [16, 12, 20, 8, 14]
    cs = [_cfg(E0=e, ncycles=n, G=g, ellip=0.3 + 0.1 * i) for i, (e, n, g) in enumerate(zip(e0, cyc, grids))]
    ms = [MODEL] * len(cs)
    print("case: batch of 5 adaptive solves, 1..6 cycles, runs of 2048 and 4096 samples", flush=True)
    with SBEBatch(cs, ms) as b:
        runs = []
        for _ in range(2):
            b.run(2048)
            b.run(4096)
            b.run()
            runs.append([(b.series(i), b.step_stats(i)) for i in range(len(cs))])
            b.reset()
    for i, (c, m) in enumerate(zip(cs, ms)):
        s, st = _gpu(c, m)
        for gs, gst in (runs[0][i], runs[1][i]):
            for key in s:
                assert np.array_equal(gs[key], s[key]), (i, key)
            assert (gst.accepted, gst.rejected) == (st.accepted, st.rejected), i


def test_step_size_underflow_is_reported():
    print("case: rtol 1e-30 (steps shrink below 1e-6 dt)", flush=True)
    with SBESolver(_cfg(rtol=1e-30, G=8), MODEL) as s:
        s.run()
        with pytest.raises(SBEError) as e:
            s.series()
    assert e.value.code == A.SBE_ERR_NAN
