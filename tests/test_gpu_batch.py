"""Batched solves: every solve of a mixed batch (different models, pulses,
grids and lengths) returns exactly the series it returns alone."""
import numpy as np
import pytest

from antelope_amd import SBEBatch, SBESolver
from antelope_amd.batch_workloads import haldane_batch

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("precision", ["fp64", "fp32"])
def test_batch_equals_alone(precision):
    cfgs, models = haldane_batch(n=5, grid=24, cycles=(1, 3), precision=precision, seed=11)
    cfgs[1].npoints = [17, 9]           # ragged grid in the mix
    cfgs[2].T2 = -1.0                   # one solve without dephasing
    with SBEBatch(cfgs, models) as b:
        b.run(700)                      # partial advance, then the rest
        b.run()
        got = [b.series(i) for i in range(len(cfgs))]
        assert [b.info(i).steps_done for i in range(len(cfgs))] == [b.info(i).nt for i in range(len(cfgs))]
    for i, (c, m) in enumerate(zip(cfgs, models)):
        with SBESolver(c, m) as s:
            s.run()
            alone = s.series()
        for k in alone:
            assert np.array_equal(got[i][k], alone[k]), (i, k)


def test_batch_member_large_fp64_grid():
    """A FP64 member of 65,536 k points (alone it takes two k points per lane) in a
    plain batch: the batch kernel runs one k point per lane for every member, so
    the member is planned that way; its series match its solve alone to rounding
    (the wave sums group k points differently) and the small member's bit for bit."""
    cfgs, models = haldane_batch(n=2, grid=20, cycles=(1, 1), precision="fp64", seed=5)
    cfgs[0].npoints = [256, 256]
    with SBEBatch(cfgs, models) as b:
        b.run()
        got = [b.series(i) for i in range(2)]
    for i in range(2):
        with SBESolver(cfgs[i], models[i]) as s:
            s.run()
            alone = s.series()
        for k in alone:
            if i == 1:
                assert np.array_equal(got[i][k], alone[k]), (i, k)
            else:
                den = np.linalg.norm(alone[k])
                err = np.linalg.norm(got[i][k] - alone[k]) / den if den > 0 else np.abs(got[i][k]).max()
                assert err < 1e-12, (k, err)


@pytest.mark.parametrize("rk_order", [4, 45])
def test_batch_timing_is_busy_time(rk_order):
    """The batch runs its solves in two groups whose step launches overlap; the
    reported step time is the union of the launch intervals, so it is positive
    and no longer than the wall time around the run (a sum over overlapping
    launches could exceed it). Running it again gives the same series."""
    import time
    cfgs, models = haldane_batch(n=6, grid=40, cycles=(2, 4), rk_order=rk_order, precision="fp64", seed=3)
    with SBEBatch(cfgs, models) as b:
        b.timing()                      # clears anything recorded before
        t0 = time.perf_counter()
        b.run()
        b.synchronize()
        wall_ms = (time.perf_counter() - t0) * 1e3
        tim = b.timing()
        first = [b.series(i) for i in range(len(cfgs))]
        b.reset()
        b.run()
        again = [b.series(i) for i in range(len(cfgs))]
    print("case: batch timing rk", rk_order, tim, "wall ms", wall_ms, flush=True)
    assert tim["launches"] >= 2
    assert 0.0 < tim["step_ms"] <= wall_ms
    for i in range(len(cfgs)):
        for k in first[i]:
            assert np.array_equal(first[i][k], again[i][k]), (i, k)
