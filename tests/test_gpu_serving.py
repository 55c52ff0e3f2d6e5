"""Serving loop: requests that arrive while others run are admitted as slots
and budget free up, run overlapped, and each returns exactly what it returns
when run alone."""
import numpy as np
import pytest

from antelope_amd import SBESolver
from antelope_amd import _abi as A
from antelope_amd.serve import SBEServer, run_stream, serving_stream
from antelope_amd.solver import SBEError, query_footprint

pytestmark = pytest.mark.gpu


def test_mixed_requests_match_alone_runs():
    reqs = serving_stream(8, rate=200.0, seed=3, wannier_share=0.3, grid_range=(16, 40), cycles=(1, 3),
                          wannier_cycles=(1, 1))
    assert {r.kind for r in reqs} >= {"haldane"}
    budget = 2 * max(query_footprint(r.config, r.model) for r in reqs)
    with SBEServer(hbm_budget_bytes=budget, max_active=3, slice_steps=200) as srv:
        out = run_stream(srv, reqs, realtime=True, keep_results=True)
    assert out["completed"] == len(reqs) and out["failed"] == 0
    assert out["max_active"] <= 3
    print({k: out[k] for k in ("wall_s", "p50_s", "p99_s", "k_steps_per_s", "max_active")}, flush=True)
    for r, rec in zip(reqs, out["records"]):
        with SBESolver(r.config, r.model) as s:
            s.run()
            alone = s.series()
        for k in alone:
            assert np.array_equal(rec["series"][k], alone[k]), (r.kind, k)


def test_admission_respects_budget():
    reqs = serving_stream(6, rate=1000.0, seed=9, wannier_share=0.0, grid_range=(30, 30), cycles=(1, 1))
    fp = query_footprint(reqs[0].config, reqs[0].model)
    with SBEServer(hbm_budget_bytes=int(1.5 * fp), max_active=8, slice_steps=300) as srv:
        out = run_stream(srv, reqs, realtime=False)
        assert out["max_active"] == 1  # only one fits at a time
        assert out["completed"] == 6
    with SBEServer(hbm_budget_bytes=fp // 2, max_active=8, slice_steps=300) as srv:
        with pytest.raises(SBEError) as e:
            srv.submit(reqs[0].config, reqs[0].model)
        assert e.value.code == A.SBE_ERR_OOM
