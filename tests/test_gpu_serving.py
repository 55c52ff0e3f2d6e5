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


def _served(srv_args, reqs):
    """Series of every request, served together through one server."""
    with SBEServer(**srv_args) as srv:
        out = run_stream(srv, reqs, realtime=False, keep_results=True)
    return out


def test_serving_configuration_sizes():
    """The serving configuration's grids: Haldane requests drawn from 100^2..400^2
    (pulses shortened to 1..2 cycles, 1795..3590 steps, so the test runs in seconds)
    and Wannier requests with N = 8, 16 and 32 (reduced as in the bench stream: 8^2..16^2
    grids, 1 cycle, 5 x 5 x 1 R). Every served result equals the same request served
    alone, bit for bit; the smallest Haldane request and the N = 8 request match the FP64
    oracle to 1e-10 (relative L2)."""
    from oracle import pyoracle
    from tests.cases import rel_l2
    reqs = serving_stream(9, rate=100.0, seed=5, wannier_share=0.34, grid_range=(100, 400), cycles=(1, 2),
                          wannier_cycles=(1, 1))
    kinds = [r.kind for r in reqs]
    assert {"wannier8", "wannier16", "wannier32"} <= set(kinds) and kinds.count("haldane") >= 4
    args = dict(max_active=6, slice_steps=256)
    print("serving: 9 requests together", flush=True)
    mix = _served(args, reqs)
    assert mix["completed"] == len(reqs) and mix["failed"] == 0, mix["records"]
    print({k: mix[k] for k in ("wall_s", "p50_s", "p99_s", "k_steps_per_s", "max_active")}, flush=True)
    for i, r in enumerate(reqs):
        print(f"serving: request {i} ({r.kind}, {r.nk} k) alone", flush=True)
        alone = _served(args, [r])["records"][0]["series"]
        for k in alone:
            assert np.array_equal(mix["records"][i]["series"][k], alone[k]), (i, r.kind, k)
    small = min((i for i in range(len(reqs)) if reqs[i].kind == "haldane"), key=lambda i: reqs[i].nk)
    for i in (small, kinds.index("wannier8")):
        print(f"serving: request {i} ({reqs[i].kind}) against the oracle", flush=True)
        ref = pyoracle.run(reqs[i].config, reqs[i].model, nthreads=16)
        got = mix["records"][i]["series"]
        jt = got["j_inter"] + got["j_intra"]
        assert rel_l2(jt, ref["j_inter"] + ref["j_intra"]) < 1e-10, i
        assert rel_l2(got["band_pop"], ref["band_pop"]) < 1e-10, i


def test_nonfinite_request_fails_alone():
    """One request whose field overflows (E0 = 1e308: A = inf, so its phases and state
    become NaN in the first step) fails with SBE_ERR_NAN after its first slice; the
    requests sharing its launches finish with exactly the series they have without it."""
    import dataclasses
    reqs = serving_stream(4, rate=1000.0, seed=3, wannier_share=0.0, grid_range=(24, 40), cycles=(2, 3))
    bad_cfg = dataclasses.replace(reqs[1].config, pulses=[dataclasses.replace(reqs[1].config.pulses[0], E0=1e308)])
    bad = type(reqs[1])(reqs[1].arrival, bad_cfg, reqs[1].model, "haldane-nonfinite", reqs[1].nk)
    args = dict(max_active=8, slice_steps=64)
    print("serving: 4 requests and one with a non-finite field", flush=True)
    with SBEServer(**args) as srv:
        ids = [srv.submit(r.config, r.model) for r in reqs]
        bid = srv.submit(bad.config, bad.model)
        states = {}
        for _ in range(100000):
            srv.step()
            states = {i: srv.poll(i) for i in ids + [bid]}
            if all(st.state in (A.SBE_REQ_DONE, A.SBE_REQ_FAILED) for st in states.values()):
                break
        st = states[bid]
        assert st.state == A.SBE_REQ_FAILED and st.status == A.SBE_ERR_NAN, (st.state, st.status)
        msg = srv.error(bid)
        assert "non-finite" in msg
        assert int(msg.split("first ")[1].split()[0]) < st.nt  # caught before the end of its pulse
        got = [srv.series(i, 2) for i in ids]
    clean = _served(args, reqs)
    for i in range(len(reqs)):
        assert states[ids[i]].state == A.SBE_REQ_DONE
        for k in got[i]:
            assert np.array_equal(got[i][k], clean["records"][i]["series"][k]), (i, k)


def test_cancel_queued_and_running():
    """A running request cancelled between rounds and a queued one cancelled before
    admission end as SBE_REQ_CANCELLED and release their memory; the others finish with
    their served-alone series."""
    reqs = serving_stream(4, rate=1000.0, seed=5, wannier_share=0.0, grid_range=(60, 80), cycles=(2, 3))
    args = dict(max_active=2, slice_steps=100)
    print("serving: cancel one running and one queued request", flush=True)
    with SBEServer(**args) as srv:
        ids = [srv.submit(r.config, r.model) for r in reqs]
        srv.step()
        assert srv.poll(ids[0]).state == A.SBE_REQ_RUNNING and srv.poll(ids[3]).state == A.SBE_REQ_QUEUED
        srv.cancel(ids[0])
        srv.cancel(ids[3])
        assert srv.poll(ids[3]).state == A.SBE_REQ_CANCELLED
        for _ in range(100000):
            srv.step()
            if all(srv.poll(i).state in (A.SBE_REQ_DONE, A.SBE_REQ_CANCELLED) for i in ids):
                break
        assert [srv.poll(i).state for i in ids] == [A.SBE_REQ_CANCELLED, A.SBE_REQ_DONE, A.SBE_REQ_DONE,
                                                   A.SBE_REQ_CANCELLED]
        q, a, used, budget = srv.counts()
        assert q == 0 and a == 0
        with pytest.raises(SBEError) as e:
            srv.cancel(ids[1])
        assert e.value.code == A.SBE_ERR_STATE
        got = {i: srv.series(ids[i], 2) for i in (1, 2)}
    for i in (1, 2):
        alone = _served(args, [reqs[i]])["records"][0]["series"]
        for k in alone:
            assert np.array_equal(got[i][k], alone[k]), (i, k)
