"""Serving: a Poisson stream of SBE solves through the continuous-batching
request loop of include/antelope/sbe_server.h.

`serving_stream` draws the request mix of the serving configuration from a
seed: arrivals at `rate` requests/s; 2-band Haldane solves on square grids
of 100^2..400^2 points with 4..20-cycle pulses and E0 log-uniform in
[0.001, 0.006]; and a share of synthetic Wannier solves (N = 8, 16 or 32). The
Wannier solves are reduced in size (k grids of 8^2..16^2, 2..4 cycles, a
5 x 5 x 1 R range) because at N = 32 and 400^2 points a single request would
run for hours on one GPU; their band counts cycle through 8, 16, 32, so every
count occurs once three Wannier requests are drawn.

`run_stream` replays the arrivals in real time: requests are submitted when
their arrival time passes, the scheduler runs rounds in between, and the
latency of a request is done - submit on the host monotonic clock.
"""
from __future__ import annotations

import ctypes as C
import math
import time
from dataclasses import dataclass, field

import numpy as np

from . import _abi as A
from .models import SplitMix64, synthetic_wannier
from .params import HaldaneModel, Pulse, RunConfig
from .solver import SBEError, _ptr, load_library


class SBEServer:
    def __init__(self, hbm_budget_bytes: int = 0, max_active: int = 16, slice_steps: int = 512):
        self._lib = load_library()
        h = C.c_void_p()
        rc = self._lib.sbe_server_create(hbm_budget_bytes, max_active, slice_steps, C.byref(h))
        if rc != A.SBE_OK:
            raise SBEError(rc, "server creation failed")
        self._h = h

    def _chk(self, rc):
        if rc != A.SBE_OK:
            raise SBEError(rc, self._lib.sbe_server_last_error(self._h).decode())

    def submit(self, config, model) -> int:
        cfg, k1 = config.to_c()
        mdl, k2 = model.to_c()
        rid = C.c_int64()
        self._chk(self._lib.sbe_server_submit(self._h, C.byref(cfg), C.byref(mdl), C.byref(rid)))
        return rid.value

    def step(self) -> None:
        self._chk(self._lib.sbe_server_step(self._h))

    def poll(self, rid: int) -> A.sbe_request_status:
        st = A.sbe_request_status()
        self._chk(self._lib.sbe_server_poll(self._h, rid, C.byref(st)))
        return st

    def counts(self):
        q, a = C.c_int32(), C.c_int32()
        used, budget = C.c_size_t(), C.c_size_t()
        self._chk(self._lib.sbe_server_counts(self._h, C.byref(q), C.byref(a), C.byref(used), C.byref(budget)))
        return q.value, a.value, used.value, budget.value

    def series(self, rid: int, nband: int) -> dict:
        st = self.poll(rid)
        if st.state != A.SBE_REQ_DONE:
            raise SBEError(A.SBE_ERR_STATE, self._lib.sbe_server_request_error(self._h, rid).decode())
        nt = st.nt
        out = {"j_inter": np.zeros((nt, 3)), "j_intra": np.zeros((nt, 3)),
               "rho_bar": np.zeros((nt, nband, nband, 2)), "band_pop": np.zeros((nt, nband))}
        self._chk(self._lib.sbe_server_get_series(self._h, rid, _ptr(out["j_inter"]), _ptr(out["j_intra"]),
                                                  _ptr(out["rho_bar"]), _ptr(out["band_pop"])))
        rb = out.pop("rho_bar")
        out["rho_bar"] = rb[..., 0] + 1j * rb[..., 1]
        return out

    def release(self, rid: int) -> None:
        self._chk(self._lib.sbe_server_release(self._h, rid))

    def cancel(self, rid: int) -> None:
        """Drop a queued request, or retire a running one at the next round."""
        self._chk(self._lib.sbe_server_cancel(self._h, rid))

    def error(self, rid: int) -> str:
        return self._lib.sbe_server_request_error(self._h, rid).decode()

    def close(self) -> None:
        if getattr(self, "_h", None):
            self._lib.sbe_server_destroy(self._h)
            self._h = None

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


@dataclass
class ServingRequest:
    arrival: float
    config: RunConfig
    model: object
    kind: str
    nk: int = 0
    meta: dict = field(default_factory=dict)


def serving_stream(n: int, rate: float = 20.0, seed: int = 5, wannier_share: float = 0.1,
                   grid_range=(100, 400), cycles=(4, 20), wannier_cycles=(2, 4), precision: str = "fp64"):
    rng = SplitMix64(seed)
    t = 0.0
    reqs = []
    t1 = 0.075
    nw = 0
    for i in range(n):
        t += -math.log(rng.uniform()) / rate
        E0 = math.exp(math.log(0.001) + rng.uniform() * (math.log(0.006) - math.log(0.001)))
        if rng.uniform() < wannier_share:
            N = (8, 16, 32)[nw % 3]
            nw += 1
            rng.uniform()  # keeps the draws of the stream as they were
            g = 8 + int(rng.uniform() * 9)
            ncyc = wannier_cycles[0] + int(rng.uniform() * (wannier_cycles[1] - wannier_cycles[0] + 1))
            model = synthetic_wannier(N, nmax=(2, 2, 0), seed=1000 + i)
            cfg = RunConfig(npoints=[g, g], pulses=[Pulse(E0=E0, w0=0.014, ncycles=float(ncyc), envelope="sin2")],
                            dt=0.25, T2=220.0, rk_order=4, inter_intra=True)
            reqs.append(ServingRequest(t, cfg, model, f"wannier{N}", g * g))
        else:
            g = grid_range[0] + int(rng.uniform() * (grid_range[1] - grid_range[0] + 1))
            ncyc = cycles[0] + int(rng.uniform() * (cycles[1] - cycles[0] + 1))
            model = HaldaneModel(a0=2.683, t1=t1, t2=t1 / 3, phi0=0.5 * math.pi * rng.uniform(),
                                 M0=4.0 * rng.uniform() * t1 / 3)
            cfg = RunConfig(npoints=[g, g], pulses=[Pulse(E0=E0, w0=0.014, ncycles=float(ncyc), ellip=rng.uniform(),
                                                          envelope="sin2")],
                            dt=0.25, T2=220.0, rk_order=4, inter_intra=True, precision=precision)
            reqs.append(ServingRequest(t, cfg, model, "haldane", g * g))
    return reqs


def run_stream(server: SBEServer, reqs, realtime: bool = True, keep_results: bool = False, poll_s: float = 0.0005):
    """Replays the arrivals; returns per-request records and aggregate numbers."""
    t0 = time.monotonic()
    ids = {}
    pending = list(range(len(reqs)))
    done = {}
    nxt = 0
    max_active = 0
    while len(done) < len(reqs):
        now = time.monotonic() - t0
        while nxt < len(reqs) and (not realtime or reqs[nxt].arrival <= now):
            ids[nxt] = server.submit(reqs[nxt].config, reqs[nxt].model)
            nxt += 1
        server.step()
        q, a, used, budget = server.counts()
        max_active = max(max_active, a)
        for i in list(pending):
            if i not in ids:
                continue
            st = server.poll(ids[i])
            if st.state in (A.SBE_REQ_DONE, A.SBE_REQ_FAILED, A.SBE_REQ_CANCELLED):
                rec = {"kind": reqs[i].kind, "nk": reqs[i].nk, "nt": st.nt, "state": st.state,
                       "latency_s": st.t_done - st.t_submit, "queue_s": (st.t_admit - st.t_submit) if st.t_admit else None,
                       "footprint": st.footprint}
                if keep_results and st.state == A.SBE_REQ_DONE:
                    rec["series"] = server.series(ids[i], reqs[i].model.nband)
                done[i] = rec
                server.release(ids[i])
                pending.remove(i)
        if poll_s:
            time.sleep(poll_s)
    wall = time.monotonic() - t0
    lat = np.array([done[i]["latency_s"] for i in range(len(reqs)) if done[i]["state"] == A.SBE_REQ_DONE])
    ksteps = sum(done[i]["nk"] * done[i]["nt"] for i in range(len(reqs)) if done[i]["state"] == A.SBE_REQ_DONE)
    return {
        "records": [done[i] for i in range(len(reqs))],
        "wall_s": wall,
        "p50_s": float(np.percentile(lat, 50)) if lat.size else None,
        "p99_s": float(np.percentile(lat, 99)) if lat.size else None,
        "k_steps_per_s": ksteps / wall,
        "completed": int(lat.size),
        "failed": int(sum(1 for r in done.values() if r["state"] != A.SBE_REQ_DONE)),
        "max_active": max_active,
    }
