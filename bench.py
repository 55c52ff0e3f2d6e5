"""Throughput benchmark of the SBE time propagation on one MI355X.

Default workload (--workload batch): the batched configuration, 64 seeded
Haldane solves on 200 x 200 k grids (2.56 M k points), per solve E0
log-uniform in [0.001, 0.006], ellipticity U[0, 1], CEP U[0, 2 pi), phi0
U[0, pi/2], M/t2 U[0, 4], 6..16-cycle sin^2 pulses, T2 = 220, samples every
dt = 0.25. One "step" is one pass over the batch: every solve from its
initial state through its whole pulse (10771..28723 samples), including the
per-sample k sums. Four forms run the same batch: plain FP64 RK4 (the
reference's method), RK4 with FP32 state, and the adaptive Cash-Karp stepper
in FP64 and with FP32 state; each form's j(t) error against the plain FP64
pass is measured per solve, and `value` is the fastest form whose worst solve
is within 1e-6. Throughput is in oracle-equivalent k-point steps (nk x Nt at
dt = 0.25 per solve).

Other workloads: --workload haldane (300 x 300 Haldane, elliptic 12-cycle
pulse, FP64 and FP32 state; a step is one RK4 step of all 90,000 k points),
wannier16 (N = 16 Wannier model on 48 x 48 x 8), velocity (the velocity
gauge), serving (a seeded request stream) and wannier-requests (the stream's
Wannier requests solved one at a time, fused and separate kernels).

Prints ONE JSON line. Synthetic, seeded inputs only.

    python bench.py [--steps K] [--warmup W] [--workload ...] [--no-cpu-baseline]
"""
from __future__ import annotations

import argparse
import dataclasses
import json
import math
import os
import sys
import time
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT))

from antelope_amd import HaldaneModel, Pulse, RunConfig, SBESolver  # noqa: E402
from antelope_amd.solver import autotune, batch_autotune  # noqa: E402
from antelope_amd.spectrum import harmonic_error_db, hhg_spectrum  # noqa: E402

METRIC = "k-point·RK-steps/s at ≤1e-6 rel err vs FP64 oracle; p50/p99 solve latency"
PEAK_HBM_GBS = 8000.0
# Algorithmic bytes per k-point RK step (the compulsory state traffic of a
# kernel that streams rho through HBM every step: read + write of the 2x2
# complex density matrix): 128 B in FP64, 64 B with FP32 state.
ALGO_BYTES_PER_KSTEP = {"fp64": 128.0, "fp32": 64.0, "adaptive_fp64": 128.0}
# Executed flops per k-point RK4 step of the step kernel, from the rocprofv3
# counters SQ_INSTS_VALU_{FMA,MUL,ADD}_{F64,F32} (FMA = 2) on this workload:
# FP64 385; FP32 state 365 FP32 + 26 FP64 (the k sums).
FLOPS_PER_KSTEP = {"fp64": 385.0, "fp32": 391.0}
KERNEL_NAME = {"fp64": "haldane_steps_kernel<double, 4, false, true, 2>",
               "fp32": "haldane_steps_kernel<float, 4, false, true, 1>",
               "adaptive_fp64": "haldane_adaptive_kernel<double, true, 1>"}
PEAK_TFLOPS = {"fp64": 78.6, "fp32": 157.3}
# Flops per k-point RK4 step of the velocity-gauge step kernel (hand count:
# sample 36, four stages of input 6 + h x y 12 + dephasing 4, h 10 at the three
# distinct stage times, update 27)
VEL_FLOPS_PER_KSTEP = {"fp64": 181.0, "fp32": 181.0}
# Executed FP64 flops per accepted k-point step of the adaptive Cash-Karp FP64
# kernel (one k point per lane), from SQ_INSTS_VALU_{FMA,MUL,ADD}_F64 (FMA = 2)
# on the 64 x 200^2 batch (profiles/haldane_batch_64x200x200_traffic.json);
# includes rejected attempts and the per-wave laser, controller, reduction and
# sample work counted on all 64 lanes
ADAPT_FLOPS_PER_KSTEP = {"fp64": 929.8}


def wannier_workload(precision: str = "fp64"):
    """Synthetic Wannier N = 16 model (147 R), 48 x 48 x 8 grid, linear
    sin^2 pulse E0 = 0.003, w0 = 0.014, 8 cycles, dt = 0.1 (Nt = 35903),
    RK4, T2 = 220, inter/intra split on."""
    from antelope_amd.models import synthetic_wannier
    model = synthetic_wannier(16, nmax=(3, 3, 1), seed=2024)
    pulse = Pulse(E0=0.003, w0=0.014, ncycles=8.0, envelope="sin2")
    cfg = RunConfig(npoints=[48, 48, 8], pulses=[pulse], dt=0.1, T2=220.0, rk_order=4,
                    inter_intra=True, precision=precision)
    return cfg, model


def workload(grid: int, precision: str) -> tuple[RunConfig, HaldaneModel]:
    t1 = 0.075
    model = HaldaneModel(a0=2.683, t1=t1, t2=t1 / 3, phi0=0.3 * math.pi, M0=2.54 * t1 / 3)
    pulse = Pulse(E0=0.0045, w0=0.014, ncycles=12.0, ellip=0.5, envelope="sin2")
    cfg = RunConfig(npoints=[grid, grid], pulses=[pulse], dt=0.25, T2=220.0, rk_order=4,
                    inter_intra=True, precision=precision)
    return cfg, model


def gpu_sync():
    try:
        import torch
        if torch.cuda.is_available():
            torch.cuda.synchronize()
    except Exception:
        pass


def timed_solve(cfg, model, steps: int, warmup: int):
    """Warm up in a throwaway context, then time `steps` steps of a fresh one."""
    with SBESolver(cfg, model) as w:
        w.run(0, min(warmup, w.nt))
        w.synchronize()
    s = SBESolver(cfg, model)
    steps = min(steps, s.nt)
    s.set_timing(True)
    gpu_sync()
    s.synchronize()
    t0 = time.perf_counter()
    s.run(0, steps)
    s.synchronize()
    gpu_sync()
    wall = time.perf_counter() - t0
    tim = s.timing()
    series = s.series()
    nk = s.info.k_end - s.info.k_begin
    info = s.info
    stats = s.step_stats() if cfg.rk_order == 45 else None
    s.close()
    return {"wall_s": wall, "steps": steps, "nk": nk, "timing": tim, "series": series, "info": info,
            "stats": stats}


def cpu_baseline(cfg, model, budget_s: float,
                 method: str = "per-stage Jacobi eigensolve, generic N^3 products"):
    """FP64 CPU oracle (faithful-cost restatement of the reference per-k loop)
    on a bounded sample: a contiguous k block of the same workload for the
    first steps of the pulse, all host threads (capped at 16)."""
    from oracle import pyoracle
    threads = min(os.cpu_count() or 1, 16)
    nsteps = 200
    ntot = cfg.npoints[0] * cfg.npoints[1]
    nk = min(ntot, 64 * threads)
    while True:
        r = pyoracle.run(cfg, model, 0, nk, nsteps, threads)
        sec = r["info"].seconds
        if sec >= 0.5 * budget_s:
            break
        grow = budget_s / max(sec, 1e-3)
        if nk < ntot:
            nk = int(min(ntot, max(2 * nk, nk * grow)))
        else:
            nsteps = int(nsteps * grow)
    return {
        "value": nk * nsteps / sec,
        "unit": "k-point·RK-steps/s",
        "cores": threads,
        "kind": "port",
        "sample": f"first {nsteps} steps of k points [0,{nk}) of the same workload, FP64 oracle "
                  f"({method}), {sec:.1f} s wall",
    }


def _box_passes(model, npoints, nk):
    """Complex multiply-adds per column of one separable Fourier evaluation, or 0
    where the library takes the dense GEMM (ANTELOPE_NB_DENSE=1, box sides over
    32, a box under a quarter full, or no 2x saving)."""
    if os.environ.get("ANTELOPE_NB_DENSE") == "1":
        return 0
    rv = np.asarray(model.rvec)
    M = [int(rv[:, i].max() - rv[:, i].min() + 1) for i in range(3)]
    K = [max(int(x), 1) for x in npoints]
    if max(M) > 32 or M[0] * M[1] * M[2] > 4 * len(rv):
        return 0
    cost = M[0] * M[1] * K[2] * M[2] + M[0] * K[1] * K[2] * M[1] + K[0] * K[1] * K[2] * M[0]
    return cost if 2 * cost <= nk * len(rv) else 0


def bench_wannier(args) -> int:
    """N-band path on the Wannier workload: times `steps` steps (default 200) of
    the FP64 form (Fourier sums as separable FP64 passes on this product grid) and
    of the SPLIT16 form (Fourier sums on FP16 MFMA with hi/lo operands, FP32
    accumulation), and reports the
    SPLIT16 series' deviation from the FP64 one over those steps."""
    steps = args.steps if args.steps > 0 else 200
    forms, res = {}, {}
    for prec in ("fp64", "split16"):
        cfg, model = wannier_workload(prec)
        tuning = autotune(cfg, model, 8)
        r = timed_solve(cfg, model, steps, min(args.warmup, 5))
        res[prec] = r
        forms[prec] = {"k_steps_per_s": r["nk"] * r["steps"] / r["wall_s"], "wall_s": r["wall_s"],
                       "ms_per_step": r["wall_s"] * 1e3 / r["steps"], "tuning": tuning}
        print(f"[bench] wannier16 {prec}: {forms[prec]['ms_per_step']:.2f} ms/step", file=sys.stderr, flush=True)
    a, b = res["fp64"]["series"], res["split16"]["series"]
    n = res["fp64"]["steps"]
    forms["split16"]["err_j_rel_l2_vs_fp64"] = rel_l2((b["j_inter"] + b["j_intra"])[:n], (a["j_inter"] + a["j_intra"])[:n])
    forms["split16"]["err_pop_rel_l2_vs_fp64"] = rel_l2(b["band_pop"][:n], a["band_pop"][:n])
    forms["split16"]["err_rho_bar_rel_l2_vs_fp64"] = rel_l2(b["rho_bar"][:n], a["rho_bar"][:n])
    inf = res["fp64"]["info"]
    if n == inf.nt:  # the whole pulse: harmonic intensities too
        o, S64 = hhg_spectrum(a["j_inter"] * inf.dV, a["j_intra"] * inf.dV, inf.dt, 0.014)
        _, S16 = hhg_spectrum(b["j_inter"] * inf.dV, b["j_intra"] * inf.dV, inf.dt, 0.014)
        db, hcut = harmonic_error_db(o, S16, S64)
        forms["split16"]["err_hhg_db_vs_fp64"] = db
        forms["split16"]["hhg_cutoff_harmonic"] = hcut
    r = res["fp64"]
    ks = forms["fp64"]["k_steps_per_s"]
    # flops per k-point RK4 step of the fused step (nb16_step_kernel, one launch per step):
    # 20 complex 16 x 16 products (8 N^3 each; rocprofv3 SQ_INSTS_MFMA = 320 per wave and
    # step, 16 MFMAs a product: 2 for the sample's band-diagonal part, 3 per stage
    # (H~ y and the band-diagonal part of y) past the first's 1, 3 per eigen refinement,
    # 2 for the last Newton-Schulz step), the separable Fourier sums' passes along a3 and
    # a2 for 7 matrices (M1 M2 K3 M3 + M1 K2 K3 M2 complex multiply-adds per column for the
    # whole grid) and their pass along a1 inside the step kernel (8 tiles of M1 terms per
    # k point: J_x, J_y, J_z and H~ at t, H~ and H at t + dt/2 and at t + dt). The
    # reference's algorithm for comparison: 4 stages x (H + 3 D sums + H rho) + the sample's
    # H + 3 J sums and its 8 products.
    model = wannier_workload("fp64")[1]
    N, nr = 16, int(model.rvec.shape[0])
    fsum = 8 * N * N * nr
    nk = r["nk"]
    box = _box_passes(model, [48, 48, 8], nk)
    rv = np.asarray(model.rvec)
    Mb = [int(rv[:, i].max() - rv[:, i].min() + 1) for i in range(3)]
    if box:
        pass23 = Mb[0] * Mb[1] * 8 * Mb[2] + Mb[0] * 48 * 8 * Mb[1]
        flops = 20 * 8 * N ** 3 + 7 * 8 * N * N * pass23 / nk + 8 * 8 * N * N * Mb[0]
    else:
        flops = 7 * fsum + 20 * 8 * N ** 3
    flops_ref = 4 * (4 * fsum + 8 * N ** 3) + 4 * fsum + 8 * 8 * N ** 3
    line = {"metric": METRIC, "value": ks, "unit": "k-point·RK-steps/s", "n_gpus": 1,
            "steps": r["steps"], "warmup": min(args.warmup, 5), "ms_per_step": r["wall_s"] * 1e3 / r["steps"],
            "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f64",
            "data": "synthetic (seeded SplitMix64 Wannier N=16 model, 147 R)",
            "config": {"workload": f"Wannier N=16, 147 R, 48x48x8 k grid, sin^2 8-cycle pulse, RK4 dt=0.1, "
                                   f"T2=220, Nt={r['info'].nt} (first {r['steps']} steps timed)", "nk": r["nk"]},
            "forms": forms,
            "executed_flops_per_kstep": flops, "achieved_tflops": ks * flops / 1e12,
            "reference_algorithm_flops_per_kstep": flops_ref,
            "reference_equivalent_tflops": ks * flops_ref / 1e12,
            "roofline": {"bound": "mfma", "achieved": ks * flops / 1e12, "peak": PEAK_TFLOPS["fp64"],
                         "unit": "TFLOP/s", "frac": ks * flops / 1e12 / PEAK_TFLOPS["fp64"], "traffic": None,
                         "note": "executed FP64 flops of the whole step (MFMA products as counted by "
                                 "rocprofv3, the separable Fourier passes) over the step time; one fused "
                                 "launch per step, HBM traffic about 0.5 GB per step"},
            "fourier_form": "separable box passes" if box else "dense GEMM",
            "timing": r["timing"]}
    print(json.dumps(line), flush=True)
    return 0


def bench_velocity(args) -> int:
    """VelocityHamiltonian gauge (SBEs.h:787-871): the Haldane workload above
    with gauge = Velocity, FP64 and FP32 state over the whole pulse (the FP32
    error measured against the FP64 run), and the Wannier N = 16 workload
    (48 x 48 x 8) in the same gauge for `steps` steps (default 2000). Per-k
    E(k) and pMatrix = U^dag J U are built once; a step has no eigensolve and
    no Fourier sum."""
    import dataclasses
    forms, res = {}, {}
    for p in ("fp64", "fp32"):
        cfg, model = workload(args.grid, p)
        cfg = dataclasses.replace(cfg, gauge="VelocityHamiltonian")
        r = timed_solve(cfg, model, args.steps if args.steps > 0 else 10 ** 9, args.warmup)
        res[p] = (r, cfg, model)
        forms[f"haldane_{p}"] = {"k_steps_per_s": r["nk"] * r["steps"] / r["wall_s"], "wall_s": r["wall_s"],
                                 "steps": r["steps"], "step_kernel_ms": r["timing"]["step_ms"],
                                 "step_launches": r["timing"]["step_launches"],
                                 "reduce_kernel_ms": r["timing"]["reduce_ms"]}
        print(f"[bench] velocity haldane {p}: {r['steps']} steps in {r['wall_s']:.3f} s", file=sys.stderr, flush=True)
    a, b = res["fp64"][0]["series"], res["fp32"][0]["series"]
    n = min(res["fp64"][0]["steps"], res["fp32"][0]["steps"])
    forms["haldane_fp32"]["err_j_rel_l2_vs_fp64"] = rel_l2((b["j_inter"] + b["j_intra"])[:n], (a["j_inter"] + a["j_intra"])[:n])
    forms["haldane_fp32"]["err_pop_rel_l2_vs_fp64"] = rel_l2(b["band_pop"][:n], a["band_pop"][:n])
    wsteps = args.steps if args.steps > 0 else 2000
    wcfg, wmodel = wannier_workload("fp64")
    wcfg = dataclasses.replace(wcfg, gauge="VelocityHamiltonian")
    w = timed_solve(wcfg, wmodel, wsteps, min(args.warmup, 5))
    forms["wannier16_fp64"] = {"k_steps_per_s": w["nk"] * w["steps"] / w["wall_s"], "wall_s": w["wall_s"],
                               "steps": w["steps"], "nk": w["nk"], "ms_per_step": w["wall_s"] * 1e3 / w["steps"]}
    print(f"[bench] velocity wannier16: {forms['wannier16_fp64']['ms_per_step']:.3f} ms/step", file=sys.stderr, flush=True)
    best = "fp64"
    if forms["haldane_fp32"]["err_j_rel_l2_vs_fp64"] <= 1e-6 and \
            forms["haldane_fp32"]["k_steps_per_s"] > forms["haldane_fp64"]["k_steps_per_s"]:
        best = "fp32"
    r, cfg, model = res[best]
    kst = r["nk"] * r["steps"]
    sec = r["timing"]["step_ms"] * 1e-3
    fl = kst * VEL_FLOPS_PER_KSTEP[best] / sec / 1e12
    cpu = None
    if not args.no_cpu_baseline:
        cpu = cpu_baseline(res["fp64"][1], res["fp64"][2], args.cpu_seconds,
                           "velocity gauge: per-k E, U, pMatrix once, generic N^3 products per stage")
    line = {"metric": METRIC, "value": forms[f"haldane_{best}"]["k_steps_per_s"], "unit": "k-point·RK-steps/s",
            "n_gpus": 1, "steps": r["steps"], "warmup": args.warmup, "ms_per_step": r["wall_s"] * 1e3 / r["steps"],
            "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
            "dtype": "f64" if best == "fp64" else "f32 state, f64 sums",
            "data": "synthetic (seeded analytic Haldane model and laser pulse; seeded Wannier N=16 model)",
            "config": {"workload": f"VelocityHamiltonian gauge: Haldane 2-band {args.grid}x{args.grid}, elliptic 0.5 "
                                   f"sin^2 12-cycle pulse, T2=220, RK4 dt=0.25, Nt={r['info'].nt}; Wannier N=16 "
                                   f"48x48x8 dt=0.1 (first {w['steps']} steps)", "nk": r["nk"], "form": best},
            "forms": forms,
            "roofline": {"bound": "valu", "kernel": f"haldane_vel_kernel<{'double' if best == 'fp64' else 'float'}, 4, "
                                                    f"false, true, 1>",
                         "achieved": fl, "peak": PEAK_TFLOPS["fp64"], "unit": "TFLOP/s", "frac": fl / PEAK_TFLOPS["fp64"],
                         "traffic": None, "flops_per_kstep": VEL_FLOPS_PER_KSTEP[best],
                         "avg_launch_ms": r["timing"]["step_ms"] / max(r["timing"]["step_launches"], 1),
                         "note": "state and per-k constants stay in registers for a launch; HBM sees the state "
                                 "once per launch and one row of partial sums per wave and step"},
            "cpu_baseline": cpu}
    print(json.dumps(line), flush=True)
    return 0


def bench_serving(args) -> int:
    """Continuous-batching request loop on a seeded Poisson stream (20 requests/s):
    Haldane solves on 100^2..400^2 grids with 4..20-cycle pulses, 10% synthetic
    Wannier solves (N in {8, 16, 32}) at reduced grid sizes. Reports aggregate
    k-point RK steps/s over the whole stream and p50/p99 submit-to-done latency."""
    # one hardware queue per in-flight request (HIP's default is 4 per process): requests on
    # a shared queue wait behind each other's kernels; must be set before the first HIP call
    os.environ["GPU_MAX_HW_QUEUES"] = str(min(16, args.max_active))
    from antelope_amd.serve import SBEServer, run_stream, serving_stream

    n = args.requests
    reqs = serving_stream(n, rate=20.0, seed=5, wannier_share=0.1)
    # warm the kernels (first-launch code object loading) outside the stream
    warm = serving_stream(2, rate=1e3, seed=11, wannier_share=0.5, grid_range=(16, 16), cycles=(1, 1),
                          wannier_cycles=(1, 1))
    with SBEServer(max_active=args.max_active, slice_steps=args.slice_steps) as srv:
        run_stream(srv, warm, realtime=False)
        out = run_stream(srv, reqs, realtime=True)
    kinds = {}
    for r in out["records"]:
        kinds.setdefault(r["kind"], []).append(r["latency_s"])
    print(f"[bench] serving: {out['completed']} done, {out['failed']} failed in {out['wall_s']:.2f} s, "
          f"max {out['max_active']} active; " +
          ", ".join(f"{k}: {len(v)} (p50 {np.percentile(v, 50):.3f} s)" for k, v in sorted(kinds.items())),
          file=sys.stderr, flush=True)
    line = {"metric": METRIC, "value": out["k_steps_per_s"], "unit": "k-point·RK-steps/s", "n_gpus": 1,
            "steps": n, "warmup": len(warm), "ms_per_step": out["wall_s"] * 1e3 / n,
            "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f64",
            "data": "synthetic (seeded SplitMix64 request stream, models and pulses)",
            "config": {"workload": f"serving: {n} requests, Poisson 20/s, Haldane 100^2..400^2 grids, "
                                   f"4..20 cycles, E0 log-uniform [0.001, 0.006]; 10% Wannier N in {{8,16,32}} "
                                   f"on 8^2..16^2 grids, 2..4 cycles", "max_active": args.max_active,
                       "slice_steps": args.slice_steps, "hw_queues": os.environ.get("GPU_MAX_HW_QUEUES")},
            "latency_p50_s": out["p50_s"], "latency_p99_s": out["p99_s"],
            "latency_by_kind": {k: {"n": len(v), "p50_s": float(np.percentile(v, 50)),
                                    "p99_s": float(np.percentile(v, 99))} for k, v in sorted(kinds.items())},
            "completed": out["completed"], "failed": out["failed"], "max_active_seen": out["max_active"],
            "wall_s": out["wall_s"]}
    print(json.dumps(line), flush=True)
    return 0 if out["failed"] == 0 else 1


# Executed FP64 flops per k-point step of the batch step kernels, from the rocprofv3
# counters SQ_INSTS_VALU_{FMA,MUL,ADD}_F64 (FMA = 2, x 64 lanes) on this workload's
# generator: per RK4 step for the plain kernel, per accepted step for the adaptive one
# (rejected attempts and the per-wave laser, controller and reduction work included).
BATCH_FLOPS = {"plain_fp64_rk4": 392.6, "adaptive_fp64": ADAPT_FLOPS_PER_KSTEP["fp64"]}
BATCH_KERNEL = {"plain_fp64_rk4": "haldane_batch_kernel<double, 4, false, true>",
                "plain_fp32_rk4": "haldane_batch_kernel<float, 4, false, true>",
                "adaptive_fp64": "haldane_adaptive_kernel<double, true, 1>",
                "adaptive_fp32": "haldane_adaptive_kernel<float, true, 1>"}


def _batch_passes(cfgs, models, warmup: int, steps: int):
    """One batch object; `warmup` untimed passes, then `steps` timed passes. A pass
    (one bench step) resets every solve to its initial state and propagates the
    whole batch through its pulses. Returns the timed wall clock, the step-kernel
    HIP-event time of the timed passes and the results of the last pass."""
    from antelope_amd import SBEBatch
    with SBEBatch(cfgs, models) as b:
        for _ in range(warmup):
            b.reset()
            b.run()
            b.synchronize()
        b.timing()  # drop the warm-up events
        gpu_sync()
        b.synchronize()
        t0 = time.perf_counter()
        for _ in range(steps):
            b.reset()
            b.run()
        b.synchronize()
        gpu_sync()
        wall = time.perf_counter() - t0
        tim = b.timing()
        series = [b.series(i) for i in range(len(cfgs))]
        infos = [b.info(i) for i in range(len(cfgs))]
        stats = [b.step_stats(i) for i in range(len(cfgs))]
    return {"wall_s": wall, "timing": tim, "series": series, "info": infos, "stats": stats}


def bench_batch(args) -> int:
    """The batched configuration (the default workload): 64 seeded Haldane solves
    on 200^2 grids with uneven pulses (6..16 cycles). One bench step is one pass
    over the batch: every solve from its initial state through its whole pulse.
    Timed per form on the same batch: the plain FP64 fixed-step RK4 (dt = 0.25,
    the reference's method), the same RK4 with FP32 state (FP64 sums), and the
    adaptive Cash-Karp forms, FP64 (rtol 1e-8, atol --atol) and FP32 state with
    FP64 error control and sums (atol --atol32). Throughput counts oracle-equivalent
    steps, sum_i nk_i Nt_i per pass, for every form. Each form's error is measured
    per solve on its own timed pass against the plain FP64 pass (which meets the
    FP64 CPU oracle to 1e-10 on the parity cases), and `value` is the fastest form
    whose worst solve stays within 1e-6 relative L2 in j(t)."""
    from antelope_amd.batch_workloads import haldane_batch
    n, grid = args.instances, args.grid if args.grid != 300 else 200
    steps = args.steps if args.steps > 0 else 3
    warmup = args.warmup
    forms, res = {}, {}
    plain_c, models = haldane_batch(n=n, grid=grid, rk_order=4, precision="fp64")
    eq = None
    for name, prec, atol in (("plain_fp64_rk4", "fp64", None), ("plain_fp32_rk4", "fp32", None),
                             ("adaptive_fp64", "fp64", args.atol), ("adaptive_fp32", "fp32", args.atol32)):
        if atol is None:
            cfgs, _ = haldane_batch(n=n, grid=grid, rk_order=4, precision=prec)
        else:
            cfgs, _ = haldane_batch(n=n, grid=grid, rk_order=45, precision=prec, rtol=args.rtol, atol=atol)
        tuning = batch_autotune(cfgs, models, 2048)  # plan for these members, cached for the batch below
        r = _batch_passes(cfgs, models, warmup, steps)
        res[name] = r
        if eq is None:  # oracle-equivalent k-point steps of one pass
            eq = sum((inf.k_end - inf.k_begin) * inf.nt for inf in r["info"])
        ks = eq * steps / r["wall_s"]
        f = {"k_steps_per_s": ks, "wall_s": r["wall_s"], "ms_per_pass": r["wall_s"] * 1e3 / steps,
             "step_kernel_ms_per_pass": r["timing"]["step_ms"] / steps,
             "step_launches_per_pass": r["timing"]["launches"] / steps, "kernel": BATCH_KERNEL[name],
             "tuning": tuning}
        if atol is not None:
            acc = sum(st.accepted for st in r["stats"])
            kpb = r["stats"][0].k_per_block
            f.update({"rtol": args.rtol, "atol": atol, "accepted_block_steps": acc,
                      "rejected_block_steps": sum(st.rejected for st in r["stats"]),
                      "blocks": sum(st.blocks for st in r["stats"]), "k_per_block": kpb,
                      "mean_step_over_dt": eq / max(acc * kpb, 1)})
        forms[name] = f
        print(f"[bench] batch {name}: {steps} passes in {r['wall_s']:.3f} s ({f['ms_per_pass']:.1f} ms/pass)",
              file=sys.stderr, flush=True)
    ref = res["plain_fp64_rk4"]
    v_plain = forms["plain_fp64_rk4"]["k_steps_per_s"]
    for name, f in forms.items():
        f["speedup_vs_plain_fp64"] = f["k_steps_per_s"] / v_plain
        if name == "plain_fp64_rk4":
            continue
        r = res[name]
        errs, perrs = [], []
        for i in range(n):
            a, b = ref["series"][i], r["series"][i]
            errs.append(rel_l2(b["j_inter"] + b["j_intra"], a["j_inter"] + a["j_intra"]))
            perrs.append(rel_l2(b["band_pop"], a["band_pop"]))
        worst = int(np.argmax(errs))
        dbs = []
        for i in sorted({0, 1, worst}):
            inf = ref["info"][i]
            a, b = ref["series"][i], r["series"][i]
            o, S64 = hhg_spectrum(a["j_inter"] * inf.dV, a["j_intra"] * inf.dV, inf.dt, 0.014)
            _, Sf = hhg_spectrum(b["j_inter"] * inf.dV, b["j_intra"] * inf.dV, inf.dt, 0.014)
            dbs.append({"solve": i, "db": harmonic_error_db(o, Sf, S64)[0]})
        f["err_j_rel_l2_vs_fp64_plain"] = {"max": max(errs), "median": float(np.median(errs)), "worst_solve": worst,
                                           "worst_E0": plain_c[worst].pulses[0].E0}
        f["err_pop_rel_l2_vs_fp64_plain"] = {"max": max(perrs), "median": float(np.median(perrs))}
        f["err_hhg_db_vs_fp64_plain"] = dbs
        f["within_1e-6"] = max(errs) <= 1e-6
        print(f"[bench] batch {name}: {f['speedup_vs_plain_fp64']:.2f}x plain, worst j error {max(errs):.2e}",
              file=sys.stderr, flush=True)
    ok = ["plain_fp64_rk4"] + [k for k, f in forms.items() if f.get("within_1e-6")]
    best = max(ok, key=lambda k: forms[k]["k_steps_per_s"])
    f = forms[best]
    r = res[best]
    value = f["k_steps_per_s"]
    # step-kernel busy time of the timed passes: the union of the launch intervals (HIP events; a
    # batch runs its solves in two groups whose launches overlap)
    kt = r["timing"]["step_ms"] * 1e-3
    launches = max(r["timing"]["launches"], 1)
    # the bound that limits the step kernels: FP64 VALU issue (the state stays in registers for a
    # window; HBM carries the state once per launch and one partial row per wave and sample)
    if best.startswith("adaptive"):
        ksteps = f["accepted_block_steps"] * f["k_per_block"] * steps
    else:
        ksteps = eq * steps
    fpk = BATCH_FLOPS.get(best)
    ach = ksteps * fpk / kt / 1e12 if fpk else None
    rfl = {"bound": "valu", "kernel": f["kernel"], "achieved": ach, "peak": PEAK_TFLOPS["fp64"], "unit": "TFLOP/s",
           "frac": ach / PEAK_TFLOPS["fp64"] if ach else None, "traffic": None,
           "flops_per_kstep": fpk, "ksteps_per_pass": ksteps / steps,
           "busy_ms_per_pass": r["timing"]["step_ms"] / steps, "launches": launches,
           "note": "executed FP64 flops (rocprofv3 SQ_INSTS_VALU_*_F64 counts, FMA = 2) per "
                   + ("accepted " if best.startswith("adaptive") else "") + "k-point step x the k-point steps of "
                   "the timed passes, over the step kernels' HIP-event busy time (union of the launch "
                   "intervals); peak = FP64 vector rate"}
    abytes = ALGO_BYTES_PER_KSTEP["fp32" if "fp32" in best else "fp64"]
    hbm = eq * steps * abytes / kt / 1e9
    prof = ROOT / "profiles" / "haldane_batch_64x200x200_traffic.json"
    if prof.exists() and n == 64 and grid == 200:
        try:
            t = json.loads(prof.read_text())[best]
            rfl["traffic"] = t["hbm_bytes_per_launch"]
            rfl["traffic_source"] = str(prof.relative_to(ROOT))
            # VALU issue: wave instructions (PMC) x 4 cycles (a wave64 VALU op on a 16-lane
            # SIMD) over 1024 SIMDs x 2.4 GHz x the step-kernel time
            wave_steps = ksteps / 64.0 if best.startswith("plain") else \
                (f["accepted_block_steps"] + f["rejected_block_steps"]) * steps
            insts = t["valu_insts_per_wave_step"] * wave_steps
            rfl["valu_issue_frac"] = insts * 4.0 / (1024 * 2.4e9 * kt)
        except Exception:
            pass
    cpu = None
    if not args.no_cpu_baseline:
        cpu = cpu_baseline(plain_c[0], models[0], args.cpu_seconds)
        cpu["sample"] = "solve 0 of the batch (200x200, RK4 dt=0.25): " + cpu["sample"]
    line = {"metric": METRIC, "value": value, "unit": "k-point·RK-steps/s",
            "n_gpus": 1, "steps": steps, "warmup": warmup, "ms_per_step": r["wall_s"] * 1e3 / steps,
            "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
            "dtype": {"adaptive_fp32": "f32 state, f64 error control and sums",
                      "plain_fp32_rk4": "f32 state, f64 sums"}.get(best, "f64"),
            "data": "synthetic (seeded SplitMix64 batch of Haldane models and pulses)",
            "config": {"workload": f"batched: {n} Haldane solves on {grid}x{grid} grids, per solve E0 log-U[0.001,0.006], "
                                   f"ellipticity U[0,1], CEP U[0,2pi), phi0 U[0,pi/2], M/t2 U[0,4], 6..16-cycle sin^2 "
                                   f"pulses, T2=220; one step = one pass over the whole batch",
                       "instances": n, "grid": grid, "oracle_equivalent_ksteps_per_pass": eq, "form": best,
                       "dt_oracle": 0.25},
            "count": "k-point steps are oracle-equivalent: sum over solves of nk x Nt at dt = 0.25 per pass, "
                     "for every form (an adaptive step spans several dt)",
            "forms": forms,
            "roofline": rfl,
            "hbm_algorithmic": {"achieved": hbm, "peak": PEAK_HBM_GBS, "unit": "GB/s", "frac": hbm / PEAK_HBM_GBS,
                                "algo_bytes_per_kstep": abytes,
                                "note": "oracle-equivalent k-steps x state bytes (read + write of rho) over the "
                                        "step-kernel time; above 1 since rho stays in registers"},
            "cpu_baseline": cpu}
    print(json.dumps(line), flush=True)
    return 0


def rel_l2(x, ref):
    d = np.linalg.norm(ref)
    return float(np.linalg.norm(x - ref) / d) if d > 0 else float(np.linalg.norm(x))


def bench_wannier_requests(args) -> int:
    """The serving stream's Wannier requests (synthetic N = 8, 16, 32 models with 25 R on
    8^2..16^2 grids, 2..4-cycle pulses, dt = 0.25) solved one at a time over the whole
    pulse: the fused RK4 step kernels beside the separate stage and sample kernels
    (ANTELOPE_NB_FUSED=0). Reports the solve time of each; value is the N = 32 request's
    k-point RK steps/s with the fused kernels."""
    from antelope_amd.serve import serving_stream
    reqs, seen = [], set()
    for r in serving_stream(args.requests, rate=20.0, seed=5, wannier_share=0.1):
        if r.kind.startswith("wannier") and r.kind not in seen:
            seen.add(r.kind)
            reqs.append(r)
    forms = {}
    for fused in ("1", "0"):
        os.environ["ANTELOPE_NB_FUSED"] = fused
        for r in reqs:
            with SBESolver(r.config, r.model) as w:  # first-launch code loading outside the timing
                w.run(0, min(8, w.nt))
                w.synchronize()
            with SBESolver(r.config, r.model) as s:
                gpu_sync()
                t0 = time.perf_counter()
                s.run()
                s.synchronize()
                wall = time.perf_counter() - t0
                nk = s.info.k_end - s.info.k_begin
                key = f"{r.kind}_{'fused' if fused == '1' else 'separate'}"
                forms[key] = {"nk": nk, "steps": s.nt, "solve_s": wall, "us_per_step": wall * 1e6 / s.nt,
                              "k_steps_per_s": nk * s.nt / wall}
                print(f"[bench] {key}: nk={nk} {s.nt} steps in {wall:.3f} s ({wall * 1e6 / s.nt:.1f} us/step)",
                      file=sys.stderr, flush=True)
    os.environ.pop("ANTELOPE_NB_FUSED", None)
    top = forms["wannier32_fused"]
    line = {"metric": METRIC, "value": top["k_steps_per_s"], "unit": "k-point·RK-steps/s", "n_gpus": 1,
            "steps": top["steps"], "warmup": 8, "ms_per_step": top["us_per_step"] * 1e-3,
            "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f64",
            "data": "synthetic (seeded Wannier models and pulses of the serving stream)",
            "config": {"workload": "the serving stream's Wannier requests (N = 8, 16, 32, 25 R), one at a time, "
                                   "whole pulse, RK4 dt = 0.25, T2 = 220", "nk": top["nk"]},
            "forms": forms, "roofline": None, "cpu_baseline": None}
    print(json.dumps(line), flush=True)
    return 0


def main(argv=None) -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=0, help="timed steps (default: the whole pulse)")
    ap.add_argument("--warmup", type=int, default=None,
                    help="untimed steps (default: 1 batch pass; 200 time steps for --workload haldane)")
    ap.add_argument("--grid", type=int, default=300)
    ap.add_argument("--workload", choices=["batch", "haldane", "wannier16", "serving", "velocity", "wannier-requests"],
                    default="batch")
    ap.add_argument("--precision", choices=["both", "fp64", "fp32"], default="both")
    ap.add_argument("--requests", type=int, default=100, help="serving: requests in the stream")
    ap.add_argument("--instances", type=int, default=64, help="batch: solves in the batch")
    ap.add_argument("--rtol", type=float, default=1e-8, help="batch: adaptive tolerance")
    ap.add_argument("--atol", type=float, default=1e-9, help="batch: FP64 adaptive absolute tolerance")
    ap.add_argument("--atol32", type=float, default=1e-9, help="batch: FP32-state adaptive absolute tolerance")
    ap.add_argument("--max-active", type=int, default=16, help="serving: admitted requests at once")
    ap.add_argument("--slice-steps", type=int, default=256, help="serving: steps per scheduling slice")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--cpu-seconds", type=float, default=15.0)
    ap.add_argument("--copy-calibration-mib", type=int, default=0,
                    help="also run one device copy of this size (HBM counter calibration)")
    args = ap.parse_args(argv)
    if args.warmup is None:
        args.warmup = 1 if args.workload == "batch" else 200
    if args.copy_calibration_mib > 0:
        import torch
        n = args.copy_calibration_mib * (1 << 20) // 4
        src = torch.ones(n, dtype=torch.float32, device="cuda")
        dst = torch.empty_like(src)
        dst.copy_(src)
        torch.cuda.synchronize()
        print(f"[bench] copy calibration: {args.copy_calibration_mib} MiB read + written", file=sys.stderr)
        del src, dst

    world = int(os.environ.get("WORLD_SIZE", "1"))
    if args.gpus != 1 or world != 1:
        print("bench.py measures one GPU only: this build has no multi-GPU path "
              "(k blocks would shard across GPUs with one final sum of the k-summed series).", file=sys.stderr)
        return 2

    if args.workload == "wannier16":
        return bench_wannier(args)
    if args.workload == "serving":
        return bench_serving(args)
    if args.workload == "wannier-requests":
        return bench_wannier_requests(args)
    if args.workload == "batch":
        return bench_batch(args)
    if args.workload == "velocity":
        return bench_velocity(args)
    forms = ["fp64", "fp32"] if args.precision == "both" else [args.precision]
    if "fp64" in forms:
        forms.append("adaptive_fp64")  # Cash-Karp, the batch's tolerances; gated on its j error below
    res = {}
    tunings = {}
    for p in forms:
        cfg, model = workload(args.grid, "fp64" if p == "adaptive_fp64" else p)
        if p == "adaptive_fp64":
            cfg = dataclasses.replace(cfg, rk_order=45, rtol=args.rtol, atol=args.atol)
        steps = args.steps if args.steps > 0 else 10 ** 9
        tunings[p] = autotune(cfg, model, 256)  # cached for the timed context below
        r = timed_solve(cfg, model, steps, args.warmup)
        print(f"[bench] {p}: {r['steps']} steps x {r['nk']} k in {r['wall_s']:.3f} s "
              f"(step kernels {r['timing']['step_ms']:.1f} ms / {r['timing']['step_launches']} launches, "
              f"reductions {r['timing']['reduce_ms']:.1f} ms)", file=sys.stderr, flush=True)
        res[p] = r

    out_forms = {}
    for p, r in res.items():
        ks = r["nk"] * r["steps"] / r["wall_s"]
        out_forms[p] = {"k_steps_per_s": ks, "wall_s": r["wall_s"],
                        "step_kernel_ms": r["timing"]["step_ms"], "reduce_kernel_ms": r["timing"]["reduce_ms"],
                        "tuning": tunings[p]}
    for f in ("fp32", "adaptive_fp64"):
        if "fp64" not in res or f not in res:
            continue
        a, b = res["fp64"]["series"], res[f]["series"]
        n = min(res["fp64"]["steps"], res[f]["steps"])
        j64 = (a["j_inter"] + a["j_intra"])[:n]
        jf = (b["j_inter"] + b["j_intra"])[:n]
        out_forms[f]["err_j_rel_l2_vs_fp64"] = rel_l2(jf, j64)
        out_forms[f]["err_pop_rel_l2_vs_fp64"] = rel_l2(b["band_pop"][:n], a["band_pop"][:n])
        if n == res["fp64"]["info"].nt:
            dV, dt = res["fp64"]["info"].dV, res["fp64"]["info"].dt
            o, S64 = hhg_spectrum(a["j_inter"] * dV, a["j_intra"] * dV, dt, 0.014)
            _, Sf = hhg_spectrum(b["j_inter"] * dV, b["j_intra"] * dV, dt, 0.014)
            db, hcut = harmonic_error_db(o, Sf, S64)
            out_forms[f]["err_hhg_db_vs_fp64"] = db
            out_forms[f]["hhg_cutoff_harmonic"] = hcut
        print(f"[bench] {f}: {out_forms[f]['k_steps_per_s'] / out_forms['fp64']['k_steps_per_s']:.2f}x fp64, "
              f"j error {out_forms[f]['err_j_rel_l2_vs_fp64']:.2e}", file=sys.stderr, flush=True)
    if "adaptive_fp64" in res:
        st = res["adaptive_fp64"]["stats"]
        out_forms["adaptive_fp64"].update({"rtol": args.rtol, "atol": args.atol, "accepted_block_steps": st.accepted,
                                           "rejected_block_steps": st.rejected, "k_per_block": st.k_per_block,
                                           "mean_step_over_dt": res["adaptive_fp64"]["steps"] * st.blocks
                                           / max(st.accepted, 1)})

    # the reported value: the fastest form within the 1e-6 error bar
    best = "fp64"
    for f in ("fp32", "adaptive_fp64"):
        if f in out_forms and out_forms[f].get("err_j_rel_l2_vs_fp64", 1.0) <= 1e-6 and \
                (best not in out_forms or out_forms[f]["k_steps_per_s"] > out_forms[best]["k_steps_per_s"]):
            best = f
    if best not in out_forms:
        best = forms[0]
    r = res[best]
    value = out_forms[best]["k_steps_per_s"]
    # roofline of the dominant kernel (the fused step kernel of the reported
    # form), timed with HIP events on the library's stream
    tim = r["timing"]
    kst = r["nk"] * r["steps"]  # oracle-equivalent k-point steps (every form)
    sec = tim["step_ms"] * 1e-3
    kname = KERNEL_NAME[best]
    if best == "adaptive_fp64":  # the launch plan's k points per lane
        kname = f"haldane_adaptive_kernel<double, true, {max(1, tunings[best].get('kpt', 1))}>"
    ach = kst * ALGO_BYTES_PER_KSTEP[best] / sec / 1e9
    hbm_rfl = {"bound": "hbm", "kernel": kname, "achieved": ach, "peak": PEAK_HBM_GBS,
               "unit": "GB/s", "frac": ach / PEAK_HBM_GBS, "traffic": None,
               "algo_bytes_per_kstep": ALGO_BYTES_PER_KSTEP[best],
               "note": "time-fused kernel: rho stays in registers for a whole launch, so the measured HBM "
                       "traffic is far below the algorithmic bytes and HBM is not what limits it"}
    # the bound that limits it: FP64 VALU issue. Executed FP64 flops per k-point step of the
    # fixed-step kernels, per accepted k-point step of the adaptive one (counted with one k
    # point per lane on the 64 x 200^2 batch)
    if best == "adaptive_fp64":
        fpk = ADAPT_FLOPS_PER_KSTEP["fp64"]
        kfl = out_forms[best]["accepted_block_steps"] * out_forms[best]["k_per_block"]
    else:
        fpk = FLOPS_PER_KSTEP[best]
        kfl = kst
    fl = kfl * fpk / sec / 1e12
    rfl = {"bound": "valu", "kernel": kname, "achieved": fl, "peak": PEAK_TFLOPS["fp64"], "unit": "TFLOP/s",
           "frac": fl / PEAK_TFLOPS["fp64"], "flops_per_kstep": fpk, "traffic": None,
           "avg_launch_ms": tim["step_ms"] / max(tim["step_launches"], 1),
           "launches": tim["step_launches"],
           "note": "executed FP64 flops (rocprofv3 SQ_INSTS_VALU_*_F64, FMA = 2) per "
                   + ("accepted " if best == "adaptive_fp64" else "") + "k-point step over the step kernels' "
                   "HIP-event time; peak = FP64 vector rate (packed FP32 is not used, so FP32 issues at it too)"
                   + ("; flops counted with one k point per lane" if best == "adaptive_fp64" else "")}
    prof = ROOT / "profiles" / "haldane_steps_300x300_traffic.json"
    if prof.exists() and args.grid == 300:
        try:
            t = json.loads(prof.read_text())[best]
            rfl["traffic"] = hbm_rfl["traffic"] = t["hbm_bytes_per_launch"]
            rfl["traffic_source"] = str(prof.relative_to(ROOT))
        except Exception:
            pass

    # every form's step kernel against the same algorithmic bytes (the plain FP64 path is
    # the reference's method; `roofline` above is the reported form's)
    for f, rr in res.items():
        if f in out_forms and rr["timing"]["step_ms"] > 0:
            k_s = rr["nk"] * rr["steps"] / (rr["timing"]["step_ms"] * 1e-3)
            out_forms[f]["kernel"] = KERNEL_NAME[f] if f != "adaptive_fp64" else \
                f"haldane_adaptive_kernel<double, true, {max(1, tunings[f].get('kpt', 1))}>"
            out_forms[f]["hbm_roofline_frac_algorithmic"] = k_s * ALGO_BYTES_PER_KSTEP[f] / 1e9 / PEAK_HBM_GBS
            if f in FLOPS_PER_KSTEP:
                out_forms[f]["valu_frac_fp64_peak"] = k_s * FLOPS_PER_KSTEP[f] / 1e12 / PEAK_TFLOPS["fp64"]

    cpu = None
    if not args.no_cpu_baseline:
        cfg, model = workload(args.grid, "fp64")
        cpu = cpu_baseline(cfg, model, args.cpu_seconds)

    line = {
        "metric": METRIC,
        "value": value,
        "unit": "k-point·RK-steps/s",
        "n_gpus": 1,
        "steps": r["steps"],
        "warmup": args.warmup,
        "ms_per_step": r["wall_s"] * 1e3 / r["steps"],
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32 state, f64 sums" if best == "fp32" else "f64",
        "data": "synthetic (seeded analytic Haldane model and laser pulse)",
        "config": {"workload": f"Haldane 2-band, {args.grid}x{args.grid} k grid, elliptic 0.5 sin^2 12-cycle "
                               f"pulse, T2=220, RK4 dt=0.25, Nt={res[best]['info'].nt}",
                   "nk": r["nk"], "form": best},
        "count": "k-point steps are oracle-equivalent: nk x samples at dt = 0.25 for every form "
                 "(an adaptive step spans several dt)",
        "forms": out_forms,
        "roofline": rfl,
        "hbm_roofline": hbm_rfl,
        "cpu_baseline": cpu,
    }
    print(json.dumps(line), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
