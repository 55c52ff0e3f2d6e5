"""The fused RK4 step of the N-band LengthWannier path (nb16_step_kernel: the
sample, the four stages and the last pass of the separable Fourier sums in one
launch per step) against the FP64 oracle and against the separate stage and
sample kernels (ANTELOPE_NB_FUSED=0). Bar against the oracle: 1e-10 relative
L2 (j(t), band populations, k-summed rho, final state); against the separate
kernels: 1e-12 (the same arithmetic, summed in another order).
"""
import numpy as np
import pytest

from antelope_amd import Pulse, RunConfig, SBESolver, query_footprint
from antelope_amd.models import synthetic_wannier
from tests.cases import rel_l2

pytestmark = pytest.mark.gpu
TOL64 = 1e-10


def _cfg(npts, ncycles=0.5, T2=80.0, inter_intra=True, dt=0.5, E0=0.01, rk=4):
    return RunConfig(npoints=list(npts), pulses=[Pulse(E0=E0, w0=0.05, ncycles=ncycles, envelope="sin2")],
                     dt=dt, T2=T2, rk_order=rk, inter_intra=inter_intra)


def _errs(got, ref):
    e = {"j": rel_l2(got["j_inter"] + got["j_intra"], ref["j_inter"] + ref["j_intra"]),
         "j_inter": rel_l2(got["j_inter"], ref["j_inter"]),
         "band_pop": rel_l2(got["band_pop"], ref["band_pop"]),
         "rho_bar": rel_l2(got["rho_bar"], ref["rho_bar"])}
    if np.linalg.norm(ref["j_intra"]) > 0:
        e["j_intra"] = rel_l2(got["j_intra"], ref["j_intra"])
    return e


def _solve(cfg, model, monkeypatch, fused, splits=None, peek=None):
    monkeypatch.setenv("ANTELOPE_NB_FUSED", "1" if fused else "0")
    with SBESolver(cfg, model) as s:
        assert s.info.device_bytes == query_footprint(cfg, model)
        if splits is None:
            s.run()
        else:
            b = 0
            for e in list(splits) + [s.nt]:
                s.run(b, e)
                if peek is not None and b == 0:
                    peek(s)
                b = e
        s.synchronize()
        return s.series(), s.state(), s.info.device_bytes


@pytest.mark.parametrize("label,N,npts,kw", [
    ("n8_box5x5x3", 8, (6, 5, 2), dict()),
    ("n8_8x16x16_eight_waves_per_block", 8, (8, 16, 16), dict(ncycles=0.25)),
    ("n16_ragged_a1_no_dephasing_total_current", 16, (12, 3, 2), dict(T2=-1.0, inter_intra=False)),
    ("n3_one_wave_per_block_used", 3, (5, 4, 3), dict()),
])
def test_fused_step_against_oracle_and_separate_kernels(oracle, monkeypatch, label, N, npts, kw):
    """Blocks of 4 waves on grids that give fewer blocks of 8 than CUs (the small cases),
    of 8 on the 8 x 16 x 16 grid (256 blocks)."""
    model = synthetic_wannier(N, nmax=(2, 2, 1) if N <= 8 else (1, 1, 1), seed=40 + N)
    cfg = _cfg(npts, **kw)
    print(f"case: {label}", flush=True)
    ref = oracle.run(cfg, model, nthreads=8, want_state=True)
    fz, st_f, bytes_f = _solve(cfg, model, monkeypatch, fused=True)
    sp, st_s, bytes_s = _solve(cfg, model, monkeypatch, fused=False)
    assert bytes_f > bytes_s  # the fused plan's buffers were allocated: it ran fused
    e_or = _errs(fz, ref)
    e_sp = _errs(fz, sp)
    print(label, "vs oracle", {k: f"{v:.2e}" for k, v in e_or.items()},
          "vs separate kernels", {k: f"{v:.2e}" for k, v in e_sp.items()}, flush=True)
    assert all(v < TOL64 for v in e_or.values()), e_or
    assert all(v < 1e-12 for v in e_sp.values()), e_sp
    assert rel_l2(st_f, ref["state"]) < TOL64
    assert rel_l2(st_f, st_s) < 1e-12


def test_fused_split_runs_and_state_reads_are_bitwise(monkeypatch):
    """Runs split anywhere, with a band-basis state read in between, give the bits of one run;
    so does any number of steps whose separable passes run in one batch of launches (1 and
    7: the ring of end pairs wraps)."""
    model = synthetic_wannier(8, nmax=(2, 2, 1), seed=9)
    cfg = _cfg((6, 5, 2), ncycles=0.5)
    print("case: split runs", flush=True)
    one, st1, _ = _solve(cfg, model, monkeypatch, fused=True)
    seen = []
    two, st2, _ = _solve(cfg, model, monkeypatch, fused=True, splits=(37, 38, 90),
                         peek=lambda s: seen.append(s.state("band")))
    assert seen and np.all(np.isfinite(seen[0]))
    for k in one:
        assert np.array_equal(one[k], two[k]), k
    assert np.array_equal(st1, st2)
    for S in ("1", "7"):
        monkeypatch.setenv("ANTELOPE_NB_EVAL_STEPS", S)
        three, st3, _ = _solve(cfg, model, monkeypatch, fused=True, splits=(37,))
        for k in one:
            assert np.array_equal(one[k], three[k]), (S, k)
        assert np.array_equal(st1, st3)


def test_fused_rk5_keeps_separate_kernels(oracle, monkeypatch):
    """The fused step is the classical RK4 tableau; RK5 plans the separate kernels."""
    model = synthetic_wannier(6, nmax=(2, 2, 1), seed=12)
    cfg = _cfg((6, 5, 2), rk=5, ncycles=0.25)
    print("case: rk5", flush=True)
    a, _, bytes_a = _solve(cfg, model, monkeypatch, fused=True)
    b, _, bytes_b = _solve(cfg, model, monkeypatch, fused=False)
    assert bytes_a == bytes_b
    ref = oracle.run(cfg, model, nthreads=8)
    assert all(v < TOL64 for v in _errs(a, ref).values())


@pytest.mark.parametrize("label,N,kw", [
    ("n8_velocity_t2", 8, dict()),
    ("n16_velocity_t1t2_total_current", 16, dict(T2=60.0, inter_intra=False)),
    ("n24_velocity_t2_four_waves_per_k", 24, dict()),
    ("n32_velocity_t1t2_total_current", 32, dict(T1=150.0, T2=60.0, inter_intra=False)),
    ("n8_velocity_2048k_eight_waves_per_block", 8, dict(npts=(16, 16, 8), ncycles=0.25)),
])
def test_time_fused_velocity_gauge(oracle, monkeypatch, label, N, kw):
    """The time-fused velocity-gauge kernels (all steps of a chunk in one launch; one wave
    per k point for N <= 16, a block of four waves per k point above) against the oracle
    and the one-launch-per-step kernel, and split runs."""
    import dataclasses
    model = synthetic_wannier(N, nmax=(1, 1, 1), seed=70 + N)
    T1 = kw.pop("T1", 150.0 if N == 16 else -1.0)
    npts = kw.pop("npts", (5, 4, 2))  # blocks of 4 waves for N <= 16 below 2048 k points, of 8 above
    cfg = dataclasses.replace(_cfg(npts, **kw), gauge="VelocityHamiltonian", T1=T1)
    print(f"case: {label}", flush=True)
    ref = oracle.run(cfg, model, nthreads=8, want_state=False)
    fz, _, bytes_f = _solve(cfg, model, monkeypatch, fused=True)
    sp, _, bytes_s = _solve(cfg, model, monkeypatch, fused=False)
    assert bytes_f > bytes_s  # the stage-time table was allocated: it ran time-fused
    e_or, e_sp = _errs(fz, ref), _errs(fz, sp)
    print(label, "vs oracle", {k: f"{v:.2e}" for k, v in e_or.items()},
          "vs one launch per step", {k: f"{v:.2e}" for k, v in e_sp.items()}, flush=True)
    assert all(v < TOL64 for v in e_or.values()), e_or
    assert all(v < 1e-12 for v in e_sp.values()), e_sp
    two, _, _ = _solve(cfg, model, monkeypatch, fused=True, splits=(17, 18, 60))
    for k in fz:
        assert np.array_equal(fz[k], two[k]), k


@pytest.mark.parametrize("label,N,npts,nmax,kw,dense", [
    ("n24_box6x5x2_75R_separable_passes", 24, (6, 5, 2), (2, 2, 1), dict(), False),
    ("n24_27R_one_dense_launch", 24, (6, 5, 2), (1, 1, 1), dict(), False),
    ("n32_grid8x8_no_dephasing_total_current", 32, (8, 8), (2, 2, 0), dict(T2=-1.0, inter_intra=False), False),
    ("n20_dense_sums_t1t2", 20, (5, 4, 3), (1, 1, 1), dict(T1=150.0), True),
    ("n17_300k_more_k_than_blocks", 17, (20, 15), (1, 1, 0), dict(), False),
])
def test_fused_step_n32_against_oracle_and_separate_kernels(oracle, monkeypatch, label, N, npts, nmax, kw, dense):
    """The fused RK4 step for 16 < N <= 32 (nb32_step_kernel: four waves per k point, one
    16 x 16 quadrant of every matrix each, after one batch of Fourier launches per step)
    against the oracle and the separate stage and sample kernels."""
    import dataclasses
    model = synthetic_wannier(N, nmax=nmax, seed=50 + N)
    T1 = kw.pop("T1", -1.0)
    cfg = dataclasses.replace(_cfg(npts, **kw), T1=T1)
    if dense:
        monkeypatch.setenv("ANTELOPE_NB_DENSE", "1")
    print(f"case: {label}", flush=True)
    ref = oracle.run(cfg, model, nthreads=8, want_state=True)
    fz, st_f, bytes_f = _solve(cfg, model, monkeypatch, fused=True)
    sp, st_s, bytes_s = _solve(cfg, model, monkeypatch, fused=False)
    assert bytes_f > bytes_s  # the band energies were allocated: it ran fused
    e_or, e_sp = _errs(fz, ref), _errs(fz, sp)
    print(label, "vs oracle", {k: f"{v:.2e}" for k, v in e_or.items()},
          "vs separate kernels", {k: f"{v:.2e}" for k, v in e_sp.items()}, flush=True)
    assert all(v < TOL64 for v in e_or.values()), e_or
    assert all(v < 1e-12 for v in e_sp.values()), e_sp
    assert rel_l2(st_f, ref["state"]) < TOL64
    assert rel_l2(st_f, st_s) < 1e-12


@pytest.mark.parametrize("label,nmax", [("n24_27R_steps_per_launch", (1, 1, 1)),
                                         ("n24_75R_separable_passes", (2, 2, 1))])
def test_fused_step_n32_split_runs_and_state_reads_are_bitwise(monkeypatch, label, nmax):
    """Runs of the N <= 32 fused step split anywhere, with a band-basis state read in
    between (it reuses matrix slots the step may keep H~ and H in), give the bits of one
    run; with the dense sums of many steps per launch, so does any number of steps per
    launch (1 and 5: the ring of end pairs wraps)."""
    model = synthetic_wannier(24, nmax=nmax, seed=19)
    cfg = _cfg((6, 5, 2), ncycles=0.5)
    print(f"case: {label} split runs", flush=True)
    one, st1, _ = _solve(cfg, model, monkeypatch, fused=True)
    seen = []
    two, st2, _ = _solve(cfg, model, monkeypatch, fused=True, splits=(37, 38, 90),
                         peek=lambda s: seen.append(s.state("band")))
    assert seen and np.all(np.isfinite(seen[0]))
    for k in one:
        assert np.array_equal(one[k], two[k]), k
    assert np.array_equal(st1, st2)
    if nmax == (1, 1, 1):
        for S in ("1", "5"):
            monkeypatch.setenv("ANTELOPE_NB_EVAL_STEPS", S)
            three, st3, _ = _solve(cfg, model, monkeypatch, fused=True, splits=(37,))
            for k in one:
                assert np.array_equal(one[k], three[k]), (S, k)
            assert np.array_equal(st1, st3)


@pytest.mark.parametrize("label,nmax", [("n20_27R_dense_steps", (1, 1, 1)), ("n20_75R_separable_passes", (2, 2, 1))])
def test_fused_step_n32_k_blocks(oracle, monkeypatch, label, nmax):
    """The N <= 32 fused step on k blocks of the grid (one rank's share each, ragged):
    the blocks' series add up to the oracle's whole-grid series."""
    model = synthetic_wannier(20, nmax=nmax, seed=31)
    cfg = _cfg((6, 5, 2), ncycles=0.25)
    monkeypatch.setenv("ANTELOPE_NB_FUSED", "1")
    print(f"case: {label} k blocks", flush=True)
    ref = oracle.run(cfg, model, nthreads=8)
    tot = None
    for b, e in [(0, 23), (23, 41), (41, 60)]:
        with SBESolver(cfg, model, b, e) as s:
            assert s.info.device_bytes == query_footprint(cfg, model, b, e)
            s.run(0, 40)
            s.run(40, s.nt)
            g = s.series()
        tot = g if tot is None else {k: tot[k] + g[k] for k in g}
    e_or = _errs(tot, ref)
    print(label, "k blocks vs oracle", {k: f"{v:.2e}" for k, v in e_or.items()}, flush=True)
    assert all(v < TOL64 for v in e_or.values()), e_or


def test_fused_step_on_k_blocks(oracle, monkeypatch):
    """The N <= 16 fused step on ragged k blocks of the grid (one rank's share each: the
    blocks' boxes have holes, idle waves): each block runs fused (its footprint holds the
    fused buffers) and the blocks' series add up to the oracle's whole-grid series."""
    model = synthetic_wannier(8, nmax=(2, 2, 1), seed=44)
    cfg = _cfg((8, 6, 2), ncycles=0.25)  # blocks of at least 26 k points: the separable sums pay
    print("case: n8 k blocks fused", flush=True)
    ref = oracle.run(cfg, model, nthreads=8)
    tot = None
    for b, e in [(0, 40), (40, 70), (70, 96)]:
        monkeypatch.setenv("ANTELOPE_NB_FUSED", "0")
        sep = query_footprint(cfg, model, b, e)
        monkeypatch.setenv("ANTELOPE_NB_FUSED", "1")
        with SBESolver(cfg, model, b, e) as s:
            assert s.info.device_bytes == query_footprint(cfg, model, b, e) > sep
            s.run(0, 30)
            s.run(30, s.nt)
            g = s.series()
        tot = g if tot is None else {k: tot[k] + g[k] for k in g}
    e_or = _errs(tot, ref)
    print("n8 k blocks fused vs oracle", {k: f"{v:.2e}" for k, v in e_or.items()}, flush=True)
    assert all(v < TOL64 for v in e_or.values()), e_or
