"""Parity of the HIP path (through the C ABI) against the FP64 CPU oracle.

The bar for the FP64 path: j(t) (inter, intra, total), band populations and
the orbital-basis k-summed density matrix each within 1e-10 relative L2 over
the whole time series. The oracle is the generic N x N restatement of
SBEs.h (Jacobi eigensolver per stage); the GPU uses the 2-band Bloch form,
so agreement checks the algebra as well as the kernels.
"""
import numpy as np
import pytest

from antelope_amd import HaldaneModel, Pulse, RunConfig, SBESolver
from tests.cases import haldane_cfg1_model, rel_l2, small_config

pytestmark = pytest.mark.gpu

TOL64 = 1e-10


def _compare(gpu, ref, tol, label=""):
    errs = {}
    jt_g = gpu["j_inter"] + gpu["j_intra"]
    jt_r = ref["j_inter"] + ref["j_intra"]
    errs["j_total"] = rel_l2(jt_g, jt_r)
    errs["j_inter"] = rel_l2(gpu["j_inter"], ref["j_inter"])
    if np.linalg.norm(ref["j_intra"]) > 0:
        errs["j_intra"] = rel_l2(gpu["j_intra"], ref["j_intra"])
    else:
        errs["j_intra"] = float(np.abs(gpu["j_intra"]).max())
    errs["band_pop"] = rel_l2(gpu["band_pop"], ref["band_pop"])
    errs["rho_bar"] = rel_l2(gpu["rho_bar"], ref["rho_bar"])
    print(label, {k: f"{v:.2e}" for k, v in errs.items()}, flush=True)
    bad = {k: v for k, v in errs.items() if not v <= tol}
    assert not bad, f"{label}: parity failed {bad}"
    return errs


def _run_gpu(cfg, model, k_begin=-1, k_end=-1, splits=None):
    with SBESolver(cfg, model, k_begin, k_end) as s:
        if splits:
            b = 0
            for e in splits + [s.nt]:
                s.run(b, e)
                b = e
        else:
            s.run()
        s.synchronize()
        return s.series(), s.state(), s.info


@pytest.mark.parametrize(
    "label,kw",
    [
        ("rk4_t2_interintra", dict()),
        ("rk5_fehlberg", dict(rk=5)),
        ("no_dephasing", dict(T2=-1.0)),
        ("total_current_only", dict(inter_intra=False)),
        ("t1_and_t2", dict(T1=300.0)),
        ("cos4_elliptic", dict(env="cos4", ellip=0.5)),
        ("gauss", dict(env="gauss", ncycles=0.7)),
        ("mp_grid", dict(grid="mp")),
        ("ragged_grid", dict(n=(7, 5))),
        ("ksfactor", dict(ksfactor=(0.5, 2.0, 1.0))),
    ],
)
def test_parity_small(oracle, label, kw):
    cfg = small_config(**kw)
    model = haldane_cfg1_model()
    ref = oracle.run(cfg, model, nthreads=8, want_state=True)
    gpu, state, info = _run_gpu(cfg, model)
    assert info.nt == ref["info"].nt
    assert abs(info.dV - ref["info"].dV) <= 1e-15 * ref["info"].dV
    _compare(gpu, ref, TOL64, label)
    assert rel_l2(state, ref["state"]) < TOL64


def test_parity_fermi_dirac_init(oracle):
    cfg = small_config(init="fermi", fermi_energy=0.0)
    model = haldane_cfg1_model()
    ref = oracle.run(cfg, model, nthreads=8)
    gpu, _, _ = _run_gpu(cfg, model)
    _compare(gpu, ref, TOL64, "fermi")


def test_k_blocks_add_up(oracle):
    """Three rank-like k blocks (ParaRange split) sum to the whole grid."""
    cfg = small_config(n=(11, 9))
    model = haldane_cfg1_model()
    ref = oracle.run(cfg, model, nthreads=8)
    total = None
    for b, e in [(0, 33), (33, 66), (66, 99)]:
        g, _, _ = _run_gpu(cfg, model, b, e)
        total = g if total is None else {k: total[k] + g[k] for k in g}
    _compare(total, ref, TOL64, "k_blocks")


def test_split_runs_equal_single_run():
    cfg = small_config()
    model = haldane_cfg1_model()
    a, sa, _ = _run_gpu(cfg, model)
    b, sb, _ = _run_gpu(cfg, model, splits=[1, 700, 701, 2000])
    for k in a:
        assert np.array_equal(a[k], b[k]), k  # deterministic reduction order
    assert np.array_equal(sa, sb)


def test_large_grid_long_launches_split_equal():
    """2,359,296 k points with FP32 state: one step's partial rows (2.4 MB) fit
    fewer than 64 times in 128 MiB, so a launch covers 64 steps in a larger
    buffer, and the series take the two-pass reduction (36,864 rows per step).
    A run split at odd steps gives the same bits as one run, and the single run
    of 130 steps is three launches (64 + 64 + 2)."""
    cfg = small_config(n=1536, ncycles=0.5, precision="fp32")
    model = haldane_cfg1_model()

    def run(bounds):
        with SBESolver(cfg, model) as s:
            s.set_timing(True)
            b = 0
            for e in bounds:
                s.run(b, e)
                b = e
            s.synchronize()
            return s.series(), s.state(), s.timing()

    a, sa, ta = run([130])
    assert ta["step_launches"] == 3, ta
    b, sb, _ = run([1, 64, 65, 130])
    for k in a:
        assert np.array_equal(a[k], b[k]), k
    assert np.array_equal(sa, sb)
    assert np.abs(a["j_inter"][1:130]).max() > 0  # the pulse has started


def test_single_k_point(oracle):
    cfg = small_config(n=(1, 1))
    model = haldane_cfg1_model()
    ref = oracle.run(cfg, model)
    gpu, _, _ = _run_gpu(cfg, model)
    _compare(gpu, ref, TOL64, "single_k")


def test_parity_plain_reference_case(oracle):
    """60 x 60 grid, 10-cycle sin^2 pulse, T2 = 220, RK4 dt = 0.25 (17951 steps)."""
    cfg = small_config(n=60, ncycles=10.0)
    model = haldane_cfg1_model()
    gpu, _, info = _run_gpu(cfg, model)
    assert info.nt == 17951
    ref = oracle.run(cfg, model, nthreads=16)
    _compare(gpu, ref, TOL64, "plain_60x60")


def test_fp32_state_error_budget():
    """FP32 state + FP64 sums against the FP64 GPU path (same discretisation)."""
    model = haldane_cfg1_model()
    c64 = small_config(n=60, ncycles=10.0, ellip=0.5)
    c32 = small_config(n=60, ncycles=10.0, ellip=0.5, precision="fp32")
    g64, _, _ = _run_gpu(c64, model)
    g32, _, _ = _run_gpu(c32, model)
    jt64 = g64["j_inter"] + g64["j_intra"]
    jt32 = g32["j_inter"] + g32["j_intra"]
    e_j = rel_l2(jt32, jt64)
    e_p = rel_l2(g32["band_pop"], g64["band_pop"])
    print(f"fp32 vs fp64: j {e_j:.2e} pop {e_p:.2e}", flush=True)
    assert e_j < 1e-5 and e_p < 1e-6
