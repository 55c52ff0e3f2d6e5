"""The C-ABI boundary on the CPU: the library loads, exports every function
declared in include/antelope/sbe_c.h, its structs have the layout the
ctypes mirror assumes, and invalid inputs come back as status codes (the
reference exits the process instead, e.g. SBEs.h:472-480, laser.h:81-85)."""
import ctypes as C
import dataclasses
import re
import subprocess
from pathlib import Path

import numpy as np
import pytest

from antelope_amd import _abi as A
from antelope_amd.models import synthetic_wannier
from antelope_amd import HaldaneModel, Pulse, RunConfig, SBEBatch, SBESolver, laser_eval, query_footprint
from antelope_amd.solver import SBEError, library_path, load_library
from tests.cases import haldane_cfg1_model, small_config

ROOT = Path(__file__).resolve().parent.parent
HEADERS = sorted((ROOT / "include").rglob("*.h"))


def declared_functions():
    names = set()
    for h in HEADERS:
        text = re.sub(r"/\*.*?\*/", "", h.read_text(), flags=re.S)
        names |= set(re.findall(r"\b(sbe_[a-z_0-9]+)\s*\(", text))
    return sorted(names)


def test_library_exports_every_declared_function():
    lib = load_library()
    names = declared_functions()
    assert len(names) >= 10
    for n in names:
        assert hasattr(lib, n), n
    assert set(names) <= set(A.SIGNATURES) | {"sbe_batch_create"} or set(names) >= set(A.SIGNATURES)


def test_struct_layout_matches_header(tmp_path):
    probe = tmp_path / "probe.c"
    fields = {
        "sbe_pulse": [f for f, _ in A.sbe_pulse._fields_],
        "sbe_config": [f for f, _ in A.sbe_config._fields_],
        "sbe_model": [f for f, _ in A.sbe_model._fields_],
        "sbe_info": [f for f, _ in A.sbe_info._fields_],
        "sbe_request_status": [f for f, _ in A.sbe_request_status._fields_],
        "sbe_step_stats": [f for f, _ in A.sbe_step_stats._fields_],
        "sbe_tuning": [f for f, _ in A.sbe_tuning._fields_],
    }
    lines = ['#include <stdio.h>', '#include <stddef.h>', '#include "antelope/sbe_c.h"',
             '#include "antelope/sbe_server.h"', "int main(void){"]
    for s, fs in fields.items():
        lines.append(f'printf("{s} %zu\\n", sizeof({s}));')
        for f in fs:
            lines.append(f'printf("{s}.{f} %zu\\n", offsetof({s}, {f}));')
    lines.append("return 0;}")
    probe.write_text("\n".join(lines))
    exe = tmp_path / "probe"
    subprocess.run(["gcc", "-std=c99", f"-I{ROOT / 'include'}", str(probe), "-o", str(exe)], check=True)
    out = dict(ln.rsplit(" ", 1) for ln in subprocess.run([str(exe)], capture_output=True, text=True,
                                                         check=True).stdout.splitlines())
    for s, fs in fields.items():
        cls = getattr(A, s)
        assert int(out[s]) == C.sizeof(cls), s
        for f in fs:
            assert int(out[f"{s}.{f}"]) == getattr(cls, f).offset, f"{s}.{f}"


def _expect(code, fn):
    with pytest.raises(SBEError) as e:
        fn()
    assert e.value.code == code, str(e.value)


def test_invalid_inputs_return_status_codes():
    m = haldane_cfg1_model()
    cfg = small_config()
    _expect(A.SBE_ERR_CONFIG, lambda: SBESolver(dataclasses.replace(cfg, dt=0.0), m))
    _expect(A.SBE_ERR_CONFIG, lambda: SBESolver(dataclasses.replace(cfg, rk_order=3), m))
    _expect(A.SBE_ERR_CONFIG, lambda: SBESolver(dataclasses.replace(cfg, npoints=[0, 4]), m))
    _expect(A.SBE_ERR_CONFIG, lambda: SBESolver(dataclasses.replace(cfg, pulses=[]), m))
    _expect(A.SBE_ERR_UNSUPPORTED, lambda: SBESolver(dataclasses.replace(cfg, gauge="LengthHamiltonian"), m))
    _expect(A.SBE_ERR_MODEL, lambda: SBESolver(cfg, dataclasses.replace(m, a0=-1.0)))
    _expect(A.SBE_ERR_RANGE, lambda: SBESolver(cfg, m, 0, 10_000))
    _expect(A.SBE_ERR_RANGE, lambda: SBESolver(cfg, m, 50, 40))
    ck = dataclasses.replace(cfg, rk_order=45)
    _expect(A.SBE_ERR_UNSUPPORTED, lambda: SBESolver(dataclasses.replace(ck, T1=100.0), m))
    _expect(A.SBE_ERR_CONFIG, lambda: SBESolver(dataclasses.replace(ck, rtol=0.0), m))
    _expect(A.SBE_ERR_UNSUPPORTED, lambda: SBESolver(ck, synthetic_wannier(4, nmax=(1, 1, 0))))
    # the velocity gauge runs the fixed-step tableaus, FP64 or FP32, not batched
    vel = dataclasses.replace(cfg, gauge="VelocityHamiltonian")
    _expect(A.SBE_ERR_UNSUPPORTED, lambda: SBESolver(dataclasses.replace(vel, rk_order=45), m))
    _expect(A.SBE_ERR_UNSUPPORTED,
            lambda: SBESolver(dataclasses.replace(vel, precision="split16"), synthetic_wannier(4, nmax=(1, 1, 0))))
    _expect(A.SBE_ERR_UNSUPPORTED, lambda: SBEBatch([vel, vel], [m, m]))
    assert query_footprint(vel, m) == query_footprint(cfg, m)


def test_laser_eval_and_footprint_without_gpu():
    cfg = small_config(n=300, ncycles=12.0, ellip=0.5)
    m = haldane_cfg1_model()
    Av, Ev, atmin, nt = laser_eval(cfg, [0.0])
    assert nt == 21542
    assert Ev[0, 0] == pytest.approx(0.0045 / (1.25 ** 0.5))
    b = query_footprint(cfg, m)
    assert 90_000 * 32 < b < 512 * 2 ** 20


def _footprint_env(cfg, m, dense, fused=False):
    import os
    env = {"ANTELOPE_NB_DENSE": "1" if dense else "0", "ANTELOPE_NB_FUSED": "1" if fused else "0"}
    old = {k: os.environ.get(k) for k in env}
    os.environ.update(env)
    try:
        return query_footprint(cfg, m)
    finally:
        for k, v in old.items():
            if v is None:
                del os.environ[k]
            else:
                os.environ[k] = v


def test_separable_sums_footprint_without_gpu():
    """A Wannier model with R on a 5 x 5 x 3 box on a 6 x 5 x 2 grid takes the separable
    Fourier sums; the footprint adds exactly their buffers: B on the box, the two pass
    outputs (columns padded to 4 N^2 rounded to 64, complex FP64), the axis factors
    [K_i][M_i], the R map and the k map, each 256-byte aligned, and drops the dense sums'
    e^{ik.R} table [nk_stride][N_R] and their B buffer [N_R padded to 32][columns]. Where
    the passes would not cost under half the dense sum (2 x 2 x 1 grid) nothing is added."""
    import dataclasses

    from antelope_amd import Pulse, RunConfig
    from antelope_amd.models import synthetic_wannier

    def cfg_for(npts):
        return RunConfig(npoints=list(npts), pulses=[Pulse(E0=0.01, w0=0.05, ncycles=0.25, envelope="sin2")],
                         dt=0.5, T2=80.0, rk_order=4, inter_intra=True)

    N, M, K = 8, (5, 5, 3), (6, 5, 2)
    m = synthetic_wannier(N, nmax=(2, 2, 1), seed=21)
    cfg = cfg_for(K)
    al = lambda x: (x + 255) // 256 * 256
    col = 16 * ((4 * N * N + 63) // 64 * 64)
    cells = M[0] * M[1] * M[2]
    want = (al(cells * col) + al(M[0] * M[1] * K[2] * col) + al(M[0] * K[1] * K[2] * col)
            + sum(al(K[i] * M[i] * 16) for i in range(3)) + al(cells * 4) + al(K[0] * K[1] * K[2] * 4))
    nk_stride, nr = (K[0] * K[1] * K[2] + 63) // 64 * 64, cells  # the dense e^{ik.R} table is not kept
    want -= nk_stride * nr * 16 + (nr + 31) // 32 * 32 * col
    assert _footprint_env(cfg, m, dense=False) - _footprint_env(cfg, m, dense=True) == want
    # the fused RK4 step evaluates the separable sums of S steps per batch of launches: it
    # adds their second-pass outputs [M1][K2][K3][columns] (J and the midpoint pair per step,
    # the end pairs in a ring of S + 1), the band energies [nk_stride][16] (FP64), B on the
    # box and a first-pass output per evaluation (three a step), and A, E at the three times
    # of every step; S for about 256 MB of them, at most 64; its 10 partial rows per step fit
    # in the 60 the sample kernel has
    nt = laser_eval(cfg, [0.0])[3]
    t2, b, t1 = al(M[0] * K[1] * K[2] * col), al(cells * col), al(M[0] * M[1] * K[2] * col)
    S = max(1, min(nt, 21845, 64, (256 << 20) // (3 * (t2 + b + t1))))
    fused = (3 * S + 1) * t2 + al(nk_stride * 16 * 8) + 3 * S * (b + t1) + al(nt * 18 * 8)
    assert _footprint_env(cfg, m, dense=False, fused=True) - _footprint_env(cfg, m, dense=False) == fused
    assert _footprint_env(cfg, m, dense=True, fused=True) == _footprint_env(cfg, m, dense=True)
    rk5 = dataclasses.replace(cfg, rk_order=5)  # the fused step is the classical RK4 tableau only
    assert _footprint_env(rk5, m, dense=False, fused=True) == _footprint_env(rk5, m, dense=False)
    small = cfg_for((2, 2, 1))
    m1 = synthetic_wannier(N, nmax=(1, 1, 1), seed=21)
    assert _footprint_env(small, m1, dense=False) == _footprint_env(small, m1, dense=True)


def test_fused_step_n32_footprint_without_gpu():
    """For 16 < N <= 32 the fused RK4 step keeps its matrices in the mats and stage slots
    the separate kernels have and adds the band energies [nk_stride][32] (FP64). With at
    most 32 lattice vectors and at most 1024 k points the evaluations of a run of steps
    are one dense launch, which reads the e^{ik.R} table [nk_stride][N_R] (kept by the
    dense sums anyway); otherwise, with the separable sums, each of up to four batched
    evaluations has B on the box and both pass outputs. RK5 and the SPLIT16 sums keep
    the separate kernels."""
    import dataclasses

    from antelope_amd import Pulse, RunConfig
    from antelope_amd.models import synthetic_wannier

    def cfg_for(npts):
        return RunConfig(npoints=list(npts), pulses=[Pulse(E0=0.003, w0=0.014, ncycles=2.0, envelope="sin2")],
                         dt=0.25, T2=220.0, rk_order=4, inter_intra=True)

    al = lambda x: (x + 255) // 256 * 256
    N = 24
    col = 16 * ((4 * N * N + 63) // 64 * 64)
    # 25 lattice vectors, 144 k points: the matrices of S steps per dense launch (J, the
    # midpoint pair, the end pair in a ring of S + 1), S for about 256 MB of them (at most
    # 64), and the pulse's A and E at the three evaluation times of every step
    m = synthetic_wannier(N, nmax=(2, 2, 0), seed=3)
    cfg = cfg_for((12, 12))
    nt = laser_eval(cfg, [0.0])[3]
    nk_stride = (144 + 63) // 64 * 64
    lam = al(nk_stride * 32 * 8)
    mat = nk_stride * N * N * 16
    S = max(1, min(nt, 64, (256 << 20) // (7 * mat)))
    steps = al((7 * S + 2) * mat) + al(nt * 18 * 8)
    assert _footprint_env(cfg, m, dense=False, fused=True) - _footprint_env(cfg, m, dense=False) == \
        lam + steps + nk_stride * 25 * 16
    assert _footprint_env(cfg, m, dense=True, fused=True) - _footprint_env(cfg, m, dense=True) == lam + steps
    # 75 lattice vectors on a 5 x 5 x 3 box, 6 x 5 x 2 grid: the separable passes, batched
    M, K = (5, 5, 3), (6, 5, 2)
    m75 = synthetic_wannier(N, nmax=(2, 2, 1), seed=3)
    cfg75 = cfg_for(K)
    nk_stride = (60 + 63) // 64 * 64
    scratch = 4 * (al(M[0] * M[1] * M[2] * col) + al(M[0] * M[1] * K[2] * col) + al(M[0] * K[1] * K[2] * col))
    assert _footprint_env(cfg75, m75, dense=False, fused=True) - _footprint_env(cfg75, m75, dense=False) == \
        al(nk_stride * 32 * 8) + scratch
    for other in (dataclasses.replace(cfg, rk_order=5), dataclasses.replace(cfg, precision="split16")):
        assert _footprint_env(other, m, dense=False, fused=True) == _footprint_env(other, m, dense=False)


def test_separable_sums_eligibility_without_gpu():
    """The separable form needs theta_j = k.a_j to depend on grid index j alone. On a
    hexagonal lattice, calc.BZaxes given as Cartesian axes break that (the footprint is
    the dense one); the reciprocal vectors given explicitly keep it; an R listed twice
    (the dense sums add both) also keeps the dense form."""
    import dataclasses
    import math

    from antelope_amd import Pulse, RunConfig
    from antelope_amd.models import synthetic_wannier

    a = 5.0
    hexl = np.array([[a, 0, 0], [a / 2, a * math.sqrt(3) / 2, 0], [0, 0, 10.0]])
    m = synthetic_wannier(8, nmax=(2, 2, 1), seed=21, lattice=hexl)
    base = RunConfig(npoints=[6, 5, 2], pulses=[Pulse(E0=0.01, w0=0.05, ncycles=0.25, envelope="sin2")],
                     dt=0.5, T2=80.0, rk_order=4, inter_intra=True)
    box, dense = _footprint_env(base, m, dense=False), _footprint_env(base, m, dense=True)
    assert box != dense
    cart = dataclasses.replace(base, bz_axes=[1.2, 0, 0, 0, 1.2, 0, 0, 0, 0.6])
    assert _footprint_env(cart, m, dense=False) == _footprint_env(cart, m, dense=True)
    recip = 2 * math.pi * np.linalg.inv(hexl).T  # rows b_i with b_i . a_j = 2 pi delta_ij
    rec = dataclasses.replace(base, bz_axes=list(recip.ravel()))
    assert _footprint_env(rec, m, dense=False) == box
    dup = dataclasses.replace(m, rvec=np.concatenate([m.rvec, m.rvec[:1]]), rweight=np.concatenate([m.rweight, [1.0]]),
                              ham=np.concatenate([m.ham, m.ham[:1]]), pos=np.concatenate([m.pos, m.pos[:1]]))
    assert _footprint_env(base, dup, dense=False) == _footprint_env(base, dup, dense=True)


def test_server_bookkeeping_without_gpu():
    """Submission validates and prices a request on the host; nothing runs until a round."""
    import math

    from antelope_amd import HaldaneModel, Pulse, RunConfig
    from antelope_amd.serve import SBEServer
    from antelope_amd.solver import query_footprint

    t1 = 0.075
    model = HaldaneModel(a0=2.683, t1=t1, t2=t1 / 3, phi0=0.3 * math.pi, M0=2.54 * t1 / 3)
    cfg = RunConfig(npoints=[40, 40], pulses=[Pulse(E0=0.0045, w0=0.014, ncycles=2.0)], dt=0.25, rk_order=4)
    fp = query_footprint(cfg, model)
    srv = SBEServer(hbm_budget_bytes=3 * fp, max_active=2, slice_steps=100)
    a, b = srv.submit(cfg, model), srv.submit(cfg, model)
    assert b == a + 1
    st = srv.poll(a)
    assert st.state == A.SBE_REQ_QUEUED and st.footprint == fp and st.t_submit > 0
    assert srv.counts() == (2, 0, 0, 3 * fp)
    _expect(A.SBE_ERR_STATE, lambda: srv.release(a))        # still queued
    _expect(A.SBE_ERR_RANGE, lambda: srv.poll(12345))
    srv.cancel(b)                                            # queued: dropped at once, nothing ran
    assert srv.poll(b).state == A.SBE_REQ_CANCELLED and srv.counts()[0] == 1
    _expect(A.SBE_ERR_STATE, lambda: srv.cancel(b))
    _expect(A.SBE_ERR_RANGE, lambda: srv.cancel(12345))
    srv.release(b)
    small = SBEServer(hbm_budget_bytes=fp - 1)
    _expect(A.SBE_ERR_OOM, lambda: small.submit(cfg, model))  # can never fit
    bad = RunConfig(npoints=[40, 40], pulses=[Pulse(E0=0.0045, w0=0.014, ncycles=2.0)], dt=-1.0)
    _expect(A.SBE_ERR_CONFIG, lambda: srv.submit(bad, model))
    small.close()
    srv.close()


def test_tuning_cache_without_gpu():
    """The plan cache: empty means the default rules; a stored plan is returned for its
    problem class (k points to the nearest power of two) and not for another; values out of
    range are refused; clear empties it."""
    from antelope_amd.solver import tuning_clear, tuning_get, tuning_set
    tuning_clear()
    m = haldane_cfg1_model()
    cfg = small_config(n=64)
    assert tuning_get(cfg, m)["tuned"] == 0
    tuning_set(cfg, m, kpt=2, chunk_steps=37)
    t = tuning_get(cfg, m)
    assert (t["tuned"], t["kpt"], t["chunk_steps"]) == (1, 2, 37)
    assert tuning_get(small_config(n=63), m)["kpt"] == 2          # 3969 k: the same power of two as 4096
    assert tuning_get(small_config(n=128), m)["tuned"] == 0       # another class
    assert tuning_get(small_config(n=64, precision="fp32"), m)["tuned"] == 0
    _expect(A.SBE_ERR_CONFIG, lambda: tuning_set(cfg, m, kpt=3))
    tuning_clear()
    assert tuning_get(cfg, m)["tuned"] == 0
