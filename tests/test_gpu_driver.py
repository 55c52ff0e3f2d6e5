"""End to end through the drop-in executable: inputParam.cfg in, the
reference's .dat files out (antelope.cpp:514-548), checked against the FP64
oracle run on the same parsed configuration."""
import math
import subprocess
from pathlib import Path

import numpy as np
import pytest

from antelope_amd.frontend import parse_config
from antelope_amd.params import Pulse, RunConfig
from antelope_amd.solver import library_path
from tests.cases import rel_l2

pytestmark = pytest.mark.gpu

CFG = """
target = "Haldane";
laser = { pulses = ( { E0 = 0.0045; w0 = 0.014; ncycles = 2.; ellip = 0.3; cep = 30.0; t0 = 0.0;
                       phix = 0.; env_name = "sin2"; } ); };
calc = { Npoints = [12, 10]; dt = 0.25; T2 = 220.; ksfactor = [1.0, 1.0, 1.0]; RKorder = 4; InterIntra = true; };
Haldane = { a0 = 1.41978; t1 = 0.075; t2 = 0.025; phi0 = 0.9424777960769379; Mt2 = 2.54; };
"""


def test_driver_writes_reference_files(tmp_path, oracle):
    (tmp_path / "inputParam.cfg").write_text(CFG)
    exe = library_path().parent / "antelope_mi355x"
    r = subprocess.run([str(exe), "--quiet"], cwd=tmp_path, capture_output=True, text=True, timeout=120)
    print(r.stdout[-500:], r.stderr[-500:], flush=True)
    assert r.returncode == 0
    spec = parse_config(path=tmp_path / "inputParam.cfg")
    ref = oracle.run(spec, spec_model(spec), nthreads=8)
    dV, nk = ref["info"].dV, 120
    inter = np.loadtxt(tmp_path / "interband_dipole_full_evol.dat")
    intra = np.loadtxt(tmp_path / "intraband_current_full_evol.dat")
    occ = np.loadtxt(tmp_path / "occupation__full__evol.dat")
    assert inter.shape == (ref["info"].nt, 3)
    assert rel_l2(inter + intra, (ref["j_inter"] + ref["j_intra"]) * dV) < 1e-10
    assert rel_l2(intra, ref["j_intra"] * dV) < 1e-10
    diag = np.stack([ref["rho_bar"][:, 0, 0].real, ref["rho_bar"][:, 1, 1].real], 1) / nk
    assert rel_l2(occ[:, :2], diag) < 1e-10
    assert (tmp_path / "outlaserdata.dat").exists() and (tmp_path / "coherence__full__evol.dat").exists()


class _M:
    """Adapter: the oracle binding takes objects with to_c() and nband."""

    def __init__(self, m):
        self.m, self.nband = m, m.nband

    def to_c(self):
        return self.m, []


def spec_model(spec):
    return _M(spec.model)


ATL = """
BEGIN GLOBAL_VARIABLES
    pi  = acos(-1)
    E0  = 0.0045 au
    la0 = 1.41978 ang
    t1  = 0.075 au
    t2  = t1/3.0
END GLOBAL_VARIABLES
BEGIN SIMULATION_SETUP
    kgrid.n1 = 12
    kgrid.n2 = 10
    tgrid.step = 0.25
    T2 = 220
    InterIntra = TRUE
END SIMULATION_SETUP
BEGIN SYSTEM
    E(k).model   = MODEL_HALDANE_BANDS( la0, t1, t2, 0.3*pi, 2.54*t2 )
    U_f(t).model = LASER_SIN2( E0, 0.014, 2, 0.3, 30 deg )
END SYSTEM
BEGIN SOLVER
    tsolver = RUNGE_KUTTA_4()
END SOLVER
"""


def test_driver_reads_atl(tmp_path):
    """The same solve written as .atl gives the same files as inputParam.cfg."""
    exe = library_path().parent / "antelope_mi355x"
    (tmp_path / "cfg").mkdir()
    (tmp_path / "atl").mkdir()
    (tmp_path / "cfg" / "inputParam.cfg").write_text(CFG)
    (tmp_path / "atl" / "run.atl").write_text(ATL)
    for d, args in (("cfg", []), ("atl", ["run.atl"])):
        r = subprocess.run([str(exe), *args, "--quiet"], cwd=tmp_path / d, capture_output=True, text=True, timeout=120)
        print(d, r.stderr[-300:], flush=True)
        assert r.returncode == 0
    for name in ("interband_dipole_full_evol.dat", "intraband_current_full_evol.dat", "occupation__full__evol.dat"):
        a = np.loadtxt(tmp_path / "cfg" / name)
        b = np.loadtxt(tmp_path / "atl" / name)
        assert a.shape == b.shape and rel_l2(b, a) < 1e-12, name


def test_wannier_tb_target_runs(tmp_path):
    """A Wannier90 _tb.dat target in inputParam.cfg gives the series of the model it holds."""
    from antelope_amd import SBESolver
    from antelope_amd.frontend import write_tb
    from antelope_amd.models import synthetic_wannier
    model = synthetic_wannier(4, nmax=(1, 1, 0), seed=77)
    write_tb(tmp_path / "w_tb.dat", model)
    text = CFG.replace('target = "Haldane"', 'target = "W90"').replace("Npoints = [12, 10]", "Npoints = [4, 4]")
    (tmp_path / "inputParam.cfg").write_text(text + '\nW90 = { dataName = "w_tb.dat"; Nval = 1; };\n')
    spec = parse_config(path=tmp_path / "inputParam.cfg")
    with SBESolver(spec, spec.solver_model()) as s:
        s.run()
        a = s.series()
    cfg = RunConfig(npoints=[4, 4], pulses=[Pulse(E0=0.0045, w0=0.014, ncycles=2.0, ellip=0.3, cep=math.radians(30.0),
                                                  envelope="sin2")], dt=0.25, T2=220.0, rk_order=4, inter_intra=True)
    model.nval = 1
    with SBESolver(cfg, model) as s:
        s.run()
        b = s.series()
    assert rel_l2(a["j_inter"] + a["j_intra"], b["j_inter"] + b["j_intra"]) < 1e-12
    assert rel_l2(a["band_pop"], b["band_pop"]) < 1e-12


def _band(oracle, model, k, A, rho):
    """U^dag rho U with U = eig(H(k + A)), columns phase-fixed as the library does."""
    ev, U = oracle.eigh(oracle.hamiltonian(model, k + A))
    N = U.shape[0]
    for j in range(N):
        p = next(r for r in range(N) if abs(U[r, j]) ** 2 * 2 * N >= 1.0)
        U[:, j] *= np.conj(U[p, j]) / abs(U[p, j])
    return U.conj().T @ rho @ U


@pytest.mark.parametrize("which", ["haldane", "wannier", "kanemele"])
def test_band_basis_export_length_gauge(oracle, which):
    """sbe_get_state(SBE_BASIS_BAND) in the LengthWannier gauge: U^dag rho U at
    k + A(t) of the next sample time (WannierToHamiltonian, SBEs.h:927-946)."""
    from antelope_amd import AnalyticModel, SBESolver
    from antelope_amd.models import synthetic_wannier
    from tests.cases import haldane_cfg1_model, small_config
    if which == "haldane":
        model = haldane_cfg1_model()
    elif which == "wannier":
        model = synthetic_wannier(6, nmax=(1, 1, 0), seed=31)
        model.nval = 3
    else:
        model = AnalyticModel.from_reference("KaneMele", t1=0.075, t2=0.025, tR=0.01, Mt2=1.5, phi0=0.9, a0=1.42)
    cfg = small_config(n=(5, 4), ncycles=1.0, ellip=0.3, grid="mp")
    nsteps = 600
    ref = oracle.run(cfg, model, nsteps=nsteps, want_state=True, nthreads=8)
    with SBESolver(cfg, model) as s:
        s.run(0, nsteps)
        got = s.state("band")
        orb = s.state("orbital")
    k, *_ = oracle.kgrid(cfg, model)
    A, _, atmin, *_ = oracle.laser_eval(cfg, np.array([ref["info"].atmin + ref["info"].dt * nsteps]))
    want = np.array([_band(oracle, model, k[i], A[0], ref["state"][i]) for i in range(len(k))])
    assert rel_l2(orb, ref["state"]) < 1e-10
    assert rel_l2(got, want) < 1e-10, rel_l2(got, want)


H2L_CFG = """
target = "Haldane2L";
laser = { pulses = ( { E0 = 0.004; w0 = 0.014; ncycles = 2.; ellip = 0.3; cep = 0.0; t0 = 0.0; phix = 0.;
                       env_name = "sin2"; } ); };
calc = { Npoints = [6, 5]; dt = 0.5; T2 = 150.; ksfactor = [1.0, 1.0, 1.0]; RKorder = 4; InterIntra = true;
         grid = "mp"; shotNumber = 2; };
Haldane2L = { a0 = 1.42; t1 = 0.075; t2 = 0.025; Mt2 = 1.0; phi0 = 0.9; t11 = 0.01; polytype = "AB"; };
"""


def test_driver_haldane2l_with_snapshots(tmp_path, oracle):
    """A Haldane2L config through the drop-in executable: the .dat series match the
    oracle, and the snapshots (calc.shotNumber = 2: every int(T/2/dt) steps, before
    the step) hold the band-basis conduction occupation and coherence of each k."""
    from antelope_amd.frontend import snapshot_interval
    (tmp_path / "inputParam.cfg").write_text(H2L_CFG)
    exe = library_path().parent / "antelope_mi355x"
    r = subprocess.run([str(exe), "--quiet"], cwd=tmp_path, capture_output=True, text=True, timeout=120)
    print(r.stdout[-300:], r.stderr[-300:], flush=True)
    assert r.returncode == 0
    spec = parse_config(path=tmp_path / "inputParam.cfg")
    model = spec.solver_model()
    ref = oracle.run(spec, model, nthreads=8)
    dV = ref["info"].dV
    inter = np.loadtxt(tmp_path / "interband_dipole_full_evol.dat")
    intra = np.loadtxt(tmp_path / "intraband_current_full_evol.dat")
    assert rel_l2(inter + intra, (ref["j_inter"] + ref["j_intra"]) * dV) < 1e-10
    shot = snapshot_interval(spec, spec.shot_number)
    nt = ref["info"].nt
    steps = list(range(0, nt, shot))
    assert shot == int(2 * math.pi / 0.014 / 2 / 0.5) and len(steps) == 5
    nk = 30
    nc = np.fromfile(tmp_path / "fisrt_conduction.dat").reshape(len(steps), nk)
    pi = np.fromfile(tmp_path / "fisrt_coherence.dat").reshape(len(steps), nk, 2)
    occ = np.loadtxt(tmp_path / "occupation_1d_integration.dat")
    k, w, *_ = oracle.kgrid(spec, model)
    for i, n in enumerate(steps):
        st = oracle.run(spec, model, nsteps=n, want_state=True, nthreads=8)["state"] if n else None
        A, *_ = oracle.laser_eval(spec, np.array([ref["info"].atmin + ref["info"].dt * n]))
        if st is None:  # initial state
            st = oracle.run(spec, model, nsteps=0, want_state=True)["state"]
        band = np.array([_band(oracle, model, k[j], A[0], st[j]) for j in range(nk)])
        np.testing.assert_allclose(nc[i], band[:, 2, 2].real, atol=1e-10)
        np.testing.assert_allclose(pi[i, :, 0] + 1j * pi[i, :, 1], band[:, 2, 1], atol=1e-10)
        want = np.array([np.sum((w * band[:, 2, 2].real)[j * 5:(j + 1) * 5]) for j in range(6)]) / 5
        np.testing.assert_allclose(occ[i], want, rtol=1e-6, atol=1e-12)
