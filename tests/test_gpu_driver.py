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
