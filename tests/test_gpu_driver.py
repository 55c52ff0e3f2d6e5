"""End to end through the drop-in executable: inputParam.cfg in, the
reference's .dat files out (antelope.cpp:514-548), checked against the FP64
oracle run on the same parsed configuration."""
import subprocess
from pathlib import Path

import numpy as np
import pytest

from antelope_amd.frontend import parse_config
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
