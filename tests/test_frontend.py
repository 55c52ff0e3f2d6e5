"""Input/output front end on the CPU: inputParam.cfg reading (antelope.cpp:90-138,
SBEs.h:432-561, Materials/Haldane.h:46-84) and the reference's output files
(antelope.cpp:208-235, 514-548)."""
import math
import re
from pathlib import Path

import numpy as np
import pytest

from antelope_amd import _abi as A
from antelope_amd.frontend import parse_config, write_outputs
from antelope_amd.solver import SBEError, laser_eval

GOLDEN = Path(__file__).parent / "golden" / "configs"
REF_CFG = Path("/root/reference/src/GeneralModel/inputParam.cfg")


def test_parse_plain_reference_case():
    s = parse_config(path=GOLDEN / "haldane_plain_60.cfg")
    c, m = s.config, s.model
    assert list(c.npoints) == [60, 60, 1]
    assert c.dt == 0.25 and c.T2 == 220.0 and c.T1 == -1.0
    assert c.rk_order == 4 and c.inter_intra == 1 and c.gauge == A.SBE_GAUGE_LENGTH_WANNIER
    assert c.npulses == 1
    p = c.pulses[0]
    assert (p.E0, p.w0, p.ncycles, p.envelope) == (0.0045, 0.014, 10.0, A.SBE_ENV_SIN2)
    assert m.kind == A.SBE_MODEL_HALDANE and m.nband == 2 and m.nval == 1
    assert m.a0 == pytest.approx(1.41978 / 0.529177210903, rel=1e-15)
    assert m.M0 == pytest.approx(2.54 * 0.025, rel=1e-15)
    assert s.warnings == ""


BASE = """
target = "Haldane";
laser = { pulses = ( { E0 = 0.01; w0 = 0.05; ncycles = 4; ellip = 0.5; cep = 90.0; t0 = 1.0;
                       phix = 180.0; env_name = "cos4"; thetaz = 45.0; } ); };
calc = { Npoints = [8, 4]; %s };
Haldane = { a0 = 1.0; t1 = 0.075; t2 = 0.025; phi0 = 1.16; M0 = 0.01; };
"""


def test_defaults_units_and_warnings():
    s = parse_config(BASE % "")
    c = s.config
    assert c.dt == 0.1 and c.rk_order == 5 and c.inter_intra == 0       # SBEs.h:438-455
    assert c.T1 == -1.0 and c.T2 == -1.0 and list(c.ksfactor) == [1.0, 1.0, 1.0]
    assert "ksfactor" in s.warnings and "gauge" in s.warnings
    p = c.pulses[0]
    assert p.cep == pytest.approx(math.pi / 2) and p.phix == pytest.approx(math.pi)
    assert p.thetaz == pytest.approx(math.pi / 4) and p.phiz == 0.0 and p.ncycles == 4.0
    assert s.model.M0 == 0.01


def test_libconfig_syntax_features():
    text = """
    // colon separators, hex, string concatenation, all comment kinds, trailing commas
    target : "Hal" "dane";  # concatenated
    gauge = "Wannier"
    laser : { offset = [100, 50.5]; efield = "consistent";
              pulses = ( { E0 = 1e-3; w0 = .05; ncycles = 2; ellip = 0; cep = 0; t0 = 0; phix = 0; env_name = "gauss" },
                         { E0 = 2e-3; w0 = 0.1; ncycles = 3; ellip = 1; cep = 0; t0 = 10; phix = 0;
                           env_name = "gauss_clipped"; ncycles_clipped = 2.0 }, ) };
    calc : { Npoints = [0x10, 4L, 2]; dt = 5e-2; InterIntra = TRUE; RKorder = 4; ksfactor = [1, 2.0, 0.5];
             BZaxes = [1., 0., 0., 0., 1., 0., 0., 0., 1.]; BZorigin = [0.1, 0.2]; grid = "mp";
             precision = "fp32"; T2 = 220 };
    /* block
       comment */
    Haldane : { a0 = 1; t1 = 0.1; t2 = 0.02; phi0 = 0; Mt2 = 3; Ef = -0.1; };
    """
    c = parse_config(text).config
    assert list(c.npoints) == [16, 4, 2] and c.dt == 0.05 and c.inter_intra == 1
    assert (c.offset_before, c.offset_after) == (100.0, 50.5)
    assert c.efield_mode == A.SBE_EFIELD_CONSISTENT and c.grid == A.SBE_GRID_MP and c.precision == A.SBE_PREC_FP32
    assert c.npulses == 2 and c.pulses[1].envelope == A.SBE_ENV_GAUSS_CLIPPED and c.pulses[1].ncycles_clipped == 2.0
    assert c.has_bz_axes == 1 and c.bz_axes[8] == 1.0 and c.has_bz_origin == 1 and c.bz_origin[1] == 0.2
    assert list(c.ksfactor) == [1.0, 2.0, 0.5] and c.T2 == 220.0
    assert c.init_type == A.SBE_INIT_FERMI_DIRAC and c.fermi_energy == -0.1


@pytest.mark.parametrize("text,code", [
    ('target = "Haldane"; laser = {pulses = ();};', A.SBE_ERR_CONFIG),
    (BASE.replace('env_name = "cos4"', 'env_name = "cos2"') % "", A.SBE_ERR_CONFIG),
    (BASE.replace("ellip = 0.5;", "") % "", A.SBE_ERR_CONFIG),
    (BASE.replace("M0 = 0.01;", "") % "", A.SBE_ERR_MODEL),
    (BASE.replace('target = "Haldane"', 'target = "Graphene"') % "", A.SBE_ERR_UNSUPPORTED),
    (BASE.replace('target = "Haldane"', 'target = "KaneMele"') % "", A.SBE_ERR_MODEL),
    ("calc = { Npoints = [4, 4] ", A.SBE_ERR_CONFIG),
    ("calc = { Npoints = [4, 4]; dt = 1.2.3; };", A.SBE_ERR_CONFIG),
])
def test_parse_errors_are_status_codes(text, code):
    with pytest.raises(SBEError) as e:
        parse_config(text)
    assert e.value.code == code, str(e.value)


def test_syntax_error_reports_line():
    with pytest.raises(SBEError) as e:
        parse_config("a = 1;\nb = 2;\nc = = 3;\n")
    assert "line 3" in str(e.value)


@pytest.mark.skipif(not REF_CFG.exists(), reason="reference checkout not present")
def test_reference_inputparam_cfg():
    """The reference's own config parses, its Haldane2L target included: t1_bottom
    is present, so t2_bottom and phi0_bottom are never read (Haldane2L.h:83-88)
    and the bottom layer keeps the top's t2 and phi0; Mt2_bottom scales that t2."""
    text = REF_CFG.read_text()
    p = parse_config(text).solver_model().params
    assert p[4:8] == [0.075, 0.025, -2.54 * 0.025, 1.16] and p[9] == 0.0
    assert p[10:16] == [1.0, -1.0, 0.0, 0.01, 0.01, 0.0]      # polytype AA': h01 = h10 = t11, bottom conjugated
    s = parse_config(text.replace('target = "Haldane2L"', 'target = "Haldane"'))
    c, m = s.config, s.model
    assert list(c.npoints) == [201, 201, 1] and c.dt == 0.1 and c.T2 == 220.0 and c.inter_intra == 1
    assert (c.offset_before, c.offset_after) == (100.0, 100.0) and c.pulses[0].envelope == A.SBE_ENV_GAUSS
    assert c.pulses[0].ellip == 1.0 and m.t2 == 0.0 and m.M0 == pytest.approx(0.07202867146428378)


def test_output_files(tmp_path):
    s = parse_config(path=GOLDEN / "haldane_plain_60.cfg")
    nt, N = 50, 2
    rng = np.random.default_rng(0)
    series = {"j_inter": rng.normal(size=(nt, 3)), "j_intra": rng.normal(size=(nt, 3)),
              "rho_bar": rng.normal(size=(nt, N, N)) + 1j * rng.normal(size=(nt, N, N)),
              "band_pop": rng.uniform(size=(nt, N))}
    info = A.sbe_info()
    info.nt, info.nband, info.nk_total, info.dV, info.total_weight = nt, N, 3600, 2.5e-3, 3481.0
    write_outputs(tmp_path, s.config, s.model, info, series)
    inter = np.loadtxt(tmp_path / "interband_dipole_full_evol.dat")
    intra = np.loadtxt(tmp_path / "intraband_current_full_evol.dat")
    np.testing.assert_allclose(inter, series["j_inter"] * 2.5e-3, rtol=1e-15)
    np.testing.assert_allclose(intra, series["j_intra"] * 2.5e-3, rtol=1e-15)
    occ = np.loadtxt(tmp_path / "occupation__full__evol.dat")
    diag = np.stack([series["rho_bar"][:, 0, 0].real, series["rho_bar"][:, 1, 1].real], 1) / 3600
    np.testing.assert_allclose(occ[:, :2], diag, rtol=1e-15)
    np.testing.assert_allclose(occ[:, 2], diag.sum(1), rtol=1e-14)
    coh = np.loadtxt(tmp_path / "coherence__full__evol.dat")
    np.testing.assert_allclose(coh[:, 0] + 1j * coh[:, 1], series["rho_bar"][:, 0, 1] / 3600, rtol=1e-15)
    bp = np.loadtxt(tmp_path / "band_population__full__evol.dat")
    np.testing.assert_allclose(bp[:, :2], series["band_pop"] / 3481.0, rtol=1e-15)
    assert (tmp_path / "laserParameters.dat").stat().st_size == 0
    first = (tmp_path / "interband_dipole_full_evol.dat").read_text().splitlines()[0]
    assert re.fullmatch(r"(-?\d\.\d{16}e[+-]\d{2}    ){3}", first), first
    laser = np.loadtxt(tmp_path / "outlaserdata.dat")
    assert laser.shape == (nt, 7)
    _, _, atmin, _ = laser_eval(s, np.zeros(1))
    Av, Ev, _, _ = laser_eval(s, atmin + 0.25 * np.arange(nt))
    np.testing.assert_allclose(laser[:, 1:4], Ev, rtol=1e-5, atol=1e-20)  # %e keeps 6 digits
    np.testing.assert_allclose(laser[:, 4:7], Av, rtol=1e-5, atol=1e-20)
    assert laser[0, 0] == pytest.approx(atmin, rel=1e-6)
    line = (tmp_path / "outlaserdata.dat").read_text().splitlines()[1]
    assert re.fullmatch(r"-?\d\.\d{6}e[+-]\d{2}    (-?\d\.\d{6}e[+-]\d{2}      ){6}", line), line


def _w90_model():
    from antelope_amd.models import synthetic_wannier
    return synthetic_wannier(4, nmax=(1, 1, 0), seed=77)


def test_wannier_tb_round_trip(tmp_path):
    """seedname_tb.dat as Wannier90 writes it (Wannier90.h:107-218): eV / Angstrom in the
    file, Hartree / bohr in the model, weights 1/degeneracy, position elements re + i im."""
    from antelope_amd.frontend import read_tb, write_tb
    m = _w90_model()
    m.rweight = np.where(np.arange(len(m.rweight)) % 3 == 0, 0.5, 1.0)   # some degeneracy 2
    f = tmp_path / "wannier90_tb.dat"
    write_tb(f, m, header="synthetic N=4")
    lines = f.read_text().splitlines()
    assert lines[0] == "synthetic N=4" and int(lines[4]) == 4 and int(lines[5]) == len(m.rweight)
    np.testing.assert_allclose(float(lines[1].split()[0]), m.lattice[0, 0] * 0.529177210903, rtol=1e-15)
    # the first Hamiltonian line is element (1, 1) of the first R in eV
    first = next(l for l in lines[8:] if len(l.split()) == 4)
    mm, nn, re, im = first.split()
    assert (mm, nn) == ("1", "1")
    np.testing.assert_allclose(float(re), m.ham[0, 0, 0].real * 27.211386245988, rtol=1e-15)
    back = read_tb(f, nval=2)
    np.testing.assert_array_equal(back.rvec, m.rvec)
    np.testing.assert_allclose(back.rweight, m.rweight, rtol=1e-15)
    np.testing.assert_allclose(back.lattice, m.lattice, rtol=1e-14)
    np.testing.assert_allclose(back.ham, m.ham, rtol=1e-14, atol=1e-18)
    np.testing.assert_allclose(back.pos, m.pos, rtol=1e-14, atol=1e-18)
    assert back.nval == 2 and not back.is_dipole_zero


def test_wannier_target_in_config(tmp_path):
    from antelope_amd.frontend import write_tb
    write_tb(tmp_path / "model_tb.dat", _w90_model())
    cfg = BASE.replace('target = "Haldane"', 'target = "W90"') % "" + '\nW90 = { dataName = "model_tb.dat"; Nval = 2; };\n'
    (tmp_path / "inputParam.cfg").write_text(cfg)
    s = parse_config(path=tmp_path / "inputParam.cfg")
    assert s.model.kind == A.SBE_MODEL_WANNIER and s.model.nband == 4 and s.model.nval == 2
    wm = s.solver_model()
    np.testing.assert_allclose(wm.ham, _w90_model().ham, rtol=1e-14, atol=1e-18)
    fd = parse_config(path=tmp_path / "inputParam.cfg").config  # Nval path
    assert fd.init_type == A.SBE_INIT_UNIFORM_VALENCE
    (tmp_path / "fd.cfg").write_text(cfg.replace("Nval = 2;", "FermiE = 0.5;"))
    c = parse_config(path=tmp_path / "fd.cfg").config
    assert c.init_type == A.SBE_INIT_FERMI_DIRAC and c.fermi_energy == pytest.approx(0.5 / 27.211386245988)
    with pytest.raises(SBEError) as e:
        parse_config(cfg.replace('"model_tb.dat"', '"missing_tb.dat"'))
    assert e.value.code == A.SBE_ERR_CONFIG


def test_wannier_tb_errors(tmp_path):
    from antelope_amd.frontend import read_tb
    f = tmp_path / "bad_tb.dat"
    f.write_text("header\n1 0 0\n0 1 0\n0 0 1\n2\n1\n1\n\n0 0 0\n1 1 0.1 0\n1 2 0 0\n2 1 0 0\n")
    with pytest.raises(SBEError) as e:     # one Hamiltonian element short
        read_tb(f)
    assert e.value.code == A.SBE_ERR_MODEL
    f.write_text("header\n1 0 0\n0 1 0\n0 0 1\n2\n1\n1\n\n0 0 0\n1 1 0.1 0\n1 2 0 0\n2 1 0 0\n2 2 -0.1 0\n")
    m = read_tb(f, nval=1)                 # no position part: dipoles zero
    assert m.is_dipole_zero and m.pos is None
    np.testing.assert_allclose(m.ham[0].real, np.diag([0.1, -0.1]) / 27.211386245988, rtol=1e-15)


ANALYTIC_CFG = """
target = "%s";
gauge = "%s";
laser = { pulses = ( { E0 = 0.003; w0 = 0.014; ncycles = 2; ellip = 0.0; cep = 0.0; t0 = 0.0; phix = 0.0;
                       env_name = "sin2"; } ); };
calc = { Npoints = [6, 6]; dt = 0.25; T2 = 150.; ksfactor = [1.0, 1.0, 1.0]; RKorder = 4; InterIntra = true; };
%s
"""


@pytest.mark.parametrize("target,group,keys", [
    ("WilsonMass", "WilsonMass = { t = 0.02; delta = 0.03; mu = 0.01; a0 = 4.0; };",
     dict(t=0.02, delta=0.03, mu=0.01, a0=4.0)),
    ("SquareSBand", "SquareSBand = { t1 = 0.03; M0 = 0.05; a0 = 2.5; };", dict(t1=0.03, M0=0.05, a0=2.5)),
    ("KaneMele", "KaneMele = { t1 = 0.075; t2 = 0.025; tR = 0.01; Mt2 = 1.5; phi0 = 0.9; a0 = 1.42; };",
     dict(t1=0.075, t2=0.025, tR=0.01, Mt2=1.5, phi0=0.9, a0=1.42)),
    ("XYHlattice", "XYHlattice = { tl = 0.05; tr = 0.03; lambda = 0.01; a0 = 2.46; };",
     {"tl": 0.05, "tr": 0.03, "lambda": 0.01, "a0": 2.46}),
    ("TMDCs", 'TMDCs = { Material = "WSe2"; isSOCEnabled = true; };', dict(Material="WSe2", isSOCEnabled=True)),
    ("Haldane2L", 'Haldane2L = { t1 = 0.075; t2 = 0.025; phi0 = 0.9; a0 = 1.42; t11 = 0.01; polytype = "2H"; '
                  'Mt2 = 1.0; t1_bottom = 0.07; t2_bottom = 0.02; Mt2_bottom = 2.0; };',
     dict(t1=0.075, t2=0.025, phi0=0.9, a0=1.42, t11=0.01, polytype="2H", Mt2=1.0, t1_bottom=0.07, t2_bottom=0.02,
          Mt2_bottom=2.0)),
    ("MultiWSM", "MultiWSM = { n = 2.0; v = 0.05; alphan = 0.3; phi0 = 0.0; a0 = 3.0; };",
     dict(n=2.0, v=0.05, alphan=0.3, phi0=0.0, a0=3.0)),
    ("WDSM_TaAs", "WDSM_TaAs = { t = 0.02; delta = 0.5; my = 0.8; mz = 0.9; a0 = 3.4; };",
     dict(t=0.02, delta=0.5, my=0.8, mz=0.9, a0=3.4)),
])
def test_analytic_targets(target, group, keys):
    """The analytic classes' config groups (their constructors' keys and
    Angstrom conversions) give the same model as the Python mirror."""
    from antelope_amd import AnalyticModel
    s = parse_config(ANALYTIC_CFG % (target, "Velocity", group))
    assert s.config.gauge == A.SBE_GAUGE_VELOCITY_HAMILTONIAN
    got = s.solver_model()
    want = AnalyticModel.from_reference(target, **keys)
    assert got.name == want.name and got.nband == want.nband and got.nval == want.nval
    np.testing.assert_allclose(got.params, want.params, rtol=1e-15)
    with pytest.raises(SBEError) as e:
        parse_config(ANALYTIC_CFG % (target, "Wannier", f"{target} = {{ }};"))
    assert "missing" in str(e.value) or "Unknowm" in str(e.value)


def test_bi2se3_target_and_haldane2l_bottom_keys():
    """Bi2Se3surf reads no group (its parameters are hard-coded); Haldane2L
    reads only the first present of t1_bottom / t2_bottom / phi0_bottom, as
    the short-circuit test of Haldane2L.h:83-88 does."""
    from antelope_amd import AnalyticModel
    s = parse_config(ANALYTIC_CFG % ("Bi2Se3surf", "Wannier", ""))
    assert s.solver_model().name == "Bi2Se3surf" and s.model.nband == 2
    g = ('Haldane2L = { t1 = 0.075; t2 = 0.025; phi0 = 0.9; a0 = 1.42; t11 = 0.01; polytype = "AB"; M0 = 0.02; '
         't1_bottom = 0.07; t2_bottom = 0.02; };')
    p = parse_config(ANALYTIC_CFG % ("Haldane2L", "Wannier", g)).solver_model().params
    assert p[4] == 0.07 and p[5] == 0.025 and p[9] == 0.0   # t2_bottom skipped



def test_snapshot_files(tmp_path):
    """sbe_write_snapshot (antelope.cpp:340-421): one row of the 1-D integrated
    conduction occupation per snapshot (sum over the other k indices of
    w Re rho_H[Nval][Nval], / N2, "%16e    "), Ntotal doubles of Re
    rho_H[Nval][Nval] and Ntotal complex rho_H[Nval][Nval-1] per snapshot."""
    from antelope_amd import AnalyticModel
    from antelope_amd.frontend import snapshot_interval, write_snapshot
    from oracle import pyoracle
    m = AnalyticModel.from_reference("KaneMele", t1=0.075, t2=0.025, tR=0.01, Mt2=1.5, phi0=0.9, a0=1.42)
    from tests.cases import small_config
    cfg = small_config(n=(4, 3), ncycles=2.0)
    assert snapshot_interval(cfg, 4) == int(2 * math.pi / 0.014 / 4 / 0.25) and snapshot_interval(cfg, 0) == 0
    rng = np.random.default_rng(1)
    states = [rng.normal(size=(12, 4, 4)) + 1j * rng.normal(size=(12, 4, 4)) for _ in range(2)]
    for i, st in enumerate(states):
        write_snapshot(tmp_path, cfg, m, st, append=i > 0)
    k, w, dV, tw = pyoracle.kgrid(cfg, m)
    occ = np.loadtxt(tmp_path / "occupation_1d_integration.dat")
    assert occ.shape == (2, 4)
    nc = np.fromfile(tmp_path / "fisrt_conduction.dat").reshape(2, 12)
    pi = np.fromfile(tmp_path / "fisrt_coherence.dat").reshape(2, 12, 2)
    for i, st in enumerate(states):
        want = np.array([np.sum((w * st[:, 2, 2].real)[j * 3:(j + 1) * 3]) for j in range(4)]) / 3
        np.testing.assert_allclose(occ[i], want, rtol=1e-6)      # %16e keeps 7 digits
        np.testing.assert_array_equal(nc[i], st[:, 2, 2].real)
        np.testing.assert_array_equal(pi[i, :, 0] + 1j * pi[i, :, 1], st[:, 2, 1])
    assert (tmp_path / "snapshot_log.dat").read_text() == ""
