"""The .atl input format (src/inputFile.atl, read by nothing in the reference):
expressions with units, the Haldane model, the time grid and the RK4 solver map
onto the same structs as inputParam.cfg; directives of solvers this library
lacks are errors that name them."""
import math
from pathlib import Path

import pytest

from antelope_amd import _abi as A
from antelope_amd.frontend import parse_atl, parse_config
from antelope_amd.solver import SBEError

REF_ATL = Path("/root/reference/src/inputFile.atl")
AU_ANG = 0.529177210903

ATL = """
BEGIN GLOBAL_VARIABLES
    pi = acos(-1)
    c  = 137.0359996287515
    lambda = 1.27 microm      # wavelength
    T      = lambda/c
    omega  = 2.0*pi/T
    E0     = 0.0045 au
    la0    = 1.41978 ang
    t1     = 0.075 au
    t2     = t1/3.0
    phi0   = 0.3*pi rad
    M0     = 2.54*t2
END GLOBAL_VARIABLES
BEGIN SIMULATION_SETUP
    kgrid.size = 60
    tgrid.min  = 0.0
    tgrid.max  = 12*T
    tgrid.step = 0.25
    T2 = 220
    InterIntra = TRUE
END SIMULATION_SETUP
BEGIN PROPERTIES
    energy = TRUE
END PROPERTIES
BEGIN SYSTEM
    E(k).model   = MODEL_HALDANE_BANDS( la0, t1, t2, phi0, M0 )
    U_f(t).model = LASER_SIN2( E0, omega, 10, 0.5, 0.0 )
END SYSTEM
BEGIN SOLVER
    tsolver = RUNGE_KUTTA_4()
END SOLVER
"""


def test_units_and_expressions():
    s = parse_atl(ATL)
    p = s.config.pulses[0]
    # photon energy of a 1.27 micron wave: hc / lambda = 1239.84198 eV nm / 1270 nm
    assert p.w0 == pytest.approx(1239.841984 / 1270.0 / 27.211386245988, rel=1e-8)
    assert (p.E0, p.ncycles, p.ellip, p.envelope) == (0.0045, 10.0, 0.5, A.SBE_ENV_SIN2)
    m = s.model
    assert m.kind == A.SBE_MODEL_HALDANE and m.a0 == pytest.approx(1.41978 / AU_ANG, rel=1e-15)
    assert m.t2 == pytest.approx(0.025) and m.phi0 == pytest.approx(0.3 * math.pi) and m.M0 == pytest.approx(2.54 * 0.025)
    c = s.config
    assert list(c.npoints) == [60, 60, 1] and c.dt == 0.25 and c.rk_order == A.SBE_RK4
    assert c.T2 == 220.0 and c.inter_intra == 1 and "PROPERTIES.energy" in s.warnings


def test_same_structs_as_inputparam_cfg():
    a = parse_atl(ATL)
    cfg = f"""
    target = "Haldane";
    laser = {{ pulses = ( {{ E0 = 0.0045; w0 = {a.config.pulses[0].w0!r}; ncycles = 10; ellip = 0.5; cep = 0.0;
                             t0 = 0.0; phix = 0.0; env_name = "sin2"; }} ); }};
    calc = {{ Npoints = [60, 60]; dt = 0.25; T2 = 220; RKorder = 4; InterIntra = true; ksfactor = [1.0, 1.0, 1.0]; }};
    Haldane = {{ a0 = 1.41978; t1 = 0.075; t2 = {0.075 / 3.0!r}; phi0 = {0.3 * math.pi!r}; M0 = {2.54 * 0.025 / 1.0!r}; }};
    """
    b = parse_config(cfg)
    for f, _ in A.sbe_config._fields_:
        if f in ("pulses", "gauge", "ksfactor", "rtol", "atol"):
            continue
        va, vb = getattr(a.config, f), getattr(b.config, f)
        assert (list(va) == list(vb)) if hasattr(va, "__len__") else va == pytest.approx(vb, rel=1e-14), f
    for f in ("a0", "t1", "t2", "phi0", "M0", "nband", "nval", "kind"):
        assert getattr(a.model, f) == pytest.approx(getattr(b.model, f), rel=1e-14), f


@pytest.mark.parametrize("edit,code,word", [
    (("U_f(t).model = LASER_SIN2( E0, omega, 10, 0.5, 0.0 )", "U_f(t).model = LASER_TOPHAT_SIN( E0, 1, 2, 3, omega )"),
     A.SBE_ERR_UNSUPPORTED, "LASER_TOPHAT_SIN"),
    (("tsolver = RUNGE_KUTTA_4()", "ksolver = CRANK_NICOLSON()"), A.SBE_ERR_UNSUPPORTED, "CRANK_NICOLSON"),
    (("kgrid.size = 60", "kbox.size = 101"), A.SBE_ERR_UNSUPPORTED, "kbox"),
    (("t2     = t1/3.0", "t2     = t1/3.0 parsec"), A.SBE_ERR_CONFIG, "parsec"),
    (("t2     = t1/3.0", "t2     = t1/q"), A.SBE_ERR_CONFIG, "undefined variable q"),
    (("END SOLVER", ""), A.SBE_ERR_CONFIG, "missing END SOLVER"),
])
def test_errors_name_the_directive(edit, code, word):
    with pytest.raises(SBEError) as e:
        parse_atl(ATL.replace(*edit))
    assert e.value.code == code and word in str(e.value), str(e.value)


@pytest.mark.skipif(not REF_ATL.exists(), reason="reference checkout not present")
def test_reference_inputfile_atl():
    """The reference's own file: its variables evaluate; its flat-top laser, k box and
    Crank-Nicolson k solver are reported as unsupported, first the k box (file order)."""
    text = REF_ATL.read_text()
    with pytest.raises(SBEError) as e:
        parse_atl(text)
    assert e.value.code == A.SBE_ERR_UNSUPPORTED and "kbox" in str(e.value)
    lines = [l for l in text.splitlines() if "kbox" not in l]
    text2 = "\n".join(lines).replace("tgrid.min = 0.0", "kgrid.size = 101\n\ttgrid.min = 0.0")
    with pytest.raises(SBEError) as e:
        parse_atl(text2)
    assert "LASER_TOPHAT_SIN" in str(e.value)
    text3 = text2.replace("LASER_TOPHAT_SIN( E0, tc, w, dw, omega )", "LASER_SIN2( E0, omega, 6 )")
    with pytest.raises(SBEError) as e:
        parse_atl(text3)
    assert "CRANK_NICOLSON" in str(e.value)
    s = parse_atl("\n".join(l for l in text3.splitlines() if "ksolver" not in l))
    m, c = s.model, s.config
    assert m.a0 == pytest.approx(1.0 / AU_ANG) and m.t1 == 0.075 and m.t2 == pytest.approx(0.025)
    assert m.phi0 == 0.1 and m.M0 == pytest.approx(0.025)
    assert c.dt == 0.1 and c.rk_order == A.SBE_RK4 and list(c.npoints) == [101, 101, 1]
    # E0 = 0.024 au at lambda = 1.27 micron
    assert c.pulses[0].E0 == 0.024 and c.pulses[0].w0 == pytest.approx(0.0358764, rel=1e-5)
