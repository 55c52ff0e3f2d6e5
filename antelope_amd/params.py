"""Run configuration and material models, mirroring the reference's inputs.

`RunConfig` holds the `calc` and `laser` groups of inputParam.cfg
(SBEs.h:432-561); `Pulse` the arguments of Pulse::Pulse (laser.h:68-141);
`HaldaneModel` the `Haldane` group (Materials/Haldane.h:46-84); `WannierModel`
the tight-binding tables of Materials/Wannier90.h:107-218 (read from a
file there, given as arrays here). All values are in atomic units and
radians; `HaldaneModel.from_reference` and `Pulse.from_reference` take the
reference's Angstrom / degree inputs.
"""
from __future__ import annotations

import ctypes as C
import math
from dataclasses import dataclass, field
from typing import Optional, Sequence

import numpy as np

from . import _abi as A

AU_ANGSTROM = 0.529177210903  # constant.h:34
AU_EV = 27.211386245988       # constant.h:35


@dataclass
class Pulse:
    E0: float
    w0: float
    ncycles: float
    ellip: float = 0.0
    cep: float = 0.0            # rad
    t0: float = 0.0
    phix: float = 0.0           # rad
    envelope: str = "sin2"
    ncycles_clipped: float = 0.0
    thetaz: float = 0.0         # rad
    phiz: float = 0.0           # rad

    @classmethod
    def from_reference(cls, E0, w0, ncycles, ellip, cep_deg, t0, phix_deg, env_name,
                       thetaz_deg=0.0, phiz_deg=0.0, ncycles_clipped=0.0) -> "Pulse":
        """Degrees -> radians as SBEs.h:531-537."""
        d = math.pi / 180.0
        return cls(E0=E0, w0=w0, ncycles=ncycles, ellip=ellip, cep=cep_deg * d, t0=t0,
                   phix=phix_deg * d, envelope=env_name, ncycles_clipped=ncycles_clipped,
                   thetaz=thetaz_deg * d, phiz=phiz_deg * d)

    def to_c(self) -> A.sbe_pulse:
        if self.envelope not in A.ENVELOPES:
            raise ValueError(f"invalid envelope type {self.envelope!r}")
        return A.sbe_pulse(self.E0, self.w0, self.ellip, self.ncycles, self.cep, self.t0,
                           self.phix, self.ncycles_clipped, self.thetaz, self.phiz,
                           A.ENVELOPES[self.envelope], 0)


_RK = {4: A.SBE_RK4, 5: A.SBE_RK5_FEHLBERG, 45: A.SBE_RK_CASH_KARP, "cash_karp": A.SBE_RK_CASH_KARP}


@dataclass
class RunConfig:
    npoints: Sequence[int]
    pulses: Sequence[Pulse]
    dt: float = 0.1               # SBEs.h:455
    T1: float = -1.0
    T2: float = -1.0
    rk_order: object = 5          # SBEs.h:438
    inter_intra: bool = False     # SBEs.h:442
    ksfactor: Sequence[float] = (1.0, 1.0, 1.0)
    grid: str = "reference"       # or "mp"
    bz_axes: Optional[Sequence[float]] = None
    bz_origin: Optional[Sequence[float]] = None
    init: str = "uniform"         # or "fermi"
    fermi_energy: float = 0.0
    precision: str = "fp64"       # or "fp32"
    offset: Sequence[float] = (0.0, 0.0)
    efield: str = "reference"     # or "consistent"
    gauge: str = "LengthWannier"
    rtol: float = 1e-8
    atol: float = 1e-12

    def to_c(self):
        """Returns (sbe_config, keepalive)."""
        pulses = (A.sbe_pulse * len(self.pulses))(*[p.to_c() for p in self.pulses])
        npts = list(self.npoints) + [1] * (3 - len(self.npoints))
        gauges = {"LengthWannier": 0, "Wannier": 0, "LengthHamiltonian": 1, "Hamiltonian": 1,
                  "VelocityHamiltonian": 2, "Velocity": 2}
        c = A.sbe_config()
        c.abi_version = A.SBE_ABI_VERSION
        c.gauge = gauges[self.gauge]
        c.npoints = (C.c_int32 * 3)(*[int(x) for x in npts])
        c.rk_order = _RK.get(self.rk_order, self.rk_order if isinstance(self.rk_order, int) else -1)
        c.dt, c.T1, c.T2 = self.dt, self.T1, self.T2
        ks = list(self.ksfactor) + [1.0] * (3 - len(self.ksfactor))
        c.ksfactor = (C.c_double * 3)(*ks)
        c.inter_intra = int(bool(self.inter_intra))
        c.grid = {"reference": A.SBE_GRID_REFERENCE, "mp": A.SBE_GRID_MP}[self.grid]
        if self.bz_axes is not None:
            c.has_bz_axes = 1
            ax = list(self.bz_axes) + [0.0] * (9 - len(self.bz_axes))
            c.bz_axes = (C.c_double * 9)(*ax)
        if self.bz_origin is not None:
            c.has_bz_origin = 1
            org = list(self.bz_origin) + [0.0] * (3 - len(self.bz_origin))
            c.bz_origin = (C.c_double * 3)(*org)
        c.init_type = {"uniform": A.SBE_INIT_UNIFORM_VALENCE, "fermi": A.SBE_INIT_FERMI_DIRAC}[self.init]
        c.precision = {"fp64": A.SBE_PREC_FP64, "fp32": A.SBE_PREC_FP32, "split16": A.SBE_PREC_SPLIT16}[self.precision]
        c.fermi_energy = self.fermi_energy
        c.npulses = len(self.pulses)
        c.efield_mode = {"reference": A.SBE_EFIELD_REFERENCE, "consistent": A.SBE_EFIELD_CONSISTENT}[self.efield]
        c.pulses = C.cast(pulses, C.POINTER(A.sbe_pulse))
        c.offset_before, c.offset_after = float(self.offset[0]), float(self.offset[1])
        c.rtol, c.atol = self.rtol, self.atol
        return c, [pulses]


@dataclass
class HaldaneModel:
    a0: float          # bohr
    t1: float
    t2: float
    phi0: float        # rad
    M0: float
    nval: int = 1      # Haldane.h:48

    @classmethod
    def from_reference(cls, a0_angstrom, t1, t2, phi0, M0=None, Mt2=None) -> "HaldaneModel":
        """Haldane.h:49-67: a0 in Angstrom, M0 given directly or as Mt2*t2."""
        if M0 is None:
            if Mt2 is None:
                raise ValueError("M0 param missing")
            M0 = Mt2 * t2
        return cls(a0=a0_angstrom / AU_ANGSTROM, t1=t1, t2=t2, phi0=phi0, M0=M0)

    @property
    def nband(self) -> int:
        return 2

    def to_c(self):
        m = A.sbe_model()
        m.kind = A.SBE_MODEL_HALDANE
        m.nband, m.nval, m.is_dipole_zero = 2, self.nval, 1
        m.a0, m.t1, m.t2, m.phi0, m.M0 = self.a0, self.t1, self.t2, self.phi0, self.M0
        return m, []


@dataclass
class WannierModel:
    lattice: np.ndarray      # [3][3] rows a1,a2,a3 (bohr)
    rvec: np.ndarray         # [nr][3] int
    rweight: np.ndarray      # [nr] 1/degeneracy
    ham: np.ndarray          # [nr][N][N] complex (Hartree)
    pos: Optional[np.ndarray] = None   # [nr][3][N][N] complex (bohr)
    nval: int = 1
    is_dipole_zero: bool = False

    @property
    def nband(self) -> int:
        return int(self.ham.shape[1])

    def to_c(self):
        nr, N = self.ham.shape[0], self.ham.shape[1]
        rv = np.ascontiguousarray(self.rvec, dtype=np.int32)
        rw = np.ascontiguousarray(self.rweight, dtype=np.float64)
        hm = np.ascontiguousarray(np.stack([self.ham.real, self.ham.imag], -1), dtype=np.float64)
        keep = [rv, rw, hm]
        m = A.sbe_model()
        m.kind = A.SBE_MODEL_WANNIER
        m.nband, m.nval, m.is_dipole_zero = N, self.nval, int(self.is_dipole_zero)
        m.nrpts = nr
        m.lattice = (C.c_double * 9)(*np.asarray(self.lattice, dtype=np.float64).ravel())
        m.rvec = rv.ctypes.data_as(A.c_int32_p)
        m.rweight = rw.ctypes.data_as(A.c_double_p)
        m.ham = hm.ctypes.data_as(A.c_double_p)
        if self.pos is not None:
            ps = np.ascontiguousarray(np.stack([self.pos.real, self.pos.imag], -1), dtype=np.float64)
            keep.append(ps)
            m.pos = ps.ctypes.data_as(A.c_double_p)
        return m, keep


_ANALYTIC = {  # name -> (kind, bands, valence bands, parameter names)
    "WilsonMass": (A.SBE_MODEL_WILSON_MASS, 2, 1, ("t", "delta", "mu", "a0", "a1")),
    "SquareSBand": (A.SBE_MODEL_SQUARE_SBAND, 2, 1, ("t1", "M0", "a0")),
    "KaneMele": (A.SBE_MODEL_KANE_MELE, 4, 2, ("t1", "t2", "tR", "M0", "phi0", "a0")),
    "XYHlattice": (A.SBE_MODEL_XYH_LATTICE, 4, 2, ("tl", "tr", "lambda", "a0")),
    "TMDCs": (A.SBE_MODEL_TMDC, 3, 1, ("a0", "soc", "e1", "e2", "t0", "t1", "t2", "t11", "t12", "t22", "r0", "r1",
                                       "r2", "r11", "r12", "u0", "u1", "u2", "u11", "u12", "u22")),
    "Haldane2L": (A.SBE_MODEL_HALDANE_2L, 4, 2, ("t1", "t2", "M0", "phi0", "t1_b", "t2_b", "M0_b", "phi0_b", "a0",
                                                 "same", "s0", "s1", "h0", "h1", "h2", "h3")),
    "Bi2Se3surf": (A.SBE_MODEL_BI2SE3_SURF, 2, 1, ()),
    "MultiWSM": (A.SBE_MODEL_MULTI_WSM, 2, 1, ("n", "v", "alphan", "phi0", "a0")),
    "WDSM_TaAs": (A.SBE_MODEL_WDSM_TAAS, 4, 1, ("t", "delta", "my", "mz", "a0")),
}

# Haldane2L interlayer blocks and B2 signs per polytype (Haldane2L.h:107-147)
_H2L_POLY = {"AA": ((1, 0, 0, 1), (1, 1)), "AB": ((0, 1, 0, 0), (1, 1)), "AB1": ((0, 1, 0, 0), (1, 1)),
             "AB2": ((0, 0, 1, 0), (1, 1)), "3R": ((0, 0, 1, 0), (1, 1)), "AA'": ((0, 1, 1, 0), (1, -1)),
             "2H": ((0, 1, 1, 0), (1, -1)), "A'B": ((0, 0, 0, 1), (-1, 1)), "AB'": ((1, 0, 0, 0), (1, -1))}


@dataclass
class AnalyticModel:
    """One of the analytic model classes of Materials/ (WilsonMass,
    SquareSBand, KaneMele, XYHlattice, TMDCs, Haldane2L, Bi2Se3surf, MultiWSM,
    WDSM_TaAs) with its parameters in atomic
    units (lengths in bohr; the `from_reference` constructors take the
    classes' config keys and convert Angstrom where the class does)."""
    name: str
    params: Sequence[float]
    nband: int = 0
    nval: int = -1

    def __post_init__(self):
        kind, nb, nv, names = _ANALYTIC[self.name]
        if self.nband == 0:
            self.nband = nb
        if self.nval < 0:
            self.nval = nv if self.nband == nb else 2

    @property
    def kind(self) -> int:
        return _ANALYTIC[self.name][0]

    @classmethod
    def from_reference(cls, name: str, **keys) -> "AnalyticModel":
        """Parameters as the class constructor reads them from its config group."""
        if name == "WilsonMass":  # WilsonMass.h:40-63 (a.u.)
            return cls(name, [keys["t"], keys["delta"], keys["mu"], keys["a0"], keys.get("a1", keys["a0"])])
        if name == "SquareSBand":  # SquareSBand.h:39-59
            return cls(name, [keys["t1"], keys["M0"], keys["a0"] / AU_ANGSTROM])
        if name == "KaneMele":  # KaneMele.h:47-75: M0 = Mt2 t2
            return cls(name, [keys["t1"], keys["t2"], keys["tR"], keys["Mt2"] * keys["t2"], keys["phi0"],
                              keys["a0"] / AU_ANGSTROM])
        if name == "XYHlattice":  # XYHlattice.h:49-73
            return cls(name, [keys["tl"], keys["tr"], keys["lambda"], keys["a0"] / AU_ANGSTROM])
        if name == "TMDCs":  # TMDCs.h:48-149: parameter table by material, 6 bands with spin-orbit
            return tmdc(keys["Material"], bool(keys.get("isSOCEnabled", False)))
        if name == "Haldane2L":  # Haldane2L.h:52-148
            t1, t2, phi = keys["t1"], keys["t2"], keys["phi0"]
            M0 = keys["M0"] if "M0" in keys else keys["Mt2"] * t2
            b = [t1, t2, M0, phi]
            same = True
            for i, key in ((0, "t1_bottom"), (1, "t2_bottom"), (3, "phi0_bottom")):  # first present key only
                if key in keys:
                    b[i], same = keys[key], False
                    break
            if "M0_bottom" in keys:
                b[2], same = keys["M0_bottom"], False
            elif "Mt2_bottom" in keys:
                b[2], same = keys["Mt2_bottom"] * b[1], False
            h, sg = _H2L_POLY[keys["polytype"]]
            t11 = keys["t11"]
            return cls(name, [t1, t2, M0, phi, *b, keys["a0"] / AU_ANGSTROM, 1.0 if same else 0.0, *sg,
                              *[x * t11 for x in h]])
        if name == "Bi2Se3surf":  # Bi2Se3surf.h:51-82: hard-coded
            return cls(name, [])
        if name == "MultiWSM":  # MultiWSM.h:47-74
            return cls(name, [keys["n"], keys["v"], keys["alphan"], keys["phi0"], keys["a0"] / AU_ANGSTROM])
        if name == "WDSM_TaAs":  # WDSM_TaAs.h:41-67
            return cls(name, [keys["t"], keys["delta"], keys["my"], keys["mz"], keys["a0"] / AU_ANGSTROM])
        raise KeyError(name)

    def to_c(self):
        m = A.sbe_model()
        m.kind = self.kind
        m.nband, m.nval, m.is_dipole_zero = self.nband, self.nval, 1
        p = list(self.params) + [0.0] * (A.SBE_MODEL_NPARAMS - len(self.params))
        m.params = (C.c_double * A.SBE_MODEL_NPARAMS)(*p)
        return m, []


def tmdc(material: str, soc: bool = False) -> AnalyticModel:
    """TMDC three-band model of `material` (MoS2, WS2, MoSe2, WSe2, MoTe2,
    WTe2) from the library's parameter table (sbe_tmdc_params)."""
    from .solver import load_library
    m = A.sbe_model()
    rc = load_library().sbe_tmdc_params(material.encode(), int(soc), C.byref(m))
    if rc != 0:
        raise ValueError(f"unknown TMDC material {material!r}")
    return AnalyticModel("TMDCs", list(m.params)[:21], nband=m.nband, nval=m.nval)

