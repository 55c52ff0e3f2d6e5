/*
 * sbe_frontend.h - input and output front end around sbe_c.h: the parts of
 * the reference's driver that sit either side of the hot path.
 *
 *   sbe_spec_parse_file / _string  read a libconfig-format inputParam.cfg
 *       the way antelope.cpp:90-138, SBEs.h:125-215, SBEs.h:432-561 and
 *       Materials/Haldane.h:46-84 read it (defaults, degree and Angstrom
 *       conversions, required keys) into sbe_config + sbe_model; a
 *       Wannier90 target reads its _tb.dat file (Wannier90.h:107-218).
 *   sbe_write_outputs  writes the reference's files (antelope.cpp:208-235,
 *       514-548): outlaserdata.dat, laserParameters.dat (empty, as the
 *       reference leaves it), interband_dipole_full_evol.dat,
 *       intraband_current_full_evol.dat, occupation__full__evol.dat,
 *       coherence__full__evol.dat, plus band_population__full__evol.dat.
 *   sbe_hhg_spectrum / sbe_write_hhg  the HHG spectrum of the currents, as
 *       AnalysisReader_PyDir.py:163-167,249-382 computes it.
 * Host-only: none of these touch the GPU.
 */
#ifndef ANTELOPE_SBE_FRONTEND_H
#define ANTELOPE_SBE_FRONTEND_H

#include "sbe_c.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct sbe_spec sbe_spec;

/* Parse inputParam.cfg text / file. Targets: Haldane, WilsonMass, SquareSBand,
 * KaneMele, XYHlattice, TMDCs (the classes' own keys and unit conversions)
 * and Wannier90 groups with dataName. Extensions beyond the reference's keys:
 * calc.grid = "mp", calc.precision = "fp32", calc.method = "cash_karp",
 * calc.rtol/atol, laser.efield = "consistent", env_name = "sin2". */
int sbe_spec_parse_string(const char *text, sbe_spec **out);
int sbe_spec_parse_file(const char *path, sbe_spec **out);
/* Pointers into the spec (valid until sbe_spec_free). */
int sbe_spec_get(const sbe_spec *spec, const sbe_config **cfg, const sbe_model **model);
const char *sbe_spec_warnings(const sbe_spec *spec);
void sbe_spec_free(sbe_spec *spec);
const char *sbe_frontend_last_error(void);

/* The .atl format of src/inputFile.atl (which the reference ships but does
 * not read): BEGIN/END sections of `name = expression [unit]` lines.
 * GLOBAL_VARIABLES are arithmetic expressions (units au, rad, ang, bohr,
 * nm, microm, fs, as, eV, deg); SYSTEM E(k).model = MODEL_HALDANE_BANDS(la0,
 * t1, t2, phi0, M0) and U_f(t).model = LASER_SIN2 | LASER_COS4 |
 * LASER_GAUSS(E0, omega, ncycles [, ellip [, cep [, t0]]]); SETUP tgrid.step,
 * kgrid.size (or kgrid.n1/n2/n3), T1, T2, InterIntra, offset.before/after;
 * SOLVER tsolver = RUNGE_KUTTA_4() | RUNGE_KUTTA_5() | CASH_KARP(rtol, atol).
 * Directives of solvers this library does not have (LASER_TOPHAT_SIN, ksolver
 * = CRANK_NICOLSON(), kbox.*) return SBE_ERR_UNSUPPORTED naming them. */
int sbe_spec_parse_atl_string(const char *text, sbe_spec **out);
int sbe_spec_parse_atl_file(const char *path, sbe_spec **out);
const char *sbe_atl_last_error(void);

/* Wannier90 seedname_tb.dat (Materials/Wannier90.h:107-218): lattice,
 * degeneracies, <0m|H|Rn> and <0m|r|Rn>, converted to Hartree and bohr; the
 * position elements are read as re + i*im (the reference has re * i*im,
 * Wannier90.h:173). A file without the position part gives
 * is_dipole_zero = 1. sbe_wannier_tb_model points a model into the reader's
 * arrays (valid until sbe_wannier_tb_free). In an inputParam.cfg, a target
 * group with `dataName` (path; relative to the config file's directory when
 * parsed from a file) and `Nval`, or `FermiE` in eV (Fermi-Dirac initial
 * state, Nval ignored), selects such a model. sbe_wannier_write_tb writes a
 * Wannier model in the same format (eV, Angstrom). */
typedef struct sbe_wannier_tb sbe_wannier_tb;
int sbe_wannier_read_tb(const char *path, sbe_wannier_tb **out);
int sbe_wannier_tb_model(const sbe_wannier_tb *tb, int32_t nval, sbe_model *m);
void sbe_wannier_tb_free(sbe_wannier_tb *tb);
int sbe_wannier_write_tb(const char *path, const sbe_model *m, const char *header);
const char *sbe_wannier_last_error(void);

/* Fill a TMDC model (SBE_MODEL_TMDC, sbe_c.h) for material "MoS2", "WS2",
 * "MoSe2", "WSe2", "MoTe2" or "WTe2" from the GGA parameter table the
 * reference's TMDCs class hard-codes (TMDCs.h:48-149): lattice constant,
 * spin-orbit coupling and the 19 hoppings, converted to bohr and Hartree;
 * soc != 0 gives the 6-band model. SBE_ERR_MODEL for an unknown material. */
int sbe_tmdc_params(const char *material, int32_t soc, sbe_model *m);

/* Snapshots (antelope.cpp:267-292,340-421). calc.shotNumber of the config;
 * the step interval int(2 pi / w0 / shotNumber / dt) of pulse 0 (0: none).
 * sbe_write_snapshot appends one snapshot of the whole grid from the
 * band-basis state (sbe_get_state(ctx, SBE_BASIS_BAND, ...) of a context
 * holding all k points): a row of occupation_1d_integration.dat (per first k
 * index, sum w Re rho_H[Nval][Nval] / N2, "%16e    "), Ntotal doubles
 * Re rho_H[Nval][Nval] to fisrt_conduction.dat and Ntotal complex
 * rho_H[Nval][Nval-1] to fisrt_coherence.dat (the reference's file names);
 * snapshot_log.dat is created empty as the reference leaves it. append = 0
 * truncates the files. The reference fills the per-k arrays with a stale
 * index (one element, or the last k point's values everywhere); these are
 * the values it intends. */
int32_t sbe_spec_shot_number(const sbe_spec *spec);
int64_t sbe_snapshot_interval(const sbe_config *cfg, int32_t shot_number);
int sbe_write_snapshot(const char *dir, const sbe_config *cfg, const sbe_model *model, const double *rho_band,
                       int32_t append);

/* HHG spectrum of a current series, as the reference's post-processing
 * computes it (src/GeneralModel/AlsisPythonShort/AnalysisReader_PyDir.py:
 * 163-167, 249-382, which does it in numpy): with w = linspace(-pi/dt,
 * pi/dt - dw, nt), dw = 2 pi / (nt dt), and a Blackman window over the whole
 * series, F_a(w) = -i w fftshift(fft(window j_a) dt) for the inter and intra
 * series separately, their sum per axis a < dim, and S = sum_a |F_a|^2 +
 * 0.98e-25. j_inter, j_intra are [nt][3] as written to the .dat files (x dV).
 * Outputs cover the script's half w[nt/2 : nt-1], sbe_hhg_points(nt) =
 * nt - nt/2 - 1 points; each may be NULL:
 *   order [n]          w / w0 (harmonic order)
 *   F     [n][dim][2]  inter + intra F_a (re, im)
 *   S, S_inter, S_intra [n]  total, inter-only and intra-only spectra
 * sbe_write_hhg writes the script's HHG__*.dat layout to `path`: per point
 * w/w0, Re F_a, Im F_a for a < dim, then S, in %1.16e. Host-only. */
int64_t sbe_hhg_points(int64_t nt);
int sbe_hhg_spectrum(int64_t nt, const double *j_inter, const double *j_intra, double dt, double w0, int32_t dim,
                     double *order, double *F, double *S, double *S_inter, double *S_intra);
int sbe_write_hhg(const char *path, int64_t nt, const double *j_inter, const double *j_intra, double dt, double w0,
                  int32_t dim);
const char *sbe_spectrum_last_error(void);

/* Write the reference's output files into `dir` from summed series (the
 * layout of sbe_get_series; several ranks' series added together). */
int sbe_write_outputs(const char *dir, const sbe_config *cfg, const sbe_model *model, const sbe_info *info,
                      const double *j_inter, const double *j_intra, const double *rho_bar,
                      const double *band_pop);

#ifdef __cplusplus
}
#endif

#endif /* ANTELOPE_SBE_FRONTEND_H */
