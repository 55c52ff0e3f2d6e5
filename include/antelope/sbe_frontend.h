/*
 * sbe_frontend.h - input and output front end around sbe_c.h: the parts of
 * the reference's driver that sit either side of the hot path.
 *
 *   sbe_spec_parse_file / _string  read a libconfig-format inputParam.cfg
 *       the way antelope.cpp:90-138, SBEs.h:125-215, SBEs.h:432-561 and
 *       Materials/Haldane.h:46-84 read it (defaults, degree and Angstrom
 *       conversions, required keys) into sbe_config + sbe_model.
 *   sbe_write_outputs  writes the reference's files (antelope.cpp:208-235,
 *       514-548): outlaserdata.dat, laserParameters.dat (empty, as the
 *       reference leaves it), interband_dipole_full_evol.dat,
 *       intraband_current_full_evol.dat, occupation__full__evol.dat,
 *       coherence__full__evol.dat, plus band_population__full__evol.dat.
 * Host-only: none of these touch the GPU.
 */
#ifndef ANTELOPE_SBE_FRONTEND_H
#define ANTELOPE_SBE_FRONTEND_H

#include "sbe_c.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct sbe_spec sbe_spec;

/* Parse inputParam.cfg text / file. Extensions beyond the reference's keys:
 * calc.grid = "mp", calc.precision = "fp32", calc.method = "cash_karp",
 * calc.rtol/atol, laser.efield = "consistent", env_name = "sin2". */
int sbe_spec_parse_string(const char *text, sbe_spec **out);
int sbe_spec_parse_file(const char *path, sbe_spec **out);
/* Pointers into the spec (valid until sbe_spec_free). */
int sbe_spec_get(const sbe_spec *spec, const sbe_config **cfg, const sbe_model **model);
const char *sbe_spec_warnings(const sbe_spec *spec);
void sbe_spec_free(sbe_spec *spec);
const char *sbe_frontend_last_error(void);

/* Write the reference's output files into `dir` from summed series (the
 * layout of sbe_get_series; several ranks' series added together). */
int sbe_write_outputs(const char *dir, const sbe_config *cfg, const sbe_model *model, const sbe_info *info,
                      const double *j_inter, const double *j_intra, const double *rho_bar,
                      const double *band_pop);

#ifdef __cplusplus
}
#endif

#endif /* ANTELOPE_SBE_FRONTEND_H */
