/*
 * sbe_oracle.h - FP64 CPU restatement of Antelope's SBE time propagation.
 *
 * TEST INFRASTRUCTURE ONLY. This is the checker that the HIP path is
 * compared against; it is never linked into, loaded by or called from the
 * product library. Only tests/, __graft_entry__.smoke() and bench.py's CPU
 * baseline leg use it.
 *
 * It follows the reference's own per-k structure ("faithful cost"): generic
 * N^3 complex matrix products, a Hermitian eigensolve for every RK stage and
 * every current sample, and per-k laser evaluation. Every function cites the
 * reference file:line it restates (paths relative to src/GeneralModel/ of
 * the reference unless given in full).
 */
#ifndef ANTELOPE_SBE_ORACLE_H
#define ANTELOPE_SBE_ORACLE_H

#include <stdint.h>
#include "../include/antelope/sbe_c.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct orc_info {
    int64_t nt;
    int64_t nk_total;
    int64_t nk_local;
    double atmin, atmax, dt, dV, total_weight;
    double seconds;   /* wall time of the time loop only */
} orc_info;

/* Laser A(t), E(t) at n times; also returns atmin, Nt (laser.h:278-322). */
int orc_laser_eval(const sbe_config *cfg, int64_t n, const double *t,
                   double *A, double *E, double *atmin, double *atmax,
                   int64_t *nt);

/* k-grid of block [k_begin,k_end): coordinates [nk][3], weights [nk],
 * dV and the block's total weight (momaxis.h:147-240,349-383 + the BZ
 * selection of SBEs.h:219-244). */
int orc_kgrid(const sbe_config *cfg, const sbe_model *model, int64_t k_begin,
              int64_t k_end, double *kpts, double *w, double *dV,
              double *total_weight);

/* Model matrices at one k (row-major [N][N][re,im]). */
int orc_hamiltonian(const sbe_model *m, const double k[3], double *H);
int orc_jmatrix(const sbe_model *m, const double k[3], double *J /*[3][N][N][2]*/);
int orc_dipole(const sbe_model *m, const double k[3], double *D /*[3][N][N][2]*/);

/* Hermitian eigensolver (cyclic complex Jacobi): ascending eigenvalues,
 * U[i][j] = component i of eigenvector j (the zheevr row-major convention
 * of SBEs.h:874-898). */
int orc_eigh(int n, const double *H, double *evals, double *U);

/* Right-hand side of the SBE at one k: LengthWannier (SBEs.h:762-861) or
 * VelocityHamiltonian (SBEs.h:787-805,814-829,862-871, with the per-k
 * constants of orc_velocity_setup; rho is then band-basis). */
int orc_rhs(const sbe_config *cfg, const sbe_model *model, const double k[3],
            double t, const double *rho, double *out);

/* VelocityHamiltonian per-k constants (SBEs.h:284-373): E[N] ascending,
 * P[3][N][N][2] = U^dag J_a(k) U and rho0[N][N][2] = diag(occupations),
 * with the eigenvector phases fixed as sbe_oracle.c's phase_fix states. */
int orc_velocity_setup(const sbe_config *cfg, const sbe_model *model, const double k[3],
                       double *E, double *P, double *rho0);

/* Full propagation of block [k_begin,k_end) over steps [0, nsteps)
 * (nsteps < 0: all Nt), antelope.cpp:297-338. Series as sbe_get_series;
 * rho_final [nk][N][N][2] may be NULL. nthreads >= 1 splits the block into
 * contiguous ParaRange sub-blocks (utility.h:104-115); partial sums are added
 * in sub-block order. */
int orc_run(const sbe_config *cfg, const sbe_model *model, int64_t k_begin,
            int64_t k_end, int64_t nsteps, int nthreads, double *j_inter,
            double *j_intra, double *rho_bar, double *band_pop,
            double *rho_final, orc_info *info);

/* Adaptive Cash-Karp 5(4) restatement (see sbe_oracle.c): blocks of `block`
 * consecutive k points (<= 0: 64) choose their own steps; windows of
 * `window` samples (<= 0: 2048, as the product) end on their sample time.
 * 2-band models, T1 off. accepted / rejected [nblocks] may be NULL. */
int orc_run_adaptive(const sbe_config *cfg, const sbe_model *model, int64_t k_begin,
                     int64_t k_end, int64_t block, int64_t window, int nthreads,
                     double *j_inter, double *j_intra, double *rho_bar, double *band_pop,
                     double *rho_final, int64_t *accepted, int64_t *rejected,
                     orc_info *info);
/* The Cash-Karp coefficients the restatement uses. */
int orc_cash_karp(double c[6], double a[6][5], double b[6], double bhat[6]);
/* 2-band observables o[8] = (inter_x, inter_y, intra_x, intra_y, s_x, s_y,
 * s_z, s.dhat) of rho at (k, t) and their time derivatives along drho/dt. */
int orc_observe(const sbe_config *cfg, const sbe_model *model, const double k[3], double t,
                const double *rho, const double *drho, double *o, double *od);
/* Currents and band populations of the generic path (SBEs.h:626-760). */
int orc_currents(const sbe_config *cfg, const sbe_model *model, const double k[3], double t,
                 const double *rho, double *inter, double *intra, double *pop);

/* Last error message of the calling thread. */
const char *orc_last_error(void);

#ifdef __cplusplus
}
#endif
#endif
