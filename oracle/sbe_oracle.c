/*
 * sbe_oracle.c - FP64 CPU restatement of Antelope's SBE propagation.
 *
 * TEST INFRASTRUCTURE ONLY (the checker, never the product). See
 * sbe_oracle.h. Reference paths are relative to src/GeneralModel/.
 *
 * Pinning (what checks each part independently of this restatement) is
 * listed per function; the tests live in tests/test_oracle_*.py.
 */
#define _GNU_SOURCE
#include "sbe_oracle.h"

#include <complex.h>
#include <math.h>
#include <pthread.h>
#include <stdarg.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <time.h>

typedef double complex cplx;

static __thread char g_err[512];

static int fail(int code, const char *fmt, ...)
{
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(g_err, sizeof g_err, fmt, ap);
    va_end(ap);
    return code;
}

const char *orc_last_error(void) { return g_err; }

/* constant.h:26-34 */
static const double PI = 3.14159265358979323846;

/* ------------------------------------------------------------------ laser */
/* Pulse derived quantities, laser.h:68-141 (+ the sin^2 envelope). */
typedef struct {
    double E0x, E0y, A0x, A0y, w0, cep, t0;
    double twidth, twidth_clipped, envfactor, period0;
    double xdir[3], ydir[3];
    int env;
} opulse;

typedef struct {
    int np;
    opulse *p;
    double atmin, atmax, dt;
    double A0[3], E0[3]; /* values at atmin, subtracted (laser.h:294-295) */
    int64_t nt;
    int consistent;
} olaser;

/* Pulse::f_envelope, laser.h:150-206; SBE_ENV_SIN2 = cos^2(t*w0/(2 ncycles))
 * with twidth = ncycles*T/2 (the cos-n convention of laser.h:102-104). */
static double f_env(const opulse *p, double t, int dtflag)
{
    double ef = p->envfactor, to;
    switch (p->env) {
    case SBE_ENV_COS4:
        if (fabs(t) > p->twidth) return 0.0;
        if (dtflag) return -4.0 * ef * sin(t * ef) * pow(cos(t * ef), 3.0);
        return pow(cos(t * ef), 4.0);
    case SBE_ENV_SIN2:
        if (fabs(t) > p->twidth) return 0.0;
        if (dtflag) return -2.0 * ef * sin(t * ef) * cos(t * ef);
        return cos(t * ef) * cos(t * ef);
    case SBE_ENV_COS4_CLIPPED:
        if (fabs(t) > p->twidth) return 0.0;
        if (t > p->twidth_clipped) to = t - p->twidth_clipped;
        else if (t > -p->twidth_clipped) return dtflag ? 0.0 : 1.0;
        else to = t + p->twidth_clipped;
        if (dtflag) return -4.0 * ef * sin(to * ef) * pow(cos(to * ef), 3.0);
        return pow(cos(to * ef), 4.0);
    case SBE_ENV_GAUSS:
        if (dtflag) return 2.0 * t * ef * exp(ef * t * t);
        return exp(ef * t * t);
    case SBE_ENV_GAUSS_CLIPPED:
        if (t > p->twidth_clipped) to = t - p->twidth_clipped;
        else if (t > -p->twidth_clipped) return dtflag ? 0.0 : 1.0;
        else to = t + p->twidth_clipped;
        if (dtflag) return 2.0 * to * ef * exp(ef * to * to);
        return exp(ef * to * to);
    }
    return 0.0;
}

/* Pulse::avlaser_2d/elaser_2d + avlaser/elaser, laser.h:209-238 */
static void pulse_eval(const opulse *p, double t, int consistent, double A[3],
                       double E[3])
{
    double kt = t - p->t0;
    double ph = kt * p->w0 - p->cep;
    double f = f_env(p, kt, 0), fd = f_env(p, kt, 1);
    double ax = -p->A0x * sin(ph) * f;
    double ay = p->A0y * cos(ph) * f;
    double ex, ey;
    if (!consistent) { /* reference sign of the f' term (quirk Q1) */
        ex = p->E0x * cos(ph) * f - p->A0x * sin(ph) * fd;
        ey = p->E0y * sin(ph) * f + p->A0y * cos(ph) * fd;
    } else {           /* E = -dA/dt */
        ex = p->E0x * cos(ph) * f + p->A0x * sin(ph) * fd;
        ey = p->E0y * sin(ph) * f - p->A0y * cos(ph) * fd;
    }
    for (int i = 0; i < 3; ++i) {
        A[i] = ax * p->xdir[i] + ay * p->ydir[i];
        E[i] = ex * p->xdir[i] + ey * p->ydir[i];
    }
}

/* laser::avlaser / laser::elaser, laser.h:301-322 */
static void laser_eval(const olaser *L, double t, double A[3], double E[3])
{
    double a[3] = {0, 0, 0}, e[3] = {0, 0, 0}, pa[3], pe[3];
    for (int i = 0; i < L->np; ++i) {
        pulse_eval(&L->p[i], t, L->consistent, pa, pe);
        for (int j = 0; j < 3; ++j) { a[j] += pa[j]; e[j] += pe[j]; }
    }
    for (int j = 0; j < 3; ++j) {
        A[j] = a[j] - L->A0[j];
        E[j] = e[j] - L->E0[j];
    }
}

/* Pulse::Pulse laser.h:68-141 and laser::Initialize laser.h:278-298 */
static int laser_init(const sbe_config *cfg, olaser *L)
{
    if (cfg->npulses < 1 || !cfg->pulses) return fail(SBE_ERR_CONFIG, "no laser pulses");
    if (!(cfg->dt > 0)) return fail(SBE_ERR_CONFIG, "dt must be > 0");
    L->np = cfg->npulses;
    L->p = (opulse *)calloc((size_t)L->np, sizeof(opulse));
    L->consistent = cfg->efield_mode == SBE_EFIELD_CONSISTENT;
    double smin = 0, emax = 0;
    for (int i = 0; i < L->np; ++i) {
        const sbe_pulse *s = &cfg->pulses[i];
        opulse *p = &L->p[i];
        p->w0 = s->w0; p->cep = s->cep; p->t0 = s->t0; p->env = s->envelope;
        p->period0 = 2.0 * PI / s->w0;
        switch (s->envelope) {
        case SBE_ENV_GAUSS:
            p->twidth = 2.0 * s->ncycles * p->period0;
            p->envfactor = -4.0 * log(2.0) / (s->ncycles * p->period0 * s->ncycles * p->period0);
            break;
        case SBE_ENV_GAUSS_CLIPPED:
            p->twidth = 2.0 * s->ncycles * p->period0 + s->ncycles_clipped * p->period0 / 2;
            p->twidth_clipped = s->ncycles_clipped * p->period0 / 2;
            p->envfactor = -4.0 * log(2.0) / (s->ncycles * p->period0 * s->ncycles * p->period0);
            break;
        case SBE_ENV_COS4:
        case SBE_ENV_SIN2:
            p->twidth = s->ncycles * p->period0 / 2;
            p->envfactor = s->w0 / (2 * s->ncycles);
            break;
        case SBE_ENV_COS4_CLIPPED:
            p->twidth = (s->ncycles + s->ncycles_clipped) * p->period0 / 2;
            p->twidth_clipped = s->ncycles_clipped * p->period0 / 2;
            p->envfactor = s->w0 / (2 * s->ncycles);
            break;
        default:
            free(L->p); L->p = NULL;
            return fail(SBE_ERR_CONFIG, "invalid envelope type %d", s->envelope);
        }
        double factor = sqrt(1.0 / (1.0 + s->ellip * s->ellip));
        p->E0x = factor * s->E0;
        p->E0y = s->ellip * factor * s->E0;
        if (s->w0 > 1.0e-30) { p->A0x = p->E0x / s->w0; p->A0y = p->E0y / s->w0; }
        double tx0[3] = {cos(s->thetaz) * cos(s->phiz), cos(s->thetaz) * sin(s->phiz), -sin(s->thetaz)};
        double ty0[3] = {-sin(s->phiz), cos(s->phiz), 0.0};
        for (int j = 0; j < 3; ++j) {
            p->xdir[j] = cos(s->phix) * tx0[j] + sin(s->phix) * ty0[j];
            p->ydir[j] = -sin(s->phix) * tx0[j] + cos(s->phix) * ty0[j];
        }
        double st = s->t0 - p->twidth, et = s->t0 + p->twidth;
        if (i == 0 || st < smin) smin = st;
        if (i == 0 || et > emax) emax = et;
    }
    L->dt = cfg->dt;
    L->atmin = smin - cfg->offset_before;
    L->atmax = emax + cfg->offset_after;
    L->A0[0] = L->A0[1] = L->A0[2] = 0;
    L->E0[0] = L->E0[1] = L->E0[2] = 0;
    double a[3], e[3];
    laser_eval(L, L->atmin, a, e);
    memcpy(L->A0, a, sizeof a);
    memcpy(L->E0, e, sizeof e);
    L->nt = (int64_t)floor((L->atmax - L->atmin) / L->dt);
    return SBE_OK;
}

int orc_laser_eval(const sbe_config *cfg, int64_t n, const double *t, double *A,
                   double *E, double *atmin, double *atmax, int64_t *nt)
{
    olaser L;
    int rc = laser_init(cfg, &L);
    if (rc) return rc;
    for (int64_t i = 0; i < n; ++i) laser_eval(&L, t[i], A + 3 * i, E + 3 * i);
    if (atmin) *atmin = L.atmin;
    if (atmax) *atmax = L.atmax;
    if (nt) *nt = L.nt;
    free(L.p);
    return SBE_OK;
}

/* ------------------------------------------------------------------ model */
typedef struct {
    const sbe_model *m;
    int N;
    /* Haldane (Haldane.h:134-143) */
    double va[3][2], vb[3][2];
    /* Wannier */
    double *Rv;    /* [nr][3] cartesian R */
    double *tau;   /* [N][3] Re <0n|r|0n> (R = 0 block) */
} omodel;

static int model_init(const sbe_model *m, omodel *om)
{
    memset(om, 0, sizeof *om);
    om->m = m;
    if (m->kind == SBE_MODEL_HALDANE) {
        if (m->nband != 2) return fail(SBE_ERR_MODEL, "Haldane model has nband = 2");
        double a0 = m->a0;
        if (!(a0 > 0)) return fail(SBE_ERR_MODEL, "Haldane a0 must be > 0");
        om->N = 2;
        /* Haldane::SetBasis, Haldane.h:134-143 */
        om->va[0][0] = 0.;                  om->va[0][1] = a0;
        om->va[1][0] = -sqrt(3.) / 2 * a0;  om->va[1][1] = -0.5 * a0;
        om->va[2][0] = sqrt(3.) / 2 * a0;   om->va[2][1] = -0.5 * a0;
        om->vb[0][0] = sqrt(3.) * a0;       om->vb[0][1] = 0.;
        om->vb[1][0] = -sqrt(3.) / 2 * a0;  om->vb[1][1] = 3. / 2. * a0;
        om->vb[2][0] = -sqrt(3.) / 2 * a0;  om->vb[2][1] = -3. / 2. * a0;
        return SBE_OK;
    }
    if (m->kind == SBE_MODEL_WANNIER) {
        int N = m->nband, nr = m->nrpts;
        if (N < 1 || nr < 1 || !m->rvec || !m->rweight || !m->ham)
            return fail(SBE_ERR_MODEL, "incomplete Wannier model");
        if (!m->is_dipole_zero && !m->pos)
            return fail(SBE_ERR_MODEL, "Wannier model with dipoles needs pos");
        om->N = N;
        om->Rv = (double *)calloc((size_t)nr * 3, sizeof(double));
        om->tau = (double *)calloc((size_t)N * 3, sizeof(double));
        int r0 = -1;
        /* Wannier90.h:193-204: R = n1 a1 + n2 a2 + n3 a3 */
        for (int r = 0; r < nr; ++r) {
            const int32_t *n = m->rvec + 3 * r;
            if (n[0] == 0 && n[1] == 0 && n[2] == 0) r0 = r;
            for (int i = 0; i < 3; ++i)
                om->Rv[3 * r + i] = n[0] * m->lattice[0 * 3 + i] + n[1] * m->lattice[1 * 3 + i] +
                                    n[2] * m->lattice[2 * 3 + i];
        }
        /* Wannier90.h:178-191: tau_n = Re <0n|r|0n> enters r = R + tau_n - tau_m */
        if (m->pos && r0 >= 0)
            for (int n = 0; n < N; ++n)
                for (int i = 0; i < 3; ++i)
                    om->tau[3 * n + i] = m->pos[((((size_t)r0 * 3 + i) * N + n) * N + n) * 2];
        return SBE_OK;
    }
    if (m->kind >= SBE_MODEL_WILSON_MASS && m->kind <= SBE_MODEL_WDSM_TAAS) {
        int want = (m->kind == SBE_MODEL_KANE_MELE || m->kind == SBE_MODEL_XYH_LATTICE ||
                    m->kind == SBE_MODEL_HALDANE_2L || m->kind == SBE_MODEL_WDSM_TAAS) ? 4
                   : m->kind == SBE_MODEL_TMDC ? (m->nband == 6 ? 6 : 3) : 2;
        if (m->nband != want) return fail(SBE_ERR_MODEL, "model class has %d bands", want);
        om->N = want;
        return SBE_OK;
    }
    return fail(SBE_ERR_MODEL, "unknown model kind %d", m->kind);
}

static void model_free(omodel *om)
{
    free(om->Rv);
    free(om->tau);
}

/* Haldane::GenBcomp / GenHamiltonian, Haldane.h:86-92,157-170 */
static void haldane_H(const omodel *om, const double k[3], cplx *H)
{
    const sbe_model *m = om->m;
    double B[4] = {0, 0, 0, 0};
    for (int i = 0; i < 3; ++i) {
        double aa = k[0] * om->va[i][0] + k[1] * om->va[i][1];
        double ab = k[0] * om->vb[i][0] + k[1] * om->vb[i][1];
        B[0] += cos(ab); B[1] += cos(aa); B[2] += sin(aa); B[3] += sin(ab);
    }
    B[0] *= 2. * m->t2 * cos(m->phi0); B[1] *= m->t1;
    B[2] *= m->t1;                     B[3] = m->M0 - 2. * m->t2 * sin(m->phi0) * B[3];
    H[0] = B[0] + B[3];      H[1] = B[1] - I * B[2];
    H[2] = B[1] + I * B[2];  H[3] = B[0] - B[3];
}

/* Haldane::GenJMatrix, Haldane.h:94-127 (J_z stays zero) */
static void haldane_J(const omodel *om, const double k[3], cplx *J)
{
    const sbe_model *m = om->m;
    double xd[4] = {0, 0, 0, 0}, yd[4] = {0, 0, 0, 0};
    for (int i = 0; i < 3; ++i) {
        double aa = k[0] * om->va[i][0] + k[1] * om->va[i][1];
        double ab = k[0] * om->vb[i][0] + k[1] * om->vb[i][1];
        xd[0] += om->vb[i][0] * sin(ab); yd[0] += om->vb[i][1] * sin(ab);
        xd[1] += om->va[i][0] * sin(aa); yd[1] += om->va[i][1] * sin(aa);
        xd[2] += om->va[i][0] * cos(aa); yd[2] += om->va[i][1] * cos(aa);
        xd[3] += om->vb[i][0] * cos(ab); yd[3] += om->vb[i][1] * cos(ab);
    }
    double c0 = -2. * m->t2 * cos(m->phi0), c3 = -2. * m->t2 * sin(m->phi0);
    xd[0] *= c0; yd[0] *= c0; xd[1] *= -m->t1; yd[1] *= -m->t1;
    xd[2] *= m->t1; yd[2] *= m->t1; xd[3] *= c3; yd[3] *= c3;
    cplx *Jx = J, *Jy = J + 4, *Jz = J + 8;
    Jx[0] = -(xd[0] + xd[3]);       Jy[0] = -(yd[0] + yd[3]);
    Jx[1] = -(xd[1] - I * xd[2]);   Jy[1] = -(yd[1] - I * yd[2]);
    Jx[2] = -(xd[1] + I * xd[2]);   Jy[2] = -(yd[1] + I * yd[2]);
    Jx[3] = -(xd[0] - xd[3]);       Jy[3] = -(yd[0] - yd[3]);
    for (int i = 0; i < 4; ++i) Jz[i] = 0;
}

static inline cplx wham(const sbe_model *m, int r, int a, int b)
{
    int N = m->nband;
    const double *p = m->ham + (((size_t)r * N + a) * N + b) * 2;
    return p[0] + I * p[1];
}

static inline cplx wpos(const sbe_model *m, int r, int ax, int a, int b)
{
    int N = m->nband;
    const double *p = m->pos + ((((size_t)r * 3 + ax) * N + a) * N + b) * 2;
    return p[0] + I * p[1];
}

/* phase vector of element (a,b) of block r: R (dipoles kept) or
 * R + tau_b - tau_a (is_dipole_zero), Wannier90.h:178-204,297-323 */
static inline void wvec(const omodel *om, int r, int a, int b, double v[3])
{
    for (int i = 0; i < 3; ++i) {
        v[i] = om->Rv[3 * r + i];
        if (om->m->is_dipole_zero) v[i] += om->tau[3 * b + i] - om->tau[3 * a + i];
    }
}

/* Wannier90::GenHamiltonianPrimitive, Wannier90.h:297-323 */
static void wannier_H(const omodel *om, const double k[3], cplx *H)
{
    const sbe_model *m = om->m;
    int N = m->nband;
    for (int a = 0; a < N; ++a)
        for (int b = 0; b < N; ++b) {
            cplx s = 0;
            for (int r = 0; r < m->nrpts; ++r) {
                double v[3];
                wvec(om, r, a, b, v);
                double ph = k[0] * v[0] + k[1] * v[1] + k[2] * v[2];
                s += m->rweight[r] * wham(m, r, a, b) * cexp(I * ph);
            }
            H[a * N + b] = s;
        }
}

/* Wannier90::GenDipolePrimitive, Wannier90.h:324-345 */
static void wannier_D(const omodel *om, const double k[3], cplx *D)
{
    const sbe_model *m = om->m;
    int N = m->nband;
    memset(D, 0, sizeof(cplx) * 3 * N * N);
    if (m->is_dipole_zero) return;
    for (int ax = 0; ax < 3; ++ax)
        for (int a = 0; a < N; ++a)
            for (int b = 0; b < N; ++b) {
                cplx s = 0;
                for (int r = 0; r < m->nrpts; ++r) {
                    const double *R = om->Rv + 3 * r;
                    double ph = k[0] * R[0] + k[1] * R[1] + k[2] * R[2];
                    s += m->rweight[r] * wpos(m, r, ax, a, b) * cexp(I * ph);
                }
                D[(ax * N + a) * N + b] = s;
            }
}

/* Wannier90::GenJMatrixPrimitive, Wannier90.h:347-387 (no i[D,H] term, as in
 * the reference) */
static void wannier_J(const omodel *om, const double k[3], cplx *J)
{
    const sbe_model *m = om->m;
    int N = m->nband;
    for (int ax = 0; ax < 3; ++ax)
        for (int a = 0; a < N; ++a)
            for (int b = 0; b < N; ++b) {
                cplx s = 0;
                for (int r = 0; r < m->nrpts; ++r) {
                    double v[3];
                    wvec(om, r, a, b, v);
                    double ph = k[0] * v[0] + k[1] * v[1] + k[2] * v[2];
                    s -= I * v[ax] * wham(m, r, a, b) * m->rweight[r] * cexp(I * ph);
                }
                J[(ax * N + a) * N + b] = s;
            }
}

/* ---------------------------------------------- analytic model classes */
/* 2x2 from Pauli components (b0, b1, b2, b3): [[b0+b3, b1-ib2], [b1+ib2, b0-b3]]
 * at rows/cols (o, o+1) of an N x N matrix */
static void pauli_block(cplx *M, int N, int o, double b0, double b1, double b2, double b3)
{
    M[o * N + o] = b0 + b3;
    M[o * N + o + 1] = b1 - I * b2;
    M[(o + 1) * N + o] = b1 + I * b2;
    M[(o + 1) * N + o + 1] = b0 - b3;
}

/* WilsonMass::GenHamiltonian / GenJMatrix, WilsonMass.h:65-92,110-116 */
static void wilson_mass_HJ(const double *p, const double k[3], cplx *H, cplx *J)
{
    double t = p[0], dl = p[1], mu = p[2], a0 = p[3], a1 = p[4];
    double B[4] = {0., dl * sin(a0 * k[0]), dl * sin(a1 * k[1]), -2. * t * (cos(a0 * k[0]) + cos(a1 * k[1])) - mu};
    if (H) pauli_block(H, 2, 0, B[0], B[1], B[2], B[3]);
    if (J) {
        double X[4] = {0., a0 * dl * cos(a0 * k[0]), 0., 2. * t * a0 * sin(a0 * k[0])};
        double Y[4] = {0., 0., a1 * dl * cos(a1 * k[1]), 2. * t * a1 * sin(a1 * k[1])};
        pauli_block(J, 2, 0, -X[0], -X[1], -X[2], -X[3]);
        pauli_block(J + 4, 2, 0, -Y[0], -Y[1], -Y[2], -Y[3]);
    }
}

/* SquareSBand.h:61-78,95-101: J has only the (0,1) element */
static void square_sband_HJ(const double *p, const double k[3], cplx *H, cplx *J)
{
    double t1 = p[0], M0 = p[1], a0 = p[2];
    if (H) pauli_block(H, 2, 0, 0., 2 * t1 * (cos(a0 * k[0]) + cos(a0 * k[1])), 0., M0);
    if (J) {
        J[1] = 2 * t1 * a0 * sin(a0 * k[0]);
        J[4 + 1] = 2 * t1 * a0 * sin(a0 * k[1]);
    }
}

static void honeycomb_vectors(double a0, double va[3][2], double vb[3][2])
{
    va[0][0] = 0.;                 va[0][1] = a0;
    va[1][0] = -sqrt(3.) / 2 * a0; va[1][1] = -0.5 * a0;
    va[2][0] = sqrt(3.) / 2 * a0;  va[2][1] = -0.5 * a0;
    vb[0][0] = sqrt(3.) * a0;      vb[0][1] = 0.;
    vb[1][0] = -sqrt(3.) / 2 * a0; vb[1][1] = 3. / 2. * a0;
    vb[2][0] = -sqrt(3.) / 2 * a0; vb[2][1] = -3. / 2. * a0;
}

/* KaneMele::GenHamiltonian / GenJMatrix / GenBcomp / GenRcomp, KaneMele.h:77-248 */
static void kane_mele_HJ(const double *p, const double k[3], cplx *H, cplx *J)
{
    double t1 = p[0], t2 = p[1], tR = p[2], M0 = p[3], phi0 = p[4], a0 = p[5];
    double va[3][2], vb[3][2];
    honeycomb_vectors(a0, va, vb);
    const double cR[3][2] = {{1., 0.}, {-0.5, sqrt(3.) / 2}, {-0.5, -sqrt(3.) / 2}};
    double aa[3], ab[3];
    for (int i = 0; i < 3; ++i) {
        aa[i] = k[0] * va[i][0] + k[1] * va[i][1];
        ab[i] = k[0] * vb[i][0] + k[1] * vb[i][1];
    }
    if (H) {
        for (int blk = 0; blk < 2; ++blk) {
            double ph = blk ? -phi0 : phi0, B[4] = {0, 0, 0, 0};
            for (int i = 0; i < 3; ++i) {
                B[0] += cos(ab[i]); B[1] += cos(aa[i]); B[2] += sin(aa[i]); B[3] += sin(ab[i]);
            }
            B[0] *= 2. * t2 * cos(ph); B[1] *= t1; B[2] *= t1; B[3] = M0 - 2. * t2 * sin(ph) * B[3];
            pauli_block(H, 4, 2 * blk, B[0], B[1], B[2], B[3]);
        }
        cplx R0 = 0, R1 = 0;
        for (int i = 0; i < 3; ++i) {
            R0 += (cR[i][0] - I * cR[i][1]) * cexp(I * aa[i]);
            R1 += (cR[i][0] - I * cR[i][1]) * cexp(-I * aa[i]);
        }
        R0 *= I * tR; R1 *= I * tR;
        H[3] = R0; H[6] = R1; H[9] = conj(R1); H[12] = conj(R0);
    }
    if (J) {
        for (int ax = 0; ax < 2; ++ax) {
            cplx *Ja = J + 16 * ax;
            double d[4] = {0, 0, 0, 0};
            for (int i = 0; i < 3; ++i) {
                d[0] += vb[i][ax] * sin(ab[i]); d[1] += va[i][ax] * sin(aa[i]);
                d[2] += va[i][ax] * cos(aa[i]); d[3] += vb[i][ax] * cos(ab[i]);
            }
            d[0] *= -2. * t2 * cos(phi0); d[1] *= -t1; d[2] *= t1; d[3] *= -2. * t2 * sin(phi0);
            pauli_block(Ja, 4, 0, -d[0], -d[1], -d[2], -d[3]);
            pauli_block(Ja, 4, 2, -d[0], -d[1], -d[2], d[3]); /* -phi0: sin(phi0) -> -sin(phi0) */
            cplx R[4] = {0, 0, 0, 0};
            for (int i = 0; i < 3; ++i) {
                cplx cm = cR[i][0] - I * cR[i][1], cp = cR[i][0] + I * cR[i][1];
                R[0] += cm * cexp(I * aa[i]) * I * va[i][ax];
                R[1] += cm * cexp(-I * aa[i]) * -I * va[i][ax];
                R[2] -= cp * cexp(I * aa[i]) * I * va[i][ax];
                R[3] -= cp * cexp(-I * aa[i]) * -I * va[i][ax];
            }
            for (int i = 0; i < 3; ++i) R[i] *= I * tR; /* not R[3] (KaneMele.h:155,174) */
            Ja[3] = R[0]; Ja[6] = R[1]; Ja[9] = R[2]; Ja[12] = R[3];
        }
    }
}

/* XYHlattice::GenHamiltonian / GenJMatrix / GenBcomp, XYHlattice.h:75-183 */
static void xyh_fill(cplx *M, double tl, double tr, cplx p0, cplx p1, cplx p2)
{
    cplx ap = p0 + p1, am = p0 - p1;
    M[2] = (3. / 4 * tl + 1. / 4 * tr) * ap + tr * p2;
    M[3] = sqrt(3.) / 4 * (tl - tr) * am;
    M[6] = M[3];
    M[7] = (1. / 4 * tl + 3. / 4 * tr) * ap + tl * p2;
    M[8] = conj(M[2]); M[9] = conj(M[6]); M[12] = conj(M[3]); M[13] = conj(M[7]);
}

static void xyh_lattice_HJ(const double *p, const double k[3], cplx *H, cplx *J)
{
    double tl = p[0], tr = p[1], lam = p[2], a0 = p[3];
    double va[3][2] = {{0., a0 / sqrt(3.)}, {-0.5 * a0, -0.5 * a0 / sqrt(3.)}, {0.5 * a0, -0.5 * a0 / sqrt(3.)}};
    cplx ph[3];
    for (int i = 0; i < 3; ++i) ph[i] = cexp(-I * (k[0] * va[i][0] + k[1] * va[i][1]));
    if (H) {
        H[0] = lam; H[5] = lam; H[10] = -lam; H[15] = -lam;
        xyh_fill(H, tl, tr, ph[0], ph[1], ph[2]);
    }
    if (J)
        for (int ax = 0; ax < 2; ++ax)
            xyh_fill(J + 16 * ax, tl, tr, -I * va[0][ax] * ph[0], -I * va[1][ax] * ph[1], -I * va[2][ax] * ph[2]);
}

/* TMDCs::GenHamiltonian / GenJMatrix, TMDCs.h:151-283 */
static void tmdc_fill(cplx *M, int N, int o, double v0, cplx v1, cplx v2, double v11, cplx v12, double v22, double s)
{
    M[(o + 0) * N + o + 0] = v0;       M[(o + 0) * N + o + 1] = v1;            M[(o + 0) * N + o + 2] = v2;
    M[(o + 1) * N + o + 0] = conj(v1); M[(o + 1) * N + o + 1] = v11;           M[(o + 1) * N + o + 2] = v12 + I * s;
    M[(o + 2) * N + o + 0] = conj(v2); M[(o + 2) * N + o + 1] = conj(v12) - I * s; M[(o + 2) * N + o + 2] = v22;
}

static void tmdc_HJ(const double *p, int N, const double k[3], cplx *H, cplx *J)
{
    double a0 = p[0], soc = p[1], e1 = p[2], e2 = p[3], t0 = p[4], t1 = p[5], t2 = p[6], t11 = p[7], t12 = p[8],
           t22 = p[9], r0 = p[10], r1 = p[11], r2 = p[12], r11 = p[13], r12 = p[14], u0 = p[15], u1 = p[16],
           u2 = p[17], u11 = p[18], u12 = p[19], u22 = p[20];
    double r3 = sqrt(3.), al = k[0] * a0 / 2, be = k[1] * a0 * r3 / 2;
    double ca[4], sa[4], cb[4], sb[4], dca[4], dsa[4], dcb[4], dsb[4];
    for (int i = 0; i < 4; ++i) {
        ca[i] = cos((i + 1) * al); cb[i] = cos((i + 1) * be);
        sa[i] = sin((i + 1) * al); sb[i] = sin((i + 1) * be);
        dca[i] = -(i + 1) * a0 / 2 * sin((i + 1) * al); dcb[i] = -(i + 1) * a0 * r3 / 2 * sin((i + 1) * be);
        dsa[i] = (i + 1) * a0 / 2 * cos((i + 1) * al);  dsb[i] = (i + 1) * a0 * r3 / 2 * cos((i + 1) * be);
    }
    int soc_on = N == 6;
    if (H) {
        double V0 = e1 + 2 * t0 * (2 * ca[0] * cb[0] + ca[1]) + 2 * r0 * (2 * ca[2] * cb[0] + cb[1]) +
                    2 * u0 * (2 * ca[1] * cb[1] + ca[3]);
        cplx V1 = (-2 * r3 * t2 * sa[0] * sb[0] + 2 * (r1 + r2) * sa[2] * sb[0] - 2 * r3 * u2 * sa[1] * sb[1]) +
                  I * (2 * t1 * sa[0] * (2 * ca[0] + cb[0]) + 2 * (r1 - r2) * sa[2] * cb[0] +
                       2 * u1 * sa[1] * (2 * ca[1] + cb[1]));
        cplx V2 = (2 * t2 * (ca[1] - ca[0] * cb[0]) - 2 / r3 * (r1 + r2) * (ca[2] * cb[0] - cb[1]) +
                   2 * u2 * (ca[3] - ca[1] * cb[1])) +
                  I * (2 * r3 * t1 * ca[0] * sb[0] + 2 / r3 * sb[0] * (r1 - r2) * (ca[2] + 2 * cb[0]) +
                       2 * r3 * u1 * ca[1] * sb[1]);
        double V11 = e2 + (t11 + 3 * t22) * ca[0] * cb[0] + 2 * t11 * ca[1] + 4 * r11 * ca[2] * cb[0] +
                     2 * (r11 + r3 * r12) * cb[1] + (u11 + 3 * u22) * ca[1] * cb[1] + 2 * u11 * ca[3];
        cplx V12 = (r3 * (t22 - t11) * sa[0] * sb[0] + 4 * r12 * sa[2] * sb[0] + r3 * (u22 - u11) * sa[1] * sb[1]) +
                   I * (4 * t12 * sa[0] * (ca[0] - cb[0]) + 4 * u12 * sa[1] * (ca[1] - cb[1]));
        double V22 = e2 + (3 * t11 + t22) * ca[0] * cb[0] + 2 * t22 * ca[1] + 2 * r11 * (2 * ca[2] * cb[0] + cb[1]) +
                     2 / r3 * r12 * (4 * ca[2] * cb[0] - cb[1]) + (3 * u11 + u22) * ca[1] * cb[1] + 2 * u22 * ca[3];
        tmdc_fill(H, N, 0, V0, V1, V2, V11, V12, V22, soc_on ? soc : 0.);
        if (soc_on) tmdc_fill(H, N, 3, V0, V1, V2, V11, V12, V22, -soc);
    }
    if (J) {
        double dxV0 = 2 * t0 * (2 * dca[0] * cb[0] + dca[1]) + 2 * r0 * (2 * dca[2] * cb[0]) +
                      2 * u0 * (2 * dca[1] * cb[1] + dca[3]);
        cplx dxV1 = (-2 * r3 * t2 * dsa[0] * sb[0] + 2 * (r1 + r2) * dsa[2] * sb[0] - 2 * r3 * u2 * dsa[1] * sb[1]) +
                    I * (2 * t1 * dsa[0] * (2 * ca[0] + cb[0]) + 2 * t1 * sa[0] * (2 * dca[0]) +
                         2 * (r1 - r2) * dsa[2] * cb[0] + 2 * u1 * dsa[1] * (2 * ca[1] + cb[1]) +
                         2 * u1 * sa[1] * (2 * dca[1]));
        cplx dxV2 = (2 * t2 * (dca[1] - dca[0] * cb[0]) - 2 / r3 * (r1 + r2) * (dca[2] * cb[0]) +
                     2 * u2 * (dca[3] - dca[1] * cb[1])) +
                    I * (2 * r3 * t1 * dca[0] * sb[0] + 2 / r3 * sb[0] * (r1 - r2) * (dca[2]) +
                         2 * r3 * u1 * dca[1] * sb[1]);
        double dxV11 = (t11 + 3 * t22) * dca[0] * cb[0] + 2 * t11 * dca[1] + 4 * r11 * dca[2] * cb[0] +
                       (u11 + 3 * u22) * dca[1] * cb[1] + 2 * u11 * dca[3];
        cplx dxV12 = (r3 * (t22 - t11) * dsa[0] * sb[0] + 4 * r12 * dsa[2] * sb[0] + r3 * (u22 - u11) * dsa[1] * sb[1]) +
                     I * (4 * t12 * dsa[0] * (ca[0] - cb[0]) + 4 * t12 * sa[0] * (dca[0]) +
                          4 * u12 * dsa[1] * (ca[1] - cb[1]) + 4 * u12 * sa[1] * (dca[1]));
        double dxV22 = (3 * t11 + t22) * dca[0] * cb[0] + 2 * t22 * dca[1] + 2 * r11 * (2 * dca[2] * cb[0]) +
                       2 / r3 * r12 * (4 * dca[2] * cb[0]) + (3 * u11 + u22) * dca[1] * cb[1] + 2 * u22 * dca[3];
        double dyV0 = 2 * t0 * (2 * ca[0] * dcb[0]) + 2 * r0 * (2 * ca[2] * dcb[0] + dcb[1]) +
                      2 * u0 * (2 * ca[1] * dcb[1]);
        cplx dyV1 = (-2 * r3 * t2 * sa[0] * dsb[0] + 2 * (r1 + r2) * sa[2] * dsb[0] - 2 * r3 * u2 * sa[1] * dsb[1]) +
                    I * (2 * t1 * sa[0] * (dcb[0]) + 2 * (r1 - r2) * sa[2] * dcb[0] + 2 * u1 * sa[1] * (dcb[1]));
        cplx dyV2 = (2 * t2 * (-ca[0] * dcb[0]) - 2 / r3 * (r1 + r2) * (ca[2] * dcb[0] - dcb[1]) +
                     2 * u2 * (-ca[1] * dcb[1])) +
                    I * (2 * r3 * t1 * ca[0] * dsb[0] + 2 / r3 * dsb[0] * (r1 - r2) * (ca[2] + 2 * cb[0]) +
                         2 / r3 * sb[0] * (r1 - r2) * (2 * dcb[0]) + 2 * r3 * u1 * ca[1] * dsb[1]);
        double dyV11 = (t11 + 3 * t22) * ca[0] * dcb[0] + 4 * r11 * ca[2] * dcb[0] + 2 * (r11 + r3 * r12) * dcb[1] +
                       (u11 + 3 * u22) * ca[1] * dcb[1];
        cplx dyV12 = (r3 * (t22 - t11) * sa[0] * dsb[0] + 4 * r12 * sa[2] * dsb[0] + r3 * (u22 - u11) * sa[1] * dsb[1]) +
                     I * (4 * t12 * sa[0] * (-dcb[0]) + 4 * u12 * sa[1] * (-dcb[1]));
        double dyV22 = (3 * t11 + t22) * ca[0] * dcb[0] + 2 * r11 * (2 * ca[2] * dcb[0] + dcb[1]) +
                       2 / r3 * r12 * (4 * ca[2] * dcb[0] - dcb[1]) + (3 * u11 + u22) * ca[1] * dcb[1];
        int nn = N * N;
        tmdc_fill(J, N, 0, dxV0, dxV1, dxV2, dxV11, dxV12, dxV22, 0.);
        tmdc_fill(J + nn, N, 0, dyV0, dyV1, dyV2, dyV11, dyV12, dyV22, 0.);
        if (soc_on) {
            tmdc_fill(J, N, 3, dxV0, dxV1, dxV2, dxV11, dxV12, dxV22, 0.);
            tmdc_fill(J + nn, N, 3, dyV0, dyV1, dyV2, dyV11, dyV12, dyV22, 0.);
        }
    }
}

/* Haldane2L::GenHamiltonian / GenJMatrix / GenBcomp, Haldane2L.h:166-311 */
static void haldane2l_HJ(const double *p, const double k[3], cplx *H, cplx *J)
{
    double va[3][2], vb[3][2];
    honeycomb_vectors(p[8], va, vb);
    int same = p[9] != 0.;
    double sgn[2] = {p[10], p[11]};
    if (H) {
        double Bt[4] = {0, 0, 0, 0}, Bb[4];
        for (int i = 0; i < 3; ++i) {
            double aa = k[0] * va[i][0] + k[1] * va[i][1], ab = k[0] * vb[i][0] + k[1] * vb[i][1];
            Bt[0] += cos(ab); Bt[1] += cos(aa); Bt[2] += sin(aa); Bt[3] += sin(ab);
        }
        for (int c = 0; c < 4; ++c) Bb[c] = Bt[c]; /* raw sums for the bottom layer */
        Bt[0] *= 2. * p[1] * cos(p[3]); Bt[1] *= p[0]; Bt[2] *= p[0]; Bt[3] = p[2] - 2. * p[1] * sin(p[3]) * Bt[3];
        if (same) {
            for (int c = 0; c < 4; ++c) Bb[c] = Bt[c];
        } else {
            Bb[0] *= 2. * p[5] * cos(p[7]); Bb[1] *= p[4]; Bb[2] *= p[4];
            Bb[3] = p[6] - 2. * p[5] * sin(p[7]) * Bt[3]; /* the top's B3, Haldane2L.h:307 */
        }
        H[0] = Bt[0] + Bt[3];                 H[1] = Bt[1] - sgn[0] * I * Bt[2];
        H[4] = Bt[1] + sgn[0] * I * Bt[2];    H[5] = Bt[0] - Bt[3];
        H[10] = Bb[0] + Bb[3];                H[11] = Bb[1] - sgn[1] * I * Bb[2];
        H[14] = Bb[1] + sgn[1] * I * Bb[2];   H[15] = Bb[0] - Bb[3];
        H[2] = p[12]; H[3] = p[13]; H[6] = p[14]; H[7] = p[15];
        H[8] = p[12]; H[9] = p[14]; H[12] = p[13]; H[13] = p[15];
    }
    if (J) {
        for (int l = 0; l < 2; ++l) {
            const double *q = p + 4 * l;
            double X[4] = {0, 0, 0, 0}, Y[4] = {0, 0, 0, 0};
            for (int i = 0; i < 3; ++i) {
                double aa = k[0] * va[i][0] + k[1] * va[i][1], ab = k[0] * vb[i][0] + k[1] * vb[i][1];
                X[0] += vb[i][0] * sin(ab); Y[0] += vb[i][1] * sin(ab);
                X[1] += va[i][0] * sin(aa); Y[1] += va[i][1] * sin(aa);
                X[2] += va[i][0] * cos(aa); Y[2] += va[i][1] * cos(aa);
                X[3] += vb[i][0] * cos(ab); Y[3] += vb[i][1] * cos(ab);
            }
            double f[4] = {-2. * q[1] * cos(q[3]), -q[0], q[0], -2. * q[1] * sin(q[3])};
            for (int c = 0; c < 4; ++c) { X[c] *= f[c]; Y[c] *= f[c]; }
            int o = 10 * l; /* 0 or 10: first element of the layer block */
            for (int ax = 0; ax < 2; ++ax) {
                double *D = ax ? Y : X;
                cplx *Ja = J + 16 * ax;
                Ja[o + 0] = -(D[0] + D[3]);
                Ja[o + 1] = -(D[1] - sgn[l] * I * D[2]);
                Ja[o + 4] = -(D[1] + sgn[l] * I * D[2]);
                Ja[o + 5] = -(D[0] - D[3]);
            }
        }
    }
}

/* BieSe3surf::GenHamiltonian / GenJMatrix / GenBcomp, Bi2Se3surf.h:51-277,
 * with the sixth derivative component kept in bounds (the class writes it
 * past five-element arrays) */
static void bi2se3_HJ(const double k[3], cplx *H, cplx *J)
{
    const double ev = 27.211386245988;
    double A0 = -0.0255 / ev, A11 = 0.1937 / ev, A12 = 0.2240 / ev, A14 = 0.0551 / ev;
    double B0 = 0.0164 / ev, B11 = 0.1203 / ev, B12 = 0.3264 / ev, B14 = 0.0 / ev, m11 = -1.6978 / ev;
    (void)B12; /* enters h4, which the class computes and never uses */
    double a = 4.14 / 0.529177210903;
    double va[3][2] = {{a, 0.}, {-a / 2, sqrt(3.) / 2 * a}, {-a / 2, -sqrt(3.) / 2 * a}};
    double vb[3][2] = {{0., sqrt(3.) * a / 3}, {-a / 2, -sqrt(3.) / 6 * a}, {a / 2, -sqrt(3.) / 6 * a}};
    double aa[3], ab[3];
    for (int i = 0; i < 3; ++i) {
        aa[i] = k[0] * va[i][0] + k[1] * va[i][1];
        ab[i] = k[0] * vb[i][0] + k[1] * vb[i][1];
    }
    double w = -2. * PI / 3.;
    double h5g = 6. * A11 + 6. * B11 + m11, bco = sqrt(1. - B0 * B0 / (B11 * B11));
    if (H) {
        double cA = 0, sA = 0, cB = 0, sB = 0;
        for (int i = 0; i < 3; ++i) { cA += cos(aa[i]); sA += sin(aa[i]); cB += cos(ab[i]); sB += sin(ab[i]); }
        double h0 = 2. * A0 * cA + 2. * B0 * cB;
        double h1 = -2. * A14 * sin(w) * (sin(aa[1]) - sin(aa[2])) +
                    2. * B14 * (sin(ab[0]) + cos(w) * (sin(ab[1]) - sin(ab[2])));
        double h2 = -2. * B14 * sin(w) * (sin(ab[1]) - sin(ab[2])) -
                    2. * A14 * (sin(aa[0]) + cos(w) * (sin(aa[1]) + sin(aa[2])));
        double h3 = 2. * A12 * sA;
        double h5 = 2. * A11 * cA + 2. * B11 * cB + m11;
        pauli_block(H, 2, 0, h0 + B0 / B11 * (-h5 + h5g), bco * h1, bco * h2, bco * h3);
    }
    if (J) {
        for (int ax = 0; ax < 2; ++ax) {
            double dcA = 0, dsA = 0, dcB = 0, dsB = 0;
            for (int i = 0; i < 3; ++i) {
                dcA += -va[i][ax] * sin(aa[i]); dsA += va[i][ax] * cos(aa[i]);
                dcB += -vb[i][ax] * sin(ab[i]); dsB += vb[i][ax] * cos(ab[i]);
            }
            double d0 = 2. * A0 * dcA + 2. * B0 * dcB;
            double d1 = -2. * A14 * sin(w) * (va[1][ax] * cos(aa[1]) - va[2][ax] * cos(aa[2])) +
                        2. * B14 * (vb[0][ax] * cos(ab[0]) + cos(w) * (vb[1][ax] * cos(ab[1]) + vb[2][ax] * cos(ab[2])));
            double d2 = -2. * B14 * sin(w) * (vb[1][ax] * cos(ab[1]) - vb[2][ax] * cos(ab[2])) -
                        2. * A14 * (va[0][ax] * cos(aa[0]) + cos(w) * (va[1][ax] * cos(aa[1]) + va[2][ax] * cos(aa[2])));
            double d3 = 2. * A12 * dsA;
            double d5 = 2. * A11 * dcA + 2. * B11 * dcB;
            pauli_block(J + 4 * ax, 2, 0, -(d0 - B0 / B11 * d5), -bco * d1, -bco * d2, -bco * d3);
        }
    }
}

/* MultiWSM::GenHamiltonian / GenJMatrix / GenBcomp, MultiWSM.h:77-200 */
static void multi_wsm_HJ(const double *p, const double k[3], cplx *H, cplx *J)
{
    double n = p[0], v = p[1], al = p[2];
    double phip = atan2(k[1], k[0]), ps = k[0] * k[0] + k[1] * k[1];
    if (H) {
        double b1 = al * pow(ps, n / 2.) * cos(n * phip), b2 = al * pow(ps, n / 2.) * sin(n * phip);
        pauli_block(H, 2, 0, 0., b1, b2, v * k[2]);
    }
    if (J) {
        double X[4] = {0, 0, 0, 0}, Y[4] = {0, 0, 0, 0}, Z[4] = {0, 0, 0, v};
        X[1] = n * al * k[0] * cos(n * phip) * pow(ps, n / 2. - 1) + n * al * sin(n * phip) * pow(ps, n / 2.) * k[1] / ps;
        Y[1] = n * al * k[1] * cos(n * phip) * pow(ps, n / 2. - 1) - n * al * sin(n * phip) * pow(ps, n / 2.) * k[0] / ps;
        X[2] = n * al * k[0] * sin(n * phip) * pow(ps, n / 2. - 1) - n * al * k[1] * cos(n * phip) * pow(ps, n / 2.) / ps;
        Y[2] = n * al * k[1] * sin(n * phip) * pow(ps, n / 2. - 1) + n * al * k[0] * cos(n * phip) * pow(ps, n / 2.) / ps;
        pauli_block(J, 2, 0, -X[0], -X[1], -X[2], -X[3]);
        pauli_block(J + 4, 2, 0, -Y[0], -Y[1], -Y[2], -Y[3]);
        pauli_block(J + 8, 2, 0, -Z[0], -Z[1], -Z[2], -Z[3]);
    }
}

/* WDSM_TaAs::GenHamiltonian / GenJMatrix / GenBcomp, WDSM_TaAs.h:69-255:
 * J's y slot ends up holding -d/dkz (written last), its z slot stays zero */
static void wdsm_HJ(const double *p, const double k[3], cplx *H, cplx *J)
{
    double t = p[0], dl = p[1], my = p[2], mz = p[3], a0 = p[4];
    if (H) {
        double B[16] = {0}, Bc[16] = {0};
        B[1] += t * sin(k[2] * a0); B[4] += t * sin(k[2] * a0);
        B[11] += -t * sin(k[2] * a0); B[14] += -t * sin(k[2] * a0);
        double m = t * cos(k[0] * a0) + t * my * (1 - cos(k[1] * a0)) + t * mz * (1 - cos(k[2] * a0));
        B[2] += m; B[7] += m; B[8] += m; B[13] += m;
        Bc[2] -= t * sin(k[1] * a0); Bc[3] -= t * cos(k[1] * a0) * dl;
        Bc[6] -= t * cos(k[1] * a0) * dl; Bc[7] -= t * sin(k[1] * a0);
        Bc[8] += t * sin(k[1] * a0); Bc[9] += t * cos(k[1] * a0) * dl;
        Bc[12] += t * cos(k[1] * a0) * dl; Bc[13] += t * sin(k[1] * a0);
        for (int i = 0; i < 16; ++i) H[i] = B[i] + I * Bc[i];
    }
    if (J) {
        double X[16] = {0}, Z[16] = {0};
        X[2] += -t * sin(k[0] * a0); X[7] += -t * sin(k[0] * a0);
        X[8] += -t * sin(k[0] * a0); X[13] += -t * sin(k[0] * a0);
        Z[2] += t * mz * sin(k[2] * a0); Z[7] += t * mz * sin(k[2] * a0);
        Z[8] += t * mz * sin(k[2] * a0); Z[13] += t * mz * sin(k[2] * a0);
        Z[1] += t * cos(k[2] * a0); Z[4] += t * cos(k[2] * a0);
        Z[11] += -t * cos(k[2] * a0); Z[14] += -t * cos(k[2] * a0);
        for (int i = 0; i < 16; ++i) {
            J[i] = -X[i];
            J[16 + i] = -Z[i];
        }
    }
}

static int is_analytic(int kind) { return kind >= SBE_MODEL_WILSON_MASS && kind <= SBE_MODEL_WDSM_TAAS; }

/* H (if H) and J[3] (if J) of an analytic class; J_z stays zero */
static void analytic_HJ(const omodel *om, const double k[3], cplx *H, cplx *J)
{
    const sbe_model *m = om->m;
    int N = om->N;
    if (H) memset(H, 0, sizeof(cplx) * N * N);
    if (J) memset(J, 0, sizeof(cplx) * 3 * N * N);
    switch (m->kind) {
    case SBE_MODEL_WILSON_MASS: wilson_mass_HJ(m->params, k, H, J); break;
    case SBE_MODEL_SQUARE_SBAND: square_sband_HJ(m->params, k, H, J); break;
    case SBE_MODEL_KANE_MELE: kane_mele_HJ(m->params, k, H, J); break;
    case SBE_MODEL_XYH_LATTICE: xyh_lattice_HJ(m->params, k, H, J); break;
    case SBE_MODEL_TMDC: tmdc_HJ(m->params, N, k, H, J); break;
    case SBE_MODEL_HALDANE_2L: haldane2l_HJ(m->params, k, H, J); break;
    case SBE_MODEL_BI2SE3_SURF: bi2se3_HJ(k, H, J); break;
    case SBE_MODEL_MULTI_WSM: multi_wsm_HJ(m->params, k, H, J); break;
    case SBE_MODEL_WDSM_TAAS: wdsm_HJ(m->params, k, H, J); break;
    default: break;
    }
}

static void model_H(const omodel *om, const double k[3], cplx *H)
{
    if (om->m->kind == SBE_MODEL_HALDANE) haldane_H(om, k, H);
    else if (is_analytic(om->m->kind)) analytic_HJ(om, k, H, NULL);
    else wannier_H(om, k, H);
}

static void model_J(const omodel *om, const double k[3], cplx *J)
{
    if (om->m->kind == SBE_MODEL_HALDANE) haldane_J(om, k, J);
    else if (is_analytic(om->m->kind)) analytic_HJ(om, k, NULL, J);
    else wannier_J(om, k, J);
}

static void model_D(const omodel *om, const double k[3], cplx *D)
{
    int N = om->N;
    if (om->m->kind != SBE_MODEL_WANNIER || om->m->is_dipole_zero)
        memset(D, 0, sizeof(cplx) * 3 * N * N); /* the analytic classes' GenDipole do nothing */
    else wannier_D(om, k, D);
}

int orc_hamiltonian(const sbe_model *m, const double k[3], double *H)
{
    omodel om;
    int rc = model_init(m, &om);
    if (rc) return rc;
    model_H(&om, k, (cplx *)H);
    model_free(&om);
    return SBE_OK;
}

int orc_jmatrix(const sbe_model *m, const double k[3], double *J)
{
    omodel om;
    int rc = model_init(m, &om);
    if (rc) return rc;
    model_J(&om, k, (cplx *)J);
    model_free(&om);
    return SBE_OK;
}

int orc_dipole(const sbe_model *m, const double k[3], double *D)
{
    omodel om;
    int rc = model_init(m, &om);
    if (rc) return rc;
    model_D(&om, k, (cplx *)D);
    model_free(&om);
    return SBE_OK;
}

/* ------------------------------------------------------------ linear alg */
/* MatrixMult, utility.h:78-92: C = A B (zero fill, then i-j-l loop) */
static void matmul(cplx *C, const cplx *A, const cplx *B, int N)
{
    for (int i = 0; i < N * N; ++i) C[i] = 0;
    for (int i = 0; i < N; ++i)
        for (int j = 0; j < N; ++j)
            for (int l = 0; l < N; ++l) C[i * N + j] += A[i * N + l] * B[l * N + j];
}

static void adjoint(cplx *Ad, const cplx *A, int N)
{
    for (int m = 0; m < N; ++m)
        for (int n = 0; n < N; ++n) Ad[m * N + n] = conj(A[n * N + m]);
}

/* Hermitian eigensolver standing in for LAPACKE_zheevr (SBEs.h:874-898):
 * cyclic Jacobi with complex rotations G = diag(1, e^{-i phi}) R(theta),
 * eigenvalues sorted ascending, eigenvectors in columns. Pinned by the
 * invariants U^dag U = 1, U^dag H U = diag(ascending) and by numpy eigh
 * fixtures (tests). */
static void eigh(int N, const cplx *Hin, double *ev, cplx *U, cplx *A)
{
    for (int i = 0; i < N * N; ++i) A[i] = Hin[i];
    for (int i = 0; i < N; ++i)
        for (int j = 0; j < N; ++j) U[i * N + j] = (i == j) ? 1.0 : 0.0;
    for (int i = 0; i < N; ++i) A[i * N + i] = creal(A[i * N + i]);
    for (int sweep = 0; sweep < 60; ++sweep) {
        double off = 0, diag = 0;
        for (int p = 0; p < N; ++p) {
            diag += creal(A[p * N + p]) * creal(A[p * N + p]);
            for (int q = p + 1; q < N; ++q) off += creal(A[p * N + q] * conj(A[p * N + q]));
        }
        if (off <= 1e-30 * diag || off < 1e-300) break;
        for (int p = 0; p < N - 1; ++p)
            for (int q = p + 1; q < N; ++q) {
                cplx b = A[p * N + q];
                double ab = cabs(b);
                double app = creal(A[p * N + p]), aqq = creal(A[q * N + q]);
                if (ab <= 1e-300) continue;
                cplx ph = b / ab; /* e^{i phi} */
                double theta = (aqq - app) / (2.0 * ab);
                double t = (theta >= 0 ? 1.0 : -1.0) / (fabs(theta) + sqrt(theta * theta + 1.0));
                double c = 1.0 / sqrt(t * t + 1.0), s = t * c;
                /* G = [[c, s], [-s e^{-i phi}, c e^{-i phi}]] acting on (p,q) */
                cplx gpp = c, gpq = s, gqp = -s * conj(ph), gqq = c * conj(ph);
                for (int r = 0; r < N; ++r) { /* columns: A <- A G */
                    cplx ap = A[r * N + p], aq = A[r * N + q];
                    A[r * N + p] = ap * gpp + aq * gqp;
                    A[r * N + q] = ap * gpq + aq * gqq;
                    cplx up = U[r * N + p], uq = U[r * N + q];
                    U[r * N + p] = up * gpp + uq * gqp;
                    U[r * N + q] = up * gpq + uq * gqq;
                }
                for (int r = 0; r < N; ++r) { /* rows: A <- G^dag A */
                    cplx ap = A[p * N + r], aq = A[q * N + r];
                    A[p * N + r] = conj(gpp) * ap + conj(gqp) * aq;
                    A[q * N + r] = conj(gpq) * ap + conj(gqq) * aq;
                }
                A[p * N + q] = 0; A[q * N + p] = 0;
                A[p * N + p] = creal(A[p * N + p]);
                A[q * N + q] = creal(A[q * N + q]);
            }
    }
    for (int i = 0; i < N; ++i) ev[i] = creal(A[i * N + i]);
    /* ascending order (selection sort, stable for ties) */
    for (int i = 0; i < N - 1; ++i) {
        int mi = i;
        for (int j = i + 1; j < N; ++j)
            if (ev[j] < ev[mi]) mi = j;
        if (mi != i) {
            double te = ev[i]; ev[i] = ev[mi]; ev[mi] = te;
            for (int r = 0; r < N; ++r) {
                cplx tu = U[r * N + i]; U[r * N + i] = U[r * N + mi]; U[r * N + mi] = tu;
            }
        }
    }
}

int orc_eigh(int n, const double *H, double *evals, double *U)
{
    if (n < 1 || n > 256) return fail(SBE_ERR_RANGE, "eigh size %d", n);
    cplx *A = (cplx *)malloc(sizeof(cplx) * n * n);
    eigh(n, (const cplx *)H, evals, (cplx *)U, A);
    free(A);
    return SBE_OK;
}

/* ------------------------------------------------------------------- grid */
/* material BZ (Haldane.h:77-81, Wannier90.h:508-521), ksfactor and the
 * calc.BZaxes/BZorigin overrides (SBEs.h:219-241) */
static void bz_select(const sbe_config *cfg, const sbe_model *m, double ax[9], double org[3])
{
    memset(ax, 0, 9 * sizeof(double));
    memset(org, 0, 3 * sizeof(double));
    if (m->kind == SBE_MODEL_HALDANE) {
        double kx = PI / sqrt(3.) / m->a0, ky = 2. * PI / 3. / m->a0;
        ax[0] = 2 * kx; ax[4] = 2 * ky;
    } else if (m->kind != SBE_MODEL_WANNIER) {
        /* GenBrillouinzone of the classes: rectangles 2 kxMax x 2 kyMax at 0 */
        const double *p = m->params;
        double kx = 0, ky = 0;
        if (m->kind == SBE_MODEL_WILSON_MASS) { kx = PI / p[3]; ky = PI / p[4]; }            /* WilsonMass.h:56 */
        if (m->kind == SBE_MODEL_SQUARE_SBAND) { kx = PI / p[2]; ky = PI / p[2]; }           /* SquareSBand.h:54 */
        if (m->kind == SBE_MODEL_KANE_MELE) { kx = PI / sqrt(3.) / p[5]; ky = 2. * PI / 3. / p[5]; } /* KaneMele.h:68 */
        if (m->kind == SBE_MODEL_XYH_LATTICE) { kx = PI / p[3]; ky = 2. * PI / sqrt(3.) / p[3]; }    /* XYHlattice.h:65 */
        if (m->kind == SBE_MODEL_TMDC) { kx = PI / p[0]; ky = 2. * PI / sqrt(3.) / p[0]; }          /* TMDCs.h:143 */
        if (m->kind == SBE_MODEL_HALDANE_2L) { kx = PI / sqrt(3.) / p[8]; ky = 2. * PI / 3. / p[8]; } /* Haldane2L.h:151 */
        if (m->kind == SBE_MODEL_BI2SE3_SURF) {                                                       /* Bi2Se3surf.h:71 */
            double a = 4.14 / 0.529177210903;
            kx = PI / a; ky = 2. * PI / sqrt(3.) / a;
        }
        if (m->kind == SBE_MODEL_MULTI_WSM || m->kind == SBE_MODEL_WDSM_TAAS) { /* MultiWSM.h:65, WDSM_TaAs.h:61 */
            kx = ky = PI / p[4];
            ax[8] = 2 * PI / p[4];
        }
        ax[0] = 2 * kx; ax[4] = 2 * ky;
    } else {
        const double *a = m->lattice;
        double vol = a[0] * (a[4] * a[8] - a[5] * a[7]) + a[1] * (a[5] * a[6] - a[3] * a[8]) +
                     a[2] * (a[3] * a[7] - a[4] * a[6]);
        for (int i = 0; i < 3; ++i) {
            const double *u = a + 3 * ((i + 1) % 3), *v = a + 3 * ((i + 2) % 3);
            ax[3 * i + 0] = (u[1] * v[2] - u[2] * v[1]) * 2 * PI / vol;
            ax[3 * i + 1] = (u[2] * v[0] - u[0] * v[2]) * 2 * PI / vol;
            ax[3 * i + 2] = (u[0] * v[1] - v[0] * u[1]) * 2 * PI / vol;
        }
    }
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) ax[3 * i + j] *= cfg->ksfactor[i];
    if (cfg->has_bz_axes) memcpy(ax, cfg->bz_axes, 9 * sizeof(double));
    if (cfg->has_bz_origin) memcpy(org, cfg->bz_origin, 3 * sizeof(double));
}

/* momaxis::steps_sizes volume rule, momaxis.h:208-238: over the axes with
 * N != 1: |b1| (1D), |b1 x b2| (2D), sqrt((b1 x b2).b3) (3D, as written). */
static double bz_volume(const int N[3], const double ax[9])
{
    const double *v[3];
    int na = 0;
    for (int i = 0; i < 3; ++i)
        if (N[i] != 1) v[na++] = ax + 3 * i;
    if (na == 1) return sqrt(v[0][0] * v[0][0] + v[0][1] * v[0][1] + v[0][2] * v[0][2]);
    if (na >= 2) {
        double c[3] = {v[0][1] * v[1][2] - v[0][2] * v[1][1], v[0][2] * v[1][0] - v[0][0] * v[1][2],
                       v[0][0] * v[1][1] - v[1][0] * v[0][1]};
        if (na == 2) return sqrt(c[0] * c[0] + c[1] * c[1] + c[2] * c[2]);
        return sqrt(c[0] * v[2][0] + c[1] * v[2][1] + c[2] * v[2][2]);
    }
    return 1.0;
}

int orc_kgrid(const sbe_config *cfg, const sbe_model *model, int64_t k_begin, int64_t k_end,
              double *kpts, double *w, double *dV, double *total_weight)
{
    int N[3] = {cfg->npoints[0], cfg->npoints[1], cfg->npoints[2]};
    if (N[0] <= 0 || N[1] <= 0 || N[2] <= 0) return fail(SBE_ERR_CONFIG, "Npoints must be > 0");
    int64_t nt = (int64_t)N[0] * N[1] * N[2];
    if (k_begin < 0) { k_begin = 0; k_end = nt; }
    if (k_end > nt || k_begin >= k_end) return fail(SBE_ERR_RANGE, "bad k range");
    double ax[9], org[3];
    bz_select(cfg, model, ax, org);
    double tw = 0;
    for (int64_t g = k_begin; g < k_end; ++g) {
        int i = (int)(g / ((int64_t)N[1] * N[2])), j = (int)((g / N[2]) % N[1]), l = (int)(g % N[2]);
        int idx[3] = {i, j, l};
        double fr[3], wt = 1.0;
        for (int d = 0; d < 3; ++d) {
            if (cfg->grid == SBE_GRID_MP) {
                fr[d] = (idx[d] + 0.5) / N[d] - 0.5;
            } else {
                /* momaxis.h:353-370 and trapezoid weights :162-174 */
                fr[d] = (N[d] == 1) ? 0.0 : (double)idx[d] / (N[d] - 1) - 0.5;
                if (N[d] != 1 && (idx[d] == 0 || idx[d] == N[d] - 1)) wt /= 2;
            }
        }
        int64_t o = g - k_begin;
        for (int c = 0; c < 3; ++c)
            kpts[3 * o + c] = org[c] + fr[0] * ax[0 * 3 + c] + fr[1] * ax[1 * 3 + c] + fr[2] * ax[2 * 3 + c];
        w[o] = wt;
        tw += wt;
    }
    if (dV) *dV = fabs(bz_volume(N, ax)) / (double)nt;
    if (total_weight) *total_weight = tw;
    return SBE_OK;
}

/* ------------------------------------------------------------------ solver */
typedef struct {
    const sbe_config *cfg;
    omodel om;
    olaser L;
    int N, nstage;
    double c[6], b[6], a[6][6];
    double gam_intra, gam_inter;
    int velocity; /* GaugeType::VelocityHamiltonian */
} osolver;

/* Per-k constants of the VelocityHamiltonian gauge, built once at set-up as
 * SBEs.h:284-373 does: edispersion E[N], pMatrix P[3][N][N] = U^dag J(k) U
 * and initMatrix rho0 = diag(occupations) (band basis, no rotation). */
typedef struct {
    double *E;
    cplx *P, *rho0;
} ovel;

/* SBEs::InitializeGeneral post-processing (SBEs.h:484-508) and the RK
 * tableaus of SBEs.h:395-416 */
static int solver_init(const sbe_config *cfg, const sbe_model *model, osolver *S)
{
    memset(S, 0, sizeof *S);
    S->cfg = cfg;
    if (cfg->abi_version != SBE_ABI_VERSION) return fail(SBE_ERR_CONFIG, "abi_version mismatch");
    if (cfg->gauge != SBE_GAUGE_LENGTH_WANNIER && cfg->gauge != SBE_GAUGE_VELOCITY_HAMILTONIAN)
        return fail(SBE_ERR_UNSUPPORTED, "the LengthWannier and VelocityHamiltonian gauges are restated");
    S->velocity = cfg->gauge == SBE_GAUGE_VELOCITY_HAMILTONIAN;
    int rc = model_init(model, &S->om);
    if (rc) return rc;
    rc = laser_init(cfg, &S->L);
    if (rc) { model_free(&S->om); return rc; }
    S->N = S->om.N;
    S->gam_inter = cfg->T2 <= 1e-16 ? 0.0 : 1.0 / cfg->T2;
    S->gam_intra = cfg->T1 <= 1e-16 ? 0.0 : 1.0 / cfg->T1;
    if (cfg->rk_order == SBE_RK4) {
        S->nstage = 4;
        S->c[0] = 0.5; S->c[1] = 0.5; S->c[2] = 1.0;
        S->b[0] = 1.0 / 6.0; S->b[1] = 2.0 / 6.0; S->b[2] = 2.0 / 6.0; S->b[3] = 1.0 / 6.0;
        S->a[0][0] = 0.5;
        S->a[1][0] = 0.0; S->a[1][1] = 0.5;
        S->a[2][0] = 0.0; S->a[2][1] = 0.0; S->a[2][2] = 1.0;
    } else if (cfg->rk_order == SBE_RK5_FEHLBERG) {
        S->nstage = 6;
        S->c[0] = 0.25; S->c[1] = 3. / 8.; S->c[2] = 12. / 13.; S->c[3] = 1.0; S->c[4] = 0.5;
        S->b[0] = 16. / 135.; S->b[1] = 0; S->b[2] = 6656. / 12825.; S->b[3] = 28561. / 56430.;
        S->b[4] = -9. / 50.; S->b[5] = 2. / 55.;
        S->a[0][0] = 1.0 / 4.0;
        S->a[1][0] = 3. / 32.; S->a[1][1] = 9. / 32.;
        S->a[2][0] = 1932. / 2197.; S->a[2][1] = -7200. / 2197.; S->a[2][2] = 7296. / 2197.;
        S->a[3][0] = 439. / 216.; S->a[3][1] = -8.; S->a[3][2] = 3680. / 513.; S->a[3][3] = -845. / 4104.;
        S->a[4][0] = -8. / 27.; S->a[4][1] = 2.; S->a[4][2] = -3544. / 2565; S->a[4][3] = 1859. / 4104;
        S->a[4][4] = -11. / 40.;
    } else if (cfg->rk_order == SBE_RK_CASH_KARP) {
        S->nstage = 6; /* coefficients live in orc_run_adaptive */
    } else {
        model_free(&S->om); free(S->L.p);
        return fail(SBE_ERR_UNSUPPORTED, "oracle restates RKorder 4 and 5 (got %d)", cfg->rk_order);
    }
    return SBE_OK;
}

static void solver_free(osolver *S)
{
    model_free(&S->om);
    free(S->L.p);
}

typedef struct {
    cplx *H, *Hd, *T, *U, *Ud, *t1, *t2, *D, *J, *work;
    double *ev;
} oscratch;

static void scratch_alloc(oscratch *w, int N)
{
    size_t nn = (size_t)N * N;
    w->H = calloc(nn, sizeof(cplx)); w->Hd = calloc(nn, sizeof(cplx));
    w->T = calloc(nn, sizeof(cplx)); w->U = calloc(nn, sizeof(cplx));
    w->Ud = calloc(nn, sizeof(cplx)); w->t1 = calloc(nn, sizeof(cplx));
    w->t2 = calloc(nn, sizeof(cplx)); w->D = calloc(3 * nn, sizeof(cplx));
    w->J = calloc(3 * nn, sizeof(cplx)); w->work = calloc(nn, sizeof(cplx));
    w->ev = calloc((size_t)N, sizeof(double));
}

static void scratch_free(oscratch *w)
{
    free(w->H); free(w->Hd); free(w->T); free(w->U); free(w->Ud);
    free(w->t1); free(w->t2); free(w->D); free(w->J); free(w->work); free(w->ev);
}

/* SBEs::GenUMatrix, SBEs.h:874-898 */
static void gen_u(const osolver *S, const double kp[3], oscratch *w)
{
    model_H(&S->om, kp, w->Hd);
    eigh(S->N, w->Hd, w->ev, w->U, w->work);
    adjoint(w->Ud, w->U, S->N);
}

/* Eigenvector phases of the velocity gauge. The reference keeps whatever
 * phases zheevr returns (SBEs.h:284-316); LAPACK does not specify them, and
 * the off-diagonal band-basis elements of rho depend on them (currents and
 * populations do not). Fixed here: column j is multiplied by
 * conj(U[p][j]) / |U[p][j]| with p the first row where |U[p][j]|^2 >= 1/(2N)
 * (one exists, the column being a unit vector). */
static void phase_fix(int N, cplx *U)
{
    for (int j = 0; j < N; ++j) {
        int p = 0;
        while (p < N - 1 && creal(U[p * N + j] * conj(U[p * N + j])) * (2.0 * N) < 1.0) ++p;
        const double a = cabs(U[p * N + j]);
        if (a == 0.0) continue;
        const cplx ph = conj(U[p * N + j]) / a;
        for (int r = 0; r < N; ++r) U[r * N + j] *= ph;
        U[p * N + j] = a;
    }
}

/* Initial occupations of the bands (SBEs.h:318-340): the first Nval bands,
 * or Fermi-Dirac at kT = 300 K / 3.15775024804e5 normalised to sum 1. */
static void occupations(const osolver *S, const double *ev, double *occ)
{
    int N = S->N;
    if (S->cfg->init_type == SBE_INIT_FERMI_DIRAC) {
        double kT = 300. / 3.15775024804e5, sum = 0;
        for (int m = 0; m < N; ++m) { occ[m] = 1.0 / (exp((ev[m] - S->cfg->fermi_energy) / kT) + 1.0); sum += occ[m]; }
        for (int m = 0; m < N; ++m) occ[m] /= sum;
    } else {
        for (int m = 0; m < N; ++m) occ[m] = m < S->om.m->nval ? 1.0 : 0.0;
    }
}

/* Velocity-gauge constants at one k (SBEs.h:284-373): E(k), U(k) (phases
 * fixed as above), pMatrix_a = U^dag J_a(k) U with J = GenJMatrix, and
 * rho0 = diag(occupations). */
static void vel_setup(const osolver *S, const double k[3], double *E, cplx *P, cplx *rho0, oscratch *w)
{
    int N = S->N, nn = N * N;
    double occ[256];
    model_H(&S->om, k, w->Hd);
    eigh(N, w->Hd, E, w->U, w->work);
    phase_fix(N, w->U);
    adjoint(w->Ud, w->U, N);
    model_J(&S->om, k, w->J);
    for (int ax = 0; ax < 3; ++ax) {
        matmul(w->t1, w->J + ax * nn, w->U, N);
        matmul(P + ax * nn, w->Ud, w->t1, N);
    }
    occupations(S, E, occ);
    for (int i = 0; i < nn; ++i) rho0[i] = 0;
    for (int m = 0; m < N; ++m) rho0[m * N + m] = occ[m];
}

/* SBEs::GenDifferentialDM, VelocityHamiltonian branch, SBEs.h:787-805,
 * 814-829 and the relaxation toward initMatrix of SBEs.h:862-871:
 * out = -i((H rho) - (H rho)^dag) - (rho - rho0) o Gamma with
 * H = sum_a A_a(t) pMatrix_a + diag(E). */
static void rhs_velocity(const osolver *S, const ovel *V, double t, const cplx *rho, cplx *out, oscratch *w)
{
    int N = S->N, nn = N * N;
    double A[3], E[3];
    laser_eval(&S->L, t, A, E);
    for (int i = 0; i < nn; ++i) {
        w->H[i] = 0;
        for (int ax = 0; ax < 3; ++ax) w->H[i] += A[ax] * V->P[ax * nn + i];
    }
    for (int m = 0; m < N; ++m) w->H[m * N + m] += V->E[m];
    matmul(out, w->H, rho, N);
    for (int i = 0; i < nn; ++i) w->T[i] = out[i];
    for (int m = 0; m < N; ++m)
        for (int n = 0; n < N; ++n) {
            out[m * N + n] -= conj(w->T[n * N + m]);
            out[m * N + n] *= -I;
        }
    for (int m = 0; m < N; ++m)
        for (int n = 0; n < N; ++n)
            out[m * N + n] -= (rho[m * N + n] - V->rho0[m * N + n]) * ((m == n) ? S->gam_intra : S->gam_inter);
}

/* SBEs::GenDifferentialDM, LengthWannier branch, SBEs.h:762-786,814-861 */
static void rhs(const osolver *S, const double k[3], const ovel *V, double t, const cplx *rho, cplx *out,
                oscratch *w)
{
    if (S->velocity) {
        rhs_velocity(S, V, t, rho, out, w);
        return;
    }
    int N = S->N;
    double A[3], E[3];
    laser_eval(&S->L, t, A, E);
    double kp[3] = {k[0] + A[0], k[1] + A[1], k[2] + A[2]};
    model_H(&S->om, kp, w->H);
    if (!S->om.m->is_dipole_zero && S->om.m->kind == SBE_MODEL_WANNIER) {
        model_D(&S->om, kp, w->D);
        for (int i = 0; i < N * N; ++i)
            w->H[i] += E[0] * w->D[i] + E[1] * w->D[N * N + i] + E[2] * w->D[2 * N * N + i];
    }
    matmul(out, w->H, rho, N);
    for (int i = 0; i < N * N; ++i) w->T[i] = out[i];
    for (int m = 0; m < N; ++m)
        for (int n = 0; n < N; ++n) {
            out[m * N + n] -= conj(w->T[n * N + m]);
            out[m * N + n] *= -I;
        }
    gen_u(S, kp, w);
    matmul(w->t1, rho, w->U, N);
    matmul(w->t2, w->Ud, w->t1, N);
    for (int m = 0; m < N; ++m)
        for (int n = 0; n < N; ++n) w->t2[m * N + n] *= (m == n) ? S->gam_intra : S->gam_inter;
    matmul(w->t1, w->U, w->t2, N);
    matmul(w->t2, w->t1, w->Ud, N);
    for (int i = 0; i < N * N; ++i) out[i] -= w->t2[i];
}

/* velocity-gauge constants of one k in freshly allocated arrays */
static void vel_alloc(ovel *V, int N, size_t nk)
{
    size_t nn = (size_t)N * N;
    V->E = calloc(nk * N, sizeof(double));
    V->P = calloc(nk * 3 * nn, sizeof(cplx));
    V->rho0 = calloc(nk * nn, sizeof(cplx));
}

static void vel_free(ovel *V) { free(V->E); free(V->P); free(V->rho0); }

static ovel vel_at(const ovel *V, int N, int64_t k)
{
    size_t nn = (size_t)N * N;
    ovel v = {V->E + (size_t)k * N, V->P + (size_t)k * 3 * nn, V->rho0 + (size_t)k * nn};
    return v;
}

int orc_rhs(const sbe_config *cfg, const sbe_model *model, const double k[3], double t,
            const double *rho, double *out)
{
    osolver S;
    int rc = solver_init(cfg, model, &S);
    if (rc) return rc;
    oscratch w;
    scratch_alloc(&w, S.N);
    ovel V = {0};
    if (S.velocity) {
        vel_alloc(&V, S.N, 1);
        vel_setup(&S, k, V.E, V.P, V.rho0, &w);
    }
    rhs(&S, k, &V, t, (const cplx *)rho, (cplx *)out, &w);
    if (S.velocity) vel_free(&V);
    scratch_free(&w);
    solver_free(&S);
    return SBE_OK;
}

int orc_velocity_setup(const sbe_config *cfg, const sbe_model *model, const double k[3], double *E, double *P,
                       double *rho0)
{
    osolver S;
    int rc = solver_init(cfg, model, &S);
    if (rc) return rc;
    oscratch w;
    scratch_alloc(&w, S.N);
    vel_setup(&S, k, E, (cplx *)P, (cplx *)rho0, &w);
    scratch_free(&w);
    solver_free(&S);
    return SBE_OK;
}

/* SBEs::RunSBEs, SBEs.h:596-624 */
static void rk_step(const osolver *S, const double k[3], const ovel *V, double t, cplx *rho, cplx **stg,
                    cplx *tmp, oscratch *w)
{
    int N = S->N, nn = N * N;
    double dt = S->L.dt;
    rhs(S, k, V, t, rho, stg[0], w);
    for (int irk = 0; irk < S->nstage - 1; ++irk) {
        for (int i = 0; i < nn; ++i) {
            tmp[i] = rho[i];
            for (int j = 0; j < irk + 1; ++j) tmp[i] += dt * S->a[irk][j] * stg[j][i];
        }
        rhs(S, k, V, t + S->c[irk] * dt, tmp, stg[irk + 1], w);
    }
    for (int i = 0; i < nn; ++i)
        for (int irk = 0; irk < S->nstage; ++irk) rho[i] += dt * S->b[irk] * stg[irk][i];
}

/* SBEs::GenInterIntraCurrent SBEs.h:694-760 / SBEs::GenCurrent SBEs.h:626-657,
 * plus the band populations diag(U^dag rho U) that the reference builds as
 * newdMatrix but does not output. */
static void currents(const osolver *S, const double k[3], const ovel *V, double t, const cplx *rho,
                     double inter[3], double intra[3], double *pop, oscratch *w)
{
    int N = S->N;
    double A[3], E[3];
    laser_eval(&S->L, t, A, E);
    if (S->velocity) {
        /* _tkp = k: rho is already band-basis and U^dag J U is pMatrix; then
         * j = -(p + A(t)) adds Nval A(t) (SBEs.h:640-642,676-690,746-757) */
        for (int m = 0; m < N; ++m) pop[m] = creal(rho[m * N + m]);
        for (int ax = 0; ax < 3; ++ax) {
            const cplx *P = V->P + ax * N * N;
            inter[ax] = 0; intra[ax] = 0;
            for (int i = 0; i < N; ++i)
                for (int j = 0; j < N; ++j) {
                    double v = creal(rho[N * i + j] * P[N * j + i]);
                    if (i == j && S->cfg->inter_intra) intra[ax] += v;
                    else inter[ax] += v;
                }
            if (S->cfg->inter_intra) intra[ax] += S->om.m->nval * A[ax];
            else inter[ax] += S->om.m->nval * A[ax];
        }
        return;
    }
    double kp[3] = {k[0] + A[0], k[1] + A[1], k[2] + A[2]};
    model_J(&S->om, kp, w->J);
    gen_u(S, kp, w);
    matmul(w->t1, rho, w->U, N);
    matmul(w->T, w->Ud, w->t1, N); /* rho_H */
    for (int m = 0; m < N; ++m) pop[m] = creal(w->T[m * N + m]);
    for (int ax = 0; ax < 3; ++ax) {
        inter[ax] = 0; intra[ax] = 0;
        const cplx *J = w->J + ax * N * N;
        if (S->cfg->inter_intra) {
            matmul(w->t1, J, w->U, N);
            matmul(w->t2, w->Ud, w->t1, N);
            for (int i = 0; i < N; ++i)
                for (int j = 0; j < N; ++j) {
                    double v = creal(w->T[N * i + j] * w->t2[N * j + i]);
                    if (i == j) intra[ax] += v;
                    else inter[ax] += v;
                }
        } else {
            for (int i = 0; i < N; ++i)
                for (int j = 0; j < N; ++j) inter[ax] += creal(rho[N * i + j] * J[N * j + i]);
        }
    }
}

/* SBEs::SBEs per-k initial state, SBEs.h:284-364 (UniformValence and
 * FermiDirac, kT = 300 K / 3.15775024804e5) */
static void init_state(const osolver *S, const double k[3], cplx *rho, oscratch *w)
{
    int N = S->N;
    double occ[256];
    model_H(&S->om, k, w->Hd);
    eigh(N, w->Hd, w->ev, w->U, w->work);
    adjoint(w->Ud, w->U, N);
    for (int i = 0; i < N * N; ++i) rho[i] = 0;
    occupations(S, w->ev, occ);
    for (int m = 0; m < N; ++m) rho[m * N + m] = occ[m];
    matmul(w->t1, rho, w->Ud, N);
    matmul(rho, w->U, w->t1, N);
}

typedef struct {
    const osolver *S;
    const double *kp, *wk;
    int64_t k0, k1, nsteps;
    double *ji, *ja, *rb, *bp; /* per-thread series */
    cplx *rho;                 /* [nk][N*N] slice */
    const ovel *V;             /* velocity gauge: per-k constants [nk] */
} othread;

/* antelope.cpp:297-338 for one contiguous k block */
static void *thread_main(void *arg)
{
    othread *T = (othread *)arg;
    const osolver *S = T->S;
    int N = S->N, nn = N * N;
    oscratch w;
    scratch_alloc(&w, N);
    cplx *stg[6], *tmp = calloc((size_t)nn, sizeof(cplx));
    for (int i = 0; i < 6; ++i) stg[i] = calloc((size_t)nn, sizeof(cplx));
    double inter[3], intra[3], pop[256];
    for (int64_t n = 0; n < T->nsteps; ++n) {
        double t = S->L.atmin + S->L.dt * n;
        for (int64_t k = T->k0; k < T->k1; ++k) {
            const double *kk = T->kp + 3 * k;
            double wk = T->wk[k];
            cplx *rho = T->rho + (size_t)k * nn;
            ovel v = {0};
            if (S->velocity) v = vel_at(T->V, N, k);
            currents(S, kk, &v, t, rho, inter, intra, pop, &w);
            for (int ax = 0; ax < 3; ++ax) {
                T->ja[3 * n + ax] += intra[ax] * wk;
                T->ji[3 * n + ax] += inter[ax] * wk;
            }
            for (int i = 0; i < nn; ++i) {
                T->rb[(n * nn + i) * 2 + 0] += creal(rho[i]) * wk;
                T->rb[(n * nn + i) * 2 + 1] += cimag(rho[i]) * wk;
            }
            for (int m = 0; m < N; ++m) T->bp[n * N + m] += pop[m] * wk;
            rk_step(S, kk, &v, t, rho, stg, tmp, &w);
        }
    }
    for (int i = 0; i < 6; ++i) free(stg[i]);
    free(tmp);
    scratch_free(&w);
    return NULL;
}

int orc_run(const sbe_config *cfg, const sbe_model *model, int64_t k_begin, int64_t k_end,
            int64_t nsteps, int nthreads, double *j_inter, double *j_intra, double *rho_bar,
            double *band_pop, double *rho_final, orc_info *info)
{
    osolver S;
    int rc = solver_init(cfg, model, &S);
    if (rc) return rc;
    int N = S.N, nn = N * N;
    if (cfg->rk_order == SBE_RK_CASH_KARP) {
        solver_free(&S);
        return fail(SBE_ERR_UNSUPPORTED, "the adaptive stepper is restated by orc_run_adaptive");
    }
    if (N > 256) { solver_free(&S); return fail(SBE_ERR_MODEL, "N too large"); }
    int64_t ntot = (int64_t)cfg->npoints[0] * cfg->npoints[1] * cfg->npoints[2];
    if (k_begin < 0) { k_begin = 0; k_end = ntot; }
    int64_t nk = k_end - k_begin;
    if (nk <= 0 || k_end > ntot) { solver_free(&S); return fail(SBE_ERR_RANGE, "bad k range"); }
    if (nsteps < 0 || nsteps > S.L.nt) nsteps = S.L.nt;
    double *kp = malloc(sizeof(double) * 3 * nk), *wk = malloc(sizeof(double) * nk);
    double dV, tw;
    rc = orc_kgrid(cfg, model, k_begin, k_end, kp, wk, &dV, &tw);
    if (rc) { free(kp); free(wk); solver_free(&S); return rc; }
    cplx *rho = calloc((size_t)nk * nn, sizeof(cplx));
    ovel V = {0};
    {
        oscratch w;
        scratch_alloc(&w, N);
        if (S.velocity) {
            vel_alloc(&V, N, (size_t)nk);
            for (int64_t k = 0; k < nk; ++k) {
                ovel v = vel_at(&V, N, k);
                vel_setup(&S, kp + 3 * k, v.E, v.P, v.rho0, &w);
                memcpy(rho + (size_t)k * nn, v.rho0, sizeof(cplx) * nn); /* dmatrix = initMatrix */
            }
        } else {
            for (int64_t k = 0; k < nk; ++k) init_state(&S, kp + 3 * k, rho + (size_t)k * nn, &w);
        }
        scratch_free(&w);
    }
    if (nthreads < 1) nthreads = 1;
    if (nthreads > nk) nthreads = (int)nk;
    othread *th = calloc((size_t)nthreads, sizeof(othread));
    pthread_t *tid = calloc((size_t)nthreads, sizeof(pthread_t));
    int64_t ns = nsteps > 0 ? nsteps : 1;
    for (int r = 0; r < nthreads; ++r) {
        /* ParaRange, utility.h:104-115 */
        int64_t w1 = nk / nthreads, w2 = nk % nthreads;
        int64_t s = w1 * r + (r < w2 ? r : w2), e = s + w1 + (w2 > r ? 1 : 0);
        th[r].S = &S; th[r].kp = kp; th[r].wk = wk; th[r].k0 = s; th[r].k1 = e;
        th[r].nsteps = nsteps; th[r].rho = rho; th[r].V = &V;
        th[r].ji = calloc((size_t)ns * 3, sizeof(double));
        th[r].ja = calloc((size_t)ns * 3, sizeof(double));
        th[r].rb = calloc((size_t)ns * nn * 2, sizeof(double));
        th[r].bp = calloc((size_t)ns * N, sizeof(double));
    }
    struct timespec t0, t1;
    clock_gettime(CLOCK_MONOTONIC, &t0);
    for (int r = 1; r < nthreads; ++r) pthread_create(&tid[r], NULL, thread_main, &th[r]);
    thread_main(&th[0]);
    for (int r = 1; r < nthreads; ++r) pthread_join(tid[r], NULL);
    clock_gettime(CLOCK_MONOTONIC, &t1);
    for (int64_t n = 0; n < nsteps; ++n) {
        for (int ax = 0; ax < 3; ++ax) {
            double si = 0, sa = 0;
            for (int r = 0; r < nthreads; ++r) { si += th[r].ji[3 * n + ax]; sa += th[r].ja[3 * n + ax]; }
            if (j_inter) j_inter[3 * n + ax] = si;
            if (j_intra) j_intra[3 * n + ax] = sa;
        }
        if (rho_bar)
            for (int i = 0; i < nn * 2; ++i) {
                double s = 0;
                for (int r = 0; r < nthreads; ++r) s += th[r].rb[n * nn * 2 + i];
                rho_bar[n * nn * 2 + i] = s;
            }
        if (band_pop)
            for (int m = 0; m < N; ++m) {
                double s = 0;
                for (int r = 0; r < nthreads; ++r) s += th[r].bp[n * N + m];
                band_pop[n * N + m] = s;
            }
    }
    if (rho_final) memcpy(rho_final, rho, sizeof(cplx) * (size_t)nk * nn);
    if (info) {
        info->nt = S.L.nt; info->nk_total = ntot; info->nk_local = nk;
        info->atmin = S.L.atmin; info->atmax = S.L.atmax; info->dt = S.L.dt;
        info->dV = dV; info->total_weight = tw;
        info->seconds = (t1.tv_sec - t0.tv_sec) + 1e-9 * (t1.tv_nsec - t0.tv_nsec);
    }
    for (int r = 0; r < nthreads; ++r) { free(th[r].ji); free(th[r].ja); free(th[r].rb); free(th[r].bp); }
    free(th); free(tid); free(rho); free(kp); free(wk);
    if (S.velocity) vel_free(&V);
    solver_free(&S);
    return SBE_OK;
}

/* ------------------------------------------------------ adaptive stepper */
/* Restatement of the build's adaptive Cash-Karp 5(4) stepper (no reference
 * counterpart; the reference steps with a fixed dt, SBEs.h:596-624). Same
 * model, right-hand side and observables as orc_run; the step sequence is:
 *  - blocks of `block` consecutive k points share one step size; a step is
 *    accepted when err = max over the block's k points and the Bloch
 *    components s_x = 2 Re rho01, s_y = -2 Im rho01, s_z = rho00 - rho11 of
 *    |y5 - y4| / (atol + rtol max(|y|, |y5|)) is <= 1, and the next size is
 *    h min(5, max(0.2, 0.9 err^-1/5)) (after a rejection max(0.2, ...));
 *  - step sizes start at dt and are capped at 32 dt; a step that would end
 *    within 1e-4 of it before a window end is stretched to end there; each
 *    window of `window` samples (the last one shorter) ends exactly on its
 *    sample time, and a shortened last step leaves the proposal unchanged;
 *  - the stage at c = 1 is evaluated at the step end itself (t + h, or the
 *    window end for a shortened last step);
 *  - at every step end the block sums of the eight observables (inter and
 *    intra current x, y; the Bloch vector; s.dhat) are formed; a sample
 *    t_n = atmin + n dt is the Lagrange interpolant through the last six
 *    step ends of its window (ring_at above, Newton form), centred where
 *    the window allows and one-sided at its ends. The Hermite form with time
 *    derivatives at two step ends, and this form with values only, were
 *    compared on the batched workload's worst solve: equal accuracy (5e-7
 *    and 4e-7 at atol 1e-9), without the derivatives' cost.
 * Cash-Karp coefficients: J.R. Cash, A.H. Karp, ACM TOMS 16 (1990) 201. */
static const double ck_c[6] = {0.0, 0.2, 0.3, 0.6, 1.0, 0.875};
static const double ck_a[6][5] = {
    {0, 0, 0, 0, 0},
    {0.2, 0, 0, 0, 0},
    {3. / 40., 9. / 40., 0, 0, 0},
    {0.3, -0.9, 1.2, 0, 0},
    {-11. / 54., 2.5, -70. / 27., 35. / 27., 0},
    {1631. / 55296., 175. / 512., 575. / 13824., 44275. / 110592., 253. / 4096.}};
static const double ck_b[6] = {37. / 378., 0, 250. / 621., 125. / 594., 0, 512. / 1771.};
static const double ck_bhat[6] = {2825. / 27648., 0, 18575. / 48384., 13525. / 55296., 277. / 14336., 0.25};

int orc_cash_karp(double c[6], double a[6][5], double b[6], double bhat[6])
{
    if (c) memcpy(c, ck_c, sizeof ck_c);
    if (a) memcpy(a, ck_a, sizeof ck_a);
    if (b) memcpy(b, ck_b, sizeof ck_b);
    if (bhat) memcpy(bhat, ck_bhat, sizeof ck_bhat);
    return SBE_OK;
}

static void bloch(const cplx *r, double s[3])
{
    s[0] = 2.0 * creal(r[1]);
    s[1] = -2.0 * cimag(r[1]);
    s[2] = creal(r[0]) - creal(r[3]);
}

/* Haldane B-components of Haldane.h:86-92 (v[0] = B0 = d0, v[1..3] = d) at q
 * with their first (g[a]) and second (hs[a][b]) derivatives in k_x, k_y. */
static void haldane_d2(const omodel *om, const double q[3], double v[4], double g[2][4], double hs[2][2][4])
{
    const sbe_model *m = om->m;
    const double c0 = 2. * m->t2 * cos(m->phi0), c3 = 2. * m->t2 * sin(m->phi0);
    memset(v, 0, 4 * sizeof(double));
    memset(g, 0, 8 * sizeof(double));
    memset(hs, 0, 16 * sizeof(double));
    v[3] = m->M0;
    for (int i = 0; i < 3; ++i) {
        const double *va = om->va[i], *vb = om->vb[i];
        const double aa = q[0] * va[0] + q[1] * va[1], ab = q[0] * vb[0] + q[1] * vb[1];
        const double ca = cos(aa), sa = sin(aa), cb = cos(ab), sb = sin(ab);
        v[0] += c0 * cb;
        v[1] += m->t1 * ca;
        v[2] += m->t1 * sa;
        v[3] -= c3 * sb;
        for (int x = 0; x < 2; ++x) {
            g[x][0] -= c0 * vb[x] * sb;
            g[x][1] -= m->t1 * va[x] * sa;
            g[x][2] += m->t1 * va[x] * ca;
            g[x][3] -= c3 * vb[x] * cb;
            for (int y = 0; y < 2; ++y) {
                hs[x][y][0] -= c0 * vb[x] * vb[y] * cb;
                hs[x][y][1] -= m->t1 * va[x] * va[y] * ca;
                hs[x][y][2] -= m->t1 * va[x] * va[y] * sa;
                hs[x][y][3] += c3 * vb[x] * vb[y] * sb;
            }
        }
    }
}

/* dA/dt at t: the derivative of the summed pulses (the offset at atmin is a
 * constant), i.e. minus the consistent field of laser.h's formulas. */
static void laser_rate(const olaser *L, double t, double u[3])
{
    u[0] = u[1] = u[2] = 0.0;
    for (int p = 0; p < L->np; ++p) {
        double A[3], E[3];
        pulse_eval(&L->p[p], t, 1, A, E);
        for (int i = 0; i < 3; ++i) u[i] -= E[i];
    }
}

/* Observables of one k point with Bloch vector s (n = trace) and their time
 * derivatives along ds/dt = f: o = (inter_x, inter_y, intra_x, intra_y, s_x,
 * s_y, s_z, s.dhat) in the GenInterIntraCurrent / GenCurrent conventions of
 * SBEs.h:626-760 (inter carries the total when InterIntra is off). */
static void bloch_observe(const osolver *S, const double k[3], double t, double n, const double s[3],
                          const double f[3], double o[8], double od[8])
{
    double A[3], E[3], u[3];
    laser_eval(&S->L, t, A, E);
    laser_rate(&S->L, t, u);
    const double q[3] = {k[0] + A[0], k[1] + A[1], k[2] + A[2]};
    double v[4], g[2][4], hs[2][2][4];
    haldane_d2(&S->om, q, v, g, hs);
    const double *d = v + 1;
    double td[4], gt[2][4];
    for (int c = 0; c < 4; ++c) {
        td[c] = g[0][c] * u[0] + g[1][c] * u[1];
        for (int x = 0; x < 2; ++x) gt[x][c] = hs[x][0][c] * u[0] + hs[x][1][c] * u[1];
    }
    const double dd = d[0] * d[0] + d[1] * d[1] + d[2] * d[2];
    const double inv = dd > 0 ? 1.0 / dd : 0.0;
    const double ddt = d[0] * td[1] + d[1] * td[2] + d[2] * td[3];
    const double inv_t = -2.0 * inv * inv * ddt;
    const double sd = s[0] * d[0] + s[1] * d[1] + s[2] * d[2];
    const double sd_t = f[0] * d[0] + f[1] * d[1] + f[2] * d[2] + s[0] * td[1] + s[1] * td[2] + s[2] * td[3];
    const int ii = S->cfg->inter_intra != 0;
    for (int x = 0; x < 2; ++x) {
        const double *gd = g[x] + 1, *hd = gt[x] + 1;
        const double tt = s[0] * gd[0] + s[1] * gd[1] + s[2] * gd[2];
        const double tt_t = f[0] * gd[0] + f[1] * gd[1] + f[2] * gd[2] + s[0] * hd[0] + s[1] * hd[1] + s[2] * hd[2];
        const double qq = d[0] * gd[0] + d[1] * gd[1] + d[2] * gd[2];
        const double qq_t = td[1] * gd[0] + td[2] * gd[1] + td[3] * gd[2] + d[0] * hd[0] + d[1] * hd[1] + d[2] * hd[2];
        const double pp = sd * qq * inv, pp_t = (sd_t * qq + sd * qq_t) * inv + sd * qq * inv_t;
        /* Tr(rho J_x) = -(n d_x d0 + s.d_x d); intra = -(n d_x d0 + (s.dh)(dh.d_x d)) */
        if (ii) {
            o[x] = -(tt - pp);
            od[x] = -(tt_t - pp_t);
            o[2 + x] = -(n * g[x][0] + pp);
            od[2 + x] = -(n * gt[x][0] + pp_t);
        } else {
            o[x] = -(n * g[x][0] + tt);
            od[x] = -(n * gt[x][0] + tt_t);
            o[2 + x] = od[2 + x] = 0.0;
        }
    }
    for (int c = 0; c < 3; ++c) {
        o[4 + c] = s[c];
        od[4 + c] = f[c];
    }
    if (dd > 0) {
        const double r = 1.0 / sqrt(dd);
        o[7] = sd * r;
        od[7] = sd_t * r - sd * r * inv * ddt;
    } else {
        o[7] = -s[2];
        od[7] = -f[2];
    }
}

int orc_observe(const sbe_config *cfg, const sbe_model *model, const double k[3], double t, const double *rho,
                const double *drho, double *o, double *od)
{
    osolver S;
    int rc = solver_init(cfg, model, &S);
    if (rc) return rc;
    if (S.N != 2 || S.velocity) { solver_free(&S); return fail(SBE_ERR_UNSUPPORTED, "orc_observe is for N = 2, LengthWannier"); }
    double s[3], f[3];
    bloch((const cplx *)rho, s);
    bloch((const cplx *)drho, f);
    const cplx *r = (const cplx *)rho;
    bloch_observe(&S, k, t, creal(r[0]) + creal(r[3]), s, f, o, od);
    solver_free(&S);
    return SBE_OK;
}

int orc_currents(const sbe_config *cfg, const sbe_model *model, const double k[3], double t, const double *rho,
                 double *inter, double *intra, double *pop)
{
    osolver S;
    int rc = solver_init(cfg, model, &S);
    if (rc) return rc;
    oscratch w;
    scratch_alloc(&w, S.N);
    ovel V = {0};
    if (S.velocity) {
        vel_alloc(&V, S.N, 1);
        vel_setup(&S, k, V.E, V.P, V.rho0, &w);
    }
    currents(&S, k, &V, t, (const cplx *)rho, inter, intra, pop, &w);
    if (S.velocity) vel_free(&V);
    scratch_free(&w);
    solver_free(&S);
    return SBE_OK;
}

typedef struct {
    const osolver *S;
    const double *kp, *wk;
    int64_t b0, b1, nk, block, nsteps, window;
    double *ob; /* [nt][8] block-summed observables (this thread) */
    double *nsum;
    cplx *rho;
    int64_t *acc, *rej;
    int failed;
} oathread;

static void block_observe(const oathread *T, int64_t k0, int64_t k1, const cplx *rho, const cplx *drho, double t,
                          double O[8], double D[8])
{
    for (int i = 0; i < 8; ++i) O[i] = D[i] = 0.0;
    for (int64_t k = k0; k < k1; ++k) {
        const cplx *r = rho + (size_t)(k - k0) * 4, *dr = drho + (size_t)(k - k0) * 4;
        double s[3], f[3], o[8], od[8];
        bloch(r, s);
        bloch(dr, f);
        bloch_observe(T->S, T->kp + 3 * k, t, creal(r[0]) + creal(r[3]), s, f, o, od);
        for (int i = 0; i < 8; ++i) {
            O[i] += T->wk[k] * o[i];
            D[i] += T->wk[k] * od[i];
        }
    }
}

/* Samples of the adaptive stepper: the k sums at the step ends of a block
 * go into a ring of the last six; a sample at tau is the polynomial through
 * the ring's nodes (fewer than six right after a window start) in Newton
 * form on the newest nodes, p(tau) = D0 + (tau - t_n)(D1 + (tau - t_n-1)(D2
 * + ...)), D_k = f[t_n, ..., t_n-k] (backward divided differences, updated
 * in place as nodes arrive). Samples up to node c - 2 are written when node
 * c (c >= 5, from 0) arrives, and all pending ones at the window end. */
typedef struct {
    double t[6], D[6][8]; /* newest node first */
    int cnt;
} oring;

static void ring_push(oring *R, double tn, const double val[8])
{
    for (int i = 0; i < 8; ++i) {
        double prev = R->D[0][i];
        R->D[0][i] = val[i];
        for (int k = 1; k < 6; ++k) {
            const double old = R->D[k][i];
            R->D[k][i] = k <= R->cnt ? (R->D[k - 1][i] - prev) / (tn - R->t[k - 1]) : 0.0;
            prev = old;
        }
    }
    for (int k = 5; k > 0; --k) R->t[k] = R->t[k - 1];
    R->t[0] = tn;
    ++R->cnt;
}

static void ring_at(const oring *R, double tau, double out[8])
{
    for (int i = 0; i < 8; ++i) {
        double acc = R->D[5][i];
        for (int k = 4; k >= 0; --k) acc = acc * (tau - R->t[k]) + R->D[k][i];
        out[i] = acc;
    }
}

static void *adaptive_main(void *arg)
{
    oathread *T = (oathread *)arg;
    const osolver *S = T->S;
    const int N = 2, nn = 4;
    const double dt = S->L.dt, atmin = S->L.atmin;
    const double rtol = S->cfg->rtol, atol = S->cfg->atol, hmax = 32.0 * dt, hmin = 1e-6 * dt;
    oscratch w;
    scratch_alloc(&w, N);
    const int64_t B = T->block;
    cplx *st = calloc((size_t)B * 7 * nn, sizeof(cplx)); /* stages f_1..f_6 and f(t+h, y5) per k */
    cplx *y5 = calloc((size_t)B * nn, sizeof(cplx)), *tmp = calloc((size_t)nn, sizeof(cplx));
    cplx *cur = calloc((size_t)B * nn, sizeof(cplx)), *der = calloc((size_t)B * nn, sizeof(cplx));
    for (int64_t blk = T->b0; blk < T->b1; ++blk) {
        const int64_t k0 = blk * B, k1 = k0 + B < T->nk ? k0 + B : T->nk;
        double h = dt;
        int64_t nacc = 0, nrej = 0;
        for (int64_t k = k0; k < k1; ++k) {
            const cplx *r = T->rho + (size_t)k * nn;
            *T->nsum += T->wk[k] * (creal(r[0]) + creal(r[3]));
        }
        for (int64_t w0 = 0; w0 < T->nsteps && !T->failed; w0 += T->window) {
            const int64_t w1 = w0 + T->window < T->nsteps ? w0 + T->window : T->nsteps;
            double t = atmin + dt * (double)w0;
            const double t_end = atmin + dt * (double)w1;
            int64_t m = w0;
            double O0[8], D0[8], smp[8];
            oring R;
            memset(&R, 0, sizeof R);
            for (int64_t k = k0; k < k1; ++k) {
                const cplx *r = T->rho + (size_t)k * nn;
                rhs(S, T->kp + 3 * k, NULL, t, r, st + ((k - k0) * 7 + 0) * nn, &w);
                memcpy(cur + (k - k0) * nn, r, sizeof(cplx) * nn);
                memcpy(der + (k - k0) * nn, st + (k - k0) * 7 * nn, sizeof(cplx) * nn);
            }
            block_observe(T, k0, k1, cur, der, t, O0, D0);
            for (int i = 0; i < 8; ++i) T->ob[m * 8 + i] += O0[i];
            ring_push(&R, t, O0);
            ++m;
            int64_t iter = 0;
            while (t < t_end) {
                if (++iter > 64 * (w1 - w0) + 1024 || !(h >= hmin)) { T->failed = 1; break; }
                double hh = h < hmax ? h : hmax;
                const int last = t + 1.0001 * hh >= t_end;
                if (last) hh = t_end - t;
                double err = 0.0;
                const double tn = last ? t_end : t + hh;
                for (int64_t k = k0; k < k1; ++k) {
                    const cplx *r = T->rho + (size_t)k * nn;
                    cplx *f = st + (k - k0) * 7 * nn;
                    for (int i = 1; i < 6; ++i) {
                        for (int e = 0; e < nn; ++e) {
                            tmp[e] = r[e];
                            for (int j = 0; j < i; ++j) tmp[e] += hh * ck_a[i][j] * f[j * nn + e];
                        }
                        /* the stage at c = 1 is taken at the step end itself */
                        rhs(S, T->kp + 3 * k, NULL, ck_c[i] == 1.0 ? tn : t + ck_c[i] * hh, tmp, f + i * nn, &w);
                    }
                    cplx *o = y5 + (k - k0) * nn, ev[4];
                    for (int e = 0; e < nn; ++e) {
                        o[e] = r[e];
                        ev[e] = 0;
                        for (int i = 0; i < 6; ++i) {
                            o[e] += hh * ck_b[i] * f[i * nn + e];
                            ev[e] += hh * (ck_b[i] - ck_bhat[i]) * f[i * nn + e];
                        }
                    }
                    double sy[3], s5[3], se[3];
                    bloch(r, sy);
                    bloch(o, s5);
                    bloch(ev, se);
                    for (int c = 0; c < 3; ++c) {
                        const double sc = atol + rtol * fmax(fabs(sy[c]), fabs(s5[c]));
                        const double q = se[c] == 0.0 ? 0.0 : fabs(se[c]) / sc;
                        err = q > err || isnan(q) ? q : err;
                    }
                }
                if (err <= 1.0) {
                    for (int64_t k = k0; k < k1; ++k) {
                        cplx *f = st + (k - k0) * 7 * nn;
                        rhs(S, T->kp + 3 * k, NULL, tn, y5 + (k - k0) * nn, f + 6 * nn, &w);
                        memcpy(der + (k - k0) * nn, f + 6 * nn, sizeof(cplx) * nn);
                    }
                    block_observe(T, k0, k1, y5, der, tn, O0, D0);
                    ring_push(&R, tn, O0);
                    if (R.cnt >= 6) /* samples up to node c - 2 */
                        for (; m < w1; ++m) {
                            const double tm = atmin + dt * (double)m;
                            if (tm > R.t[2]) break;
                            ring_at(&R, tm, smp);
                            for (int i = 0; i < 8; ++i) T->ob[m * 8 + i] += smp[i];
                        }
                    for (int64_t k = k0; k < k1; ++k) {
                        memcpy(T->rho + (size_t)k * nn, y5 + (k - k0) * nn, sizeof(cplx) * nn);
                        cplx *f = st + (k - k0) * 7 * nn;
                        memcpy(f, f + 6 * nn, sizeof(cplx) * nn); /* f(t+h, y5) is the next first stage */
                    }
                    t = tn;
                    ++nacc;
                    const double fac = err > 0.0 ? fmin(5.0, fmax(0.2, 0.9 * pow(err, -0.2))) : 5.0;
                    if (!last) h = hh * fac;
                } else {
                    h = hh * fmax(0.2, 0.9 * pow(err, -0.2));
                    ++nrej;
                }
            }
            if (!T->failed) /* the window end: every pending sample */
                for (; m < w1; ++m) {
                    ring_at(&R, atmin + dt * (double)m, smp);
                    for (int i = 0; i < 8; ++i) T->ob[m * 8 + i] += smp[i];
                }
        }
        if (T->acc) T->acc[blk] = nacc;
        if (T->rej) T->rej[blk] = nrej;
    }
    free(st); free(y5); free(tmp); free(cur); free(der);
    scratch_free(&w);
    return NULL;
}

int orc_run_adaptive(const sbe_config *cfg, const sbe_model *model, int64_t k_begin, int64_t k_end,
                     int64_t block, int64_t window, int nthreads, double *j_inter, double *j_intra,
                     double *rho_bar, double *band_pop, double *rho_final, int64_t *accepted,
                     int64_t *rejected, orc_info *info)
{
    if (cfg->rk_order != SBE_RK_CASH_KARP) return fail(SBE_ERR_CONFIG, "orc_run_adaptive needs SBE_RK_CASH_KARP");
    if (!(cfg->rtol > 0.0) || !(cfg->atol >= 0.0)) return fail(SBE_ERR_CONFIG, "rtol must be > 0, atol >= 0");
    if (cfg->T1 > 1e-16) return fail(SBE_ERR_UNSUPPORTED, "T1 is not restated for the adaptive stepper");
    if (cfg->gauge != SBE_GAUGE_LENGTH_WANNIER)
        return fail(SBE_ERR_UNSUPPORTED, "the adaptive stepper is restated for the LengthWannier gauge");
    osolver S;
    int rc = solver_init(cfg, model, &S);
    if (rc) return rc;
    const int N = S.N, nn = N * N;
    if (N != 2) { solver_free(&S); return fail(SBE_ERR_UNSUPPORTED, "the adaptive stepper is restated for N = 2"); }
    int64_t ntot = (int64_t)cfg->npoints[0] * cfg->npoints[1] * cfg->npoints[2];
    if (k_begin < 0) { k_begin = 0; k_end = ntot; }
    int64_t nk = k_end - k_begin;
    if (nk <= 0 || k_end > ntot) { solver_free(&S); return fail(SBE_ERR_RANGE, "bad k range"); }
    if (block < 1) block = 64;
    const int64_t nt = S.L.nt;
    if (window < 1) window = 2048; /* the product's fixed window grid */
    if (window > nt) window = nt;
    double *kp = malloc(sizeof(double) * 3 * nk), *wk = malloc(sizeof(double) * nk);
    double dV, tw;
    rc = orc_kgrid(cfg, model, k_begin, k_end, kp, wk, &dV, &tw);
    if (rc) { free(kp); free(wk); solver_free(&S); return rc; }
    cplx *rho = calloc((size_t)nk * nn, sizeof(cplx));
    ovel V = {0};
    {
        oscratch w;
        scratch_alloc(&w, N);
        if (S.velocity) {
            vel_alloc(&V, N, (size_t)nk);
            for (int64_t k = 0; k < nk; ++k) {
                ovel v = vel_at(&V, N, k);
                vel_setup(&S, kp + 3 * k, v.E, v.P, v.rho0, &w);
                memcpy(rho + (size_t)k * nn, v.rho0, sizeof(cplx) * nn); /* dmatrix = initMatrix */
            }
        } else {
            for (int64_t k = 0; k < nk; ++k) init_state(&S, kp + 3 * k, rho + (size_t)k * nn, &w);
        }
        scratch_free(&w);
    }
    const int64_t nblk = (nk + block - 1) / block;
    if (nthreads < 1) nthreads = 1;
    if (nthreads > nblk) nthreads = (int)nblk;
    oathread *th = calloc((size_t)nthreads, sizeof(oathread));
    pthread_t *tid = calloc((size_t)nthreads, sizeof(pthread_t));
    double *nsums = calloc((size_t)nthreads, sizeof(double));
    for (int r = 0; r < nthreads; ++r) {
        int64_t w1 = nblk / nthreads, w2 = nblk % nthreads;
        int64_t s = w1 * r + (r < w2 ? r : w2), e = s + w1 + (w2 > r ? 1 : 0);
        th[r].S = &S; th[r].kp = kp; th[r].wk = wk; th[r].b0 = s; th[r].b1 = e; th[r].nk = nk;
        th[r].block = block; th[r].nsteps = nt; th[r].window = window; th[r].rho = rho;
        th[r].acc = accepted; th[r].rej = rejected;
        th[r].ob = calloc((size_t)nt * 8, sizeof(double));
        th[r].nsum = nsums + r;
    }
    struct timespec t0, t1;
    clock_gettime(CLOCK_MONOTONIC, &t0);
    for (int r = 1; r < nthreads; ++r) pthread_create(&tid[r], NULL, adaptive_main, &th[r]);
    adaptive_main(&th[0]);
    for (int r = 1; r < nthreads; ++r) pthread_join(tid[r], NULL);
    clock_gettime(CLOCK_MONOTONIC, &t1);
    int failed = 0;
    double nsum = 0.0;
    for (int r = 0; r < nthreads; ++r) { failed |= th[r].failed; nsum += nsums[r]; }
    for (int64_t n = 0; n < nt; ++n) {
        double O[8] = {0, 0, 0, 0, 0, 0, 0, 0};
        for (int r = 0; r < nthreads; ++r)
            for (int i = 0; i < 8; ++i) O[i] += th[r].ob[n * 8 + i];
        if (j_inter) { j_inter[3 * n] = O[0]; j_inter[3 * n + 1] = O[1]; j_inter[3 * n + 2] = 0.0; }
        if (j_intra) { j_intra[3 * n] = O[2]; j_intra[3 * n + 1] = O[3]; j_intra[3 * n + 2] = 0.0; }
        if (rho_bar) { /* rho = (n + s.sigma)/2 */
            double *o = rho_bar + n * 8;
            o[0] = 0.5 * (nsum + O[6]); o[1] = 0.0;
            o[2] = 0.5 * O[4];          o[3] = -0.5 * O[5];
            o[4] = 0.5 * O[4];          o[5] = 0.5 * O[5];
            o[6] = 0.5 * (nsum - O[6]); o[7] = 0.0;
        }
        if (band_pop) { band_pop[2 * n] = 0.5 * (nsum - O[7]); band_pop[2 * n + 1] = 0.5 * (nsum + O[7]); }
    }
    if (rho_final) memcpy(rho_final, rho, sizeof(cplx) * (size_t)nk * nn);
    if (info) {
        info->nt = nt; info->nk_total = ntot; info->nk_local = nk;
        info->atmin = S.L.atmin; info->atmax = S.L.atmax; info->dt = S.L.dt;
        info->dV = dV; info->total_weight = tw;
        info->seconds = (t1.tv_sec - t0.tv_sec) + 1e-9 * (t1.tv_nsec - t0.tv_nsec);
    }
    for (int r = 0; r < nthreads; ++r) free(th[r].ob);
    free(nsums); free(th); free(tid); free(rho); free(kp); free(wk);
    solver_free(&S);
    return failed ? fail(SBE_ERR_NAN, "adaptive step size underflow") : SBE_OK;
}
