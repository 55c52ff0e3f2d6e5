/*
 * sbe_c.h - C ABI of the MI355X Semiconductor Bloch Equation propagator.
 *
 * This is the drop-in boundary for Antelope's per-k-point time loop. One
 * context owns one contiguous block of the flattened k-grid (exactly what one
 * MPI rank owns in the reference) and keeps its density matrices resident in
 * HBM for the whole pulse. The caller above (the C++ driver, an MPI/RCCL
 * layer, the serving loop or a ctypes test) only sees plain structs, plain
 * pointers and integer status codes: no HIP or torch types cross it.
 *
 * Reference interfaces replaced (paths relative to the reference repo):
 *   sbe_create        SBEs::SBEs(cfg, gauge, idx_start, idx_end)
 *                     src/GeneralModel/SBEs.h:102,125-430 (state init, tableau,
 *                     dephasing matrix) + momaxis(...) momaxis.h:52-61 +
 *                     laser::Initialize laser.h:278-298
 *   sbe_run           the time x k loop src/GeneralModel/antelope.cpp:297-338,
 *                     i.e. SBEs::GenInterIntraCurrent SBEs.h:694-760 (or
 *                     SBEs::GenCurrent SBEs.h:626-693) sampled before each
 *                     SBEs::RunSBEs SBEs.h:596-624 step
 *   sbe_get_series    the rank-local accumulators inter_rad / intra_rad /
 *                     density_matrix_integrated, antelope.cpp:196-205,313-333,
 *                     before the MPI_Reduce of antelope.cpp:482-512 and before
 *                     the x dV / x 1/Ntotal scaling of the writers :514-548
 *   sbe_get_state     the public SBEs::dmatrix (SBEs.h:38, AoS [k][N*N]) and
 *                     SBEs::WannierToHamiltonian SBEs.h:927-941
 *   sbe_get_info      SBEs::kmesh->{Ntotal,Npartial,dV,total_weight}
 *                     (momaxis.h:40-49) and SBEs::fpulses->{Nt,dt,atmin}
 *                     (laser.h:256-264)
 *   sbe_laser_eval    laser::avlaser / laser::elaser laser.h:301-322
 *   sbe_destroy       SBEs::~SBEs SBEs.h:563-594
 *   sbe_last_error    replaces std::cerr + std::exit(EXIT_FAILURE) error paths
 *                     (e.g. SBEs.h:292-308, laser.h:81-85): nothing here exits
 *                     the process.
 *   sbe_query_footprint  (no reference counterpart; used for admission control
 *                     by the serving loop, which admits solves by HBM bytes)
 *
 * The material model (BaseMaterial.h:22-78 virtual GenHamiltonian, GenDipole,
 * GenJMatrix, GenBrillouinzone) cannot be a virtual call per k on the GPU, so
 * it crosses the boundary as data: sbe_model below.
 *
 * Units are atomic units throughout (Hartree, bohr, a.u. of time), angles in
 * radians. Front ends that read inputParam.cfg convert the reference's degree
 * and Angstrom inputs before filling these structs (SBEs.h:531-537,
 * Haldane.h:55).
 */
#ifndef ANTELOPE_SBE_C_H
#define ANTELOPE_SBE_C_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define SBE_ABI_VERSION 1

/* ---- status codes: 0 ok, negative = failure (message via sbe_last_error) */
enum {
    SBE_OK = 0,
    SBE_ERR_CONFIG = -1,      /* invalid sbe_config                          */
    SBE_ERR_MODEL = -2,       /* invalid sbe_model                           */
    SBE_ERR_OOM = -3,         /* device or host allocation failed            */
    SBE_ERR_HIP = -4,         /* HIP runtime error                           */
    SBE_ERR_NAN = -5,         /* non-finite value detected in the state      */
    SBE_ERR_STATE = -6,       /* call out of order (e.g. non-contiguous run) */
    SBE_ERR_RANGE = -7,       /* index / size out of range                   */
    SBE_ERR_UNSUPPORTED = -8, /* valid in the reference, not built here      */
    SBE_ERR_NODEVICE = -9     /* no GPU visible                              */
};

/* Pulse envelopes, laser.h:18-24, plus the sin^2 envelope (cos^2 of the
 * cos-n family, twidth = ncycles*T/2) that the reference lacks. */
enum {
    SBE_ENV_GAUSS = 0,
    SBE_ENV_GAUSS_CLIPPED = 1,
    SBE_ENV_COS4 = 2,
    SBE_ENV_COS4_CLIPPED = 3,
    SBE_ENV_SIN2 = 4
};

/* GaugeType, constant.h:41-46 */
enum {
    SBE_GAUGE_LENGTH_WANNIER = 0,
    SBE_GAUGE_LENGTH_HAMILTONIAN = 1,  /* unimplemented in the reference too */
    SBE_GAUGE_VELOCITY_HAMILTONIAN = 2
};

/* k-grid kind: the reference's endpoint-inclusive trapezoid grid
 * (momaxis.h:147-186,349-383) or a centred Monkhorst-Pack grid (equal
 * weights, frac = (i+1/2)/N - 1/2). */
enum { SBE_GRID_REFERENCE = 0, SBE_GRID_MP = 1 };

/* InitialValueType, BaseMaterial.h:10-15 (Custom is not supported) */
enum { SBE_INIT_UNIFORM_VALENCE = 0, SBE_INIT_FERMI_DIRAC = 1 };

/* Arithmetic of the propagated state. The k-sums of currents and
 * populations and the laser tables are FP64 in every mode. FP32: 2-band
 * models, FP32 state (Kahan-compensated). SPLIT16: Wannier models, the
 * Fourier sums H(k+A), D, J on FP16 MFMA with split (hi + lo) operands and
 * FP32 accumulation; state, eigensolver and currents stay FP64. */
enum { SBE_PREC_FP64 = 0, SBE_PREC_FP32 = 1, SBE_PREC_SPLIT16 = 2 };

/* Time stepper: RKorder = 4 (classical RK4) or 5 (fixed-step 6-stage
 * Fehlberg weights) as in SBEs.h:395-416. SBE_RK_CASH_KARP is the adaptive
 * embedded 5(4) pair of Cash & Karp (no reference counterpart; 2-band
 * models with T1 off): every block of 64 k points chooses its own steps
 * under rtol/atol, and the series is still sampled at t_n = atmin + n dt,
 * from the stepper's dense output, so it has the layout of a fixed-step run.
 * dt then only sets the sample spacing (and the first step size). The
 * stepper lands exactly on the sample time at the end of every window of
 * 2048 samples and at every sbe_run step_end; runs split at multiples of
 * 2048 samples give the same bits as one run, other splits agree to the
 * tolerance. */
enum { SBE_RK4 = 4, SBE_RK5_FEHLBERG = 5, SBE_RK_CASH_KARP = 45 };

/* Electric field convention. REFERENCE reproduces laser.h:217-224, whose
 * envelope-derivative term has the opposite sign to -dA/dt; CONSISTENT uses
 * E = -dA/dt. Only matters when dipoles are non-zero and for the laser file. */
enum { SBE_EFIELD_REFERENCE = 0, SBE_EFIELD_CONSISTENT = 1 };

/* Models (material classes of Materials/ that this build carries as data).
 * The analytic tight-binding classes take their parameters in
 * sbe_model.params, already in Hartree / bohr / rad (the config readers
 * convert Angstrom as the classes do), and run on the N-band path:
 *   WILSON_MASS   Materials/WilsonMass.h   N=2 Nval=1  {t, delta, mu, a0, a1}
 *   SQUARE_SBAND  Materials/SquareSBand.h  N=2 Nval=1  {t1, M0, a0}
 *   KANE_MELE     Materials/KaneMele.h     N=4 Nval=2  {t1, t2, tR, M0, phi0, a0}
 *   XYH_LATTICE   Materials/XYHlattice.h   N=4 Nval=2  {tl, tr, lambda, a0}
 *   TMDC          Materials/TMDCs.h        N=3 Nval=1 (N=6 Nval=2 with spin-orbit)
 *                 {a0, soc, e1, e2, t0, t1, t2, t11, t12, t22, r0, r1, r2, r11,
 *                  r12, u0, u1, u2, u11, u12, u22}; sbe_tmdc_params
 *                 (sbe_frontend.h) fills them from the class's GGA table.
 *   HALDANE_2L    Materials/Haldane2L.h    N=4 Nval=2  {t1, t2, M0, phi0 (top),
 *                 t1, t2, M0, phi0 (bottom), a0, same, s0, s1, h0, h1, h2, h3}:
 *                 same = 1 when no bottom key was given, s_l = +-1 the B2 sign
 *                 of layer l, h the interlayer block of the polytype
 *   BI2SE3_SURF   Materials/Bi2Se3surf.h   N=2 Nval=1  {} (hard-coded)
 *   MULTI_WSM     Materials/MultiWSM.h     N=2 Nval=1  {n, v, alphan, phi0, a0}
 *   WDSM_TAAS     Materials/WDSM_TaAs.h    N=4 Nval=1  {t, delta, my, mz, a0} */
enum {
    SBE_MODEL_HALDANE = 1, SBE_MODEL_WANNIER = 2, SBE_MODEL_WILSON_MASS = 3, SBE_MODEL_SQUARE_SBAND = 4,
    SBE_MODEL_KANE_MELE = 5, SBE_MODEL_XYH_LATTICE = 6, SBE_MODEL_TMDC = 7, SBE_MODEL_HALDANE_2L = 8,
    SBE_MODEL_BI2SE3_SURF = 9, SBE_MODEL_MULTI_WSM = 10, SBE_MODEL_WDSM_TAAS = 11
};
#define SBE_MODEL_NPARAMS 24

/* Basis for sbe_get_state. */
enum { SBE_BASIS_ORBITAL = 0, SBE_BASIS_BAND = 1 };

/* One pulse, the arguments of Pulse::Pulse (laser.h:68-141). */
typedef struct sbe_pulse {
    double E0;              /* peak field (a.u.)                          */
    double w0;              /* carrier frequency (a.u.)                   */
    double ellip;           /* ellipticity                                */
    double ncycles;         /* cycles (FWHM cycles for gauss)             */
    double cep;             /* carrier-envelope phase (rad)               */
    double t0;              /* pulse centre (a.u.)                        */
    double phix;            /* semi-major axis azimuth (rad)              */
    double ncycles_clipped; /* flat-top cycles of the *_CLIPPED envelopes */
    double thetaz;          /* propagation polar angle (rad)              */
    double phiz;            /* propagation azimuth (rad)                  */
    int32_t envelope;       /* SBE_ENV_*                                  */
    int32_t reserved;
} sbe_pulse;

/* Run configuration: the `calc` and `laser` groups of inputParam.cfg
 * (SBEs.h:432-561) after unit conversion. */
typedef struct sbe_config {
    int32_t abi_version;    /* must be SBE_ABI_VERSION                    */
    int32_t gauge;          /* SBE_GAUGE_*                                */
    int32_t npoints[3];     /* calc.Npoints, missing axes = 1             */
    int32_t rk_order;       /* SBE_RK4 | SBE_RK5_FEHLBERG | SBE_RK_CASH_KARP */
    double dt;              /* calc.dt (reference default 0.1)            */
    double T1;              /* calc.T1, <= 1e-16 disables                 */
    double T2;              /* calc.T2, <= 1e-16 disables                 */
    double ksfactor[3];     /* calc.ksfactor (scales BZ axes per row)     */
    int32_t inter_intra;    /* calc.InterIntra                            */
    int32_t grid;           /* SBE_GRID_*                                 */
    int32_t has_bz_axes;    /* calc.BZaxes given -> overrides model axes  */
    int32_t has_bz_origin;  /* calc.BZorigin given                        */
    double bz_axes[9];
    double bz_origin[3];
    int32_t init_type;      /* SBE_INIT_*                                 */
    int32_t precision;      /* SBE_PREC_*                                 */
    double fermi_energy;    /* Ef (Hartree) for SBE_INIT_FERMI_DIRAC      */
    int32_t npulses;
    int32_t efield_mode;    /* SBE_EFIELD_*                               */
    const sbe_pulse *pulses;
    double offset_before;   /* laser.offset[0]                            */
    double offset_after;    /* laser.offset[1]                            */
    double rtol, atol;      /* SBE_RK_CASH_KARP tolerances                */
} sbe_config;

/* Material model as data. Haldane: Materials/Haldane.h:46-170. Wannier
 * tight-binding: the exact Fourier sums of Materials/Wannier90.h:297-387
 * (GenHamiltonianPrimitive / GenDipolePrimitive / GenJMatrixPrimitive). */
typedef struct sbe_model {
    int32_t kind;           /* SBE_MODEL_*                                 */
    int32_t nband;          /* N (2 for Haldane)                           */
    int32_t nval;           /* number of valence bands                     */
    int32_t is_dipole_zero; /* 1: no E.D term; Wannier phases use R+tau_n-tau_m */
    /* Haldane (a0 in bohr, t1/t2/M0 in Hartree, phi0 in rad) */
    double a0, t1, t2, phi0, M0;
    /* Wannier */
    int32_t nrpts;          /* number of lattice vectors R                 */
    int32_t reserved;
    double lattice[9];      /* rows a1, a2, a3 (bohr)                      */
    const int32_t *rvec;    /* [nrpts][3] integer R = n1 a1 + n2 a2 + n3 a3 */
    const double *rweight;  /* [nrpts] 1/degeneracy                        */
    const double *ham;      /* [nrpts][N][N][re,im] <0m|H|Rn> (Hartree)    */
    const double *pos;      /* [nrpts][3][N][N][re,im] <0m|r_i|Rn> (bohr)  */
    /* analytic tight-binding classes (layouts above) */
    double params[SBE_MODEL_NPARAMS];
} sbe_model;

/* Read-only facts about a context. */
typedef struct sbe_info {
    int64_t nt;             /* number of time samples / steps (laser.h:297) */
    int64_t nk_total;       /* Ntotal                                      */
    int64_t k_begin, k_end; /* owned block of the flattened k index        */
    int32_t nband;
    int32_t nstage;         /* RK stages per step                          */
    double dt, atmin, atmax;
    double dV;              /* |Volume| / Ntotal (momaxis.h:238)           */
    double total_weight;    /* sum of the owned k weights                  */
    int64_t steps_done;     /* steps already propagated                    */
    int64_t device_bytes;   /* device memory held by the context           */
} sbe_info;

typedef struct sbe_ctx sbe_ctx;

/* Create a solver for k indices [k_begin, k_end) of the flattened grid
 * (k_begin = k_end = -1 means the whole grid). Builds the k-grid, the
 * laser tables and the initial state on the device. */
int sbe_create(const sbe_config *cfg, const sbe_model *model,
               int64_t k_begin, int64_t k_end, sbe_ctx **out);

/* Propagate steps [step_begin, step_end). Steps must be run in order:
 * step_begin must equal the number of steps already done. For each step n
 * the currents, populations and the weighted density matrix are sampled at
 * t_n = atmin + n*dt from the state before the step (antelope.cpp:299-334).
 * Asynchronous with respect to the host; sbe_synchronize waits. */
int sbe_run(sbe_ctx *ctx, int64_t step_begin, int64_t step_end);
int sbe_synchronize(sbe_ctx *ctx);
/* 1 when no work of the context is pending on the GPU, else 0 (non-blocking;
 * lets a scheduler overlap many contexts). */
int sbe_is_idle(const sbe_ctx *ctx);

/* Copy the per-step k-sums of this context's k block to host buffers.
 * Each pointer may be NULL. Values are sum_k w_k x (...) before any dV or
 * 1/Ntotal factor, so partial sums from several contexts (ranks) add up.
 *   j_inter  [nt][3]          Re sum w Tr(rho J) inter-band part (or the total
 *                             current if inter_intra == 0, as GenCurrent)
 *   j_intra  [nt][3]          intra-band part (0 if inter_intra == 0)
 *   rho_bar  [nt][N][N][2]    sum w rho in the orbital (Wannier) basis
 *   band_pop [nt][N]          sum w (U^dag rho U)_mm at k + A(t_n) */
int sbe_get_series(sbe_ctx *ctx, double *j_inter, double *j_intra,
                   double *rho_bar, double *band_pop);

/* Copy the current density matrices [k_local][N][N][re,im] (row-major,
 * the reference's AoS dmatrix layout). basis = SBE_BASIS_BAND returns
 * U^dag rho U at k + A(t_now) (WannierToHamiltonian). */
int sbe_get_state(sbe_ctx *ctx, int32_t basis, double *rho);

int sbe_get_info(const sbe_ctx *ctx, sbe_info *info);

/* Device bytes a context for this problem would allocate. */
int sbe_query_footprint(const sbe_config *cfg, const sbe_model *model,
                        int64_t k_begin, int64_t k_end, size_t *bytes);

/* Laser A(t), E(t) of a configuration (the k-independent tables fed to the
 * kernels), for the laser output file and for checks. */
int sbe_laser_eval(const sbe_config *cfg, int64_t n, const double *t,
                   double *A /*[n][3]*/, double *E /*[n][3]*/,
                   double *atmin, int64_t *nt);

/* ---- batched solves (no reference counterpart: the reference runs one
 * solve per MPI job). n independent 2-band solves, each with its own
 * configuration, model, k grid and pulse length, advanced together: every
 * launch of the step kernel covers the unfinished solves of one of two groups
 * (of about equal work, k points x samples; their launch windows overlap), and
 * solves that are done drop out of the launch grid. All solves must share rk_order and precision. Results are read
 * per solve exactly as from a single context. */
typedef struct sbe_batch sbe_batch;
int sbe_batch_create(int32_t n, const sbe_config *cfgs, const sbe_model *models, sbe_batch **out);
/* Advance every unfinished solve by up to max_steps steps (<= 0: to the end). */
int sbe_batch_run(sbe_batch *b, int64_t max_steps);
int sbe_batch_synchronize(sbe_batch *b);
/* Every solve back to its initial state (as after sbe_batch_create), with
 * no allocation or table rebuild: a batch can be run again. */
int sbe_batch_reset(sbe_batch *b);
int sbe_batch_get_info(const sbe_batch *b, int32_t i, sbe_info *info);
int sbe_batch_get_series(sbe_batch *b, int32_t i, double *j_inter, double *j_intra, double *rho_bar,
                         double *band_pop);
/* Time during which at least one batch step launch ran since the last call (the
 * union of their HIP-event intervals: the two groups' launches overlap), and
 * the number of launches. */
int sbe_batch_get_timing(sbe_batch *b, double *ms, int64_t *launches);
const char *sbe_batch_last_error(const sbe_batch *b);
void sbe_batch_destroy(sbe_batch *b);

/* Step statistics (no reference counterpart). The 2-band path advances its
 * k points in blocks of k_per_block (one wave each); with the adaptive
 * Cash-Karp stepper every block chooses its own steps. accepted / rejected
 * sum the step attempts over the blocks; max_accepted / min_accepted are the
 * extremes of one block. Fixed-step solves report steps_done per block and
 * no rejections. */
typedef struct sbe_step_stats {
    int64_t blocks, k_per_block;
    int64_t accepted, rejected;
    int64_t max_accepted, min_accepted;
} sbe_step_stats;
int sbe_get_step_stats(sbe_ctx *ctx, sbe_step_stats *out);
int sbe_batch_get_step_stats(sbe_batch *b, int32_t i, sbe_step_stats *out);

/* Measurement (no reference counterpart): when enabled, HIP events are
 * recorded on the context's stream around every kernel sbe_run launches.
 * sbe_get_timing waits for them and returns the summed durations since the
 * previous call (then resets): ms[0] / launches[0] for the step kernels,
 * ms[1] / launches[1] for the series reductions. */
int sbe_set_timing(sbe_ctx *ctx, int32_t enable);
int sbe_get_timing(sbe_ctx *ctx, double *ms /*[2]*/, int64_t *launches /*[2]*/);

void sbe_destroy(sbe_ctx *ctx);

/* ---- launch-plan autotuner (no reference counterpart). A context's plan has
 * choices that do not change the method: k points per lane of the 2-band
 * kernels, steps per launch of the fixed-step kernels, the N-band stage
 * kernels (MFMA register tiles for N <= 16, or LDS) and the N-band Fourier
 * sums (separable passes on product grids, or the dense GEMM). Each has a
 * default rule. sbe_autotune times every candidate of the problem's path on
 * throwaway contexts (trial_steps steps each, after a warm-up) and caches the
 * fastest for this device and problem class (path, precision, tableau, bands,
 * gauge, k points to the nearest power of two); sbe_create then plans
 * contexts of that class with it (environment switches still win).
 * Steps per launch never change results; k points per lane and the N-band
 * choices change the order of sums (results agree to rounding). */
typedef struct sbe_tuning {
    int32_t kpt;          /* 2-band k points per lane (1, 2; 0 = default rule)      */
    int32_t chunk_steps;  /* fixed-step steps per launch cap (0 = by memory)         */
    int32_t nb_kernels;   /* N-band stage kernels: 0 default, 1 tiles, 2 LDS         */
    int32_t nb_sums;      /* N-band Fourier sums: 0 default, 1 separable, 2 dense    */
    int32_t tuned;        /* 1: from the cache (set or measured), 0: default rule    */
    int32_t candidates;   /* plans timed by sbe_autotune                            */
    double ms_per_step;   /* the chosen plan's trial time per step (sbe_autotune)    */
} sbe_tuning;
int sbe_autotune(const sbe_config *cfg, const sbe_model *model, int64_t trial_steps, sbe_tuning *out);
/* The cached plan of this problem class on the current device (tuned = 0 and
 * zeros: the default rules apply). */
int sbe_tuning_get(const sbe_config *cfg, const sbe_model *model, sbe_tuning *out);
/* Store a plan for this problem class on the current device (as sbe_autotune would). */
int sbe_tuning_set(const sbe_config *cfg, const sbe_model *model, const sbe_tuning *tuning);
void sbe_tuning_clear(void);
/* The same for a batch (sbe_batch_create's arguments): the candidates are
 * timed on throwaway batches over trial_steps samples (adaptive batches: one
 * or two k points per lane; fixed-step members always run one), and the
 * fastest is cached for every member's problem class. */
int sbe_batch_autotune(int32_t n, const sbe_config *cfgs, const sbe_model *models, int64_t trial_steps,
                       sbe_tuning *out);

/* ---- propagation with no device allocation. No counterpart in the
 * reference: for callers that cannot allocate device memory (code under HIP
 * graph capture, serving loops with a fixed memory budget). The 2-band
 * (Haldane) fixed-step path of sbe_create + sbe_run (FP64 state, RK4 or
 * RK5, LengthWannier gauge) over caller-owned device buffers:
 *   kpts    [4][nk_stride] FP64: k_x, k_y, k_z, weight; padding k points
 *           (index >= nk) must carry weight 0
 *   state   [4][nk_stride] FP64 Bloch vector (n, s_x, s_y, s_z) per k point,
 *           rho = (n I + s.sigma)/2; read, propagated, written back.
 *           init_state != 0 first writes the initial state of sbe_create.
 *   series  [step_end - step_begin][8] FP64, written: per step n the k sums
 *           before the step, sum_k w (j_inter_x, j_inter_y, j_intra_x,
 *           j_intra_y, s_x, s_y, s_z, s.d/|d|) at k + A(t_n); sbe_get_series
 *           forms j_inter, j_intra, rho_bar and band_pop from the same values.
 * The laser rows and the partial sums of the k reduction use one static
 * device array of SBE_NOALLOC_SCRATCH_BYTES per device (held until the
 * process ends); k sets whose partial rows of one step exceed it are
 * rejected with SBE_ERR_RANGE (about 16 M k points). The call only enqueues
 * kernels on `stream` (a hipStream_t; NULL = the default stream) and returns;
 * the caller synchronizes. Calls that overlap on one GPU race on the shared
 * array. Results equal sbe_run's to rounding (the laser rows are evaluated
 * on the device instead of the host). */
#define SBE_NOALLOC_SCRATCH_BYTES (16u << 20)
#define SBE_NOALLOC_K_ALIGN 512
int sbe_run_haldane_noalloc(const sbe_config *cfg, const sbe_model *model, const double *kpts, double *state,
                            int64_t nk, int64_t nk_stride, int64_t step_begin, int64_t step_end,
                            int32_t init_state, double *series, void *stream);

/* ---- materials given as code (BaseMaterial.h:22-78). The reference's
 * materials are classes with virtual GenHamiltonian(complex*, k) and
 * GenDipole(complex**, k); the GPU path takes tight-binding blocks. This
 * samples a material's callbacks on the (2 nmax_i + 1)-point grid of each
 * reciprocal axis and inverts the discrete Fourier transform into an
 * SBE_MODEL_WANNIER model with hoppings |n_i| <= nmax_i along a_i. Models
 * whose hoppings reach no further are reproduced exactly (to rounding); the
 * largest deviation of the sampled H(k) (and D(k)) from the callbacks at 16
 * k points off the grid is returned in *max_dev, so a caller sees a range
 * too short. Callbacks write row-major N x N complex matrices as
 * [re, im] pairs: gen_h writes H(k) (Hartree) to out[N][N][2], gen_d the
 * Wannier-basis dipole D_i(k) (bohr) to out[3][N][N][2]. is_dipole_zero
 * takes no gen_d (NULL) and uses the offsets tau ([N][3] bohr, NULL = 0):
 * H_mn(k) = sum_R h_mn(R) e^{ik.(R + tau_n - tau_m)} (Wannier90.h:297-310).
 * The model's arrays are owned by the library until
 * sbe_free_sampled_material. Host only: no device is touched. */
typedef void (*sbe_gen_matrix_fn)(void *user, const double k[3], double *out);
int sbe_sample_material(int32_t nband, int32_t nval, int32_t is_dipole_zero, const double lattice[9],
                        const double *tau, const int32_t nmax[3], sbe_gen_matrix_fn gen_h,
                        sbe_gen_matrix_fn gen_d, void *user, sbe_model *out, double *max_dev);
void sbe_free_sampled_material(sbe_model *model);

/* Last error of a context (ctx may be NULL: last error of this thread,
 * e.g. a failed sbe_create). Never NULL. */
const char *sbe_last_error(const sbe_ctx *ctx);

#ifdef __cplusplus
}
#endif

#endif /* ANTELOPE_SBE_C_H */
