// nband.hpp - N-band (Wannier) path: kernel arguments and launchers.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

namespace sbe {

// Matrix kinds an evaluation of the Fourier sums can produce.
enum : int { kMatHtilde = 0, kMatH0 = 1, kMatJx = 2, kMatJy = 3, kMatJz = 4 };

constexpr int kNbMaxBands = 32;

struct NbArgs {
    int N;                 // bands
    int nr;                // lattice vectors R
    int nval;
    int fermi;             // Fermi-Dirac initial occupations
    double fermi_energy;
    int inter_intra;
    int deph;              // any dephasing on (T1 or T2)
    double gam1, gam2;     // 1/T1, 1/T2 (0 = off)
    int64_t nk, nk_stride;
    int nv;                // values per sample row: 6 + N + 2 N^2
    int sample_waves;      // rows of the per-step partial sums
    const double *kpts;    // [nk_stride][3]
    const double *weight;  // [nk_stride]
    const double *rcart;   // [nr][3] cartesian R (bohr)
    const double *rweight; // [nr] 1/degeneracy
    const double2 *ham;    // [nr][N][N]
    const double2 *pos;    // [nr][3][N][N]
    // FP16-split Fourier sums (SBE_PREC_SPLIT16): halves as raw 16-bit words
    int split16;
    int nr_pad;            // nr rounded up to the MFMA K (32)
    int64_t bcols;         // columns of the B planes (4 N^2 rounded up to 64)
    const uint16_t *p16;   // e^{ik.R}: planes [re_hi, re_lo, im_hi, im_lo][nk_stride][nr_pad]
    uint16_t *b16;         // B of one evaluation: planes [re_hi, re_lo, im_hi, im_lo][bcols][nr_pad]
    double2 *b64;          // FP64 form: B of one evaluation [nr_pad][bcols]
    double hmax, pmax, rmax;  // max |H(R)|, max |r(R)|, max |R_a| (scales of B)
    int velocity;          // VelocityHamiltonian gauge (band basis at k)
    // analytic tight-binding class (SBE_MODEL_WILSON_MASS ... SBE_MODEL_TMDC):
    // H and J come from analytic.hpp instead of the Fourier sums; 0 = Wannier
    int model_kind;
    double mparams[24];
    int tile16;            // N <= 16: stage and sample kernels with the matrices in MFMA registers
    // Separable Fourier sums (FP64 form, product k grids): R = n1 a1 + n2 a2 + n3 a3 on the
    // lattice box lo_i <= n_i < lo_i + M_i, theta_i = k.a_i taking K_i values per axis
    int box;
    int boxM[3], boxK[3];
    const int32_t *box_r;     // [M1][M2][M3] index into the R list, -1 where the box has no R
    const double2 *boxE[3];   // [K_i][M_i] e^{i n theta}, n = lo_i + m
    const int32_t *box_kinv;  // [K1][K2][K3] k index of this block, -1 where absent
    double2 *box_b;           // B on the box [M1 M2 M3][bcols]
    double2 *box_t1;          // first pass  [M1][M2][K3][bcols]
    double2 *box_t2;          // second pass [M1][K2][K3][bcols]
    // isDipoleZero (Wannier90.h:180-194): the sums take the phase of R + tau_n - tau_m
    // (tau_n = Re <0n|r|0n>) and no E.D term
    int dzero;
    double tau[kNbMaxBands][3];
};

// One evaluation of the Fourier sums at k + A(t) with field E(t).
struct NbEval {
    double A[3], E[3];
    int nmat;
    int mats[5];
};

// One RK stage: coefficients already multiplied by dt.
struct NbStage {
    int stage, nstage, last;
    int ureuse;   // the stored eigenvectors belong to this stage's H (same time)
    double a[8];  // dt a_{stage,j}
    double b[8];  // dt b_j (used when last)
};

// One whole RK4 step of the LengthWannier gauge for N <= 16 on a product k grid
// with separable Fourier sums (nb16_step_kernel): the last pass of the sums
// (along a1) runs inside the step kernel, from the second pass's output of
// each evaluation time, so the matrices never reach HBM. t2_* are that output,
// [M1][K2][K3][bcols] with the evaluation's matrices side by side in the columns.
constexpr int kNbFusedWaves = 8;  // k points along a1 per block (one wave each); 4 on small grids
constexpr int kNbFusedMaxM1 = 7;  // lattice vectors along a1 staged in LDS
struct NbFused {
    const double2 *t2_prev;  // H~, H at t_n
    const double2 *t2_j;     // J_x, J_y, J_z at t_n
    const double2 *t2_half;  // H~, H at t_n + dt/2
    const double2 *t2_end;   // H~, H at t_n + dt
    double *lam;             // [nk_stride][16] band energies at the stored eigenvectors' time
    int u_valid;             // the stored eigenvectors belong to H(t_n)
    int lam_valid;           // lam belongs to them
    double dt;
    int64_t step_local;      // row block of the partial sums
    int g1;                  // blocks per (k2, k3) column: ceil(K1 / waves)
    int nblocks;             // g1 K2 K3: partial rows per step
    int waves;               // k points along a1 per block: 8, or 4 where 8 gives fewer blocks than CUs
};
// bytes of the block's LDS (N bands, M1 lattice vectors along a1, nv values per sample row,
// waves per block)
size_t nb16_step_lds_bytes(int N, int M1, int nv, int waves);
hipError_t nb16_launch_step(const NbArgs &a, const NbFused &f, double2 *rho, double2 *U, double *partials,
                            hipStream_t s);
// the separable sums of one evaluation up to the second pass: t2 [M1][K2][K3][bcols]
hipError_t nb_launch_box_t2(const NbArgs &a, const NbEval &ev, double2 *t2, hipStream_t s);
// the same for nev <= 4 evaluations in three launches, each with its own scratch b
// [cells][bcols], t1 [M1][M2][K3][bcols] and output t2
hipError_t nb_launch_box_t2_multi(const NbArgs &a, int nev, const NbEval *evs, double2 *const *b, double2 *const *t1,
                                  double2 *const *t2, hipStream_t s);
// The separable sums up to the second pass for a run of RK4 steps [s0, s0 + ns) in three
// launches (bform, the passes along a3 and a2), evaluation e = 3 sl + typ: typ 0 J at t_n
// into tj + sl t2_stride, typ 1 H~ H at t_n + dt/2 into th + sl t2_stride, typ 2 H~ H at
// t_n + dt into ring + (step mod nring) t2_stride (the pair of the step before s0 stays in the
// ring when nring > ns); per evaluation scratch b + e b_stride, t1 + e t1_stride. laser:
// [nt][3][6], A and E at t_n, t_n + dt/2, t_n + dt.
struct NbBoxSteps {
    const double *laser;
    int64_t s0;
    int ns, nring;
    double2 *b, *t1, *tj, *th, *ring;
    int64_t b_stride, t1_stride, t2_stride;  // elements
};
hipError_t nb_launch_box_steps(const NbArgs &a, const NbBoxSteps &d, hipStream_t s);
// The same with the pass along a1 as well, each evaluation's matrices into out[e]
// ([nmat][nk_stride][N][N]; the padding k points are not written)
hipError_t nb_launch_box_mats_multi(const NbArgs &a, int nev, const NbEval *evs, double2 *const *b, double2 *const *t1,
                                    double2 *const *t2, double2 *const *out, hipStream_t s);

// Up to four evaluations of the dense sums in one launch (Wannier models with at most
// kNbDenseMultiMaxR lattice vectors; P = e^{ik.R} [nk_stride][nr]), each into out[e]
// ([nmat][nk_stride][N][N]; the padding k points are not written)
constexpr int kNbDenseMultiMaxR = 32;
hipError_t nb_launch_dense_multi(const NbArgs &a, int nev, const NbEval *evs, const double2 *P, double2 *const *out,
                                 hipStream_t s);

// The matrices of steps [s0, s0 + ns) of the RK4 LengthWannier gauge in one launch
// (the dense sums, the same R lists): J at t_n into j [ns][3][slot], H~ (and H with
// dephasing) at t_n + dt/2 into half [ns][2][slot], H~ and H at t_n + dt into the ring
// [nring][2][slot] at step mod nring (the pair of the step before s0 is kept there when
// nring > ns). laser: [nt][3][6], A and E at t_n, t_n + dt/2, t_n + dt.
struct NbDenseSteps {
    const double *laser;
    int64_t s0;
    int ns, nring, deph;
    int64_t slot;  // nk_stride N^2
    double2 *j, *half, *ring;
};
hipError_t nb_launch_dense_steps(const NbArgs &a, const NbDenseSteps &d, const double2 *P, hipStream_t s);

// One whole RK4 step of the LengthWannier gauge for 16 < N <= 32 (nb32_step_kernel):
// four waves per k point, each one 16x16 quadrant of every matrix; the matrices of the
// step's times from the Fourier-sum outputs ([nk_stride][N][N] each, j: [3][nk_stride][N][N]).
constexpr int kNb32MaxBlocks = 256;  // blocks of the N <= 32 step: one per CU, k points walked in turn
struct NbStep32 {
    const double2 *ht_prev, *h_prev;  // H~, H at t_n (h_prev read when u_valid or lam_valid is 0)
    const double2 *j;                 // J_x, J_y, J_z at t_n
    const double2 *ht_half, *h_half;  // at t_n + dt/2
    const double2 *ht_end, *h_end;    // at t_n + dt
    double *lam;                      // [nk_stride][32] band energies at the stored eigenvectors' time
    int u_valid, lam_valid;
    double dt;
    int64_t step_local;
};
hipError_t nb32_launch_step(const NbArgs &a, const NbStep32 &f, int nblocks, double2 *rho, double2 *U, double *partials,
                            hipStream_t s);

// Time-fused VelocityHamiltonian steps for N <= 16 with the classical RK4 tableau
// (nb16_vel_fused_kernel): one wave per k point for steps [s0, s1), one partial row
// per block of kNbFusedWaves k points and step.
struct NbVelFused {
    const double *A;  // [nt][4][3]: A(t_n + c_s dt)
    int64_t s0, s1;
    double dt;
    int nblocks;      // partial rows per step: ceil(nk / waves) for N <= 16, nk above
    int waves;        // N <= 16: k points per block, 8, or 4 where 8 gives fewer blocks than CUs
};
size_t nb16_vel_fused_lds_bytes(int nv, int waves);
hipError_t nb16_launch_vel_fused(const NbArgs &a, const NbVelFused &f, const double2 *P, const double *Eocc,
                                 double2 *rho, double *partials, hipStream_t s);
// the same for 16 < N <= 32: one block of four waves per k point (nblocks = nk)
constexpr int64_t kNb32VelMaxK = 1024;
hipError_t nb32_launch_vel_fused(const NbArgs &a, const NbVelFused &f, const double2 *P, const double *Eocc,
                                 double2 *rho, double *partials, hipStream_t s);

// One time step of the VelocityHamiltonian gauge: A at the stage times and
// the tableau (coefficients already multiplied by dt).
struct NbVelStep {
    int nstage;
    double A[7][3];  // A(t_n + c_i dt), i < nstage
    double a[7][7];  // dt a_ij
    double b[7];     // dt b_i
};

size_t nb_lds_bytes(int N);
// LDS of the velocity step kernel; p_lds keeps the k point's pMatrix in LDS
size_t nb_vel_lds_bytes(int N, int nstage, bool p_lds);
// per k: U, E and pMatrix_a = U^dag J_a U from mats {H, J_x, J_y, J_z} at k;
// rho = diag(occupations). P [3][nk_stride][N][N], Eocc [nk_stride][2N] (E, occ)
hipError_t nb_launch_vel_init(const NbArgs &a, const double2 *mats, double2 *P, double *Eocc, double2 *U,
                              double2 *rho, hipStream_t s);
// U^dag rho U per k with U = eig(H) (H: [nk_stride][N][N]), the phase rule of the velocity set-up
hipError_t nb_launch_band_export(const NbArgs &a, const double2 *H, const double2 *rho, double2 *out, hipStream_t s);
// sample at t_n, then one RK step, for every k (partials row step_local)
hipError_t nb_launch_vel_step(const NbArgs &a, const NbVelStep &v, const double2 *P, const double *Eocc, double2 *rho,
                              double *partials, int64_t step_local, bool p_lds, hipStream_t s);
hipError_t nb_launch_phase(const NbArgs &a, double2 *P, hipStream_t s);
// SPLIT16: e^{ik.R} split into FP16 hi/lo planes once per context
hipError_t nb_launch_psplit(const NbArgs &a, const double2 *P, uint16_t *p16, hipStream_t s);
hipError_t nb_launch_fourier(const NbArgs &a, const NbEval &ev, const double2 *P, double2 *out, hipStream_t s);
// analytic classes: the same matrices [mat][nk_stride][N][N] at k + A (analytic.hip)
hipError_t nb_launch_model(const NbArgs &a, const NbEval &ev, double2 *out, hipStream_t s);
hipError_t nb_launch_init(const NbArgs &a, const double2 *H0, double2 *rho, double2 *U, hipStream_t s);
hipError_t nb_launch_sample(const NbArgs &a, const double2 *mats, const double2 *rho, double *partials,
                            int64_t step_local, double2 *U, int ureuse, hipStream_t s);
hipError_t nb_launch_rhs(const NbArgs &a, const NbStage &sg, const double2 *mats, double2 *rho, double2 *kbuf,
                         double2 *U, hipStream_t s);
// N <= 16 forms of the two above (nband16.hip)
hipError_t nb16_launch_rhs(const NbArgs &a, const NbStage &sg, const double2 *mats, double2 *rho, double2 *kbuf,
                           double2 *U, hipStream_t s);
hipError_t nb16_launch_sample(const NbArgs &a, const double2 *mats, const double2 *rho, double *partials,
                              int64_t step_local, double2 *U, int ureuse, hipStream_t s);
// two consecutive stages at the same time in one launch (RK4's c = 1/2, 1/2)
hipError_t nb16_launch_rhs_pair(const NbArgs &a, const NbStage &sg, const NbStage &sg2, const double2 *mats,
                                double2 *rho, double2 *kbuf, double2 *U, hipStream_t s);
hipError_t nb16_launch_vel_step(const NbArgs &a, const NbVelStep &v, const double2 *P, const double *Eocc,
                                double2 *rho, double *partials, int64_t step_local, hipStream_t s);

}  // namespace sbe
