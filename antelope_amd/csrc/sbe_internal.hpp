// sbe_internal.hpp - host-side context and kernel launchers of the MI355X
// SBE propagator. Not part of the C ABI (include/antelope/sbe_c.h is).
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <string>
#include <utility>
#include <vector>

#include "antelope/sbe_c.h"
#include "laser_core.hpp"
#include "nband.hpp"

namespace sbe {

// ---------------------------------------------------------------- laser --
// Sum of pulses minus its value at atmin (laser.h:249-322), FP64 on the host.
// The fixed-step kernels read per-stage tables built from this class; the
// adaptive stepper evaluates the same pulses (laser_core.hpp) on the device.
struct Laser {
    std::vector<PulseConst> pulses;
    double atmin = 0, atmax = 0, dt = 0;
    double A_at_min[3] = {0, 0, 0}, E_at_min[3] = {0, 0, 0};
    int64_t nt = 0;
    bool consistent = false;

    int init(const sbe_config &cfg, std::string &err);
    void eval(double t, double A[3], double E[3]) const;
};

// ----------------------------------------------------------------- grid --
// Block [k_begin,k_end) of the flattened k grid (momaxis.h) on the host.
struct Grid {
    int n[3] = {1, 1, 1};
    int64_t ntotal = 0, k_begin = 0, k_end = 0;
    double axes[9] = {0}, origin[3] = {0};
    double dV = 0, total_weight = 0;
    std::vector<double> k;  // [nk][3]
    std::vector<double> w;  // [nk]

    int build(const sbe_config &cfg, const sbe_model &m, int64_t kb, int64_t ke,
              std::string &err);
};

// Explicit Runge-Kutta tableau (SBEs.h:395-416; Cash-Karp 1990 for the
// adaptive pair).
struct Tableau {
    int nstage = 0;
    double c[7] = {0}, b[7] = {0}, bhat[7] = {0}, a[7][7] = {{0}};
    int order = 0;
    bool embedded = false;
    int init(int rk_order, std::string &err);
};

// Number of FP64 quantities reduced over k per time sample for the 2-band
// path: inter x,y; intra x,y; sum w s_x, s_y, s_z; sum w s.dhat.
constexpr int kHaldaneNV = 8;

// Kernel parameters of the 2-band (Haldane) path, passed by value.
struct HaldaneArgs {
    // model: d(k) = (t1 sum cos(k.a_j), t1 sum sin(k.a_j),
    //                M0 - c3 sum sin(k.b_j)),  d0 = c0 sum cos(k.b_j)
    double t1, c0, c3, M0;
    double va[3][2], vb[3][2];
    double gam1, gam2;  // 1/T1, 1/T2 (0 = off)
    double dt;
    int64_t nk;         // k points in this block
    int64_t nk_stride;  // padded SoA stride
    int64_t step0;      // first step of this launch (index into the tables)
    int32_t nsteps;     // steps in this launch
    int32_t nwaves;     // waves in the grid (partials row length)
    int32_t inter_intra;
    int32_t precision;
    const double *kpts;     // [3][nk_stride]
    const double *weight;   // [nk_stride] (0 on padding)
    const double *phase_tab;  // [nt][nstage][12]: t1 e^{i A.a_j} (j<3), e^{i A.b_j}, at t_n + c_s dt
                              // (velocity gauge: [nt][nstage][2] = A_x, A_y at t_n + c_s dt)
    const float *phase_tab32; // same table rounded to FP32 (FP32 state only)
    int32_t nval;             // occupied bands of the initial state
    int32_t kpt;              // k points per lane in the step kernel (1 or 2)
    int32_t fermi;            // Fermi-Dirac initial occupations (velocity gauge kernels)
    double fermi_energy;
    void *state;            // SoA [4][nk_stride] (fp64) or [8][nk_stride] (fp32 + compensation)
    double *partials;       // [nsteps][nwaves][kHaldaneNV]
};

// Kernel parameters of the adaptive Cash-Karp stepper (2-band path). Each
// wave (64 k points) controls its own step size; samples are taken at the
// fixed times t_n = atmin + n dt_out by the stepper's dense output. The
// laser is evaluated on the device at the stage times.
constexpr int kMaxDevicePulses = 8;
// The struct lives in device memory and does not change after set-up; a
// launch covers the samples [*cursor, min(*cursor + nwin, nt)) and a second
// one-thread-per-solve kernel then advances the cursor, so nothing is copied
// per launch and one kernel serves single solves and batches alike.
struct AdaptiveArgs {
    HaldaneArgs h;             // model, grid, state, partials (phase_tab: rows at t_n only)
    double atmin, dt_out, inv_dt_out;
    int64_t nt;                // < 2^31 (the kernels index samples with 32 bits)
    int64_t *cursor;           // next sample; the state is at t_cursor
    double rtol, atol, hmax, hmin;
    int32_t npulses;
    double A_at_min[3];
    const PulseConst *pulses;  // [npulses] in device memory
    const double *a_tab;       // [nt][4]: A - A(t_min) (x, y) and dA/dt (x, y) at t_n
    double *hstep;             // [nwaves] proposed step size carried between launches
    int64_t *counts;           // [nwaves][2] accepted and rejected step attempts
    int32_t *fail;             // set when a wave exhausts its attempts or its step falls below hmin
};

// ------------------------------------------------------------------ ctx --
enum class PathKind { Haldane2 = 0, NBand = 1 };

struct DeviceBuffer {
    void *ptr = nullptr;
    size_t bytes = 0;
};

}  // namespace sbe

struct sbe_ctx {
    sbe_config cfg{};
    std::vector<sbe_pulse> pulses;
    sbe_model model{};
    sbe::Laser laser;
    sbe::Grid grid;
    sbe::Tableau tab;
    sbe::PathKind path = sbe::PathKind::Haldane2;
    bool velocity = false;  // VelocityHamiltonian gauge: the state is band-basis at k
    int N = 2;
    int64_t nk = 0, nk_stride = 0;
    int64_t steps_done = 0;
    int device = 0;
    hipStream_t stream = nullptr;
    std::vector<sbe::DeviceBuffer> owned;
    int64_t device_bytes = 0;

    // device data
    double *d_kpts = nullptr;       // [3][nk_stride]
    double *d_weight = nullptr;     // [nk_stride]
    double *d_phase_tab = nullptr;  // per-stage laser phase table
    void *d_state = nullptr;
    double *d_partials = nullptr;   // [chunk][nwaves][nv]
    double *d_series = nullptr;     // [nt][nv]
    // set to 1 by the series reductions when a k sum is not finite (pinned
    // host memory mapped for the device: readable without a copy once the
    // reduction that wrote it has completed)
    int32_t *nonfinite_h = nullptr, *nonfinite_d = nullptr;
    int64_t chunk_steps = 0;
    int32_t nwaves = 0;
    int nv = 0;
    double n_sum = 0;  // sum_k w_k n_k (constant when T1 is off)

    sbe::HaldaneArgs hargs{};
    // adaptive (Cash-Karp) 2-band path
    bool adaptive = false;
    sbe::AdaptiveArgs aargs{};
    sbe::PulseConst *d_pulses = nullptr;
    sbe::AdaptiveArgs *d_aargs = nullptr;  // device copy of aargs (constant)
    int64_t *d_cursor = nullptr;
    double *d_hstep = nullptr;
    int64_t *d_counts = nullptr;
    int32_t *d_fail = nullptr;
    // N-band path
    sbe::NbArgs nargs{};
    double2 *nb_P = nullptr;     // e^{ik.R} [nk_stride][nr]
    double2 *nb_rho = nullptr;   // [nk_stride][N][N]
    double2 *nb_mats = nullptr;  // Fourier-sum outputs [5][nk_stride][N][N]: H~, H, J_x, J_y, J_z
    double2 *nb_k = nullptr;     // RK stage derivatives [nstage][nk_stride][N][N]
    double2 *nb_U = nullptr;     // last eigenvectors per k [nk_stride][N][N]
    double nb_u_time = 0;        // time whose H the stored eigenvectors diagonalise
    double nb_ht_time = 0;       // time of the H~ in mats slot 0
    bool nb_ht_valid = false;
    double nb_h_time = 0;        // time of the H in mats slot 1
    bool nb_h_valid = false;
    bool nb_vel_p_lds = false;   // velocity gauge: pMatrix of the k point staged in LDS
    // fused RK4 steps (nb16_step_kernel): second-pass outputs of the separable sums
    bool nbf = false;
    // second-pass outputs of up to nbf_S steps: J at t_n and H~ H at t_n + dt/2 per step, the
    // end pairs H~ H at t_n + dt in a ring of nbf_S + 1 (step mod ring)
    int nbf_S = 0;
    double2 *nbf_tj = nullptr, *nbf_th = nullptr, *nbf_ring = nullptr;
    int64_t nbf_t2_stride = 0;
    double *nbf_lam = nullptr;   // [nk_stride][16] band energies at the stored eigenvectors' time
    double2 *nbf_bs = nullptr, *nbf_t1s = nullptr;  // per evaluation: B on the box, first-pass output
    int64_t nbf_b_stride = 0, nbf_t1_stride = 0;
    const double *nbf_laser = nullptr;  // [nt][3][6] A, E at t_n, t_n + dt/2, t_n + dt
    int nbf_g1 = 0, nbf_nblocks = 0, nbf_waves = 8;
    bool nbf_prev_valid = false;  // the ring holds H~, H at nbf_prev_time for the step that starts there
    double nbf_prev_time = 0;
    bool nbf_lam_valid = false;
    double nbf_lam_time = 0;
    // fused RK4 steps for 16 < N <= 32 (nb32_step_kernel): the matrices of a step's
    // times in mats and the stage-derivative slots, H~ H pairs rotating
    bool nb32 = false;
    double2 *nb32_pair[3] = {nullptr, nullptr, nullptr};  // H~ H at t_n, t_n + dt/2, t_n + dt
    double *nb32_lam = nullptr;  // [nk_stride][32] band energies at the stored eigenvectors' time
    double2 *nb32_b[4] = {nullptr, nullptr, nullptr, nullptr};   // separable sums: per evaluation B,
    double2 *nb32_t1[4] = {nullptr, nullptr, nullptr, nullptr};  // first and second pass
    double2 *nb32_t2[4] = {nullptr, nullptr, nullptr, nullptr};
    int nb32_blocks = 0;
    // short R lists, few k points: the matrices of up to nb32_S steps in one dense launch,
    // J and the midpoint pair per step, the end pairs in a ring of nb32_S + 1
    bool nb32_dm = false;
    int nb32_S = 0;
    double2 *nb32_j = nullptr, *nb32_half = nullptr, *nb32_ring = nullptr;
    const double *nb32_laser = nullptr;  // [nt][3][6] A, E at t_n, t_n + dt/2, t_n + dt
    // time-fused velocity-gauge steps (nb16_vel_fused_kernel)
    bool nbv = false;
    const double *nbv_A = nullptr;  // [nt][4][3] A at the RK4 stage times
    int nbv_nblocks = 0, nbv_waves = 8;
    std::string err;

    // optional kernel timing: (start, stop) event pairs per kernel class
    bool timing = false;
    std::vector<std::pair<hipEvent_t, hipEvent_t>> ev[2];
};

namespace sbe {
// Launchers (haldane.hip, series.hip).
hipError_t launch_haldane_init(const HaldaneArgs &a, int init_type, double ef, hipStream_t s);
hipError_t launch_haldane_steps(const HaldaneArgs &a, int rk_order, bool t1_on, hipStream_t s);
// flag (may be null): set to 1 when a sum of the launch is not finite; group_rows
// (0: the default) is the first pass's group size for steps of many rows
hipError_t launch_series_reduce(double *partials, int64_t nsteps, int32_t nrows, int nv,
                                double *series, hipStream_t s, int32_t *flag = nullptr, int32_t group_rows = 0);
// One k block's share of a multi-block reduction: steps [first, first +
// nsteps) of the launch grid sum partials[nsteps][nrows][nv] into series.
struct SeriesJob {
    const double *partials;
    double *series;
    int32_t nrows;
    int32_t nsteps;
    int64_t first;
    int32_t *flag;  // set to 1 on a non-finite sum (may be null)
};
// true when one pass suffices (what the multi-block launch runs)
bool series_multi_ok(int32_t nrows);
hipError_t launch_series_reduce_multi(const SeriesJob *d_jobs, int njobs, int64_t total_steps, int nv,
                                      hipStream_t s);
hipError_t launch_haldane_export(const HaldaneArgs &a, double *out, hipStream_t s);
// VelocityHamiltonian gauge (haldane_velocity.hip): band-basis state
hipError_t launch_haldane_vel_init(const HaldaneArgs &a, hipStream_t s);
hipError_t launch_haldane_vel_steps(const HaldaneArgs &a, int rk_order, bool t1_on, hipStream_t s);
hipError_t launch_haldane_batch(const HaldaneArgs *d_insts, const int32_t *d_block_inst, const int32_t *d_block_first,
                                int nblocks, int rk_order, int precision, bool t1_on, bool deph, hipStream_t s);
// Adaptive stepper: d_block_inst / d_block_first may be null for a single
// solve (d_insts[0], grid = its blocks). Then advance every solve's cursor.
hipError_t launch_haldane_adaptive(const AdaptiveArgs *d_insts, const int32_t *d_block_inst,
                                   const int32_t *d_block_first, int nblocks, int64_t nwin, int precision,
                                   bool deph, int kpt, hipStream_t s);
hipError_t launch_adaptive_advance(const AdaptiveArgs *d_insts, int n, int64_t nwin, hipStream_t s);
int haldane_waves(int64_t nk);
// N-band host side (nband_host.cpp)
int nb_validate(const sbe_model &m, std::string &err);
size_t nb_footprint(const sbe_config &cfg, const sbe_model &m, int64_t nk, int64_t nt, int nstage);
int nb_setup(sbe_ctx *ctx);
int nb_run(sbe_ctx *ctx, int64_t step_begin, int64_t step_end);
int nb_series(sbe_ctx *ctx, double *j_inter, double *j_intra, double *rho_bar, double *band_pop);
int nb_state(sbe_ctx *ctx, int basis, double *rho);
constexpr int kBlock = 256;
// The adaptive stepper's launch unit: one wave per block, so a wave slot frees
// as soon as its wave ends (the waves of a window take different step counts).
constexpr int kAdaptiveBlock = 64;
// sbe_api.cpp: argument checks shared by the entry points, the thread-local
// message of sbe_last_error(NULL), and the fixed-step plan's k points per lane
int validate_inputs(const sbe_config *cfg, const sbe_model *m, std::string &err);
// Shared launches (the serving loop): 2-band fixed-step contexts planned as
// batch members (one k point per lane) advance together in one launch of the
// batch step kernel and one reduction launch, each by its own step count, in
// stream order on s. A context gives the same bits in any shared launch and
// alone (launch boundaries do not change the fixed-step results).
struct SharedLaunch {
    HaldaneArgs *d_insts = nullptr;
    int32_t *d_map = nullptr;
    SeriesJob *d_jobs = nullptr;
    int cap_insts = 0, cap_map = 0;
};
int create_batch_member(const sbe_config *cfg, const sbe_model *model, sbe_ctx **out);
bool shared_eligible(const sbe_ctx *c);
int shared_key(const sbe_ctx *c);  // members of one launch share it (tableau, precision)
int launch_shared(sbe_ctx *const *cs, const int64_t *nsteps, int n, SharedLaunch &buf, hipStream_t s, std::string &err);
void free_shared(SharedLaunch &buf);
int set_thread_error(int code, const std::string &msg);
bool fixed_step_two_k_per_lane(const sbe_config &cfg, const sbe_model &m, int64_t nk);
// tuner.cpp: the cached plan of this problem class on the current device, if any
bool tuning_lookup(const sbe_config &cfg, const sbe_model &m, int64_t nk, sbe_tuning &out);
}  // namespace sbe
