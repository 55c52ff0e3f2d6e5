// nband_host.cpp - host orchestration of the N-band (Wannier) path: device
// tables, the initial state and the per-step launch sequence
//   sample:  Fourier sums {H, J_x, J_y, J_z} at t_n  -> nb_sample_kernel
//   stage i: Fourier sums {H + E.D, H} at t_n + c_i dt -> nb_rhs_kernel
// (antelope.cpp:297-338 with SBEs.h:596-624).
#include <algorithm>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include "analytic.hpp"
#include "nband.hpp"
#include "sbe_internal.hpp"

namespace sbe {

namespace {

int hipc(sbe_ctx *ctx, hipError_t e, const char *what)
{
    if (e == hipSuccess) return SBE_OK;
    ctx->err = std::string(what) + ": " + hipGetErrorString(e);
    return (e == hipErrorOutOfMemory || e == hipErrorMemoryAllocation) ? SBE_ERR_OOM : SBE_ERR_HIP;
}

template <typename T>
T *nalloc(sbe_ctx *ctx, size_t bytes, int &rc)
{
    if (rc) return nullptr;
    void *p = nullptr;
    const hipError_t e = hipMalloc(&p, bytes ? bytes : 1);
    if (e != hipSuccess) {
        rc = hipc(ctx, e, "hipMalloc");
        return nullptr;
    }
    ctx->owned.push_back({p, bytes});
    ctx->device_bytes += static_cast<int64_t>(bytes);
    return static_cast<T *>(p);
}

void nfree(sbe_ctx *ctx, void *ptr)
{
    for (size_t i = 0; i < ctx->owned.size(); ++i)
        if (ctx->owned[i].ptr == ptr) {
            (void)hipFree(ptr);
            ctx->device_bytes -= static_cast<int64_t>(ctx->owned[i].bytes);
            ctx->owned.erase(ctx->owned.begin() + static_cast<std::ptrdiff_t>(i));
            return;
        }
}

constexpr size_t kNbPartialBytes = size_t(128) << 20;
// rows per group of the first reduction pass for the N-band sample rows (hundreds of
// values per row: small groups spread a step's sum over many blocks)
constexpr int32_t kNbReduceGroup = 128;

// the matrices {H~, H, J} at k + A: Fourier sums (Wannier) or the analytic classes
// two sample/stage times that differ only by rounding (t_n + dt against t_n+1)
bool same_time(double x, double y, double dt) { return std::fabs(x - y) <= 1e-9 * dt; }

hipError_t eval_mats(const NbArgs &a, const NbEval &ev, const double2 *P, double2 *out, hipStream_t s)
{
    return a.model_kind ? nb_launch_model(a, ev, out, s) : nb_launch_fourier(a, ev, P, out, s);
}

// Shape of the separable Fourier sums: the lattice box of the R list and the
// k grid's sides (distinct k.a_i per axis are at most npoints[i]).
struct BoxShape {
    int ok = 0;
    int lo[3] = {0, 0, 0}, M[3] = {1, 1, 1}, K[3] = {1, 1, 1};
    int64_t cells() const { return static_cast<int64_t>(M[0]) * M[1] * M[2]; }
    int64_t kcells() const { return static_cast<int64_t>(K[0]) * K[1] * K[2]; }
};

constexpr int kBoxMaxSide = 32;

// theta_j = k.a_j of a grid point k = origin + sum_i f_i b_i depends on f_j alone
// when every BZ axis b_i (of an axis with more than one point) is normal to the
// lattice vectors a_j, j != i: then any block of the grid is a product grid in
// theta. The model's own axes (reciprocal vectors, scaled per row by ksfactor)
// are; an explicit calc.BZaxes is checked.
bool axes_separable(const sbe_config &cfg, const sbe_model &m)
{
    if (!cfg.has_bz_axes) return true;
    for (int i = 0; i < 3; ++i) {
        if (cfg.npoints[i] <= 1) continue;
        const double *b = cfg.bz_axes + 3 * i;
        const double nb = std::sqrt(b[0] * b[0] + b[1] * b[1] + b[2] * b[2]);
        for (int j = 0; j < 3; ++j) {
            if (j == i) continue;
            const double *av = m.lattice + 3 * j;
            const double na = std::sqrt(av[0] * av[0] + av[1] * av[1] + av[2] * av[2]);
            if (std::fabs(b[0] * av[0] + b[1] * av[1] + b[2] * av[2]) > 1e-12 * nb * na) return false;
        }
    }
    return true;
}

// The box form is taken when the model is a Wannier model in the FP64 form, the
// grid is a product grid in theta (axes_separable), its R are distinct and fill a
// lattice box of sides <= 32 at least a quarter full, and the three passes cost
// under half the dense sums (nk x nr multiply-adds per column). The footprint
// query and the set-up take the same decision. ANTELOPE_NB_DENSE=1 keeps the
// dense GEMM.
BoxShape box_shape(const sbe_config &cfg, const sbe_model &m, int64_t nk, bool split16)
{
    BoxShape b;
    const int32_t *npoints = cfg.npoints;
    if (am::is_analytic(m.kind) || split16 || m.nrpts < 1 || !m.rvec) return b;
    if (const char *e = std::getenv("ANTELOPE_NB_DENSE"))
        if (std::atoi(e) == 1) return b;
    sbe_tuning tu{};
    if (!std::getenv("ANTELOPE_NB_DENSE") && tuning_lookup(cfg, m, nk, tu) && tu.nb_sums == 2) return b;
    if (!axes_separable(cfg, m)) return b;
    int hi[3];
    for (int i = 0; i < 3; ++i) b.lo[i] = hi[i] = m.rvec[i];
    for (int r = 0; r < m.nrpts; ++r)
        for (int i = 0; i < 3; ++i) {
            b.lo[i] = std::min(b.lo[i], m.rvec[3 * r + i]);
            hi[i] = std::max(hi[i], m.rvec[3 * r + i]);
        }
    for (int i = 0; i < 3; ++i) {
        b.M[i] = hi[i] - b.lo[i] + 1;
        b.K[i] = npoints[i] > 0 ? npoints[i] : 1;
        if (b.M[i] > kBoxMaxSide) return b;
    }
    if (b.cells() > 4 * static_cast<int64_t>(m.nrpts)) return b;
    std::vector<char> seen(static_cast<size_t>(b.cells()), 0);
    for (int r = 0; r < m.nrpts; ++r) {  // the same R twice: the dense sums add both
        const int32_t *n = m.rvec + 3 * r;
        const int64_t cell = (static_cast<int64_t>(n[0] - b.lo[0]) * b.M[1] + (n[1] - b.lo[1])) * b.M[2] + (n[2] - b.lo[2]);
        if (seen[cell]) return b;
        seen[cell] = 1;
    }
    const double box = static_cast<double>(b.M[0]) * b.M[1] * b.K[2] * b.M[2] +
                       static_cast<double>(b.M[0]) * b.K[1] * b.K[2] * b.M[1] +
                       static_cast<double>(b.K[0]) * b.K[1] * b.K[2] * b.M[0];
    if (2.0 * box > static_cast<double>(nk) * m.nrpts) return b;
    b.ok = 1;
    return b;
}

// The fused RK4 step (nb16_step_kernel) replaces the stage and sample kernels when
// one launch per step can hold everything: FP64 LengthWannier with the classical
// RK4 tableau, N <= 16 on the register-tile kernels, a Wannier model with
// dipoles (the dipole-free phases are an epilogue of the full sums) whose k points
// (the whole product grid or a k block of it) take the separable sums with at most
// kNbFusedMaxM1 lattice vectors along a1. ANTELOPE_NB_FUSED=0 keeps the separate kernels.
bool fused_eligible(const sbe_config &cfg, const sbe_model &m, int64_t nk, const BoxShape &bx)
{
    if (const char *e = std::getenv("ANTELOPE_NB_FUSED"))
        if (std::atoi(e) == 0) return false;
    if (!bx.ok || cfg.rk_order != SBE_RK4 || cfg.gauge != SBE_GAUGE_LENGTH_WANNIER || cfg.precision != SBE_PREC_FP64)
        return false;
    if (am::is_analytic(m.kind) || m.is_dipole_zero || m.nband > 16 || bx.M[0] > kNbFusedMaxM1) return false;
    if (const char *e = std::getenv("ANTELOPE_NB_LDS"))
        if (std::atoi(e) == 1) return false;
    sbe_tuning tu{};
    if (tuning_lookup(cfg, m, nk, tu) && tu.nb_kernels == 2) return false;
    return true;  // the whole grid, or a k block of it (holes in the box are idle waves)
}

// The time-fused velocity-gauge steps (nb16_vel_fused_kernel for N <= 16 on the
// register-tile kernels, nb32_vel_fused_kernel for 16 < N <= 32 up to kNb32VelMaxK k
// points, one block per k point): FP64 with the classical RK4 tableau; any model (the step
// has no Fourier sums). ANTELOPE_NB_FUSED=0 keeps one launch per step.
bool vel_fused_eligible(const sbe_config &cfg, const sbe_model &m, int64_t nk)
{
    if (const char *e = std::getenv("ANTELOPE_NB_FUSED"))
        if (std::atoi(e) == 0) return false;
    if (cfg.gauge != SBE_GAUGE_VELOCITY_HAMILTONIAN || cfg.rk_order != SBE_RK4 || cfg.precision != SBE_PREC_FP64 ||
        m.nband > 32 || (m.nband > 16 && nk > kNb32VelMaxK))
        return false;
    if (const char *e = std::getenv("ANTELOPE_NB_LDS"))
        if (std::atoi(e) == 1) return false;
    sbe_tuning tu{};
    return !(tuning_lookup(cfg, m, nk, tu) && tu.nb_kernels == 2);
}

// The fused RK4 step for 16 < N <= 32 (nb32_step_kernel): FP64 LengthWannier with the
// classical RK4 tableau and a Wannier model with dipoles (dense or separable sums);
// ANTELOPE_NB_FUSED=0 keeps the separate kernels.
bool nb32_eligible(const sbe_config &cfg, const sbe_model &m)
{
    if (const char *e = std::getenv("ANTELOPE_NB_FUSED"))
        if (std::atoi(e) == 0) return false;
    return cfg.rk_order == SBE_RK4 && cfg.gauge == SBE_GAUGE_LENGTH_WANNIER && cfg.precision == SBE_PREC_FP64 &&
           !am::is_analytic(m.kind) && !m.is_dipole_zero && m.nband > 16 && m.nband <= 32;
}

// The N <= 32 fused step's matrices in one dense launch for all evaluations of a step
// (nb_launch_dense_multi) when the R list is short and the k points few: there the
// separable passes cost more in launches than they save in arithmetic.
bool nb32_dense_multi_n(int nr, int64_t nk) { return nr >= 1 && nr <= kNbDenseMultiMaxR && nk <= 1024; }

// Steps whose Fourier evaluations one batch of launches covers (the fused RK4 steps):
// at most kNbEvalSteps (a few launches per hundred steps) and about kNbEvalBytes of outputs
// and scratch, at least one, at most the pulse; ANTELOPE_NB_EVAL_STEPS sets it (the results
// do not depend on it).
constexpr int64_t kNbEvalSteps = 64;
constexpr size_t kNbEvalBytes = size_t(256) << 20;
int64_t eval_steps(size_t bytes_per_step, int64_t nt)
{
    int64_t S = std::min<int64_t>(kNbEvalSteps, static_cast<int64_t>(kNbEvalBytes / bytes_per_step));
    if (const char *e = std::getenv("ANTELOPE_NB_EVAL_STEPS"))
        if (std::atoi(e) > 0) S = std::atoi(e);
    return std::max<int64_t>(1, std::min<int64_t>(nt, S));
}

// N <= 16: the separable passes up to the second of three evaluations per step (at most
// 21845 steps: three evaluations a step in one grid dimension)
int nb16_steps_per_eval(size_t t2, size_t b, size_t t1, int64_t nt)
{
    return static_cast<int>(std::min<int64_t>(eval_steps(3 * (t2 + b + t1), nt), 21845));
}

// 16 < N <= 32, one dense launch: seven N x N matrices per k point and step
int nb32_steps_per_eval(int64_t nk_stride, int N, int64_t nt)
{
    return static_cast<int>(eval_steps(7 * static_cast<size_t>(nk_stride) * N * N * sizeof(double2), nt));
}

// k points along a1 per block of the fused N <= 16 step: 8, or 4 where 8 would give fewer
// blocks than the GPU has CUs (small grids: a wave per SIMD shortens the step, whose time is
// then one k point's chain of products)
constexpr int kNbCUs = 256;
int fused_waves(const BoxShape &bx)
{
    const int64_t b8 = static_cast<int64_t>((bx.K[0] + 7) / 8) * bx.K[1] * bx.K[2];
    return b8 < kNbCUs ? 4 : 8;
}

// the same rule for the time-fused velocity steps (one k point per wave, k in order)
int vel_waves(int64_t nk) { return (nk + 7) / 8 < kNbCUs ? 4 : 8; }

int fused_blocks(const BoxShape &bx, int &g1)
{
    const int w = fused_waves(bx);
    g1 = (bx.K[0] + w - 1) / w;
    return g1 * bx.K[1] * bx.K[2];
}

struct NbPlan {
    int64_t nk_stride;
    int sample_waves, nv;
    int64_t chunk;
    int nr_pad;
    int64_t bcols;
    size_t bytes_P, bytes_rho, bytes_mats, bytes_k, bytes_tab, bytes_partials, bytes_series, bytes_split, bytes_b64;
    size_t bytes_box;
    size_t bytes_fused;  // fused steps: four second-pass outputs and the band energies
    size_t bytes_vel;    // time-fused velocity steps: A at the stage times of the whole pulse
    // offsets inside the box buffer: B, T1, T2, E_1..E_3, R map, k map
    size_t off_t1, off_t2, off_e[3], off_r, off_kinv;
};

size_t align256(size_t x) { return (x + 255) / 256 * 256; }

NbPlan nb_plan(int N, int nr, int64_t nk, int64_t nt, int nstage, bool split16, const BoxShape &bx, bool fused,
               bool vel_fused = false, bool nb32 = false)
{
    NbPlan p{};
    p.nk_stride = (nk + 63) / 64 * 64;
    p.sample_waves = static_cast<int>(std::min<int64_t>(nk, 2048));
    p.nv = 6 + N + 2 * N * N;
    const size_t nn = static_cast<size_t>(N) * N;
    int g1 = 0;
    int64_t rows = fused ? std::max<int64_t>(p.sample_waves, fused_blocks(bx, g1)) : p.sample_waves;
    if (vel_fused) rows = std::max<int64_t>(rows, N > 16 ? nk : (nk + vel_waves(nk) - 1) / vel_waves(nk));
    p.bytes_vel = vel_fused ? align256(static_cast<size_t>(nt) * 12 * sizeof(double)) : 0;
    const size_t row = static_cast<size_t>(rows) * p.nv * sizeof(double);
    p.chunk = std::max<int64_t>(1, std::min<int64_t>(nt, static_cast<int64_t>(kNbPartialBytes / row)));
    // e^{ik.R} for the dense sums (also the N <= 32 fused step's one-launch sums); the
    // separable form does not read it
    const bool dm32 = nb32 && nb32_dense_multi_n(nr, nk);
    p.bytes_P = bx.ok && !dm32 ? 0 : static_cast<size_t>(p.nk_stride) * nr * sizeof(double2);
    p.bytes_rho = static_cast<size_t>(p.nk_stride) * nn * sizeof(double2);
    p.bytes_mats = 5 * p.bytes_rho;  // H~, H, J_x, J_y, J_z
    p.bytes_k = static_cast<size_t>(nstage + 1) * p.bytes_rho;  // stage derivatives + eigenvectors
    p.bytes_tab = static_cast<size_t>(p.nk_stride) * 4 * sizeof(double) + static_cast<size_t>(nr) * 4 * sizeof(double) +
                  static_cast<size_t>(nr) * 4 * nn * sizeof(double2);
    p.bytes_partials = static_cast<size_t>(p.chunk) * row;
    p.bytes_series = static_cast<size_t>(nt) * p.nv * sizeof(double);
    p.nr_pad = (nr + 31) / 32 * 32;
    p.bcols = (4 * static_cast<int64_t>(nn) + 63) / 64 * 64;
    // FP16 hi/lo planes of e^{ik.R} (per context) and of B (per evaluation)
    p.bytes_split = split16 ? 4 * sizeof(uint16_t) * static_cast<size_t>(p.nr_pad) *
                                  static_cast<size_t>(p.nk_stride + p.bcols)
                            : 0;
    // FP64 dense form: B of one evaluation, formed once per evaluation (Wannier models only)
    p.bytes_b64 = !split16 && nr > 0 && !bx.ok ? sizeof(double2) * static_cast<size_t>(p.nr_pad) * p.bcols : 0;
    p.bytes_box = 0;
    if (bx.ok) {
        const size_t col = sizeof(double2) * static_cast<size_t>(p.bcols);
        size_t o = align256(static_cast<size_t>(bx.cells()) * col);
        p.off_t1 = o;
        o += align256(static_cast<size_t>(bx.M[0]) * bx.M[1] * bx.K[2] * col);
        p.off_t2 = o;
        o += align256(static_cast<size_t>(bx.M[0]) * bx.K[1] * bx.K[2] * col);
        for (int i = 0; i < 3; ++i) {
            p.off_e[i] = o;
            o += align256(static_cast<size_t>(bx.K[i]) * bx.M[i] * sizeof(double2));
        }
        p.off_r = o;
        o += align256(static_cast<size_t>(bx.cells()) * sizeof(int32_t));
        p.off_kinv = o;
        o += align256(static_cast<size_t>(bx.kcells()) * sizeof(int32_t));
        p.bytes_box = o;
    }
    p.bytes_fused = 0;
    if (fused) {  // second-pass outputs of S steps (J, the midpoint pair, the end pairs in a ring
        // of S + 1), the energies, per evaluation B and the first pass, A and E of the pulse
        const size_t col = sizeof(double2) * static_cast<size_t>(p.bcols);
        const size_t t2 = align256(col * bx.M[0] * bx.K[1] * bx.K[2]);
        const size_t b = align256(col * static_cast<size_t>(bx.cells()));
        const size_t t1 = align256(col * bx.M[0] * bx.M[1] * bx.K[2]);
        const size_t S = static_cast<size_t>(nb16_steps_per_eval(t2, b, t1, nt));
        p.bytes_fused = (3 * S + 1) * t2 + align256(static_cast<size_t>(p.nk_stride) * 16 * sizeof(double)) +
                        3 * S * (b + t1) + align256(static_cast<size_t>(nt) * 18 * sizeof(double));
    }
    if (nb32) {  // the energies, and with the separable sums per evaluation B and both passes
        p.bytes_fused = align256(static_cast<size_t>(p.nk_stride) * 32 * sizeof(double));
        if (dm32) {  // the matrices of a run of steps (7 per step, 2 more in the ring), A and E of the pulse
            const int S = nb32_steps_per_eval(p.nk_stride, N, nt);
            p.bytes_fused += align256((7 * static_cast<size_t>(S) + 2) * p.bytes_rho) +
                             align256(static_cast<size_t>(nt) * 18 * sizeof(double));
        }
        if (bx.ok && !dm32) {
            const size_t col = sizeof(double2) * static_cast<size_t>(p.bcols);
            p.bytes_fused += 4 * (align256(col * static_cast<size_t>(bx.cells())) +
                                  align256(col * bx.M[0] * bx.M[1] * bx.K[2]) + align256(col * bx.M[0] * bx.K[1] * bx.K[2]));
        }
    }
    return p;
}

// theta_i = k.a_i of this block's k points grouped per axis (values within
// 1e-9 relative are one grid line); fills E_i, the R map and the k map and
// returns false when the k points do not form a product grid within the
// planned sides (the dense sums are used then).
bool box_tables(const sbe_model &m, const std::vector<double> &hk, int64_t nk, const BoxShape &bx, int Kact[3],
                std::vector<double> &E, size_t e_off[3], std::vector<int32_t> &rmap, std::vector<int32_t> &kinv)
{
    std::vector<int32_t> gi[3];
    std::vector<double> rep[3];
    for (int i = 0; i < 3; ++i) {
        const double *av = m.lattice + 3 * i;
        std::vector<double> th(static_cast<size_t>(nk));
        double big = 0.0;
        for (int64_t k = 0; k < nk; ++k) {
            th[k] = hk[3 * k] * av[0] + hk[3 * k + 1] * av[1] + hk[3 * k + 2] * av[2];
            big = std::max(big, std::fabs(th[k]));
        }
        const double tol = 1e-9 * (1.0 + big);
        std::vector<int64_t> ord(static_cast<size_t>(nk));
        for (int64_t k = 0; k < nk; ++k) ord[k] = k;
        std::sort(ord.begin(), ord.end(), [&](int64_t x, int64_t y) { return th[x] < th[y]; });
        gi[i].assign(static_cast<size_t>(nk), 0);
        for (int64_t q = 0; q < nk; ++q) {
            const int64_t k = ord[q];
            if (rep[i].empty() || th[k] - rep[i].back() > tol) rep[i].push_back(th[k]);
            gi[i][k] = static_cast<int32_t>(rep[i].size() - 1);
        }
        Kact[i] = static_cast<int>(rep[i].size());
        if (Kact[i] > bx.K[i]) return false;
    }
    // E_i[g][m] = e^{i (lo_i + m) theta_g}, stored at the planned offsets (in units of double)
    for (int i = 0; i < 3; ++i)
        for (int g = 0; g < Kact[i]; ++g)
            for (int mm = 0; mm < bx.M[i]; ++mm) {
                const double ph = static_cast<double>(bx.lo[i] + mm) * rep[i][g];
                E[e_off[i] + 2 * (static_cast<size_t>(g) * bx.M[i] + mm)] = std::cos(ph);
                E[e_off[i] + 2 * (static_cast<size_t>(g) * bx.M[i] + mm) + 1] = std::sin(ph);
            }
    rmap.assign(static_cast<size_t>(bx.cells()), -1);
    for (int r = 0; r < m.nrpts; ++r) {
        const int32_t *n = m.rvec + 3 * r;
        const int64_t cell = (static_cast<int64_t>(n[0] - bx.lo[0]) * bx.M[1] + (n[1] - bx.lo[1])) * bx.M[2] + (n[2] - bx.lo[2]);
        if (rmap[cell] >= 0) return false;  // the same R twice
        rmap[cell] = r;
    }
    kinv.assign(static_cast<size_t>(bx.kcells()), -1);
    for (int64_t k = 0; k < nk; ++k) {
        const int64_t cell = (static_cast<int64_t>(gi[0][k]) * Kact[1] + gi[1][k]) * Kact[2] + gi[2][k];
        if (kinv[cell] >= 0) return false;  // two k points on one grid node
        kinv[cell] = static_cast<int32_t>(k);
    }
    return true;
}

}  // namespace

size_t nb_footprint(const sbe_config &cfg, const sbe_model &m, int64_t nk, int64_t nt, int nstage)
{
    const bool split16 = cfg.precision == SBE_PREC_SPLIT16;
    const BoxShape bx = box_shape(cfg, m, nk, split16);
    const bool fused = fused_eligible(cfg, m, nk, bx);
    const bool vfused = vel_fused_eligible(cfg, m, nk);
    const bool f32 = nb32_eligible(cfg, m);
    const NbPlan p =
        nb_plan(m.nband, am::is_analytic(m.kind) ? 0 : m.nrpts, nk, nt, nstage, split16, bx, fused, vfused, f32);
    return p.bytes_P + p.bytes_rho + p.bytes_mats + p.bytes_k + p.bytes_tab + p.bytes_partials + p.bytes_series +
           p.bytes_split + p.bytes_b64 + p.bytes_box + p.bytes_fused + p.bytes_vel;
}

int nb_validate(const sbe_model &m, std::string &err)
{
    if (am::is_analytic(m.kind)) {
        int want = 2;
        if (m.kind == SBE_MODEL_KANE_MELE || m.kind == SBE_MODEL_XYH_LATTICE || m.kind == SBE_MODEL_HALDANE_2L ||
            m.kind == SBE_MODEL_WDSM_TAAS)
            want = 4;
        if (m.kind == SBE_MODEL_TMDC) want = m.nband == 6 ? 6 : 3;
        if (m.nband != want) {
            err = "this model class has " + std::to_string(want) + " bands (TMDC: 3, or 6 with spin-orbit)";
            return SBE_ERR_MODEL;
        }
        for (double v : m.params)
            if (!std::isfinite(v)) {
                err = "model parameters must be finite";
                return SBE_ERR_MODEL;
            }
        int ia0 = 0;  // index of the lattice constant in params (-1: none)
        switch (m.kind) {
        case SBE_MODEL_WILSON_MASS: ia0 = 3; break;
        case SBE_MODEL_SQUARE_SBAND: ia0 = 2; break;
        case SBE_MODEL_KANE_MELE: ia0 = 5; break;
        case SBE_MODEL_XYH_LATTICE: ia0 = 3; break;
        case SBE_MODEL_HALDANE_2L: ia0 = 8; break;
        case SBE_MODEL_BI2SE3_SURF: ia0 = -1; break;
        case SBE_MODEL_MULTI_WSM: ia0 = 4; break;
        case SBE_MODEL_WDSM_TAAS: ia0 = 4; break;
        default: ia0 = 0; break;
        }
        if ((ia0 >= 0 && !(m.params[ia0] > 0.0)) || (m.kind == SBE_MODEL_WILSON_MASS && !(m.params[4] > 0.0))) {
            err = "lattice constants must be > 0";
            return SBE_ERR_MODEL;
        }
        if (m.nval < 0 || m.nval > m.nband) {
            err = "nval must be within 0..nband";
            return SBE_ERR_MODEL;
        }
        return SBE_OK;
    }
    if (m.nband < 1 || m.nband > kNbMaxBands) {
        err = "Wannier path supports 1..32 bands";
        return SBE_ERR_UNSUPPORTED;
    }
    if (m.nrpts < 1 || !m.rvec || !m.rweight || !m.ham) {
        err = "Wannier model needs nrpts >= 1, rvec, rweight and ham";
        return SBE_ERR_MODEL;
    }
    if (!m.pos && !m.is_dipole_zero) {
        err = "Wannier model with dipoles needs pos";
        return SBE_ERR_MODEL;
    }
    if (m.nval < 0 || m.nval > m.nband) {
        err = "nval must be within 0..nband";
        return SBE_ERR_MODEL;
    }
    return SBE_OK;
}

int nb_setup(sbe_ctx *ctx)
{
    const sbe_model &m = ctx->model;
    const bool analytic = am::is_analytic(m.kind);
    const int N = m.nband, nr = analytic ? 0 : m.nrpts;
    const int64_t nk = ctx->grid.k_end - ctx->grid.k_begin;
    const bool split16 = ctx->cfg.precision == SBE_PREC_SPLIT16;
    const BoxShape bx = box_shape(ctx->cfg, m, nk, split16);
    const bool fused = fused_eligible(ctx->cfg, m, nk, bx);
    const bool vfused = vel_fused_eligible(ctx->cfg, m, nk);
    const bool f32 = nb32_eligible(ctx->cfg, m);
    const NbPlan p = nb_plan(N, nr, nk, ctx->laser.nt, ctx->tab.nstage, split16, bx, fused, vfused, f32);
    const size_t nn = static_cast<size_t>(N) * N;
    ctx->nk = nk;
    ctx->nk_stride = p.nk_stride;
    ctx->nv = p.nv;
    ctx->chunk_steps = p.chunk;
    int rc = SBE_OK;
    ctx->nb_P = p.bytes_P ? nalloc<double2>(ctx, p.bytes_P, rc) : nullptr;
    ctx->nb_rho = nalloc<double2>(ctx, p.bytes_rho, rc);
    ctx->nb_mats = nalloc<double2>(ctx, p.bytes_mats, rc);
    ctx->nb_k = nalloc<double2>(ctx, p.bytes_k, rc);
    if (!rc) ctx->nb_U = ctx->nb_k + static_cast<size_t>(ctx->tab.nstage) * p.nk_stride * nn;
    char *tab = nalloc<char>(ctx, p.bytes_tab, rc);
    ctx->d_partials = nalloc<double>(ctx, p.bytes_partials, rc);
    ctx->d_series = nalloc<double>(ctx, p.bytes_series, rc);
    uint16_t *split = split16 ? nalloc<uint16_t>(ctx, p.bytes_split, rc) : nullptr;
    double2 *b64 = p.bytes_b64 ? nalloc<double2>(ctx, p.bytes_b64, rc) : nullptr;
    char *boxbuf = p.bytes_box ? nalloc<char>(ctx, p.bytes_box, rc) : nullptr;
    char *fbuf = p.bytes_fused ? nalloc<char>(ctx, p.bytes_fused, rc) : nullptr;
    double *vtab = p.bytes_vel ? nalloc<double>(ctx, p.bytes_vel, rc) : nullptr;
    if (rc) return rc;

    // host copies of the tables: k [nk_stride][3], w, R cartesian, weights, H, r
    std::vector<double> hk(static_cast<size_t>(p.nk_stride) * 4, 0.0);
    for (int64_t k = 0; k < nk; ++k) {
        for (int c = 0; c < 3; ++c) hk[3 * k + c] = ctx->grid.k[3 * k + c];
        hk[3 * p.nk_stride + k] = ctx->grid.w[k];
    }
    std::vector<double> hr(static_cast<size_t>(nr) * 4, 0.0);
    for (int r = 0; r < nr; ++r) {
        const int32_t *n = m.rvec + 3 * r;  // R = n1 a1 + n2 a2 + n3 a3 (Wannier90.h:193-204)
        for (int i = 0; i < 3; ++i)
            hr[3 * r + i] = n[0] * m.lattice[i] + n[1] * m.lattice[3 + i] + n[2] * m.lattice[6 + i];
        hr[3 * nr + r] = m.rweight[r];
    }
    const size_t bytes_k = hk.size() * sizeof(double), bytes_r = hr.size() * sizeof(double);
    double *dk = reinterpret_cast<double *>(tab);
    double *dr = reinterpret_cast<double *>(tab + bytes_k);
    double2 *dham = reinterpret_cast<double2 *>(tab + bytes_k + bytes_r);
    double2 *dpos = dham + static_cast<size_t>(nr) * nn;
    rc = hipc(ctx, hipMemcpy(dk, hk.data(), bytes_k, hipMemcpyHostToDevice), "upload k");
    if (!rc && nr) rc = hipc(ctx, hipMemcpy(dr, hr.data(), bytes_r, hipMemcpyHostToDevice), "upload R");
    if (!rc && nr) rc = hipc(ctx, hipMemcpy(dham, m.ham, static_cast<size_t>(nr) * nn * sizeof(double2), hipMemcpyHostToDevice), "upload H");
    // isDipoleZero: no E.D term (zero positions on the device; the R = 0 diagonal gives tau)
    if (!rc && nr && m.is_dipole_zero)
        rc = hipc(ctx, hipMemset(dpos, 0, static_cast<size_t>(nr) * 3 * nn * sizeof(double2)), "clear r");
    else if (!rc && nr)
        rc = hipc(ctx, hipMemcpy(dpos, m.pos, static_cast<size_t>(nr) * 3 * nn * sizeof(double2), hipMemcpyHostToDevice), "upload r");
    if (rc) return rc;

    NbArgs &a = ctx->nargs;
    a.N = N;
    a.nr = nr;
    a.nval = m.nval;
    a.fermi = ctx->cfg.init_type == SBE_INIT_FERMI_DIRAC;
    a.fermi_energy = ctx->cfg.fermi_energy;
    a.inter_intra = ctx->cfg.inter_intra ? 1 : 0;
    a.gam2 = ctx->cfg.T2 <= 1e-16 ? 0.0 : 1.0 / ctx->cfg.T2;  // SBEs.h:484-493
    a.gam1 = ctx->cfg.T1 <= 1e-16 ? 0.0 : 1.0 / ctx->cfg.T1;
    a.deph = (a.gam1 != 0.0 || a.gam2 != 0.0) ? 1 : 0;
    a.nk = nk;
    a.nk_stride = p.nk_stride;
    a.nv = p.nv;
    a.sample_waves = p.sample_waves;
    // N <= 16: stage and sample kernels on MFMA register tiles (ANTELOPE_NB_LDS=1 keeps the
    // LDS kernels, for comparison)
    {
        sbe_tuning tu{};
        const bool tuned = tuning_lookup(ctx->cfg, m, nk, tu) && tu.nb_kernels > 0;
        a.tile16 = N <= 16 && !(tuned && tu.nb_kernels == 2);
        if (const char *e = std::getenv("ANTELOPE_NB_LDS")) a.tile16 = N <= 16 && std::atoi(e) != 1;
    }
    a.kpts = dk;
    a.weight = dk + 3 * p.nk_stride;
    a.rcart = dr;
    a.rweight = dr + 3 * nr;
    a.ham = dham;
    a.pos = dpos;
    a.split16 = split16 ? 1 : 0;
    a.nr_pad = p.nr_pad;
    a.bcols = p.bcols;
    a.p16 = split;
    a.b16 = split ? split + 4 * static_cast<size_t>(p.nr_pad) * p.nk_stride : nullptr;
    a.b64 = b64;
    a.hmax = a.pmax = a.rmax = 0.0;
    for (size_t e = 0; e < static_cast<size_t>(nr) * nn; ++e)
        a.hmax = std::max(a.hmax, std::hypot(m.ham[2 * e], m.ham[2 * e + 1]));
    if (!m.is_dipole_zero)
        for (size_t e = 0; e < static_cast<size_t>(nr) * 3 * nn; ++e)
            a.pmax = std::max(a.pmax, std::hypot(m.pos[2 * e], m.pos[2 * e + 1]));
    // isDipoleZero (Wannier90.h:178-194): tau_n = Re <0n|r|0n> of the R = 0 block (0 without positions)
    a.dzero = !analytic && m.is_dipole_zero ? 1 : 0;
    for (int n = 0; n < kNbMaxBands; ++n)
        for (int c = 0; c < 3; ++c) a.tau[n][c] = 0.0;
    if (a.dzero && m.pos)
        for (int r = 0; r < nr; ++r) {
            const int32_t *nv = m.rvec + 3 * r;
            if (nv[0] || nv[1] || nv[2]) continue;
            for (int n = 0; n < N; ++n)
                for (int c = 0; c < 3; ++c) a.tau[n][c] = m.pos[2 * ((static_cast<size_t>(r) * 3 + c) * nn + n * N + n)];
        }
    for (int e = 0; e < 3 * nr; ++e) a.rmax = std::max(a.rmax, std::fabs(hr[e]));
    a.model_kind = analytic ? m.kind : 0;
    for (int i = 0; i < 24; ++i) a.mparams[i] = analytic ? m.params[i] : 0.0;
    a.box = 0;
    if (boxbuf) {  // separable Fourier sums when this block's k points form a product grid
        int Kact[3];
        std::vector<double> E(p.off_r / sizeof(double), 0.0);  // covers E_1..E_3 at their offsets
        size_t e_off[3];
        for (int i = 0; i < 3; ++i) e_off[i] = p.off_e[i] / sizeof(double);
        std::vector<int32_t> rmap, kinv;
        if (box_tables(m, hk, nk, bx, Kact, E, e_off, rmap, kinv)) {
            const size_t e_bytes = p.off_r - p.off_e[0];
            rc = hipc(ctx, hipMemcpy(boxbuf + p.off_e[0], E.data() + e_off[0], e_bytes, hipMemcpyHostToDevice), "upload box");
            if (!rc) rc = hipc(ctx, hipMemcpy(boxbuf + p.off_r, rmap.data(), rmap.size() * sizeof(int32_t), hipMemcpyHostToDevice), "upload box");
            if (!rc) rc = hipc(ctx, hipMemcpy(boxbuf + p.off_kinv, kinv.data(), kinv.size() * sizeof(int32_t), hipMemcpyHostToDevice), "upload box");
            if (rc) return rc;
            a.box = 1;
            for (int i = 0; i < 3; ++i) {
                a.boxM[i] = bx.M[i];
                a.boxK[i] = Kact[i];
                a.boxE[i] = reinterpret_cast<const double2 *>(boxbuf + p.off_e[i]);
            }
            a.box_r = reinterpret_cast<const int32_t *>(boxbuf + p.off_r);
            a.box_kinv = reinterpret_cast<const int32_t *>(boxbuf + p.off_kinv);
            a.box_b = reinterpret_cast<double2 *>(boxbuf);
            a.box_t1 = reinterpret_cast<double2 *>(boxbuf + p.off_t1);
            a.box_t2 = reinterpret_cast<double2 *>(boxbuf + p.off_t2);
        }
    }

    // fused RK4 steps: the planned grid must have grouped into the whole box (every node one k point)
    ctx->nbf = false;
    if (fbuf && a.box && a.tile16 && nb16_step_lds_bytes(N, bx.M[0], a.nv, fused_waves(bx)) <= (size_t(160) << 10)) {
        const size_t col = sizeof(double2) * static_cast<size_t>(p.bcols);
        const size_t t2 = align256(col * bx.M[0] * bx.K[1] * bx.K[2]);
        const size_t b = align256(col * static_cast<size_t>(bx.cells()));
        const size_t t1 = align256(col * bx.M[0] * bx.M[1] * bx.K[2]);
        const int64_t nt = ctx->laser.nt;
        const int S = nb16_steps_per_eval(t2, b, t1, nt);
        char *q = fbuf;
        ctx->nbf_S = S;
        ctx->nbf_t2_stride = static_cast<int64_t>(t2 / sizeof(double2));
        ctx->nbf_tj = reinterpret_cast<double2 *>(q);
        ctx->nbf_th = reinterpret_cast<double2 *>(q + S * t2);
        ctx->nbf_ring = reinterpret_cast<double2 *>(q + 2 * S * t2);
        q += (3 * static_cast<size_t>(S) + 1) * t2;
        ctx->nbf_lam = reinterpret_cast<double *>(q);
        q += align256(static_cast<size_t>(p.nk_stride) * 16 * sizeof(double));
        ctx->nbf_bs = reinterpret_cast<double2 *>(q);
        ctx->nbf_b_stride = static_cast<int64_t>(b / sizeof(double2));
        q += 3 * static_cast<size_t>(S) * b;
        ctx->nbf_t1s = reinterpret_cast<double2 *>(q);
        ctx->nbf_t1_stride = static_cast<int64_t>(t1 / sizeof(double2));
        q += 3 * static_cast<size_t>(S) * t1;
        double *lt = reinterpret_cast<double *>(q);
        std::vector<double> hl(static_cast<size_t>(nt) * 18);
        for (int64_t n = 0; n < nt; ++n) {  // the times of the step loop's evaluations
            const double tn = ctx->laser.atmin + ctx->laser.dt * static_cast<double>(n);  // antelope.cpp:299
            const double ts[3] = {tn, tn + 0.5 * ctx->laser.dt, tn + ctx->laser.dt};
            for (int j = 0; j < 3; ++j) ctx->laser.eval(ts[j], hl.data() + 18 * n + 6 * j, hl.data() + 18 * n + 6 * j + 3);
        }
        rc = hipc(ctx, hipMemcpy(lt, hl.data(), hl.size() * sizeof(double), hipMemcpyHostToDevice), "upload laser");
        if (rc) return rc;
        ctx->nbf_laser = lt;
        // blocks over the k points' own box (a k block of the grid may span fewer k1, k2, k3)
        ctx->nbf_waves = fused_waves(bx);
        ctx->nbf_g1 = (a.boxK[0] + ctx->nbf_waves - 1) / ctx->nbf_waves;
        ctx->nbf_nblocks = ctx->nbf_g1 * a.boxK[1] * a.boxK[2];
        ctx->nbf = true;
    }
    ctx->nbf_prev_valid = false;
    ctx->nbf_lam_valid = false;
    // fused RK4 steps for 16 < N <= 32: H~ H pairs in mats slots 0-1 and stage-derivative
    // slots 0-1, 2-3 (J in mats slots 2-4); with the separable sums, per evaluation scratch
    ctx->nb32 = false;
    if (f32 && fbuf && ctx->tab.nstage == 4) {
        const size_t slot = static_cast<size_t>(p.nk_stride) * nn;
        ctx->nb32_pair[0] = ctx->nb_mats;
        ctx->nb32_pair[1] = ctx->nb_k;
        ctx->nb32_pair[2] = ctx->nb_k + 2 * slot;
        ctx->nb32_lam = reinterpret_cast<double *>(fbuf);
        ctx->nb32_dm = nb32_dense_multi_n(nr, nk);
        if (ctx->nb32_dm) {
            const int64_t nt = ctx->laser.nt;
            const int S = nb32_steps_per_eval(p.nk_stride, N, nt);
            char *mb = fbuf + align256(static_cast<size_t>(p.nk_stride) * 32 * sizeof(double));
            ctx->nb32_S = S;
            ctx->nb32_j = reinterpret_cast<double2 *>(mb);
            ctx->nb32_half = ctx->nb32_j + 3 * static_cast<size_t>(S) * slot;
            ctx->nb32_ring = ctx->nb32_half + 2 * static_cast<size_t>(S) * slot;
            double *lt = reinterpret_cast<double *>(mb + align256((7 * static_cast<size_t>(S) + 2) * p.bytes_rho));
            std::vector<double> hl(static_cast<size_t>(nt) * 18);
            for (int64_t n = 0; n < nt; ++n) {  // the times of the step loop's evaluations
                const double tn = ctx->laser.atmin + ctx->laser.dt * static_cast<double>(n);  // antelope.cpp:299
                const double ts[3] = {tn, tn + 0.5 * ctx->laser.dt, tn + ctx->laser.dt};
                for (int j = 0; j < 3; ++j) ctx->laser.eval(ts[j], hl.data() + 18 * n + 6 * j, hl.data() + 18 * n + 6 * j + 3);
            }
            rc = hipc(ctx, hipMemcpy(lt, hl.data(), hl.size() * sizeof(double), hipMemcpyHostToDevice), "upload laser");
            if (rc) return rc;
            ctx->nb32_laser = lt;
        }
        for (int i = 0; i < 4; ++i) ctx->nb32_b[i] = ctx->nb32_t1[i] = ctx->nb32_t2[i] = nullptr;
        if (a.box && !ctx->nb32_dm) {
            const size_t col = sizeof(double2) * static_cast<size_t>(p.bcols);
            const size_t b = align256(col * static_cast<size_t>(bx.cells()));
            const size_t t1 = align256(col * bx.M[0] * bx.M[1] * bx.K[2]);
            const size_t t2 = align256(col * bx.M[0] * bx.K[1] * bx.K[2]);
            char *sc = fbuf + align256(static_cast<size_t>(p.nk_stride) * 32 * sizeof(double));
            for (int i = 0; i < 4; ++i) {
                ctx->nb32_b[i] = reinterpret_cast<double2 *>(sc + i * (b + t1 + t2));
                ctx->nb32_t1[i] = reinterpret_cast<double2 *>(sc + i * (b + t1 + t2) + b);
                ctx->nb32_t2[i] = reinterpret_cast<double2 *>(sc + i * (b + t1 + t2) + b + t1);
            }
        }
        ctx->nb32_blocks = static_cast<int>(std::min<int64_t>(nk, kNb32MaxBlocks));
        ctx->nb32 = true;
    }

    a.velocity = ctx->velocity ? 1 : 0;
    // e^{ik.R}, then rho(k) from H(k) (A(atmin) = 0)
    if (!analytic && !a.box && boxbuf) {
        // planned for the separable sums, but the k points did not group into a product
        // grid (box_shape rules this out for the grids this library builds): the dense
        // sums' e^{ik.R} table and B buffer replace the box buffer, as the dense plan has them
        nfree(ctx, boxbuf);
        boxbuf = nullptr;
        if (!ctx->nb_P) ctx->nb_P = nalloc<double2>(ctx, static_cast<size_t>(p.nk_stride) * nr * sizeof(double2), rc);
        a.b64 = nalloc<double2>(ctx, sizeof(double2) * static_cast<size_t>(p.nr_pad) * p.bcols, rc);
        if (rc) return rc;
    }
    if (!analytic && ctx->nb_P) rc = hipc(ctx, nb_launch_phase(a, ctx->nb_P, ctx->stream), "phase kernel");
    if (!rc && split16) rc = hipc(ctx, nb_launch_psplit(a, ctx->nb_P, split, ctx->stream), "phase split kernel");
    NbEval ev{};
    ev.nmat = 1;
    ev.mats[0] = kMatH0;
    if (ctx->velocity) {
        // VelocityHamiltonian: U, E, pMatrix at k once (SBEs.h:284-373); the
        // stage derivatives live in LDS, so nb_k holds P [3] and (E, occ) instead
        ev.nmat = 4;
        ev.mats[1] = kMatJx;
        ev.mats[2] = kMatJy;
        ev.mats[3] = kMatJz;
        double2 *Pv = ctx->nb_k;
        double *Eocc = reinterpret_cast<double *>(ctx->nb_k + 3 * static_cast<size_t>(p.nk_stride) * nn);
        if (!rc) rc = hipc(ctx, eval_mats(a, ev, ctx->nb_P, ctx->nb_mats, ctx->stream), "fourier kernel");
        if (!rc) rc = hipc(ctx, nb_launch_vel_init(a, ctx->nb_mats, Pv, Eocc, ctx->nb_U, ctx->nb_rho, ctx->stream), "velocity init kernel");
        const size_t lds_p = nb_vel_lds_bytes(N, ctx->tab.nstage, true);
        ctx->nb_vel_p_lds = lds_p <= (size_t(64) << 10);
        ctx->nbv = false;
        if (vtab && (a.tile16 || N > 16)) {  // A at the four RK4 stage times of every step, as the host evaluates them
            const int64_t nt = ctx->laser.nt;
            std::vector<double> hv(static_cast<size_t>(nt) * 12);
            for (int64_t n = 0; n < nt; ++n) {
                const double tn = ctx->laser.atmin + ctx->laser.dt * static_cast<double>(n);  // antelope.cpp:299
                for (int st = 0; st < 4; ++st) {
                    double E[3];
                    ctx->laser.eval(tn + ctx->tab.c[st] * ctx->laser.dt, hv.data() + 12 * n + 3 * st, E);  // SBEs.h:612
                }
            }
            if (!rc) rc = hipc(ctx, hipMemcpy(vtab, hv.data(), hv.size() * sizeof(double), hipMemcpyHostToDevice), "upload A");
            ctx->nbv_A = vtab;
            ctx->nbv_waves = vel_waves(nk);
            ctx->nbv_nblocks = static_cast<int>(N > 16 ? nk : (nk + ctx->nbv_waves - 1) / ctx->nbv_waves);
            ctx->nbv = true;
        }
        if (!rc) rc = hipc(ctx, hipStreamSynchronize(ctx->stream), "init sync");
        return rc;
    }
    // H(k) = H(k + A(atmin)) into slot 1, where the first sample finds it
    double2 *h_slot = ctx->nb_mats + static_cast<size_t>(p.nk_stride) * nn;
    if (!rc) rc = hipc(ctx, eval_mats(a, ev, ctx->nb_P, h_slot, ctx->stream), "fourier kernel");
    if (!rc) rc = hipc(ctx, nb_launch_init(a, h_slot, ctx->nb_rho, ctx->nb_U, ctx->stream), "init kernel");
    ctx->nb_u_time = ctx->laser.atmin;  // U diagonalises H(k) = H(k + A(atmin))
    ctx->nb_h_time = ctx->laser.atmin;
    ctx->nb_h_valid = true;
    ctx->nb_ht_valid = false;
    if (!rc) rc = hipc(ctx, hipStreamSynchronize(ctx->stream), "init sync");
    return rc;
}

namespace {

// VelocityHamiltonian gauge: one fused sample + RK-step launch per time step
int nb_run_velocity(sbe_ctx *ctx, int64_t step_begin, int64_t step_end)
{
    const NbArgs &a = ctx->nargs;
    const Tableau &tb = ctx->tab;
    const double dt = ctx->laser.dt;
    const size_t nn = static_cast<size_t>(a.N) * a.N;
    const double2 *Pv = ctx->nb_k;
    const double *Eocc = reinterpret_cast<const double *>(ctx->nb_k + 3 * static_cast<size_t>(a.nk_stride) * nn);
    if (ctx->nbv) {  // the steps of a chunk in one launch
        int rc = SBE_OK;
        for (int64_t c0 = step_begin; c0 < step_end && !rc; c0 += ctx->chunk_steps) {
            const int64_t n = std::min<int64_t>(ctx->chunk_steps, step_end - c0);
            NbVelFused f{};
            f.A = ctx->nbv_A;
            f.s0 = c0;
            f.s1 = c0 + n;
            f.dt = dt;
            f.nblocks = ctx->nbv_nblocks;
            f.waves = ctx->nbv_waves;
            rc = hipc(ctx,
                      a.N > 16 ? nb32_launch_vel_fused(a, f, Pv, Eocc, ctx->nb_rho, ctx->d_partials, ctx->stream)
                               : nb16_launch_vel_fused(a, f, Pv, Eocc, ctx->nb_rho, ctx->d_partials, ctx->stream),
                      "velocity step kernel");
            if (!rc)
                rc = hipc(ctx, launch_series_reduce(ctx->d_partials, n, ctx->nbv_nblocks, a.nv, ctx->d_series + c0 * a.nv,
                                                    ctx->stream, ctx->nonfinite_d, kNbReduceGroup),
                          "series kernel");
            if (!rc) ctx->steps_done = c0 + n;
        }
        return rc;
    }
    NbVelStep v{};
    v.nstage = tb.nstage;
    for (int i = 0; i < tb.nstage; ++i) {
        for (int j = 0; j < i; ++j) v.a[i][j] = dt * tb.a[i][j];
        v.b[i] = dt * tb.b[i];
    }
    int rc = SBE_OK;
    for (int64_t c0 = step_begin; c0 < step_end && !rc; c0 += ctx->chunk_steps) {
        const int64_t n = std::min<int64_t>(ctx->chunk_steps, step_end - c0);
        for (int64_t s = 0; s < n && !rc; ++s) {
            const double tn = ctx->laser.atmin + dt * static_cast<double>(c0 + s);  // antelope.cpp:299
            for (int i = 0; i < tb.nstage; ++i) {
                double E[3];
                ctx->laser.eval(tn + tb.c[i] * dt, v.A[i], E);  // SBEs.h:612
            }
            rc = hipc(ctx, nb_launch_vel_step(a, v, Pv, Eocc, ctx->nb_rho, ctx->d_partials, s, ctx->nb_vel_p_lds, ctx->stream),
                      "velocity step kernel");
        }
        if (!rc)
            rc = hipc(ctx, launch_series_reduce(ctx->d_partials, n, a.sample_waves, a.nv, ctx->d_series + c0 * a.nv, ctx->stream,
                                          ctx->nonfinite_d, kNbReduceGroup),
                      "series kernel");
        if (!rc) ctx->steps_done = c0 + n;
    }
    return rc;
}

}  // namespace

namespace {

// LengthWannier RK4 steps as one nb16_step_kernel launch each (after the
// Fourier passes of the step's times)
int nb_run_fused(sbe_ctx *ctx, int64_t step_begin, int64_t step_end)
{
    const NbArgs &a = ctx->nargs;
    const double dt = ctx->laser.dt;
    const int nring = ctx->nbf_S + 1;
    const int64_t t2s = ctx->nbf_t2_stride;
    auto pair = [&](int64_t step) { return ctx->nbf_ring + ((step % nring + nring) % nring) * t2s; };
    int rc = SBE_OK;
    for (int64_t c0 = step_begin; c0 < step_end && !rc; c0 += ctx->chunk_steps) {
        const int64_t n = std::min<int64_t>(ctx->chunk_steps, step_end - c0);
        for (int64_t s = 0; s < n && !rc;) {
            // the separable sums of up to nbf_S steps' times (they do not depend on rho), then
            // one step launch per step
            const int ns = static_cast<int>(std::min<int64_t>(ctx->nbf_S, n - s));
            const int64_t step0 = c0 + s;
            const double t0 = ctx->laser.atmin + dt * static_cast<double>(step0);  // antelope.cpp:299
            if (!(ctx->nbf_prev_valid && same_time(ctx->nbf_prev_time, t0, dt))) {  // H~, H at t_n of the first step
                NbEval ev{};
                ctx->laser.eval(t0, ev.A, ev.E);
                ev.nmat = 2;
                ev.mats[0] = kMatHtilde;
                ev.mats[1] = kMatH0;
                double2 *out = pair(step0 - 1);
                rc = hipc(ctx, nb_launch_box_t2_multi(a, 1, &ev, &ctx->nbf_bs, &ctx->nbf_t1s, &out, ctx->stream),
                          "fourier passes");
            }
            NbBoxSteps d{};
            d.laser = ctx->nbf_laser;
            d.s0 = step0;
            d.ns = ns;
            d.nring = nring;
            d.b = ctx->nbf_bs;
            d.t1 = ctx->nbf_t1s;
            d.tj = ctx->nbf_tj;
            d.th = ctx->nbf_th;
            d.ring = ctx->nbf_ring;
            d.b_stride = ctx->nbf_b_stride;
            d.t1_stride = ctx->nbf_t1_stride;
            d.t2_stride = t2s;
            if (!rc) rc = hipc(ctx, nb_launch_box_steps(a, d, ctx->stream), "fourier passes");
            for (int i = 0; i < ns && !rc; ++i) {
                const int64_t step = step0 + i;
                const double tn = ctx->laser.atmin + dt * static_cast<double>(step);
                NbFused f{};
                f.t2_prev = pair(step - 1);
                f.t2_j = ctx->nbf_tj + i * t2s;
                f.t2_half = ctx->nbf_th + i * t2s;
                f.t2_end = pair(step);
                f.lam = ctx->nbf_lam;
                f.u_valid = same_time(ctx->nb_u_time, tn, dt) ? 1 : 0;
                f.lam_valid = f.u_valid && ctx->nbf_lam_valid && same_time(ctx->nbf_lam_time, tn, dt) ? 1 : 0;
                f.dt = dt;
                f.step_local = s + i;
                f.g1 = ctx->nbf_g1;
                f.nblocks = ctx->nbf_nblocks;
                f.waves = ctx->nbf_waves;
                rc = hipc(ctx, nb16_launch_step(a, f, ctx->nb_rho, ctx->nb_U, ctx->d_partials, ctx->stream), "step kernel");
                // the end of this step is the start of the next: its H~, H serve there
                ctx->nbf_prev_time = tn + dt;
                ctx->nbf_prev_valid = true;
                ctx->nb_u_time = tn + dt;
                ctx->nbf_lam_time = tn + dt;
                ctx->nbf_lam_valid = true;
                ctx->nb_ht_valid = false;  // the separate kernels' matrices are not kept up to date
                ctx->nb_h_valid = false;
            }
            s += ns;
        }
        if (!rc)
            rc = hipc(ctx, launch_series_reduce(ctx->d_partials, n, ctx->nbf_nblocks, a.nv, ctx->d_series + c0 * a.nv,
                                                ctx->stream, ctx->nonfinite_d, kNbReduceGroup),
                      "series kernel");
        if (!rc) ctx->steps_done = c0 + n;
    }
    return rc;
}

// LengthWannier RK4 steps for 16 < N <= 32: the matrices of the step's times (one
// batch of launches with the separable sums), then one nb32_step_kernel launch
int nb_run_fused32(sbe_ctx *ctx, int64_t step_begin, int64_t step_end)
{
    const NbArgs &a = ctx->nargs;
    const double dt = ctx->laser.dt;
    const size_t slot = static_cast<size_t>(a.nk_stride) * a.N * a.N;
    int rc = SBE_OK;
    auto eval = [&](double t, bool hmats, bool h0) {
        NbEval ev{};
        ctx->laser.eval(t, ev.A, ev.E);
        if (hmats) {
            ev.nmat = h0 ? 2 : 1;
            ev.mats[0] = kMatHtilde;
            ev.mats[1] = kMatH0;
        } else {
            ev.nmat = 3;
            ev.mats[0] = kMatJx;
            ev.mats[1] = kMatJy;
            ev.mats[2] = kMatJz;
        }
        return ev;
    };
    double2 *jm = ctx->nb_mats + 2 * slot;
    auto launch_step = [&](const NbStep32 &f0, double tn) {
        NbStep32 f = f0;
        f.lam = ctx->nb32_lam;
        f.u_valid = same_time(ctx->nb_u_time, tn, dt) ? 1 : 0;
        f.lam_valid = f.u_valid && ctx->nbf_lam_valid && same_time(ctx->nbf_lam_time, tn, dt) ? 1 : 0;
        f.dt = dt;
        const int r = hipc(ctx, nb32_launch_step(a, f, ctx->nb32_blocks, ctx->nb_rho, ctx->nb_U, ctx->d_partials, ctx->stream),
                           "step kernel");
        ctx->nbf_prev_time = tn + dt;
        ctx->nbf_prev_valid = true;
        ctx->nb_u_time = tn + dt;
        ctx->nbf_lam_time = tn + dt;
        ctx->nbf_lam_valid = true;
        ctx->nb_ht_valid = false;  // the separate kernels' matrices are not kept up to date
        ctx->nb_h_valid = false;
        return r;
    };
    if (ctx->nb32_dm) {  // the matrices of up to nb32_S steps per launch; one step launch per step
        const int nring = ctx->nb32_S + 1;
        auto pair = [&](int64_t step) { return ctx->nb32_ring + ((step % nring + nring) % nring) * 2 * slot; };
        for (int64_t c0 = step_begin; c0 < step_end && !rc; c0 += ctx->chunk_steps) {
            const int64_t n = std::min<int64_t>(ctx->chunk_steps, step_end - c0);
            for (int64_t s = 0; s < n && !rc;) {
                const int ns = static_cast<int>(std::min<int64_t>(ctx->nb32_S, n - s));
                const int64_t step0 = c0 + s;
                const double t0 = ctx->laser.atmin + dt * static_cast<double>(step0);  // antelope.cpp:299
                if (!(ctx->nbf_prev_valid && same_time(ctx->nbf_prev_time, t0, dt))) {
                    NbEval ev = eval(t0, true, true);
                    double2 *out = pair(step0 - 1);
                    rc = hipc(ctx, nb_launch_dense_multi(a, 1, &ev, ctx->nb_P, &out, ctx->stream), "fourier kernel");
                }
                NbDenseSteps d{};
                d.laser = ctx->nb32_laser;
                d.s0 = step0;
                d.ns = ns;
                d.nring = nring;
                d.deph = a.deph;
                d.slot = static_cast<int64_t>(slot);
                d.j = ctx->nb32_j;
                d.half = ctx->nb32_half;
                d.ring = ctx->nb32_ring;
                if (!rc) rc = hipc(ctx, nb_launch_dense_steps(a, d, ctx->nb_P, ctx->stream), "fourier kernel");
                for (int i = 0; i < ns && !rc; ++i) {
                    const int64_t step = step0 + i;
                    const double tn = ctx->laser.atmin + dt * static_cast<double>(step);
                    NbStep32 f{};
                    f.ht_prev = pair(step - 1);
                    f.h_prev = f.ht_prev + slot;
                    f.j = ctx->nb32_j + 3 * static_cast<size_t>(i) * slot;
                    f.ht_half = ctx->nb32_half + 2 * static_cast<size_t>(i) * slot;
                    f.h_half = f.ht_half + slot;
                    f.ht_end = pair(step);
                    f.h_end = f.ht_end + slot;
                    f.step_local = s + i;
                    rc = launch_step(f, tn);
                }
                s += ns;
            }
            if (!rc)
                rc = hipc(ctx, launch_series_reduce(ctx->d_partials, n, ctx->nb32_blocks, a.nv, ctx->d_series + c0 * a.nv,
                                                    ctx->stream, ctx->nonfinite_d, kNbReduceGroup),
                          "series kernel");
            if (!rc) ctx->steps_done = c0 + n;
        }
        return rc;
    }
    for (int64_t c0 = step_begin; c0 < step_end && !rc; c0 += ctx->chunk_steps) {
        const int64_t n = std::min<int64_t>(ctx->chunk_steps, step_end - c0);
        for (int64_t s = 0; s < n && !rc; ++s) {
            const double tn = ctx->laser.atmin + dt * static_cast<double>(c0 + s);  // antelope.cpp:299
            NbEval evs[4];
            double2 *outs[4];
            int nev = 0;
            if (!(ctx->nbf_prev_valid && same_time(ctx->nbf_prev_time, tn, dt))) {
                evs[nev] = eval(tn, true, true);
                outs[nev++] = ctx->nb32_pair[0];
            }
            evs[nev] = eval(tn, false, false);
            outs[nev++] = jm;
            evs[nev] = eval(tn + 0.5 * dt, true, a.deph != 0);  // H at the midpoint only for dephasing
            outs[nev++] = ctx->nb32_pair[1];
            evs[nev] = eval(tn + dt, true, true);
            outs[nev++] = ctx->nb32_pair[2];
            if (a.box) {
                rc = hipc(ctx, nb_launch_box_mats_multi(a, nev, evs, ctx->nb32_b, ctx->nb32_t1, ctx->nb32_t2, outs,
                                                        ctx->stream),
                          "fourier passes");
            } else {
                for (int e = 0; e < nev && !rc; ++e)
                    rc = hipc(ctx, eval_mats(a, evs[e], ctx->nb_P, outs[e], ctx->stream), "fourier kernel");
            }
            if (rc) break;
            NbStep32 f{};
            f.ht_prev = ctx->nb32_pair[0];
            f.h_prev = ctx->nb32_pair[0] + slot;
            f.j = jm;
            f.ht_half = ctx->nb32_pair[1];
            f.h_half = ctx->nb32_pair[1] + slot;
            f.ht_end = ctx->nb32_pair[2];
            f.h_end = ctx->nb32_pair[2] + slot;
            f.step_local = s;
            rc = launch_step(f, tn);
            // the end of this step is the start of the next: its H~, H serve there
            std::swap(ctx->nb32_pair[0], ctx->nb32_pair[2]);
        }
        if (!rc)
            rc = hipc(ctx, launch_series_reduce(ctx->d_partials, n, ctx->nb32_blocks, a.nv, ctx->d_series + c0 * a.nv,
                                                ctx->stream, ctx->nonfinite_d, kNbReduceGroup),
                      "series kernel");
        if (!rc) ctx->steps_done = c0 + n;
    }
    return rc;
}

}  // namespace

int nb_run(sbe_ctx *ctx, int64_t step_begin, int64_t step_end)
{
    if (ctx->velocity) return nb_run_velocity(ctx, step_begin, step_end);
    if (ctx->nbf) return nb_run_fused(ctx, step_begin, step_end);
    if (ctx->nb32) return nb_run_fused32(ctx, step_begin, step_end);
    const NbArgs &a = ctx->nargs;
    const Tableau &tb = ctx->tab;
    const double dt = ctx->laser.dt;
    int rc = SBE_OK;
    for (int64_t c0 = step_begin; c0 < step_end && !rc; c0 += ctx->chunk_steps) {
        const int64_t n = std::min<int64_t>(ctx->chunk_steps, step_end - c0);
        for (int64_t s = 0; s < n && !rc; ++s) {
            const int64_t step = c0 + s;
            const double tn = ctx->laser.atmin + dt * static_cast<double>(step);  // antelope.cpp:299
            // matrix slots: 0 H~, 1 H, 2..4 J. The last stage of a step is at the next sample's
            // time, so its H (and H~, and the eigenvectors) serve the sample and the first
            // stage of the next step; times match to rounding (t_n + dt against t_n+1)
            const size_t slot = static_cast<size_t>(a.nk_stride) * a.N * a.N;
            NbEval ev{};
            double Ev[3];
            ctx->laser.eval(tn, ev.A, Ev);
            const bool h_ok = ctx->nb_h_valid && same_time(ctx->nb_h_time, tn, dt);
            ev.nmat = 0;
            if (!h_ok) ev.mats[ev.nmat++] = kMatH0;
            ev.mats[ev.nmat++] = kMatJx;
            ev.mats[ev.nmat++] = kMatJy;
            ev.mats[ev.nmat++] = kMatJz;
            // eigenvectors: reused when they already belong to H(k + A(t_n)),
            // otherwise warm-started from the stored ones
            const int ureuse = same_time(ctx->nb_u_time, tn, dt);
            rc = hipc(ctx, eval_mats(a, ev, ctx->nb_P, ctx->nb_mats + (h_ok ? 2 : 1) * slot, ctx->stream), "fourier kernel");
            if (!rc)
                rc = hipc(ctx, nb_launch_sample(a, ctx->nb_mats, ctx->nb_rho, ctx->d_partials, s, ctx->nb_U, ureuse, ctx->stream),
                          "sample kernel");
            ctx->nb_u_time = tn;
            ctx->nb_h_time = tn;
            ctx->nb_h_valid = true;
            for (int i = 0; i < tb.nstage && !rc; ++i) {
                NbEval se{};
                const double ts = tn + tb.c[i] * dt;
                ctx->laser.eval(ts, se.A, se.E);  // SBEs.h:612
                NbStage sg{};
                sg.stage = i;
                sg.nstage = tb.nstage;
                sg.last = i == tb.nstage - 1;
                sg.ureuse = same_time(ctx->nb_u_time, ts, dt);
                for (int j = 0; j < i; ++j) sg.a[j] = dt * tb.a[i][j];
                for (int j = 0; j < tb.nstage; ++j) sg.b[j] = dt * tb.b[j];
                // H~ unless slot 0 already holds it for this time (RK4 stages 1 and 2 share a
                // time; the last stage's serves the next step's first), and H when new
                // eigenvectors are needed and slot 1 does not hold it
                const bool need_ht = !(ctx->nb_ht_valid && same_time(ctx->nb_ht_time, ts, dt));
                const bool need_h = a.deph && !sg.ureuse && !(ctx->nb_h_valid && same_time(ctx->nb_h_time, ts, dt));
                if (need_ht || need_h) {
                    se.nmat = 0;
                    if (need_ht) se.mats[se.nmat++] = kMatHtilde;
                    if (need_h) se.mats[se.nmat++] = kMatH0;
                    rc = hipc(ctx, eval_mats(a, se, ctx->nb_P, ctx->nb_mats + (need_ht ? 0 : slot), ctx->stream),
                              "fourier kernel");
                    if (need_ht) {
                        ctx->nb_ht_time = ts;
                        ctx->nb_ht_valid = true;
                    }
                    if (need_h) {
                        ctx->nb_h_time = ts;
                        ctx->nb_h_valid = true;
                    }
                }
                // the next stage at the same time (RK4's c = 1/2, 1/2) runs in the same pass on
                // the register-tile kernels, sharing H~, U and rho
                const bool pair = a.tile16 && i + 1 < tb.nstage && tb.c[i + 1] == tb.c[i];
                if (!rc && pair) {
                    NbStage sg2{};
                    sg2.stage = i + 1;
                    sg2.nstage = tb.nstage;
                    sg2.last = i + 1 == tb.nstage - 1;
                    sg2.ureuse = 1;
                    for (int j = 0; j <= i; ++j) sg2.a[j] = dt * tb.a[i + 1][j];
                    for (int j = 0; j < tb.nstage; ++j) sg2.b[j] = dt * tb.b[j];
                    rc = hipc(ctx, nb16_launch_rhs_pair(a, sg, sg2, ctx->nb_mats, ctx->nb_rho, ctx->nb_k, ctx->nb_U,
                                                        ctx->stream),
                              "rhs kernel");
                    ++i;
                } else if (!rc) {
                    rc = hipc(ctx, nb_launch_rhs(a, sg, ctx->nb_mats, ctx->nb_rho, ctx->nb_k, ctx->nb_U, ctx->stream),
                              "rhs kernel");
                }
                if (a.deph) ctx->nb_u_time = ts;
            }
        }
        if (!rc)
            rc = hipc(ctx, launch_series_reduce(ctx->d_partials, n, a.sample_waves, a.nv, ctx->d_series + c0 * a.nv, ctx->stream,
                                          ctx->nonfinite_d, kNbReduceGroup),
                      "series kernel");
        if (!rc) ctx->steps_done = c0 + n;
    }
    return rc;
}

int nb_series(sbe_ctx *ctx, double *j_inter, double *j_intra, double *rho_bar, double *band_pop)
{
    const int N = ctx->N, nv = ctx->nv;
    const int64_t nt = ctx->laser.nt;
    std::vector<double> s(static_cast<size_t>(nt) * nv);
    int rc = hipc(ctx, hipMemcpyAsync(s.data(), ctx->d_series, s.size() * sizeof(double), hipMemcpyDeviceToHost, ctx->stream), "copy series");
    if (!rc) rc = hipc(ctx, hipStreamSynchronize(ctx->stream), "sync series");
    if (rc) return rc;
    bool finite = true;
    for (int64_t n = 0; n < nt; ++n) {
        const double *r = s.data() + n * nv;
        if (n < ctx->steps_done)
            for (int i = 0; i < nv; ++i) finite = finite && std::isfinite(r[i]);
        double dia[3] = {0.0, 0.0, 0.0};  // velocity gauge: sum_k w_k Nval A(t_n) (SBEs.h:683-690,750-757)
        if (ctx->velocity && n < ctx->steps_done) {
            double A[3], E[3];
            ctx->laser.eval(ctx->laser.atmin + ctx->laser.dt * static_cast<double>(n), A, E);
            for (int ax = 0; ax < 3; ++ax) dia[ax] = ctx->model.nval * A[ax] * ctx->grid.total_weight;
        }
        for (int ax = 0; ax < 3; ++ax) {
            if (j_inter) j_inter[3 * n + ax] = r[ax] + (ctx->cfg.inter_intra ? 0.0 : dia[ax]);
            if (j_intra) j_intra[3 * n + ax] = ctx->cfg.inter_intra ? r[3 + ax] + dia[ax] : 0.0;
        }
        if (band_pop)
            for (int m = 0; m < N; ++m) band_pop[n * N + m] = r[6 + m];
        if (rho_bar) std::memcpy(rho_bar + n * 2 * N * N, r + 6 + N, sizeof(double) * 2 * N * N);
    }
    if (!finite) {
        ctx->err = "non-finite value in the propagated state";
        return SBE_ERR_NAN;
    }
    return SBE_OK;
}

int nb_state(sbe_ctx *ctx, int basis, double *rho)
{
    if (basis != SBE_BASIS_ORBITAL && basis != SBE_BASIS_BAND) {
        ctx->err = "unknown basis";
        return SBE_ERR_CONFIG;
    }
    // LengthWannier propagates the orbital basis, VelocityHamiltonian the band basis at k
    const int native = ctx->velocity ? SBE_BASIS_BAND : SBE_BASIS_ORBITAL;
    const int N = ctx->N;
    const size_t nn = static_cast<size_t>(N) * N;
    const size_t bytes = static_cast<size_t>(ctx->nk) * nn * sizeof(double2);
    if (basis != native && !ctx->velocity) {
        // U^dag rho U at k + A(t) for the next sample time t (WannierToHamiltonian, SBEs.h:927-946):
        // H into mats slot 0, the rotated state into slot 1 (slot 0 then no longer holds H~)
        const NbArgs &a = ctx->nargs;
        NbEval ev{};
        double E[3];
        ctx->laser.eval(ctx->laser.atmin + ctx->laser.dt * static_cast<double>(ctx->steps_done), ev.A, E);
        ev.nmat = 1;
        ev.mats[0] = kMatH0;
        double2 *out = ctx->nb_mats + static_cast<size_t>(a.nk_stride) * nn;
        ctx->nb_ht_valid = false;
        ctx->nb_h_valid = false;
        if (ctx->nb32 && !ctx->nb32_dm) ctx->nbf_prev_valid = false;  // an H~ H pair may live in slots 0-1
        int rc = hipc(ctx, eval_mats(a, ev, ctx->nb_P, ctx->nb_mats, ctx->stream), "fourier kernel");
        if (!rc) rc = hipc(ctx, nb_launch_band_export(a, ctx->nb_mats, ctx->nb_rho, out, ctx->stream), "band export kernel");
        if (!rc) rc = hipc(ctx, hipMemcpyAsync(rho, out, bytes, hipMemcpyDeviceToHost, ctx->stream), "copy state");
        if (!rc) rc = hipc(ctx, hipStreamSynchronize(ctx->stream), "sync state");
        return rc;
    }
    int rc = hipc(ctx, hipMemcpyAsync(rho, ctx->nb_rho, bytes, hipMemcpyDeviceToHost, ctx->stream), "copy state");
    std::vector<double2> U;
    if (!rc && basis != native) {  // velocity gauge, orbital basis: U rho U^dag with the stored U(k)
        U.resize(static_cast<size_t>(ctx->nk) * nn);
        rc = hipc(ctx, hipMemcpyAsync(U.data(), ctx->nb_U, bytes, hipMemcpyDeviceToHost, ctx->stream), "copy U");
    }
    if (!rc) rc = hipc(ctx, hipStreamSynchronize(ctx->stream), "sync state");
    if (rc || basis == native) return rc;
    std::vector<double> T(2 * nn);
    for (int64_t k = 0; k < ctx->nk; ++k) {
        const double2 *u = U.data() + k * nn;
        double *r = rho + 2 * k * nn;
        for (int i = 0; i < N; ++i)  // T = rho U^dag
            for (int j = 0; j < N; ++j) {
                double re = 0, im = 0;
                for (int l = 0; l < N; ++l) {
                    const double ar = r[2 * (i * N + l)], ai = r[2 * (i * N + l) + 1];
                    const double br = u[j * N + l].x, bi = -u[j * N + l].y;
                    re += ar * br - ai * bi;
                    im += ar * bi + ai * br;
                }
                T[2 * (i * N + j)] = re;
                T[2 * (i * N + j) + 1] = im;
            }
        for (int i = 0; i < N; ++i)  // rho = U T
            for (int j = 0; j < N; ++j) {
                double re = 0, im = 0;
                for (int l = 0; l < N; ++l) {
                    const double ar = u[i * N + l].x, ai = u[i * N + l].y;
                    const double br = T[2 * (l * N + j)], bi = T[2 * (l * N + j) + 1];
                    re += ar * br - ai * bi;
                    im += ar * bi + ai * br;
                }
                r[2 * (i * N + j)] = re;
                r[2 * (i * N + j) + 1] = im;
            }
    }
    return SBE_OK;
}

}  // namespace sbe
