// noalloc.hip - the 2-band fixed-step propagation over caller-owned device
// buffers, with no device allocation: sbe_run_haldane_noalloc.
//
// sbe_create + sbe_run allocate the context's k points, state, laser table,
// partial-sum rows and series with hipMalloc. Callers that cannot allocate
// (code under HIP graph capture, fixed-memory serving loops) pass the k
// points, the state and the series buffer themselves; the two buffers that
// are internal to the library (the per-stage laser phase rows and the
// per-wave partial-sum rows of the k sums) live in one file-scope device
// array, `scratch`, of kScratchBytes. The laser rows are evaluated on the
// device from the pulse constants (passed by value), so the call issues no
// host-to-device copy either: it only enqueues kernels on the caller's
// stream. The step and reduction kernels are those of sbe_run.
//
// Every call on one device shares that device's array and leaves its last
// contents behind: calls that overlap on one GPU (two streams, two threads)
// race on it. Nothing in it has to start at zero: each launch writes all the
// table rows and partial rows it reads.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <mutex>
#include <string>

#include "sbe_internal.hpp"

namespace sbe {
namespace {

constexpr size_t kScratchBytes = size_t(SBE_NOALLOC_SCRATCH_BYTES);
constexpr int kMaxDevices = 64;

static __device__ __align__(256) unsigned char scratch[kScratchBytes];

// Device address of `scratch` per device, looked up once per device with
// that device current (each device holds its own copy of the array).
std::mutex g_mu;
void *g_scratch[kMaxDevices] = {};

int scratch_address(int dev, void **out, std::string &err)
{
    if (dev < 0 || dev >= kMaxDevices) {
        err = "device index outside the scratch table";
        return SBE_ERR_RANGE;
    }
    std::lock_guard<std::mutex> lk(g_mu);
    if (!g_scratch[dev]) {
        void *p = nullptr;
        const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !p) {
            err = std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e);
            return SBE_ERR_HIP;
        }
        g_scratch[dev] = p;
    }
    *out = g_scratch[dev];
    return SBE_OK;
}

// Everything the table kernel needs, by value (about 1.4 KB of kernel
// arguments): the pulses and the stage times of the tableau.
struct TableArgs {
    PulseConst pulses[kMaxDevicePulses];
    int32_t npulses;
    int32_t consistent;
    int32_t nstage;
    double A_at_min[3];
    double atmin, dt, t1;
    double c[7];
    double va[3][2], vb[3][2];
    int64_t step0, nsteps;
};

// Row (step, stage) of the per-stage phase table, the same quantities the
// host writes for a context: t1 e^{iA.a_j} (j < 3) and e^{iA.b_j} as (re, im)
// pairs, at t_n + c_s dt (the stage at c = 1 at t_{n+1} itself), A the sum of
// the pulses less its value at atmin.
__global__ __launch_bounds__(256) void noalloc_table_kernel(const TableArgs a, double *__restrict__ tab)
{
    const int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i >= a.nsteps * a.nstage) return;
    const int64_t n = a.step0 + i / a.nstage;
    const int s = static_cast<int>(i % a.nstage);
    const double tn = a.atmin + a.dt * static_cast<double>(n);
    const double tn1 = a.atmin + a.dt * static_cast<double>(n + 1);
    const double t = a.c[s] == 1.0 ? tn1 : tn + a.c[s] * a.dt;
    double A[3] = {0.0, 0.0, 0.0};
    for (int p = 0; p < a.npulses; ++p) {
        double pa[3], pe[3];
        pulse_field(a.pulses[p], t, a.consistent != 0, pa, pe);
        for (int j = 0; j < 3; ++j) A[j] += pa[j];
    }
    for (int j = 0; j < 3; ++j) A[j] -= a.A_at_min[j];
    double *row = tab + i * 12;
    for (int j = 0; j < 3; ++j) {
        const double pa = A[0] * a.va[j][0] + A[1] * a.va[j][1];
        const double pb = A[0] * a.vb[j][0] + A[1] * a.vb[j][1];
        double sa, ca, sb, cb;
        sincos(pa, &sa, &ca);
        sincos(pb, &sb, &cb);
        row[2 * j] = a.t1 * ca;
        row[2 * j + 1] = a.t1 * sa;
        row[6 + 2 * j] = cb;
        row[7 + 2 * j] = sb;
    }
}

int fail(int code, const std::string &msg) { return set_thread_error(code, msg); }

int hip_fail(hipError_t e, const char *what)
{
    return fail(SBE_ERR_HIP, std::string(what) + ": " + hipGetErrorString(e));
}

}  // namespace
}  // namespace sbe

extern "C" int sbe_run_haldane_noalloc(const sbe_config *cfg, const sbe_model *model, const double *kpts,
                                       double *state, int64_t nk, int64_t nk_stride, int64_t step_begin,
                                       int64_t step_end, int32_t init_state, double *series, void *stream)
{
    using namespace sbe;
    std::string err;
    int rc = validate_inputs(cfg, model, err);
    if (rc) return fail(rc, err);
    if (model->kind != SBE_MODEL_HALDANE || cfg->gauge != SBE_GAUGE_LENGTH_WANNIER ||
        cfg->precision != SBE_PREC_FP64 || (cfg->rk_order != SBE_RK4 && cfg->rk_order != SBE_RK5_FEHLBERG))
        return fail(SBE_ERR_UNSUPPORTED, "sbe_run_haldane_noalloc: Haldane model, LengthWannier gauge, FP64 state, "
                                         "fixed-step RK4 or RK5 only");
    if (cfg->npulses > kMaxDevicePulses)
        return fail(SBE_ERR_UNSUPPORTED, "sbe_run_haldane_noalloc: at most " + std::to_string(kMaxDevicePulses) +
                                             " pulses");
    if (!kpts || !state || !series) return fail(SBE_ERR_CONFIG, "sbe_run_haldane_noalloc: null device buffer");
    if (nk <= 0 || nk_stride < nk || nk_stride % SBE_NOALLOC_K_ALIGN != 0)
        return fail(SBE_ERR_RANGE, "sbe_run_haldane_noalloc: need 0 < nk <= nk_stride, nk_stride a multiple of " +
                                       std::to_string(SBE_NOALLOC_K_ALIGN));
    Laser laser;
    Tableau tab;
    if ((rc = laser.init(*cfg, err)) || (rc = tab.init(cfg->rk_order, err))) return fail(rc, err);
    if (step_begin < 0 || step_end < step_begin || step_end > laser.nt)
        return fail(SBE_ERR_RANGE, "sbe_run_haldane_noalloc: steps outside [0, Nt]");

    // the plan of sbe_run for this many k points: k points per lane, waves
    const int kpt = fixed_step_two_k_per_lane(*cfg, *model, nk) ? 2 : 1;
    const int32_t nwaves = static_cast<int32_t>(nk_stride / (64 * kpt));
    const size_t row_tab = static_cast<size_t>(tab.nstage) * 12 * sizeof(double);
    const size_t row_part = static_cast<size_t>(nwaves) * kHaldaneNV * sizeof(double);
    const int64_t chunk = static_cast<int64_t>(kScratchBytes / (row_tab + row_part));
    if (chunk < 1)
        return fail(SBE_ERR_RANGE, "sbe_run_haldane_noalloc: " + std::to_string(nk_stride) +
                                       " k points need more than the " + std::to_string(kScratchBytes >> 20) +
                                       " MiB scratch holds for one step");

    int dev = 0;
    hipError_t e = hipGetDevice(&dev);
    if (e != hipSuccess) return hip_fail(e, "hipGetDevice");
    void *base = nullptr;
    if ((rc = scratch_address(dev, &base, err))) return fail(rc, err);
    const hipStream_t s = static_cast<hipStream_t>(stream);

    HaldaneArgs h{};
    const double a0 = model->a0;
    const double va[3][2] = {{0., a0}, {-std::sqrt(3.) / 2 * a0, -0.5 * a0}, {std::sqrt(3.) / 2 * a0, -0.5 * a0}};
    const double vb[3][2] = {{std::sqrt(3.) * a0, 0.}, {-std::sqrt(3.) / 2 * a0, 1.5 * a0},
                             {-std::sqrt(3.) / 2 * a0, -1.5 * a0}};
    std::memcpy(h.va, va, sizeof va);
    std::memcpy(h.vb, vb, sizeof vb);
    h.t1 = model->t1;
    h.c0 = 2.0 * model->t2 * std::cos(model->phi0);
    h.c3 = 2.0 * model->t2 * std::sin(model->phi0);
    h.M0 = model->M0;
    h.gam1 = cfg->T1 <= 1e-16 ? 0.0 : 1.0 / cfg->T1;
    h.gam2 = cfg->T2 <= 1e-16 ? 0.0 : 1.0 / cfg->T2;
    h.dt = laser.dt;
    h.nk = nk;
    h.nk_stride = nk_stride;
    h.nwaves = nwaves;
    h.inter_intra = cfg->inter_intra ? 1 : 0;
    h.precision = SBE_PREC_FP64;
    h.nval = model->nval;
    h.fermi = cfg->init_type == SBE_INIT_FERMI_DIRAC ? 1 : 0;
    h.fermi_energy = cfg->fermi_energy;
    h.kpt = kpt;
    h.kpts = kpts;
    h.weight = kpts + 3 * nk_stride;
    h.state = state;
    if (init_state) {
        e = launch_haldane_init(h, cfg->init_type, cfg->fermi_energy, s);
        if (e != hipSuccess) return hip_fail(e, "init kernel");
    }

    TableArgs ta{};
    for (size_t p = 0; p < laser.pulses.size(); ++p) ta.pulses[p] = laser.pulses[p];
    ta.npulses = static_cast<int32_t>(laser.pulses.size());
    ta.consistent = laser.consistent ? 1 : 0;
    ta.nstage = tab.nstage;
    for (int j = 0; j < 3; ++j) ta.A_at_min[j] = laser.A_at_min[j];
    ta.atmin = laser.atmin;
    ta.dt = laser.dt;
    ta.t1 = model->t1;
    for (int i = 0; i < 7; ++i) ta.c[i] = tab.c[i];
    std::memcpy(ta.va, va, sizeof va);
    std::memcpy(ta.vb, vb, sizeof vb);

    // scratch: [chunk][nstage][12] table rows, then [chunk][nwaves][8] partial rows
    double *d_tab = static_cast<double *>(base);
    double *d_part = reinterpret_cast<double *>(static_cast<unsigned char *>(base) + chunk * row_tab);
    const bool t1_on = h.gam1 != 0.0;
    for (int64_t s0 = step_begin; s0 < step_end; s0 += chunk) {
        const int64_t n = std::min<int64_t>(chunk, step_end - s0);
        ta.step0 = s0;
        ta.nsteps = n;
        const int64_t rows = n * tab.nstage;
        noalloc_table_kernel<<<static_cast<unsigned>((rows + 255) / 256), 256, 0, s>>>(ta, d_tab);
        if ((e = hipGetLastError()) != hipSuccess) return hip_fail(e, "table kernel");
        HaldaneArgs a = h;
        a.phase_tab = d_tab;
        a.step0 = 0;  // the table starts at this launch's first step
        a.nsteps = static_cast<int32_t>(n);
        a.partials = d_part;
        if ((e = launch_haldane_steps(a, cfg->rk_order, t1_on, s)) != hipSuccess) return hip_fail(e, "step kernel");
        e = launch_series_reduce(d_part, n, nwaves, kHaldaneNV, series + (s0 - step_begin) * kHaldaneNV, s);
        if (e != hipSuccess) return hip_fail(e, "series kernel");
    }
    return SBE_OK;
}
