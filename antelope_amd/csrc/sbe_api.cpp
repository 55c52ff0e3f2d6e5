// sbe_api.cpp - the C ABI of include/antelope/sbe_c.h.
//
// Host-side orchestration only: validation, the k-independent tables, device
// memory and the launch sequence. All per-k arithmetic runs in the HIP
// kernels (haldane.hip, series.hip); there is no CPU fallback. Every error
// is reported through the return code and sbe_last_error, never by exiting.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <climits>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <new>
#include <string>
#include <vector>

#include "analytic.hpp"
#include "band_frame.hpp"
#include "sbe_internal.hpp"

using sbe::kBlock;

namespace {

thread_local std::string g_thread_error;

int set_error(sbe_ctx *ctx, int code, const std::string &msg)
{
    if (ctx) ctx->err = msg;
    g_thread_error = msg;
    return code;
}

int hip_check(sbe_ctx *ctx, hipError_t e, const char *what)
{
    if (e == hipSuccess) return SBE_OK;
    const int code = (e == hipErrorOutOfMemory || e == hipErrorMemoryAllocation) ? SBE_ERR_OOM : SBE_ERR_HIP;
    return set_error(ctx, code, std::string(what) + ": " + hipGetErrorString(e));
}

// Upper bound of the per-chunk partial-sum buffer (per context; the members
// of a batch use a smaller one).
constexpr size_t kPartialBytes = size_t(128) << 20;
constexpr size_t kBatchPartialBytes = size_t(24) << 20;
// Large grids: every fixed-step launch reads the state and the k points once
// and recomputes e^{ik.v} (six sincos per k point), so a launch keeps at least
// kMinChunkSteps steps while the buffer stays under kPartialBytesMax (a 4096^2
// grid in FP32: 64 steps, 1.07 GB, against 8 steps at the 128 MiB bound).
constexpr int64_t kMinChunkSteps = 64;
constexpr size_t kPartialBytesMax = size_t(2) << 30;
// The adaptive stepper ends every launch exactly on a sample time, so launch
// boundaries shape its step sequence. Its windows are therefore a fixed
// property of the method, [0, W), [W, 2W), ..., for single solves, batches and
// served requests alike; a run split at multiples of W gives the same bits as
// one run.
constexpr int64_t kAdaptiveWindow = 2048;
thread_local size_t t_partial_bytes = kPartialBytes;
// set while a batch creates its members: the batch step kernel runs one k point
// per lane for every member, so a member's rows and block count must be planned so
thread_local bool t_batch_member = false;

struct Plan {
    int kpt = 1;
    int64_t nk = 0, nk_stride = 0;
    int32_t nwaves = 0;
    int nv = 0;
    int64_t chunk = 0;
    int64_t nt = 0;
    int nstage = 0;
    size_t bytes_state = 0, bytes_kpts = 0, bytes_tab = 0, bytes_series = 0, bytes_partials = 0, bytes_adaptive = 0;
    size_t total() const
    {
        return bytes_state + bytes_kpts + bytes_tab + bytes_series + bytes_partials + bytes_adaptive;
    }
};

int validate(const sbe_config *cfg, const sbe_model *m, std::string &err)
{
    if (!cfg || !m) {
        err = "null config or model";
        return SBE_ERR_CONFIG;
    }
    if (cfg->abi_version != SBE_ABI_VERSION) {
        err = "sbe_config.abi_version mismatch (expected " + std::to_string(SBE_ABI_VERSION) + ")";
        return SBE_ERR_CONFIG;
    }
    if (cfg->gauge == SBE_GAUGE_LENGTH_HAMILTONIAN) {
        err = "LengthHamiltonian gauge is not implemented (as in the reference, SBEs.h:806-810)";
        return SBE_ERR_UNSUPPORTED;
    }
    if (cfg->gauge != SBE_GAUGE_LENGTH_WANNIER && cfg->gauge != SBE_GAUGE_VELOCITY_HAMILTONIAN) {
        err = "unknown gauge";
        return SBE_ERR_CONFIG;
    }
    if (cfg->gauge == SBE_GAUGE_VELOCITY_HAMILTONIAN && cfg->rk_order == SBE_RK_CASH_KARP) {
        err = "the adaptive Cash-Karp stepper is built for the LengthWannier gauge";
        return SBE_ERR_UNSUPPORTED;
    }
    if (cfg->precision != SBE_PREC_FP64 && cfg->precision != SBE_PREC_FP32 && cfg->precision != SBE_PREC_SPLIT16) {
        err = "unknown precision";
        return SBE_ERR_CONFIG;
    }
    if (cfg->precision == SBE_PREC_SPLIT16 && cfg->gauge == SBE_GAUGE_VELOCITY_HAMILTONIAN) {
        err = "SPLIT16 precision applies to the per-stage Fourier sums, which the velocity gauge does not have";
        return SBE_ERR_UNSUPPORTED;
    }
    if (cfg->precision == SBE_PREC_SPLIT16 && m->kind != SBE_MODEL_WANNIER) {
        err = "SPLIT16 precision applies to the Wannier Fourier sums";
        return SBE_ERR_UNSUPPORTED;
    }
    if (cfg->precision == SBE_PREC_FP32 && m->kind != SBE_MODEL_HALDANE) {
        err = "the N-band path runs FP64 (or SPLIT16 Fourier sums for Wannier models)";
        return SBE_ERR_UNSUPPORTED;
    }
    if (m->kind == SBE_MODEL_HALDANE) {
        if (m->nband != 2) {
            err = "Haldane model needs nband = 2";
            return SBE_ERR_MODEL;
        }
        if (!(m->a0 > 0.0) || !std::isfinite(m->t1) || !std::isfinite(m->t2) || !std::isfinite(m->phi0) ||
            !std::isfinite(m->M0)) {
            err = "Haldane parameters must be finite with a0 > 0";
            return SBE_ERR_MODEL;
        }
        if (m->nval < 0 || m->nval > 2) {
            err = "Haldane nval must be 0..2";
            return SBE_ERR_MODEL;
        }
        return SBE_OK;
    }
    if (m->kind == SBE_MODEL_WANNIER || sbe::am::is_analytic(m->kind)) return sbe::nb_validate(*m, err);
    err = "unknown model kind";
    return SBE_ERR_MODEL;
}

int make_plan(const sbe_config &cfg, const sbe_model &m, int64_t nk, int64_t nt, int nstage, Plan &p)
{
    p.nk = nk;
    // k points per lane of the step kernel: 2 gives each lane two independent
    // dependency chains. Measured: FP64 gains 6-10 % with two once there are a
    // thousand waves or more (300^2: 36.9 against 41.0 ms; 1024^2: 37.6 against
    // 39.9 ms for 3000 steps) and loses below that (200^2: 5.1 against 4.0 ms);
    // FP32 does best with one throughout
    p.kpt = cfg.precision == SBE_PREC_FP64 && nk >= 65536 ? 2 : 1;
    sbe_tuning tu{};
    const bool tuned = !t_batch_member && sbe::tuning_lookup(cfg, m, nk, tu);
    if (tuned && tu.kpt > 0 && cfg.rk_order != SBE_RK_CASH_KARP) p.kpt = tu.kpt;
    if (const char *e = std::getenv("ANTELOPE_HALDANE_KPT")) p.kpt = std::atoi(e) == 2 ? 2 : 1;
    if (t_batch_member) p.kpt = 1;
    if (cfg.rk_order == SBE_RK_CASH_KARP) {  // adaptive: 64 kpt consecutive k points per wave (step-control block)
        // one k point per lane: the FP64 kernel then fits two waves per SIMD (1.1x over
        // two chains per lane at one wave on the batched workload) and the 64-point
        // step-control blocks take longer steps than 128-point ones
        p.kpt = tuned && tu.kpt > 0 ? tu.kpt : 1;
        if (const char *e = std::getenv("ANTELOPE_ADAPTIVE_KPT")) p.kpt = std::atoi(e) == 1 ? 1 : 2;
    }
    const int64_t unit = static_cast<int64_t>(kBlock) * p.kpt;
    p.nk_stride = (nk + unit - 1) / unit * unit;
    p.nwaves = static_cast<int32_t>(p.nk_stride / (64 * p.kpt));
    p.nv = sbe::kHaldaneNV;
    p.nt = nt;
    p.nstage = nstage;
    const size_t row = static_cast<size_t>(p.nwaves) * p.nv * sizeof(double);
    int64_t by_bytes = static_cast<int64_t>(t_partial_bytes / row);
    if (by_bytes < kMinChunkSteps) by_bytes = std::min<int64_t>(kMinChunkSteps, static_cast<int64_t>(kPartialBytesMax / row));
    p.chunk = std::max<int64_t>(1, std::min<int64_t>(nt, by_bytes));
    if (tuned && tu.chunk_steps > 0) p.chunk = std::min<int64_t>(p.chunk, tu.chunk_steps);
    if (cfg.rk_order == SBE_RK_CASH_KARP) p.chunk = std::min<int64_t>(nt, kAdaptiveWindow);
    const size_t real = cfg.precision == SBE_PREC_FP32 ? sizeof(float) : sizeof(double);
    const int comps = cfg.precision == SBE_PREC_FP32 ? 8 : 4;
    p.bytes_state = static_cast<size_t>(p.nk_stride) * comps * real;
    p.bytes_kpts = static_cast<size_t>(p.nk_stride) * 4 * sizeof(double);
    // the adaptive stepper reads table rows at the sample times only
    const bool adaptive = cfg.rk_order == SBE_RK_CASH_KARP;
    p.bytes_tab = static_cast<size_t>(nt) * (adaptive ? 1 : nstage) * 12 * (sizeof(double) + sizeof(float));
    if (adaptive) p.bytes_tab += static_cast<size_t>(std::max<int64_t>(nt, 2)) * 4 * sizeof(double);  // A, dA/dt at t_n
    p.bytes_adaptive = adaptive ? static_cast<size_t>(p.nwaves) * (sizeof(double) + 2 * sizeof(int64_t)) + 256 +
                                      sizeof(sbe::PulseConst) * sbe::kMaxDevicePulses
                                : 0;
    p.bytes_series = static_cast<size_t>(nt) * p.nv * sizeof(double);
    p.bytes_partials = static_cast<size_t>(p.chunk) * row;
    return SBE_OK;
}

template <typename T>
T *dalloc(sbe_ctx *ctx, size_t bytes, int &rc)
{
    void *p = nullptr;
    if (rc != SBE_OK) return nullptr;
    hipError_t e = hipMalloc(&p, bytes ? bytes : 1);
    if (e != hipSuccess) {
        rc = hip_check(ctx, e, "hipMalloc");
        return nullptr;
    }
    ctx->owned.push_back({p, bytes});
    ctx->device_bytes += static_cast<int64_t>(bytes);
    return static_cast<T *>(p);
}

// RK amplification of dn/dt = -g1 n over one step (population decay with T1
// on is the same for every k: n_k(t_m) = n_k(0) R^m).
double rk_linear_factor(const sbe::Tableau &tab, double z)
{
    double kk[7];
    for (int i = 0; i < tab.nstage; ++i) {
        double y = 1.0;
        for (int j = 0; j < i; ++j) y += tab.a[i][j] * kk[j];
        kk[i] = z * y;
    }
    double r = 1.0;
    for (int i = 0; i < tab.nstage; ++i) r += tab.b[i] * kk[i];
    return r;
}

void drop_events(sbe_ctx *ctx)
{
    for (auto &v : ctx->ev) {
        for (auto &p : v) {
            (void)hipEventDestroy(p.first);
            (void)hipEventDestroy(p.second);
        }
        v.clear();
    }
}

// Launch through `fn`, bracketed by events when timing is on.
template <typename F>
int timed(sbe_ctx *ctx, int cls, F &&fn, const char *what)
{
    hipEvent_t a = nullptr, b = nullptr;
    if (ctx->timing) {
        int rc = hip_check(ctx, hipEventCreate(&a), "hipEventCreate");
        if (!rc) rc = hip_check(ctx, hipEventCreate(&b), "hipEventCreate");
        if (!rc) rc = hip_check(ctx, hipEventRecord(a, ctx->stream), "hipEventRecord");
        if (rc) return rc;
    }
    int rc = hip_check(ctx, fn(), what);
    if (ctx->timing) {
        if (!rc) rc = hip_check(ctx, hipEventRecord(b, ctx->stream), "hipEventRecord");
        ctx->ev[cls].push_back({a, b});
    }
    return rc;
}

// Back to the initial state of a 2-band context, in stream order on s: the
// state of sbe_create (SBEs.h:284-364), an empty series, step 0, and for the
// adaptive stepper the cursor, the first step proposals (dt), the counters
// and the failure flag. Used by sbe_create and sbe_batch_reset.
int reset_2band(sbe_ctx *ctx, hipStream_t s)
{
    ctx->steps_done = 0;
    *ctx->nonfinite_h = 0;  // no reduction of this context is in flight here
    const size_t series_bytes = static_cast<size_t>(ctx->laser.nt) * ctx->nv * sizeof(double);
    int rc = hip_check(ctx, hipMemsetAsync(ctx->d_series, 0, series_bytes, s), "clear series");
    if (ctx->adaptive && !rc) {
        const std::vector<double> h0(static_cast<size_t>(ctx->nwaves), ctx->laser.dt);  // first proposal: dt
        rc = hip_check(ctx, hipMemcpyAsync(ctx->d_hstep, h0.data(), h0.size() * sizeof(double), hipMemcpyHostToDevice, s),
                       "upload step sizes");
        if (!rc) rc = hip_check(ctx, hipMemsetAsync(ctx->d_counts, 0, 2 * sizeof(int64_t) * ctx->nwaves, s), "clear counts");
        if (!rc) rc = hip_check(ctx, hipMemsetAsync(ctx->d_fail, 0, sizeof(int32_t), s), "clear flag");
        if (!rc) rc = hip_check(ctx, hipMemsetAsync(ctx->d_cursor, 0, sizeof(int64_t), s), "clear cursor");
        // the host vector must outlive the copy
        if (!rc) rc = hip_check(ctx, hipStreamSynchronize(s), "reset sync");
    }
    if (rc) return rc;
    if (ctx->velocity) return hip_check(ctx, sbe::launch_haldane_vel_init(ctx->hargs, s), "init kernel");
    return hip_check(ctx, sbe::launch_haldane_init(ctx->hargs, ctx->cfg.init_type, ctx->cfg.fermi_energy, s),
                     "init kernel");
}

}  // namespace

namespace sbe {
int create_batch_member(const sbe_config *cfg, const sbe_model *model, sbe_ctx **out)
{
    t_partial_bytes = kBatchPartialBytes;
    t_batch_member = true;
    const int rc = sbe_create(cfg, model, -1, -1, out);
    t_partial_bytes = kPartialBytes;
    t_batch_member = false;
    return rc;
}

bool shared_eligible(const sbe_ctx *c)
{
    return c->path == PathKind::Haldane2 && !c->adaptive && !c->velocity && c->hargs.kpt == 1;
}

int shared_key(const sbe_ctx *c) { return c->cfg.rk_order * 16 + c->cfg.precision; }

void free_shared(SharedLaunch &buf)
{
    if (buf.d_insts) (void)hipFree(buf.d_insts);
    if (buf.d_map) (void)hipFree(buf.d_map);
    if (buf.d_jobs) (void)hipFree(buf.d_jobs);
    buf = SharedLaunch{};
}

int launch_shared(sbe_ctx *const *cs, const int64_t *nsteps, int n, SharedLaunch &buf, hipStream_t s, std::string &err)
{
    if (n <= 0) return SBE_OK;
    int blocks = 0;
    bool t1 = false, deph = false, multi = true;
    for (int i = 0; i < n; ++i) {
        const sbe_ctx *c = cs[i];
        if (!shared_eligible(c) || shared_key(c) != shared_key(cs[0]) || nsteps[i] <= 0 ||
            nsteps[i] > c->chunk_steps || c->steps_done + nsteps[i] > c->laser.nt) {
            err = "shared launch: member " + std::to_string(i) + " cannot take " + std::to_string(nsteps[i]) + " steps";
            return SBE_ERR_STATE;
        }
        blocks += static_cast<int>(c->nk_stride / kBlock);
        t1 = t1 || c->hargs.gam1 != 0.0;
        deph = deph || c->hargs.gam1 != 0.0 || c->hargs.gam2 != 0.0;
        multi = multi && series_multi_ok(c->nwaves);
    }
    hipError_t e = hipSuccess;
    if (n > buf.cap_insts || blocks + n > buf.cap_map) {  // grow: in-flight launches read the old buffers
        e = hipStreamSynchronize(s);
        const int ci = std::max(n, 2 * buf.cap_insts), cm = std::max(blocks + n, 2 * buf.cap_map);
        free_shared(buf);
        if (e == hipSuccess) e = hipMalloc(&buf.d_insts, sizeof(HaldaneArgs) * ci);
        if (e == hipSuccess) e = hipMalloc(&buf.d_map, sizeof(int32_t) * cm);
        if (e == hipSuccess) e = hipMalloc(&buf.d_jobs, sizeof(SeriesJob) * ci);
        if (e != hipSuccess) {
            free_shared(buf);
            err = std::string("shared launch buffers: ") + hipGetErrorString(e);
            return SBE_ERR_OOM;
        }
        buf.cap_insts = ci;
        buf.cap_map = cm;
    }
    std::vector<HaldaneArgs> insts(n);
    std::vector<int32_t> map(static_cast<size_t>(blocks) + n);
    std::vector<SeriesJob> jobs;
    int nb = 0;
    int64_t total = 0;
    for (int i = 0; i < n; ++i) {
        const sbe_ctx *c = cs[i];
        insts[i] = c->hargs;
        insts[i].step0 = c->steps_done;
        insts[i].nsteps = static_cast<int32_t>(nsteps[i]);
        map[blocks + i] = nb;
        for (int64_t q = 0; q < c->nk_stride / kBlock; ++q) map[nb++] = i;
        jobs.push_back({c->d_partials, c->d_series + c->steps_done * c->nv, c->nwaves, insts[i].nsteps, total,
                        c->nonfinite_d});
        total += nsteps[i];
    }
    e = hipMemcpyAsync(buf.d_insts, insts.data(), sizeof(HaldaneArgs) * n, hipMemcpyHostToDevice, s);
    if (e == hipSuccess) e = hipMemcpyAsync(buf.d_map, map.data(), sizeof(int32_t) * map.size(), hipMemcpyHostToDevice, s);
    if (e == hipSuccess)
        e = launch_haldane_batch(buf.d_insts, buf.d_map, buf.d_map + blocks, blocks, cs[0]->cfg.rk_order,
                                 cs[0]->cfg.precision, t1, deph, s);
    if (e == hipSuccess && multi) {
        e = hipMemcpyAsync(buf.d_jobs, jobs.data(), sizeof(SeriesJob) * n, hipMemcpyHostToDevice, s);
        if (e == hipSuccess) e = launch_series_reduce_multi(buf.d_jobs, n, total, kHaldaneNV, s);
    }
    for (int i = 0; i < n && e == hipSuccess; ++i) {
        sbe_ctx *c = cs[i];
        if (!multi)
            e = launch_series_reduce(c->d_partials, nsteps[i], c->nwaves, c->nv, c->d_series + c->steps_done * c->nv, s,
                                     c->nonfinite_d);
        c->steps_done += nsteps[i];
    }
    if (e != hipSuccess) {
        err = std::string("shared launch: ") + hipGetErrorString(e);
        return SBE_ERR_HIP;
    }
    return SBE_OK;
}

int validate_inputs(const sbe_config *cfg, const sbe_model *m, std::string &err) { return validate(cfg, m, err); }
int set_thread_error(int code, const std::string &msg) { return set_error(nullptr, code, msg); }
bool fixed_step_two_k_per_lane(const sbe_config &cfg, const sbe_model &m, int64_t nk)
{
    Plan p;
    make_plan(cfg, m, nk, 1, 4, p);
    return p.kpt == 2;
}
}  // namespace sbe

extern "C" {

int sbe_set_timing(sbe_ctx *ctx, int32_t enable)
{
    if (!ctx) return set_error(nullptr, SBE_ERR_STATE, "null context");
    ctx->timing = enable != 0;
    return SBE_OK;
}

int sbe_get_timing(sbe_ctx *ctx, double *ms, int64_t *launches)
{
    if (!ctx) return set_error(nullptr, SBE_ERR_STATE, "null context");
    int rc = hip_check(ctx, hipStreamSynchronize(ctx->stream), "hipStreamSynchronize");
    for (int c = 0; c < 2 && !rc; ++c) {
        double sum = 0.0;
        for (auto &p : ctx->ev[c]) {
            float t = 0.f;
            rc = hip_check(ctx, hipEventElapsedTime(&t, p.first, p.second), "hipEventElapsedTime");
            if (rc) break;
            sum += t;
        }
        if (ms) ms[c] = sum;
        if (launches) launches[c] = static_cast<int64_t>(ctx->ev[c].size());
    }
    drop_events(ctx);
    return rc;
}

const char *sbe_last_error(const sbe_ctx *ctx)
{
    if (ctx) return ctx->err.c_str();
    return g_thread_error.c_str();
}

int sbe_laser_eval(const sbe_config *cfg, int64_t n, const double *t, double *A, double *E, double *atmin,
                   int64_t *nt)
{
    if (!cfg) return set_error(nullptr, SBE_ERR_CONFIG, "null config");
    sbe::Laser L;
    std::string err;
    int rc = L.init(*cfg, err);
    if (rc) return set_error(nullptr, rc, err);
    for (int64_t i = 0; i < n; ++i) {
        double a[3], e[3];
        L.eval(t[i], a, e);
        for (int j = 0; j < 3; ++j) {
            if (A) A[3 * i + j] = a[j];
            if (E) E[3 * i + j] = e[j];
        }
    }
    if (atmin) *atmin = L.atmin;
    if (nt) *nt = L.nt;
    return SBE_OK;
}

int sbe_query_footprint(const sbe_config *cfg, const sbe_model *model, int64_t k_begin, int64_t k_end,
                        size_t *bytes)
{
    std::string err;
    int rc = validate(cfg, model, err);
    if (rc) return set_error(nullptr, rc, err);
    sbe::Laser L;
    rc = L.init(*cfg, err);
    if (rc) return set_error(nullptr, rc, err);
    sbe::Tableau tab;
    rc = tab.init(cfg->rk_order, err);
    if (rc) return set_error(nullptr, rc, err);
    int64_t ntot = 1;
    for (int d = 0; d < 3; ++d) ntot *= cfg->npoints[d] > 0 ? cfg->npoints[d] : 1;
    if (k_begin == -1 && k_end == -1) {
        k_begin = 0;
        k_end = ntot;
    }
    if (k_begin < 0 || k_end > ntot || k_begin >= k_end) return set_error(nullptr, SBE_ERR_RANGE, "bad k range");
    if (model->kind != SBE_MODEL_HALDANE) {
        if (bytes) *bytes = sbe::nb_footprint(*cfg, *model, k_end - k_begin, L.nt, tab.nstage);
        return SBE_OK;
    }
    Plan p;
    make_plan(*cfg, *model, k_end - k_begin, L.nt, tab.nstage, p);
    if (bytes) *bytes = p.total();
    return SBE_OK;
}

int sbe_create(const sbe_config *cfg, const sbe_model *model, int64_t k_begin, int64_t k_end, sbe_ctx **out)
{
    if (!out) return set_error(nullptr, SBE_ERR_CONFIG, "null output pointer");
    *out = nullptr;
    std::string err;
    int rc = validate(cfg, model, err);
    if (rc) return set_error(nullptr, rc, err);

    sbe_ctx *ctx = new (std::nothrow) sbe_ctx();
    if (!ctx) return set_error(nullptr, SBE_ERR_OOM, "host allocation failed");
    ctx->cfg = *cfg;
    ctx->pulses.assign(cfg->pulses, cfg->pulses + cfg->npulses);
    ctx->cfg.pulses = ctx->pulses.data();
    ctx->model = *model;
    ctx->N = model->nband;
    ctx->velocity = cfg->gauge == SBE_GAUGE_VELOCITY_HAMILTONIAN;

    if ((rc = ctx->laser.init(ctx->cfg, err)) || (rc = ctx->tab.init(ctx->cfg.rk_order, err)) ||
        (rc = ctx->grid.build(ctx->cfg, ctx->model, k_begin, k_end, err))) {
        set_error(nullptr, rc, err);
        delete ctx;
        return rc;
    }
    if (ctx->tab.embedded) {
        std::string why;
        rc = SBE_ERR_UNSUPPORTED;
        if (model->kind != SBE_MODEL_HALDANE) why = "the adaptive Cash-Karp stepper is built for the 2-band (Haldane) path";
        else if (ctx->cfg.T1 > 1e-16) why = "T1 population decay is not built for the adaptive Cash-Karp stepper";
        else if (ctx->cfg.npulses > sbe::kMaxDevicePulses)
            why = "the adaptive Cash-Karp stepper takes at most " + std::to_string(sbe::kMaxDevicePulses) + " pulses";
        else if (ctx->laser.nt >= (int64_t(1) << 31))
            why = "the adaptive Cash-Karp stepper takes fewer than 2^31 samples", rc = SBE_ERR_RANGE;
        else if (!(ctx->cfg.rtol > 0.0) || !(ctx->cfg.atol >= 0.0) || !std::isfinite(ctx->cfg.rtol) ||
                 !std::isfinite(ctx->cfg.atol))
            why = "calc.rtol must be > 0 and calc.atol >= 0", rc = SBE_ERR_CONFIG;
        if (!why.empty()) {
            set_error(nullptr, rc, why);
            delete ctx;
            return rc;
        }
        ctx->adaptive = true;
    }

    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev < 1) {
        set_error(nullptr, SBE_ERR_NODEVICE, "no HIP device visible");
        delete ctx;
        return SBE_ERR_NODEVICE;
    }
    (void)hipGetDevice(&ctx->device);
    if (hipHostMalloc(reinterpret_cast<void **>(&ctx->nonfinite_h), sizeof(int32_t), hipHostMallocMapped) != hipSuccess ||
        hipHostGetDevicePointer(reinterpret_cast<void **>(&ctx->nonfinite_d), ctx->nonfinite_h, 0) != hipSuccess) {
        set_error(nullptr, SBE_ERR_OOM, "pinned host flag");
        sbe_destroy(ctx);
        return SBE_ERR_OOM;
    }
    *ctx->nonfinite_h = 0;

    if (model->kind != SBE_MODEL_HALDANE) {
        ctx->path = sbe::PathKind::NBand;
        rc = hip_check(ctx, hipStreamCreateWithFlags(&ctx->stream, hipStreamNonBlocking), "hipStreamCreate");
        if (!rc) rc = sbe::nb_setup(ctx);
        if (rc) {
            set_error(nullptr, rc, ctx->err);
            sbe_destroy(ctx);
            return rc;
        }
        *out = ctx;
        return SBE_OK;
    }
    Plan p;
    const int64_t nk = ctx->grid.k_end - ctx->grid.k_begin;
    make_plan(ctx->cfg, ctx->model, nk, ctx->laser.nt, ctx->tab.nstage, p);
    ctx->nk = p.nk;
    ctx->nk_stride = p.nk_stride;
    ctx->nwaves = p.nwaves;
    ctx->nv = p.nv;
    ctx->chunk_steps = p.chunk;

    rc = hip_check(ctx, hipStreamCreateWithFlags(&ctx->stream, hipStreamNonBlocking), "hipStreamCreate");
    ctx->d_state = dalloc<void>(ctx, p.bytes_state, rc);
    ctx->d_kpts = dalloc<double>(ctx, p.bytes_kpts, rc);
    ctx->d_phase_tab = dalloc<double>(ctx, p.bytes_tab, rc);
    ctx->d_series = dalloc<double>(ctx, p.bytes_series, rc);
    ctx->d_partials = dalloc<double>(ctx, p.bytes_partials, rc);
    if (ctx->adaptive) {
        ctx->d_hstep = dalloc<double>(ctx, sizeof(double) * ctx->nwaves, rc);
        ctx->d_counts = dalloc<int64_t>(ctx, 2 * sizeof(int64_t) * ctx->nwaves, rc);
        ctx->d_fail = dalloc<int32_t>(ctx, sizeof(int32_t), rc);
        ctx->d_pulses = dalloc<sbe::PulseConst>(ctx, sizeof(sbe::PulseConst) * sbe::kMaxDevicePulses, rc);
        ctx->d_aargs = dalloc<sbe::AdaptiveArgs>(ctx, sizeof(sbe::AdaptiveArgs), rc);
        ctx->d_cursor = dalloc<int64_t>(ctx, sizeof(int64_t), rc);
    }
    if (rc) {
        set_error(nullptr, rc, ctx->err);
        sbe_destroy(ctx);
        return rc;
    }
    ctx->d_weight = ctx->d_kpts + 3 * ctx->nk_stride;

    // k points (SoA, zero-padded) and weights
    {
        std::vector<double> hk(static_cast<size_t>(ctx->nk_stride) * 4, 0.0);
        for (int64_t k = 0; k < nk; ++k) {
            for (int c = 0; c < 3; ++c) hk[c * ctx->nk_stride + k] = ctx->grid.k[3 * k + c];
            hk[3 * ctx->nk_stride + k] = ctx->grid.w[k];
        }
        rc = hip_check(ctx, hipMemcpy(ctx->d_kpts, hk.data(), p.bytes_kpts, hipMemcpyHostToDevice), "upload k");
    }

    // Haldane geometry (Haldane.h:134-143) and per-stage laser phase tables
    sbe::HaldaneArgs &h = ctx->hargs;
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
    h.gam1 = ctx->cfg.T1 <= 1e-16 ? 0.0 : 1.0 / ctx->cfg.T1;  // SBEs.h:484-493
    h.gam2 = ctx->cfg.T2 <= 1e-16 ? 0.0 : 1.0 / ctx->cfg.T2;
    h.dt = ctx->laser.dt;
    h.nk = nk;
    h.nk_stride = ctx->nk_stride;
    h.nwaves = ctx->nwaves;
    h.inter_intra = ctx->cfg.inter_intra ? 1 : 0;
    h.precision = ctx->cfg.precision;
    h.nval = model->nval;
    h.fermi = ctx->cfg.init_type == SBE_INIT_FERMI_DIRAC ? 1 : 0;
    h.fermi_energy = ctx->cfg.fermi_energy;
    h.kpt = p.kpt;
    h.kpts = ctx->d_kpts;
    h.weight = ctx->d_weight;
    h.state = ctx->d_state;
    h.partials = ctx->d_partials;
    {
        const int S = ctx->adaptive ? 1 : ctx->tab.nstage;  // adaptive: rows at the sample times t_n
        const int64_t nt = ctx->laser.nt;
        const size_t n64 = static_cast<size_t>(nt) * S * 12;
        std::vector<double> tab(n64);
        std::vector<float> tab32(n64);
        for (int64_t n = 0; n < nt; ++n) {
            const double tn = ctx->laser.atmin + ctx->laser.dt * static_cast<double>(n);  // antelope.cpp:299
            const double tn1 = ctx->laser.atmin + ctx->laser.dt * static_cast<double>(n + 1);
            for (int s = 0; s < S; ++s) {
                double A[3], E[3];
                // SBEs.h:612; the stage at c = 1 is taken at t_{n+1} itself, so its row is the
                // next step's first row bit for bit (the step kernels carry it across steps)
                ctx->laser.eval(ctx->tab.c[s] == 1.0 ? tn1 : tn + ctx->tab.c[s] * ctx->laser.dt, A, E);
                if (ctx->velocity) {  // A at the stage times, [nt][S][2]
                    double *row = tab.data() + (static_cast<size_t>(n) * S + s) * 2;
                    row[0] = A[0];
                    row[1] = A[1];
                    continue;
                }
                double *row = tab.data() + (static_cast<size_t>(n) * S + s) * 12;
                for (int j = 0; j < 3; ++j) {
                    const double pa = A[0] * va[j][0] + A[1] * va[j][1];
                    const double pb = A[0] * vb[j][0] + A[1] * vb[j][1];
                    row[2 * j] = model->t1 * std::cos(pa);  // NN part carries t1
                    row[2 * j + 1] = model->t1 * std::sin(pa);
                    row[6 + 2 * j] = std::cos(pb);
                    row[6 + 2 * j + 1] = std::sin(pb);
                }
            }
        }
        for (size_t i = 0; i < n64; ++i) tab32[i] = static_cast<float>(tab[i]);
        float *d32 = reinterpret_cast<float *>(ctx->d_phase_tab + n64);
        if (!rc) rc = hip_check(ctx, hipMemcpy(ctx->d_phase_tab, tab.data(), n64 * sizeof(double), hipMemcpyHostToDevice), "upload table");
        if (!rc) rc = hip_check(ctx, hipMemcpy(d32, tab32.data(), n64 * sizeof(float), hipMemcpyHostToDevice), "upload table");
        h.phase_tab = ctx->d_phase_tab;
        h.phase_tab32 = d32;
        if (ctx->adaptive && !rc) {  // A - A(t_min) and the consistent dA/dt = -E at t_n, summed in pulse order
            const int64_t na = std::max<int64_t>(nt, 2);  // an interval to interpolate in
            std::vector<double> at(static_cast<size_t>(na) * 4);
            for (int64_t n = 0; n < na; ++n) {
                const double tn = ctx->laser.atmin + ctx->laser.dt * static_cast<double>(n);
                double sx = 0.0, sy = 0.0, gx = 0.0, gy = 0.0;
                for (const sbe::PulseConst &pc : ctx->laser.pulses) {
                    double pa[3], pe[3];
                    sbe::pulse_field(pc, tn, true, pa, pe);
                    sx += pa[0];
                    sy += pa[1];
                    gx -= pe[0];
                    gy -= pe[1];
                }
                at[4 * n] = sx - ctx->laser.A_at_min[0];
                at[4 * n + 1] = sy - ctx->laser.A_at_min[1];
                at[4 * n + 2] = gx;
                at[4 * n + 3] = gy;
            }
            double *d_at = reinterpret_cast<double *>(d32 + n64);
            rc = hip_check(ctx, hipMemcpy(d_at, at.data(), at.size() * sizeof(double), hipMemcpyHostToDevice), "upload table");
            ctx->aargs.a_tab = d_at;
        }
    }
    if (ctx->adaptive && !rc) {
        sbe::AdaptiveArgs &A = ctx->aargs;
        A.h = h;
        A.atmin = ctx->laser.atmin;
        A.dt_out = ctx->laser.dt;
        A.inv_dt_out = 1.0 / ctx->laser.dt;
        A.nt = ctx->laser.nt;
        A.cursor = ctx->d_cursor;
        A.rtol = ctx->cfg.rtol;
        A.atol = ctx->cfg.atol;
        A.hmax = 32.0 * ctx->laser.dt;
        A.hmin = 1e-6 * ctx->laser.dt;
        A.npulses = static_cast<int32_t>(ctx->laser.pulses.size());
        for (int j = 0; j < 3; ++j) A.A_at_min[j] = ctx->laser.A_at_min[j];
        A.pulses = ctx->d_pulses;
        A.hstep = ctx->d_hstep;
        A.counts = ctx->d_counts;
        A.fail = ctx->d_fail;
        rc = hip_check(ctx, hipMemcpy(ctx->d_pulses, ctx->laser.pulses.data(),
                                      sizeof(sbe::PulseConst) * ctx->laser.pulses.size(), hipMemcpyHostToDevice),
                       "upload pulses");
        if (!rc)
            rc = hip_check(ctx, hipMemcpy(ctx->d_aargs, &A, sizeof A, hipMemcpyHostToDevice), "upload adaptive arguments");
    }
    if (!rc) rc = reset_2band(ctx, ctx->stream);
    if (!rc) rc = hip_check(ctx, hipStreamSynchronize(ctx->stream), "init sync");
    if (rc) {
        set_error(nullptr, rc, ctx->err);
        sbe_destroy(ctx);
        return rc;
    }
    // occupied-band count per k is the same for every k (FD is normalised to 1)
    const double n0 = ctx->cfg.init_type == SBE_INIT_FERMI_DIRAC ? 1.0 : static_cast<double>(std::min(model->nval, 2));
    ctx->n_sum = n0 * ctx->grid.total_weight;
    *out = ctx;
    return SBE_OK;
}

int sbe_run(sbe_ctx *ctx, int64_t step_begin, int64_t step_end)
{
    if (!ctx) return set_error(nullptr, SBE_ERR_STATE, "null context");
    if (step_begin != ctx->steps_done)
        return set_error(ctx, SBE_ERR_STATE,
                         "steps run in order: next step is " + std::to_string(ctx->steps_done));
    if (step_end > ctx->laser.nt || step_end < step_begin)
        return set_error(ctx, SBE_ERR_RANGE, "step_end outside [step_begin, Nt]");
    if (ctx->path == sbe::PathKind::NBand) return sbe::nb_run(ctx, step_begin, step_end);
    if (ctx->adaptive) {
        const bool deph = ctx->hargs.gam2 != 0.0;
        const int nblocks = static_cast<int>(ctx->nk_stride / (sbe::kAdaptiveBlock * ctx->hargs.kpt));
        for (int64_t s = step_begin; s < step_end; s += ctx->chunk_steps) {
            const int64_t n = std::min<int64_t>(ctx->chunk_steps, step_end - s);
            // the kernel reads the window start from the device cursor (== s)
            int rc = timed(ctx, 0, [&] {
                return sbe::launch_haldane_adaptive(ctx->d_aargs, nullptr, nullptr, nblocks, n, ctx->cfg.precision,
                                                    deph, ctx->hargs.kpt, ctx->stream);
            }, "adaptive step kernel");
            if (!rc) rc = hip_check(ctx, sbe::launch_adaptive_advance(ctx->d_aargs, 1, n, ctx->stream), "advance kernel");
            if (!rc)
                rc = timed(ctx, 1, [&] {
                    return sbe::launch_series_reduce(ctx->d_partials, n, ctx->nwaves, ctx->nv,
                                                     ctx->d_series + s * ctx->nv, ctx->stream, ctx->nonfinite_d);
                }, "series kernel");
            if (rc) return rc;
            ctx->steps_done = s + n;
        }
        return SBE_OK;
    }
    const bool t1_on = ctx->hargs.gam1 != 0.0;
    for (int64_t s = step_begin; s < step_end; s += ctx->chunk_steps) {
        const int64_t n = std::min<int64_t>(ctx->chunk_steps, step_end - s);
        sbe::HaldaneArgs a = ctx->hargs;
        a.step0 = s;
        a.nsteps = static_cast<int32_t>(n);
        int rc = timed(ctx, 0, [&] {
            return ctx->velocity ? sbe::launch_haldane_vel_steps(a, ctx->cfg.rk_order, t1_on, ctx->stream)
                                 : sbe::launch_haldane_steps(a, ctx->cfg.rk_order, t1_on, ctx->stream);
        }, "step kernel");
        if (!rc)
            rc = timed(ctx, 1, [&] {
                return sbe::launch_series_reduce(ctx->d_partials, n, ctx->nwaves, ctx->nv, ctx->d_series + s * ctx->nv,
                                                 ctx->stream, ctx->nonfinite_d);
            }, "series kernel");
        if (rc) return rc;
        ctx->steps_done = s + n;
    }
    return SBE_OK;
}

int sbe_get_step_stats(sbe_ctx *ctx, sbe_step_stats *out)
{
    if (!ctx || !out) return set_error(ctx, SBE_ERR_STATE, "null argument");
    std::memset(out, 0, sizeof *out);
    if (ctx->path == sbe::PathKind::NBand) {
        out->blocks = 1;
        out->k_per_block = ctx->nk;
        out->accepted = out->max_accepted = out->min_accepted = ctx->steps_done;
        return SBE_OK;
    }
    out->blocks = ctx->nwaves;
    out->k_per_block = 64 * (ctx->adaptive ? ctx->hargs.kpt : 1);
    if (!ctx->adaptive) {
        out->accepted = ctx->steps_done * ctx->nwaves;
        out->max_accepted = out->min_accepted = ctx->steps_done;
        return SBE_OK;
    }
    std::vector<int64_t> c(2 * static_cast<size_t>(ctx->nwaves));
    int rc = hip_check(ctx, hipMemcpyAsync(c.data(), ctx->d_counts, c.size() * sizeof(int64_t), hipMemcpyDeviceToHost, ctx->stream), "copy counts");
    if (!rc) rc = hip_check(ctx, hipStreamSynchronize(ctx->stream), "sync counts");
    if (rc) return rc;
    out->min_accepted = INT64_MAX;
    for (int32_t w = 0; w < ctx->nwaves; ++w) {
        out->accepted += c[2 * w];
        out->rejected += c[2 * w + 1];
        out->max_accepted = std::max(out->max_accepted, c[2 * w]);
        out->min_accepted = std::min(out->min_accepted, c[2 * w]);
    }
    return SBE_OK;
}

int sbe_synchronize(sbe_ctx *ctx)
{
    if (!ctx) return set_error(nullptr, SBE_ERR_STATE, "null context");
    return hip_check(ctx, hipStreamSynchronize(ctx->stream), "hipStreamSynchronize");
}

int sbe_is_idle(const sbe_ctx *ctx)
{
    if (!ctx || !ctx->stream) return 1;
    return hipStreamQuery(ctx->stream) == hipSuccess ? 1 : 0;
}

int sbe_get_series(sbe_ctx *ctx, double *j_inter, double *j_intra, double *rho_bar, double *band_pop)
{
    if (!ctx) return set_error(nullptr, SBE_ERR_STATE, "null context");
    if (ctx->path == sbe::PathKind::NBand) return sbe::nb_series(ctx, j_inter, j_intra, rho_bar, band_pop);
    const int64_t nt = ctx->laser.nt;
    std::vector<double> s(static_cast<size_t>(nt) * ctx->nv);
    int rc = hip_check(ctx, hipMemcpyAsync(s.data(), ctx->d_series, s.size() * sizeof(double), hipMemcpyDeviceToHost, ctx->stream), "copy series");
    int32_t failed = 0;
    if (!rc && ctx->adaptive)
        rc = hip_check(ctx, hipMemcpyAsync(&failed, ctx->d_fail, sizeof failed, hipMemcpyDeviceToHost, ctx->stream), "copy flag");
    if (!rc) rc = hip_check(ctx, hipStreamSynchronize(ctx->stream), "sync series");
    if (rc) return rc;
    if (failed)
        return set_error(ctx, SBE_ERR_NAN, "adaptive stepper: a step size fell below 1e-6 dt or a window needed more "
                                           "than 64 attempts per sample (non-finite state or rtol too tight)");
    // n(t_m) = n(0) R^m for every k when T1 is on (R: RK factor of dn/dt = -n/T1);
    // in the velocity gauge the diagonal relaxes toward rho0 and n stays n(0)
    const double R = ctx->hargs.gam1 != 0.0 && !ctx->velocity
                         ? rk_linear_factor(ctx->tab, -ctx->hargs.gam1 * ctx->laser.dt) : 1.0;
    double nsum = ctx->n_sum;
    bool finite = true;
    for (int64_t n = 0; n < nt; ++n, nsum *= R) {
        const double *r = s.data() + n * ctx->nv;
        const bool done = n < ctx->steps_done;
        for (int i = 0; i < ctx->nv && done; ++i) finite = finite && std::isfinite(r[i]);
        const double ns = done ? nsum : 0.0;
        double dia[3] = {0.0, 0.0, 0.0};  // velocity gauge: sum_k w_k Nval A(t_n) (SBEs.h:683-690,750-757)
        if (ctx->velocity && done) {
            double A[3], E[3];
            ctx->laser.eval(ctx->laser.atmin + ctx->laser.dt * static_cast<double>(n), A, E);
            for (int ax = 0; ax < 3; ++ax) dia[ax] = ctx->model.nval * A[ax] * ctx->grid.total_weight;
        }
        double *jd = ctx->cfg.inter_intra ? j_intra : j_inter;
        if (j_inter) { j_inter[3 * n] = r[0]; j_inter[3 * n + 1] = r[1]; j_inter[3 * n + 2] = 0.0; }
        if (j_intra) { j_intra[3 * n] = r[2]; j_intra[3 * n + 1] = r[3]; j_intra[3 * n + 2] = 0.0; }
        if (jd)
            for (int ax = 0; ax < 3; ++ax) jd[3 * n + ax] += dia[ax];
        if (rho_bar) {
            double *o = rho_bar + n * 8;  // [2][2][re,im], rho = (n + s.sigma)/2
            o[0] = 0.5 * (ns + r[6]); o[1] = 0.0;
            o[2] = 0.5 * r[4];        o[3] = -0.5 * r[5];
            o[4] = 0.5 * r[4];        o[5] = 0.5 * r[5];
            o[6] = 0.5 * (ns - r[6]); o[7] = 0.0;
        }
        if (band_pop) {
            band_pop[2 * n] = 0.5 * (ns - r[7]);
            band_pop[2 * n + 1] = 0.5 * (ns + r[7]);
        }
    }
    if (!finite) return set_error(ctx, SBE_ERR_NAN, "non-finite value in the propagated state");
    return SBE_OK;
}

int sbe_get_state(sbe_ctx *ctx, int32_t basis, double *rho)
{
    if (!ctx) return set_error(nullptr, SBE_ERR_STATE, "null context");
    if (!rho) return set_error(ctx, SBE_ERR_CONFIG, "null output");
    if (ctx->path == sbe::PathKind::NBand) return sbe::nb_state(ctx, basis, rho);
    const size_t bytes = static_cast<size_t>(ctx->nk) * 8 * sizeof(double);
    double *d_out = nullptr;
    int rc = hip_check(ctx, hipMalloc(&d_out, bytes), "hipMalloc");
    if (!rc) rc = hip_check(ctx, sbe::launch_haldane_export(ctx->hargs, d_out, ctx->stream), "export kernel");
    if (!rc) rc = hip_check(ctx, hipMemcpyAsync(rho, d_out, bytes, hipMemcpyDeviceToHost, ctx->stream), "copy state");
    if (!rc) rc = hip_check(ctx, hipStreamSynchronize(ctx->stream), "sync state");
    if (d_out) (void)hipFree(d_out);
    if (rc) return rc;
    if (ctx->velocity) {
        // the state is band-basis at k (phases of band_frame.hpp); orbital: s_orb = sum_i s_i f_i
        if (basis == SBE_BASIS_ORBITAL) {
            const sbe::HaldaneArgs &h = ctx->hargs;
            for (int64_t k = 0; k < ctx->nk; ++k) {
                const double kx = ctx->grid.k[3 * k], ky = ctx->grid.k[3 * k + 1];
                double ca = 0, sa = 0, sb = 0;
                for (int j = 0; j < 3; ++j) {
                    ca += std::cos(kx * h.va[j][0] + ky * h.va[j][1]);
                    sa += std::sin(kx * h.va[j][0] + ky * h.va[j][1]);
                    sb += std::sin(kx * h.vb[j][0] + ky * h.vb[j][1]);
                }
                const sbe::BandFrame F = sbe::band_frame(h.t1 * ca, h.t1 * sa, h.M0 - h.c3 * sb);
                double *o = rho + 8 * k;
                const double n = o[0] + o[6], b[3] = {2 * o[2], 2 * o[5], o[0] - o[6]};
                double s[3];
                for (int c = 0; c < 3; ++c) s[c] = b[0] * F.fx[c] + b[1] * F.fy[c] + b[2] * F.fz[c];
                o[0] = 0.5 * (n + s[2]); o[1] = 0.0;
                o[2] = 0.5 * s[0];       o[3] = -0.5 * s[1];
                o[4] = 0.5 * s[0];       o[5] = 0.5 * s[1];
                o[6] = 0.5 * (n - s[2]); o[7] = 0.0;
            }
        } else if (basis != SBE_BASIS_BAND) {
            return set_error(ctx, SBE_ERR_CONFIG, "unknown basis");
        }
        return SBE_OK;
    }
    if (basis == SBE_BASIS_BAND) {
        // U^dag rho U at k + A(t_now) with the eigenvectors of d.sigma (t_now: the next sample time)
        const double t = ctx->laser.atmin + ctx->laser.dt * static_cast<double>(ctx->steps_done);
        double A[3], E[3];
        ctx->laser.eval(t, A, E);
        const sbe::HaldaneArgs &h = ctx->hargs;
        for (int64_t k = 0; k < ctx->nk; ++k) {
            const double kx = ctx->grid.k[3 * k] + A[0], ky = ctx->grid.k[3 * k + 1] + A[1];
            double sa = 0, ca = 0, sb = 0;
            for (int j = 0; j < 3; ++j) {
                ca += std::cos(kx * h.va[j][0] + ky * h.va[j][1]);
                sa += std::sin(kx * h.va[j][0] + ky * h.va[j][1]);
                sb += std::sin(kx * h.vb[j][0] + ky * h.vb[j][1]);
            }
            // eigenvectors with the phase rule of band_frame.hpp: s_band,i = s . f_i
            const sbe::BandFrame F = sbe::band_frame(h.t1 * ca, h.t1 * sa, h.M0 - h.c3 * sb);
            double *o = rho + 8 * k;
            const double n = o[0] + o[6], sv[3] = {2 * o[2], 2 * o[5], o[0] - o[6]};
            double b[3] = {0, 0, 0};
            for (int c = 0; c < 3; ++c) {
                b[0] += sv[c] * F.fx[c];
                b[1] += sv[c] * F.fy[c];
                b[2] += sv[c] * F.fz[c];
            }
            o[0] = 0.5 * (n + b[2]); o[1] = 0.0;
            o[2] = 0.5 * b[0];       o[3] = -0.5 * b[1];
            o[4] = 0.5 * b[0];       o[5] = 0.5 * b[1];
            o[6] = 0.5 * (n - b[2]); o[7] = 0.0;
        }
    } else if (basis != SBE_BASIS_ORBITAL) {
        return set_error(ctx, SBE_ERR_CONFIG, "unknown basis");
    }
    return SBE_OK;
}

int sbe_get_info(const sbe_ctx *ctx, sbe_info *info)
{
    if (!ctx || !info) return set_error(nullptr, SBE_ERR_STATE, "null context or info");
    info->nt = ctx->laser.nt;
    info->nk_total = ctx->grid.ntotal;
    info->k_begin = ctx->grid.k_begin;
    info->k_end = ctx->grid.k_end;
    info->nband = ctx->N;
    info->nstage = ctx->tab.nstage;
    info->dt = ctx->laser.dt;
    info->atmin = ctx->laser.atmin;
    info->atmax = ctx->laser.atmax;
    info->dV = ctx->grid.dV;
    info->total_weight = ctx->grid.total_weight;
    info->steps_done = ctx->steps_done;
    info->device_bytes = ctx->device_bytes;
    return SBE_OK;
}

}  // extern "C"

// A group of a batch's solves with its own streams: its launch windows follow
// each other on `stream`, its k-sum reductions run on `rstream` while its next
// window steps, and the windows alternate between two partial-row buffers per
// solve (a window waits for the reduction that last read its buffer). A batch
// of two or more solves runs two groups of about equal work, so that one
// group's launch tail overlaps the other's next window; each solve still takes
// the same window sequence, so the groups change no result.
struct BatchGroup {
    std::vector<int> solves;        // the group's solves (batch indices)
    int arg0 = 0;                   // its first argument in each parity half of d_ainsts
    int nblocks = 0;                // adaptive: its blocks
    int32_t *d_map = nullptr;       // adaptive: [nblocks] block -> solve (in the group), then [n_g] first blocks;
                                    // fixed step: [max_blocks] block -> solve, then [n] first blocks (per window)
    sbe::HaldaneArgs *d_insts = nullptr;   // fixed step: [n] the window's arguments (the group's solves)
    std::vector<sbe::HaldaneArgs> h_insts;  // fixed step: host copies
    std::vector<int32_t> h_map;
    hipStream_t stream = nullptr, rstream = nullptr;
    hipEvent_t win_done[2] = {nullptr, nullptr}, red_done[2] = {nullptr, nullptr};
    bool red_pending[2] = {false, false};
    int64_t windows = 0;                 // windows launched (parity = windows & 1)
    sbe::SeriesJob *h_jobs = nullptr;    // [2][n_g] pinned job lists of the reductions
    sbe::SeriesJob *d_jobs = nullptr;    // [2][n_g]
};

struct sbe_batch {
    std::vector<sbe_ctx *> ctx;
    hipStream_t stream = nullptr;  // the first group's step stream; every run ends with it joined to the others
    int max_blocks = 0;
    int rk = 0, prec = 0;
    bool adaptive = false;
    int kpt = 1;
    sbe::AdaptiveArgs *d_ainsts = nullptr;  // adaptive: [2][n], by window parity, group by group
    bool t1 = false, deph = false;
    bool multi = false;  // every member's rows fit one reduction pass
    std::vector<BatchGroup> groups;
    std::vector<double *> part2;  // second partial-row buffer per member (odd windows)
    hipEvent_t run_start = nullptr;
    std::vector<std::pair<hipEvent_t, hipEvent_t>> ev;
    std::string err;
};

namespace {
int batch_fail(sbe_batch *b, int code, const std::string &m)
{
    if (b) b->err = m;
    g_thread_error = m;
    return code;
}
}  // namespace

extern "C" {

const char *sbe_batch_last_error(const sbe_batch *b) { return b ? b->err.c_str() : g_thread_error.c_str(); }

void sbe_batch_destroy(sbe_batch *b)
{
    if (!b) return;
    for (BatchGroup &g : b->groups) {
        if (g.stream) (void)hipStreamSynchronize(g.stream);
        if (g.rstream) (void)hipStreamSynchronize(g.rstream);
    }
    for (auto &p : b->ev) {
        (void)hipEventDestroy(p.first);
        (void)hipEventDestroy(p.second);
    }
    for (sbe_ctx *c : b->ctx) sbe_destroy(c);
    for (double *q : b->part2)
        if (q) (void)hipFree(q);
    if (b->d_ainsts) (void)hipFree(b->d_ainsts);
    if (b->run_start) (void)hipEventDestroy(b->run_start);
    for (BatchGroup &g : b->groups) {
        if (g.d_map) (void)hipFree(g.d_map);
        if (g.d_insts) (void)hipFree(g.d_insts);
        if (g.d_jobs) (void)hipFree(g.d_jobs);
        if (g.h_jobs) (void)hipHostFree(g.h_jobs);
        for (int p = 0; p < 2; ++p) {
            if (g.win_done[p]) (void)hipEventDestroy(g.win_done[p]);
            if (g.red_done[p]) (void)hipEventDestroy(g.red_done[p]);
        }
        if (g.rstream) (void)hipStreamDestroy(g.rstream);
        if (g.stream) (void)hipStreamDestroy(g.stream);
    }
    delete b;
}

int sbe_batch_create(int32_t n, const sbe_config *cfgs, const sbe_model *models, sbe_batch **out)
{
    if (!out || !cfgs || !models || n < 1) return batch_fail(nullptr, SBE_ERR_CONFIG, "bad batch arguments");
    *out = nullptr;
    for (int i = 0; i < n; ++i) {
        if (models[i].kind != SBE_MODEL_HALDANE)
            return batch_fail(nullptr, SBE_ERR_UNSUPPORTED, "batched solves take 2-band (Haldane) models");
        if (cfgs[i].rk_order != cfgs[0].rk_order || cfgs[i].precision != cfgs[0].precision)
            return batch_fail(nullptr, SBE_ERR_CONFIG, "all solves of a batch share rk_order and precision");
        if (cfgs[i].gauge != SBE_GAUGE_LENGTH_WANNIER)
            return batch_fail(nullptr, SBE_ERR_UNSUPPORTED, "batched solves run the LengthWannier gauge");
    }
    sbe_batch *b = new (std::nothrow) sbe_batch();
    if (!b) return batch_fail(nullptr, SBE_ERR_OOM, "host allocation failed");
    b->adaptive = cfgs[0].rk_order == SBE_RK_CASH_KARP;
    t_partial_bytes = kBatchPartialBytes;
    t_batch_member = !b->adaptive;  // the adaptive batch takes each member's kpt
    for (int i = 0; i < n; ++i) {
        sbe_ctx *c = nullptr;
        const int rc = sbe_create(&cfgs[i], &models[i], -1, -1, &c);
        if (rc) {
            t_partial_bytes = kPartialBytes;
            t_batch_member = false;
            const std::string m = "solve " + std::to_string(i) + ": " + g_thread_error;
            sbe_batch_destroy(b);
            return batch_fail(nullptr, rc, m);
        }
        b->ctx.push_back(c);
        b->max_blocks += static_cast<int>(b->adaptive ? c->nk_stride / (sbe::kAdaptiveBlock * c->hargs.kpt)
                                                     : c->nk_stride / kBlock);
        b->kpt = c->hargs.kpt;
        b->t1 = b->t1 || c->hargs.gam1 != 0.0;
        b->deph = b->deph || c->hargs.gam1 != 0.0 || c->hargs.gam2 != 0.0;
    }
    t_partial_bytes = kPartialBytes;
    t_batch_member = false;
    b->rk = cfgs[0].rk_order;
    b->prec = cfgs[0].precision;
    // groups of about equal work (k points x samples): longest first, each solve to the lighter
    // group (one group for a single solve)
    const int ng = n >= 2 ? 2 : 1;
    b->groups.resize(ng);
    {
        std::vector<int> order(n);
        std::vector<double> work(n);
        for (int i = 0; i < n; ++i) {
            order[i] = i;
            work[i] = static_cast<double>(b->ctx[i]->nk_stride) * static_cast<double>(b->ctx[i]->laser.nt);
        }
        std::stable_sort(order.begin(), order.end(), [&](int x, int y) { return work[x] > work[y]; });
        double load[2] = {0.0, 0.0};
        for (int i : order) {
            const int g = ng == 2 && load[1] < load[0] ? 1 : 0;
            load[g] += work[i];
            b->groups[g].solves.push_back(i);
        }
        // within a group, the strongest fields first (they take the most steps per window), so
        // the slowest blocks are dispatched early rather than left for the launch's tail
        auto field = [&](int i) {
            double e = 0.0;
            const sbe_config &cf = cfgs[i];
            for (int q = 0; q < cf.npulses; ++q) e = std::max(e, std::fabs(cf.pulses[q].E0));
            return e;
        };
        for (BatchGroup &g : b->groups)
            std::stable_sort(g.solves.begin(), g.solves.end(), [&](int x, int y) { return field(x) > field(y); });
    }
    hipError_t e = hipEventCreateWithFlags(&b->run_start, hipEventDisableTiming);
    int arg0 = 0;
    for (BatchGroup &g : b->groups) {
        g.arg0 = arg0;
        arg0 += static_cast<int>(g.solves.size());
        const size_t nj = 2 * g.solves.size();
        if (e == hipSuccess) e = hipStreamCreateWithFlags(&g.stream, hipStreamNonBlocking);
        if (e == hipSuccess) e = hipStreamCreateWithFlags(&g.rstream, hipStreamNonBlocking);
        for (int p = 0; p < 2 && e == hipSuccess; ++p) {
            e = hipEventCreateWithFlags(&g.win_done[p], hipEventDisableTiming);
            if (e == hipSuccess) e = hipEventCreateWithFlags(&g.red_done[p], hipEventDisableTiming);
        }
        if (e == hipSuccess) e = hipMalloc(&g.d_jobs, sizeof(sbe::SeriesJob) * nj);
        if (e == hipSuccess) e = hipHostMalloc(&g.h_jobs, sizeof(sbe::SeriesJob) * nj, hipHostMallocDefault);
    }
    b->stream = b->groups[0].stream;
    if (!b->adaptive)
        for (BatchGroup &g : b->groups) {
            g.h_insts.resize(n);
            g.h_map.resize(b->max_blocks + n);
            if (e == hipSuccess) e = hipMalloc(&g.d_insts, sizeof(sbe::HaldaneArgs) * n);
            if (e == hipSuccess) e = hipMalloc(&g.d_map, sizeof(int32_t) * (b->max_blocks + n));
        }
    b->part2.assign(n, nullptr);
    for (int i = 0; i < n && e == hipSuccess; ++i) {
        const sbe_ctx *c = b->ctx[i];
        e = hipMalloc(&b->part2[i], static_cast<size_t>(c->chunk_steps) * c->nwaves * c->nv * sizeof(double));
    }
    b->multi = true;
    for (sbe_ctx *c : b->ctx) b->multi = b->multi && sbe::series_multi_ok(c->nwaves);
    if (e == hipSuccess && b->adaptive) {
        // constant for the batch's lifetime: every solve's arguments, [2][n] by window parity
        // (odd windows write part2) and group by group, and each group's block -> solve map
        // (solves with no samples left return at once)
        e = hipMalloc(&b->d_ainsts, sizeof(sbe::AdaptiveArgs) * 2 * n);
        std::vector<sbe::AdaptiveArgs> ai(2 * n);
        for (BatchGroup &g : b->groups) {
            const int ngs = static_cast<int>(g.solves.size());
            std::vector<int32_t> map;
            std::vector<int32_t> first(ngs);
            for (int r = 0; r < ngs; ++r) {
                const int i = g.solves[r];
                ai[g.arg0 + r] = b->ctx[i]->aargs;
                ai[n + g.arg0 + r] = b->ctx[i]->aargs;
                ai[n + g.arg0 + r].h.partials = b->part2[i];
                first[r] = static_cast<int32_t>(map.size());
                for (int64_t q = 0; q < b->ctx[i]->nk_stride / (sbe::kAdaptiveBlock * b->kpt); ++q) map.push_back(r);
            }
            g.nblocks = static_cast<int>(map.size());
            map.insert(map.end(), first.begin(), first.end());
            if (e == hipSuccess) e = hipMalloc(&g.d_map, sizeof(int32_t) * map.size());
            if (e == hipSuccess) e = hipMemcpy(g.d_map, map.data(), sizeof(int32_t) * map.size(), hipMemcpyHostToDevice);
        }
        if (e == hipSuccess) e = hipMemcpy(b->d_ainsts, ai.data(), sizeof(sbe::AdaptiveArgs) * 2 * n, hipMemcpyHostToDevice);
    }
    if (e != hipSuccess) {
        const std::string m = std::string("batch set-up: ") + hipGetErrorString(e);
        sbe_batch_destroy(b);
        return batch_fail(nullptr, SBE_ERR_HIP, m);
    }
    *out = b;
    return SBE_OK;
}

namespace {
// One launch window of group g: the step launch on g.stream (its partial rows in
// this window's parity buffers), then the group's k sums on g.rstream.
hipError_t batch_window(sbe_batch *b, BatchGroup &g, int64_t launch, std::vector<sbe::HaldaneArgs> &args)
{
    const int n = static_cast<int>(b->ctx.size());
    // the fixed-step path uploads this group's arguments and block map from its own host copies
    std::vector<sbe::HaldaneArgs> &insts = b->adaptive ? args : g.h_insts;
    std::vector<int32_t> &map = g.h_map;
    const int par = static_cast<int>(g.windows & 1);
    int nb = 0;
    for (int i : g.solves) {
        sbe_ctx *c = b->ctx[i];
        insts[i] = c->hargs;
        if (par) insts[i].partials = b->part2[i];
        insts[i].step0 = c->steps_done;
        insts[i].nsteps = static_cast<int32_t>(std::min<int64_t>(launch, c->laser.nt - c->steps_done));
        if (b->adaptive) continue;  // its block maps are static (set up at creation)
        map[b->max_blocks + i] = nb;
        if (insts[i].nsteps > 0)
            for (int64_t q = 0; q < c->nk_stride / kBlock; ++q) map[nb++] = i;
    }
    // this parity's buffers are free once the reduction that last read them has run
    hipError_t e = g.red_pending[par] ? hipStreamWaitEvent(g.stream, g.red_done[par], 0) : hipSuccess;
    // solve -> first block table sits after the block -> solve table
    int32_t *d_first = b->adaptive ? nullptr : g.d_map + b->max_blocks;
    if (e == hipSuccess && !b->adaptive) {  // the adaptive batch keeps its arguments on the device
        e = hipMemcpyAsync(g.d_insts, insts.data(), sizeof(sbe::HaldaneArgs) * n, hipMemcpyHostToDevice, g.stream);
        if (e == hipSuccess) e = hipMemcpyAsync(g.d_map, map.data(), sizeof(int32_t) * nb, hipMemcpyHostToDevice, g.stream);
        if (e == hipSuccess)
            e = hipMemcpyAsync(d_first, map.data() + b->max_blocks, sizeof(int32_t) * n, hipMemcpyHostToDevice, g.stream);
    }
    hipEvent_t ea = nullptr, eb = nullptr;
    if (e == hipSuccess) e = hipEventCreate(&ea);
    if (e == hipSuccess) e = hipEventCreate(&eb);
    if (e == hipSuccess) e = hipEventRecord(ea, g.stream);
    const int ngs = static_cast<int>(g.solves.size());
    sbe::AdaptiveArgs *d_ai = b->d_ainsts + static_cast<int64_t>(par) * n + g.arg0;
    if (e == hipSuccess && b->adaptive) {
        e = sbe::launch_haldane_adaptive(d_ai, g.d_map, g.d_map + g.nblocks, g.nblocks, launch, b->prec, b->deph,
                                         b->kpt, g.stream);
        if (e == hipSuccess) e = sbe::launch_adaptive_advance(d_ai, ngs, launch, g.stream);
    } else if (e == hipSuccess) {
        e = sbe::launch_haldane_batch(g.d_insts, g.d_map, d_first, nb, b->rk, b->prec, b->t1, b->deph, g.stream);
    }
    if (e == hipSuccess) e = hipEventRecord(eb, g.stream);
    if (ea && eb) b->ev.push_back({ea, eb});
    else {
        if (ea) (void)hipEventDestroy(ea);
        if (eb) (void)hipEventDestroy(eb);
    }
    if (e == hipSuccess) e = hipEventRecord(g.win_done[par], g.stream);
    // the k sums of this window: one launch for the group when each member fits one pass (the
    // same per-step arithmetic as a member's own reduction, so the same bits)
    if (e == hipSuccess) e = hipStreamWaitEvent(g.rstream, g.win_done[par], 0);
    // the pinned job list of this parity was last read by the reduction two windows back
    if (e == hipSuccess && g.red_pending[par]) e = hipEventSynchronize(g.red_done[par]);
    sbe::SeriesJob *h_jobs = g.h_jobs + static_cast<int64_t>(par) * ngs;
    sbe::SeriesJob *d_jobs = g.d_jobs + static_cast<int64_t>(par) * ngs;
    int njobs = 0;
    int64_t total = 0;
    for (int i : g.solves) {
        if (e != hipSuccess) break;
        sbe_ctx *c = b->ctx[i];
        if (insts[i].nsteps <= 0) continue;
        double *rows = par ? b->part2[i] : c->d_partials;
        if (b->multi) {
            h_jobs[njobs++] = {rows, c->d_series + c->steps_done * c->nv, c->nwaves, insts[i].nsteps, total,
                               c->nonfinite_d};
            total += insts[i].nsteps;
        } else {
            e = sbe::launch_series_reduce(rows, insts[i].nsteps, c->nwaves, c->nv, c->d_series + c->steps_done * c->nv,
                                          g.rstream, c->nonfinite_d);
        }
        c->steps_done += insts[i].nsteps;
    }
    if (e == hipSuccess && njobs > 0) {
        e = hipMemcpyAsync(d_jobs, h_jobs, sizeof(sbe::SeriesJob) * njobs, hipMemcpyHostToDevice, g.rstream);
        if (e == hipSuccess) e = sbe::launch_series_reduce_multi(d_jobs, njobs, total, sbe::kHaldaneNV, g.rstream);
    }
    if (e == hipSuccess) e = hipEventRecord(g.red_done[par], g.rstream);
    if (e == hipSuccess) {
        g.red_pending[par] = true;
        ++g.windows;
    }
    return e;
}
}  // namespace

int sbe_batch_run(sbe_batch *b, int64_t max_steps)
{
    if (!b) return batch_fail(nullptr, SBE_ERR_STATE, "null batch");
    const int n = static_cast<int>(b->ctx.size());
    int64_t budget = max_steps > 0 ? max_steps : INT64_MAX;
    std::vector<sbe::HaldaneArgs> insts(n);
    // work enqueued on the first stream before this run (a reset) precedes every group
    hipError_t e = hipEventRecord(b->run_start, b->stream);
    for (size_t gi = 1; gi < b->groups.size() && e == hipSuccess; ++gi)
        e = hipStreamWaitEvent(b->groups[gi].stream, b->run_start, 0);
    while (budget > 0 && e == hipSuccess) {
        int64_t launch = INT64_MAX;
        for (sbe_ctx *c : b->ctx)
            if (c->steps_done < c->laser.nt) launch = std::min<int64_t>(launch, c->chunk_steps);
        if (launch == INT64_MAX) break;  // all done
        if (b->adaptive) launch = kAdaptiveWindow;  // every solve on the fixed window grid
        launch = std::min(launch, budget);
        for (BatchGroup &g : b->groups) {
            bool active = false;
            for (int i : g.solves) active = active || b->ctx[i]->steps_done < b->ctx[i]->laser.nt;
            if (active && e == hipSuccess) e = batch_window(b, g, launch, insts);
        }
        budget -= launch;
    }
    // the first stream completes no earlier than every group's last reduction (synchronize and
    // reset wait on it)
    for (BatchGroup &g : b->groups) {
        if (e != hipSuccess || g.windows == 0) continue;
        e = hipStreamWaitEvent(b->stream, g.red_done[(g.windows + 1) & 1], 0);
    }
    if (e != hipSuccess) return batch_fail(b, SBE_ERR_HIP, std::string("batch launch: ") + hipGetErrorString(e));
    return SBE_OK;
}

int sbe_batch_synchronize(sbe_batch *b)
{
    if (!b) return batch_fail(nullptr, SBE_ERR_STATE, "null batch");
    const hipError_t e = hipStreamSynchronize(b->stream);
    return e == hipSuccess ? SBE_OK : batch_fail(b, SBE_ERR_HIP, hipGetErrorString(e));
}

int sbe_batch_reset(sbe_batch *b)
{
    if (!b) return batch_fail(nullptr, SBE_ERR_STATE, "null batch");
    int rc = sbe_batch_synchronize(b);
    if (rc) return rc;
    for (sbe_ctx *c : b->ctx) {
        rc = reset_2band(c, b->stream);
        if (rc) return batch_fail(b, rc, c->err);
    }
    return sbe_batch_synchronize(b);
}

int sbe_batch_get_info(const sbe_batch *b, int32_t i, sbe_info *info)
{
    if (!b || i < 0 || i >= static_cast<int32_t>(b->ctx.size())) return batch_fail(nullptr, SBE_ERR_RANGE, "bad solve index");
    return sbe_get_info(b->ctx[i], info);
}

int sbe_batch_get_series(sbe_batch *b, int32_t i, double *j_inter, double *j_intra, double *rho_bar, double *band_pop)
{
    if (!b || i < 0 || i >= static_cast<int32_t>(b->ctx.size())) return batch_fail(nullptr, SBE_ERR_RANGE, "bad solve index");
    int rc = sbe_batch_synchronize(b);
    if (rc) return rc;
    rc = sbe_get_series(b->ctx[i], j_inter, j_intra, rho_bar, band_pop);
    if (rc) return batch_fail(b, rc, b->ctx[i]->err);
    return SBE_OK;
}

int sbe_batch_get_step_stats(sbe_batch *b, int32_t i, sbe_step_stats *out)
{
    if (!b || i < 0 || i >= static_cast<int32_t>(b->ctx.size())) return batch_fail(nullptr, SBE_ERR_RANGE, "bad solve index");
    int rc = sbe_batch_synchronize(b);
    if (rc) return rc;
    rc = sbe_get_step_stats(b->ctx[i], out);
    if (rc) return batch_fail(b, rc, b->ctx[i]->err);
    return SBE_OK;
}

int sbe_batch_get_timing(sbe_batch *b, double *ms, int64_t *launches)
{
    if (!b) return batch_fail(nullptr, SBE_ERR_STATE, "null batch");
    int rc = sbe_batch_synchronize(b);
    // the time during which at least one step launch ran: the union of the launch
    // intervals (the groups' launches overlap), from their offsets to the first start
    std::vector<std::pair<double, double>> iv;
    iv.reserve(b->ev.size());
    for (auto &p : b->ev) {
        float t0 = 0.f, t1 = 0.f;
        if (!rc && (hipEventElapsedTime(&t0, b->ev[0].first, p.first) != hipSuccess ||
                    hipEventElapsedTime(&t1, b->ev[0].first, p.second) != hipSuccess))
            rc = batch_fail(b, SBE_ERR_HIP, "event timing");
        iv.push_back({t0, t1});
    }
    std::sort(iv.begin(), iv.end());
    double busy = 0.0, hi = -1e300;
    for (const auto &v : iv) {
        if (v.second <= hi) continue;
        busy += v.second - std::max(v.first, hi);
        hi = v.second;
    }
    for (auto &p : b->ev) {
        (void)hipEventDestroy(p.first);
        (void)hipEventDestroy(p.second);
    }
    if (ms) *ms = rc ? 0.0 : busy;
    if (launches) *launches = static_cast<int64_t>(b->ev.size());
    b->ev.clear();
    return rc;
}

void sbe_destroy(sbe_ctx *ctx)
{
    if (!ctx) return;
    if (ctx->stream) (void)hipStreamSynchronize(ctx->stream);
    drop_events(ctx);
    for (auto &b : ctx->owned) (void)hipFree(b.ptr);
    if (ctx->nonfinite_h) (void)hipHostFree(ctx->nonfinite_h);
    if (ctx->stream) (void)hipStreamDestroy(ctx->stream);
    delete ctx;
}

}  // extern "C"
