// tuner.cpp - the launch-plan autotuner and its per-device cache.
//
// A context's plan has a few free choices that do not change the method:
// k points per lane of the 2-band kernels (1 or 2: how many independent
// chains a lane carries, and how many waves a grid makes), steps per launch
// of the fixed-step kernels (launch overhead against the partial-sum buffer),
// the N-band stage kernels (MFMA register tiles, N <= 16, or the LDS
// kernels) and the N-band Fourier sums (separable passes on product grids,
// or the dense GEMM). Each has a default rule; sbe_autotune times every
// candidate on throwaway contexts of the caller's problem and keeps the
// fastest in a cache keyed by device and problem class (path, precision,
// tableau, bands, gauge, k points to the nearest power of two).
// sbe_create consults the cache (a GPU's cache applies to the device the
// context is created on); the ANTELOPE_* environment switches still win.
// Steps per launch never change a result; k points per lane change the
// grouping of the wave sums (results agree to rounding); the N-band choices
// agree to rounding too (tested).
#include <hip/hip_runtime.h>

#include <chrono>
#include <cmath>
#include <map>
#include <mutex>
#include <string>
#include <tuple>
#include <vector>

#include "analytic.hpp"
#include "sbe_internal.hpp"

namespace sbe {
namespace {

using Key = std::tuple<int, int, int, int, int, int, int>;  // device, path, precision, rk, N, log2 nk, gauge
std::mutex g_mu;
std::map<Key, sbe_tuning> g_cache;

int current_device()
{
    int d = 0;
    return hipGetDevice(&d) == hipSuccess ? d : -1;
}

Key make_key(const sbe_config &cfg, const sbe_model &m, int64_t nk)
{
    const int path = m.kind == SBE_MODEL_HALDANE ? (cfg.rk_order == SBE_RK_CASH_KARP ? 1 : 0) : 2;
    int lg = 0;
    while (lg < 62 && (int64_t(1) << (lg + 1)) <= nk) ++lg;
    // the nearest power of two (2^lg or 2^(lg+1))
    if (nk - (int64_t(1) << lg) > (int64_t(1) << (lg + 1)) - nk) ++lg;
    return Key{current_device(), path, cfg.precision, cfg.rk_order, m.nband, lg, cfg.gauge};
}

int64_t grid_points(const sbe_config &cfg)
{
    int64_t n = 1;
    for (int i = 0; i < 3; ++i) n *= cfg.npoints[i] > 0 ? cfg.npoints[i] : 1;
    return n;
}

double now_s()
{
    return std::chrono::duration<double>(std::chrono::steady_clock::now().time_since_epoch()).count();
}

// seconds per step of a throwaway context with tuning t (trial steps after a warm-up launch)
int trial(const sbe_config &cfg, const sbe_model &m, const sbe_tuning &t, int64_t steps, double &sec, std::string &err)
{
    const Key key = make_key(cfg, m, grid_points(cfg));
    {
        std::lock_guard<std::mutex> lk(g_mu);
        g_cache[key] = t;
    }
    sbe_ctx *ctx = nullptr;
    int rc = sbe_create(&cfg, &m, -1, -1, &ctx);
    if (rc) {
        err = sbe_last_error(nullptr);
        return rc;
    }
    sbe_info inf{};
    sbe_get_info(ctx, &inf);
    const int64_t warm = std::min<int64_t>(std::max<int64_t>(1, steps / 4), inf.nt);
    const int64_t n = std::min<int64_t>(steps, inf.nt - warm);
    rc = sbe_run(ctx, 0, warm);
    if (!rc) rc = sbe_synchronize(ctx);
    const double t0 = now_s();
    if (!rc && n > 0) rc = sbe_run(ctx, warm, warm + n);
    if (!rc) rc = sbe_synchronize(ctx);
    sec = (now_s() - t0) / static_cast<double>(std::max<int64_t>(n, 1));
    if (rc) err = sbe_last_error(ctx);
    sbe_destroy(ctx);
    return rc;
}

}  // namespace

bool tuning_lookup(const sbe_config &cfg, const sbe_model &m, int64_t nk, sbe_tuning &out)
{
    std::lock_guard<std::mutex> lk(g_mu);
    auto it = g_cache.find(make_key(cfg, m, nk));
    if (it == g_cache.end()) return false;
    out = it->second;
    return true;
}

}  // namespace sbe

extern "C" {

int sbe_tuning_get(const sbe_config *cfg, const sbe_model *model, sbe_tuning *out)
{
    if (!cfg || !model || !out) return sbe::set_thread_error(SBE_ERR_CONFIG, "null argument");
    std::string err;
    const int rc = sbe::validate_inputs(cfg, model, err);
    if (rc) return sbe::set_thread_error(rc, err);
    if (!sbe::tuning_lookup(*cfg, *model, sbe::grid_points(*cfg), *out)) *out = sbe_tuning{};
    return SBE_OK;
}

int sbe_tuning_set(const sbe_config *cfg, const sbe_model *model, const sbe_tuning *t)
{
    if (!cfg || !model || !t) return sbe::set_thread_error(SBE_ERR_CONFIG, "null argument");
    std::string err;
    const int rc = sbe::validate_inputs(cfg, model, err);
    if (rc) return sbe::set_thread_error(rc, err);
    if (t->kpt < 0 || t->kpt > 2 || t->chunk_steps < 0 || t->nb_kernels < 0 || t->nb_kernels > 2 || t->nb_sums < 0 ||
        t->nb_sums > 2)
        return sbe::set_thread_error(SBE_ERR_CONFIG, "tuning values out of range");
    sbe_tuning v = *t;
    v.tuned = 1;
    std::lock_guard<std::mutex> lk(sbe::g_mu);
    sbe::g_cache[sbe::make_key(*cfg, *model, sbe::grid_points(*cfg))] = v;
    return SBE_OK;
}

void sbe_tuning_clear(void)
{
    std::lock_guard<std::mutex> lk(sbe::g_mu);
    sbe::g_cache.clear();
}

int sbe_autotune(const sbe_config *cfg, const sbe_model *model, int64_t trial_steps, sbe_tuning *out)
{
    if (!cfg || !model || !out) return sbe::set_thread_error(SBE_ERR_CONFIG, "null argument");
    std::string err;
    int rc = sbe::validate_inputs(cfg, model, err);
    if (rc) return sbe::set_thread_error(rc, err);
    if (trial_steps < 1) trial_steps = 64;
    const sbe::Key key = sbe::make_key(*cfg, *model, sbe::grid_points(*cfg));
    {
        std::lock_guard<std::mutex> lk(sbe::g_mu);
        sbe::g_cache.erase(key);  // the default rule for the reference plan below
    }
    // candidates of this problem's path (0 = the default rule)
    std::vector<sbe_tuning> cand;
    if (model->kind == SBE_MODEL_HALDANE) {
        for (int kpt = 1; kpt <= 2; ++kpt) {
            sbe_tuning t{};
            t.kpt = kpt;
            cand.push_back(t);
            if (cfg->rk_order != SBE_RK_CASH_KARP && cfg->gauge == SBE_GAUGE_LENGTH_WANNIER) {
                t.chunk_steps = 32;  // short launches: less partial-sum memory, more launches
                cand.push_back(t);
            }
        }
    } else {
        const bool tiles = model->nband <= 16;
        const bool wannier = model->kind == SBE_MODEL_WANNIER && cfg->precision != SBE_PREC_SPLIT16;
        for (int k = tiles ? 1 : 2; k <= 2; ++k)
            for (int s = wannier ? 1 : 0; s <= (wannier ? 2 : 0); ++s) {
                sbe_tuning t{};
                t.nb_kernels = k;
                t.nb_sums = s;
                cand.push_back(t);
            }
    }
    sbe_tuning best{};
    double best_s = 0.0;
    for (sbe_tuning t : cand) {
        t.tuned = 1;
        double sec = 0.0;
        rc = sbe::trial(*cfg, *model, t, trial_steps, sec, err);
        if (rc) break;
        if (best_s == 0.0 || sec < best_s) best = t, best_s = sec;
    }
    {
        std::lock_guard<std::mutex> lk(sbe::g_mu);
        if (rc) {
            sbe::g_cache.erase(key);
        } else {
            best.ms_per_step = best_s * 1e3;
            best.candidates = static_cast<int32_t>(cand.size());
            sbe::g_cache[key] = best;
        }
    }
    if (rc) return sbe::set_thread_error(rc, "autotune trial: " + err);
    *out = best;
    return SBE_OK;
}

int sbe_batch_autotune(int32_t n, const sbe_config *cfgs, const sbe_model *models, int64_t trial_steps,
                       sbe_tuning *out)
{
    if (n < 1 || !cfgs || !models || !out) return sbe::set_thread_error(SBE_ERR_CONFIG, "bad batch arguments");
    if (trial_steps < 1) trial_steps = 64;
    std::string err;
    for (int i = 0; i < n; ++i) {
        const int rc = sbe::validate_inputs(&cfgs[i], &models[i], err);
        if (rc) return sbe::set_thread_error(rc, "solve " + std::to_string(i) + ": " + err);
    }
    // members of a fixed-step batch always run one k point per lane in the batch kernel;
    // the adaptive batch runs each member's plan: one or two k points per lane
    const bool adaptive = cfgs[0].rk_order == SBE_RK_CASH_KARP;
    std::vector<sbe_tuning> cand;
    for (int kpt = 1; kpt <= (adaptive ? 2 : 1); ++kpt) {
        sbe_tuning t{};
        t.kpt = kpt;
        t.tuned = 1;
        cand.push_back(t);
    }
    auto set_all = [&](const sbe_tuning &t) {
        std::lock_guard<std::mutex> lk(sbe::g_mu);
        for (int i = 0; i < n; ++i) sbe::g_cache[sbe::make_key(cfgs[i], models[i], sbe::grid_points(cfgs[i]))] = t;
    };
    sbe_tuning best{};
    double best_s = 0.0;
    int rc = SBE_OK;
    for (const sbe_tuning &t : cand) {
        set_all(t);
        sbe_batch *b = nullptr;
        rc = sbe_batch_create(n, cfgs, models, &b);
        if (rc) {
            err = sbe_batch_last_error(nullptr);
            break;
        }
        rc = sbe_batch_run(b, std::max<int64_t>(1, trial_steps / 4));  // warm-up (code objects, caches)
        if (!rc) rc = sbe_batch_synchronize(b);
        const double t0 = sbe::now_s();
        if (!rc) rc = sbe_batch_run(b, trial_steps);
        if (!rc) rc = sbe_batch_synchronize(b);
        const double sec = (sbe::now_s() - t0) / static_cast<double>(trial_steps);
        if (rc) err = sbe_batch_last_error(b);
        sbe_batch_destroy(b);
        if (rc) break;
        if (best_s == 0.0 || sec < best_s) best = t, best_s = sec;
    }
    if (rc) {
        std::lock_guard<std::mutex> lk(sbe::g_mu);
        for (int i = 0; i < n; ++i) sbe::g_cache.erase(sbe::make_key(cfgs[i], models[i], sbe::grid_points(cfgs[i])));
        return sbe::set_thread_error(rc, "batch autotune trial: " + err);
    }
    best.ms_per_step = best_s * 1e3;
    best.candidates = static_cast<int32_t>(cand.size());
    set_all(best);
    *out = best;
    return SBE_OK;
}

}  // extern "C"
