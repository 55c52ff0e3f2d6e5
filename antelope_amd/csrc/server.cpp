// server.cpp - continuous-batching request loop over SBE contexts.
//
// Requests (a config and a model each) queue on submit. Every scheduling
// round (sbe_server_step):
//   1. retire requests whose last slice has finished (an event query, no
//      wait): copy their series to host memory and park the context; parked
//      contexts are freed in groups, because freeing device memory waits for
//      the whole device, and their memory stays charged to the budget until
//      then;
//   2. admit queued requests, first fit in arrival order, while their device
//      footprint (sbe_query_footprint) fits the remaining HBM budget and a
//      slot is free; admission creates the context, so its density matrices
//      live on the GPU from then on;
//   3. keep up to kDepth slices of steps in flight per admitted request, each
//      request on its own stream, so the device never waits for the host and
//      slices of different requests overlap on the device. A slice is sized
//      by work (k points x steps), so a small grid advances more steps per
//      slice than a large one.
// Requests never share a context, so a result does not depend on what else
// was running (the k sums are reduced in a fixed order per context).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cstring>
#include <deque>
#include <map>
#include <memory>
#include <string>
#include <vector>

#include "antelope/sbe_server.h"
#include "sbe_internal.hpp"

namespace {

double now_s()
{
    using clk = std::chrono::steady_clock;
    return std::chrono::duration<double>(clk::now().time_since_epoch()).count();
}

constexpr int kDepth = 2;                    // slices in flight per request
constexpr double kSliceKSteps = 67108864.0;  // 2-band work per slice, about 1.5 ms of device time
constexpr size_t kParkMax = 8;               // parked contexts freed together

struct Request {
    int64_t id = 0;
    int32_t state = SBE_REQ_QUEUED;
    sbe_config cfg{};
    std::vector<sbe_pulse> pulses;
    sbe_model model{};
    std::vector<int32_t> rvec;
    std::vector<double> rweight, ham, pos;
    size_t footprint = 0;
    sbe_ctx *ctx = nullptr;
    hipEvent_t ev[kDepth] = {};
    int64_t slice = 0;
    int64_t nt = 0;
    int nband = 0;
    double t_submit = 0, t_admit = 0, t_done = 0;
    std::vector<double> j_inter, j_intra, rho_bar, band_pop;
    std::string err;
    int status = SBE_OK;
};

}  // namespace

struct sbe_server {
    size_t budget = 0, in_use = 0;
    int32_t max_active = 0;
    int64_t slice = 0;
    int64_t next_id = 1;
    std::map<int64_t, std::unique_ptr<Request>> reqs;
    std::deque<int64_t> queue;
    std::vector<int64_t> active;
    std::vector<std::pair<sbe_ctx *, size_t>> parked;  // retired contexts not yet freed, their footprints
    std::string err;
    int64_t rounds = 0;
};

namespace {

int srv_fail(sbe_server *s, int code, const std::string &m)
{
    if (s) s->err = m;
    return code;
}

void flush_parked(sbe_server *s)
{
    for (auto &p : s->parked) {
        sbe_destroy(p.first);
        s->in_use -= p.second;
    }
    s->parked.clear();
}

void drop_events(Request &r)
{
    for (auto &e : r.ev)
        if (e) {
            (void)hipEventDestroy(e);
            e = nullptr;
        }
}

int in_flight(const Request &r)
{
    int n = 0;
    for (auto e : r.ev)
        if (e && hipEventQuery(e) == hipErrorNotReady) ++n;
    return n;
}

void finish(sbe_server *s, Request &r, int status, const std::string &msg)
{
    if (r.ctx) {
        if (status == SBE_OK) {
            const int N = r.nband;
            r.j_inter.assign(3 * r.nt, 0.0);
            r.j_intra.assign(3 * r.nt, 0.0);
            r.rho_bar.assign(static_cast<size_t>(2) * N * N * r.nt, 0.0);
            r.band_pop.assign(static_cast<size_t>(N) * r.nt, 0.0);
            status = sbe_get_series(r.ctx, r.j_inter.data(), r.j_intra.data(), r.rho_bar.data(), r.band_pop.data());
            if (status != SBE_OK) r.err = sbe_last_error(r.ctx);
        } else {
            (void)sbe_synchronize(r.ctx);
        }
        drop_events(r);
        s->parked.emplace_back(r.ctx, r.footprint);
        r.ctx = nullptr;
        if (s->parked.size() >= kParkMax) flush_parked(s);
    }
    r.status = status;
    if (status != SBE_OK && r.err.empty()) r.err = msg;
    r.state = status == SBE_OK ? SBE_REQ_DONE : SBE_REQ_FAILED;
    r.t_done = now_s();
}

}  // namespace

extern "C" {

int sbe_server_create(size_t hbm_budget_bytes, int32_t max_active, int64_t slice_steps, sbe_server **out)
{
    if (!out) return SBE_ERR_CONFIG;
    auto *s = new (std::nothrow) sbe_server();
    if (!s) return SBE_ERR_OOM;
    if (hbm_budget_bytes == 0) {
        size_t fr = 0, tot = 0;
        if (hipMemGetInfo(&fr, &tot) != hipSuccess) {
            delete s;
            return SBE_ERR_NODEVICE;
        }
        hbm_budget_bytes = fr / 10 * 8;  // leave a fifth of the free memory to others
    }
    s->budget = hbm_budget_bytes;
    s->max_active = max_active > 0 ? max_active : 16;
    s->slice = slice_steps > 0 ? slice_steps : 512;
    *out = s;
    return SBE_OK;
}

void sbe_server_destroy(sbe_server *s)
{
    if (!s) return;
    for (auto &kv : s->reqs)
        if (kv.second->ctx) {
            drop_events(*kv.second);
            sbe_destroy(kv.second->ctx);
        }
    flush_parked(s);
    delete s;
}

const char *sbe_server_last_error(const sbe_server *s) { return s ? s->err.c_str() : ""; }

int sbe_server_submit(sbe_server *s, const sbe_config *cfg, const sbe_model *model, int64_t *id)
{
    if (!s || !cfg || !model || !id) return srv_fail(s, SBE_ERR_CONFIG, "null argument");
    auto r = std::make_unique<Request>();
    r->cfg = *cfg;
    r->pulses.assign(cfg->pulses, cfg->pulses + cfg->npulses);
    r->cfg.pulses = r->pulses.data();
    r->model = *model;
    if (model->kind == SBE_MODEL_WANNIER) {  // deep copy of the tables
        const size_t nr = static_cast<size_t>(model->nrpts), nn = static_cast<size_t>(model->nband) * model->nband;
        r->rvec.assign(model->rvec, model->rvec + 3 * nr);
        r->rweight.assign(model->rweight, model->rweight + nr);
        r->ham.assign(model->ham, model->ham + 2 * nr * nn);
        if (model->pos) r->pos.assign(model->pos, model->pos + 6 * nr * nn);
        r->model.rvec = r->rvec.data();
        r->model.rweight = r->rweight.data();
        r->model.ham = r->ham.data();
        r->model.pos = model->pos ? r->pos.data() : nullptr;
    }
    const int rc = sbe_query_footprint(&r->cfg, &r->model, -1, -1, &r->footprint);
    if (rc != SBE_OK) return srv_fail(s, rc, sbe_last_error(nullptr));
    if (r->footprint > s->budget)
        return srv_fail(s, SBE_ERR_OOM, "request needs " + std::to_string(r->footprint) + " bytes, budget is " +
                                            std::to_string(s->budget));
    r->id = s->next_id++;
    r->t_submit = now_s();
    *id = r->id;
    s->queue.push_back(r->id);
    s->reqs[r->id] = std::move(r);
    return SBE_OK;
}

int sbe_server_step(sbe_server *s)
{
    if (!s) return SBE_ERR_STATE;
    ++s->rounds;
    // 1. retire finished requests (every step enqueued, no slice in flight)
    for (auto it = s->active.begin(); it != s->active.end();) {
        Request &r = *s->reqs[*it];
        if (r.ctx->steps_done >= r.nt && in_flight(r) == 0) {
            finish(s, r, SBE_OK, "");
            it = s->active.erase(it);
        } else {
            ++it;
        }
    }
    // 2. admission: first fit in arrival order
    for (auto it = s->queue.begin(); it != s->queue.end();) {
        if (static_cast<int32_t>(s->active.size()) >= s->max_active) break;
        Request &r = *s->reqs[*it];
        if (s->in_use + r.footprint > s->budget && !s->parked.empty()) {
            size_t parked = 0;
            for (auto &q : s->parked) parked += q.second;
            if (s->in_use - parked + r.footprint <= s->budget) flush_parked(s);  // freeing them makes room
        }
        if (s->in_use + r.footprint > s->budget) {
            ++it;
            continue;
        }
        const int rc = sbe_create(&r.cfg, &r.model, -1, -1, &r.ctx);
        if (rc != SBE_OK) {
            r.err = sbe_last_error(nullptr);
            r.state = SBE_REQ_FAILED;
            r.status = rc;
            r.t_done = now_s();
            it = s->queue.erase(it);
            continue;
        }
        s->in_use += r.footprint;
        sbe_info inf;
        sbe_get_info(r.ctx, &inf);
        r.nt = inf.nt;
        r.nband = inf.nband;
        if (r.ctx->path == sbe::PathKind::NBand) {
            // every N-band step is a dozen small launches: short slices keep the hardware
            // queue this request shares with others from filling with thousands of them
            r.slice = std::max<int64_t>(1, s->slice / 32);
        } else if (r.ctx->adaptive) {
            // whole windows of the adaptive stepper, so the result equals a run on its own
            const int64_t w = r.ctx->chunk_steps;
            r.slice = std::max<int64_t>(1, (s->slice + w - 1) / w) * w;
        } else {
            const double by_work = kSliceKSteps / static_cast<double>(std::max<int64_t>(1, inf.nk_total));
            r.slice = std::max<int64_t>(s->slice, static_cast<int64_t>(by_work));
        }
        for (auto &e : r.ev)
            if (!r.status && hipEventCreateWithFlags(&e, hipEventDisableTiming) != hipSuccess)
                r.status = SBE_ERR_HIP, r.err = "hipEventCreate failed";
        r.state = SBE_REQ_RUNNING;
        r.t_admit = now_s();
        s->active.push_back(r.id);
        it = s->queue.erase(it);
    }
    // 3. keep kDepth slices in flight per admitted request
    for (int64_t id : s->active) {
        Request &r = *s->reqs[id];
        if (r.status) continue;
        for (int slot = 0; slot < kDepth && r.ctx->steps_done < r.nt; ++slot) {
            if (hipEventQuery(r.ev[slot]) == hipErrorNotReady) continue;
            const int64_t b = r.ctx->steps_done, e = std::min(r.nt, b + r.slice);
            int rc = sbe_run(r.ctx, b, e);
            if (rc != SBE_OK) {
                r.err = sbe_last_error(r.ctx), r.status = rc;
                break;
            }
            if (hipEventRecord(r.ev[slot], r.ctx->stream) != hipSuccess) {
                r.err = "hipEventRecord failed", r.status = SBE_ERR_HIP;
                break;
            }
        }
    }
    // failed requests leave the active set
    for (auto it = s->active.begin(); it != s->active.end();) {
        Request &r = *s->reqs[*it];
        if (r.status != SBE_OK) {
            finish(s, r, r.status, r.err);
            it = s->active.erase(it);
        } else {
            ++it;
        }
    }
    return SBE_OK;
}

int sbe_server_poll(const sbe_server *s, int64_t id, sbe_request_status *st)
{
    if (!s || !st) return SBE_ERR_CONFIG;
    auto it = s->reqs.find(id);
    if (it == s->reqs.end()) return SBE_ERR_RANGE;
    const Request &r = *it->second;
    st->state = r.state;
    st->status = r.status;
    st->nt = r.nt;
    st->nband = r.nband;
    st->footprint = static_cast<int64_t>(r.footprint);
    st->t_submit = r.t_submit;
    st->t_admit = r.t_admit;
    st->t_done = r.t_done;
    return SBE_OK;
}

int sbe_server_counts(const sbe_server *s, int32_t *queued, int32_t *active, size_t *in_use, size_t *budget)
{
    if (!s) return SBE_ERR_CONFIG;
    if (queued) *queued = static_cast<int32_t>(s->queue.size());
    if (active) *active = static_cast<int32_t>(s->active.size());
    if (in_use) *in_use = s->in_use;
    if (budget) *budget = s->budget;
    return SBE_OK;
}

int sbe_server_get_series(const sbe_server *s, int64_t id, double *j_inter, double *j_intra, double *rho_bar,
                          double *band_pop)
{
    if (!s) return SBE_ERR_CONFIG;
    auto it = s->reqs.find(id);
    if (it == s->reqs.end()) return SBE_ERR_RANGE;
    const Request &r = *it->second;
    if (r.state != SBE_REQ_DONE) return SBE_ERR_STATE;
    auto cp = [](double *dst, const std::vector<double> &v) {
        if (dst) std::memcpy(dst, v.data(), v.size() * sizeof(double));
    };
    cp(j_inter, r.j_inter);
    cp(j_intra, r.j_intra);
    cp(rho_bar, r.rho_bar);
    cp(band_pop, r.band_pop);
    return SBE_OK;
}

const char *sbe_server_request_error(const sbe_server *s, int64_t id)
{
    if (!s) return "";
    auto it = s->reqs.find(id);
    return it == s->reqs.end() ? "" : it->second->err.c_str();
}

int sbe_server_release(sbe_server *s, int64_t id)
{
    if (!s) return SBE_ERR_CONFIG;
    auto it = s->reqs.find(id);
    if (it == s->reqs.end()) return SBE_ERR_RANGE;
    if (it->second->state == SBE_REQ_RUNNING || it->second->state == SBE_REQ_QUEUED) return SBE_ERR_STATE;
    s->reqs.erase(it);
    return SBE_OK;
}

}  // extern "C"
