// server.cpp - continuous-batching request loop over SBE contexts.
//
// Requests (a config and a model each) queue on submit. Every scheduling
// round (sbe_server_step):
//   1. retire requests: finished ones (every step enqueued and completed:
//      series copied to host memory), cancelled ones and ones whose k sums
//      went non-finite (checked per completed slice, from a flag the series
//      reduction sets); their contexts are parked and freed in groups,
//      because freeing device memory waits for the whole device, and their
//      memory stays charged to the budget until then;
//   2. admit queued requests, first fit in arrival order, while their device
//      footprint (sbe_query_footprint) fits the remaining HBM budget and a
//      slot is free; admission creates the context, so its density matrices
//      live on the GPU from then on;
//   3. enqueue work. 2-band fixed-step requests of one tableau and precision
//      form a lane: each launch of the batch step kernel on the lane's stream
//      advances every request of the lane by its next slice (requests join
//      and leave the launches as they are admitted and retired), with one
//      reduction launch for all of them. Every other request (N-band, the
//      adaptive stepper, the velocity gauge) runs its slices on its own
//      stream. Up to kDepth launches stay in flight per lane or request, so
//      the device does not wait for the host.
// Results do not depend on the mix: a lane member is planned with one k
// point per lane like a batch member, its blocks and its k sums (fixed order
// per request) are the same in any launch, and fixed-step results do not
// depend on where a run is split into launches; the other requests have
// their own contexts and streams.
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

constexpr int kDepth = 2;                    // launches in flight per lane or request
constexpr double kSliceKSteps = 67108864.0;  // 2-band work per own-stream slice, about 1.5 ms of device time
constexpr int64_t kNbSliceKSteps = 65536;    // fused N-band work per slice (256 steps of 256 k points)
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
    int lane = -1;                 // shared lane key, or -1: own stream
    int64_t last_seq = 0;          // lane launch that carries this request's last enqueued slice
    hipEvent_t ev[kDepth] = {};    // own-stream requests: slices in flight
    int64_t slice = 0;
    int64_t nt = 0;
    int nband = 0;
    bool cancel = false;
    double t_submit = 0, t_admit = 0, t_done = 0;
    std::vector<double> j_inter, j_intra, rho_bar, band_pop;
    std::string err;
    int status = SBE_OK;
};

// Requests that share launches: one stream, a ring of launch events
struct Lane {
    hipStream_t stream = nullptr;
    hipEvent_t ev[kDepth] = {};
    int64_t ev_seq[kDepth] = {};  // launch number recorded by each event (0: none)
    int64_t next_seq = 1;
    int64_t done_seq = 0;         // every launch up to this one has completed
    sbe::SharedLaunch buf;
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
    std::map<int, Lane> lanes;
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

// advance done_seq over the launches whose events have completed (in order: one stream)
void poll_lane(Lane &L)
{
    for (;;) {
        int next = -1;
        for (int i = 0; i < kDepth; ++i)
            if (L.ev_seq[i] == L.done_seq + 1) next = i;
        if (next < 0 || hipEventQuery(L.ev[next]) != hipSuccess) return;
        ++L.done_seq;
    }
}

// no enqueued work of r is pending on the device
bool quiet(sbe_server *s, const Request &r)
{
    if (r.lane < 0) return in_flight(r) == 0;
    Lane &L = s->lanes[r.lane];
    poll_lane(L);
    return L.done_seq >= r.last_seq;
}

// wait until no enqueued work of r is pending
void drain(sbe_server *s, const Request &r)
{
    if (r.lane < 0) {
        (void)sbe_synchronize(r.ctx);
        return;
    }
    Lane &L = s->lanes[r.lane];
    for (int i = 0; i < kDepth; ++i)
        if (L.ev_seq[i] >= r.last_seq && L.ev_seq[i] > L.done_seq) (void)hipEventSynchronize(L.ev[i]);
    poll_lane(L);
}

void finish(sbe_server *s, Request &r, int state, int status, const std::string &msg)
{
    if (r.ctx) {
        drain(s, r);
        if (state == SBE_REQ_DONE) {
            const int N = r.nband;
            r.j_inter.assign(3 * r.nt, 0.0);
            r.j_intra.assign(3 * r.nt, 0.0);
            r.rho_bar.assign(static_cast<size_t>(2) * N * N * r.nt, 0.0);
            r.band_pop.assign(static_cast<size_t>(N) * r.nt, 0.0);
            status = sbe_get_series(r.ctx, r.j_inter.data(), r.j_intra.data(), r.rho_bar.data(), r.band_pop.data());
            if (status != SBE_OK) {
                r.err = sbe_last_error(r.ctx);
                state = SBE_REQ_FAILED;
            }
        }
        drop_events(r);
        s->parked.emplace_back(r.ctx, r.footprint);
        r.ctx = nullptr;
        if (s->parked.size() >= kParkMax) flush_parked(s);
    }
    r.status = status;
    if (status != SBE_OK && r.err.empty()) r.err = msg;
    r.state = state;
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
    for (auto &kv : s->lanes)
        if (kv.second.stream) (void)hipStreamSynchronize(kv.second.stream);
    for (auto &kv : s->reqs)
        if (kv.second->ctx) {
            drop_events(*kv.second);
            sbe_destroy(kv.second->ctx);
        }
    flush_parked(s);
    for (auto &kv : s->lanes) {
        Lane &L = kv.second;
        for (auto &e : L.ev)
            if (e) (void)hipEventDestroy(e);
        sbe::free_shared(L.buf);
        if (L.stream) (void)hipStreamDestroy(L.stream);
    }
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

int sbe_server_cancel(sbe_server *s, int64_t id)
{
    if (!s) return SBE_ERR_CONFIG;
    auto it = s->reqs.find(id);
    if (it == s->reqs.end()) return srv_fail(s, SBE_ERR_RANGE, "unknown request");
    Request &r = *it->second;
    if (r.state == SBE_REQ_QUEUED) {  // never admitted: nothing on the device
        s->queue.erase(std::find(s->queue.begin(), s->queue.end(), id));
        r.state = SBE_REQ_CANCELLED;
        r.status = SBE_OK;
        r.t_done = now_s();
        return SBE_OK;
    }
    if (r.state == SBE_REQ_RUNNING) {
        r.cancel = true;  // retired by the next round, once its enqueued work has drained
        return SBE_OK;
    }
    return srv_fail(s, SBE_ERR_STATE, "request already finished");
}

int sbe_server_step(sbe_server *s)
{
    if (!s) return SBE_ERR_STATE;
    ++s->rounds;
    // 1. retire: cancelled, non-finite, or every step enqueued and completed
    for (auto it = s->active.begin(); it != s->active.end();) {
        Request &r = *s->reqs[*it];
        bool out = true;
        if (r.cancel) {
            finish(s, r, SBE_REQ_CANCELLED, SBE_OK, "");
        } else if (*r.ctx->nonfinite_h) {
            finish(s, r, SBE_REQ_FAILED, SBE_ERR_NAN,
                   "non-finite k sum within the first " + std::to_string(r.ctx->steps_done) + " steps");
        } else if (r.status != SBE_OK) {
            finish(s, r, SBE_REQ_FAILED, r.status, r.err);
        } else if (r.ctx->steps_done >= r.nt && quiet(s, r)) {
            finish(s, r, SBE_REQ_DONE, SBE_OK, "");
        } else {
            out = false;
        }
        it = out ? s->active.erase(it) : it + 1;
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
        // 2-band fixed-step LengthWannier requests are planned as batch members, to share launches
        const bool lane_kind = r.model.kind == SBE_MODEL_HALDANE && r.cfg.gauge == SBE_GAUGE_LENGTH_WANNIER &&
                               (r.cfg.rk_order == SBE_RK4 || r.cfg.rk_order == SBE_RK5_FEHLBERG);
        const int rc = lane_kind ? sbe::create_batch_member(&r.cfg, &r.model, &r.ctx)
                                 : sbe_create(&r.cfg, &r.model, -1, -1, &r.ctx);
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
        if (sbe::shared_eligible(r.ctx)) {
            r.lane = sbe::shared_key(r.ctx);
            Lane &L = s->lanes[r.lane];
            if (!L.stream && hipStreamCreateWithFlags(&L.stream, hipStreamNonBlocking) != hipSuccess)
                r.status = SBE_ERR_HIP, r.err = "hipStreamCreate failed";
            for (auto &e : L.ev)
                if (!e && !r.status && hipEventCreateWithFlags(&e, hipEventDisableTiming) != hipSuccess)
                    r.status = SBE_ERR_HIP, r.err = "hipEventCreate failed";
            r.slice = std::min(s->slice, r.ctx->chunk_steps);
        } else {
            if (r.ctx->path == sbe::PathKind::NBand && !(r.ctx->nbf || r.ctx->nb32 || r.ctx->nbv)) {
                // every N-band step on the separate kernels is a dozen small launches: short
                // slices keep the hardware queue this request shares with others from filling
                // with thousands of them (the fused steps take one or two launches per step)
                r.slice = std::max<int64_t>(1, s->slice / 32);
            } else if (r.ctx->path == sbe::PathKind::NBand) {
                // fused steps: whole slices on the serving-size grids, fewer steps per slice on
                // large ones (about kNbSliceKSteps k-point steps, a few ms of device time)
                const int64_t by_work = kNbSliceKSteps / std::max<int64_t>(1, inf.nk_total);
                r.slice = std::max<int64_t>(1, std::min<int64_t>(s->slice, by_work));
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
        }
        r.state = SBE_REQ_RUNNING;
        r.t_admit = now_s();
        s->active.push_back(r.id);
        it = s->queue.erase(it);
    }
    // 3a. own-stream requests: keep kDepth slices in flight each
    for (int64_t id : s->active) {
        Request &r = *s->reqs[id];
        if (r.status || r.lane >= 0) continue;
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
    // 3b. lanes: one shared launch per free event slot, over every member with steps left
    for (auto &kv : s->lanes) {
        Lane &L = kv.second;
        poll_lane(L);
        for (int slot = 0; slot < kDepth; ++slot) {
            if (L.ev_seq[slot] > L.done_seq) continue;  // still in flight
            std::vector<sbe_ctx *> cs;
            std::vector<int64_t> ns;
            std::vector<Request *> rs;
            for (int64_t id : s->active) {
                Request &r = *s->reqs[id];
                if (r.lane != kv.first || r.status || r.cancel || r.ctx->steps_done >= r.nt) continue;
                cs.push_back(r.ctx);
                ns.push_back(std::min(r.slice, r.nt - r.ctx->steps_done));
                rs.push_back(&r);
            }
            if (cs.empty()) break;
            std::string err;
            const int rc = sbe::launch_shared(cs.data(), ns.data(), static_cast<int>(cs.size()), L.buf, L.stream, err);
            const int64_t seq = L.next_seq++;
            const bool rec = rc == SBE_OK && hipEventRecord(L.ev[slot], L.stream) == hipSuccess;
            if (rec) L.ev_seq[slot] = seq;
            for (Request *r : rs) {
                r->last_seq = seq;
                if (!rec) r->status = rc ? rc : SBE_ERR_HIP, r->err = rc ? err : "hipEventRecord failed";
            }
            if (!rec) {  // the launch is lost for its members: wait out the lane, then account it as done
                (void)hipStreamSynchronize(L.stream);
                L.done_seq = seq;
                break;
            }
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
