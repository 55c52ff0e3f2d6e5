/*
 * sbe_server.h - continuous-batching request loop over SBE contexts (no
 * reference counterpart: the reference runs one solve per MPI job).
 *
 * Submitted solves queue; each scheduling round admits queued requests
 * first fit by device footprint against an HBM budget and a slot cap
 * (their density matrices then stay on the GPU until they finish),
 * enqueues the next slice of steps of every admitted request, and retires
 * requests (series copied to host, device memory freed). 2-band fixed-step
 * requests of one tableau and precision share launches: every launch of
 * the batch step kernel advances all of them, and requests join and leave
 * between launches. Other requests run their slices on their own streams,
 * overlapping on the device. A request whose k sums become non-finite fails
 * at the next round after the slice that produced them (SBE_ERR_NAN) and
 * the others go on; a request can be cancelled. A result is the same alone
 * and in any mix (bit for bit). Single-threaded: call from one host thread.
 */
#ifndef ANTELOPE_SBE_SERVER_H
#define ANTELOPE_SBE_SERVER_H

#include "sbe_c.h"

#ifdef __cplusplus
extern "C" {
#endif

enum { SBE_REQ_QUEUED = 0, SBE_REQ_RUNNING = 1, SBE_REQ_DONE = 2, SBE_REQ_FAILED = 3, SBE_REQ_CANCELLED = 4 };

typedef struct sbe_request_status {
    int32_t state;       /* SBE_REQ_* */
    int32_t status;      /* SBE_OK or the failure code */
    int64_t nt;
    int32_t nband;
    int32_t reserved;
    int64_t footprint;   /* device bytes the request holds while admitted */
    double t_submit, t_admit, t_done;  /* host monotonic clock, seconds */
} sbe_request_status;

typedef struct sbe_server sbe_server;

/* hbm_budget_bytes = 0: 80 % of the device memory free at creation.
 * slice_steps: steps enqueued per request and round. */
int sbe_server_create(size_t hbm_budget_bytes, int32_t max_active, int64_t slice_steps, sbe_server **out);
/* Copies the inputs (including Wannier tables); fails with SBE_ERR_OOM if the
 * request could never fit the budget. */
int sbe_server_submit(sbe_server *s, const sbe_config *cfg, const sbe_model *model, int64_t *id);
/* One scheduling round (retire, admit, enqueue). Non-blocking apart from the
 * host-side launch work. */
int sbe_server_step(sbe_server *s);
int sbe_server_poll(const sbe_server *s, int64_t id, sbe_request_status *st);
int sbe_server_counts(const sbe_server *s, int32_t *queued, int32_t *active, size_t *in_use, size_t *budget);
/* Series of a finished request, laid out as sbe_get_series. */
int sbe_server_get_series(const sbe_server *s, int64_t id, double *j_inter, double *j_intra, double *rho_bar,
                          double *band_pop);
const char *sbe_server_request_error(const sbe_server *s, int64_t id);
/* Cancel a queued or running request: a queued one is dropped at once, a
 * running one leaves at the next round, after its enqueued work drains, and
 * frees its device memory. Its state becomes SBE_REQ_CANCELLED; the other
 * requests are unaffected. SBE_ERR_STATE if it has already finished. */
int sbe_server_cancel(sbe_server *s, int64_t id);
int sbe_server_release(sbe_server *s, int64_t id);
const char *sbe_server_last_error(const sbe_server *s);
void sbe_server_destroy(sbe_server *s);

#ifdef __cplusplus
}
#endif

#endif /* ANTELOPE_SBE_SERVER_H */
