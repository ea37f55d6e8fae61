/*
 * vaikkari_gpe_serve.h -- continuous-batching solve server.
 *
 * No counterpart in the reference (one solve per process, its state never
 * leaves the GPU until the run ends).  Requests enter a queue; every tick the
 * server admits queued requests (policy below) while their resident bytes
 * fit the HBM budget, advances every resident solve by one chunk of steps in
 * shared persistent launches (vaikkari_gpe_batch.h), and retires finished
 * solves, recording time-to-result.  A solve's result does not depend on
 * what else is resident.
 *
 * Admission policy (gpe_server_set_policy):
 *   GPE_SERVE_FIFO            queued requests in arrival order, first fit (a
 *                             request that does not fit does not block smaller
 *                             ones behind it); everything that fits runs at once.
 *   GPE_SERVE_SHORTEST_FIRST  queued requests in order of their work
 *                             (12 n^3 steps, ties by arrival), first fit.
 * With max_active_nodes > 0 a request is admitted only while the resident
 * solves hold fewer than that many nodes in total (always when none is
 * resident).  Every resident solve advances each tick, so an uncapped FIFO
 * server under overload shares the GPU among all of them and finishes few;
 * a cap near the node count that saturates the GPU keeps per-solve progress
 * high.
 *
 * Times are seconds on the server's clock (gpe_server_now).
 */
#ifndef VAIKKARI_GPE_SERVE_H
#define VAIKKARI_GPE_SERVE_H

#include "vaikkari_gpe.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct gpe_server gpe_server;

typedef struct gpe_request {
    uint32_t n;          /* n^3 cubic cells */
    uint32_t mode;       /* gpe_mode: ground state (imaginary time) or real-time evolution */
    uint64_t steps;
    double h;            /* cube edge; block_scale = h/sqrt(8) */
    double dt;           /* time step (dtau for imaginary time) */
    double g;
    double trap_ratio;   /* V = 0.5(x^2 + y^2 + ratio^2 z^2) */
    double sigma, noise, mu0; /* Gaussian initial state, see gpe_init_kind */
    uint64_t seed;
    uint32_t keep_result;     /* 1: keep the final state resident until gpe_server_fetch */
    uint32_t reserved;
} gpe_request;

typedef struct gpe_completion {
    uint64_t ticket;
    double arrival, admitted, finished; /* server clock, seconds */
    double node_steps;                  /* 12 n^3 * steps */
    double norm, energy;                /* FP64 observables of the final state */
    int32_t status;                     /* GPE_OK, or why the request was dropped: at admission,
                                           cancelled (GPE_ECANCELED) or timed out (GPE_ETIMEDOUT);
                                           then node_steps is 0 and norm, energy are NaN */
    uint32_t reserved;
} gpe_completion;

typedef struct gpe_server_stats {
    uint64_t submitted, admitted, completed, queued, active;
    uint64_t resident_bytes;
    double node_steps_done;  /* over all ticks */
    double device_ms;        /* device time of all chunks */
    double ticks;
    uint64_t failed;         /* requests dropped at admission (a failed completion each) */
    uint64_t cancelled;      /* requests cancelled by gpe_server_cancel */
    uint64_t timed_out;      /* requests dropped by the server's timeout */
} gpe_server_stats;

gpe_status gpe_server_create(uint32_t precision, uint64_t hbm_budget_bytes, uint32_t chunk_steps,
                             gpe_server **out);
void gpe_server_destroy(gpe_server *s);
double gpe_server_now(const gpe_server *s);

/* Queues a request that arrived at `arrival` (server clock).  GPE_EINVAL for a
 * malformed request or one larger than the whole HBM budget, GPE_EUNSTABLE
 * when dt*(6/b^2 + max V) >= 2 on its box. */
gpe_status gpe_server_submit(gpe_server *s, const gpe_request *req, double arrival, uint64_t *ticket);
/* Admit, run one chunk, retire.  *completed = solves finished in this tick,
 * failed admissions included.  A request whose admission fails for a reason of
 * its own (GPE_EINVAL, GPE_EUNSTABLE, or GPE_ENOMEM with nothing else resident)
 * is dropped from the queue with a completion whose status says why; the tick
 * goes on with the others.  A device error ends the tick with that status after
 * the queue has been brought up to date. */
gpe_status gpe_server_tick(gpe_server *s, uint32_t *completed);
/* Completions since the last call (at most max). */
gpe_status gpe_server_poll(gpe_server *s, gpe_completion *out, uint32_t max, uint32_t *n);
/* Final state of a finished keep_result solve (odd may be NULL); frees it. */
gpe_status gpe_server_fetch(gpe_server *s, uint64_t ticket, double *even, double *odd);
/* Withdraws a queued or resident request: it leaves the queue, or its solve
 * leaves the shared launches and frees its memory at once, and a completion
 * with status GPE_ECANCELED is posted.  GPE_EINVAL for a ticket that is
 * neither queued nor running (finished, dropped, or unknown).  No counterpart
 * in the reference. */
gpe_status gpe_server_cancel(gpe_server *s, uint64_t ticket);
/* Time limit in seconds on the server clock from a request's arrival to its
 * result (0: none, the default).  At the start of every tick each queued or
 * resident request older than that is dropped with a GPE_ETIMEDOUT
 * completion; a resident one frees its memory.  No counterpart in the
 * reference. */
gpe_status gpe_server_set_timeout(gpe_server *s, double seconds);
gpe_status gpe_server_get_stats(const gpe_server *s, gpe_server_stats *st);

typedef enum gpe_serve_policy {
    GPE_SERVE_FIFO = 0,
    GPE_SERVE_SHORTEST_FIRST = 1
} gpe_serve_policy;
/* Admission policy and cap (see above); the default is GPE_SERVE_FIFO, no cap. */
gpe_status gpe_server_set_policy(gpe_server *s, uint32_t policy, uint64_t max_active_nodes);
/* Arrival time of the oldest queued request (server clock), or -1 when the queue is empty. */
double gpe_server_oldest_queued(const gpe_server *s);
/* How the shared chunks are launched: GPE_BATCH_GRAPH (default, one captured
 * hipGraph per chunk) or GPE_BATCH_STREAM (the same kernels enqueued one by
 * one), as gpe_batch_set_launch_mode.  Results are identical either way. */
gpe_status gpe_server_set_launch_mode(gpe_server *s, uint32_t mode);
/* Real-time tile of the shared launches, as gpe_batch_set_tile (waves |
 * GPE_BATCH_TILE_QUAD | GPE_BATCH_TILE_CACHED | GPE_BATCH_TILE_SLABS or
 * GPE_BATCH_TILE_GROUP_SLABS; the batch's default unless set);
 * only before the first request is admitted (GPE_EBUSY after).  Results do
 * not depend on it.  No counterpart in the reference. */
gpe_status gpe_server_set_tile(gpe_server *s, uint32_t tile);

#ifdef __cplusplus
}
#endif
#endif /* VAIKKARI_GPE_SERVE_H */
