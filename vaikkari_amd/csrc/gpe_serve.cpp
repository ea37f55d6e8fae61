// gpe_serve.cpp -- continuous-batching solve server over the batch engine
// (include/vaikkari_gpe_serve.h).  Host-side scheduling only; all device
// work goes through gpe_batch_*.
#include <algorithm>
#include <chrono>
#include <cmath>
#include <deque>
#include <new>
#include <unordered_map>
#include <vector>

#include "gpe_internal.hpp"
#include "../../include/vaikkari_gpe_batch.h"
#include "../../include/vaikkari_gpe_serve.h"

namespace {

struct Pending {
    uint64_t ticket;
    gpe_request req;
    double arrival;
};

struct Running {
    uint64_t ticket;
    gpe_request req;
    double arrival, admitted;
    int32_t id;
};

struct Finished {
    int32_t id;  // batch instance kept resident for gpe_server_fetch
    gpe_request req;
};

}  // namespace

struct gpe_server {
    gpe_batch *batch = nullptr;
    uint32_t chunk = 50;
    uint32_t policy = GPE_SERVE_FIFO;
    uint64_t max_active_nodes = 0;  // 0: no cap
    uint64_t active_nodes = 0;
    uint64_t budget = 0;  // HBM budget of the batch (0: none)
    double timeout = 0.0;  // seconds from arrival to result (0: none)
    std::chrono::steady_clock::time_point t0;
    uint64_t next_ticket = 1;
    std::deque<Pending> queue;
    std::vector<Running> running;
    std::vector<gpe_completion> done;  // not yet polled
    std::unordered_map<uint64_t, Finished> kept;
    gpe_server_stats st{};
    ~gpe_server() { gpe_batch_destroy(batch); }
};

static gpe_instance_config to_instance(const gpe_request &r) {
    gpe_instance_config c{};
    c.nx = c.ny = c.nz = r.n;
    c.mode = r.mode;
    c.block_scale = r.h / std::sqrt(8.0);
    c.dt = r.dt;
    c.g = r.g;
    c.trap_ratio = r.trap_ratio;
    c.origin[0] = c.origin[1] = c.origin[2] = NAN;  // centred box
    c.steps = r.steps;
    c.init = GPE_INIT_GAUSSIAN;
    c.zlayers = 0;
    c.sigma = r.sigma;
    c.noise = r.noise;
    c.mu0 = r.mu0;
    c.seed = r.seed;
    return c;
}

extern "C" {

gpe_status gpe_server_create(uint32_t precision, uint64_t budget, uint32_t chunk_steps, gpe_server **out) {
    if (!out || chunk_steps == 0) {
        gpe::set_error("gpe_server_create: invalid argument");
        return GPE_EINVAL;
    }
    *out = nullptr;
    auto *s = new (std::nothrow) gpe_server();
    if (!s) return GPE_ENOMEM;
    const gpe_status st = gpe_batch_create(precision, budget, &s->batch);
    if (st != GPE_OK) {
        delete s;
        return st;
    }
    s->chunk = chunk_steps;
    s->budget = budget;
    s->t0 = std::chrono::steady_clock::now();
    *out = s;
    return GPE_OK;
}

void gpe_server_destroy(gpe_server *s) { delete s; }

double gpe_server_now(const gpe_server *s) {
    if (!s) return 0.0;
    return std::chrono::duration<double>(std::chrono::steady_clock::now() - s->t0).count();
}

gpe_status gpe_server_submit(gpe_server *s, const gpe_request *r, double arrival, uint64_t *ticket) {
    if (!s || !r || !ticket || r->n == 0 || r->n > 4096 || r->steps == 0 || !(r->h > 0) || !(r->dt > 0) ||
        r->mode > GPE_IMAG_TIME) {
        gpe::set_error("gpe_server_submit: invalid request");
        return GPE_EINVAL;
    }
    const gpe_instance_config c = to_instance(*r);
    const uint64_t bytes = gpe_batch_instance_bytes(s->batch, &c);
    if (bytes == 0) {
        gpe::set_error("gpe_server_submit: invalid request");
        return GPE_EINVAL;
    }
    if (s->budget && bytes > s->budget) {
        // could never be admitted: it would wait in the queue for ever
        gpe::set_error("gpe_server_submit: request needs %llu bytes, more than the whole budget of %llu",
                       (unsigned long long)bytes, (unsigned long long)s->budget);
        return GPE_EINVAL;
    }
    {
        // the batch's own admission test, made here so that an unstable
        // request never enters the queue
        const uint32_t dims[3] = {r->n, r->n, r->n};
        const double a = std::sqrt(8.0) * c.block_scale;
        const double origin[3] = {-0.5 * a * r->n, -0.5 * a * r->n, -0.5 * a * r->n};
        const double vmax = gpe::trap_vmax(dims, origin, c.block_scale, r->trap_ratio);
        if (!gpe::linear_step_stable(r->dt, c.block_scale, vmax)) {
            gpe::set_error("gpe_server_submit: dt = %g violates dt*(6/b^2 + max V) < 2 (max V = %g)", r->dt, vmax);
            return GPE_EUNSTABLE;
        }
    }
    *ticket = s->next_ticket++;
    s->queue.push_back({*ticket, *r, arrival});
    s->st.submitted++;
    return GPE_OK;
}

static uint64_t request_nodes(const gpe_request &r) { return 12ull * r.n * r.n * r.n; }
static double request_work(const gpe_request &r) { return (double)request_nodes(r) * (double)r.steps; }

// A completion without a result (dropped request): node_steps 0, NaN observables.
static void post_dropped(gpe_server *s, uint64_t ticket, double arrival, double admitted, gpe_status why) {
    gpe_completion cpl{};
    cpl.ticket = ticket;
    cpl.arrival = arrival;
    cpl.admitted = admitted;
    cpl.finished = gpe_server_now(s);
    cpl.node_steps = 0.0;
    cpl.norm = cpl.energy = NAN;
    cpl.status = why;
    s->done.push_back(cpl);
}

// Removes running[q]'s solve from the batch and the running list (the order
// of running does not matter).
static gpe_status drop_running(gpe_server *s, size_t q) {
    Running r = s->running[q];
    const gpe_status e = gpe_batch_remove(s->batch, r.id);
    if (e != GPE_OK) return e;
    s->active_nodes -= request_nodes(r.req);
    s->running[q] = s->running.back();
    s->running.pop_back();
    return GPE_OK;
}

// Drops every queued or resident request older than the timeout; returns the
// number dropped in *n.
static gpe_status expire(gpe_server *s, uint32_t *n) {
    *n = 0;
    if (!(s->timeout > 0.0)) return GPE_OK;
    const double now = gpe_server_now(s);
    std::deque<Pending> rest;
    for (const Pending &p : s->queue) {
        if (now - p.arrival > s->timeout) {
            post_dropped(s, p.ticket, p.arrival, now, GPE_ETIMEDOUT);
            s->st.timed_out++;
            (*n)++;
        } else {
            rest.push_back(p);
        }
    }
    s->queue.swap(rest);
    for (size_t q = 0; q < s->running.size();) {
        const Running r = s->running[q];
        if (now - r.arrival > s->timeout) {
            const gpe_status e = drop_running(s, q);
            if (e != GPE_OK) return e;
            post_dropped(s, r.ticket, r.arrival, r.admitted, GPE_ETIMEDOUT);
            s->st.timed_out++;
            (*n)++;
        } else {
            q++;
        }
    }
    return GPE_OK;
}

gpe_status gpe_server_tick(gpe_server *s, uint32_t *completed) {
    if (!s) return GPE_EINVAL;
    uint32_t nexpired = 0;
    {
        const gpe_status e = expire(s, &nexpired);
        if (e != GPE_OK) return e;
    }
    // Admission: candidates in policy order, first fit within the HBM budget
    // and the node cap.
    std::vector<size_t> order(s->queue.size());
    for (size_t q = 0; q < order.size(); q++) order[q] = q;
    if (s->policy == GPE_SERVE_SHORTEST_FIRST)
        std::stable_sort(order.begin(), order.end(), [s](size_t a, size_t b) {
            return request_work(s->queue[a].req) < request_work(s->queue[b].req);
        });
    std::vector<char> taken(s->queue.size(), 0);
    uint32_t ndone = nexpired;
    gpe_status device_error = GPE_OK;
    for (size_t q : order) {
        const Pending &p = s->queue[q];
        const uint64_t nodes = request_nodes(p.req);
        if (s->max_active_nodes && s->active_nodes > 0 && s->active_nodes + nodes > s->max_active_nodes) continue;
        const gpe_instance_config c = to_instance(p.req);
        int32_t id = -1;
        const gpe_status st = gpe_batch_add(s->batch, &c, &id);
        if (st == GPE_EBUSY) continue;
        // out of device memory while other solves hold it: wait for them
        if (st == GPE_ENOMEM && !s->running.empty()) continue;
        if (st != GPE_OK) {
            // the request cannot run: drop it with a failed completion
            const double now = gpe_server_now(s);
            gpe_completion cpl{};
            cpl.ticket = p.ticket;
            cpl.arrival = p.arrival;
            cpl.admitted = now;
            cpl.finished = now;
            cpl.node_steps = 0.0;
            cpl.norm = cpl.energy = NAN;
            cpl.status = st;
            s->done.push_back(cpl);
            s->st.failed++;
            taken[q] = 1;
            ndone++;
            if (st == GPE_EDEVICE) {
                device_error = st;
                break;
            }
            continue;
        }
        s->running.push_back({p.ticket, p.req, p.arrival, gpe_server_now(s), id});
        s->active_nodes += nodes;
        s->st.admitted++;
        taken[q] = 1;
    }
    if (!s->queue.empty()) {
        std::deque<Pending> rest;
        for (size_t q = 0; q < s->queue.size(); q++)
            if (!taken[q]) rest.push_back(s->queue[q]);
        s->queue.swap(rest);
    }
    if (device_error != GPE_OK) {
        if (completed) *completed = ndone;
        return device_error;
    }
    if (!s->running.empty()) {
        const uint64_t ns = gpe_batch_chunk_node_steps(s->batch, s->chunk);
        double ms = 0.0;
        const gpe_status st = gpe_batch_run(s->batch, s->chunk, 1, &ms);
        if (st != GPE_OK) return st;
        s->st.node_steps_done += (double)ns;
        s->st.device_ms += ms;
        s->st.ticks += 1;
        const double now = gpe_server_now(s);
        for (size_t q = 0; q < s->running.size();) {
            Running &r = s->running[q];
            gpe_instance_info inf;
            gpe_status e = gpe_batch_instance(s->batch, r.id, &inf);
            if (e != GPE_OK) return e;
            if (inf.steps_left > 0) {
                q++;
                continue;
            }
            double obs[4] = {0, 0, 0, 0};
            e = gpe_batch_observables(s->batch, r.id, obs);
            if (e != GPE_OK) return e;
            gpe_completion cpl;
            cpl.ticket = r.ticket;
            cpl.arrival = r.arrival;
            cpl.admitted = r.admitted;
            cpl.finished = now;
            cpl.node_steps = 12.0 * r.req.n * r.req.n * (double)r.req.n * (double)r.req.steps;
            cpl.norm = obs[0];
            cpl.energy = obs[1] + obs[2] + 0.5 * obs[3];
            cpl.status = GPE_OK;
            cpl.reserved = 0;
            s->done.push_back(cpl);
            s->active_nodes -= request_nodes(r.req);
            if (r.req.keep_result) {
                s->kept[r.ticket] = {r.id, r.req};
            } else {
                e = gpe_batch_remove(s->batch, r.id);
                if (e != GPE_OK) return e;
            }
            s->st.completed++;
            ndone++;
            r = s->running.back();
            s->running.pop_back();
        }
    }
    if (completed) *completed = ndone;
    return GPE_OK;
}

gpe_status gpe_server_poll(gpe_server *s, gpe_completion *out, uint32_t max, uint32_t *n) {
    if (!s || !n || (max && !out)) return GPE_EINVAL;
    const uint32_t k = (uint32_t)std::min<size_t>(max, s->done.size());
    for (uint32_t i = 0; i < k; i++) out[i] = s->done[i];
    s->done.erase(s->done.begin(), s->done.begin() + k);
    *n = k;
    return GPE_OK;
}

gpe_status gpe_server_fetch(gpe_server *s, uint64_t ticket, double *even, double *odd) {
    if (!s) return GPE_EINVAL;
    auto it = s->kept.find(ticket);
    if (it == s->kept.end()) {
        gpe::set_error("gpe_server_fetch: ticket %llu has no kept result", (unsigned long long)ticket);
        return GPE_EINVAL;
    }
    gpe_status st = gpe_batch_get_state(s->batch, it->second.id, even, odd);
    if (st != GPE_OK) return st;
    st = gpe_batch_remove(s->batch, it->second.id);
    s->kept.erase(it);
    return st;
}

gpe_status gpe_server_cancel(gpe_server *s, uint64_t ticket) {
    if (!s) {
        gpe::set_error("gpe_server_cancel: null server");
        return GPE_EINVAL;
    }
    for (auto it = s->queue.begin(); it != s->queue.end(); ++it)
        if (it->ticket == ticket) {
            const Pending p = *it;
            s->queue.erase(it);
            post_dropped(s, p.ticket, p.arrival, gpe_server_now(s), GPE_ECANCELED);
            s->st.cancelled++;
            return GPE_OK;
        }
    for (size_t q = 0; q < s->running.size(); q++)
        if (s->running[q].ticket == ticket) {
            const Running r = s->running[q];
            const gpe_status e = drop_running(s, q);
            if (e != GPE_OK) return e;
            post_dropped(s, r.ticket, r.arrival, r.admitted, GPE_ECANCELED);
            s->st.cancelled++;
            return GPE_OK;
        }
    gpe::set_error("gpe_server_cancel: ticket %llu is neither queued nor running", (unsigned long long)ticket);
    return GPE_EINVAL;
}

gpe_status gpe_server_set_timeout(gpe_server *s, double seconds) {
    if (!s || !(seconds >= 0.0) || !std::isfinite(seconds)) {
        gpe::set_error("gpe_server_set_timeout: the timeout must be a finite number of seconds >= 0");
        return GPE_EINVAL;
    }
    s->timeout = seconds;
    return GPE_OK;
}

gpe_status gpe_server_set_policy(gpe_server *s, uint32_t policy, uint64_t max_active_nodes) {
    if (!s || policy > GPE_SERVE_SHORTEST_FIRST) {
        gpe::set_error("gpe_server_set_policy: unknown policy");
        return GPE_EINVAL;
    }
    s->policy = policy;
    s->max_active_nodes = max_active_nodes;
    return GPE_OK;
}

double gpe_server_oldest_queued(const gpe_server *s) {
    if (!s || s->queue.empty()) return -1.0;
    double t = s->queue.front().arrival;
    for (const Pending &p : s->queue) t = std::min(t, p.arrival);
    return t;
}

gpe_status gpe_server_set_launch_mode(gpe_server *s, uint32_t mode) {
    if (!s) {
        gpe::set_error("gpe_server_set_launch_mode: null server");
        return GPE_EINVAL;
    }
    return gpe_batch_set_launch_mode(s->batch, mode);
}

gpe_status gpe_server_set_tile(gpe_server *s, uint32_t tile) {
    if (!s) {
        gpe::set_error("gpe_server_set_tile: null server");
        return GPE_EINVAL;
    }
    return gpe_batch_set_tile(s->batch, tile);
}

gpe_status gpe_server_get_stats(const gpe_server *s, gpe_server_stats *out) {
    if (!s || !out) return GPE_EINVAL;
    *out = s->st;
    out->queued = s->queue.size();
    out->active = s->running.size();
    out->resident_bytes = gpe_batch_resident_bytes(s->batch);
    return GPE_OK;
}

}  // extern "C"
