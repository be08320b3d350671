// Serving: continuous batching of independent simulations on one GPU.
//
// Clients submit systems (particles, parameters, a step count) from any
// thread.  A worker thread owns the arenas the HBM budget buys: gsim_batch
// arenas of at most kBatchMaxParticles each (32-bit node indices inside an
// arena), each with its own stream, so the arenas' steps also overlap on the
// device.  Each iteration it admits queued requests first-come first-served
// into the first arena with room (while the optional active-particle cap
// allows), enqueues the same number of batched steps on every arena that
// holds systems, and retires the systems whose step budget is reached: their
// state is read back, they leave their arena, and the space admits the next
// requests.  Requests that arrive while others run join at the
// next step, so their work shares every kernel launch with the running ones.
// A system's step budget is enforced on the device, so it stops at exactly its
// step count whatever the host timing; its result is bit-identical to running
// it alone.
#include <algorithm>
#include <chrono>
#include <condition_variable>
#include <cstring>
#include <deque>
#include <map>
#include <memory>
#include <mutex>
#include <new>
#include <thread>
#include <vector>

#include "batch.hpp"
#include "common.hpp"
#include "gsim.h"

using gsim::fail;

namespace {

using Clock = std::chrono::steady_clock;

struct Request {
    uint64_t id = 0;
    gsim_params params{};
    size_t n = 0;
    uint32_t steps = 0;
    std::vector<float> pos, vel, mass;  // input, then the result
    Clock::time_point submitted, admitted, done;
    uint32_t arena = 0, batch_id = 0;
    int state = 0;  // 0 queued, 1 running, 2 done, 3 failed
    uint64_t interactions = 0;  // direct sum: N_pad^2 per step (Barnes-Hut counted per batch)
};

}  // namespace

struct gsim_server {
    std::vector<gsim_batch *> arenas;
    size_t capacity = 0, largest_arena = 0, max_active = 0, active_particles = 0;
    std::thread worker;
    std::mutex mu;
    std::condition_variable cv_work, cv_done;
    bool stop = false;
    uint64_t next_id = 1;
    std::deque<Request *> queue;
    std::map<uint64_t, std::unique_ptr<Request>> all;
    std::map<uint64_t, Request *> running;  // (arena << 32 | batch id) -> request
    std::vector<double> latencies_ms;
    uint64_t completed = 0, failed = 0, batch_steps = 0, ds_interactions = 0;
    Clock::time_point first_submit, last_done;
    bool any_submit = false;
    char worker_error[256] = {0};
};

namespace {

// Largest number of batched steps enqueued between two looks at the queue.
constexpr uint64_t kChunkSteps = 8;

size_t padded(size_t n) { return (n + 63) / 64 * 64; }

void worker_loop(gsim_server *s) {
    std::unique_lock<std::mutex> lk(s->mu);
    while (!s->stop) {
        // admit first-come first-served while there is room
        while (!s->queue.empty()) {
            Request *r = s->queue.front();
            const size_t na = padded(r->n);
            if (s->max_active && s->active_particles + na > s->max_active && !s->running.empty())
                break;
            uint32_t bid = 0, ai = 0;
            gsim_status st = GSIM_ERROR_CAPACITY;
            for (; ai < s->arenas.size(); ++ai) {  // the first arena with room
                st = gsim_batch_add(s->arenas[ai], &r->params, r->pos.data(), r->vel.data(),
                                    r->mass.data(), r->n, &bid);
                if (st != GSIM_ERROR_CAPACITY) break;
            }
            if (st == GSIM_ERROR_CAPACITY && !s->running.empty()) break;  // wait for room
            s->queue.pop_front();
            if (st != GSIM_OK || gsim_batch_set_steps(s->arenas[ai], bid, r->steps) != GSIM_OK) {
                if (st == GSIM_OK) gsim_batch_remove(s->arenas[ai], bid);
                r->state = 3;
                r->done = Clock::now();
                ++s->failed;
                std::snprintf(s->worker_error, sizeof(s->worker_error), "%s", gsim_last_error());
                s->cv_done.notify_all();
                continue;
            }
            r->arena = ai;
            r->batch_id = bid;
            r->state = 1;
            r->admitted = Clock::now();
            s->running[(uint64_t)ai << 32 | bid] = r;
            s->active_particles += na;
        }
        if (s->running.empty()) {
            s->cv_work.wait(lk, [&] { return s->stop || !s->queue.empty(); });
            continue;
        }
        // Step every running system up to the earliest completion (at most
        // kChunkSteps at a time, so new arrivals wait at most one chunk), then
        // wait for the device; outside the lock, so clients keep submitting.
        uint64_t chunk = kChunkSteps;
        std::vector<char> busy(s->arenas.size(), 0);
        for (auto &kv : s->running) {
            const Request *r = kv.second;
            busy[r->arena] = 1;
            uint64_t done = 0;
            if (gsim_batch_steps_done(s->arenas[r->arena], r->batch_id, &done) == GSIM_OK &&
                r->steps > done)
                chunk = std::min<uint64_t>(chunk, r->steps - done);
        }
        lk.unlock();
        // every busy arena's steps go out on its own stream before any wait
        gsim_status st = GSIM_OK;
        for (size_t a = 0; a < s->arenas.size() && st == GSIM_OK; ++a)
            if (busy[a]) st = gsim_batch_step(s->arenas[a], (uint32_t)chunk);
        for (size_t a = 0; a < s->arenas.size() && st == GSIM_OK; ++a)
            if (busy[a]) st = gsim_batch_synchronize(s->arenas[a]);
        lk.lock();
        if (st != GSIM_OK) {
            std::snprintf(s->worker_error, sizeof(s->worker_error), "%s", gsim_last_error());
            for (auto &kv : s->running) {
                kv.second->state = 3;
                kv.second->done = Clock::now();
                ++s->failed;
            }
            s->running.clear();
            s->active_particles = 0;
            s->cv_done.notify_all();
            continue;
        }
        s->batch_steps += chunk;
        // retire the systems whose budget is spent
        std::vector<Request *> finished;
        for (auto &kv : s->running) {
            const Request *r = kv.second;
            uint64_t done = 0;
            if (gsim_batch_steps_done(s->arenas[r->arena], r->batch_id, &done) == GSIM_OK &&
                done >= r->steps)
                finished.push_back(kv.second);
        }
        for (Request *r : finished) {
            gsim_batch *b = s->arenas[r->arena];
            const gsim_status rs = gsim_batch_read(b, r->batch_id, r->pos.data(), r->vel.data(),
                                                   r->mass.data());
            gsim_batch_remove(b, r->batch_id);
            s->running.erase((uint64_t)r->arena << 32 | r->batch_id);
            s->active_particles -= padded(r->n);
            r->done = Clock::now();
            if (rs == GSIM_OK) {
                r->state = 2;
                ++s->completed;
                s->latencies_ms.push_back(
                    std::chrono::duration<double, std::milli>(r->done - r->submitted).count());
                s->last_done = r->done;
                if (r->params.algorithm == GSIM_DIRECT_SUM) {
                    const uint64_t na = padded(r->n);
                    r->interactions = na * na * r->steps;
                    s->ds_interactions += r->interactions;
                }
            } else {
                r->state = 3;
                ++s->failed;
            }
        }
        if (!finished.empty()) s->cv_done.notify_all();
    }
}

double percentile(std::vector<double> v, double q) {
    if (v.empty()) return 0.0;
    std::sort(v.begin(), v.end());
    const double pos = q * (double)(v.size() - 1);
    const size_t lo = (size_t)pos;
    const size_t hi = std::min(lo + 1, v.size() - 1);
    return v[lo] + (v[hi] - v[lo]) * (pos - (double)lo);
}

}  // namespace

extern "C" {

gsim_status gsim_server_create_arenas(int32_t device, size_t hbm_budget_bytes,
                                      uint32_t max_systems, size_t max_active_particles,
                                      size_t arena_particles, gsim_server **out) {
    gsim::clear_error();
    if (!out) return fail(GSIM_ERROR_INVALID_ARGUMENT, "gsim_server_create: out is NULL");
    *out = nullptr;
    const size_t per = gsim_batch_bytes_per_particle();
    // the budget buys this many padded particles, in arenas of at most the
    // arena's index limit (or the caller's smaller cap)
    const size_t total = hbm_budget_bytes / per / 64 * 64;
    size_t arena = arena_particles ? std::min(arena_particles / 64 * 64, gsim::kBatchMaxParticles)
                                   : gsim::kBatchMaxParticles;
    if (total < 64 || arena < 64)
        return fail(GSIM_ERROR_INVALID_ARGUMENT, "HBM budget below one 64-particle system");
    if ((total + arena - 1) / arena > 64)
        return fail(GSIM_ERROR_INVALID_ARGUMENT, "the budget would need more than 64 arenas");
    gsim_server *s = new (std::nothrow) gsim_server();
    if (!s) return fail(GSIM_ERROR_OUT_OF_MEMORY, "host allocation failed");
    for (size_t left = total; left >= 64;) {
        const size_t cap = std::min(left, arena);
        gsim_batch *b = nullptr;
        gsim_status st = gsim_batch_create(device, cap, max_systems, &b);
        if (st != GSIM_OK) {
            for (gsim_batch *x : s->arenas) gsim_batch_destroy(x);
            delete s;
            return st;
        }
        s->arenas.push_back(b);
        s->capacity += cap;
        s->largest_arena = std::max(s->largest_arena, cap);
        left -= cap;
    }
    s->max_active = max_active_particles;
    s->worker = std::thread(worker_loop, s);
    *out = s;
    return GSIM_OK;
}

gsim_status gsim_server_create(int32_t device, size_t hbm_budget_bytes, uint32_t max_systems,
                               size_t max_active_particles, gsim_server **out) {
    return gsim_server_create_arenas(device, hbm_budget_bytes, max_systems, max_active_particles,
                                     0, out);
}

gsim_status gsim_server_submit(gsim_server *s, const gsim_params *params, const float *pos_xy,
                               const float *vel_xy, const float *mass, size_t n, uint32_t steps,
                               uint64_t *request_id) {
    gsim::clear_error();
    if (!s || !params || !pos_xy || !vel_xy || !mass || !request_id)
        return fail(GSIM_ERROR_INVALID_ARGUMENT, "gsim_server_submit: NULL argument");
    if (n == 0) return fail(GSIM_ERROR_INVALID_ARGUMENT, "The simulation must contain at least one particle!");
    if (padded(n) > s->largest_arena)
        return fail(GSIM_ERROR_CAPACITY, "a %zu-particle system exceeds the HBM budget%s", n,
                    s->arenas.size() > 1 ? " (its largest arena)" : "");
    if (params->algorithm == GSIM_DIRECT_SUM && padded(n) >= gsim::kDsSymMinParticles)
        return fail(GSIM_ERROR_INVALID_ARGUMENT,
                    "direct-sum systems of %zu or more padded particles are stepped alone (gsim_create)",
                    gsim::kDsSymMinParticles);
    std::unique_ptr<Request> r(new (std::nothrow) Request());
    if (!r) return fail(GSIM_ERROR_OUT_OF_MEMORY, "host allocation failed");
    r->params = *params;
    r->n = n;
    r->steps = steps;
    r->pos.assign(pos_xy, pos_xy + 2 * n);
    r->vel.assign(vel_xy, vel_xy + 2 * n);
    r->mass.assign(mass, mass + n);
    std::lock_guard<std::mutex> lk(s->mu);
    r->id = s->next_id++;
    r->submitted = Clock::now();
    if (!s->any_submit) {
        s->first_submit = r->submitted;
        s->any_submit = true;
    }
    *request_id = r->id;
    Request *raw = r.get();
    s->all[r->id] = std::move(r);
    if (steps == 0) {  // nothing to run: complete at once
        raw->state = 2;
        raw->done = raw->submitted;
        ++s->completed;
        s->latencies_ms.push_back(0.0);
        s->cv_done.notify_all();
        return GSIM_OK;
    }
    s->queue.push_back(raw);
    s->cv_work.notify_one();
    return GSIM_OK;
}

gsim_status gsim_server_wait(gsim_server *s, uint64_t request_id, double timeout_s, int32_t *done) {
    gsim::clear_error();
    if (!s || !done) return fail(GSIM_ERROR_INVALID_ARGUMENT, "gsim_server_wait: NULL argument");
    std::unique_lock<std::mutex> lk(s->mu);
    auto it = s->all.find(request_id);
    if (it == s->all.end()) return fail(GSIM_ERROR_INVALID_ARGUMENT, "unknown request %llu", (unsigned long long)request_id);
    Request *r = it->second.get();
    s->cv_done.wait_for(lk, std::chrono::duration<double>(timeout_s), [&] { return r->state >= 2; });
    *done = r->state >= 2 ? 1 : 0;
    if (r->state == 3) return fail(GSIM_ERROR_DEVICE, "request failed: %s", s->worker_error);
    return GSIM_OK;
}

gsim_status gsim_server_result(gsim_server *s, uint64_t request_id, float *pos_xy, float *vel_xy,
                               float *mass, gsim_request_info *info) {
    gsim::clear_error();
    if (!s) return fail(GSIM_ERROR_INVALID_ARGUMENT, "gsim_server_result: NULL handle");
    std::lock_guard<std::mutex> lk(s->mu);
    auto it = s->all.find(request_id);
    if (it == s->all.end()) return fail(GSIM_ERROR_INVALID_ARGUMENT, "unknown request %llu", (unsigned long long)request_id);
    Request *r = it->second.get();
    if (r->state != 2) return fail(GSIM_ERROR_INVALID_ARGUMENT, "request %llu has not completed", (unsigned long long)request_id);
    if (pos_xy) std::memcpy(pos_xy, r->pos.data(), sizeof(float) * 2 * r->n);
    if (vel_xy) std::memcpy(vel_xy, r->vel.data(), sizeof(float) * 2 * r->n);
    if (mass) std::memcpy(mass, r->mass.data(), sizeof(float) * r->n);
    if (info) {
        auto ms = [&](Clock::time_point a, Clock::time_point b) {
            return std::chrono::duration<double, std::milli>(b - a).count();
        };
        info->n = r->n;
        info->steps = r->steps;
        info->queue_ms = r->admitted.time_since_epoch().count() ? ms(r->submitted, r->admitted) : 0.0;
        info->latency_ms = ms(r->submitted, r->done);
    }
    return GSIM_OK;
}

gsim_status gsim_server_release(gsim_server *s, uint64_t request_id) {
    gsim::clear_error();
    if (!s) return fail(GSIM_ERROR_INVALID_ARGUMENT, "gsim_server_release: NULL handle");
    std::lock_guard<std::mutex> lk(s->mu);
    auto it = s->all.find(request_id);
    if (it == s->all.end()) return GSIM_OK;
    if (it->second->state < 2) return fail(GSIM_ERROR_INVALID_ARGUMENT, "request is still running");
    s->all.erase(it);
    return GSIM_OK;
}

gsim_status gsim_server_drain(gsim_server *s, double timeout_s, int32_t *idle) {
    gsim::clear_error();
    if (!s || !idle) return fail(GSIM_ERROR_INVALID_ARGUMENT, "gsim_server_drain: NULL argument");
    std::unique_lock<std::mutex> lk(s->mu);
    s->cv_done.wait_for(lk, std::chrono::duration<double>(timeout_s),
                        [&] { return s->queue.empty() && s->running.empty(); });
    *idle = s->queue.empty() && s->running.empty() ? 1 : 0;
    return GSIM_OK;
}

gsim_status gsim_server_query(gsim_server *s, gsim_server_info *out) {
    gsim::clear_error();
    if (!s || !out) return fail(GSIM_ERROR_INVALID_ARGUMENT, "gsim_server_query: NULL argument");
    uint64_t accepts = 0;
    {
        std::lock_guard<std::mutex> lk(s->mu);
        out->completed = s->completed;
        out->failed = s->failed;
        out->queued = s->queue.size();
        out->running = s->running.size();
        out->batch_steps = s->batch_steps;
        out->p50_ms = percentile(s->latencies_ms, 0.50);
        out->p99_ms = percentile(s->latencies_ms, 0.99);
        out->max_ms = s->latencies_ms.empty() ? 0.0 : *std::max_element(s->latencies_ms.begin(), s->latencies_ms.end());
        out->span_s = s->any_submit && s->completed
                          ? std::chrono::duration<double>(s->last_done - s->first_submit).count()
                          : 0.0;
        out->ds_interactions = s->ds_interactions;
        out->capacity_particles = s->capacity;
    }
    uint64_t visits = 0;
    for (gsim_batch *b : s->arenas) {
        uint64_t a = 0, v = 0;
        gsim_batch_accepts(b, &a);
        gsim_batch_visits(b, &v);
        accepts += a;
        visits += v;
    }
    out->bh_interactions = accepts;
    out->bh_visits = visits;
    out->arenas = (uint32_t)s->arenas.size();
    return GSIM_OK;
}

gsim_status gsim_server_destroy(gsim_server *s) {
    gsim::clear_error();
    if (!s) return GSIM_OK;
    {
        std::lock_guard<std::mutex> lk(s->mu);
        s->stop = true;
    }
    s->cv_work.notify_all();
    if (s->worker.joinable()) s->worker.join();
    for (gsim_batch *b : s->arenas) gsim_batch_destroy(b);
    delete s;
    return GSIM_OK;
}

}  // extern "C"
