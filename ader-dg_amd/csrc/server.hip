// server.hip -- the serving loop of include/aderdg.h (aderdg_server_*):
// continuous batching of solve requests on one GPU.
//
// The worker thread owns a Batch.  Each round it admits queued requests in
// arrival order (initial state computed on the device), enqueues
// steps_per_round replays of the step graph and a copy of every solve state
// into one of two pinned buffers; it then waits for the previous round's copy
// - the device meanwhile runs the rounds queued after it - and retires the
// solves that copy shows as finished.  A finished solve is inactive, so the
// steps queued behind the copy leave it untouched; its removal is ordered
// after them on the stream.
#include <hip/hip_runtime.h>

#include <atomic>
#include <chrono>
#include <condition_variable>
#include <cstring>
#include <deque>
#include <map>
#include <mutex>
#include <thread>
#include <vector>

#include "../../include/aderdg.h"
#include "batch.hpp"
#include "capi_common.hpp"

using namespace aderdg;
using Clock = std::chrono::steady_clock;

namespace {

struct Request {
    int64_t ticket;
    aderdg_request req;
    Clock::time_point submitted;
    int64_t bytes;
    int batch_id = -1;
    int64_t admit_round = 0;
};

struct Done {
    int64_t ticket;
    double latency;
    int status;
    int steps = 0;
    int reached_tend = 0;
};

struct Result {
    std::vector<double> uh;
    double time;
    int steps;
};

}  // namespace

struct aderdg_server {
    aderdg_server_desc desc{};
    Batch batch;
    std::thread worker;
    std::mutex mu;
    std::condition_variable cv, idle_cv;
    std::deque<Request> inbox;  // submitted, not yet seen by the worker (under mu)
    std::vector<Done> done;     // finished since the last poll (under mu)
    std::map<int64_t, Result> results;
    aderdg_server_stats stats{};
    int64_t next_ticket = 0;
    bool started = false, stop = false;
    int worker_error = 0;
    // worker-only state
    std::deque<Request> pending;
    std::map<int, Request> running;  // by batch id
    std::vector<SolveState *> snap[2];
    hipEvent_t snap_ev[2] = {nullptr, nullptr};
    int64_t snap_round[2] = {-1, -1};
    int64_t round = 0;
};

namespace {

// Snapshot every group's solve states into buffer k after the queued steps.
int take_snapshot(aderdg_server *s, int k)
{
    Batch &b = s->batch;
    for (size_t g = 0; g < b.groups.size(); g++)
        if (b.groups[g]->read_states_async(s->snap[k][g])) return fail(ADERDG_EHIP, "server: state copy failed");
    HIPCHK(hipEventRecord(s->snap_ev[k], b.stream));
    s->snap_round[k] = s->round;
    return 0;
}

int admit(aderdg_server *s)
{
    Batch &b = s->batch;
    while (!s->pending.empty()) {
        Request &r = s->pending.front();
        if (s->desc.hbm_budget > 0 && s->stats.bytes_in_use + r.bytes > s->desc.hbm_budget &&
            !s->running.empty())
            break;  // wait for room; a lone request larger than the budget still runs
        int id = -1;
        int rc = b.add(&r.req.solve, nullptr, &r.req.ic, 0.0, r.req.tend, r.req.max_steps, 1, &id);
        if (rc == ADERDG_EFULL) {
            if (s->running.empty()) {  // can never fit: fail it rather than wait forever
                std::lock_guard<std::mutex> lk(s->mu);
                s->done.push_back({r.ticket, std::chrono::duration<double>(Clock::now() - r.submitted).count(), rc, 0, 0});
                s->stats.failed++;
                s->stats.queued--;
                s->pending.pop_front();
                continue;
            }
            break;
        }
        if (rc) {
            std::lock_guard<std::mutex> lk(s->mu);
            s->done.push_back({r.ticket, std::chrono::duration<double>(Clock::now() - r.submitted).count(), rc, 0, 0});
            s->stats.failed++;
            s->stats.queued--;
            s->pending.pop_front();
            continue;
        }
        r.batch_id = id;
        r.admit_round = s->round;
        {
            std::lock_guard<std::mutex> lk(s->mu);
            s->stats.admitted++;
            s->stats.queued--;
            s->stats.running++;
            s->stats.bytes_in_use += r.bytes;
            if (s->stats.bytes_in_use > s->stats.peak_bytes) s->stats.peak_bytes = s->stats.bytes_in_use;
        }
        s->running.emplace(id, r);
        s->pending.pop_front();
    }
    return 0;
}

// Retire the solves that snapshot k shows as finished.
int retire(aderdg_server *s, int k)
{
    Batch &b = s->batch;
    std::vector<int> fin;
    for (auto &kv : s->running) {
        if (kv.second.admit_round >= s->snap_round[k]) continue;  // added after this snapshot was queued
        const auto &h = b.solves.at(kv.first);
        const SolveState &st = s->snap[k][h.group][h.slot];
        if (!st.active) fin.push_back(kv.first);
    }
    for (int id : fin) {
        Request r = s->running.at(id);
        const auto &h = b.solves.at(id);
        const SolveState st = s->snap[k][h.group][h.slot];
        const Group &g = *b.groups[h.group];
        const int64_t dof = (int64_t)g.V * g.n * g.n * g.mesh(h.slot).ncell * st.steps;
        Result res;
        if (r.req.keep_result) {
            res.uh.resize((size_t)g.V * g.n * g.n * g.mesh(h.slot).ncell);
            HIPCHK(hipStreamWaitEvent(b.copy, s->snap_ev[k], 0));
            if (b.groups[h.group]->download_uh(h.slot, res.uh.data(), st.parity, b.copy))
                return fail(ADERDG_EHIP, "server: result copy failed");
            HIPCHK(hipStreamSynchronize(b.copy));
            res.time = st.time;
            res.steps = st.steps;
        }
        int rc = b.remove(id);
        if (rc) return rc;
        s->running.erase(id);
        const double lat = std::chrono::duration<double>(Clock::now() - r.submitted).count();
        std::lock_guard<std::mutex> lk(s->mu);
        if (r.req.keep_result) s->results[r.ticket] = std::move(res);
        // a solve that ran out of limited-cell pool slots ends failed (ADERDG_ECAPACITY)
        s->done.push_back({r.ticket, lat, st.failed ? ADERDG_ECAPACITY : 0, st.steps, st.time < r.req.tend ? 0 : 1});
        if (st.failed)
            s->stats.failed++;
        else
            s->stats.completed++;
        s->stats.running--;
        s->stats.dof_updates += dof;
        s->stats.bytes_in_use -= r.bytes;
    }
    return 0;
}

void work(aderdg_server *s)
{
    int rc = 0;
    if (hipSetDevice(s->desc.device) != hipSuccess) rc = fail(ADERDG_EHIP, "server: hipSetDevice failed");
    const int spr = s->desc.steps_per_round > 0 ? s->desc.steps_per_round : 2;
    int cur = 0;
    bool have_prev = false;
    while (!rc) {
        {
            std::unique_lock<std::mutex> lk(s->mu);
            if (s->running.empty() && s->pending.empty() && s->inbox.empty() && !s->stop) {
                s->idle_cv.notify_all();
                s->cv.wait_for(lk, std::chrono::milliseconds(2));
            }
            if (s->stop) break;
            while (!s->inbox.empty()) {
                s->pending.push_back(s->inbox.front());
                s->inbox.pop_front();
            }
        }
        if ((rc = admit(s))) break;
        if (s->running.empty()) {
            have_prev = false;
            continue;
        }
        const auto t0 = Clock::now();
        if ((rc = s->batch.step(spr))) break;
        s->round++;
        if ((rc = take_snapshot(s, cur))) break;
        if (have_prev) {  // the device keeps running the round just queued
            if (hipEventSynchronize(s->snap_ev[cur ^ 1]) != hipSuccess) {
                rc = fail(ADERDG_EHIP, "server: step failed");
                break;
            }
            if ((rc = retire(s, cur ^ 1))) break;
        }
        have_prev = true;
        cur ^= 1;
        {
            std::lock_guard<std::mutex> lk(s->mu);
            s->stats.busy_seconds += std::chrono::duration<double>(Clock::now() - t0).count();
        }
        if (s->running.empty()) {  // last solves retired: the newest snapshot holds nothing to retire
            if (hipEventSynchronize(s->snap_ev[cur ^ 1]) != hipSuccess) rc = fail(ADERDG_EHIP, "server: step failed");
            have_prev = false;
        }
    }
    std::lock_guard<std::mutex> lk(s->mu);
    s->worker_error = rc;
    s->stop = true;
    s->idle_cv.notify_all();
}

}  // namespace

extern "C" {

int aderdg_server_create(const aderdg_server_desc *d, aderdg_server **out)
{
    if (!d || !out) return fail(ADERDG_EINVAL, "aderdg_server_create: null argument");
    *out = nullptr;
    auto *s = new aderdg_server();
    s->desc = *d;
    int r = s->batch.init(d->device, d->persist_blocks);
    if (!r) {
        for (int k = 0; k < 2 && !r; k++)
            if (hipEventCreateWithFlags(&s->snap_ev[k], hipEventDisableTiming) != hipSuccess)
                r = fail(ADERDG_EHIP, "aderdg_server_create: hipEventCreate failed");
    }
    if (r) {
        aderdg_server_destroy(s);
        return r;
    }
    *out = s;
    return 0;
}

int aderdg_server_reserve_limited(aderdg_server *s, int pde, int N, int64_t max_cells, int max_solves,
                                  double limited_fraction)
{
    if (!s) return fail(ADERDG_EINVAL, "aderdg_server_reserve: null server");
    if (s->started) return fail(ADERDG_ESTATE, "aderdg_server_reserve: server already started");
    int r = s->batch.reserve(pde, N, max_cells, max_solves, limited_fraction);
    if (r) return r;
    for (int k = 0; k < 2; k++) {
        SolveState *p = nullptr;
        if (hipHostMalloc((void **)&p, sizeof(SolveState) * max_solves, hipHostMallocDefault) != hipSuccess)
            return fail(ADERDG_EHIP, "aderdg_server_reserve: hipHostMalloc failed");
        std::memset(p, 0, sizeof(SolveState) * max_solves);
        s->snap[k].push_back(p);
    }
    return 0;
}

int aderdg_server_reserve(aderdg_server *s, int pde, int N, int64_t max_cells, int max_solves)
{
    return aderdg_server_reserve_limited(s, pde, N, max_cells, max_solves, 1.0);
}

int aderdg_server_start(aderdg_server *s)
{
    if (!s) return fail(ADERDG_EINVAL, "aderdg_server_start: null server");
    if (s->started) return fail(ADERDG_ESTATE, "aderdg_server_start: already started");
    if (s->batch.groups.empty()) return fail(ADERDG_ESTATE, "aderdg_server_start: no arena reserved");
    s->started = true;
    s->worker = std::thread(work, s);
    return 0;
}

int aderdg_server_submit(aderdg_server *s, const aderdg_request *req, int64_t *ticket)
{
    if (!s || !req || !ticket) return fail(ADERDG_EINVAL, "aderdg_server_submit: null argument");
    int r = check_desc(&req->solve, "aderdg_server_submit");
    if (r) return r;
    if (req->max_steps < 1) return fail(ADERDG_EINVAL, "aderdg_server_submit: max_steps must be positive");
    if (s->batch.group_of(req->solve.pde, req->solve.N) < 0)
        return fail(ADERDG_EINVAL, "aderdg_server_submit: no arena for (pde %d, N %d)", req->solve.pde, req->solve.N);
    Request q;
    q.req = *req;
    q.submitted = Clock::now();
    q.bytes = Batch::footprint(req->solve.pde, req->solve.N, req->solve.imax, req->solve.jmax,
                               s->batch.lh_fraction(req->solve.pde, req->solve.N));
    std::lock_guard<std::mutex> lk(s->mu);
    if (s->stop) return fail(ADERDG_ESTATE, "aderdg_server_submit: server stopped (%d)", s->worker_error);
    q.ticket = s->next_ticket++;
    *ticket = q.ticket;
    s->inbox.push_back(q);
    s->stats.submitted++;
    s->stats.queued++;
    s->cv.notify_one();
    return 0;
}

int aderdg_server_poll(aderdg_server *s, int64_t *tickets, double *latency_s, int *status, int *steps,
                       int *reached_tend, int max, int *count)
{
    if (!s || !count || max < 0 || (max > 0 && !tickets)) return fail(ADERDG_EINVAL, "aderdg_server_poll: bad argument");
    std::lock_guard<std::mutex> lk(s->mu);
    int c = 0;
    while (c < max && c < (int)s->done.size()) {
        tickets[c] = s->done[c].ticket;
        if (latency_s) latency_s[c] = s->done[c].latency;
        if (status) status[c] = s->done[c].status;
        if (steps) steps[c] = s->done[c].steps;
        if (reached_tend) reached_tend[c] = s->done[c].reached_tend;
        c++;
    }
    s->done.erase(s->done.begin(), s->done.begin() + c);
    *count = c;
    if (c == 0 && s->worker_error) return fail(s->worker_error, "aderdg_server_poll: worker stopped on an error");
    return 0;
}

int aderdg_server_result(aderdg_server *s, int64_t ticket, double *uh, double *time, int *steps)
{
    if (!s) return fail(ADERDG_EINVAL, "aderdg_server_result: null server");
    std::lock_guard<std::mutex> lk(s->mu);
    auto it = s->results.find(ticket);
    if (it == s->results.end()) return fail(ADERDG_EINVAL, "aderdg_server_result: no kept result for ticket %lld", (long long)ticket);
    if (uh) std::memcpy(uh, it->second.uh.data(), it->second.uh.size() * sizeof(double));
    if (time) *time = it->second.time;
    if (steps) *steps = it->second.steps;
    s->results.erase(it);
    return 0;
}

int aderdg_server_drain(aderdg_server *s, double timeout_s)
{
    if (!s) return fail(ADERDG_EINVAL, "aderdg_server_drain: null server");
    std::unique_lock<std::mutex> lk(s->mu);
    const auto until = Clock::now() + std::chrono::duration_cast<Clock::duration>(std::chrono::duration<double>(timeout_s));
    while (!s->stop && (s->stats.queued > 0 || s->stats.running > 0)) {
        if (timeout_s > 0 && Clock::now() >= until) return fail(ADERDG_ESTATE, "aderdg_server_drain: timed out");
        s->idle_cv.wait_for(lk, std::chrono::milliseconds(5));
    }
    if (s->worker_error) return fail(s->worker_error, "aderdg_server_drain: worker stopped on an error");
    return 0;
}

int aderdg_server_stats_get(aderdg_server *s, aderdg_server_stats *st)
{
    if (!s || !st) return fail(ADERDG_EINVAL, "aderdg_server_stats_get: null argument");
    std::lock_guard<std::mutex> lk(s->mu);
    *st = s->stats;
    return 0;
}

int aderdg_server_destroy(aderdg_server *s)
{
    if (!s) return 0;
    {
        std::lock_guard<std::mutex> lk(s->mu);
        s->stop = true;
        s->cv.notify_all();
    }
    if (s->worker.joinable()) s->worker.join();
    s->batch.release();
    for (int k = 0; k < 2; k++) {
        for (auto p : s->snap[k]) (void)hipHostFree(p);
        if (s->snap_ev[k]) (void)hipEventDestroy(s->snap_ev[k]);
    }
    delete s;
    return 0;
}

}  // extern "C"
