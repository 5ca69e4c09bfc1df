// tc_serve.cpp -- request server with continuous batching (no counterpart in
// the reference, whose two functions are one-shot and non-reentrant,
// upward.cpp:51-59).
//
// A server owns `nslots` slots; each slot is a tc_context (own HIP stream,
// own grow-only HBM buffers, own resident tree) driven by its own host
// thread.  Requests wait in one queue; whenever a slot is free a worker
// takes the first queued request whose HBM footprint fits in the budget
// still free, preferring a request whose sources are already resident in
// that slot (same source_id, kind and precision), in which case the tree is
// reused and only the new targets are evaluated.  Requests running in
// different slots overlap on the device: copies of one request run beside
// the kernels of others.  Kernels never communicate across requests and use
// no cross-request atomics, so a result is bit-identical alone and in a mix.
#include "tc_common.h"

#include "../../include/treecodes2d.h"

#include <chrono>
#include <condition_variable>
#include <deque>
#include <mutex>
#include <string>
#include <thread>
#include <unordered_map>
#include <vector>

namespace {

double now_ms()
{
    using namespace std::chrono;
    return duration<double, std::milli>(steady_clock::now().time_since_epoch()).count();
}

struct Job {
    long long ticket = 0;
    tc_request r{};
    size_t footprint = 0;
    double t_submit = 0, t_start = 0, t_done = 0;
    int status = 0;
    int slot = -1;
    bool reused = false;
    bool done = false;
    std::string err;
};

// The request's expansion order and leaf capacity (0 selects the reference's).
int req_order(const tc_request &r) { return r.order > 0 ? r.order : (r.kind == TC_REQ_POTENTIAL ? 12 : 24); }
int req_leaf(const tc_request &r) { return r.leaf > 0 ? r.leaf : (r.kind == TC_REQ_POTENTIAL ? 512 : 192); }

// Upper estimate of the HBM a request occupies while it runs.
size_t footprint(const tc_request &r)
{
    const size_t n = (size_t)(r.n > 0 ? r.n : 0);
    const int p = req_order(r);
    const int leaf = req_leaf(r);
    const size_t nodes = 8 * n / (size_t)leaf + 64;
    size_t b = n * (24 + 32 + 8 + 8 + 8 + (r.precision ? 16 : 0));
    b += nodes * (64 + 32 + 24 + 16 * (size_t)p + (r.precision ? 8 * (size_t)p : 0));
    if (r.kind == TC_REQ_POTENTIAL)
        b += (size_t)(r.m > 0 ? r.m : 0) * (16 + 8 + 8 + 8 + 8 + 8);
    else
        b += (size_t)(r.nblocks > 0 ? r.nblocks : 0) * (24 + 2 * 8 * 1024);
    return b + (1u << 20);
}

}  // namespace

struct tc_server {
    int device = 0;
    size_t budget = 0, in_use = 0, peak = 0;
    std::vector<tc_context *> ctx;
    std::vector<long long> res_src;
    std::vector<int> res_kind, res_prec, res_order, res_leaf;
    std::mutex mu;
    std::condition_variable cv_work, cv_done;
    std::deque<Job *> queue;
    std::unordered_map<long long, Job *> jobs;
    std::vector<std::thread> workers;
    bool stop = false;
    long long next_ticket = 1;
    long long completed = 0, reused = 0, failed = 0;
    int inflight = 0, max_inflight = 0;
};

namespace {

bool resident(const tc_server *s, int slot, const tc_request &r)
{
    return r.source_id != 0 && s->res_src[slot] == r.source_id && s->res_kind[slot] == r.kind &&
           s->res_prec[slot] == r.precision && s->res_order[slot] == req_order(r) && s->res_leaf[slot] == req_leaf(r);
}

void run_job(tc_server *s, int slot, Job *j)
{
    tc_context *c = s->ctx[slot];
    const tc_request &r = j->r;
    const bool pot = r.kind == TC_REQ_POTENTIAL;
    int rc = tc_set_option(c, TC_OPT_PRECISION, r.precision);
    if (rc == 0)
        rc = tc_set_option(c, pot ? TC_OPT_POT_ORDER : TC_OPT_FORCE_ORDER, req_order(r));
    if (rc == 0)
        rc = tc_set_option(c, pot ? TC_OPT_POT_LEAF : TC_OPT_FORCE_LEAF, req_leaf(r));
    if (rc == 0) {
        if (r.kind == TC_REQ_POTENTIAL)
            rc = j->reused ? tc_potential_resident(c, r.theta, r.xt, r.yt, r.m, r.out0, 0)
                           : tc_potential(c, r.theta, r.xs, r.ys, r.vs, r.n, r.xt, r.yt, r.m, r.out0, 0);
        else
            rc = j->reused ? tc_force_mrag_resident(c, r.theta, r.x0s, r.y0s, r.hs, r.nblocks, r.out0, r.out1, 0)
                           : tc_force_mrag(c, r.theta, r.xs, r.ys, r.vs, r.n, r.x0s, r.y0s, r.hs, r.nblocks, r.out0,
                                           r.out1, 0);
    }
    j->status = rc;
    if (rc)
        j->err = treecode_last_error();
}

void worker(tc_server *s, int slot)
{
    std::unique_lock<std::mutex> lk(s->mu);
    while (true) {
        Job *pick = nullptr;
        auto fits = [&](const Job *j) { return s->in_use + j->footprint <= s->budget || s->inflight == 0; };
        for (auto it = s->queue.begin(); it != s->queue.end() && !pick; ++it)
            if (resident(s, slot, (*it)->r) && fits(*it))
                pick = *it;
        for (auto it = s->queue.begin(); it != s->queue.end() && !pick; ++it)
            if (fits(*it))
                pick = *it;
        if (!pick) {
            if (s->stop)
                return;
            s->cv_work.wait(lk);
            continue;
        }
        for (auto it = s->queue.begin(); it != s->queue.end(); ++it)
            if (*it == pick) {
                s->queue.erase(it);
                break;
            }
        pick->reused = resident(s, slot, pick->r);
        pick->slot = slot;
        pick->t_start = now_ms();
        s->in_use += pick->footprint;
        s->peak = std::max(s->peak, s->in_use);
        s->inflight++;
        s->max_inflight = std::max(s->max_inflight, s->inflight);
        lk.unlock();
        run_job(s, slot, pick);
        lk.lock();
        s->in_use -= pick->footprint;
        s->inflight--;
        if (pick->status == 0) {
            s->res_src[slot] = pick->r.source_id;
            s->res_kind[slot] = pick->r.kind;
            s->res_prec[slot] = pick->r.precision;
            s->res_order[slot] = req_order(pick->r);
            s->res_leaf[slot] = req_leaf(pick->r);
            s->completed++;
            s->reused += pick->reused;
        } else {
            s->res_src[slot] = 0;
            s->failed++;
        }
        pick->t_done = now_ms();
        pick->done = true;
        s->cv_done.notify_all();
        s->cv_work.notify_all();
    }
}

}  // namespace

extern "C" {

TC_API tc_server *tc_server_create(int device, int nslots, double hbm_budget_gib)
{
    if (nslots < 1 || nslots > 64 || !(hbm_budget_gib > 0)) {
        tc_context_destroy(nullptr);
        return nullptr;
    }
    tc_server *s = new tc_server();
    s->device = device;
    s->budget = (size_t)(hbm_budget_gib * 1073741824.0);
    for (int i = 0; i < nslots; ++i) {
        tc_context *c = tc_context_create(device);
        if (!c) {
            for (tc_context *o : s->ctx)
                tc_context_destroy(o);
            delete s;
            return nullptr;
        }
        s->ctx.push_back(c);
        s->res_src.push_back(0);
        s->res_kind.push_back(-1);
        s->res_order.push_back(0);
        s->res_leaf.push_back(0);
        s->res_prec.push_back(-1);
    }
    for (int i = 0; i < nslots; ++i)
        s->workers.emplace_back(worker, s, i);
    return s;
}

TC_API void tc_server_destroy(tc_server *s)
{
    if (!s)
        return;
    {
        std::lock_guard<std::mutex> lk(s->mu);
        s->stop = true;
    }
    s->cv_work.notify_all();
    for (auto &t : s->workers)
        t.join();
    for (auto &kv : s->jobs)
        delete kv.second;
    for (tc_context *c : s->ctx)
        tc_context_destroy(c);
    delete s;
}

TC_API long long tc_server_submit(tc_server *s, const tc_request *r)
{
    if (!s || !r)
        return -1;
    Job *j = new Job();
    j->r = *r;
    j->footprint = footprint(*r);
    j->t_submit = now_ms();
    std::lock_guard<std::mutex> lk(s->mu);
    if (j->footprint > s->budget) {
        delete j;
        return -2;
    }
    j->ticket = s->next_ticket++;
    s->jobs[j->ticket] = j;
    s->queue.push_back(j);
    s->cv_work.notify_one();
    return j->ticket;
}

TC_API int tc_server_wait(tc_server *s, long long ticket, double *times_ms)
{
    if (!s)
        return -1;
    std::unique_lock<std::mutex> lk(s->mu);
    auto it = s->jobs.find(ticket);
    if (it == s->jobs.end())
        return -1;
    Job *j = it->second;
    s->cv_done.wait(lk, [&] { return j->done; });
    if (times_ms) {
        times_ms[0] = j->t_start - j->t_submit;  // queueing
        times_ms[1] = j->t_done - j->t_start;    // service
        times_ms[2] = j->t_done - j->t_submit;   // latency
        times_ms[3] = j->reused ? 1.0 : 0.0;
        times_ms[4] = (double)j->slot;
    }
    const int st = j->status;
    s->jobs.erase(it);
    delete j;
    return st;
}

TC_API int tc_server_stats(tc_server *s, long long *out)
{
    if (!s)
        return -1;
    std::lock_guard<std::mutex> lk(s->mu);
    out[0] = s->completed;
    out[1] = s->reused;
    out[2] = s->failed;
    out[3] = s->max_inflight;
    out[4] = (long long)s->peak;
    out[5] = (long long)s->budget;
    out[6] = (long long)s->queue.size();
    return 0;
}

}  // extern "C"
