// tc_serve.cpp -- request server with continuous batching (no counterpart in
// the reference, whose two functions are one-shot and non-reentrant,
// upward.cpp:51-59).
//
// A server owns `nslots` slots; each slot is a tc_context (own HIP streams,
// own scratch) driven by its own host thread.  Requests wait in one queue;
// whenever a slot is free a worker takes the first queued request that fits
// the HBM budget, preferring one whose tree is already cached.
//
// Trees live in a cache shared by all slots, keyed by the request's
// source_id (and kind, precision, order, leaf capacity): a request whose
// sources were seen before evaluates its new targets against the cached tree,
// whichever slot built it (the builder's stream records an event after the
// build; readers wait on it on the GPU).  Trees are read-only once built.
// At most 4 trees per slot stay cached: beyond, the least recently used idle
// tree's buffers are kept for the next build instead of being freed (one
// spare tree per slot; a build takes the smallest spare that fits its
// estimate, and a full spare pool keeps the larger trees), so a steady stream
// allocates little once its buffers have grown.  What is freed in steady
// state -- trees beyond the spares, the old memory of grown buffers -- is
// freed by a reaper thread, not by the slot threads: hipFree waits for the
// whole device and would stall a request behind other slots' kernels.  A
// slot hands its request's old buffers over only after synchronising the
// request's streams; until freed, their bytes count as held.
//
// HBM admission counts everything the server holds: the cached and spare
// trees and each slot's grow-only buffers, plus an upper estimate of what each
// running request may add.  A request that does not fit first frees spare
// trees, then evicts idle trees (least recently used), then releases idle
// slots' buffers (these frees, under memory pressure only, hold the server
// lock); if it still does not fit it waits for running requests to finish (a
// request alone always runs: submission rejects one larger than the whole
// budget).
//
// With TC_SERVER_BATCH, potential traversals of requests in different slots
// are batched: each slot prepares its request (copies, tree, target order) on
// its own stream, then hands the traversal to a leader -- the first slot
// that finds none running -- which launches the traversals of all waiting
// requests of the same order and precision as one kernel (k_potential_batch,
// a persistent grid draining the requests' target waves from one work
// counter, through a graph built once per size bucket and replayed) on one of
// their streams; a traversal that finds no other waiting runs unbatched.
//
// No kernel combines values across requests (the batch kernel's work counter
// only decides which wave computes an item), so a result is bit-identical
// alone and in a mix, batched or not.
#include "tc_common.h"

#include "../../include/treecodes2d.h"

#include <algorithm>
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

// The request's expansion order and leaf capacity (0 selects the reference's).
int req_order(const tc_request &r) { return r.order > 0 ? r.order : (r.kind == TC_REQ_POTENTIAL ? 12 : 24); }
int req_leaf(const tc_request &r) { return r.leaf > 0 ? r.leaf : (r.kind == TC_REQ_POTENTIAL ? 512 : 192); }

size_t pos(long long v) { return v > 0 ? (size_t)v : 0; }

// Upper estimate of a request's tree in HBM: sorted sources and mirrors, and
// nodes (at most 15 n / leaf internal cells over the 15 levels, four children
// each) with their expansions.
size_t tree_estimate(const tc_request &r)
{
    const size_t n = pos(r.n);
    const size_t p = (size_t)req_order(r);
    const size_t nodes = 4 * (15 * n / (size_t)req_leaf(r) + 1) + 1;
    const size_t per_node = 64 + 16 * p + 32 + 4 + (r.precision ? 8 * (p + 1) : 0);
    return n * (4 + 4 + 32 + (r.precision ? 16 : 0)) + nodes * per_node + (64u << 10);
}

// Upper estimate of the buffers a slot's context needs for a request besides
// the tree: host-array staging (io), sort scratch of sources and targets, and
// for the force its near-field lists and outputs.
size_t context_estimate(const tc_request &r, bool build)
{
    const size_t n = build ? pos(r.n) : 0;
    const size_t nodes = build ? 4 * (15 * n / (size_t)req_leaf(r) + 1) + 1 : 0;
    size_t b = n * (24 + 4 + 4 + 32 + 8) + nodes * 64 + (2u << 20);
    if (r.kind == TC_REQ_POTENTIAL) {
        const size_t m = pos(r.m);
        b += m * (24 + 4 + 4 + 16 + 8 + 8);
    } else {
        const size_t nb = pos(r.nblocks), bricks = 16 * nb;
        const size_t maxc = std::max<size_t>(4096, std::min<size_t>(bricks, 1u << 20) / 4);
        b += nb * (24 + 16 * 1024) + maxc * (64 * 16 + 68 * 4) + bricks * 16;
    }
    return b;
}

struct TreeEntry {
    long long source_id = 0;
    int kind = 0, precision = 0, order = 0, leaf = 0;
    TcTree *tree = nullptr;
    hipEvent_t ready = nullptr;
    int users = 0;        // requests reading or building it
    bool built = false;   // the build is enqueued and `ready` recorded after it
    unsigned long long last_use = 0;
    size_t bytes = 0;     // device bytes of its buffers (growth while building is reserved)
};

// Buffers of an evicted tree kept for the next build (no free, no allocation
// unless the next tree is larger).
struct SpareTree {
    TcTree *tree;
    size_t bytes;
};

struct Job {
    long long ticket = 0;
    tc_request r{};
    double t_submit = 0, t_start = 0, t_done = 0;
    int status = 0;
    int slot = -1;
    bool reused = false;
    bool done = false;
    size_t reserve = 0;  // admitted growth
    TreeEntry *entry = nullptr;
    std::string err;
};

// A potential traversal waiting for the batch leader.
struct BatchReq {
    TcPotBatchItem item;
    int order = 0, precision = 0;
    hipStream_t stream = nullptr;                // the owning slot's stream
    hipEvent_t ready = nullptr, done = nullptr;  // the owning slot's events
    bool launched = false;
    bool wait_done = false;  // launched on another slot's stream: wait for `done`
    int status = 0;
    std::string err;
};

}  // namespace

struct tc_server {
    int device = 0;
    int flags = 0;
    size_t budget = 0, reserved = 0, peak = 0;
    std::vector<tc_context *> ctx;
    std::vector<size_t> slot_bytes;
    std::vector<bool> slot_busy;
    std::vector<hipEvent_t> slot_ready, slot_done;
    std::vector<TreeEntry *> cache;
    std::vector<SpareTree> spare;
    size_t max_trees = 0;  // cached trees kept; beyond, the least recently used is recycled
    size_t freeing = 0;    // bytes evicted but not yet freed (still held)
    unsigned long long tick = 0;
    std::mutex mu;
    std::condition_variable cv_work, cv_done;
    std::deque<Job *> queue;
    std::unordered_map<long long, Job *> jobs;
    std::vector<std::thread> workers;
    bool stop = false;
    long long next_ticket = 1;
    long long completed = 0, reused = 0, failed = 0, evictions = 0, trims = 0;
    long long est_miss = 0;  // requests that took more than their estimate
    size_t est_over = 0;     // the largest excess
    int inflight = 0, max_inflight = 0;
    // batching
    std::mutex bmu;
    std::condition_variable cv_batch;
    std::deque<BatchReq *> bq;
    bool leader = false;
    TcBatchWs bws;
    long long batches = 0, batched = 0, merged = 0;
    // frees off the request path (their bytes are in `freeing` until done)
    std::mutex rmu;
    std::condition_variable cv_reap;
    std::vector<TcTree *> reap_trees;
    std::vector<TcDeferredFree> reap_bufs;
    size_t reap_bytes = 0;  // bytes of the queued trees and buffers
    bool reap_stop = false;
    std::thread reaper;
};

namespace {

size_t held(const tc_server *s)
{
    size_t b = 0;
    for (size_t x : s->slot_bytes)
        b += x;
    for (const TreeEntry *e : s->cache)
        b += e->bytes;
    for (const SpareTree &t : s->spare)
        b += t.bytes;
    return b + s->freeing;
}

bool matches(const TreeEntry *e, const tc_request &r)
{
    return r.source_id != 0 && e->source_id == r.source_id && e->kind == r.kind && e->precision == r.precision &&
           e->order == req_order(r) && e->leaf == req_leaf(r);
}

TreeEntry *lookup(tc_server *s, const tc_request &r)
{
    for (TreeEntry *e : s->cache)
        if (e->built && matches(e, r))
            return e;
    return nullptr;
}

// Another slot is building this request's tree: the request waits for it.
bool building(const tc_server *s, const tc_request &r)
{
    for (const TreeEntry *e : s->cache)
        if (!e->built && matches(e, r))
            return true;
    return false;
}

// Frees a tree's buffers (hipFree waits for the whole device: never called
// with s->mu held while requests run).
void free_tree(tc_server *s, TcTree *t)
{
    (void)hipSetDevice(s->device);
    t->release();
    delete t;
}

// Removes an idle entry from the cache; its tree goes to the caller.  No
// request uses the entry, and each finished request synchronized its streams,
// so no kernel still reads the tree.
TcTree *drop_entry(tc_server *s, TreeEntry *e)
{
    s->cache.erase(std::find(s->cache.begin(), s->cache.end(), e));
    TcTree *t = e->tree;
    if (e->ready)
        (void)hipEventDestroy(e->ready);
    delete e;
    return t;
}

// The least recently used idle cached tree, or nullptr.
TreeEntry *lru_idle(tc_server *s)
{
    TreeEntry *victim = nullptr;
    for (TreeEntry *e : s->cache)
        if (e->users == 0 && e->built && (!victim || e->last_use < victim->last_use))
            victim = e;
    return victim;
}

// Growth a job may add to what the server holds.  A slot's buffers grow one
// by one (each to the largest size any of its requests needed), so a slot
// already holding large buffers of one kind may still grow others: the bound
// is the request's whole estimate, not its excess over the slot's total.
size_t job_need(const tc_request &r, bool hit)
{
    return context_estimate(r, !hit) + (hit ? 0 : tree_estimate(r));
}

// Makes room for `need` more bytes: spare tree buffers first, then idle
// cached trees (least recently used), then the buffers of idle slots; takes
// nothing when even all of that would not be enough.  Evicted trees are
// appended to `victims`, their bytes counted in s->freeing until the caller
// frees them.  Caller holds s->mu.
bool make_room(tc_server *s, size_t need, std::vector<TcTree *> &victims, size_t &mine)
{
    auto fits = [&] { return held(s) - mine + s->reserved + need <= s->budget; };
    if (fits())
        return true;
    size_t idle = 0;
    for (const TreeEntry *e : s->cache)
        if (e->users == 0 && e->built)
            idle += e->bytes;
    for (const SpareTree &t : s->spare)
        idle += t.bytes;
    for (size_t i = 0; i < s->ctx.size(); ++i)
        if (!s->slot_busy[i])
            idle += s->slot_bytes[i];
    if (held(s) - mine - idle + s->reserved + need > s->budget)
        return false;
    while (!fits() && !s->spare.empty()) {
        victims.push_back(s->spare.back().tree);
        s->freeing += s->spare.back().bytes;
        mine += s->spare.back().bytes;
        s->spare.pop_back();
    }
    while (!fits()) {
        TreeEntry *victim = lru_idle(s);
        if (!victim)
            break;
        const size_t b = victim->bytes;
        victims.push_back(drop_entry(s, victim));
        s->freeing += b;
        mine += b;
        s->evictions++;
    }
    for (size_t i = 0; i < s->ctx.size() && !fits(); ++i)
        if (!s->slot_busy[i] && s->slot_bytes[i]) {
            tc_srv_context_trim(s->ctx[i]);
            s->slot_bytes[i] = 0;
            s->trims++;
        }
    return fits();
}

// Keeps the larger of a tree about to leave the cache and the smallest spare
// when the spare pool is full: the other goes to victims.  Caller holds s->mu.
void spare_or_victim(tc_server *s, TcTree *t, size_t b, std::vector<TcTree *> &victims, size_t &mine)
{
    if (s->spare.size() < s->ctx.size()) {
        s->spare.push_back({t, b});
        return;
    }
    auto small = std::min_element(s->spare.begin(), s->spare.end(),
                                  [](const SpareTree &x, const SpareTree &y) { return x.bytes < y.bytes; });
    if (small != s->spare.end() && small->bytes < b) {
        std::swap(small->tree, t);
        std::swap(small->bytes, b);
    }
    victims.push_back(t);
    s->freeing += b;
    mine += b;
}

// Keeps at most s->max_trees cached trees: the least recently used idle ones
// beyond go to the spare buffers (at most one per slot, the larger trees
// kept; the rest go to victims, which the caller hands to the reaper).
// Caller holds s->mu.
void enforce_cap(tc_server *s, std::vector<TcTree *> &victims, size_t &mine)
{
    size_t built = 0;
    for (const TreeEntry *e : s->cache)
        built += e->built;
    while (built > s->max_trees) {
        TreeEntry *victim = lru_idle(s);
        if (!victim)
            break;
        const size_t b = victim->bytes;
        TcTree *t = drop_entry(s, victim);
        s->evictions++;
        --built;
        spare_or_victim(s, t, b, victims, mine);
    }
}

// Hands trees evicted by the cache limit (no reservation counts on their
// bytes, which stay in s->freeing) and a request's old buffers to the reaper.
// Caller holds s->mu.
void reap(tc_server *s, std::vector<TcTree *> &victims, size_t &mine, std::vector<TcDeferredFree> *bufs = nullptr)
{
    size_t bb = 0;
    if (bufs)
        for (const TcDeferredFree &f : *bufs)
            bb += f.bytes;
    if (victims.empty() && !bb)
        return;
    s->freeing += bb;
    {
        std::lock_guard<std::mutex> rl(s->rmu);
        s->reap_trees.insert(s->reap_trees.end(), victims.begin(), victims.end());
        if (bufs)
            s->reap_bufs.insert(s->reap_bufs.end(), bufs->begin(), bufs->end());
        s->reap_bytes += mine + bb;
    }
    s->cv_reap.notify_one();
    victims.clear();
    mine = 0;
    if (bufs)
        bufs->clear();
}

void free_victims(tc_server *s, std::unique_lock<std::mutex> &, std::vector<TcTree *> &victims, size_t &mine)
{
    reap(s, victims, mine);
}

// The reaper thread: frees what reap() queued, then stops counting its bytes.
void reaper(tc_server *s)
{
    (void)hipSetDevice(s->device);
    std::unique_lock<std::mutex> rl(s->rmu);
    while (true) {
        s->cv_reap.wait(rl, [&] { return s->reap_stop || !s->reap_trees.empty() || !s->reap_bufs.empty(); });
        if (s->reap_trees.empty() && s->reap_bufs.empty() && s->reap_stop)
            return;
        std::vector<TcTree *> trees;
        std::vector<TcDeferredFree> bufs;
        trees.swap(s->reap_trees);
        bufs.swap(s->reap_bufs);
        const size_t bytes = s->reap_bytes;
        s->reap_bytes = 0;
        rl.unlock();
        for (TcTree *t : trees)
            free_tree(s, t);
        for (const TcDeferredFree &f : bufs)
            (void)hipFree(f.ptr);
        {
            std::lock_guard<std::mutex> lk(s->mu);
            s->freeing -= bytes;
            s->cv_work.notify_all();
        }
        rl.lock();
    }
}


// Batch leader: takes all waiting traversals of the same order and precision
// as the first one (up to TcBatchWs::MAXREQ).  One alone is launched as the
// one-request kernel on its owner's stream (idle: its thread is waiting
// here), so it costs nothing over an unbatched request.  Several are launched
// as one k_potential_batch on the first one's stream, after the others'
// `ready` events; the others then wait on their `done` events.  Launches of
// successive batches go to different streams and overlap.
void lead_batches(tc_server *s, std::unique_lock<std::mutex> &lk)
{
    while (!s->bq.empty()) {
        const int order = s->bq.front()->order, prec = s->bq.front()->precision;
        std::vector<BatchReq *> take;
        for (auto it = s->bq.begin(); it != s->bq.end() && (int)take.size() < TcBatchWs::MAXREQ;)
            if ((*it)->order == order && (*it)->precision == prec) {
                take.push_back(*it);
                it = s->bq.erase(it);
            } else {
                ++it;
            }
        lk.unlock();
        int status = 0;
        std::string err;
        try {
            TC_HIP(hipSetDevice(s->device));
            if (take.size() == 1 && !(s->flags & TC_SERVER_BATCH_ALWAYS)) {
                tc_potential_item(take[0]->item, prec, nullptr, take[0]->stream);
            } else {
                hipStream_t st = take[0]->stream;
                std::vector<TcPotBatchItem> items;
                for (BatchReq *b : take) {
                    if (b != take[0])
                        TC_HIP(hipStreamWaitEvent(st, b->ready, 0));
                    items.push_back(b->item);
                }
                tc_potential_batch(items.data(), (int)items.size(), order, prec, s->bws, nullptr, st, true);
                for (BatchReq *b : take)
                    if (b != take[0]) {
                        TC_HIP(hipEventRecord(b->done, st));
                        b->wait_done = true;
                    }
            }
        } catch (const TcError &e) {
            status = e.code;
            err = e.what();
        } catch (const std::exception &e) {
            status = -2;
            err = e.what();
        }
        lk.lock();
        s->batches++;
        s->batched += (long long)take.size();
        s->merged += take.size() > 1 ? (long long)take.size() : 0;
        for (BatchReq *b : take) {
            b->status = status;
            b->err = err;
            b->launched = true;
        }
        s->cv_batch.notify_all();
    }
}

// The traversal hand-off of one request (a slot thread).
void batch_eval(tc_server *s, int slot, int order, int prec, const TcPotBatchItem &it, hipStream_t st)
{
    BatchReq br;
    br.item = it;
    br.order = order;
    br.precision = prec;
    br.stream = st;
    br.ready = s->slot_ready[slot];
    br.done = s->slot_done[slot];
    TC_HIP(hipEventRecord(br.ready, st));
    std::unique_lock<std::mutex> lk(s->bmu);
    s->bq.push_back(&br);
    while (!br.launched) {
        if (!s->leader) {
            s->leader = true;
            lead_batches(s, lk);
            s->leader = false;
            s->cv_batch.notify_all();
        } else {
            s->cv_batch.wait(lk);
        }
    }
    lk.unlock();
    if (br.status)
        throw TcError(br.status, "treecode server: batched traversal failed: " + br.err);
    if (br.wait_done)
        TC_HIP(hipStreamWaitEvent(st, br.done, 0));
}

void run_job(tc_server *s, int slot, Job *j, bool build)
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
        std::function<void(const TcPotBatchItem &, hipStream_t)> ev = [&](const TcPotBatchItem &it, hipStream_t st) {
            batch_eval(s, slot, req_order(r), r.precision, it, st);
        };
        const bool batch = pot && (s->flags & (TC_SERVER_BATCH | TC_SERVER_BATCH_ALWAYS));
        rc = tc_srv_solve(c, r, *j->entry->tree, build, j->entry->ready, batch ? &ev : nullptr);
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
        TreeEntry *hit = nullptr;
        size_t need = 0, mine = 0;
        std::vector<TcTree *> victims;
        // cached trees first, then arrival order; each must fit the budget
        for (int pass = 0; pass < 2 && !pick; ++pass)
            for (auto it = s->queue.begin(); it != s->queue.end() && !pick; ++it) {
                TreeEntry *e = lookup(s, (*it)->r);
                if (pass == 0 && !e)
                    continue;
                if (!e && building(s, (*it)->r))
                    continue;
                const size_t nd = job_need((*it)->r, e != nullptr);
                // the candidate's own tree is not evicted to make room for it
                if (e)
                    e->users++;
                // alone, a request always fits once other threads' frees are done
                if (make_room(s, nd, victims, mine) || (s->inflight == 0 && s->freeing == mine)) {
                    pick = *it;
                    hit = e;
                    need = nd;
                } else if (e) {
                    e->users--;
                }
            }
        if (!pick) {
            free_victims(s, lk, victims, mine);
            if (s->stop)
                return;
            s->cv_work.wait(lk);
            continue;
        }
        s->queue.erase(std::find(s->queue.begin(), s->queue.end(), pick));
        const bool build = hit == nullptr;
        if (build) {
            hit = new TreeEntry();
            hit->source_id = pick->r.source_id;
            hit->kind = pick->r.kind;
            hit->precision = pick->r.precision;
            hit->order = req_order(pick->r);
            hit->leaf = req_leaf(pick->r);
            // the smallest spare buffers that hold the tree's estimate, else
            // the largest spare ones, else new ones
            if (!s->spare.empty()) {
                const size_t est = tree_estimate(pick->r);
                auto fit = s->spare.end(), big = s->spare.begin();
                for (auto it = s->spare.begin(); it != s->spare.end(); ++it) {
                    if (it->bytes >= est && (fit == s->spare.end() || it->bytes < fit->bytes))
                        fit = it;
                    if (it->bytes > big->bytes)
                        big = it;
                }
                if (fit == s->spare.end())
                    fit = big;
                hit->tree = fit->tree;
                hit->bytes = fit->bytes;
                s->spare.erase(fit);
            } else {
                hit->tree = new TcTree();
            }
            (void)hipSetDevice(s->device);
            if (hipEventCreateWithFlags(&hit->ready, hipEventDisableTiming) != hipSuccess)
                hit->ready = nullptr;
            s->cache.push_back(hit);
            hit->users++;
        }
        hit->last_use = ++s->tick;
        pick->entry = hit;
        pick->reused = !build;
        pick->slot = slot;
        pick->reserve = need;
        pick->t_start = now_ms();
        s->reserved += need;
        s->slot_busy[slot] = true;
        s->inflight++;
        s->max_inflight = std::max(s->max_inflight, s->inflight);
        const size_t tree_before = hit->bytes;
        // this request's evictions (memory pressure only) are freed before it
        // allocates, under the lock: freed outside it, their bytes would count
        // both as held and as room this request's reservation already took
        for (TcTree *t : victims)
            free_tree(s, t);
        victims.clear();
        s->freeing -= mine;
        mine = 0;
        s->peak = std::max(s->peak, held(s) + s->reserved);
        lk.unlock();
        // grown buffers' old memory is freed by the reaper after this request
        std::vector<TcDeferredFree> old_bufs;
        if (!hit->ready) {
            pick->status = -2;
            pick->err = "treecode server: event creation failed";
        } else {
            tc_defer_free = &old_bufs;
            run_job(s, slot, pick, build);
            tc_defer_free = nullptr;
        }
        const size_t cbytes = tc_srv_context_bytes(s->ctx[slot]);
        // a failed build's kernels end before its buffers are reused, and no
        // kernel of this request may still read a buffer handed to the reaper
        if ((build && pick->status != 0) || !old_bufs.empty())
            (void)tc_synchronize(s->ctx[slot]);
        const size_t tbytes = build ? tc_tree_bytes(*hit->tree) : 0;
        lk.lock();
        // the estimates admission relies on must bound what the request took
        const size_t grew = cbytes > s->slot_bytes[slot] ? cbytes - s->slot_bytes[slot] : 0;
        const size_t tgrew = tbytes > tree_before ? tbytes - tree_before : 0;
        const size_t cest = context_estimate(pick->r, build), test = build ? tree_estimate(pick->r) : 0;
        if (grew > cest || tgrew > test) {
            s->est_miss++;
            s->est_over = std::max(s->est_over, std::max(grew > cest ? grew - cest : 0, tgrew > test ? tgrew - test : 0));
        }
        s->reserved -= pick->reserve;
        s->slot_busy[slot] = false;
        s->slot_bytes[slot] = cbytes;
        hit->users--;
        if (build) {
            hit->bytes = tbytes;
            // a request without sources leaves no tree worth keeping
            if (pick->status == 0 && hit->tree->n > 0 && pick->r.source_id != 0) {
                hit->built = true;
                enforce_cap(s, victims, mine);
            } else {
                TcTree *t = drop_entry(s, hit);
                spare_or_victim(s, t, tbytes, victims, mine);
            }
        }
        s->peak = std::max(s->peak, held(s) + s->reserved);
        s->inflight--;
        if (pick->status == 0) {
            s->completed++;
            s->reused += pick->reused;
        } else {
            s->failed++;
        }
        pick->entry = nullptr;
        pick->t_done = now_ms();
        pick->done = true;
        s->cv_done.notify_all();
        s->cv_work.notify_all();
        reap(s, victims, mine, &old_bufs);
    }
}

}  // namespace

extern "C" {

TC_API tc_server *tc_server_create_ex(int device, int nslots, double hbm_budget_gib, int flags)
{
    if (nslots < 1 || nslots > 64 || !(hbm_budget_gib > 0))
        return nullptr;
    tc_server *s = new tc_server();
    s->device = device;
    s->flags = flags;
    s->budget = (size_t)(hbm_budget_gib * 1073741824.0);
    s->max_trees = 4 * (size_t)nslots;
    bool ok = true;
    for (int i = 0; i < nslots && ok; ++i) {
        tc_context *c = tc_context_create(device);
        if (!c) {
            ok = false;
            break;
        }
        s->ctx.push_back(c);
        s->slot_bytes.push_back(0);
        s->slot_busy.push_back(false);
        hipEvent_t a = nullptr, b = nullptr;
        ok = hipEventCreateWithFlags(&a, hipEventDisableTiming) == hipSuccess &&
             hipEventCreateWithFlags(&b, hipEventDisableTiming) == hipSuccess;
        s->slot_ready.push_back(a);
        s->slot_done.push_back(b);
    }
    if (!ok) {
        for (tc_context *o : s->ctx)
            tc_context_destroy(o);
        for (hipEvent_t e : s->slot_ready)
            if (e)
                (void)hipEventDestroy(e);
        for (hipEvent_t e : s->slot_done)
            if (e)
                (void)hipEventDestroy(e);
        delete s;
        return nullptr;
    }
    s->reaper = std::thread(reaper, s);
    for (int i = 0; i < nslots; ++i)
        s->workers.emplace_back(worker, s, i);
    return s;
}

TC_API tc_server *tc_server_create(int device, int nslots, double hbm_budget_gib)
{
    return tc_server_create_ex(device, nslots, hbm_budget_gib, 0);
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
    // the workers have handed everything over: the reaper frees it and ends
    {
        std::lock_guard<std::mutex> rl(s->rmu);
        s->reap_stop = true;
    }
    s->cv_reap.notify_one();
    if (s->reaper.joinable())
        s->reaper.join();
    for (auto &kv : s->jobs)
        delete kv.second;
    (void)hipSetDevice(s->device);
    s->bws.release();
    for (TreeEntry *e : s->cache) {
        if (e->ready)
            (void)hipEventDestroy(e->ready);
        free_tree(s, e->tree);
        delete e;
    }
    for (SpareTree &t : s->spare)
        free_tree(s, t.tree);
    for (tc_context *c : s->ctx)
        tc_context_destroy(c);
    for (hipEvent_t e : s->slot_ready)
        (void)hipEventDestroy(e);
    for (hipEvent_t e : s->slot_done)
        (void)hipEventDestroy(e);
    delete s;
}

TC_API long long tc_server_submit(tc_server *s, const tc_request *r)
{
    if (!s || !r)
        return -1;
    Job *j = new Job();
    j->r = *r;
    j->t_submit = now_ms();
    std::lock_guard<std::mutex> lk(s->mu);
    // a request that alone exceeds the budget can never run
    if (context_estimate(*r, true) + tree_estimate(*r) > s->budget) {
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
    if (st)
        tc_set_last_error(j->err);
    s->jobs.erase(it);
    delete j;
    return st;
}

TC_API int tc_server_stats_ex(tc_server *s, long long *out, int nout)
{
    if (!s || !out)
        return -1;
    std::lock_guard<std::mutex> lk(s->mu);
    long long batches = 0, batched = 0, merged = 0;
    {
        std::lock_guard<std::mutex> bl(s->bmu);
        batches = s->batches;
        batched = s->batched;
        merged = s->merged;
    }
    const long long v[] = {s->completed, s->reused, s->failed, s->max_inflight, (long long)s->peak,
                           (long long)s->budget, (long long)s->queue.size(), batches, batched, s->evictions, s->trims,
                           (long long)s->cache.size(), (long long)held(s), s->est_miss, (long long)s->est_over, merged};
    const int n = std::min<int>(nout, (int)(sizeof(v) / sizeof(v[0])));
    for (int i = 0; i < n; ++i)
        out[i] = v[i];
    return n;
}

TC_API int tc_server_stats(tc_server *s, long long *out)
{
    return tc_server_stats_ex(s, out, 7) == 7 ? 0 : -1;
}

}  // extern "C"
