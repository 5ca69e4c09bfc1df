// tc_api.cpp -- the C ABI: drop-in treecode_potential / treecode_force_mrag
// (drivers/treecode-potential.cpp:89-106, drivers/treecode-force.cpp:201-236)
// and the context-based extensions declared in include/treecodes2d.h.
//
// Unlike the reference (globals in upward.cpp:51-59, non-reentrant), all state
// lives in a tc_context: one HIP stream, grow-only HBM buffers and the tree
// that stays resident after a call.  The drop-in functions share one default
// context behind a mutex.
#include "tc_common.h"

#include "../../include/treecodes2d.h"
#ifdef TREECODES2D_FLOAT_DROPIN
// the float build exports the drop-in pair with the float signatures of the
// reference's real=float headers
#include "../../include/float/2d-treecodes-force.h"
#include "../../include/float/2d-treecodes-potential.h"
#endif

#include <cmath>
#include <cstdio>
#include <cstring>
#include <exception>
#include <functional>
#include <initializer_list>
#include <limits>
#include <mutex>
#include <string>
#include <thread>
#include <vector>

void tc_upload_binomials_eval(hipStream_t st);

enum TcKind { TC_KIND_NONE = 0, TC_KIND_POTENTIAL = 1, TC_KIND_FORCE = 2 };

constexpr size_t STAGE_BYTES = size_t(16) << 20;  // per staging buffer
constexpr size_t STAGE_MIN = size_t(4) << 20;     // smaller copies go direct

// Two pinned staging buffers and their events, so the CPU copy of one chunk
// overlaps the DMA of the other (allocated on first large copy).
struct TcStager {
    void *buf[2] = {nullptr, nullptr};
    hipEvent_t ev[2] = {nullptr, nullptr};
    bool busy[2] = {false, false};
    void ready()
    {
        if (buf[0])
            return;
        for (int b = 0; b < 2; ++b) {
            TC_HIP(hipHostMalloc(&buf[b], STAGE_BYTES, hipHostMallocDefault));
            TC_HIP(hipEventCreateWithFlags(&ev[b], hipEventDisableTiming));
        }
    }
    // Waits until buffer b is no longer read or written by the device.
    void wait(int b)
    {
        if (busy[b]) {
            TC_HIP(hipEventSynchronize(ev[b]));
            busy[b] = false;
        }
    }
    void release()
    {
        for (int b = 0; b < 2; ++b) {
            if (buf[b])
                (void)hipHostFree(buf[b]);
            if (ev[b])
                (void)hipEventDestroy(ev[b]);
            buf[b] = nullptr;
            ev[b] = nullptr;
            busy[b] = false;
        }
    }
};

struct tc_context {
    int device = 0;
    hipStream_t stream = nullptr;
    int precision = 0, pot_order = 12, pot_leaf = 32 * 16, force_order = 24, force_leaf = 192, force_chunks = 0;
    int pot_xcd_log2 = TC_XCD_CHUNK_LOG2, force_inline = 0;  // tiling options (0: the kernel's default)
    bool profile = false, count = false;
    int tree_kind = TC_KIND_NONE;
    int tree_precision = 0;
    TcTree tree;
    TcScratch ws;
    // potential targets: sorted in their own scratch on a second stream while
    // the source tree builds (both streams at the default priority)
    TcScratch tws;
    hipStream_t tstream = nullptr;
    hipEvent_t fork_ev = nullptr, join_ev = nullptr;
    // ends of the block ranges of a host-array force solve (their outputs are
    // copied back while the next range is evaluated)
    static constexpr int FORCE_RANGES = 4;
    hipEvent_t range_ev[FORCE_RANGES] = {};
    DevBuf io, cnt;
    hipEvent_t ev[4] = {nullptr, nullptr, nullptr, nullptr};
    // build phase boundaries: after the keys, the sorted records, the topology, the upward pass
    hipEvent_t bev[4] = {nullptr, nullptr, nullptr, nullptr};
    bool ev_valid = false, eval_has_tsort = false, bev_valid = false;
    unsigned long long counters[TC_NCOUNTERS] = {0};
    bool counters_on_host = false;  // the last call left its counters in `counters` (no-allocation path)
    // pinned staging for host arrays: one set on the context stream, one for
    // target uploads on the second stream (beside the source build)
    TcStager stg, tstg;
    // tc_potential_batch_resident: descriptors, pinned ring and graphs of its
    // launches; the target-sort descriptors of its small requests (pinned + device)
    TcBatchWs bws;
    void *sdesc_host = nullptr;
    DevBuf sdesc;
    // tc_potential_noalloc: tree, source and target scratch and counters carved
    // from the device's static scratch array (address looked up at creation)
    void *na_base = nullptr;
    TcArena na_arena;
    TcTree na_tree;
    TcScratch na_ws, na_tws;
    DevBuf na_cnt;
};

thread_local std::vector<TcDeferredFree> *tc_defer_free = nullptr;

namespace {

thread_local std::string g_last_error;
std::mutex g_init_mu;
bool g_binom_ready[64] = {false};
std::mutex g_default_mu;
tc_context *g_default = nullptr;

void set_error(const std::string &m) { g_last_error = m; }

template <class F> int guarded(tc_context *ctx, F &&f)
{
    if (!ctx) {
        set_error("treecode: null context");
        return -1;
    }
    try {
        TC_HIP(hipSetDevice(ctx->device));
        f();
        return 0;
    } catch (const TcError &e) {
        set_error(e.what());
        return e.code;
    } catch (const std::exception &e) {
        set_error(std::string("treecode: ") + e.what());
        return -9;
    }
}

void ensure_binomials(int device, hipStream_t st)
{
    std::lock_guard<std::mutex> lk(g_init_mu);
    if (device < 0 || device >= 64)
        throw TcError(-1, "treecode: device index out of range");
    if (!g_binom_ready[device]) {
        tc_upload_binomials(st);
        tc_upload_binomials_eval(st);
        g_binom_ready[device] = true;
    }
}

// memcpy split over up to 8 threads (pageable <-> pinned host memory).
void par_memcpy(void *dst, const void *src, size_t bytes)
{
    const unsigned hw = std::thread::hardware_concurrency();
    const size_t nt = bytes < (size_t(2) << 20) ? 1 : std::min<size_t>(8, hw ? hw : 1);
    if (nt <= 1) {
        std::memcpy(dst, src, bytes);
        return;
    }
    const size_t part = ((bytes + nt - 1) / nt + 63) & ~size_t(63);
    std::vector<std::thread> th;
    for (size_t t = 1; t < nt; ++t) {
        const size_t o = t * part;
        if (o >= bytes)
            break;
        th.emplace_back([=] { std::memcpy((char *)dst + o, (const char *)src + o, std::min(part, bytes - o)); });
    }
    std::memcpy(dst, src, std::min(part, bytes));
    for (auto &t : th)
        t.join();
}

// Host array -> device, ordered on stream st.  Large copies go through the
// pinned staging buffers of sg in 16 MiB chunks: the CPU fills one buffer
// while the DMA engine drains the other.
void copy_in(TcStager &sg, hipStream_t st, const double *h, double *d, long long n)
{
    const size_t bytes = sizeof(double) * (size_t)(n > 0 ? n : 0);
    if (!bytes)
        return;
    if (bytes < STAGE_MIN) {
        TC_HIP(hipMemcpyAsync(d, h, bytes, hipMemcpyHostToDevice, st));
        return;
    }
    sg.ready();
    int b = 0;
    for (size_t o = 0; o < bytes; o += STAGE_BYTES, b ^= 1) {
        const size_t len = std::min(STAGE_BYTES, bytes - o);
        sg.wait(b);
        par_memcpy(sg.buf[b], (const char *)h + o, len);
        TC_HIP(hipMemcpyAsync((char *)d + o, sg.buf[b], len, hipMemcpyHostToDevice, st));
        TC_HIP(hipEventRecord(sg.ev[b], st));
        sg.busy[b] = true;
    }
}

// Device -> host array through the staging buffers of sg on stream st;
// returns with the data in h (synchronises st).
void copy_out(TcStager &sg, hipStream_t st, const double *d, double *h, long long n)
{
    const size_t bytes = sizeof(double) * (size_t)(n > 0 ? n : 0);
    if (bytes < STAGE_MIN) {
        if (bytes)
            TC_HIP(hipMemcpyAsync(h, d, bytes, hipMemcpyDeviceToHost, st));
        TC_HIP(hipStreamSynchronize(st));
        return;
    }
    sg.ready();
    size_t pend_o[2] = {0, 0}, pend_len[2] = {0, 0};
    int b = 0;
    auto drain = [&](int k) {
        if (pend_len[k]) {
            sg.wait(k);
            par_memcpy((char *)h + pend_o[k], sg.buf[k], pend_len[k]);
            pend_len[k] = 0;
        }
    };
    for (size_t o = 0; o < bytes; o += STAGE_BYTES, b ^= 1) {
        const size_t len = std::min(STAGE_BYTES, bytes - o);
        drain(b);
        sg.wait(b);
        TC_HIP(hipMemcpyAsync(sg.buf[b], (const char *)d + o, len, hipMemcpyDeviceToHost, st));
        TC_HIP(hipEventRecord(sg.ev[b], st));
        sg.busy[b] = true;
        pend_o[b] = o;
        pend_len[b] = len;
    }
    drain(b);
    drain(b ^ 1);
    TC_HIP(hipStreamSynchronize(st));
}

void copy_out(tc_context *c, const double *d, double *h, long long n) { copy_out(c->stg, c->stream, d, h, n); }

struct HostIn {
    const double *h;
    long long n;
};
struct HostOut {
    double *h;
    long long n;
};

inline size_t arr_bytes(long long n) { return sizeof(double) * (size_t)(n > 0 ? n : 0); }

// Host arrays whose device copies lie back to back from d (in the io buffer).
// Together up to one staging buffer, they are packed into pinned memory and
// sent with one DMA (a copy from pageable memory costs a staged transfer of
// its own); larger sets go array by array.
void copy_in_all(tc_context *c, double *d, std::initializer_list<HostIn> arrs)
{
    size_t total = 0;
    for (const HostIn &a : arrs)
        total += arr_bytes(a.n);
    if (!total)
        return;
    if (total > STAGE_BYTES) {
        for (const HostIn &a : arrs) {
            copy_in(c->stg, c->stream, a.h, d, a.n);
            d += a.n > 0 ? a.n : 0;
        }
        return;
    }
    c->stg.ready();
    c->stg.wait(0);
    char *s = static_cast<char *>(c->stg.buf[0]);
    for (const HostIn &a : arrs) {
        par_memcpy(s, a.h, arr_bytes(a.n));
        s += arr_bytes(a.n);
    }
    TC_HIP(hipMemcpyAsync(d, c->stg.buf[0], total, hipMemcpyHostToDevice, c->stream));
    TC_HIP(hipEventRecord(c->stg.ev[0], c->stream));
    c->stg.busy[0] = true;
}

// Device block from d (arrays back to back) -> host arrays; returns with the
// data in place (synchronises the stream).  Up to one staging buffer in total
// it is one DMA into pinned memory.
void copy_out_all(tc_context *c, const double *d, std::initializer_list<HostOut> arrs)
{
    size_t total = 0;
    for (const HostOut &a : arrs)
        total += arr_bytes(a.n);
    if (total > STAGE_BYTES || !total) {
        for (const HostOut &a : arrs) {
            copy_out(c, d, a.h, a.n);
            d += a.n > 0 ? a.n : 0;
        }
        TC_HIP(hipStreamSynchronize(c->stream));
        return;
    }
    c->stg.ready();
    c->stg.wait(0);
    TC_HIP(hipMemcpyAsync(c->stg.buf[0], d, total, hipMemcpyDeviceToHost, c->stream));
    TC_HIP(hipStreamSynchronize(c->stream));
    const char *s = static_cast<const char *>(c->stg.buf[0]);
    for (const HostOut &a : arrs) {
        par_memcpy(a.h, s, arr_bytes(a.n));
        s += arr_bytes(a.n);
    }
}

unsigned long long *prepare_counters(tc_context *c)
{
    c->counters_on_host = false;
    if (!c->count)
        return nullptr;
    c->cnt.reserve(sizeof(unsigned long long) * TC_NCOUNTERS);
    tc_fill_u32(c->cnt.as<uint32_t>(), 0u, 2 * TC_NCOUNTERS, c->stream);
    return c->cnt.as<unsigned long long>();
}

// All counters of the last call (zeros when counting was off).
void read_counters(tc_context *c, unsigned long long *all)
{
    if (c->count && c->counters_on_host)
        std::memcpy(all, c->counters, sizeof(unsigned long long) * TC_NCOUNTERS);
    else if (c->count && c->cnt.ptr)
        TC_HIP(hipMemcpy(all, c->cnt.ptr, sizeof(unsigned long long) * TC_NCOUNTERS, hipMemcpyDeviceToHost));
    else
        for (int i = 0; i < TC_NCOUNTERS; ++i)
            all[i] = 0;
}

void record(tc_context *c, int i)
{
    if (c->profile)
        TC_HIP(hipEventRecord(c->ev[i], c->stream));
}

// A tree outside the context (the request server's shared cache): built into
// by this call (build), or built earlier by another context, whose stream
// recorded `ready` after the build.  eval, when set, replaces the one-request
// traversal launch of a potential solve (the server's batcher).
struct ExtTree {
    TcTree *tree = nullptr;
    bool build = false;
    hipEvent_t ready = nullptr;
    const std::function<void(const TcPotBatchItem &, hipStream_t)> *eval = nullptr;
};

// at_census (optional): work to queue once the build's census readback is
// queued (see tc_build_tree).
void build(tc_context *c, const double *xs, const double *ys, const double *vs, int n, int kind, const ExtTree &ext,
           const std::function<void()> *at_census = nullptr)
{
    if (ext.tree) {
        const int order = kind == TC_KIND_POTENTIAL ? c->pot_order : c->force_order;
        const int leaf = kind == TC_KIND_POTENTIAL ? c->pot_leaf : c->force_leaf;
        ext.tree->nnodes = 0;
        ext.tree->has_fp32 = false;
        tc_build_tree(*ext.tree, xs, ys, vs, n, leaf, order, c->precision == 1, kind == TC_KIND_FORCE, c->ws,
                      c->stream, nullptr, at_census);
        if (ext.ready)
            TC_HIP(hipEventRecord(ext.ready, c->stream));
        return;
    }

    const int order = kind == TC_KIND_POTENTIAL ? c->pot_order : c->force_order;
    const int leaf = kind == TC_KIND_POTENTIAL ? c->pot_leaf : c->force_leaf;
    // the resident tree is invalid from the first array tc_build_tree
    // overwrites until it returns: a build that throws part-way (non-finite
    // source, failed allocation) leaves no tree a resident call could use
    c->tree_kind = TC_KIND_NONE;
    c->tree.nnodes = 0;
    c->tree.has_fp32 = false;
    c->bev_valid = false;
    tc_build_tree(c->tree, xs, ys, vs, n, leaf, order, c->precision == 1, kind == TC_KIND_FORCE, c->ws, c->stream,
                  c->profile ? c->bev : nullptr, at_census);
    c->bev_valid = c->profile;
    c->tree_kind = kind;
    c->tree_precision = c->precision;
}

void check_resident(tc_context *c, int kind)
{
    const int order = kind == TC_KIND_POTENTIAL ? c->pot_order : c->force_order;
    const int leaf = kind == TC_KIND_POTENTIAL ? c->pot_leaf : c->force_leaf;
    if (c->tree_kind == TC_KIND_NONE || c->tree.n <= 0)
        throw TcError(-5, "treecode: no resident tree in this context");
    if (c->tree_kind != kind)
        throw TcError(-5, "treecode: resident tree was built for the other kernel (potential vs force)");
    if (c->tree.order != order || c->tree.leafcap != leaf || (c->precision == 1 && !c->tree.has_fp32))
        throw TcError(-5, "treecode: resident tree was built with other order/leaf/precision");
}

void eval_potential(tc_context *c, double theta, const double *xt, const double *yt, int m, double *out,
                    unsigned long long *cnt, bool sort_targets)
{
    tc_eval_potential(c->tree, theta, xt, yt, m, out, c->precision, cnt, c->tws, c->stream,
                      c->profile ? c->ev[2] : nullptr, c->profile ? c->ev[3] : nullptr, sort_targets,
                      c->pot_xcd_log2);
}

// An external tree's evaluation: the target sort (unless done beside the
// build) and the traversal, or the batcher's launch in place of the latter.
void eval_potential_ext(tc_context *c, const TcTree &t, const ExtTree &ext, double theta, const double *xt,
                        const double *yt, int m, double *out, unsigned long long *cnt, bool sort_targets)
{
    if (!ext.eval) {
        tc_eval_potential(t, theta, xt, yt, m, out, c->precision, cnt, c->tws, c->stream,
                          c->profile ? c->ev[2] : nullptr, c->profile ? c->ev[3] : nullptr, sort_targets,
                          c->pot_xcd_log2);
        return;
    }
    if (sort_targets)
        tc_sort_targets(xt, yt, m, c->tws, c->stream);
    record(c, 2);
    TcPotBatchItem it;
    it.tree = &t;
    it.theta = theta;
    it.xt = xt;
    it.yt = yt;
    it.tperm = c->tws.tperm.as<uint32_t>();
    it.m = m;
    it.out = out;
    (*ext.eval)(it, c->stream);
    record(c, 3);
}

int run_potential(tc_context *ctx, double theta, const double *xs, const double *ys, const double *vs, int n,
                  const double *xt, const double *yt, int m, double *out, int device_ptrs, bool rebuild,
                  const ExtTree &ext = ExtTree())
{
    return guarded(ctx, [&] {
        tc_context *c = ctx;
        c->ev_valid = false;
        if (m <= 0)
            return;
        if (rebuild && n <= 0) {
            if (device_ptrs)
                TC_HIP(hipMemsetAsync(out, 0, sizeof(double) * (size_t)m, c->stream));
            else
                std::memset(out, 0, sizeof(double) * (size_t)m);
            c->tree_kind = TC_KIND_NONE;
            return;
        }
        if (!rebuild && !ext.tree) {
            check_resident(c, TC_KIND_POTENTIAL);
            c->bev_valid = false;
        }
        if (!rebuild && ext.tree && ext.ready)
            TC_HIP(hipStreamWaitEvent(c->stream, ext.ready, 0));
        const double *dxs = xs, *dys = ys, *dvs = vs, *dxt = xt, *dyt = yt;
        double *dout = out;
        // host target arrays of 4 MiB or more (512 Ki targets) are uploaded by a
        // helper thread, through their own staging buffers on the second
        // stream, while the source tree builds (the build waits on the host only
        // for its census): the upload then overlaps the build instead of
        // preceding it
        const bool overlap = !device_ptrs && rebuild && arr_bytes(m) >= STAGE_BYTES / 4;
        if (!device_ptrs) {
            const size_t ns = rebuild ? (size_t)n : 0;
            c->io.reserve(sizeof(double) * (3 * ns + 3 * (size_t)m));
            double *base = c->io.as<double>();
            double *bxs = base, *bys = base + ns, *bvs = base + 2 * ns;
            double *bxt = base + 3 * ns, *byt = bxt + m;
            dout = byt + m;
            if (overlap)
                copy_in_all(c, bxs, {{xs, n}, {ys, n}, {vs, n}});
            else if (rebuild)
                copy_in_all(c, bxs, {{xs, n}, {ys, n}, {vs, n}, {xt, m}, {yt, m}});
            else
                copy_in_all(c, bxt, {{xt, m}, {yt, m}});
            dxs = bxs; dys = bys; dvs = bvs; dxt = bxt; dyt = byt;
        }
        unsigned long long *cnt = prepare_counters(c);
        if (rebuild) {
            // the target sort (one workgroup for small m) runs beside the
            // source build; the evaluation waits for both
            TC_HIP(hipEventRecord(c->fork_ev, c->stream));
            TC_HIP(hipStreamWaitEvent(c->tstream, c->fork_ev, 0));
            std::thread up;
            std::exception_ptr up_err;
            if (overlap) {
                up = std::thread([&] {
                    try {
                        TC_HIP(hipSetDevice(c->device));
                        copy_in(c->tstg, c->tstream, xt, const_cast<double *>(dxt), m);
                        copy_in(c->tstg, c->tstream, yt, const_cast<double *>(dyt), m);
                        tc_sort_targets(dxt, dyt, m, c->tws, c->tstream);
                    } catch (...) {
                        up_err = std::current_exception();
                    }
                    // whatever was queued stays ordered before later work on
                    // the context stream
                    (void)hipEventRecord(c->join_ev, c->tstream);
                });
            }
            // otherwise the target sort is queued once the source build's
            // census is, so it runs while the host waits for the census
            // readback (a gap in which a small solve's GPU had nothing to do)
            bool tsorted = false;
            const std::function<void()> tsort = [&] {
                tsorted = true;
                try {
                    tc_sort_targets(dxt, dyt, m, c->tws, c->tstream);
                } catch (...) {
                    // whatever part of the sort was queued stays ordered before
                    // later work on the context stream (the caller waits on join_ev)
                    (void)hipEventRecord(c->join_ev, c->tstream);
                    throw;
                }
                TC_HIP(hipEventRecord(c->join_ev, c->tstream));
            };
            record(c, 0);
            try {
                build(c, dxs, dys, dvs, n, TC_KIND_POTENTIAL, ext, overlap ? nullptr : &tsort);
                if (!overlap && !tsorted)
                    tsort();
            } catch (...) {
                // later work on the context stream (the next call's copies
                // into the io buffer) must not overtake the sort reading it
                if (up.joinable())
                    up.join();
                (void)hipStreamWaitEvent(c->stream, c->join_ev, 0);
                throw;
            }
            if (up.joinable())
                up.join();
            if (up_err) {
                (void)hipStreamWaitEvent(c->stream, c->join_ev, 0);
                std::rethrow_exception(up_err);
            }
            record(c, 1);
            TC_HIP(hipStreamWaitEvent(c->stream, c->join_ev, 0));
            if (ext.tree)
                eval_potential_ext(c, *ext.tree, ext, theta, dxt, dyt, m, dout, cnt, false);
            else
                eval_potential(c, theta, dxt, dyt, m, dout, cnt, false);
        } else {
            record(c, 0);
            record(c, 1);
            if (ext.tree)
                eval_potential_ext(c, *ext.tree, ext, theta, dxt, dyt, m, dout, cnt, true);
            else
                eval_potential(c, theta, dxt, dyt, m, dout, cnt, true);
        }
        c->ev_valid = c->profile;
        c->eval_has_tsort = true;
        if (!device_ptrs) {
            // the target sort's bounding-box pass flags non-finite coordinates
            uint32_t *bad = static_cast<uint32_t *>(c->tws.host_pinned());
            TC_HIP(hipMemcpyAsync(bad, tc_targets_bad_flag(c->tws, m), sizeof(uint32_t), hipMemcpyDeviceToHost,
                                  c->stream));
            copy_out_all(c, dout, {{out, m}});
            if (*bad & 2)
                throw TcError(-9, "treecode_potential: the target sort did not complete (a radix pass timed out)");
            if (*bad)
                throw TcError(-4, "treecode_potential: non-finite target coordinate");
        }
    });
}

#ifndef TREECODES2D_FLOAT_DROPIN
constexpr int NOALLOC_MAX_POINTS = 1 << 18;  // sources and targets of one tc_potential_noalloc call

// tc_potential with device pointers, every buffer carved from the static
// scratch: the same kernels in the same order as run_potential, so the results
// equal its results bit for bit.  Synchronous; on failure both streams are
// drained before returning, so no kernel still reads the scratch.
int run_potential_noalloc(tc_context *ctx, double theta, const double *xs, const double *ys, const double *vs,
                          int n, const double *xt, const double *yt, int m, double *out)
{
    return guarded(ctx, [&] {
        tc_context *c = ctx;
        c->ev_valid = false;
        if (!c->na_base)
            throw TcError(-1, "tc_potential_noalloc: no scratch address for this device");
        if (n > NOALLOC_MAX_POINTS || m > NOALLOC_MAX_POINTS)
            throw TcError(-3, "tc_potential_noalloc: at most " + std::to_string(NOALLOC_MAX_POINTS) +
                                  " sources and targets per call (got " + std::to_string(n) + " and " +
                                  std::to_string(m) + ")");
        if (m <= 0)
            return;
        if (n <= 0) {
            tc_fill_u32(reinterpret_cast<uint32_t *>(out), 0u, 2 * (size_t)m, c->stream);
            TC_HIP(hipStreamSynchronize(c->stream));
            return;
        }
        // every call lays its buffers out afresh from the start of the array
        c->na_arena.base = static_cast<char *>(c->na_base);
        c->na_arena.cap = tc_noalloc_scratch_bytes();
        c->na_arena.used = 0;
        c->na_tree.bind(&c->na_arena);
        c->na_ws.bind(&c->na_arena);
        c->na_tws.bind(&c->na_arena);
        c->na_cnt.bind(&c->na_arena);
        unsigned long long *cnt = nullptr;
        if (c->count) {
            c->na_cnt.reserve(sizeof(unsigned long long) * TC_NCOUNTERS);
            cnt = c->na_cnt.as<unsigned long long>();
            tc_fill_u32(c->na_cnt.as<uint32_t>(), 0u, 2 * TC_NCOUNTERS, c->stream);
        }
        bool forked = false;
        try {
            TC_HIP(hipEventRecord(c->fork_ev, c->stream));
            TC_HIP(hipStreamWaitEvent(c->tstream, c->fork_ev, 0));
            forked = true;
            tc_sort_targets(xt, yt, m, c->na_tws, c->tstream);
            TC_HIP(hipEventRecord(c->join_ev, c->tstream));
            record(c, 0);
            const bool f32 = c->precision == 1;
            tc_build_tree(c->na_tree, xs, ys, vs, n, c->pot_leaf, c->pot_order, f32, false, c->na_ws, c->stream);
            record(c, 1);
            TC_HIP(hipStreamWaitEvent(c->stream, c->join_ev, 0));
            tc_eval_potential(c->na_tree, theta, xt, yt, m, out, c->precision, cnt, c->na_tws, c->stream,
                              c->profile ? c->ev[2] : nullptr, c->profile ? c->ev[3] : nullptr, false,
                              c->pot_xcd_log2);
            TC_HIP(hipStreamSynchronize(c->stream));
        } catch (...) {
            if (forked)
                (void)hipStreamSynchronize(c->tstream);
            (void)hipStreamSynchronize(c->stream);
            throw;
        }
        c->ev_valid = c->profile;
        c->eval_has_tsort = true;
        if (cnt) {
            // the counters leave the scratch now (a later call reuses it)
            void *hp = c->na_tws.host_pinned();
            TC_HIP(hipMemcpyAsync(hp, cnt, sizeof(unsigned long long) * TC_NCOUNTERS, hipMemcpyDeviceToHost,
                                  c->stream));
            TC_HIP(hipStreamSynchronize(c->stream));
            std::memcpy(c->counters, hp, sizeof(unsigned long long) * TC_NCOUNTERS);
            c->counters_on_host = true;
        }
    });
}
#endif

int run_force(tc_context *ctx, double theta, const double *xs, const double *ys, const double *vs, int n,
              const double *x0s, const double *y0s, const double *hs, int nb, double *xr, double *yr,
              int device_ptrs, bool rebuild, const ExtTree &ext = ExtTree())
{
    return guarded(ctx, [&] {
        tc_context *c = ctx;
        c->ev_valid = false;
        if (nb <= 0)
            return;
        const size_t nt = (size_t)nb * 1024;
        if (rebuild && n <= 0) {
            if (device_ptrs) {
                TC_HIP(hipMemsetAsync(xr, 0, sizeof(double) * nt, c->stream));
                TC_HIP(hipMemsetAsync(yr, 0, sizeof(double) * nt, c->stream));
            } else {
                std::memset(xr, 0, sizeof(double) * nt);
                std::memset(yr, 0, sizeof(double) * nt);
            }
            c->tree_kind = TC_KIND_NONE;
            return;
        }
        if (!rebuild && !ext.tree) {
            check_resident(c, TC_KIND_FORCE);
            c->bev_valid = false;
        }
        if (!rebuild && ext.tree && ext.ready)
            TC_HIP(hipStreamWaitEvent(c->stream, ext.ready, 0));
        const double *dxs = xs, *dys = ys, *dvs = vs, *dx0 = x0s, *dy0 = y0s, *dh = hs;
        double *dxr = xr, *dyr = yr;
        if (!device_ptrs) {
            for (int i = 0; i < nb; ++i)
                if (!std::isfinite(x0s[i]) || !std::isfinite(y0s[i]) || !std::isfinite(hs[i]))
                    throw TcError(-4, "treecode_force_mrag: non-finite block origin or spacing");
            const size_t ns = rebuild ? (size_t)n : 0;
            c->io.reserve(sizeof(double) * (3 * ns + 3 * (size_t)nb + 2 * nt));
            double *base = c->io.as<double>();
            double *bxs = base, *bys = base + ns, *bvs = base + 2 * ns;
            double *bx0 = base + 3 * ns, *by0 = bx0 + nb, *bh = by0 + nb;
            dxr = bh + nb;
            dyr = dxr + nt;
            if (rebuild)
                copy_in_all(c, bxs, {{xs, n}, {ys, n}, {vs, n}, {x0s, nb}, {y0s, nb}, {hs, nb}});
            else
                copy_in_all(c, bx0, {{x0s, nb}, {y0s, nb}, {hs, nb}});
            dxs = bxs; dys = bys; dvs = bvs; dx0 = bx0; dy0 = by0; dh = bh;
        }
        unsigned long long *cnt = prepare_counters(c);
        record(c, 0);
        if (rebuild)
            build(c, dxs, dys, dvs, n, TC_KIND_FORCE, ext);
        record(c, 1);
        const TcTree &tree = ext.tree ? *ext.tree : c->tree;
        // Host outputs of at least FORCE_SPLIT_BLOCKS blocks: the blocks are
        // evaluated in up to FORCE_RANGES ranges of at least half that (each a
        // complete evaluation of its blocks, whose outputs are contiguous), and
        // a range's forces are copied back on the second stream, through its
        // own staging buffers, while the next range is evaluated.  Each range
        // gets the chunk pool the whole call would have.
        constexpr int FORCE_SPLIT_BLOCKS = 2048;
        const int nranges = device_ptrs || nb < FORCE_SPLIT_BLOCKS
                                ? 1
                                : std::min<int>(tc_context::FORCE_RANGES, nb / (FORCE_SPLIT_BLOCKS / 2));
        if (nranges <= 1) {
            tc_eval_force(tree, theta, dx0, dy0, dh, nb, dxr, dyr, c->precision, cnt, c->ws, c->force_chunks,
                          c->force_inline, c->stream, c->profile ? c->ev[2] : nullptr,
                          c->profile ? c->ev[3] : nullptr);
            c->ev_valid = c->profile;
            c->eval_has_tsort = false;
            if (!device_ptrs)
                copy_out_all(c, dxr, {{xr, (long long)nt}, {yr, (long long)nt}});
            return;
        }
        const int pool = c->force_chunks > 0 ? c->force_chunks : std::max(4096, std::min(nb * 16, 1 << 20) / 4);
        const int per = (nb + nranges - 1) / nranges;
        for (int k = 0, b0 = 0; b0 < nb; ++k, b0 += per) {
            const int nbk = std::min(per, nb - b0);
            const size_t o = (size_t)b0 * 1024;
            tc_eval_force(tree, theta, dx0 + b0, dy0 + b0, dh + b0, nbk, dxr + o, dyr + o, c->precision, cnt, c->ws,
                          pool, c->force_inline, c->stream, c->profile && k == 0 ? c->ev[2] : nullptr,
                          c->profile && b0 + nbk == nb ? c->ev[3] : nullptr);
            TC_HIP(hipEventRecord(c->range_ev[k], c->stream));
        }
        c->ev_valid = c->profile;
        c->eval_has_tsort = false;
        for (int k = 0, b0 = 0; b0 < nb; ++k, b0 += per) {
            const long long len = (long long)std::min(per, nb - b0) * 1024;
            const size_t o = (size_t)b0 * 1024;
            TC_HIP(hipStreamWaitEvent(c->tstream, c->range_ev[k], 0));
            copy_out(c->tstg, c->tstream, dxr + o, xr + o, len);
            copy_out(c->tstg, c->tstream, dyr + o, yr + o, len);
        }
    });
}

int run_force_targets(tc_context *ctx, double theta, const double *xs, const double *ys, const double *vs, int n,
                      const double *xt, const double *yt, int m, double *xr, double *yr, int device_ptrs)
{
    return guarded(ctx, [&] {
        tc_context *c = ctx;
        c->ev_valid = false;
        if (m <= 0)
            return;
        if (n <= 0) {
            if (device_ptrs) {
                TC_HIP(hipMemsetAsync(xr, 0, sizeof(double) * (size_t)m, c->stream));
                TC_HIP(hipMemsetAsync(yr, 0, sizeof(double) * (size_t)m, c->stream));
            } else {
                std::memset(xr, 0, sizeof(double) * (size_t)m);
                std::memset(yr, 0, sizeof(double) * (size_t)m);
            }
            c->tree_kind = TC_KIND_NONE;
            return;
        }
        const double *dxs = xs, *dys = ys, *dvs = vs, *dxt = xt, *dyt = yt;
        double *dxr = xr, *dyr = yr;
        if (!device_ptrs) {
            c->io.reserve(sizeof(double) * (3 * (size_t)n + 4 * (size_t)m));
            double *base = c->io.as<double>();
            double *bxs = base, *bys = base + n, *bvs = base + 2 * (size_t)n;
            double *bxt = base + 3 * (size_t)n, *byt = bxt + m;
            dxr = byt + m;
            dyr = dxr + m;
            copy_in_all(c, bxs, {{xs, n}, {ys, n}, {vs, n}, {xt, m}, {yt, m}});
            dxs = bxs; dys = bys; dvs = bvs; dxt = bxt; dyt = byt;
        }
        unsigned long long *cnt = prepare_counters(c);
        record(c, 0);
        build(c, dxs, dys, dvs, n, TC_KIND_FORCE, ExtTree());
        record(c, 1);
        tc_eval_force_targets(c->tree, theta, dxt, dyt, m, dxr, dyr, c->precision, cnt, c->ws, c->stream,
                              c->profile ? c->ev[2] : nullptr, c->profile ? c->ev[3] : nullptr);
        c->ev_valid = c->profile;
        if (!device_ptrs) {
            uint32_t *bad = static_cast<uint32_t *>(c->ws.host_pinned());
            TC_HIP(hipMemcpyAsync(bad, tc_targets_bad_flag(c->ws, m), sizeof(uint32_t), hipMemcpyDeviceToHost,
                                  c->stream));
            copy_out_all(c, dxr, {{xr, m}, {yr, m}});
            if (*bad & 2)
                throw TcError(-9, "treecode_force: the target sort did not complete (a radix pass timed out)");
            if (*bad)
                throw TcError(-4, "treecode_force: non-finite target coordinate");
        }
    });
}

tc_context *default_context()
{
    if (!g_default)
        g_default = tc_context_create(0);
    return g_default;
}

tc_context *g_default_f32 = nullptr;

tc_context *default_context_f32()
{
    if (!g_default_f32) {
        g_default_f32 = tc_context_create(0);
        if (g_default_f32)
            tc_set_option(g_default_f32, TC_OPT_PRECISION, 1);
    }
    return g_default_f32;
}

std::vector<double> widen(const float *p, size_t n)
{
    std::vector<double> v(n);
    for (size_t i = 0; i < n; ++i)
        v[i] = (double)p[i];
    return v;
}

void report_dropin_failure(const char *fn)
{
    static bool printed = false;
    if (!printed) {
        std::fprintf(stderr, "%s: %s\n", fn, g_last_error.c_str());
        printed = true;
    }
}

// The float drop-in pair (the reference's real=float build, libraries.Makefile:13):
// float arrays are widened, solved on the FP32-bulk path of a second default
// context, and narrowed back.
void dropin_potential_f32(const char *fn, const float theta, const float *const xsources,
                          const float *const ysources, const float *const sourcevalues, const int nsources,
                          const float *const xtargets, const float *const ytargets, const int ntargets,
                          float *const targetvalues)
{
    std::lock_guard<std::mutex> lk(g_default_mu);
    const size_t n = nsources > 0 ? (size_t)nsources : 0, m = ntargets > 0 ? (size_t)ntargets : 0;
    std::vector<double> xs = widen(xsources, n), ys = widen(ysources, n), vs = widen(sourcevalues, n);
    std::vector<double> xt = widen(xtargets, m), yt = widen(ytargets, m), out(m);
    tc_context *c = default_context_f32();
    int rc = c ? tc_potential(c, (double)theta, xs.data(), ys.data(), vs.data(), nsources, xt.data(), yt.data(),
                              ntargets, out.data(), 0)
               : -1;
    if (rc)
        report_dropin_failure(fn);
    for (size_t i = 0; i < m; ++i)
        targetvalues[i] = rc ? std::numeric_limits<float>::quiet_NaN() : (float)out[i];
}

void dropin_force_f32(const char *fn, const float theta, const float *const xsources, const float *const ysources,
                      const float *const sourcevalues, const int nsources, const float *const x0s,
                      const float *const y0s, const float *const hs, const int nblocks, float *const xresults,
                      float *const yresults)
{
    std::lock_guard<std::mutex> lk(g_default_mu);
    const size_t n = nsources > 0 ? (size_t)nsources : 0, nb = nblocks > 0 ? (size_t)nblocks : 0;
    std::vector<double> xs = widen(xsources, n), ys = widen(ysources, n), vs = widen(sourcevalues, n);
    std::vector<double> bx = widen(x0s, nb), by = widen(y0s, nb), bh = widen(hs, nb);
    std::vector<double> xr(nb * 1024), yr(nb * 1024);
    tc_context *c = default_context_f32();
    int rc = c ? tc_force_mrag(c, (double)theta, xs.data(), ys.data(), vs.data(), nsources, bx.data(), by.data(),
                               bh.data(), nblocks, xr.data(), yr.data(), 0)
               : -1;
    if (rc)
        report_dropin_failure(fn);
    for (size_t i = 0; i < nb * 1024; ++i) {
        xresults[i] = rc ? std::numeric_limits<float>::quiet_NaN() : (float)xr[i];
        yresults[i] = rc ? std::numeric_limits<float>::quiet_NaN() : (float)yr[i];
    }
}

}  // namespace

// ------------------------------------------------------------------ request-server internals
void tc_set_last_error(const std::string &m) { set_error(m); }

int tc_srv_solve(tc_context *c, const tc_request &r, TcTree &t, bool build, hipEvent_t ready,
                 const std::function<void(const TcPotBatchItem &, hipStream_t)> *eval)
{
    ExtTree ext;
    ext.tree = &t;
    ext.build = build;
    ext.ready = ready;
    ext.eval = eval;
    if (r.kind == TC_REQ_POTENTIAL)
        return run_potential(c, r.theta, r.xs, r.ys, r.vs, build ? r.n : 0, r.xt, r.yt, r.m, r.out0, 0, build, ext);
    return run_force(c, r.theta, r.xs, r.ys, r.vs, build ? r.n : 0, r.x0s, r.y0s, r.hs, r.nblocks, r.out0, r.out1,
                     0, build, ext);
}

size_t tc_tree_bytes(const TcTree &t)
{
    size_t b = 0;
    for (const DevBuf *d : {&t.keys, &t.perm, &t.src, &t.src32, &t.nodes, &t.exps, &t.exps32, &t.nodew, &t.selfid,
                            &t.box_d})
        b += d->cap;
    return b;
}

size_t tc_srv_context_bytes(tc_context *c)
{
    size_t b = tc_tree_bytes(c->tree) + c->io.cap + c->cnt.cap;
    for (const TcScratch *w : {&c->ws, &c->tws})
        for (const DevBuf *d : {&w->k2, &w->v2, &w->hist, &w->part, &w->scan_tmp, &w->istart, &w->ilevel, &w->info,
                                &w->tkeys, &w->tperm, &w->tsorted, &w->bbox_part, &w->ints, &w->packed, &w->work,
                                &w->flist})
            b += d->cap;
    return b;
}

void tc_srv_context_trim(tc_context *c)
{
    (void)hipSetDevice(c->device);
    (void)hipStreamSynchronize(c->tstream);
    (void)hipStreamSynchronize(c->stream);
    c->tree.release();
    c->tree_kind = TC_KIND_NONE;
    c->io.release();
    for (TcScratch *w : {&c->ws, &c->tws})
        for (DevBuf *d : {&w->k2, &w->v2, &w->hist, &w->part, &w->scan_tmp, &w->istart, &w->ilevel, &w->info,
                          &w->tkeys, &w->tperm, &w->tsorted, &w->bbox_part, &w->ints, &w->packed, &w->work,
                          &w->flist})
            d->release();
}

extern "C" {

TC_API const char *treecode_last_error(void) { return g_last_error.c_str(); }

TC_API tc_context *tc_context_create(int device)
{
    tc_context *c = new tc_context();
    c->device = device;
    try {
        int ndev = 0;
        TC_HIP(hipGetDeviceCount(&ndev));
        if (ndev <= 0)
            throw TcError(-1, "treecode: no HIP device available (this library has no CPU fallback)");
        if (device < 0 || device >= ndev)
            throw TcError(-1, "treecode: device " + std::to_string(device) + " out of range");
        TC_HIP(hipSetDevice(device));
        TC_HIP(hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking));
        TC_HIP(hipStreamCreateWithFlags(&c->tstream, hipStreamNonBlocking));
        for (auto &e : c->ev)
            TC_HIP(hipEventCreate(&e));
        for (auto &e : c->bev)
            TC_HIP(hipEventCreate(&e));
        TC_HIP(hipEventCreateWithFlags(&c->fork_ev, hipEventDisableTiming));
        TC_HIP(hipEventCreateWithFlags(&c->join_ev, hipEventDisableTiming));
        for (auto &e : c->range_ev)
            TC_HIP(hipEventCreateWithFlags(&e, hipEventDisableTiming));
        ensure_binomials(device, c->stream);
#ifndef TREECODES2D_FLOAT_DROPIN
        // set-up of tc_potential_noalloc: the device address of the static
        // scratch (device current) and the pinned readback buffers it uses
        c->na_base = tc_noalloc_scratch_lookup(device);
        c->na_ws.host_pinned();
        c->na_tws.host_pinned();
#endif
    } catch (const TcError &e) {
        set_error(e.what());
        tc_context_destroy(c);
        return nullptr;
    }
    return c;
}

TC_API void tc_context_destroy(tc_context *c)
{
    if (!c)
        return;
    (void)hipSetDevice(c->device);
    if (c->tstream)
        (void)hipStreamSynchronize(c->tstream);
    if (c->stream)
        (void)hipStreamSynchronize(c->stream);
    c->tree.release();
    c->bws.release();
    c->sdesc.release();
    if (c->sdesc_host)
        (void)hipHostFree(c->sdesc_host);
    c->ws.release();
    c->tws.release();
    c->io.release();
    c->cnt.release();
    c->na_tree.release();
    c->na_ws.release();
    c->na_tws.release();
    c->na_cnt.release();
    for (auto &e : c->ev)
        if (e)
            (void)hipEventDestroy(e);
    for (auto &e : c->bev)
        if (e)
            (void)hipEventDestroy(e);
    c->stg.release();
    c->tstg.release();
    for (hipEvent_t e : {c->fork_ev, c->join_ev})
        if (e)
            (void)hipEventDestroy(e);
    for (hipEvent_t e : c->range_ev)
        if (e)
            (void)hipEventDestroy(e);
    if (c->tstream)
        (void)hipStreamDestroy(c->tstream);
    if (c->stream)
        (void)hipStreamDestroy(c->stream);
    delete c;
}

TC_API void *tc_context_stream(tc_context *c) { return c ? (void *)c->stream : nullptr; }

TC_API int tc_set_option(tc_context *c, int key, int value)
{
    if (!c) {
        set_error("treecode: null context");
        return -1;
    }
    auto valid_order = [](int p) { return tc_order_supported(p); };
    switch (key) {
    case TC_OPT_PRECISION:
        if (value != 0 && value != 1) break;
        c->precision = value;
        return 0;
    case TC_OPT_POT_ORDER:
        if (!valid_order(value)) break;
        c->pot_order = value;
        return 0;
    case TC_OPT_POT_LEAF:
        if (value < 1 || value > TC_MAX_LEAF) break;
        c->pot_leaf = value;
        return 0;
    case TC_OPT_FORCE_ORDER:
        if (!valid_order(value)) break;
        c->force_order = value;
        return 0;
    case TC_OPT_FORCE_LEAF:
        if (value < 1 || value > TC_MAX_LEAF) break;
        c->force_leaf = value;
        return 0;
    case TC_OPT_PROFILE:
        c->profile = value != 0;
        return 0;
    case TC_OPT_FORCE_CHUNKS:
        if (value < 0) break;
        c->force_chunks = value;
        return 0;
    case TC_OPT_POT_XCD_CHUNK:
        if (value == 0) {
            c->pot_xcd_log2 = TC_XCD_CHUNK_LOG2;
            return 0;
        }
        if (value < 1 || value > 256 || (value & (value - 1)) != 0) break;
        c->pot_xcd_log2 = __builtin_ctz((unsigned)value);
        return 0;
    case TC_OPT_FORCE_INLINE:
        if (value < 0 || value > 4096) break;
        c->force_inline = value;
        return 0;
    case TC_OPT_COUNTERS:
        c->count = value != 0;
        return 0;
    default:
        break;
    }
    set_error("treecode: invalid option " + std::to_string(key) + " = " + std::to_string(value));
    return -3;
}

TC_API int tc_potential(tc_context *c, double theta, const double *xs, const double *ys, const double *vs, int n,
                        const double *xt, const double *yt, int m, double *out, int device_ptrs)
{
    return run_potential(c, theta, xs, ys, vs, n, xt, yt, m, out, device_ptrs, true);
}

TC_API int tc_potential_resident(tc_context *c, double theta, const double *xt, const double *yt, int m,
                                 double *out, int device_ptrs)
{
    return run_potential(c, theta, nullptr, nullptr, nullptr, 0, xt, yt, m, out, device_ptrs, false);
}

TC_API int tc_potential_batch_resident(tc_context *const *ctxs, int k, const double *theta,
                                       const double *const *xt, const double *const *yt, const int *m,
                                       double *const *out)
{
    if (!ctxs || k <= 0 || !theta || !xt || !yt || !m || !out) {
        set_error("tc_potential_batch_resident: null argument or no request");
        return -1;
    }
    tc_context *c0 = ctxs[0];
    return guarded(c0, [&] {
        if (k > TcBatchWs::MAXREQ)
            throw TcError(-3, "tc_potential_batch_resident: at most 64 requests per call");
        for (int i = 0; i < k; ++i) {
            tc_context *c = ctxs[i];
            if (!c)
                throw TcError(-1, "tc_potential_batch_resident: null context");
            for (int j = 0; j < i; ++j)
                if (ctxs[j] == c)
                    throw TcError(-3, "tc_potential_batch_resident: a context appears twice");
            if (c->device != c0->device)
                throw TcError(-3, "tc_potential_batch_resident: contexts on different devices");
            check_resident(c, TC_KIND_POTENTIAL);
            if (c->tree.order != c0->tree.order || c->precision != c0->precision)
                throw TcError(-3, "tc_potential_batch_resident: trees of different orders or precisions");
            if (m[i] < 0 || (m[i] > 0 && (!xt[i] || !yt[i] || !out[i])))
                throw TcError(-3, "tc_potential_batch_resident: bad target arrays");
            c->ev_valid = false;
        }
        hipStream_t st = c0->stream;
        // work queued on the other contexts (their trees' builds) before the batch
        for (int i = 1; i < k; ++i) {
            TC_HIP(hipEventRecord(ctxs[i]->join_ev, ctxs[i]->stream));
            TC_HIP(hipStreamWaitEvent(st, ctxs[i]->join_ev, 0));
        }
        // each request's target order in its own context's scratch: requests of
        // at most TC_SORT_SMALL_MAX targets in one launch, larger ones alone;
        // then one launch of the traversal
        if (!c0->sdesc_host)
            TC_HIP(hipHostMalloc(&c0->sdesc_host, sizeof(TcSortDesc) * TcBatchWs::MAXREQ, hipHostMallocDefault));
        c0->sdesc.reserve(sizeof(TcSortDesc) * TcBatchWs::MAXREQ);
        TcSortDesc *sd = static_cast<TcSortDesc *>(c0->sdesc_host);
        int nsmall = 0;
        for (int i = 0; i < k; ++i) {
            if (m[i] > 0 && m[i] <= TC_SORT_SMALL_MAX)
                sd[nsmall++] = tc_sort_targets_desc(xt[i], yt[i], m[i], ctxs[i]->tws);
            else if (m[i] > 0)
                tc_sort_targets(xt[i], yt[i], m[i], ctxs[i]->tws, st);
        }
        if (nsmall) {
            TC_HIP(hipMemcpyAsync(c0->sdesc.ptr, sd, sizeof(TcSortDesc) * nsmall, hipMemcpyHostToDevice, st));
            tc_sort_targets_batch(c0->sdesc.as<TcSortDesc>(), nsmall, st);
        }
        std::vector<TcPotBatchItem> items;
        for (int i = 0; i < k; ++i) {
            if (m[i] == 0)
                continue;
            TcPotBatchItem it;
            it.tree = &ctxs[i]->tree;
            it.theta = theta[i];
            it.xt = xt[i];
            it.yt = yt[i];
            it.tperm = ctxs[i]->tws.tperm.as<uint32_t>();
            it.m = m[i];
            it.out = out[i];
            items.push_back(it);
        }
        if (!items.empty())
            tc_potential_batch(items.data(), (int)items.size(), c0->tree.order, c0->precision, c0->bws, nullptr, st,
                               true);
        TC_HIP(hipStreamSynchronize(st));
    });
}

#ifndef TREECODES2D_FLOAT_DROPIN
TC_API int tc_potential_noalloc(tc_context *c, double theta, const double *xs, const double *ys, const double *vs,
                                int n, const double *xt, const double *yt, int m, double *out)
{
    return run_potential_noalloc(c, theta, xs, ys, vs, n, xt, yt, m, out);
}
#endif

TC_API int tc_force_mrag(tc_context *c, double theta, const double *xs, const double *ys, const double *vs,
                         int n, const double *x0s, const double *y0s, const double *hs, int nblocks, double *xr,
                         double *yr, int device_ptrs)
{
    return run_force(c, theta, xs, ys, vs, n, x0s, y0s, hs, nblocks, xr, yr, device_ptrs, true);
}

TC_API int tc_force_mrag_resident(tc_context *c, double theta, const double *x0s, const double *y0s,
                                  const double *hs, int nblocks, double *xr, double *yr, int device_ptrs)
{
    return run_force(c, theta, nullptr, nullptr, nullptr, 0, x0s, y0s, hs, nblocks, xr, yr, device_ptrs, false);
}

TC_API int tc_force_targets(tc_context *c, double theta, const double *xs, const double *ys, const double *vs,
                            int n, const double *xt, const double *yt, int m, double *xr, double *yr,
                            int device_ptrs)
{
    return run_force_targets(c, theta, xs, ys, vs, n, xt, yt, m, xr, yr, device_ptrs);
}

TC_API int tc_synchronize(tc_context *c)
{
    return guarded(c, [&] { TC_HIP(hipStreamSynchronize(c->stream)); });
}

TC_API int tc_get_stats(tc_context *c, double *times_ms, unsigned long long *counters, int *tree_info)
{
    return guarded(c, [&] {
        TC_HIP(hipStreamSynchronize(c->stream));
        if (times_ms) {
            for (int i = 0; i < TC_NPHASES; ++i)
                times_ms[i] = 0;
            if (c->ev_valid) {
                float a = 0, b = 0, d = 0, t = 0;
                TC_HIP(hipEventElapsedTime(&a, c->ev[0], c->ev[1]));
                TC_HIP(hipEventElapsedTime(&b, c->ev[1], c->ev[2]));
                TC_HIP(hipEventElapsedTime(&d, c->ev[2], c->ev[3]));
                TC_HIP(hipEventElapsedTime(&t, c->ev[0], c->ev[3]));
                times_ms[TC_PH_BUILD] = a;
                times_ms[TC_PH_TSORT] = b;
                times_ms[TC_PH_EVAL] = d;
                times_ms[TC_PH_TOTAL] = t;
            }
        }
        if (counters) {
            unsigned long long all[TC_NCOUNTERS];
            read_counters(c, all);
            std::memcpy(counters, all, sizeof(unsigned long long) * TC_NPUBLIC_COUNTERS);
        }
        if (tree_info) {
            tree_info[0] = c->tree.n;
            tree_info[1] = c->tree.nnodes;
            tree_info[2] = c->tree.ninternal;
            tree_info[3] = c->tree.order;
        }
    });
}

TC_API int tc_get_profile(tc_context *c, double *build_ms, unsigned long long *cycles)
{
    return guarded(c, [&] {
        TC_HIP(hipStreamSynchronize(c->stream));
        if (build_ms) {
            for (int i = 0; i < 4; ++i)
                build_ms[i] = 0;
            if (c->ev_valid && c->bev_valid) {
                const hipEvent_t marks[5] = {c->ev[0], c->bev[0], c->bev[1], c->bev[2], c->bev[3]};
                for (int i = 0; i < 4; ++i) {
                    float ms = 0;
                    TC_HIP(hipEventElapsedTime(&ms, marks[i], marks[i + 1]));
                    build_ms[i] = ms;
                }
            }
        }
        if (cycles) {
            unsigned long long all[TC_NCOUNTERS];
            read_counters(c, all);
            for (int i = 0; i < 5; ++i)
                cycles[i] = all[TC_CNT_CYC_P2P + i];
        }
    });
}

TC_API int tc_tree_export(tc_context *c, int *keys, int *perm, double *box, double *geom, int *topo, double *exps)
{
    return guarded(c, [&] {
        if (c->tree_kind == TC_KIND_NONE)
            throw TcError(-5, "treecode: no resident tree in this context");
        TC_HIP(hipStreamSynchronize(c->stream));
        const TcTree &t = c->tree;
        if (keys)
            TC_HIP(hipMemcpy(keys, t.keys.ptr, sizeof(int) * (size_t)t.n, hipMemcpyDeviceToHost));
        if (perm)
            TC_HIP(hipMemcpy(perm, t.perm.ptr, sizeof(int) * (size_t)t.n, hipMemcpyDeviceToHost));
        if (box)
            for (int i = 0; i < 3; ++i)
                box[i] = t.box[i];
        if (geom || topo) {
            std::vector<TcNode> nodes((size_t)t.nnodes);
            TC_HIP(hipMemcpy(nodes.data(), t.nodes.ptr, sizeof(TcNode) * nodes.size(), hipMemcpyDeviceToHost));
            for (size_t i = 0; i < nodes.size(); ++i) {
                if (geom) {
                    geom[4 * i + 0] = nodes[i].xcom;
                    geom[4 * i + 1] = nodes[i].ycom;
                    geom[4 * i + 2] = nodes[i].r;
                    geom[4 * i + 3] = nodes[i].mass;
                }
                if (topo) {
                    topo[4 * i + 0] = nodes[i].s;
                    topo[4 * i + 1] = nodes[i].e;
                    topo[4 * i + 2] = nodes[i].child;
                    topo[4 * i + 3] = nodes[i].level;
                }
            }
        }
        if (exps)
            TC_HIP(hipMemcpy(exps, t.exps.ptr, sizeof(double) * 2 * (size_t)t.order * t.nnodes,
                             hipMemcpyDeviceToHost));
    });
}

TC_API void *tc_device_alloc(tc_context *c, size_t bytes)
{
    void *p = nullptr;
    const int rc = guarded(c, [&] { TC_HIP(hipMalloc(&p, bytes ? bytes : 1)); });
    return rc ? nullptr : p;
}

TC_API int tc_device_free(tc_context *c, void *p)
{
    return guarded(c, [&] {
        TC_HIP(hipStreamSynchronize(c->stream));
        TC_HIP(hipFree(p));
    });
}

TC_API int tc_copy_to_device(tc_context *c, void *dst, const void *src, size_t bytes)
{
    return guarded(c, [&] {
        TC_HIP(hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, c->stream));
        TC_HIP(hipStreamSynchronize(c->stream));
    });
}

TC_API int tc_copy_to_host(tc_context *c, void *dst, const void *src, size_t bytes)
{
    return guarded(c, [&] {
        TC_HIP(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, c->stream));
        TC_HIP(hipStreamSynchronize(c->stream));
    });
}

TC_API int tc_direct_potential(tc_context *c, const double *xs, const double *ys, const double *vs, int n,
                               const double *xt, const double *yt, const int *idx, int nidx, double *out)
{
    return guarded(c, [&] {
        tc_direct_sums(xs, ys, vs, n, xt, yt, idx, nidx, 0, out, nullptr, c->stream);
        TC_HIP(hipStreamSynchronize(c->stream));
    });
}

TC_API int tc_direct_force(tc_context *c, const double *xs, const double *ys, const double *vs, int n,
                           const double *xt, const double *yt, const int *idx, int nidx, double *outx,
                           double *outy)
{
    return guarded(c, [&] {
        tc_direct_sums(xs, ys, vs, n, xt, yt, idx, nidx, 1, outx, outy, c->stream);
        TC_HIP(hipStreamSynchronize(c->stream));
    });
}

TC_API int tc_copy_kernel(tc_context *c, void *dst, const void *src, size_t bytes)
{
    return guarded(c, [&] {
        if (bytes % 16)
            throw TcError(-3, "tc_copy_kernel: size must be a multiple of 16 bytes");
        tc_copy16(src, dst, bytes, c->stream);
        TC_HIP(hipStreamSynchronize(c->stream));
    });
}

TC_API int tc_morton_sort_device(tc_context *c, const double *x, const double *y, int n, double *box,
                                 unsigned int *keys, unsigned int *perm)
{
    return guarded(c, [&] {
        if (n <= 0)
            throw TcError(-3, "tc_morton_sort_device: n must be positive");
        tc_sort_points(x, y, nullptr, n, box, keys, perm, nullptr, c->ws, c->stream, false);
    });
}

#ifndef TREECODES2D_FLOAT_DROPIN
TC_API void treecode_potential(const double theta, const double *const xsources, const double *const ysources,
                               const double *const sourcevalues, const int nsources,
                               const double *const xtargets, const double *const ytargets, const int ntargets,
                               double *const targetvalues)
{
    std::lock_guard<std::mutex> lk(g_default_mu);
    tc_context *c = default_context();
    int rc = c ? tc_potential(c, theta, xsources, ysources, sourcevalues, nsources, xtargets, ytargets, ntargets,
                              targetvalues, 0)
               : -1;
    if (rc) {
        report_dropin_failure("treecode_potential");
        for (int i = 0; i < ntargets; ++i)
            targetvalues[i] = std::numeric_limits<double>::quiet_NaN();
    }
}

TC_API void treecode_force(const double theta, const double *const xsources, const double *const ysources,
                           const double *const sourcevalues, const int nsources, const double *const xtargets,
                           const double *const ytargets, const int ntargets, double *const xresults,
                           double *const yresults)
{
    std::lock_guard<std::mutex> lk(g_default_mu);
    tc_context *c = default_context();
    int rc = c ? tc_force_targets(c, theta, xsources, ysources, sourcevalues, nsources, xtargets, ytargets,
                                  ntargets, xresults, yresults, 0)
               : -1;
    if (rc) {
        report_dropin_failure("treecode_force");
        for (int i = 0; i < ntargets; ++i)
            xresults[i] = yresults[i] = std::numeric_limits<double>::quiet_NaN();
    }
}

TC_API void treecode_velocity_mrag(const double theta, const double *const xsources,
                                   const double *const ysources, const double *const vorticity,
                                   const int nsources, const double *const x0s, const double *const y0s,
                                   const double *const hs, const int nblocks, double *const u, double *const v)
{
    treecode_force_mrag(theta, xsources, ysources, vorticity, nsources, x0s, y0s, hs, nblocks, v, u);
    const double inv2pi = 1.0 / (2.0 * M_PI);
    for (size_t i = 0; i < (size_t)(nblocks > 0 ? nblocks : 0) * 1024; ++i) {
        const double gx = v[i], gy = u[i];
        u[i] = gy * inv2pi;
        v[i] = -gx * inv2pi;
    }
}

TC_API void treecode_force_mrag(const double theta, const double *const xsources, const double *const ysources,
                                const double *const sourcevalues, const int nsources, const double *const x0s,
                                const double *const y0s, const double *const hs, const int nblocks,
                                double *const xresults, double *const yresults)
{
    std::lock_guard<std::mutex> lk(g_default_mu);
    tc_context *c = default_context();
    int rc = c ? tc_force_mrag(c, theta, xsources, ysources, sourcevalues, nsources, x0s, y0s, hs, nblocks,
                               xresults, yresults, 0)
               : -1;
    if (rc) {
        report_dropin_failure("treecode_force_mrag");
        for (size_t i = 0; i < (size_t)(nblocks > 0 ? nblocks : 0) * 1024; ++i)
            xresults[i] = yresults[i] = std::numeric_limits<double>::quiet_NaN();
    }
}

TC_API void treecode_potential_float(const float theta, const float *const xsources, const float *const ysources,
                                     const float *const sourcevalues, const int nsources,
                                     const float *const xtargets, const float *const ytargets,
                                     const int ntargets, float *const targetvalues)
{
    dropin_potential_f32("treecode_potential_float", theta, xsources, ysources, sourcevalues, nsources, xtargets,
                         ytargets, ntargets, targetvalues);
}

TC_API void treecode_force_mrag_float(const float theta, const float *const xsources,
                                      const float *const ysources, const float *const sourcevalues,
                                      const int nsources, const float *const x0s, const float *const y0s,
                                      const float *const hs, const int nblocks, float *const xresults,
                                      float *const yresults)
{
    dropin_force_f32("treecode_force_mrag_float", theta, xsources, ysources, sourcevalues, nsources, x0s, y0s, hs,
                     nblocks, xresults, yresults);
}

#else  // TREECODES2D_FLOAT_DROPIN

// drivers/treecode-potential.h:21-29 and drivers/treecode-force.h:21-31 with realtype = float
TC_API void treecode_potential(const float theta, const float *const xsources, const float *const ysources,
                               const float *const sourcevalues, const int nsources, const float *const xtargets,
                               const float *const ytargets, const int ntargets, float *const targetvalues)
{
    dropin_potential_f32("treecode_potential", theta, xsources, ysources, sourcevalues, nsources, xtargets, ytargets,
                         ntargets, targetvalues);
}

TC_API void treecode_force_mrag(const float theta, const float *const xsources, const float *const ysources,
                                const float *const sourcevalues, const int nsources, const float *const x0s,
                                const float *const y0s, const float *const hs, const int nblocks,
                                float *const xresults, float *const yresults)
{
    dropin_force_f32("treecode_force_mrag", theta, xsources, ysources, sourcevalues, nsources, x0s, y0s, hs, nblocks,
                     xresults, yresults);
}

#endif  // TREECODES2D_FLOAT_DROPIN

}  // extern "C"
