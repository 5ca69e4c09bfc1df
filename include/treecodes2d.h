/*
 * treecodes2d.h -- C ABI of the MI355X treecode library (libtreecodes2d.so).
 *
 * The two drop-in entry points keep the reference prototypes exactly:
 *   treecode_potential   drivers/treecode-potential.h:21-29 (README.md:5-13)
 *   treecode_force_mrag  drivers/treecode-force.h:21-31     (README.md:15-25)
 * and are also declared, byte-compatible with the reference's generated
 * headers, in 2d-treecodes-potential.h and 2d-treecodes-force.h.
 *
 * Everything below the drop-in pair is an extension with no counterpart in
 * the reference (whose library has no error channel, no device-resident
 * interface and no options; it is configured at compile time by
 * Makefile:13-15, drivers/Makefile:20).  All pointers are plain pointers,
 * sizes are ints, no C++ or HIP types cross this boundary (streams are void*).
 */
#ifndef TREECODES2D_H
#define TREECODES2D_H

#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

#define TC_API __attribute__((visibility("default")))

/* ---- drop-in entry points (reference semantics, FP64, host memory, synchronous) ----
 * TREECODES2D_FLOAT_DROPIN: defined by the float build (libtreecodes2d_float.so), whose
 * treecode_potential / treecode_force_mrag take float arrays under the same names, as
 * the reference's real=float build (include/float/2d-treecodes-*.h declare them). */
#ifndef TREECODES2D_FLOAT_DROPIN

/* drivers/treecode-potential.h:21-29: out[i] = sum_j v_j log|t_i - s_j| with the
 * multipole order 12 / leaf capacity 512 of the reference build. */
TC_API void treecode_potential(const double theta, const double *const xsources, const double *const ysources,
                               const double *const sourcevalues, const int nsources,
                               const double *const xtargets, const double *const ytargets, const int ntargets,
                               double *const targetvalues);

/* drivers/treecode-force.h:21-31: gradient of the potential at the 32x32 points of each
 * block, (x0s[b] + ix*hs[b], y0s[b] + iy*hs[b]) written at ix + 32*(iy + 32*b);
 * order 24 / leaf capacity 192 / 8x8 bricks as the reference build. */
TC_API void treecode_force_mrag(const double theta, const double *const xsources, const double *const ysources,
                                const double *const sourcevalues, const int nsources, const double *const x0s,
                                const double *const y0s, const double *const hs, const int nblocks,
                                double *const xresults, double *const yresults);

/* The reference's `real=float` build (libraries.Makefile:13, realtype substituted by m4)
 * has the same two functions over float arrays.  libtreecodes2d_float.so exports them
 * under the reference's names (headers in include/float/); in this library they carry a
 * _float suffix so both variants can be called from one process.  Both run the FP32-bulk
 * path (FP64 tree geometry and translations, FP32 near field and multipole evaluation). */
TC_API void treecode_potential_float(const float theta, const float *const xsources, const float *const ysources,
                                     const float *const sourcevalues, const int nsources,
                                     const float *const xtargets, const float *const ytargets,
                                     const int ntargets, float *const targetvalues);
TC_API void treecode_force_mrag_float(const float theta, const float *const xsources,
                                      const float *const ysources, const float *const sourcevalues,
                                      const int nsources, const float *const x0s, const float *const y0s,
                                      const float *const hs, const int nblocks, float *const xresults,
                                      float *const yresults);

/* Gradient at an arbitrary list of targets (the reference supports velocities only on
 * 32x32 blocks, main.cpp:191-195): (xr, yr)[i] = sum_j v_j (t_i - s_j)/(|t_i - s_j|^2 + 10 eps).
 * Order 24 / leaf 192 as the force build; per-target opening test r^2 < theta^2 d^2 of
 * treecode-potential.cpp:48-50 with the force E2P and P2P kernels. */
TC_API void treecode_force(const double theta, const double *const xsources, const double *const ysources,
                           const double *const sourcevalues, const int nsources, const double *const xtargets,
                           const double *const ytargets, const int ntargets, double *const xresults,
                           double *const yresults);

/* Velocities of an MRAG vortex solver from vorticity sources: the rotation and 2 pi
 * scaling that main.cpp:136-143 applies to compare against stored velocities,
 * u = grad_y / (2 pi), v = -grad_x / (2 pi) of treecode_force_mrag, same layout. */
TC_API void treecode_velocity_mrag(const double theta, const double *const xsources,
                                   const double *const ysources, const double *const vorticity,
                                   const int nsources, const double *const x0s, const double *const y0s,
                                   const double *const hs, const int nblocks, double *const u, double *const v);
#endif /* TREECODES2D_FLOAT_DROPIN */

/* ---- extensions (no counterpart in the reference) ---- */

/* Message of the last failure on the calling thread ("" if none).  The void
 * drop-in functions report failures here, print them once to stderr and fill
 * their outputs with NaN. */
TC_API const char *treecode_last_error(void);

typedef struct tc_context tc_context;

/* Options (tc_set_option keys) */
enum {
    TC_OPT_PRECISION = 0,      /* 0 = plain FP64 (default), 1 = FP32 bulk P2P/E2P */
    TC_OPT_POT_ORDER = 1,      /* potential expansion order (6,8,9,10,12,14,16,20,24,32), default 12 */
    TC_OPT_POT_LEAF = 2,       /* potential leaf capacity, default 512 */
    TC_OPT_FORCE_ORDER = 3,    /* force expansion order, default 24 */
    TC_OPT_FORCE_LEAF = 4,     /* force leaf capacity, default 192 */
    TC_OPT_PROFILE = 5,        /* 1 = record phase times with HIP events */
    TC_OPT_COUNTERS = 6,       /* 1 = count interactions (see tc_get_stats) */
    TC_OPT_FORCE_CHUNKS = 7,   /* force near-field pool, in chunks of 64 leaves; 0 = automatic
                                * (max(4096, bricks / 4)).  A brick with more P2P leaves than
                                * TC_OPT_FORCE_INLINE spreads the rest over chunks; bricks that find
                                * the pool full repeat their traversal on a slower path */
    /* tiling (autotuned with order and leaf capacity; see autotune.py): */
    TC_OPT_POT_XCD_CHUNK = 8,  /* potential traversal: workgroups per XCD chunk, a power of two in
                                * 1..256 (1 = the hardware's round-robin order); 0 = 16.  Results do
                                * not depend on it */
    TC_OPT_FORCE_INLINE = 9    /* force traversal: P2P leaves a brick sums itself before further
                                * ones go to chunks summed by other waves, 1..4096; 0 = 64.  Changes
                                * only the order of summation */
};

/* Phase times (tc_get_stats times_ms[]) */
enum { TC_PH_BUILD = 0, TC_PH_TSORT = 1, TC_PH_EVAL = 2, TC_PH_TOTAL = 3, TC_NPHASES = 4 };

/* A context owns one HIP stream, grow-only HBM buffers and the resident tree. */
TC_API tc_context *tc_context_create(int device);
TC_API void tc_context_destroy(tc_context *ctx);
TC_API int tc_set_option(tc_context *ctx, int key, int value);
TC_API void *tc_context_stream(tc_context *ctx);

/* Potential / force with all pointers either in host memory (device_ptrs = 0:
 * synchronous, copies in and out) or in device memory of the context's GPU
 * (device_ptrs = 1: enqueued on the context stream; tc_synchronize waits).
 * Return 0 on success, a negative code on failure (see treecode_last_error). */
TC_API int tc_potential(tc_context *ctx, double theta, const double *xs, const double *ys, const double *vs, int n,
                        const double *xt, const double *yt, int m, double *out, int device_ptrs);
TC_API int tc_force_mrag(tc_context *ctx, double theta, const double *xs, const double *ys, const double *vs,
                         int n, const double *x0s, const double *y0s, const double *hs, int nblocks, double *xr,
                         double *yr, int device_ptrs);
TC_API int tc_force_targets(tc_context *ctx, double theta, const double *xs, const double *ys, const double *vs,
                            int n, const double *xt, const double *yt, int m, double *xr, double *yr,
                            int device_ptrs);
TC_API int tc_synchronize(tc_context *ctx);

/* tc_potential with device pointers that allocates nothing.  No counterpart in the
 * reference (whose library allocates per call, upward.cpp:289-300): it is for callers
 * that cannot allocate device memory -- code under HIP graph capture, tight serving
 * loops, harnesses that forbid hipMalloc.  The sorted sources, the tree, the target
 * order and the counters live in one static device array of 64 MiB per GPU (held
 * until the process ends; its address is looked up when a context is created), so the
 * call makes no hipMalloc, hipFree or other allocating call.  Limits: n and m at most
 * 262,144 each (larger calls return -3 before any work is queued), and the tree must
 * fit what the points leave of the 64 MiB (about 140 + 24*order bytes per node; a
 * larger tree returns -3 at the build's census).  Order, leaf capacity, precision,
 * counters and timers come from the context's options; results equal tc_potential's
 * with device_ptrs = 1 on the same inputs.  Synchronous (returns with `out` written).
 * Every call on one GPU shares that GPU's array and overwrites it: calls that overlap
 * on one GPU (from several threads or contexts) race and must be serialised by the
 * caller.  The resident tree of tc_potential_resident is neither used nor changed. */
TC_API int tc_potential_noalloc(tc_context *ctx, double theta, const double *xs, const double *ys,
                                const double *vs, int n, const double *xt, const double *yt, int m, double *out);

/* Evaluate new targets against the tree left resident by the last tc_potential /
 * tc_force_mrag call of this context (the sources are not rebuilt).  The tree must
 * have been built for the same kind of solve (order, leaf capacity, precision). */
TC_API int tc_potential_resident(tc_context *ctx, double theta, const double *xt, const double *yt, int m,
                                 double *out, int device_ptrs);
TC_API int tc_force_mrag_resident(tc_context *ctx, double theta, const double *x0s, const double *y0s,
                                  const double *hs, int nblocks, double *xr, double *yr, int device_ptrs);

/* Several potential requests in one launch of the cross-request traversal.  No
 * counterpart in the reference (whose calls are one-shot, treecode-potential.cpp:99-105):
 * request i evaluates the tree resident in ctxs[i] (left by its last tc_potential call)
 * at the m[i] targets xt[i], yt[i] with theta[i], into out[i]; all arrays are device
 * pointers.  The k <= 64 contexts must be distinct, on one GPU, with resident potential
 * trees of one expansion order and one precision (their leaf capacities may differ).
 * A persistent grid drains the requests' waves of 64 targets from one work counter;
 * each result equals tc_potential_resident's for that request bit for bit.  Runs on
 * ctxs[0]'s stream through a graph per (order, precision, size bucket) and returns
 * with every out[i] written; no other call may use these contexts meanwhile.  The
 * contexts' interaction counters and phase timers are not updated by this call. */
TC_API int tc_potential_batch_resident(tc_context *const *ctxs, int k, const double *theta,
                                       const double *const *xt, const double *const *yt, const int *m,
                                       double *const *out);

/* Times of the last call's phases (requires TC_OPT_PROFILE), counters[8] (requires
 * TC_OPT_COUNTERS): {target-source pairs evaluated directly, (target, node) multipole
 * interactions (the accepted pairs, equal in both precisions), (brick, node) multipole-to-local
 * shifts, targets given a local expansion, lane slots occupied by the direct loops, lane slots
 * occupied by per-target multipole evaluations, multipole-to-local translations about a
 * wavefront's target centre (FP32 potential), (target, node) interactions those translations
 * evaluate}, and tree_info = {n, nnodes, ninternal, order}. */
TC_API int tc_get_stats(tc_context *ctx, double *times_ms, unsigned long long *counters, int *tree_info);

/* Finer profile of the last call, the counterpart of the reference's per-thread TimeDistrib
 * (treecode-force.cpp:26-48).  build_ms[4] (requires TC_OPT_PROFILE and a call that built a
 * tree): bounding box + Morton keys, radix sort + sorted records, topology (census, its
 * readback, nodes), upward pass (P2E, E2E, FP32 mirrors); zeros otherwise.  cycles[5]
 * (requires TC_OPT_COUNTERS): wavefront clock cycles of the traversal summed over waves, spent
 * in {direct sums, per-target multipole evaluations, local-expansion translations (force E2L,
 * FP32 potential wave-level M2L), local evaluation (L2P), the whole traversal}; the rest of the
 * total is the walk itself (decisions, stack, list handling).  Either pointer may be NULL. */
TC_API int tc_get_profile(tc_context *ctx, double *build_ms, unsigned long long *cycles);

/* Copy the resident tree to host memory (any pointer may be NULL):
 * keys[n], perm[n] (sorted -> input index), box[3] (xmin, ymin, ext),
 * geom[4*nn] (xcom, ycom, r, mass), topo[4*nn] (s, e, child, level),
 * exps[2*order*nn] ([Re a_1..a_p][Im a_1..a_p] per node). */
TC_API int tc_tree_export(tc_context *ctx, int *keys, int *perm, double *box, double *geom, int *topo,
                          double *exps);

/* Device memory on the context's GPU, for callers that keep inputs resident in HBM
 * (no PyTorch or other runtime needed).  Copies are synchronous. */
TC_API void *tc_device_alloc(tc_context *ctx, size_t bytes);
TC_API int tc_device_free(tc_context *ctx, void *ptr);
TC_API int tc_copy_to_device(tc_context *ctx, void *dst, const void *src, size_t bytes);
TC_API int tc_copy_to_host(tc_context *ctx, void *dst, const void *src, size_t bytes);

/* FP64 direct sums at the targets idx[0..nidx) (main.cpp:206-245 regularisation):
 * potential out0[k] = sum_j v_j log(sqrt(|t - s_j|^2 + 10 eps)); force (out0, out1)[k] =
 * sum_j v_j (t - s_j) / (|t - s_j|^2 + 10 eps).  All pointers in device memory of the
 * context's GPU; synchronous.  For measuring treecode error at full size. */
TC_API int tc_direct_potential(tc_context *ctx, const double *xs, const double *ys, const double *vs, int n,
                               const double *xt, const double *yt, const int *idx, int nidx, double *out);
TC_API int tc_direct_force(tc_context *ctx, const double *xs, const double *ys, const double *vs, int n,
                           const double *xt, const double *yt, const int *idx, int nidx, double *outx,
                           double *outy);

/* Device-to-device copy by a plain kernel (16 B per thread, bytes a multiple of 16); the
 * known-size reference for scaling hardware memory counters.  Synchronous. */
TC_API int tc_copy_kernel(tc_context *ctx, void *dst, const void *src, size_t bytes);

/* Morton keys + stable radix sort of n points in device memory (the source-sort
 * stage on its own): box[4] out (xmin, ymin, ext, scale), keys/perm out. */
TC_API int tc_morton_sort_device(tc_context *ctx, const double *x, const double *y, int n, double *box,
                                 unsigned int *keys, unsigned int *perm);

/* ---- request server: continuous batching over slots (no counterpart in the reference) ----
 * Each slot is a context with its own stream and host thread; requests in different
 * slots overlap on the GPU.  Trees are cached by source_id in one cache that every slot
 * reads (at most 4 per slot; the least recently used idle tree's buffers are recycled
 * into the next build).  Admission counts what the server holds (trees, each slot's buffers)
 * plus an estimate of each running request's growth against the HBM budget, evicting
 * idle trees (least recently used) and idle slots' buffers to make room.  All request
 * pointers are host memory and must stay valid until tc_server_wait returns for that
 * ticket. */
typedef struct tc_server tc_server;

enum { TC_REQ_POTENTIAL = 0, TC_REQ_FORCE = 1 };

typedef struct tc_request {
    int kind;                         /* TC_REQ_POTENTIAL or TC_REQ_FORCE */
    int precision;                    /* 0 plain FP64, 1 FP32 bulk */
    double theta;
    const double *xs, *ys, *vs;       /* sources */
    int n;
    const double *xt, *yt;            /* potential targets */
    int m;
    const double *x0s, *y0s, *hs;     /* force: 32x32 blocks */
    int nblocks;
    double *out0, *out1;              /* potential: out0[m]; force: out0/out1[1024*nblocks] */
    long long source_id;              /* nonzero: same id = same sources, the tree may be reused */
    int order, leaf;                  /* expansion order and leaf capacity for this request, e.g. the
                                       * autotuner's choice for its size bucket; 0 = the reference's
                                       * (potential 12 / 512, force 24 / 192) */
} tc_request;

TC_API tc_server *tc_server_create(int device, int nslots, double hbm_budget_gib);
/* TC_SERVER_BATCH: the potential traversals of requests running in different slots at
 * the same time are launched together, as one kernel over all their targets (a graph
 * built per size bucket and replayed); a traversal with no other waiting is launched
 * alone.  Results are bit-identical to unbatched. */
enum {
    TC_SERVER_BATCH = 1,
    TC_SERVER_BATCH_ALWAYS = 2  /* also a lone traversal goes through the batch kernel and its graphs
                                 * (for testing the graphs; slower) */
};
TC_API tc_server *tc_server_create_ex(int device, int nslots, double hbm_budget_gib, int flags);
TC_API void tc_server_destroy(tc_server *srv);
/* Returns a ticket > 0, or < 0 if the request can never fit the HBM budget. */
TC_API long long tc_server_submit(tc_server *srv, const tc_request *req);
/* Blocks until the request finished; times_ms[5] = {queue, service, latency, reused, slot}.
 * Returns the request's status (0 ok, < 0 error). */
TC_API int tc_server_wait(tc_server *srv, long long ticket, double *times_ms);
/* out[7] = {completed, reused trees, failed, max in flight, peak admitted bytes, budget bytes, queued} */
TC_API int tc_server_stats(tc_server *srv, long long *out);
/* The first nout of {completed, reused trees, failed, max in flight, peak admitted bytes,
 * budget bytes, queued, batch launches, batched requests, evicted trees, trimmed slots,
 * cached trees, bytes held, requests that took more HBM than admission estimated, the
 * largest such excess in bytes, requests launched in batches of two or more}; returns how
 * many were written. */
TC_API int tc_server_stats_ex(tc_server *srv, long long *out, int nout);

#ifdef __cplusplus
}
#endif
#endif
