// tc_common.h -- shared definitions of the MI355X (gfx950) treecode library.
//
// Data layout in HBM (one tree per context):
//   sorted sources   double4 {x, y, v, 0}         32 B/source, Morton order
//   sorted keys      uint32  30-bit Morton key     4 B/source
//   permutation      uint32  sorted -> input index 4 B/source
//   nodes            TcNode (64 B), breadth-first ids: root 0, children of the
//                    k-th internal node (level, Morton order) at 1 + 4k + c
//   expansions       double[2p] per node: [Re a_1..a_p][Im a_1..a_p]
//   FP32 mirrors     float4 {x-xc, y-yc, v, 0} per source (leaf-relative) and
//                    float2[p] per node of scaled coefficients a_j / r^j
//                    (j a_j / r^j for force trees)
#pragma once

#include <hip/hip_runtime.h>

#include <cstddef>
#include <cstdint>
#include <functional>
#include <stdexcept>
#include <string>
#include <vector>

#define TC_LMAX 15
#define TC_MAXORDER 32
#define TC_STACK 48  // DFS stack: 3 * LMAX + 1 entries suffice
// Largest leaf capacity: source index + capacity stays below 2^31 (n < 2^30).
#define TC_MAX_LEAF (1 << 30)

// Expansion orders with compiled kernels: X(p) for each.
#define TC_ORDERS(X) X(6) X(8) X(9) X(10) X(12) X(14) X(16) X(20) X(24) X(32)
inline bool tc_order_supported(int p)
{
#define TC_ORDER_EQ(P) || p == P
    return false TC_ORDERS(TC_ORDER_EQ);
#undef TC_ORDER_EQ
}

struct alignas(16) TcNode {
    double xcom, ycom, r, mass;
    int s, e, child, level;  // child: first child id, 0 for a leaf
    int cx, cy;              // integer cell coordinates at `level`
    double rr;               // r * r (FP64, rounded as the product is): the opening test's left side
};
static_assert(sizeof(TcNode) == 64, "TcNode must stay 64 B (one s_load_dwordx16)");

// P2P_SLOTS = 64 x leaf size per wave-level leaf visit and E2P_SLOTS = 64 per
// wave-level multipole evaluation: the lane slots the masked loops occupy
// (P2P / P2P_SLOTS and E2P / E2P_SLOTS are their SIMD efficiencies).
// M2L_WAVE: FP32 potential only, multipole-to-local translations about a wave's
// target centre; E2P_WAVE: the (target, node) interactions they evaluate (a
// part of E2P, which counts every accepted (target, node) pair in any mode).
// CYC_*: wavefront clock cycles (shader clock, summed over waves) spent in
// each category of a traversal -- direct sums, per-target multipole
// evaluations, local-expansion translations (E2L / wave-level M2L), local
// evaluation (L2P) -- and in the whole traversal; the per-category shares of
// the reference's TimeDistrib (treecode-force.cpp:26-48).  Counted with the
// interaction counters.
enum TcCounter {
    TC_CNT_P2P = 0, TC_CNT_E2P = 1, TC_CNT_E2L = 2, TC_CNT_L2P = 3, TC_CNT_P2P_SLOTS = 4, TC_CNT_E2P_SLOTS = 5,
    TC_CNT_M2L_WAVE = 6, TC_CNT_E2P_WAVE = 7, TC_NPUBLIC_COUNTERS = 8,
    TC_CNT_CYC_P2P = 8, TC_CNT_CYC_E2P = 9, TC_CNT_CYC_LOCAL = 10, TC_CNT_CYC_L2P = 11, TC_CNT_CYC_TOTAL = 12,
    TC_NCOUNTERS = 13
};

struct TcError : std::runtime_error {
    int code;
    TcError(int c, const std::string &m) : std::runtime_error(m), code(c) {}
};

#define TC_HIP(call)                                                                                  \
    do {                                                                                              \
        hipError_t _e = (call);                                                                       \
        if (_e != hipSuccess)                                                                         \
            throw TcError(-2, std::string("HIP error ") + hipGetErrorString(_e) + " at " __FILE__ ":" + \
                                  std::to_string(__LINE__) + " in " #call);                          \
    } while (0)

// ------------------------------------------------------------------ device helpers
#ifdef __HIP_DEVICE_COMPILE__
#define TC_DEV __device__ __forceinline__
#else
#define TC_DEV __device__ inline
#endif

TC_DEV int tc_lane() { return __lane_id(); }

// Shader clock for the per-category cycle counters (one scalar instruction).
TC_DEV unsigned long long tc_clock() { return __builtin_readcyclecounter(); }

// Workgroups go to the 8 XCDs (each with its own L2) round-robin: workgroup i
// to XCD i mod 8.  This renumbers them so that each XCD takes C consecutive
// logical blocks at a time, chunks rotating over the XCDs: neighbouring
// blocks (nearby targets, which visit the same tree nodes) share an L2, and
// every XCD still gets work from the whole domain (contiguous eighths would
// put the dense regions on a few XCDs).  The tail beyond whole rounds of 8 C
// keeps its numbering.  A bijection of 0 .. gridDim.x - 1; C = 2^lc (C = 1 is
// the hardware's own round-robin order).
constexpr int TC_XCD_CHUNK_LOG2 = 4;  // 16 blocks per chunk (measured best at 2^24 targets)
TC_DEV int tc_xcd_chunk_block(int lc)
{
    const int bi = blockIdx.x, span = 8 << lc, full = (int)(gridDim.x >> (3 + lc)) << (3 + lc);
    return bi < full ? (bi & ~(span - 1)) + ((bi & 7) << lc) + ((bi >> 3) & ((1 << lc) - 1)) : bi;
}

// Orders LDS traffic between lanes of one wavefront (no instruction, only a
// code-motion and memory-model fence at wavefront scope).
TC_DEV void tc_wave_sync()
{
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

TC_DEV int tc_uniform(int v) { return __builtin_amdgcn_readfirstlane(v); }

TC_DEV double tc_uniform(double v)
{
    const long long b = __double_as_longlong(v);
    const int lo = __builtin_amdgcn_readfirstlane((int)(b & 0xffffffff));
    const int hi = __builtin_amdgcn_readfirstlane((int)(b >> 32));
    return __longlong_as_double(((long long)hi << 32) | (unsigned int)lo);
}

// Value of v in lane l (wave-uniform result).
TC_DEV double tc_lane_value(double v, int l)
{
    const long long b = __double_as_longlong(v);
    const int lo = __builtin_amdgcn_readlane((int)(b & 0xffffffff), l);
    const int hi = __builtin_amdgcn_readlane((int)(b >> 32), l);
    return __longlong_as_double(((long long)hi << 32) | (unsigned int)lo);
}

TC_DEV unsigned long long tc_uniform64(unsigned long long v)
{
    const unsigned lo = __builtin_amdgcn_readfirstlane((unsigned)(v & 0xffffffffu));
    const unsigned hi = __builtin_amdgcn_readfirstlane((unsigned)(v >> 32));
    return ((unsigned long long)hi << 32) | lo;
}

TC_DEV unsigned long long tc_ballot(bool p) { return __ballot(p); }

TC_DEV unsigned long long tc_lanemask_lt() { return (1ull << tc_lane()) - 1ull; }

// Canonical pairwise tree over the 64 lanes (lane j gets p[j] + p[j ^ off] for
// off = 32..1); lane 0 ends with ((p0+p32)+(p16+p48))+..., the same order the
// oracle uses for leaf sums.
TC_DEV double tc_wave_tree_sum(double v)
{
#pragma unroll
    for (int off = 32; off >= 1; off >>= 1)
        v = v + __shfl_xor(v, off, 64);
    return v;
}

// Sum over the wavefront, butterfly order (every lane holds a sum of all 64
// values; lanes may differ in the last bit).
TC_DEV float tc_wave_sum_f32(float v)
{
#pragma unroll
    for (int off = 32; off >= 1; off >>= 1)
        v = v + __shfl_xor(v, off, 64);
    return v;
}

TC_DEV double tc_wave_max(double v)
{
#pragma unroll
    for (int off = 32; off >= 1; off >>= 1) {
        const double o = __shfl_xor(v, off, 64);
        v = v < o ? o : v;
    }
    return v;
}

TC_DEV uint32_t tc_spread15(uint32_t x)
{
    x = (x | (x << 8)) & 0x00FF00FFu;
    x = (x | (x << 4)) & 0x0F0F0F0Fu;
    x = (x | (x << 2)) & 0x33333333u;
    x = (x | (x << 1)) & 0x55555555u;
    return x;
}

TC_DEV uint32_t tc_compact15(uint32_t x)
{
    x &= 0x55555555u;
    x = (x | (x >> 1)) & 0x33333333u;
    x = (x | (x >> 2)) & 0x0F0F0F0Fu;
    x = (x | (x >> 4)) & 0x00FF00FFu;
    x = (x | (x >> 8)) & 0x0000FFFFu;
    return x;
}

// Deepest level at which the Morton prefixes of a and b agree (0..15).
TC_DEV int tc_common_level(uint32_t a, uint32_t b)
{
    const uint32_t x = a ^ b;
    if (!x)
        return TC_LMAX;
    const int hb = 31 - __clz((int)x);
    return (TC_LMAX - 1) - (hb >> 1);
}

// ------------------------------------------------------------------ host-side internals
// Bump allocator over device memory the library does not own (the static
// scratch array of the no-allocation entry point).  Buffers bound to an arena
// carve their storage from it and never call hipMalloc or hipFree.
struct TcArena {
    char *base = nullptr;
    size_t cap = 0, used = 0;
    void *carve(size_t bytes)
    {
        const size_t o = (used + 255) & ~size_t(255);
        if (o > cap || bytes > cap - o)
            throw TcError(-3, "treecode: the request needs more than the " + std::to_string(cap >> 20) +
                                  " MiB of the no-allocation scratch");
        used = o + bytes;
        return base + o;
    }
};

// Set by a thread that synchronises every stream its work used before the
// listed memory is freed (the request server's slot threads, per request):
// DevBuf::reserve then hands a grown buffer's old memory (pointer, bytes) to
// this list instead of freeing it, because hipFree waits for the whole device
// and would stall the thread behind other threads' kernels.  Null: freed at once.
struct TcDeferredFree {
    void *ptr;
    size_t bytes;
};
extern thread_local std::vector<TcDeferredFree> *tc_defer_free;

struct DevBuf {
    void *ptr = nullptr;
    size_t cap = 0;
    TcArena *arena = nullptr;  // non-null: storage carved from the arena, never freed
    template <class T> T *as() const { return static_cast<T *>(ptr); }
    void reserve(size_t bytes)
    {
        if (bytes <= cap)
            return;
        size_t b = bytes < 256 ? 256 : bytes;
        if (arena) {
            // a larger request takes a fresh piece; the old one stays unused
            // until the arena is reset
            ptr = arena->carve(b);
            cap = b;
            return;
        }
        if (ptr) {
            if (tc_defer_free)
                tc_defer_free->push_back({ptr, cap});
            else
                TC_HIP(hipFree(ptr));
        }
        ptr = nullptr;
        cap = 0;
        b = (b + 4095) & ~size_t(4095);
        TC_HIP(hipMalloc(&ptr, b));
        cap = b;
    }
    // Binds the buffer to an arena (or back to hipMalloc with nullptr); the
    // buffer must not own memory.
    void bind(TcArena *a)
    {
        arena = a;
        ptr = nullptr;
        cap = 0;
    }
    void release()
    {
        if (ptr && !arena)
            (void)hipFree(ptr);
        ptr = nullptr;
        cap = 0;
    }
};

// A built tree resident in HBM.
struct TcTree {
    int n = 0, order = 0, leafcap = 0;
    int nnodes = 0, ninternal = 0;
    int level_start[TC_LMAX + 2] = {0};  // internal ranks of level L: [level_start[L], level_start[L+1])
    double box[4] = {0, 0, 0, 0};        // xmin, ymin, ext, scale (host copy)
    bool has_fp32 = false;
    bool deriv32 = false;  // FP32 coefficient mirror holds k b_k
    DevBuf keys, perm, src, src32, nodes, exps, exps32, nodew, selfid;
    DevBuf box_d;
    void release()
    {
        for (DevBuf *b : {&keys, &perm, &src, &src32, &nodes, &exps, &exps32, &nodew, &selfid, &box_d})
            b->release();
    }
    void bind(TcArena *a)
    {
        for (DevBuf *b : {&keys, &perm, &src, &src32, &nodes, &exps, &exps32, &nodew, &selfid, &box_d})
            b->bind(a);
    }
};

// Scratch reused across calls of one context.
struct TcScratch {
    DevBuf k2, v2, hist, part, scan_tmp, istart, ilevel, info, tkeys, tperm, tsorted, bbox_part, ints, packed, work,
        flist;
    // pinned host memory for the per-solve readbacks (tree census, flags): a
    // copy into pageable memory costs a staged transfer of its own
    void *pinned = nullptr;
    static constexpr size_t PINNED_BYTES = 256;
    void *host_pinned()
    {
        if (!pinned)
            TC_HIP(hipHostMalloc(&pinned, PINNED_BYTES, hipHostMallocDefault));
        return pinned;
    }
    void release()
    {
        for (DevBuf *b : {&k2, &v2, &hist, &part, &scan_tmp, &istart, &ilevel, &info, &tkeys, &tperm, &tsorted,
                          &bbox_part, &ints, &packed, &work, &flist})
            b->release();
        if (pinned)
            (void)hipHostFree(pinned);
        pinned = nullptr;
    }
    // Device buffers from the arena; the pinned readback buffer must already
    // exist (host_pinned() at set-up), so no call made with this scratch allocates.
    void bind(TcArena *a)
    {
        for (DevBuf *b : {&k2, &v2, &hist, &part, &scan_tmp, &istart, &ilevel, &info, &tkeys, &tperm, &tsorted,
                          &bbox_part, &ints, &packed, &work, &flist})
            b->bind(a);
    }
};

// Request-server internals (tc_api.cpp, used by tc_serve.cpp).  tc_srv_solve
// runs request r (host arrays) on context c against the shared tree t: built
// here when build is set (then `ready` is recorded on c's stream after it), or
// built earlier by another context whose stream recorded `ready`.  eval, when
// set, takes the potential traversal (the server's batcher).
struct tc_context;
struct tc_request;
struct TcPotBatchItem;
int tc_srv_solve(tc_context *c, const tc_request &r, TcTree &t, bool build, hipEvent_t ready,
                 const std::function<void(const TcPotBatchItem &, hipStream_t)> *eval);
// Device bytes held by a context's scratch and staging buffers; release them.
size_t tc_srv_context_bytes(tc_context *c);
void tc_srv_context_trim(tc_context *c);
size_t tc_tree_bytes(const TcTree &t);
// Sets the calling thread's treecode_last_error message.
void tc_set_last_error(const std::string &m);

// Static device scratch of the no-allocation entry point (tc_noalloc.hip):
// its size, and its address on the current device (looked up once per device
// with hipGetSymbolAddress, then cached).
size_t tc_noalloc_scratch_bytes();
void *tc_noalloc_scratch_lookup(int device);

// Internal entry points (tc_build.hip, tc_eval.hip)
// Bounding box, keys (Morton, or Hilbert for targets) and stable sort of n
// points; with src_out also the sorted records {x, y, v, 0}.  One workgroup
// for n <= 16384, else a bounding-box partial kernel and tc_morton_sort.
// ev (optional, two events): recorded after the keys and after the sorted records.
void tc_sort_points(const double *x, const double *y, const double *v, int n, double *box_d, uint32_t *keys_out,
                    uint32_t *perm_out, double4 *src_out, TcScratch &ws, hipStream_t st, bool hilbert,
                    const hipEvent_t *ev = nullptr);
// Keys and the four one-launch radix passes.  lb: the passes' digit totals,
// tickets and published tile counts, zeroed beforehand by the bounding-box
// partial kernel, whose partials the key kernel then finishes into box_d (as
// tc_sort_points does).
void tc_morton_sort(const double *x, const double *y, const double *v, int n, const double *box_d,
                    uint32_t *keys_out, uint32_t *perm_out, double4 *packed, TcScratch &ws, hipStream_t st,
                    bool hilbert, const hipEvent_t *ev, uint32_t *lb);
// fp32: also build the FP32 mirrors; deriv32: the coefficient mirror holds
// k b_k (force trees) instead of b_k (potential trees).
// ev (optional, four events): recorded after the Morton keys, the sorted
// records, the topology and the upward pass (the build's phase boundaries).
// at_census (optional): called once the census readback is queued, before the
// host waits for it (the caller can queue independent work on another stream).
void tc_build_tree(TcTree &t, const double *x, const double *y, const double *v, int n, int leafcap, int order,
                   bool fp32, bool deriv32, TcScratch &ws, hipStream_t st, const hipEvent_t *ev = nullptr,
                   const std::function<void()> *at_census = nullptr);
void tc_sort_targets(const double *xt, const double *yt, int m, TcScratch &ws, hipStream_t st);
// Several target sets of at most TC_SORT_SMALL_MAX targets each sorted in one
// launch: tc_sort_targets_desc reserves ws's target buffers for m targets (as
// tc_sort_targets would) and returns where the sort of xt, yt goes; a device
// array of k such descriptors is then sorted by tc_sort_targets_batch, each
// with the results tc_sort_targets gives it alone.
constexpr int TC_SORT_SMALL_MAX = 16384;
struct TcSortDesc {
    const double *x, *y;
    double *box;
    uint32_t *bad, *keys, *perm;
    int n, pad;
};
TcSortDesc tc_sort_targets_desc(const double *xt, const double *yt, int m, TcScratch &ws);
void tc_sort_targets_batch(const TcSortDesc *desc_dev, int k, hipStream_t st);
// Device flag (uint32, nonzero = a non-finite coordinate) set by the last
// tc_sort_targets call on m targets; valid until the next bounding-box pass.
const uint32_t *tc_targets_bad_flag(TcScratch &ws, int m);
// sort_targets false: tc_sort_targets already ran on ws for these targets
// (e.g. on another stream, ordered before st).
// One request of a batched potential evaluation: a resident tree and targets
// already sorted (tperm), results to out (all device memory).
struct TcPotBatchItem {
    const TcTree *tree;
    double theta;
    const double *xt, *yt;
    const uint32_t *tperm;
    int m;
    double *out;
};

// Workspace of batched potential evaluations on one stream: device
// descriptors, a ring of pinned host copies of them (a ring slot is reused
// once the copy made from it has run), and the replayable graph of each
// (order, precision, ring slot).
struct TcBatchWs {
    // size buckets: 2^6 .. 2^(6 + NBUCKET - 1) waves (64 targets each)
    static constexpr int MAXREQ = 64, RING = 16, NBUCKET = 20;
    DevBuf dev;              // RING descriptor blocks
    void *host = nullptr;    // RING pinned descriptor blocks
    hipEvent_t done[RING] = {};  // recorded after the launch that read the block
    bool pending[RING] = {};
    int next = 0;
    // [(2 order + precision) * NBUCKET + bucket][ring slot]
    hipGraphExec_t graph[2 * (TC_MAXORDER + 1) * NBUCKET][RING] = {};
    void release();
};
// Evaluates k <= TcBatchWs::MAXREQ requests (same order and precision) in one
// launch of k_potential_batch on st; with graph, through the graph built for
// this (order, precision, size bucket, descriptor block).  Asynchronous; one
// caller at a time per workspace (launches may go to different streams).
void tc_potential_batch(const TcPotBatchItem *items, int k, int order, int precision, TcBatchWs &ws,
                        unsigned long long *counters, hipStream_t st, bool graph);
// One request's traversal by the one-request kernel (k_potential) on st.
void tc_potential_item(const TcPotBatchItem &it, int precision, unsigned long long *counters, hipStream_t st);

// xcd_chunk_log2: the traversal's workgroups go to the XCDs in chunks of
// 2^xcd_chunk_log2 consecutive blocks (a tiling option; results do not depend on it).
void tc_eval_potential(const TcTree &t, double theta, const double *xt, const double *yt, int m, double *out,
                       int precision, unsigned long long *counters, TcScratch &ws, hipStream_t st,
                       hipEvent_t ev0, hipEvent_t ev1, bool sort_targets = true,
                       int xcd_chunk_log2 = TC_XCD_CHUNK_LOG2);
// inline_leaves: direct-sum leaves a brick sums inside the traversal before the
// rest go to chunks summed by other waves (a tiling option; results differ only
// in summation order).
void tc_eval_force(const TcTree &t, double theta, const double *x0s, const double *y0s, const double *hs,
                   int nblocks, double *xr, double *yr, int precision, unsigned long long *counters,
                   TcScratch &ws, int max_chunks, int inline_leaves, hipStream_t st, hipEvent_t ev0, hipEvent_t ev1);
void tc_eval_force_targets(const TcTree &t, double theta, const double *xt, const double *yt, int m, double *xr,
                           double *yr, int precision, unsigned long long *counters, TcScratch &ws,
                           hipStream_t st, hipEvent_t ev0, hipEvent_t ev1);
void tc_copy16(const void *src, void *dst, size_t bytes, hipStream_t st);
void tc_direct_sums(const double *xs, const double *ys, const double *vs, int n, const double *xt,
                    const double *yt, const int *idx, int nidx, int force, double *out0, double *out1,
                    hipStream_t st);
void tc_scan_exclusive_u32(const uint32_t *in, uint32_t *out, int n, uint32_t *total_d, TcScratch &ws,
                           hipStream_t st);
void tc_fill_u32(uint32_t *p, uint32_t v, size_t n, hipStream_t st);
void tc_upload_binomials(hipStream_t st);
