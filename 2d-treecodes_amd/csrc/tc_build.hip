// tc_build.hip -- Tree::build on gfx950: bbox, Morton keys, LDS radix sort,
// quadtree topology from sorted keys, upward pass (P2E leaves, E2E parents).
//
// Replaces drivers/sort-sources.cpp:22-132, drivers/upward.cpp:61-314,
// kernels/upward-common.ispc:13-150 and kernels/upward-kernels.m4:18-149.
// Compiled with -ffp-contract=off: keys, centres of mass and radii must be
// rounded exactly as the canonical CPU restatement rounds them, because every
// traversal decision depends on them.
#include "tc_common.h"

#include <cfloat>
#include <climits>
#include <cstring>
#include <vector>

namespace {

constexpr int RS_NT = 256;     // radix sort block size (4 waves)
constexpr int RS_ITEMS = 20;   // keys per thread per tile
constexpr int RS_TILE = RS_NT * RS_ITEMS;
constexpr int RS_BINS = 256;   // 8-bit digits: passes at bits 0, 8, 16, 24 (30-bit keys)
constexpr int KH_GRID = 1024;  // most workgroups of the key kernel (each adds its digit totals once)
constexpr int SCAN_NT = 1024;
constexpr int SCAN_ITEMS = 8;
constexpr int SCAN_CHUNK = SCAN_NT * SCAN_ITEMS;
constexpr int TB_CHUNK = 1024;  // positions per wave in the internal-cell census

__constant__ double c_binom[64 * 64];  // C(n, k) rounded to double, n, k < 64

// ------------------------------------------------------------------ scans
template <int NT>
TC_DEV uint32_t block_scan_excl(uint32_t v, uint32_t *s_w, uint32_t *total)
{
    constexpr int NW = NT / 64;
    const int lane = tc_lane(), wid = threadIdx.x >> 6;
    uint32_t x = v;
#pragma unroll
    for (int off = 1; off < 64; off <<= 1) {
        const uint32_t y = __shfl_up(x, off, 64);
        if (lane >= off)
            x += y;
    }
    if (lane == 63)
        s_w[wid] = x;
    __syncthreads();
    if (wid == 0) {
        const uint32_t w = lane < NW ? s_w[lane] : 0u;
        uint32_t z = w;
#pragma unroll
        for (int off = 1; off < 64; off <<= 1) {
            const uint32_t y = __shfl_up(z, off, 64);
            if (lane >= off)
                z += y;
        }
        if (lane < NW)
            s_w[lane] = z - w;
        if (lane == NW - 1)
            s_w[NW] = z;
    }
    __syncthreads();
    const uint32_t res = s_w[wid] + x - v;
    if (total)
        *total = s_w[NW];
    __syncthreads();
    return res;
}

// Single-block exclusive scan of n items with a running carry.
__global__ __launch_bounds__(SCAN_NT) void k_scan_single(const uint32_t *in, uint32_t *out, int n, uint32_t *total)
{
    __shared__ uint32_t s_w[SCAN_NT / 64 + 1];
    uint32_t carry = 0;
    for (int base = 0; base < n; base += SCAN_CHUNK) {
        uint32_t v[SCAN_ITEMS];
        uint32_t sum = 0;
        const int i0 = base + threadIdx.x * SCAN_ITEMS;
#pragma unroll
        for (int j = 0; j < SCAN_ITEMS; ++j) {
            v[j] = (i0 + j < n) ? in[i0 + j] : 0u;
            sum += v[j];
        }
        uint32_t tot;
        uint32_t pre = block_scan_excl<SCAN_NT>(sum, s_w, &tot) + carry;
#pragma unroll
        for (int j = 0; j < SCAN_ITEMS; ++j) {
            if (i0 + j < n)
                out[i0 + j] = pre;
            pre += v[j];
        }
        carry += tot;
    }
    if (threadIdx.x == 0 && total)
        *total = carry;
}

__global__ __launch_bounds__(SCAN_NT) void k_scan_partial(const uint32_t *in, int n, uint32_t *part)
{
    __shared__ uint32_t s_w[SCAN_NT / 64 + 1];
    const int i0 = blockIdx.x * SCAN_CHUNK + threadIdx.x * SCAN_ITEMS;
    uint32_t sum = 0;
#pragma unroll
    for (int j = 0; j < SCAN_ITEMS; ++j)
        sum += (i0 + j < n) ? in[i0 + j] : 0u;
    uint32_t tot;
    block_scan_excl<SCAN_NT>(sum, s_w, &tot);
    if (threadIdx.x == 0)
        part[blockIdx.x] = tot;
}

__global__ __launch_bounds__(SCAN_NT) void k_scan_final(const uint32_t *in, uint32_t *out, int n,
                                                         const uint32_t *part)
{
    __shared__ uint32_t s_w[SCAN_NT / 64 + 1];
    const int i0 = blockIdx.x * SCAN_CHUNK + threadIdx.x * SCAN_ITEMS;
    uint32_t v[SCAN_ITEMS];
    uint32_t sum = 0;
#pragma unroll
    for (int j = 0; j < SCAN_ITEMS; ++j) {
        v[j] = (i0 + j < n) ? in[i0 + j] : 0u;
        sum += v[j];
    }
    uint32_t pre = block_scan_excl<SCAN_NT>(sum, s_w, nullptr) + part[blockIdx.x];
#pragma unroll
    for (int j = 0; j < SCAN_ITEMS; ++j) {
        if (i0 + j < n)
            out[i0 + j] = pre;
        pre += v[j];
    }
}

__global__ void k_fill_u32(uint32_t *p, uint32_t v, size_t n)
{
    for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x)
        p[i] = v;
}

// ------------------------------------------------------------------ bbox
// sort-sources.cpp:44-82 with upward-common.ispc:126-150's accumulators:
// min starts at 1e13, max at 0.  Non-finite coordinates are ORed into *bad
// (cleared before), or with pbad recorded per workgroup in pbad[blockIdx.x].
// Also zeroes zero[0 .. nzero) (the one-launch radix passes' counters, used
// after this kernel).
__global__ __launch_bounds__(256) void k_bbox_partial(const double *x, const double *y, int n, double4 *part,
                                                       uint32_t *bad, uint32_t *pbad, uint32_t *zero, int nzero)
{
    __shared__ double4 s[4];
    for (int i = blockIdx.x * 256 + threadIdx.x; i < nzero; i += gridDim.x * 256)
        zero[i] = 0u;
    double xmi = 1e13, ymi = 1e13, xma = 0, yma = 0;
    bool nonfinite = false;
    for (int i = blockIdx.x * 256 + threadIdx.x; i < n; i += gridDim.x * 256) {
        const double a = x[i], b = y[i];
        nonfinite |= !isfinite(a) || !isfinite(b);
        xmi = a < xmi ? a : xmi;
        xma = xma < a ? a : xma;
        ymi = b < ymi ? b : ymi;
        yma = yma < b ? b : yma;
    }
#pragma unroll
    for (int off = 32; off >= 1; off >>= 1) {
        double o;
        o = __shfl_xor(xmi, off, 64); xmi = o < xmi ? o : xmi;
        o = __shfl_xor(ymi, off, 64); ymi = o < ymi ? o : ymi;
        o = __shfl_xor(xma, off, 64); xma = xma < o ? o : xma;
        o = __shfl_xor(yma, off, 64); yma = yma < o ? o : yma;
    }
    __shared__ uint32_t s_bad;
    if (threadIdx.x == 0)
        s_bad = 0u;
    __syncthreads();
    if (tc_ballot(nonfinite) && tc_lane() == 0) {
        if (pbad)
            s_bad = 1u;
        else
            atomicOr(bad, 1u);
    }
    const int wid = threadIdx.x >> 6;
    if (tc_lane() == 0)
        s[wid] = make_double4(xmi, xma, ymi, yma);
    __syncthreads();
    if (threadIdx.x == 0) {
        double4 r = s[0];
        for (int w = 1; w < 4; ++w) {
            r.x = s[w].x < r.x ? s[w].x : r.x;
            r.y = r.y < s[w].y ? s[w].y : r.y;
            r.z = s[w].z < r.z ? s[w].z : r.z;
            r.w = r.w < s[w].w ? s[w].w : r.w;
        }
        part[blockIdx.x] = r;
        if (pbad)
            pbad[blockIdx.x] = s_bad;
    }
}

// The box from the partials (one wave): min / max over part[0 .. np), then
// the reference's padded extent.  box[4] = {xmin, ymin, ext, 2^15 / ext}.
TC_DEV void bbox_finish(const double4 *part, int np, double *box)
{
    double xmi = 1e13, xma = 0, ymi = 1e13, yma = 0;
    for (int i = tc_lane(); i < np; i += 64) {
        const double4 r = part[i];
        xmi = r.x < xmi ? r.x : xmi;
        xma = xma < r.y ? r.y : xma;
        ymi = r.z < ymi ? r.z : ymi;
        yma = yma < r.w ? r.w : yma;
    }
#pragma unroll
    for (int off = 32; off >= 1; off >>= 1) {
        double o;
        o = __shfl_xor(xmi, off, 64); xmi = o < xmi ? o : xmi;
        o = __shfl_xor(ymi, off, 64); ymi = o < ymi ? o : ymi;
        o = __shfl_xor(xma, off, 64); xma = xma < o ? o : xma;
        o = __shfl_xor(yma, off, 64); yma = yma < o ? o : yma;
    }
    if (tc_lane() == 0) {
        const double ext0 = xma - xmi, ext1 = yma - ymi;
        const double eps = 10000 * DBL_EPSILON;
        double ext = (ext0 < ext1 ? ext1 : ext0) * (1 + 2 * eps);
        if (!(ext > 0))
            ext = 1.0;
        box[0] = xmi - eps * ext;
        box[1] = ymi - eps * ext;
        box[2] = ext;
        box[3] = (1.0 / ext) * (double)(1 << TC_LMAX);
    }
}

// ------------------------------------------------------------------ keys
TC_DEV uint32_t quantise(double v, double vmin, double scale)
{
    double q = floor((v - vmin) * scale);
    if (!(q >= 0))
        q = 0;
    if (q > (double)((1 << TC_LMAX) - 1))
        q = (double)((1 << TC_LMAX) - 1);
    return (uint32_t)q;
}

// Hilbert index of the quantised point (the standard quadrant-rotation
// construction on a 2^15 x 2^15 grid, 30-bit result).  Used to order targets:
// a Hilbert run of 64 points is more compact than a Morton run (no jumps
// between distant quadrants), so the 64 lanes of a traversal wave agree on
// more accept/open decisions.
TC_DEV uint32_t hilbert15(uint32_t x, uint32_t y)
{
    const uint32_t n1 = (1u << TC_LMAX) - 1u;
    uint32_t d = 0;
    for (uint32_t s = 1u << (TC_LMAX - 1); s > 0; s >>= 1) {
        const uint32_t rx = (x & s) ? 1u : 0u;
        const uint32_t ry = (y & s) ? 1u : 0u;
        d += s * s * ((3u * rx) ^ ry);
        if (ry == 0) {
            if (rx == 1) {
                x = n1 - x;
                y = n1 - y;
            }
            const uint32_t t = x;
            x = y;
            y = t;
        }
    }
    return d;
}

// ------------------------------------------------------------------ radix sort
// Stable LSD radix sort of (key, index) pairs, 8-bit digits, one launch per
// pass (k_rs_pass_lb): each tile of RS_TILE keys is ranked stably in LDS with
// wavefront ballots, digit d of tile t is placed at (keys of smaller digits,
// from the digit totals the key kernel adds up) + (digit d in earlier tiles,
// by look-back over the counts the earlier tiles publish), and the tile's digit
// runs are written contiguously.
// Lanes of the wave holding the same 8-bit digit d as this lane (among lanes
// with ok set): a lane differs from this one where, for some bit, its ballot
// bit differs from this lane's own bit.  The per-lane 0 / ~0 of each bit is a
// sign-extended bit field, so each bit costs a ballot and four 32-bit logic ops.
TC_DEV unsigned long long wave_digit_peers(uint32_t d, bool ok)
{
    uint32_t diff_lo = 0, diff_hi = 0;
#pragma unroll
    for (int b = 0; b < 8; ++b) {
        const unsigned long long bb = tc_ballot((d >> b) & 1u);
        const uint32_t own = (uint32_t)((int32_t)(d << (31 - b)) >> 31);
        diff_lo |= (uint32_t)bb ^ own;
        diff_hi |= (uint32_t)(bb >> 32) ^ own;
    }
    const unsigned long long diff = ((unsigned long long)diff_hi << 32) | diff_lo;
    return tc_ballot(ok) & ~diff;
}

// One tile of a radix pass: the tile's keys are ranked stably in LDS and
// written out as contiguous digit runs.  start(tot, s_w) is called by every
// thread of the block after the ranking (it may use the block scratch s_w for
// collective scans); for digit threadIdx.x, of which the tile holds tot keys,
// it returns the global position of the tile's first key of that digit.
template <class Start>
TC_DEV void rs_tile(const uint32_t *kin, const uint32_t *vin, uint32_t *kout, uint32_t *vout, int n, int shift,
                    int tile, Start &&start)
{
    constexpr int NW = RS_NT / 64, WITEMS = 64 * RS_ITEMS;  // waves per tile, keys per wave
    __shared__ uint32_t s_k[RS_TILE];
    __shared__ uint32_t s_v[RS_TILE];
    __shared__ uint32_t s_tileoff[RS_BINS];
    __shared__ uint32_t s_gbase[RS_BINS];
    __shared__ uint32_t s_wrun[NW][RS_BINS];  // per-wave digit counts, then per-wave tile offsets
    __shared__ uint32_t s_w[NW + 1];

    const int tid = threadIdx.x, lane = tc_lane(), wid = tid >> 6;
    const int base = tile * RS_TILE;
    const int cnt = min(RS_TILE, n - base);
    for (int d = lane; d < RS_BINS; d += 64)
        s_wrun[wid][d] = 0;
    tc_wave_sync();

    // Each wave ranks its own contiguous WITEMS keys (item j of lane l is key
    // base + wid*WITEMS + 64 j + l), so the order (wave, j, lane) is the input
    // order and the ranks are stable without block barriers: peers with the
    // same digit are found with 8 ballots, the wave's running count per digit
    // lives in s_wrun[wid] (all lanes read it before the leader updates it).
    // Bounds are tested on the offset within the tile, so base + offset is
    // only formed below n (no int overflow for n near INT_MAX).
    uint32_t k[RS_ITEMS], v[RS_ITEMS], dr[RS_ITEMS];
#pragma unroll
    for (int j = 0; j < RS_ITEMS; ++j) {
        const int o = wid * WITEMS + j * 64 + lane;
        const bool ok = o < cnt;
        const int i = base + (ok ? o : 0);
        k[j] = ok ? kin[i] : 0u;
        v[j] = ok ? (vin ? vin[i] : (uint32_t)i) : 0u;
    }
#pragma unroll
    for (int j = 0; j < RS_ITEMS; ++j) {
        const bool ok = wid * WITEMS + j * 64 + lane < cnt;
        const uint32_t d = (k[j] >> shift) & (RS_BINS - 1);
        const unsigned long long peers = wave_digit_peers(d, ok);
        const uint32_t lower = __popcll(peers & tc_lanemask_lt());
        const uint32_t run = ok ? s_wrun[wid][d] : 0u;
        tc_wave_sync();
        if (ok && lower == 0)
            s_wrun[wid][d] = run + (uint32_t)__popcll(peers);
        tc_wave_sync();
        dr[j] = ((run + lower) << 8) | d;
    }
    __syncthreads();
    {
        // digit tid: its global start, then the tile offset = exclusive scan
        // over digits of the tile counts; wave w's keys of that digit start
        // after waves 0..w-1
        uint32_t c[NW], tot = 0;
#pragma unroll
        for (int w = 0; w < NW; ++w) {
            c[w] = s_wrun[w][tid];
            tot += c[w];
        }
        const uint32_t g = start(tot, s_w);
        uint32_t off = block_scan_excl<RS_NT>(tot, s_w, nullptr);
        s_gbase[tid] = g;
        s_tileoff[tid] = off;
#pragma unroll
        for (int w = 0; w < NW; ++w) {
            s_wrun[w][tid] = off;
            off += c[w];
        }
    }
    __syncthreads();
#pragma unroll
    for (int j = 0; j < RS_ITEMS; ++j) {
        if (wid * WITEMS + j * 64 + lane < cnt) {
            const uint32_t pos = s_wrun[wid][dr[j] & 255u] + (dr[j] >> 8);
            s_k[pos] = k[j];
            s_v[pos] = v[j];
        }
    }
    __syncthreads();

    for (int p = tid; p < cnt; p += RS_NT) {
        const uint32_t kk = s_k[p];
        const uint32_t d = (kk >> shift) & (RS_BINS - 1);
        const uint32_t dst = s_gbase[d] + (uint32_t)p - s_tileoff[d];
        kout[dst] = kk;
        vout[dst] = s_v[p];
    }
}

// Above SS_N points each pass is one launch.  The region (zeroed by
// k_bbox_partial before the keys) holds hist4[4][RS_BINS], every pass's digit
// totals over all keys (k_keys_hist); one ticket per pass (64 words); and for
// each pass the tiles' published digit counts, flags[ntiles][RS_BINS] as 64-bit
// words: status in the high half (LB_AGG: the tile's own count, LB_INC: its
// count plus those of all earlier tiles), the count in the low half.
constexpr int LB_HIST = 4 * RS_BINS;
constexpr int LB_TICKETS = 64;
constexpr unsigned long long LB_AGG = 1ull << 32, LB_INC = 2ull << 32;
constexpr int RS_SPIN_MAX = 1 << 22;  // polls of one unpublished count before the pass gives up
constexpr int RS_LB_BATCH = 4;        // earlier tiles' counts read per round trip

// 32-bit words of the region for a sort of ntiles tiles
static size_t lb_words(int ntiles) { return (size_t)LB_HIST + LB_TICKETS + (size_t)4 * 2 * RS_BINS * ntiles; }

// One pass of the sort (decoupled look-back).  Tiles are taken in ticket
// order; a tile publishes its digit counts (LB_AGG) as soon as it has ranked
// its keys, then walks back over the tiles before it, adding their counts,
// until it meets one that has published LB_INC, and publishes its own LB_INC.
// A tile only ever waits for an LB_AGG, which a tile with a smaller ticket
// (started earlier) publishes without waiting: every wait ends.  Digit d of
// tile t starts at (keys of smaller digits, from hist) + (digit d in tiles
// 0..t-1), exactly as with a scan of the tile histograms.  A count that stays
// unpublished for RS_SPIN_MAX polls sets bit 1 of *err (reported by the host
// as a failed sort) and ends the walk.
__global__ __launch_bounds__(RS_NT) void k_rs_pass_lb(const uint32_t *kin, const uint32_t *vin, uint32_t *kout,
                                                       uint32_t *vout, int n, int shift, const uint32_t *hist,
                                                       unsigned long long *flags, uint32_t *ticket, uint32_t *err)
{
    __shared__ int s_tile;
    if (threadIdx.x == 0)
        s_tile = (int)atomicAdd(ticket, 1u);
    __syncthreads();
    const int tile = s_tile;
    rs_tile(kin, vin, kout, vout, n, shift, tile, [&](uint32_t tot, uint32_t *s_w) {
        const int d = threadIdx.x;
        unsigned long long *fd = flags + d;  // fd[t * RS_BINS]: digit d of tile t
        __hip_atomic_store(&fd[(size_t)tile * RS_BINS], (tile ? LB_AGG : LB_INC) | tot, __ATOMIC_RELAXED,
                           __HIP_MEMORY_SCOPE_AGENT);
        uint32_t pre = 0;
        bool done = tile == 0;
        int t = tile;  // tiles t .. tile-1 are summed
        while (!done) {
            // the counts of tiles t-1, t-2, ..., RS_LB_BATCH loads in flight at once
            const int nt = min(RS_LB_BATCH, t);
            unsigned long long c[RS_LB_BATCH];
#pragma unroll
            for (int j = 0; j < RS_LB_BATCH; ++j)
                c[j] = j < nt ? __hip_atomic_load(&fd[(size_t)(t - 1 - j) * RS_BINS], __ATOMIC_RELAXED,
                                                  __HIP_MEMORY_SCOPE_AGENT)
                              : 0ull;
#pragma unroll
            for (int j = 0; j < RS_LB_BATCH; ++j) {
                if (!done && j < nt) {
                    unsigned long long w = c[j];
                    // tile t-1-j has not published yet: wait for it
                    for (int spin = 0; !(w >> 32) && spin < RS_SPIN_MAX; ++spin) {
                        __builtin_amdgcn_s_sleep(2);
                        w = __hip_atomic_load(&fd[(size_t)(t - 1 - j) * RS_BINS], __ATOMIC_RELAXED,
                                              __HIP_MEMORY_SCOPE_AGENT);
                    }
                    if (!(w >> 32))
                        atomicOr(err, 2u);
                    pre += (uint32_t)w;
                    done = (w & LB_INC) || !(w >> 32);
                }
            }
            t -= nt;
        }
        if (tile)
            __hip_atomic_store(&fd[(size_t)tile * RS_BINS], LB_INC | (pre + tot), __ATOMIC_RELAXED,
                               __HIP_MEMORY_SCOPE_AGENT);
        return block_scan_excl<RS_NT>(hist[d], s_w, nullptr) + pre;
    });
}

// Keys of the points: sources by sort-sources.cpp:87-110 (quantise +
// interleave, x in even bits), with `packed` also the source record
// {x, y, v, 0} in input order, so the gather after the sort reads one 32-byte
// record per source; targets (HILBERT) by hilbert15.  Tiles of RS_TILE points
// are taken in turn by at most KH_GRID workgroups, and each workgroup adds the
// digit totals of all four passes into hist4 (zeroed before) once.  Each
// workgroup finishes the bounding box from k_bbox_partial's np partials and
// per-workgroup flags itself (workgroup 0 writes it and the non-finite flag to
// box_out and bad_out).
template <bool HILBERT>
__global__ __launch_bounds__(256) void k_keys_hist(const double *x, const double *y, const double *v, int n,
                                                    const double4 *part, const uint32_t *pbad, int np,
                                                    double *box_out, uint32_t *bad_out, uint32_t *keys,
                                                    double4 *packed, uint32_t *hist4)
{
    __shared__ uint32_t h[4 * RS_BINS];
    __shared__ double s_box[4];
    for (int i = threadIdx.x; i < 4 * RS_BINS; i += 256)
        h[i] = 0;
    if (threadIdx.x < 64)
        bbox_finish(part, np, s_box);
    __syncthreads();
    if (blockIdx.x == 0 && threadIdx.x < 64) {
        uint32_t b = 0;
        for (int i = threadIdx.x; i < np; i += 64)
            b |= pbad[i];
        b = tc_ballot(b != 0) ? 1u : 0u;
        if (threadIdx.x == 0) {
            for (int q = 0; q < 4; ++q)
                box_out[q] = s_box[q];
            *bad_out = b;
        }
    }
    const double xmin = s_box[0], ymin = s_box[1], scale = s_box[3];
    for (long long base = (long long)blockIdx.x * RS_TILE; base < n; base += (long long)gridDim.x * RS_TILE) {
#pragma unroll 4
        for (int j = 0; j < RS_ITEMS; ++j) {
            const long long i = base + j * 256 + threadIdx.x;
            if (i < n) {
                const double xi = x[i], yi = y[i];
                const uint32_t qx = quantise(xi, xmin, scale), qy = quantise(yi, ymin, scale);
                const uint32_t k = HILBERT ? hilbert15(qx, qy) : (tc_spread15(qx) | (tc_spread15(qy) << 1));
                keys[i] = k;
                if (!HILBERT && packed)
                    packed[i] = make_double4(xi, yi, v[i], 0.0);
                atomicAdd(&h[k & 255u], 1u);
                atomicAdd(&h[RS_BINS + ((k >> 8) & 255u)], 1u);
                atomicAdd(&h[2 * RS_BINS + ((k >> 16) & 255u)], 1u);
                atomicAdd(&h[3 * RS_BINS + (k >> 24)], 1u);
            }
        }
    }
    __syncthreads();
    for (int i = threadIdx.x; i < 4 * RS_BINS; i += 256)
        if (h[i])
            atomicAdd(&hist4[i], h[i]);
}

// ------------------------------------------------------------------ small sorts
// Up to SS_N points in one workgroup: bounding box (k_bbox_partial +
// bbox_finish semantics), keys (Morton or Hilbert as k_keys_hist), the same
// stable 4-pass LSD radix sort as above with the keys in registers and each
// pass's scatter in LDS, and for sources the gather of the sorted records.
// One launch instead of seven, which is what a small solve is made of.
constexpr int SS_NT = 1024, SS_ITEMS = 16, SS_N = SS_NT * SS_ITEMS;
static_assert(SS_N == TC_SORT_SMALL_MAX, "one-workgroup sort size");

// The one-workgroup sort of one point set (the body of k_sort_small and of
// k_sort_small_batch's workgroups).
template <bool HILBERT>
TC_DEV void sort_small(const double *__restrict__ x, const double *__restrict__ y, const double *__restrict__ v,
                       int n, double *__restrict__ box_out, uint32_t *__restrict__ bad_out,
                       uint32_t *__restrict__ keys_out, uint32_t *__restrict__ perm_out,
                       double4 *__restrict__ src_out)
{
    constexpr int NW = SS_NT / 64, WITEMS = 64 * SS_ITEMS;
    __shared__ uint32_t s_k[SS_N];
    __shared__ uint32_t s_v[SS_N];
    __shared__ uint32_t s_wrun[NW][RS_BINS];
    __shared__ uint32_t s_w[NW + 1];
    __shared__ double4 s_red[NW];
    __shared__ double s_box[4];
    __shared__ uint32_t s_bad[NW];
    const int tid = threadIdx.x, lane = tc_lane(), wid = tid >> 6;

    // this thread's points, item j of lane l of wave w being position
    // w * WITEMS + 64 j + l (all loads in flight at once)
    double xr[SS_ITEMS], yr[SS_ITEMS];
#pragma unroll
    for (int j = 0; j < SS_ITEMS; ++j) {
        const int i = wid * WITEMS + j * 64 + lane;
        xr[j] = i < n ? x[i] : 0.0;
        yr[j] = i < n ? y[i] : 0.0;
    }
    // bounding box: max accumulators start at 0 as in the reference
    double xmi = 1e13, ymi = 1e13, xma = 0, yma = 0;
    bool nonfinite = false;
#pragma unroll
    for (int j = 0; j < SS_ITEMS; ++j) {
        const int i = wid * WITEMS + j * 64 + lane;
        if (i < n) {
            const double a = xr[j], b = yr[j];
            nonfinite |= !isfinite(a) || !isfinite(b);
            xmi = a < xmi ? a : xmi;
            xma = xma < a ? a : xma;
            ymi = b < ymi ? b : ymi;
            yma = yma < b ? b : yma;
        }
    }
#pragma unroll
    for (int off = 32; off >= 1; off >>= 1) {
        double o;
        o = __shfl_xor(xmi, off, 64); xmi = o < xmi ? o : xmi;
        o = __shfl_xor(ymi, off, 64); ymi = o < ymi ? o : ymi;
        o = __shfl_xor(xma, off, 64); xma = xma < o ? o : xma;
        o = __shfl_xor(yma, off, 64); yma = yma < o ? o : yma;
    }
    const bool wbad = tc_ballot(nonfinite) != 0;
    if (lane == 0) {
        s_red[wid] = make_double4(xmi, xma, ymi, yma);
        s_bad[wid] = wbad ? 1u : 0u;
    }
    __syncthreads();
    if (tid == 0) {
        double4 r = s_red[0];
        uint32_t bad = s_bad[0];
        for (int w = 1; w < NW; ++w) {
            r.x = s_red[w].x < r.x ? s_red[w].x : r.x;
            r.y = r.y < s_red[w].y ? s_red[w].y : r.y;
            r.z = s_red[w].z < r.z ? s_red[w].z : r.z;
            r.w = r.w < s_red[w].w ? s_red[w].w : r.w;
            bad |= s_bad[w];
        }
        const double ext0 = r.y - r.x, ext1 = r.w - r.z;
        const double eps = 10000 * DBL_EPSILON;
        double ext = (ext0 < ext1 ? ext1 : ext0) * (1 + 2 * eps);
        if (!(ext > 0))
            ext = 1.0;
        s_box[0] = r.x - eps * ext;
        s_box[1] = r.z - eps * ext;
        s_box[2] = ext;
        s_box[3] = (1.0 / ext) * (double)(1 << TC_LMAX);
        for (int q = 0; q < 4; ++q)
            box_out[q] = s_box[q];
        *bad_out = bad;
    }
    __syncthreads();

    const double xmin = s_box[0], ymin = s_box[1], scale = s_box[3];
    uint32_t k[SS_ITEMS], vi[SS_ITEMS], dr[SS_ITEMS];
#pragma unroll
    for (int j = 0; j < SS_ITEMS; ++j) {
        const int i = wid * WITEMS + j * 64 + lane;
        const uint32_t qx = quantise(xr[j], xmin, scale), qy = quantise(yr[j], ymin, scale);
        k[j] = i < n ? (HILBERT ? hilbert15(qx, qy) : (tc_spread15(qx) | (tc_spread15(qy) << 1))) : 0u;
        vi[j] = (uint32_t)i;
    }

#pragma unroll 1
    for (int shift = 0;; shift += 8) {
        for (int d = lane; d < RS_BINS; d += 64)
            s_wrun[wid][d] = 0;
        tc_wave_sync();
        // wave-local stable ranks, as in rs_tile
#pragma unroll
        for (int j = 0; j < SS_ITEMS; ++j) {
            const bool ok = wid * WITEMS + j * 64 + lane < n;
            const uint32_t d = (k[j] >> shift) & (RS_BINS - 1);
            const unsigned long long peers = wave_digit_peers(d, ok);
            const uint32_t lower = __popcll(peers & tc_lanemask_lt());
            const uint32_t run = ok ? s_wrun[wid][d] : 0u;
            tc_wave_sync();
            if (ok && lower == 0)
                s_wrun[wid][d] = run + (uint32_t)__popcll(peers);
            tc_wave_sync();
            dr[j] = ((run + lower) << 8) | d;
        }
        __syncthreads();
        {
            // digit tid: start = exclusive scan of the digit totals; wave w's
            // keys of that digit follow waves 0..w-1
            uint32_t c[NW], tot = 0;
#pragma unroll
            for (int w = 0; w < NW; ++w) {
                c[w] = tid < RS_BINS ? s_wrun[w][tid] : 0u;
                tot += c[w];
            }
            uint32_t off = block_scan_excl<SS_NT>(tot, s_w, nullptr);
            if (tid < RS_BINS) {
#pragma unroll
                for (int w = 0; w < NW; ++w) {
                    s_wrun[w][tid] = off;
                    off += c[w];
                }
            }
        }
        __syncthreads();
#pragma unroll
        for (int j = 0; j < SS_ITEMS; ++j) {
            if (wid * WITEMS + j * 64 + lane < n) {
                const uint32_t pos = s_wrun[wid][dr[j] & 255u] + (dr[j] >> 8);
                s_k[pos] = k[j];
                s_v[pos] = vi[j];
            }
        }
        __syncthreads();
        if (shift == 24)
            break;
#pragma unroll
        for (int j = 0; j < SS_ITEMS; ++j) {
            const int p = wid * WITEMS + j * 64 + lane;
            if (p < n) {
                k[j] = s_k[p];
                vi[j] = s_v[p];
            }
        }
    }

#pragma unroll 8
    for (int p = tid; p < n; p += SS_NT) {
        const uint32_t i = s_v[p];
        keys_out[p] = s_k[p];
        perm_out[p] = i;
        if (src_out)
            src_out[p] = make_double4(x[i], y[i], v[i], 0.0);
    }
}

template <bool HILBERT>
__global__ __launch_bounds__(SS_NT) void k_sort_small(const double *__restrict__ x, const double *__restrict__ y,
                                                      const double *__restrict__ v, int n, double *__restrict__ box_out,
                                                      uint32_t *__restrict__ bad_out, uint32_t *__restrict__ keys_out,
                                                      uint32_t *__restrict__ perm_out, double4 *__restrict__ src_out)
{
    sort_small<HILBERT>(x, y, v, n, box_out, bad_out, keys_out, perm_out, src_out);
}

// Target orders of several requests at once: workgroup b sorts desc[b]'s
// targets exactly as k_sort_small<true> does alone.
__global__ __launch_bounds__(SS_NT) void k_sort_small_batch(const TcSortDesc *__restrict__ desc)
{
    const TcSortDesc &d = desc[blockIdx.x];
    sort_small<true>(d.x, d.y, nullptr, d.n, d.box, d.bad, d.keys, d.perm, nullptr);
}

__global__ __launch_bounds__(256) void k_gather_src(const double4 *packed, const uint32_t *perm, int n,
                                                     double4 *src)
{
    const int i = blockIdx.x * 256 + threadIdx.x;
    if (i >= n)
        return;
    src[i] = packed[perm[i]];
}

// ------------------------------------------------------------------ topology
// Internal cells are exactly the cells of levels 0..14 holding more than
// `cap` sources (their ancestors then hold more too), i.e. the breadth-first
// build_tree recursion of upward.cpp:223-277.  A level-L cell starting at
// sorted position i is internal iff key[i-1] leaves the cell and key[i+cap]
// stays in it, so every internal cell is found from three key reads, and its
// breadth-first rank is a scan over (level, position).
TC_DEV void cell_range(const uint32_t *keys, int n, int cap, int i, int &lo, int &hi)
{
    lo = 1;
    hi = 0;
    if (i >= n)
        return;
    const uint32_t k = keys[i];
    lo = (i == 0) ? 0 : tc_common_level(keys[i - 1], k) + 1;
    hi = (i + cap < n) ? min(TC_LMAX - 1, tc_common_level(k, keys[i + cap])) : -1;
}

__global__ __launch_bounds__(256) void k_int_count(const uint32_t *keys, int n, int cap, uint32_t *cnt_tab,
                                                    int nblk)
{
    __shared__ uint32_t s_cnt[4][TC_LMAX];
    const int lane = tc_lane(), wid = threadIdx.x >> 6;
    const int base = (blockIdx.x * 4 + wid) * TB_CHUNK;
    uint32_t c[TC_LMAX];
#pragma unroll
    for (int L = 0; L < TC_LMAX; ++L)
        c[L] = 0;
    for (int j = 0; j < TB_CHUNK / 64; ++j) {
        int lo, hi;
        cell_range(keys, n, cap, base + j * 64 + lane, lo, hi);
        if (!tc_ballot(hi >= lo))
            continue;
#pragma unroll
        for (int L = 0; L < TC_LMAX; ++L)
            c[L] += __popcll(tc_ballot(L >= lo && L <= hi));
    }
    if (lane == 0)
#pragma unroll
        for (int L = 0; L < TC_LMAX; ++L)
            s_cnt[wid][L] = c[L];
    __syncthreads();
    if (threadIdx.x < TC_LMAX)
        cnt_tab[threadIdx.x * nblk + blockIdx.x] =
            s_cnt[0][threadIdx.x] + s_cnt[1][threadIdx.x] + s_cnt[2][threadIdx.x] + s_cnt[3][threadIdx.x];
}

// Census record read back by the host in one copy: level starts, internal
// count, non-finite flag, then (at int offset INFO_BOX) the bounding box.
constexpr int INFO_BOX = 20;
constexpr size_t INFO_BYTES = sizeof(int) * INFO_BOX + sizeof(double) * 4;

// The census scan in one workgroup (k_scan_single over the level-major
// counts, TC_LMAX rows of nblk) with k_level_starts' record: the thread that
// writes the first prefix of a level also writes that level's start.
__global__ __launch_bounds__(SCAN_NT) void k_census_single(const uint32_t *in, uint32_t *out, int nblk,
                                                            uint32_t *total, const uint32_t *bad, const double *box,
                                                            int *info)
{
    __shared__ uint32_t s_w[SCAN_NT / 64 + 1];
    const int n = TC_LMAX * nblk;
    uint32_t carry = 0;
    for (int base = 0; base < n; base += SCAN_CHUNK) {
        uint32_t v[SCAN_ITEMS];
        uint32_t sum = 0;
        const int i0 = base + threadIdx.x * SCAN_ITEMS;
#pragma unroll
        for (int j = 0; j < SCAN_ITEMS; ++j) {
            v[j] = (i0 + j < n) ? in[i0 + j] : 0u;
            sum += v[j];
        }
        uint32_t tot;
        uint32_t pre = block_scan_excl<SCAN_NT>(sum, s_w, &tot) + carry;
#pragma unroll
        for (int j = 0; j < SCAN_ITEMS; ++j) {
            const int i = i0 + j;
            if (i < n) {
                out[i] = pre;
                if (i % nblk == 0)
                    info[i / nblk] = (int)pre;
            }
            pre += v[j];
        }
        carry += tot;
    }
    if (threadIdx.x == 0) {
        *total = carry;
        info[TC_LMAX] = (int)carry;
        info[TC_LMAX + 1] = (int)*bad;
    }
    if (threadIdx.x >= 32 && threadIdx.x < 36)
        reinterpret_cast<double *>(info + INFO_BOX)[threadIdx.x - 32] = box[threadIdx.x - 32];
}

__global__ void k_level_starts(const uint32_t *offs, int nblk, const uint32_t *total, const uint32_t *bad,
                               const double *box, int *info)
{
    const int L = threadIdx.x;
    if (L < TC_LMAX)
        info[L] = (int)offs[L * nblk];
    if (L == TC_LMAX)
        info[TC_LMAX] = (int)*total;
    if (L == TC_LMAX + 1)
        info[TC_LMAX + 1] = (int)*bad;
    if (L >= 32 && L < 36)
        reinterpret_cast<double *>(info + INFO_BOX)[L - 32] = box[L - 32];
}

__global__ __launch_bounds__(256) void k_int_emit(const uint32_t *keys, int n, int cap, const uint32_t *offs,
                                                   int nblk, int *istart, int *ilevel)
{
    __shared__ uint32_t s_cnt[4][TC_LMAX];
    const int lane = tc_lane(), wid = threadIdx.x >> 6;
    const int base = (blockIdx.x * 4 + wid) * TB_CHUNK;
    uint32_t c[TC_LMAX];
#pragma unroll
    for (int L = 0; L < TC_LMAX; ++L)
        c[L] = 0;
    for (int j = 0; j < TB_CHUNK / 64; ++j) {
        int lo, hi;
        cell_range(keys, n, cap, base + j * 64 + lane, lo, hi);
        if (!tc_ballot(hi >= lo))
            continue;
#pragma unroll
        for (int L = 0; L < TC_LMAX; ++L)
            c[L] += __popcll(tc_ballot(L >= lo && L <= hi));
    }
    if (lane == 0)
#pragma unroll
        for (int L = 0; L < TC_LMAX; ++L)
            s_cnt[wid][L] = c[L];
    __syncthreads();
#pragma unroll
    for (int L = 0; L < TC_LMAX; ++L) {
        uint32_t o = offs[L * nblk + blockIdx.x];
        for (int w = 0; w < wid; ++w)
            o += s_cnt[w][L];
        c[L] = o;
    }
    for (int j = 0; j < TB_CHUNK / 64; ++j) {
        const int i = base + j * 64 + lane;
        int lo, hi;
        cell_range(keys, n, cap, i, lo, hi);
        if (!tc_ballot(hi >= lo))
            continue;
#pragma unroll
        for (int L = 0; L < TC_LMAX; ++L) {
            const bool f = L >= lo && L <= hi;
            const unsigned long long b = tc_ballot(f);
            if (f) {
                const uint32_t r = c[L] + __popcll(b & tc_lanemask_lt());
                istart[r] = i;
                ilevel[r] = L;
            }
            c[L] += __popcll(b);
        }
    }
}

// Clears every node's child field (leaves keep 0; k_int_nodes then sets the
// internal ones) and writes the root's range, level and cell.
__global__ __launch_bounds__(256) void k_topo_init(uint32_t *child, int nn, int n, int2 *se, int *level,
                                                   int2 *cell)
{
    for (int i = blockIdx.x * 256 + threadIdx.x; i < nn; i += gridDim.x * 256)
        child[i] = 0u;
    if (blockIdx.x == 0 && threadIdx.x == 0) {
        se[0] = make_int2(0, n);
        level[0] = 0;
        cell[0] = make_int2(0, 0);
    }
}

__global__ __launch_bounds__(256) void k_int_nodes(const uint32_t *keys, int n, int cap, const int *istart,
                                                    const int *ilevel, const int *info, int T, int *selfid,
                                                    uint32_t *child, int2 *se, int *level, int2 *cell)
{
    const int r = blockIdx.x * 256 + threadIdx.x;
    if (r >= T)
        return;
    const int L = ilevel[r], i = istart[r];
    const uint32_t key = keys[i];
    const int sh = 2 * (TC_LMAX - L);
    const uint32_t prefix = sh >= 32 ? 0u : (key >> sh);
    int self = 0;
    if (L > 0) {
        int lo = info[L - 1], hi = info[L];
        while (lo < hi) {
            const int mid = (lo + hi) >> 1;
            if (istart[mid] <= i)
                lo = mid + 1;
            else
                hi = mid;
        }
        self = 1 + 4 * (lo - 1) + (int)(prefix & 3u);
    }
    selfid[r] = self;
    child[self] = (uint32_t)(1 + 4 * r);

    // end of the cell: first position whose level-L prefix differs
    int e;
    if (L == 0) {
        e = n;
    } else {
        int lo = i + cap + 1, hi = n, step = 1;
        while (true) {
            const int probe = lo + step - 1;
            if (probe >= n) {
                hi = n;
                break;
            }
            if ((keys[probe] >> sh) != prefix) {
                hi = probe;
                break;
            }
            lo = probe + 1;
            step <<= 1;
        }
        while (lo < hi) {
            const int mid = (lo + hi) >> 1;
            if ((keys[mid] >> sh) == prefix)
                lo = mid + 1;
            else
                hi = mid;
        }
        e = lo;
    }
    const int csh = sh - 2;
    int b[5];
    b[0] = i;
    b[4] = e;
    for (int c = 1; c < 4; ++c) {
        const uint32_t target = ((prefix << 2) | (uint32_t)c) << csh;
        int lo = b[c - 1], hi = e;
        while (lo < hi) {
            const int mid = (lo + hi) >> 1;
            if (keys[mid] < target)
                lo = mid + 1;
            else
                hi = mid;
        }
        b[c] = lo;
    }
    const int cx = (int)tc_compact15(prefix), cy = (int)tc_compact15(prefix >> 1);
    for (int c = 0; c < 4; ++c) {
        const int k = 1 + 4 * r + c;
        se[k] = make_int2(b[c], b[c + 1]);
        level[k] = L + 1;
        cell[k] = make_int2(2 * cx + (c & 1), 2 * cy + (c >> 1));
    }
}

// Reduce-scatter form of the pairwise lane tree: at offsets 32, 16, ..., 1 a
// lane keeps half of its values and adds the partner's copies of that half,
// so after log2(NV) halvings (then plain butterfly steps) lane l holds the
// full sum of value (l >> (6 - log2 NV)).  Every sum is formed from the same
// pairs in the same order as tc_wave_tree_sum (a + b == b + a), so results are
// bit-identical to reducing each value separately, with NV - 1 + (6 - log2 NV)
// shuffles instead of 6 NV.
template <int C> TC_DEV void rs_halve(double *v, int off, bool up)
{
#pragma unroll
    for (int j = 0; j < C / 2; ++j) {
        const double mine = up ? v[C / 2 + j] : v[j];
        const double send = up ? v[j] : v[C / 2 + j];
        v[j] = mine + __shfl_xor(send, off, 64);
    }
}

template <int NV> TC_DEV double wave_reduce_scatter(double *v)
{
    static_assert(NV == 4 || NV == 32 || NV == 64, "reduce-scatter width");
    const int lane = tc_lane();
    if (NV == 64) {
        rs_halve<64>(v, 32, lane & 32);
        rs_halve<32>(v, 16, lane & 16);
        rs_halve<16>(v, 8, lane & 8);
        rs_halve<8>(v, 4, lane & 4);
        rs_halve<4>(v, 2, lane & 2);
        rs_halve<2>(v, 1, lane & 1);
    } else if (NV == 32) {
        rs_halve<32>(v, 32, lane & 32);
        rs_halve<16>(v, 16, lane & 16);
        rs_halve<8>(v, 8, lane & 8);
        rs_halve<4>(v, 4, lane & 4);
        rs_halve<2>(v, 2, lane & 2);
        v[0] = v[0] + __shfl_xor(v[0], 1, 64);
    } else {
        rs_halve<4>(v, 32, lane & 32);
        rs_halve<2>(v, 16, lane & 16);
#pragma unroll
        for (int off = 8; off >= 1; off >>= 1)
            v[0] = v[0] + __shfl_xor(v[0], off, 64);
    }
    return v[0];
}

// ------------------------------------------------------------------ upward
// One wavefront per leaf: build_leaf (upward.cpp:61-93) = upward_setup
// (upward-common.ispc:13-57) + upward_p2e (upward-kernels.m4:52-92), in the
// canonical 64-lane strided order with the pairwise lane tree.
template <int P>
__global__ __launch_bounds__(256) void k_leaf_upward(const uint32_t *child, const int2 *se, const int *level,
                                                      const int2 *cell, const double4 *src, const double *box,
                                                      int nnodes, TcNode *nodes, double *exps, double4 *nodew,
                                                      float4 *src32)
{
    const int id = tc_uniform((int)(blockIdx.x * 4 + (threadIdx.x >> 6)));
    if (id >= nnodes || child[id] != 0)
        return;
    const int lane = tc_lane();
    const int2 r = se[id];
    const int s = r.x, e = r.y, l = level[id];
    const int2 cc = cell[id];
    const double xmin = box[0], ymin = box[1], ext = box[2];
    const double h = ext / (double)(1 << l);
    const double x0 = xmin + h * (double)cc.x, y0 = ymin + h * (double)cc.y;

    double pm = 0, pw = 0, px = 0, py = 0;
    for (int i = s + lane; i < e; i += 64) {
        const double4 q = src[i];
        const double a = fabs(q.z);
        pm += q.z;
        pw += a;
        px += q.x * a;
        py += q.y * a;
    }
    // lane 0: mass, 16: w, 32: wx, 48: wy (bit-identical to four tree sums)
    double g4[4] = {pm, pw, px, py};
    const double gsum = wave_reduce_scatter<4>(g4);
    const double m = __shfl(gsum, 0, 64), w = __shfl(gsum, 16, 64);
    const double wx = __shfl(gsum, 32, 64), wy = __shfl(gsum, 48, 64);
    const double xc = w != 0 ? wx / w : (x0 + 0.5 * h);
    const double yc = w != 0 ? wy / w : (y0 + 0.5 * h);

    double r2 = 0;
    for (int i = s + lane; i < e; i += 64) {
        const double4 q = src[i];
        const double xr = q.x - xc, yr = q.y - yc;
        const double d = xr * xr + yr * yr;
        r2 = r2 < d ? d : r2;
    }
    r2 = tc_wave_max(r2);
    const double rad = sqrt(r2);

    double re[P], im[P];
#pragma unroll
    for (int k = 0; k < P; ++k)
        re[k] = im[k] = 0;
    for (int i = s + lane; i < e; i += 64) {
        const double4 q = src[i];
        const double r0 = q.x - xc, i0 = q.y - yc, v = q.z;
        double rtmp = r0 * v, itmp = i0 * v;
        re[0] -= rtmp;
        im[0] -= itmp;
        double rprod = r0, iprod = i0;
#pragma unroll
        for (int k = 1; k < P; ++k) {
            rtmp = rprod * r0 - iprod * i0;
            itmp = rprod * i0 + iprod * r0;
            const double term = v * (1.0 / (double)(k + 1));
            rprod = rtmp;
            iprod = itmp;
            rtmp = rprod * term;
            itmp = iprod * term;
            re[k] -= rtmp;
            im[k] -= itmp;
        }
        if (src32)
            src32[i] = make_float4((float)r0, (float)i0, (float)v, 0.f);
    }
    double *ex = exps + (size_t)2 * P * id;
    {
        // [Re a_1..a_p][Im a_1..a_p] reduced together; lane l ends with entry
        // l (64 entries) or l >> 1 (32 entries)
        constexpr int NV = 2 * P <= 32 ? 32 : 64;
        double v[NV];
#pragma unroll
        for (int k = 0; k < NV; ++k)
            v[k] = k < P ? re[k] : (k < 2 * P ? im[k - P] : 0.0);
        const double sum = wave_reduce_scatter<NV>(v);
        const int idx = NV == 64 ? lane : lane >> 1;
        if (idx < 2 * P && (NV == 64 || (lane & 1) == 0))
            ex[idx] = sum;
    }
    if (lane == 0) {
        TcNode nd;
        nd.xcom = xc;
        nd.ycom = yc;
        nd.r = rad;
        nd.mass = m;
        nd.s = s;
        nd.e = e;
        nd.child = 0;
        nd.level = l;
        nd.cx = cc.x;
        nd.cy = cc.y;
        nd.rr = rad * rad;
        nodes[id] = nd;
        nodew[id] = make_double4(w, wx, wy, 0.0);
    }
}

// Parent update of upward.cpp:95-218 and upward_e2e (upward-kernels.m4:95-149)
// for internal node `rank`, by one wavefront.  The children's translations run
// in parallel: lane group g (E2E_G lanes: 16 for p <= 16, else 32) takes child
// g of a round, one lane of it forms the powers z0^-j by the reference's
// recurrence, lane l the coefficient l+1.  Each child's contribution lands in
// LDS and lanes add them in order c = 0..3, so every coefficient is formed by
// the same operations as a child-by-child loop.
template <int P> constexpr int E2E_G = P <= 16 ? 16 : 32;

template <int P> struct E2eLds {
    double rinv[64 / E2E_G<P>][P + 1], iinv[64 / E2E_G<P>][P + 1];
    double rc[64 / E2E_G<P>][P + 1], ic[64 / E2E_G<P>][P + 1];
    double cre[4][P], cim[4][P];
};

template <int P>
TC_DEV void e2e_node(int rank, const int *selfid, const int2 *se, const int *level, const int2 *cell,
                     const double *box, TcNode *nodes, double4 *nodew, double *exps, E2eLds<P> &L)
{
    constexpr int G = E2E_G<P>, NCH = 64 / G;
    const int lane = tc_lane();
    const int grp = lane / G, gl = lane % G;
    const int id = selfid[rank];
    const int cb = 1 + 4 * rank;
    const int l = level[id];
    const int2 cc = cell[id];
    const double xmin = box[0], ymin = box[1], ext = box[2];

    TcNode ch[4];
    double4 cw[4];
    double msum = 0, wsum = 0, wxsum = 0, wysum = 0;
    for (int c = 0; c < 4; ++c) {
        ch[c] = nodes[cb + c];
        cw[c] = nodew[cb + c];
        msum += ch[c].mass;
        wsum += cw[c].x;
        wxsum += cw[c].y;
        wysum += cw[c].z;
    }
    const double h = ext / (double)(1 << l);
    const double xcp = wsum != 0 ? wxsum / wsum : (xmin + h * (double)cc.x) + 0.5 * h;
    const double ycp = wsum != 0 ? wysum / wsum : (ymin + h * (double)cc.y) + 0.5 * h;
    double rr = 0;
    for (int c = 0; c < 4; ++c) {
        if (cw[c].x != 0) {
            const double rx = xcp - ch[c].xcom, ry = ycp - ch[c].ycom;
            const double v = ch[c].r + sqrt(rx * rx + ry * ry);
            rr = rr < v ? v : rr;
        }
    }
    const double rcap = (double)1.4143f * ext / (double)(1 << l);
    const double rad = rcap < rr ? rcap : rr;

    int zcount = 0, nz = 0;
    for (int c = 0; c < 4; ++c) {
        const bool neg = fabs(cw[c].x) < 1.1e-16;
        zcount += neg;
        if (!neg)
            nz = c;
    }
    double dre = 0, dim = 0;
    if (zcount == 3) {
        if (lane < P) {
            dre = exps[(size_t)2 * P * (cb + nz) + lane];
            dim = exps[(size_t)2 * P * (cb + nz) + P + lane];
        }
    } else if (zcount < 3) {
        for (int c0 = 0; c0 < 4; c0 += NCH) {
            // this lane's child (the index differs between lane groups, so it
            // is loaded rather than picked out of ch[], which would go to scratch)
            const int c = c0 + grp;
            const TcNode chc = nodes[cb + c];
            const bool act = nodew[cb + c].x != 0;
            const double x0 = chc.xcom - xcp, y0 = chc.ycom - ycp;
            const bool ident = x0 == 0 && y0 == 0;
            const double *ca = exps + (size_t)2 * P * (cb + c);
            // the child's coefficients, requested now so that the load overlaps
            // the powers of 1/z0 below
            double car = 0, cai = 0;
            if (act && gl < P) {
                car = ca[gl];
                cai = ca[P + gl];
            }
            if (act && !ident && gl == 0) {
                const double r2z0 = x0 * x0 + y0 * y0;
                const double ri1 = x0 / r2z0, ii1 = -y0 / r2z0;
                double ri = ri1, ii = ii1;
                L.rinv[grp][1] = ri;
                L.iinv[grp][1] = ii;
                for (int j = 2; j <= P; ++j) {
                    const double rn = ri * ri1 - ii * ii1;
                    const double in = ri * ii1 + ii * ri1;
                    ri = rn;
                    ii = in;
                    L.rinv[grp][j] = ri;
                    L.iinv[grp][j] = ii;
                }
            }
            tc_wave_sync();
            if (act && !ident && gl < P) {
                const int jj = gl + 1;
                L.rc[grp][jj] = car * L.rinv[grp][jj] - cai * L.iinv[grp][jj];
                L.ic[grp][jj] = car * L.iinv[grp][jj] + cai * L.rinv[grp][jj];
            }
            tc_wave_sync();
            if (gl < P) {
                double cre = 0, cim = 0;
                if (act && ident) {
                    cre = car;
                    cim = cai;
                } else if (act) {
                    const int ll = gl + 1;
                    const double mass = chc.mass;
                    double rt = (ll == 1) ? -mass : mass * (-(1.0 / (double)ll));
                    double it = L.ic[grp][1];
                    rt = rt + L.rc[grp][1];
                    for (int k = 2; k <= ll; ++k) {
                        const double bk = c_binom[(ll - 1) * 64 + (k - 1)];
                        rt = rt + L.rc[grp][k] * bk;
                        it = it + L.ic[grp][k] * bk;
                    }
                    const double ri = L.rinv[grp][ll], ii = L.iinv[grp][ll];
                    const double invz2 = ri * ri + ii * ii;
                    const double invinv = invz2 != 0 ? 1 / invz2 : 0;
                    const double rz = ri * invinv;
                    const double iz = -ii * invinv;
                    cre = rt * rz - it * iz;
                    cim = rt * iz + it * rz;
                }
                L.cre[c][gl] = cre;
                L.cim[c][gl] = cim;
            }
            tc_wave_sync();
        }
        if (lane < P) {
            for (int c = 0; c < 4; ++c) {
                if (cw[c].x == 0)
                    continue;
                dre += L.cre[c][lane];
                dim += L.cim[c][lane];
            }
        }
        tc_wave_sync();
    }
    if (lane < P) {
        exps[(size_t)2 * P * id + lane] = dre;
        exps[(size_t)2 * P * id + P + lane] = dim;
    }
    if (lane == 0) {
        const int2 r = se[id];
        TcNode nd;
        nd.xcom = xcp;
        nd.ycom = ycp;
        nd.r = rad;
        nd.mass = msum;
        nd.s = r.x;
        nd.e = r.y;
        nd.child = cb;
        nd.level = l;
        nd.cx = cc.x;
        nd.cy = cc.y;
        nd.rr = rad * rad;
        nodes[id] = nd;
        nodew[id] = make_double4(wsum, wxsum, wysum, 0.0);
    }
}

// One wavefront per internal node of one level (ranks r0..r1-1).
template <int P>
__global__ __launch_bounds__(64) void k_e2e_level(const int *selfid, int r0, int r1, const int2 *se,
                                                   const int *level, const int2 *cell, const double *box,
                                                   TcNode *nodes, double4 *nodew, double *exps)
{
    __shared__ E2eLds<P> L;
    const int rank = r0 + blockIdx.x;
    if (rank >= r1)
        return;
    e2e_node<P>(rank, selfid, se, level, cell, box, nodes, nodew, exps, L);
}

// The small top levels in one launch: levels lt-1 .. 0 (ranks starts[L] ..
// starts[L+1]-1, at most 4 nodes per wave each), E2E_TOP_WAVES waves striding
// over a level's nodes, a workgroup barrier between levels (children before
// parents).
constexpr int E2E_TOP_WAVES = 16;
template <int P>
__global__ __launch_bounds__(64 * E2E_TOP_WAVES) void k_e2e_top(const int *selfid, const int *starts, int lt,
                                                                 const int2 *se, const int *level,
                                                                 const int2 *cell, const double *box,
                                                                 TcNode *nodes, double4 *nodew, double *exps)
{
    __shared__ E2eLds<P> L[E2E_TOP_WAVES];
    const int wid = threadIdx.x >> 6;
    for (int lv = lt - 1; lv >= 0; --lv) {
        const int r0 = starts[lv], r1 = starts[lv + 1];
        for (int rank = r0 + wid; rank < r1; rank += E2E_TOP_WAVES)
            e2e_node<P>(rank, selfid, se, level, cell, box, nodes, nodew, exps, L[wid]);
        // the next level reads what this one wrote; the readers are waves of
        // this workgroup, which share one CU and its vector L1, so the
        // barrier's workgroup-scope release / acquire orders the global
        // writes (device-scope fences would write back and invalidate caches
        // at every level)
        __syncthreads();
    }
}

// FP32 mirror of the expansions: b_j = a_j / s^j with s = max(r, 1e-30), so
// that the FP32 E2P evaluates sum b_j (s/z)^j with |s/z| < theta (force trees:
// j b_j, the derivative series' coefficients).  Since |a_j| <= sum|v| r^j / j
// and s >= r, scaling the coefficient itself by 1/s j times keeps every
// intermediate bounded (no overflow of s^-j for tiny radii).  Entry p of a
// node's p + 1 records holds the E2P constants (s, Q ln2/2) for potential
// trees (the FP32 E2P works with log2) and (s, Q) for force trees.
template <int P>
__global__ __launch_bounds__(256) void k_exps32(const TcNode *nodes, const double *exps, int nnodes, int deriv,
                                                float2 *ex32)
{
    // G lanes per node (p + 1 records each), 64 / G nodes per wave
    constexpr int G = P + 1 <= 16 ? 16 : (P + 1 <= 32 ? 32 : 64);
    const int lane = tc_lane();
    const int id = (blockIdx.x * 4 + (threadIdx.x >> 6)) * (64 / G) + lane / G;
    const int k = lane % G;
    if (id >= nnodes || k > P)
        return;
    const double r = nodes[id].r;
    const double sc = r > 1e-30 ? r : 1e-30;
    if (k == P) {
        const double q = nodes[id].mass;
        ex32[(size_t)(P + 1) * id + P] = make_float2((float)sc, (float)(deriv ? q : q * 0.34657359027997264));
        return;
    }
    const double *ex = exps + (size_t)2 * P * id;
    // force trees store j b_j, the coefficients of the derivative series
    const double kj = deriv ? (double)(k + 1) : 1.0;
    const double inv = 1.0 / sc;
    double a = kj * ex[k], b = kj * ex[P + k];
    for (int j = 0; j <= k; ++j) {
        a *= inv;
        b *= inv;
    }
    ex32[(size_t)(P + 1) * id + k] = make_float2((float)a, (float)b);
}

int tiles_of(int n, int t) { return (n + t - 1) / t; }

}  // namespace

// ================================================================== host side

void tc_fill_u32(uint32_t *p, uint32_t v, size_t n, hipStream_t st)
{
    if (!n)
        return;
    const int grid = (int)std::min<size_t>((n + 255) / 256, 4096);
    k_fill_u32<<<grid, 256, 0, st>>>(p, v, n);
    TC_HIP(hipGetLastError());
}

void tc_scan_exclusive_u32(const uint32_t *in, uint32_t *out, int n, uint32_t *total_d, TcScratch &ws,
                           hipStream_t st)
{
    if (n <= 4 * SCAN_CHUNK) {
        k_scan_single<<<1, SCAN_NT, 0, st>>>(in, out, n, total_d);
        TC_HIP(hipGetLastError());
        return;
    }
    const int nb = tiles_of(n, SCAN_CHUNK);
    ws.scan_tmp.reserve(sizeof(uint32_t) * (size_t)(2 * nb + 4));
    uint32_t *part = ws.scan_tmp.as<uint32_t>();
    uint32_t *partx = part + nb;
    k_scan_partial<<<nb, SCAN_NT, 0, st>>>(in, n, part);
    k_scan_single<<<1, SCAN_NT, 0, st>>>(part, partx, nb, total_d);
    k_scan_final<<<nb, SCAN_NT, 0, st>>>(in, out, n, partx);
    TC_HIP(hipGetLastError());
}

void tc_upload_binomials(hipStream_t st)
{
    std::vector<double> tab(64 * 64, 0.0);
    for (int n = 0; n < 64; ++n)
        for (int k = 0; k <= n; ++k) {
            unsigned __int128 r = 1;
            const int kk = k < n - k ? k : n - k;
            for (int i = 1; i <= kk; ++i)
                r = r * (unsigned __int128)(n - kk + i) / (unsigned __int128)i;
            tab[n * 64 + k] = (double)r;
        }
    TC_HIP(hipMemcpyToSymbolAsync(HIP_SYMBOL(c_binom), tab.data(), sizeof(double) * tab.size(), 0,
                                  hipMemcpyHostToDevice, st));
    TC_HIP(hipStreamSynchronize(st));
}

static const uint32_t *tc_bbox_bad_flag(TcScratch &ws, int n)
{
    const int np = std::max(1, std::min(1024, tiles_of(n, 256 * 8)));
    return reinterpret_cast<const uint32_t *>(ws.bbox_part.as<double4>() + np);
}

// Bounding-box partials of a sort with one-launch passes: np partials, the
// non-finite flag where tc_bbox_bad_flag finds it, per-workgroup flags after it.
struct LbBox {
    double4 *part;
    uint32_t *bad, *pbad;
    int np;
};
static LbBox lb_box(TcScratch &ws, int n)
{
    const int np = std::max(1, std::min(1024, tiles_of(n, 256 * 8)));
    ws.bbox_part.reserve(sizeof(double4) * (size_t)np + 256 + sizeof(uint32_t) * (size_t)np);
    LbBox b;
    b.part = ws.bbox_part.as<double4>();
    b.bad = reinterpret_cast<uint32_t *>(b.part + np);
    b.pbad = b.bad + 64;
    b.np = np;
    return b;
}

void tc_sort_points(const double *x, const double *y, const double *v, int n, double *box_d, uint32_t *keys_out,
                    uint32_t *perm_out, double4 *src_out, TcScratch &ws, hipStream_t st, bool hilbert,
                    const hipEvent_t *ev)
{
    if (n <= SS_N) {
        ws.bbox_part.reserve(sizeof(double4) + 256);
        uint32_t *bad = const_cast<uint32_t *>(tc_bbox_bad_flag(ws, n));
        if (hilbert)
            k_sort_small<true><<<1, SS_NT, 0, st>>>(x, y, v, n, box_d, bad, keys_out, perm_out, src_out);
        else
            k_sort_small<false><<<1, SS_NT, 0, st>>>(x, y, v, n, box_d, bad, keys_out, perm_out, src_out);
        TC_HIP(hipGetLastError());
        if (ev) {
            // one kernel does keys and sort: both boundaries fall after it
            TC_HIP(hipEventRecord(ev[0], st));
            TC_HIP(hipEventRecord(ev[1], st));
        }
        return;
    }
    // the radix passes take one launch each; their counters are zeroed by the
    // bounding-box kernel, whose partials the key kernel finishes
    const size_t words = lb_words(tiles_of(n, RS_TILE));
    if (words > (size_t)INT_MAX)
        throw TcError(-3, "tc_sort_points: too many points");
    ws.hist.reserve(sizeof(uint32_t) * words);
    uint32_t *lb = ws.hist.as<uint32_t>();
    const LbBox b = lb_box(ws, n);
    k_bbox_partial<<<b.np, 256, 0, st>>>(x, y, n, b.part, nullptr, b.pbad, lb, (int)words);
    TC_HIP(hipGetLastError());
    double4 *packed = nullptr;
    if (src_out) {
        ws.packed.reserve(sizeof(double4) * (size_t)n);
        packed = ws.packed.as<double4>();
    }
    tc_morton_sort(x, y, v, n, box_d, keys_out, perm_out, packed, ws, st, hilbert, ev, lb);
    if (src_out) {
        k_gather_src<<<tiles_of(n, 256), 256, 0, st>>>(packed, perm_out, n, src_out);
        TC_HIP(hipGetLastError());
    }
    if (ev)
        TC_HIP(hipEventRecord(ev[1], st));
}

void tc_morton_sort(const double *x, const double *y, const double *v, int n, const double *box_d,
                    uint32_t *keys_out, uint32_t *perm_out, double4 *packed, TcScratch &ws, hipStream_t st,
                    bool hilbert, const hipEvent_t *ev, uint32_t *lb)
{
    ws.k2.reserve(sizeof(uint32_t) * (size_t)n);
    ws.v2.reserve(sizeof(uint32_t) * (size_t)n);
    const int ntiles = tiles_of(n, RS_TILE);
    // ping-pong out -> scratch -> out -> scratch -> out (four 8-bit passes over 30-bit keys);
    // the first pass generates the index payload itself
    uint32_t *kin[4] = {keys_out, ws.k2.as<uint32_t>(), keys_out, ws.k2.as<uint32_t>()};
    uint32_t *vin[4] = {nullptr, ws.v2.as<uint32_t>(), perm_out, ws.v2.as<uint32_t>()};
    uint32_t *kout[4] = {ws.k2.as<uint32_t>(), keys_out, ws.k2.as<uint32_t>(), keys_out};
    uint32_t *vout[4] = {ws.v2.as<uint32_t>(), perm_out, ws.v2.as<uint32_t>(), perm_out};
    // the box and keys with all four passes' digit totals, then one launch per
    // pass (the bounding-box partials and the zeroed region lb were made by the caller)
    uint32_t *hist4 = lb, *tickets = lb + LB_HIST;
    unsigned long long *flags = reinterpret_cast<unsigned long long *>(tickets + LB_TICKETS);
    const LbBox b = lb_box(ws, n);
    uint32_t *err = b.bad;
    const int grid = std::min(ntiles, KH_GRID);
    if (hilbert)
        k_keys_hist<true><<<grid, 256, 0, st>>>(x, y, nullptr, n, b.part, b.pbad, b.np, const_cast<double *>(box_d),
                                                b.bad, keys_out, nullptr, hist4);
    else
        k_keys_hist<false><<<grid, 256, 0, st>>>(x, y, v, n, b.part, b.pbad, b.np, const_cast<double *>(box_d),
                                                 b.bad, keys_out, packed, hist4);
    TC_HIP(hipGetLastError());
    if (ev)
        TC_HIP(hipEventRecord(ev[0], st));
    for (int pass = 0; pass < 4; ++pass) {
        k_rs_pass_lb<<<ntiles, RS_NT, 0, st>>>(kin[pass], vin[pass], kout[pass], vout[pass], n, 8 * pass,
                                                hist4 + RS_BINS * pass, flags + (size_t)RS_BINS * ntiles * pass,
                                                tickets + pass, err);
        TC_HIP(hipGetLastError());
    }
}

template <int P>
static void upward_launch(TcTree &t, const uint32_t *child, const int2 *se, const int *lev, const int2 *cell,
                          const int *selfid, const int *starts_d, bool fp32, bool deriv, hipStream_t st)
{
    const double *box = t.box_d.as<double>();
    k_leaf_upward<P><<<tiles_of(t.nnodes, 4), 256, 0, st>>>(child, se, lev, cell, t.src.as<double4>(), box,
                                                             t.nnodes, t.nodes.as<TcNode>(), t.exps.as<double>(),
                                                             t.nodew.as<double4>(),
                                                             fp32 ? t.src32.as<float4>() : nullptr);
    TC_HIP(hipGetLastError());
    // levels whose node count is small run together in one launch
    int lt = 0;
    while (lt < TC_LMAX && t.level_start[lt + 1] - t.level_start[lt] <= 4 * E2E_TOP_WAVES)
        ++lt;
    for (int L = TC_LMAX - 1; L >= lt; --L) {
        const int r0 = t.level_start[L], r1 = t.level_start[L + 1];
        if (r1 > r0) {
            k_e2e_level<P><<<r1 - r0, 64, 0, st>>>(selfid, r0, r1, se, lev, cell, box, t.nodes.as<TcNode>(),
                                                    t.nodew.as<double4>(), t.exps.as<double>());
            TC_HIP(hipGetLastError());
        }
    }
    if (lt > 0 && t.level_start[lt] > 0) {
        k_e2e_top<P><<<1, 64 * E2E_TOP_WAVES, 0, st>>>(selfid, starts_d, lt, se, lev, cell, box,
                                                       t.nodes.as<TcNode>(), t.nodew.as<double4>(),
                                                       t.exps.as<double>());
        TC_HIP(hipGetLastError());
    }
    if (fp32) {
        constexpr int per_wave = P + 1 <= 16 ? 4 : (P + 1 <= 32 ? 2 : 1);
        k_exps32<P><<<tiles_of(t.nnodes, 4 * per_wave), 256, 0, st>>>(t.nodes.as<TcNode>(), t.exps.as<double>(),
                                                                        t.nnodes, deriv ? 1 : 0, t.exps32.as<float2>());
        TC_HIP(hipGetLastError());
    }
}

void tc_build_tree(TcTree &t, const double *x, const double *y, const double *v, int n, int leafcap, int order,
                   bool fp32, bool deriv32, TcScratch &ws, hipStream_t st, const hipEvent_t *ev,
                   const std::function<void()> *at_census)
{
    if (n <= 0)
        throw TcError(-3, "tc_build_tree: no sources");
    if (leafcap < 1 || leafcap > TC_MAX_LEAF)
        throw TcError(-3, "tc_build_tree: leaf capacity must be in [1, 2^30]");
    if (n >= (1 << 30))
        throw TcError(-3, "tc_build_tree: at most 2^30 - 1 sources per tree");
    if (!tc_order_supported(order))
        throw TcError(-3, "tc_build_tree: unsupported expansion order " + std::to_string(order));
    t.n = n;
    t.order = order;
    t.leafcap = leafcap;
    t.has_fp32 = fp32;
    t.deriv32 = fp32 && deriv32;
    t.box_d.reserve(sizeof(double) * 4);
    t.keys.reserve(sizeof(uint32_t) * (size_t)n);
    t.perm.reserve(sizeof(uint32_t) * (size_t)n);
    t.src.reserve(sizeof(double4) * (size_t)n);
    if (fp32)
        t.src32.reserve(sizeof(float4) * (size_t)n);

    double *box = t.box_d.as<double>();
    uint32_t *keys = t.keys.as<uint32_t>(), *perm = t.perm.as<uint32_t>();
    tc_sort_points(x, y, v, n, box, keys, perm, t.src.as<double4>(), ws, st, false, ev);

    // census of internal cells
    const int nblk = tiles_of(n, 4 * TB_CHUNK);
    ws.part.reserve(sizeof(uint32_t) * ((size_t)2 * TC_LMAX * nblk + 8));
    uint32_t *cnt = ws.part.as<uint32_t>();
    uint32_t *offs = cnt + (size_t)TC_LMAX * nblk;
    uint32_t *total = offs + (size_t)TC_LMAX * nblk;
    ws.info.reserve(sizeof(int) * 32);
    int *info = ws.info.as<int>();
    k_int_count<<<nblk, 256, 0, st>>>(keys, n, leafcap, cnt, nblk);
    if (TC_LMAX * nblk <= 4 * SCAN_CHUNK) {
        k_census_single<<<1, SCAN_NT, 0, st>>>(cnt, offs, nblk, total, tc_bbox_bad_flag(ws, n), box, info);
    } else {
        tc_scan_exclusive_u32(cnt, offs, TC_LMAX * nblk, total, ws, st);
        k_level_starts<<<1, 64, 0, st>>>(offs, nblk, total, tc_bbox_bad_flag(ws, n), box, info);
    }
    TC_HIP(hipGetLastError());
    static_assert(INFO_BYTES <= sizeof(int) * 32 && INFO_BYTES <= TcScratch::PINNED_BYTES, "census record size");
    char *hp = static_cast<char *>(ws.host_pinned());
    TC_HIP(hipMemcpyAsync(hp, info, INFO_BYTES, hipMemcpyDeviceToHost, st));
    if (at_census)
        (*at_census)();  // work queued elsewhere fills the GPU while the host waits
    TC_HIP(hipStreamSynchronize(st));
    int hinfo[TC_LMAX + 2];
    std::memcpy(hinfo, hp, sizeof(hinfo));
    std::memcpy(t.box, hp + sizeof(int) * INFO_BOX, sizeof(t.box));
    if (hinfo[TC_LMAX + 1] & 2)
        throw TcError(-9, "treecode: the source sort did not complete (a radix pass timed out)");
    if (hinfo[TC_LMAX + 1])
        throw TcError(-4, "treecode: non-finite source coordinate");
    const int T = hinfo[TC_LMAX];
    for (int L = 0; L < TC_LMAX; ++L)
        t.level_start[L] = hinfo[L];
    t.level_start[TC_LMAX] = T;
    t.level_start[TC_LMAX + 1] = T;
    t.ninternal = T;
    t.nnodes = 1 + 4 * T;
    const int nn = t.nnodes;

    t.nodes.reserve(sizeof(TcNode) * (size_t)nn);
    t.exps.reserve(sizeof(double) * 2 * (size_t)order * nn);
    t.nodew.reserve(sizeof(double4) * (size_t)nn);
    t.selfid.reserve(sizeof(int) * (size_t)(T + 1));
    if (fp32)
        t.exps32.reserve(sizeof(float2) * (size_t)(order + 1) * nn);
    ws.istart.reserve(sizeof(int) * (size_t)(T + 1));
    ws.ilevel.reserve(sizeof(int) * (size_t)(T + 1));
    // per-node topology scratch: child | se | level | cell
    ws.ints.reserve(sizeof(int) * (size_t)nn * 6 + 64);
    uint32_t *child = ws.ints.as<uint32_t>();
    int2 *se = reinterpret_cast<int2 *>(child + ((nn + 1) & ~1));
    int *lev = reinterpret_cast<int *>(se + nn);
    int2 *cell = reinterpret_cast<int2 *>(lev + ((nn + 1) & ~1));

    k_topo_init<<<std::min(tiles_of(nn, 256), 4096), 256, 0, st>>>(child, nn, n, se, lev, cell);
    if (T > 0) {
        k_int_emit<<<nblk, 256, 0, st>>>(keys, n, leafcap, offs, nblk, ws.istart.as<int>(), ws.ilevel.as<int>());
        k_int_nodes<<<tiles_of(T, 256), 256, 0, st>>>(keys, n, leafcap, ws.istart.as<int>(), ws.ilevel.as<int>(),
                                                      info, T, t.selfid.as<int>(), child, se, lev, cell);
    }
    TC_HIP(hipGetLastError());
    if (ev)
        TC_HIP(hipEventRecord(ev[2], st));

    const int *sid = t.selfid.as<int>();
    switch (order) {
#define TC_CASE(P) \
    case P: upward_launch<P>(t, child, se, lev, cell, sid, info, fp32, deriv32, st); break;
        TC_ORDERS(TC_CASE)
#undef TC_CASE
    }
    if (ev)
        TC_HIP(hipEventRecord(ev[3], st));
}

const uint32_t *tc_targets_bad_flag(TcScratch &ws, int m) { return tc_bbox_bad_flag(ws, m); }

TcSortDesc tc_sort_targets_desc(const double *xt, const double *yt, int m, TcScratch &ws)
{
    if (m < 1 || m > SS_N)
        throw TcError(-3, "tc_sort_targets_desc: 1 to 16384 targets");
    // the buffers and places tc_sort_targets uses for m targets
    ws.tkeys.reserve(sizeof(uint32_t) * (size_t)m + 64);
    ws.tperm.reserve(sizeof(uint32_t) * (size_t)m + 64);
    ws.tsorted.reserve(sizeof(double) * 2 * (size_t)m + 64);
    ws.info.reserve(sizeof(int) * 32 + sizeof(double) * 4);
    ws.bbox_part.reserve(sizeof(double4) + 256);
    TcSortDesc d;
    d.x = xt;
    d.y = yt;
    d.box = reinterpret_cast<double *>(ws.info.as<char>() + 128);
    d.bad = const_cast<uint32_t *>(tc_bbox_bad_flag(ws, m));
    d.keys = ws.tkeys.as<uint32_t>();
    d.perm = ws.tperm.as<uint32_t>();
    d.n = m;
    d.pad = 0;
    return d;
}

void tc_sort_targets_batch(const TcSortDesc *desc_dev, int k, hipStream_t st)
{
    if (k <= 0)
        return;
    k_sort_small_batch<<<k, SS_NT, 0, st>>>(desc_dev);
    TC_HIP(hipGetLastError());
}

void tc_sort_targets(const double *xt, const double *yt, int m, TcScratch &ws, hipStream_t st)
{
    ws.tkeys.reserve(sizeof(uint32_t) * (size_t)m + 64);
    ws.tperm.reserve(sizeof(uint32_t) * (size_t)m + 64);
    ws.tsorted.reserve(sizeof(double) * 2 * (size_t)m + 64);
    // the target box lives after the keys' tail
    DevBuf &boxbuf = ws.info;
    ws.info.reserve(sizeof(int) * 32 + sizeof(double) * 4);
    double *box = reinterpret_cast<double *>(boxbuf.as<char>() + 128);
    tc_sort_points(xt, yt, nullptr, m, box, ws.tkeys.as<uint32_t>(), ws.tperm.as<uint32_t>(), nullptr, ws, st, true);
}
