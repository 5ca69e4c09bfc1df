// tc_eval.hip -- theta-criterion traversals on gfx950.
//
// Potential (drivers/treecode-potential.cpp:25-86): one wavefront per 64
// Hilbert-sorted targets.  The wave walks the tree once with a stack of
// (node, lane mask) entries; every lane takes exactly the accept/open
// decisions of its own per-target DFS, so each target sees the reference's
// interaction list.  Node data and expansion coefficients are wave-uniform
// (scalar loads).  At a leaf, sources are either staged through LDS 64 at a
// time and broadcast to the lanes (dense), or, when few lanes need the leaf,
// spread over the lanes with one reduction per active target (sparse, FP32).
//
// Force (drivers/treecode-force.cpp:94-199): one wavefront per 8x8 brick of a
// 32x32 block, lane = target, bricks dealt from a device counter to a
// persistent grid.  Decisions are brick-wide, taken for four children at once
// when a node is opened; E2L is wave-cooperative (lane l of a lane group owns
// local coefficient l+1).
//
// precision 0 = plain FP64; 1 = FP32 bulk: P2P and E2P run in FP32 on
// leaf-relative offsets with per-tile flushes into FP64 accumulators, while
// decisions, E2L and L2P stay FP64 (same interaction lists as FP64).
#include "tc_common.h"

#include <algorithm>
#include <atomic>
#include <cfloat>
#include <vector>

namespace {

constexpr double EPS_REG = 10.0 * DBL_EPSILON;  // potential-kernels.m4:15, force-kernels.m4:26
constexpr float EPS_REG32 = (float)(10.0 * DBL_EPSILON);
constexpr float LN2_HALF = 0.34657359027997264f;  // ln(2)/2: log(r^2)/2 = log2(r^2) * ln2/2
constexpr int TILE = 8;

__constant__ double c_binomT[TC_MAXORDER * 64];  // c_binomT[k*64 + n] = C(n, k)

TC_DEV float fast_log2(float x) { return __builtin_amdgcn_logf(x); }
TC_DEV float fast_rcp(float x) { return __builtin_amdgcn_rcpf(x); }

// 1/d to ~1 ulp for positive normal d: hardware estimate + two Newton steps.
TC_DEV double rcp64(double d)
{
    double r = __builtin_amdgcn_rcp(d);
    double e = fma(-d, r, 1.0);
    r = fma(r, e, r);
    e = fma(-d, r, 1.0);
    return fma(r, e, r);
}

// Natural log of a positive normal double to about 2 ulp (the P2P argument is
// r^2 + 10 eps >= 2.2e-15).  x = 2^k m with m in [sqrt(1/2), sqrt(2)),
// log m = 2 atanh(s), s = (m - 1)/(m + 1), |s| <= 0.1716, and
// 2 atanh(s) = 2s (1 + z/3 + z^2/5 + ... + z^10/21) with z = s^2 <= 0.0295,
// the Taylor series truncated where the next term is below 2^-57 relative.
// k ln 2 uses ln 2 split into a 32-bit head (k * head exact) and a tail.
TC_DEV double log64(double x)
{
#pragma clang fp contract(fast)
    double m = __builtin_amdgcn_frexp_mant(x);  // [0.5, 1)
    int k = __builtin_amdgcn_frexp_exp(x);
    if (m < 0.70710678118654757) {
        m = m + m;
        k -= 1;
    }
    const double f = m - 1.0;  // exact (Sterbenz)
    const double d = m + 1.0;
    double r = __builtin_amdgcn_rcp(d);
    double e = fma(-d, r, 1.0);
    r = fma(r, e, r);
    e = fma(-d, r, 1.0);
    r = fma(r, e, r);
    double q = f * r;
    q = fma(r, fma(-q, d, f), q);  // s = f / d to ~0.5 ulp
    const double z = q * q;
    double P = 1.0 / 21.0;
    P = fma(P, z, 1.0 / 19.0);
    P = fma(P, z, 1.0 / 17.0);
    P = fma(P, z, 1.0 / 15.0);
    P = fma(P, z, 1.0 / 13.0);
    P = fma(P, z, 1.0 / 11.0);
    P = fma(P, z, 1.0 / 9.0);
    P = fma(P, z, 1.0 / 7.0);
    P = fma(P, z, 1.0 / 5.0);
    P = fma(P, z, 1.0 / 3.0);
    const double q2 = q + q;
    const double logm = fma(q2 * z, P, q2);
    const double kd = (double)k;
    return fma(kd, 0.6931471803691238, fma(kd, 1.9082149292705877e-10, logm));
}

// ------------------------------------------------------------------ potential kernels
// potential_e2p (potential-kernels.m4:59-89) for one expansion:
// Q log|z| + Re sum_{j=1..p} a_j u^j with u = 1/z, evaluated by Horner's rule
// (4 FMAs per coefficient).
template <int P>
TC_DEV double e2p_pot64(double rz, double iz, double mass, const double *__restrict__ ex)
{
    const double r2 = fma(rz, rz, iz * iz);
    const double inv = 1.0 / r2;
    const double ur = rz * inv, ui = -iz * inv;
    double ar = ex[P - 1], ai = ex[2 * P - 1];
#pragma unroll
    for (int j = P - 2; j >= 0; --j) {
        const double tr = fma(ur, ar, fma(-ui, ai, ex[j]));
        const double ti = fma(ur, ai, fma(ui, ar, ex[P + j]));
        ar = tr;
        ai = ti;
    }
    return fma(mass * 0.5, log64(r2), fma(ur, ar, -ui * ai));
}

// FP32 E2P on scaled coefficients b_j = a_j / s^j: Re sum b_j w^j, w = s/z
// (Horner), plus Q log|z| as (Q ln2/2) log2|z|^2; ex[P] = (s, Q ln2/2).
template <int P>
TC_DEV float e2p_pot32(float rz, float iz, const float2 *__restrict__ ex)
{
    const float2 cq = ex[P];
    const float r2 = fmaf(rz, rz, iz * iz);
    const float inv = fast_rcp(r2);
    const float wr = cq.x * rz * inv, wi = -cq.x * iz * inv;
    float ar = ex[P - 1].x, ai = ex[P - 1].y;
#pragma unroll
    for (int j = P - 2; j >= 0; --j) {
        const float tr = fmaf(wr, ar, fmaf(-wi, ai, ex[j].x));
        const float ti = fmaf(wr, ai, fmaf(wi, ar, ex[j].y));
        ar = tr;
        ai = ti;
    }
    return fmaf(cq.y, fast_log2(r2), fmaf(wr, ar, -wi * ai));
}

// force_e2p_8x8 (force-kernels.m4:57-98) for one target: dphi/dz =
// Q u - sum_k k a_k u^(k+1) = u (Q - u T), T = sum_k k a_k u^(k-1) by Horner;
// fx += Re dphi/dz, fy -= Im dphi/dz.
template <int P>
TC_DEV void e2p_force64(double rz, double iz, double mass, const double *__restrict__ ex, double &fx, double &fy)
{
    const double inv = 1.0 / fma(rz, rz, iz * iz);
    const double ur = rz * inv, ui = -iz * inv;
    double ar = (double)P * ex[P - 1], ai = (double)P * ex[2 * P - 1];
#pragma unroll
    for (int j = P - 2; j >= 0; --j) {
        const double k = (double)(j + 1);
        const double tr = fma(ur, ar, fma(-ui, ai, k * ex[j]));
        const double ti = fma(ur, ai, fma(ui, ar, k * ex[P + j]));
        ar = tr;
        ai = ti;
    }
    const double vr = mass - fma(ur, ar, -ui * ai);
    const double vi = -fma(ur, ai, ui * ar);
    fx += fma(ur, vr, -ui * vi);
    fy -= fma(ur, vi, ui * vr);
}

// FP32 force E2P on scaled derivative coefficients d_k = k b_k (the force
// tree's mirror): dphi/dz = u (Q - w T'), w = s u, T' = sum_k d_k w^(k-1) by
// Horner; ex[P] = (s, Q).
template <int P>
TC_DEV void e2p_force32(float zr, float zi, const float2 *__restrict__ ex, float &gx, float &gy)
{
    const float2 cq = ex[P];
    const float inv = fast_rcp(fmaf(zr, zr, zi * zi));
    const float ur = zr * inv, ui = -zi * inv;
    const float wr = cq.x * ur, wi = cq.x * ui;
    float ar = ex[P - 1].x, ai = ex[P - 1].y;
#pragma unroll
    for (int j = P - 2; j >= 0; --j) {
        const float tr = fmaf(wr, ar, fmaf(-wi, ai, ex[j].x));
        const float ti = fmaf(wr, ai, fmaf(wi, ar, ex[j].y));
        ar = tr;
        ai = ti;
    }
    const float vr = cq.y - fmaf(wr, ar, -wi * ai);
    const float vi = -fmaf(wr, ai, wi * ar);
    gx = fmaf(ur, vr, -ui * vi);
    gy = -fmaf(ur, vi, ui * vr);
}

// ------------------------------------------------------------------ FP32 wave-level local expansions
// A node that every target of a wave accepts, and that lies far from the
// wave's targets compared with their spread, is evaluated for the whole wave
// at once: its multipole expansion is translated into a local (Taylor)
// expansion about the centre c of the wave's targets (multipole-to-local),
// the wave sums these local expansions, and each lane evaluates the sum at
// its own target once at the end.  The accept decisions -- and so every
// target's interaction list -- are exactly those of the per-target walk; only
// the evaluation of those far interactions changes.
//
// Local coefficients, scaled by the wave radius R (max |t - c|) so that the
// evaluation point zeta = (t - c) / R has |zeta| <= 1, with z0 = com - c,
// w = s / z0 (s: the node's coefficient scale), v = R / z0, b_k the node's
// scaled coefficients (a_k = b_k s^k):
//   L_0 = Q log|z0| + Re sum_k (-1)^k b_k w^k
//   L_l = v^l (-Q / l + sum_k (-1)^k C(l + k - 1, l) b_k w^k),   l = 1 .. 15
// (the Taylor coefficients of Q log(z - com) + sum_k a_k (z - com)^-k about c).
// The series converges for |t - c| < |z0| - r; a node qualifies when
// R <= 0.3 (|z0| - r), so truncating at l = 15 leaves at most 0.3^16 ~ 4e-9 of
// the node's field.  A node qualifies only if the whole wave reaches it and
// |z0| - R > r / theta' with theta' = theta (1 - 1e-9): then every target's
// distance d >= |z0| - R satisfies r < theta' d, so every lane's own FP64
// test r^2 < theta^2 d^2 accepts it (the margin covers the roundings).
constexpr int M2L_G = 16;            // lanes per node; local order M2L_G - 1
constexpr double M2L_RHO = 0.3;      // R <= M2L_RHO (|z0| - r)
constexpr double M2L_SLACK = 1.0 + 1e-9;  // relative margin of the squared distance test
constexpr int M2L_Q = 128;           // queued translations per wave (a full queue: nodes take the per-lane path)

// One pass of four queued nodes: lane (g, l) = (lane / 16, lane % 16) adds
// L_l of node g (coefficients staged in co, z0 in z) to (lre, lim); qs is
// (2 / ln 2) / l, so cq.y qs = Q / l (0 on lanes l = 0).
template <int P>
TC_DEV void m2l_wave_pass(int nq, const float2 *z, const float2 *co, float rw, float qs, const float *s_beta,
                          double &lre, double &lim)
{
#pragma clang fp contract(fast)
    const int lane = tc_lane();
    const int g = lane / M2L_G, l = lane % M2L_G;
    if (g >= nq)
        return;
    const float2 zz = z[g];
    const float z2 = fmaf(zz.x, zz.x, zz.y * zz.y);
    const float inv = fast_rcp(z2);
    const float ur = zz.x * inv, ui = -zz.y * inv;
    co += g * (P + 1);
    const float2 cq = co[P];  // (s, Q ln2 / 2)
    const float wr = cq.x * ur, wi = cq.x * ui;
    const float *be = s_beta + l * P;
    float tr = 0.f, ti = 0.f;
#pragma unroll 2
    for (int k = P; k >= 1; --k) {
        const float ar = fmaf(be[k - 1], co[k - 1].x, tr), ai = fmaf(be[k - 1], co[k - 1].y, ti);
        tr = fmaf(ar, wr, -ai * wi);
        ti = fmaf(ar, wi, ai * wr);
    }
    float re, im;
    if (l == 0) {
        re = fmaf(cq.y, fast_log2(z2), tr);
        im = 0.f;
    } else {
        // v^l by binary powering, v = R / z0
        float br = rw * ur, bi = rw * ui, pr = 1.f, pi = 0.f;
#pragma unroll
        for (int e = 1; e < M2L_G; e <<= 1) {
            if (l & e) {
                const float t = fmaf(pr, br, -pi * bi);
                pi = fmaf(pr, bi, pi * br);
                pr = t;
            }
            const float t = fmaf(br, br, -bi * bi);
            bi = 2.f * br * bi;
            br = t;
        }
        const float cr = fmaf(-cq.y, qs, tr);
        re = fmaf(pr, cr, -pi * ti);
        im = fmaf(pr, ti, pi * cr);
    }
    lre += (double)re;
    lim += (double)im;
}

// EvaluatePotential::evaluate (treecode-potential.cpp:25-86) for targets
// tperm[64 w .. 64 w + 63] on wave w: masked DFS with the accept test
// r^2 < theta^2 d^2 per lane, E2P for accepting lanes, P2P at leaves.
// FP32: nodes that qualify (above) go through the wave's local expansion.
// CNT: count interactions (a separate instance, so uncounted solves carry no
// counter arithmetic)
// Per-request arguments of a potential traversal (one request of a batch:
// its waves are wave0 .. wave0 + ceil(m / 64) - 1 of the batch).
struct TcPotReq {
    const TcNode *nodes;
    const double *exps;
    const float2 *exps32;
    const double4 *src;
    const float4 *src32;
    const double *xt, *yt;
    const uint32_t *tperm;
    double *out;
    double th2;
    int m, wave0;
};

// A DFS stack entry: the node and the lanes that reach it (one 16-byte LDS
// access per push or pop).
struct alignas(16) TcStackEntry {
    unsigned long long mask;
    int node, pad;
};

// LDS of one workgroup (four waves) of the potential traversal.
template <int P, bool F32> struct TcPotShared {
    TcStackEntry stk[4][TC_STACK];
    double4 src[4][64];
    // FP32 wave-level local expansions: beta[l P + k - 1] = (-1)^k C(l + k - 1, l),
    // queued node ids, their z0 = com - centre, and each wave's {centre, radius, 1 / theta'}
    float beta[F32 ? M2L_G * P : 1];
    int q[4][F32 ? M2L_Q : 1];
    double4 wc[4];
    float2 qz[4][F32 ? M2L_Q : 1];
};

template <int P, bool F32> TC_DEV void potential_shared_init(TcPotShared<P, F32> &sh)
{
    if constexpr (F32) {
        for (int i = threadIdx.x; i < M2L_G * P; i += 256) {
            const int l = i / P, kk = i - l * P + 1;
            const double c = c_binomT[l * 64 + l + kk - 1];
            sh.beta[i] = (float)((kk & 1) ? -c : c);
        }
        __syncthreads();
    }
}

// One wave's traversal: targets tperm[k0 .. k0 + 63] of request rq, k0 = 64 x
// (the wave's index within the request); shared by the one-request kernel
// and the batched one, so a request's results do not depend on which runs it.
template <int P, bool F32, bool CNT>
TC_DEV void potential_wave(TcPotShared<P, F32> &sh, const TcNode *__restrict__ nodes,
                           const double *__restrict__ exps, const float2 *__restrict__ exps32,
                           const double4 *__restrict__ src, const float4 *__restrict__ src32,
                           const double *__restrict__ xt, const double *__restrict__ yt,
                           const uint32_t *__restrict__ tperm, double *__restrict__ out, const double th2,
                           const int m, int k0, unsigned long long *counters)
{
    TcStackEntry(&s_stk)[4][TC_STACK] = sh.stk;
    double4(&s_src)[4][64] = sh.src;
    const float *s_beta = sh.beta;
    int(&s_q)[4][F32 ? M2L_Q : 1] = sh.q;
    double4 *s_wc = sh.wc;
    float2(&s_qz)[4][F32 ? M2L_Q : 1] = sh.qz;
    static_assert(!F32 || 4 * (P + 1) * sizeof(float2) <= sizeof(s_src[0]), "staged coefficients fit a source tile");
    const int wid = threadIdx.x >> 6, lane = tc_lane();
    const int k = k0 + lane;
    const bool valid = k < m;
    const uint32_t ti = valid ? tperm[k] : 0u;
    const double x = valid ? xt[ti] : 0.0, y = valid ? yt[ti] : 0.0;
    const unsigned long long live = tc_ballot(valid);
    if (!live)
        return;
    // FP32: centre (bounding-box middle) and radius of the wave's targets
    int nq = 0;
    unsigned long long nm2l = 0, ne2pw = 0;
    if constexpr (F32) {
        const int first = __builtin_ctzll(live);
        const double fx = __shfl(x, first, 64), fy = __shfl(y, first, 64);
        const double lx = valid ? x : fx, ly = valid ? y : fy;
        double mnx = lx, mxx = lx, mny = ly, mxy = ly;
#pragma unroll
        for (int off = 32; off >= 1; off >>= 1) {
            const double a = __shfl_xor(mnx, off, 64), b = __shfl_xor(mxx, off, 64);
            const double c = __shfl_xor(mny, off, 64), d = __shfl_xor(mxy, off, 64);
            mnx = a < mnx ? a : mnx;
            mxx = b > mxx ? b : mxx;
            mny = c < mny ? c : mny;
            mxy = d > mxy ? d : mxy;
        }
        const double cwx = 0.5 * (mnx + mxx), cwy = 0.5 * (mny + mxy);
        const double ddx = lx - cwx, ddy = ly - cwy;
        const double rwave = sqrt(tc_wave_max(ddx * ddx + ddy * ddy));
        if (lane == 0)
            s_wc[wid] = make_double4(cwx, cwy, rwave, th2 > 0 ? 1.0 / (sqrt(th2) * (1.0 - 1e-9)) : 0.0);
    }
    double acc = 0;
    // FP32 E2P results are summed in FP32 over 8 of the wave's evaluations, then
    // flushed into the FP64 accumulator (each term is already an FP32 value)
    float acc_e2p = 0.f;
    int n_e2p = 0;
    unsigned long long np2p = 0, ne2p = 0, nslots = 0, neslots = 0;
    // CNT: wave clock cycles per category (direct sums, per-target multipoles,
    // wave-level translations, local evaluation) and in total
    unsigned long long cyc0 = 0, cp2p = 0, ce2p = 0, cloc = 0, cl2p = 0, tc0 = 0;
    if constexpr (CNT)
        cyc0 = tc_clock();
    int sp = 0;
    if (lane == 0)
        s_stk[wid][0] = TcStackEntry{live, 0, 0};
    sp = 1;
    while (sp > 0) {
        --sp;
        tc_wave_sync();
        // the entry in one LDS read, made scalar before the node record's loads
        // are issued (LDS and scalar loads share a wait counter: a wait for
        // the LDS read behind the record's loads would split them in two rounds)
        const uint4 top = *reinterpret_cast<const uint4 *>(&s_stk[wid][sp]);
        const int id = tc_uniform((int)top.z);
        const unsigned long long mask =
            ((unsigned long long)(unsigned)tc_uniform((int)top.y) << 32) | (unsigned)tc_uniform((int)top.x);
        // no early exit for empty nodes: all fields of the record are needed
        // before any branch, so they arrive in one round of scalar loads (an
        // empty node leaves no lane accepting or opening it)
        const TcNode nd = nodes[id];
        const bool nonempty = nd.s != nd.e;
        const double dx = x - nd.xcom, dy = y - nd.ycom;
        const double r2 = dx * dx + dy * dy;
        // the lanes that accept, as a wave mask: the compare's own lane mask
        // combined on the scalar unit; the lane predicate is that mask itself
        const unsigned long long amask = tc_ballot(nd.rr < th2 * r2) & mask & (nonempty ? ~0ull : 0ull);
        const bool acc_ok = __builtin_amdgcn_inverse_ballot_w64(amask);
        if constexpr (CNT)
            tc0 = tc_clock();
        if (acc_ok) {
            if (F32)
                acc_e2p += e2p_pot32<P>((float)dx, (float)dy, exps32 + (size_t)(P + 1) * id);
            else
                acc += e2p_pot64<P>(dx, dy, nd.mass, exps + (size_t)2 * P * id);
        }
        if (F32 && amask && ++n_e2p == 8) {
            acc += (double)acc_e2p;
            acc_e2p = 0.f;
            n_e2p = 0;
        }
        if constexpr (CNT) {
            ne2p += __popcll(amask);
            neslots += amask ? 64ull : 0ull;
            if (amask)
                ce2p += tc_clock() - tc0;
        }
        const unsigned long long omask = nonempty ? mask & ~amask : 0ull;
        if (!omask)
            continue;
        if (nd.child == 0) {
            if constexpr (CNT)
                tc0 = tc_clock();
            const bool act = __builtin_amdgcn_inverse_ballot_w64(omask);
            const int nact = __popcll(omask), nsrc = nd.e - nd.s, nchunk = (nsrc + 63) >> 6;
            if constexpr (CNT)
                np2p += (unsigned long long)nact * (unsigned long long)nsrc;
            // Few lanes need this leaf: lanes take sources instead and the wave
            // loops over the active targets (8 at a time), one reduction per
            // target.  Chosen when k (C 9.5 + 16.5) < n 9.8, the measured
            // nanoseconds per SIMD of the two forms (FP32 op 1.18 ns, log 3.6 ns;
            // C chunks of 64 of the n sources, k active lanes).
            if (F32 && (long long)nact * (nchunk * 19 + 33) < (long long)nsrc * 20) {
                if constexpr (CNT)
                    nslots += 64ull * (unsigned long long)nchunk * (unsigned long long)nact;
                const float txl = (float)(x - nd.xcom), tyl = (float)(y - nd.ycom);
                unsigned long long rest = omask;
                double part = 0;
                while (rest) {
                    int tl[8];
                    float ttx[8], tty[8], a[8];
                    int g = 0;
#pragma unroll
                    for (int j = 0; j < 8; ++j) {
                        a[j] = 0.f;
                        tl[j] = -1;
                        ttx[j] = tty[j] = 0.f;
                        if (rest) {
                            const int t = __builtin_ctzll(rest);
                            rest &= rest - 1;
                            tl[j] = t;
                            ttx[j] = __int_as_float(__builtin_amdgcn_readlane(__float_as_int(txl), t));
                            tty[j] = __int_as_float(__builtin_amdgcn_readlane(__float_as_int(tyl), t));
                            g = j + 1;
                        }
                    }
                    for (int b0 = nd.s; b0 < nd.e; b0 += 64) {
                        const float4 q = b0 + lane < nd.e ? src32[b0 + lane] : make_float4(0.f, 0.f, 0.f, 0.f);
#pragma unroll
                        for (int j = 0; j < 8; ++j) {
                            if (j < g) {
#pragma clang fp contract(fast)
                                const float ax = ttx[j] - q.x, ay = tty[j] - q.y;
                                a[j] += q.z * fast_log2(ax * ax + (ay * ay + EPS_REG32));
                            }
                        }
                    }
#pragma unroll
                    for (int j = 0; j < 8; ++j) {
                        if (j < g) {
                            const float sj = tc_wave_sum_f32(a[j]);
                            if (lane == tl[j])
                                part = (double)sj;
                        }
                    }
                }
                acc += part * (double)LN2_HALF;
            } else if (F32) {
                // Dense: sources staged 64 at a time in LDS.  Lanes l and l^32
                // split each tile (even / odd sources) and each computes both
                // targets, its own and its partner's: one broadcast LDS read
                // (768 bytes of LDS return per wave) feeds two pair evaluations,
                // which keeps the LDS well below the VALU's rate even when few
                // waves are resident.  Partner sums are exchanged at the end of
                // the leaf.
                if constexpr (CNT)
                    nslots += 64ull * (unsigned long long)nsrc;
                const float tx = (float)(x - nd.xcom), ty = (float)(y - nd.ycom);
                const float ux = __shfl_xor(tx, 32, 64), uy = __shfl_xor(ty, 32, 64);
                const int half = lane >> 5;
                const float4 zero4 = make_float4(0.f, 0.f, 0.f, 0.f);
                float4 *tile = reinterpret_cast<float4 *>(&s_src[wid][0]);
                double own = 0, other = 0;
                for (int b0 = nd.s; b0 < nd.e; b0 += 64) {
                    const int cnt = min(64, nd.e - b0);
                    if (lane < cnt)
                        tile[lane] = src32[b0 + lane];
                    tc_wave_sync();
                    float s0 = 0, s1 = 0, o0 = 0, o1 = 0;
                    int j = 0;
                    for (; 2 * j + 3 < cnt; j += 2) {
#pragma clang fp contract(fast)
                        const float4 qa = tile[2 * j + half], qb = tile[2 * j + 2 + half];
                        float ax = tx - qa.x, ay = ty - qa.y;
                        s0 += qa.z * fast_log2(ax * ax + (ay * ay + EPS_REG32));
                        ax = ux - qa.x;
                        ay = uy - qa.y;
                        o0 += qa.z * fast_log2(ax * ax + (ay * ay + EPS_REG32));
                        ax = tx - qb.x;
                        ay = ty - qb.y;
                        s1 += qb.z * fast_log2(ax * ax + (ay * ay + EPS_REG32));
                        ax = ux - qb.x;
                        ay = uy - qb.y;
                        o1 += qb.z * fast_log2(ax * ax + (ay * ay + EPS_REG32));
                    }
                    for (; 2 * j < cnt; ++j) {
#pragma clang fp contract(fast)
                        const int idx = 2 * j + half;
                        const float4 qa = idx < cnt ? tile[idx] : zero4;
                        float ax = tx - qa.x, ay = ty - qa.y;
                        s0 += qa.z * fast_log2(ax * ax + (ay * ay + EPS_REG32));
                        ax = ux - qa.x;
                        ay = uy - qa.y;
                        o0 += qa.z * fast_log2(ax * ax + (ay * ay + EPS_REG32));
                    }
                    own += (double)(s0 + s1);
                    other += (double)(o0 + o1);
                    tc_wave_sync();
                }
                const double leafsum = own + __shfl_xor(other, 32, 64);
                if (act)
                    acc += leafsum * (double)LN2_HALF;
            } else {
                if constexpr (CNT)
                    nslots += 64ull * (unsigned long long)nsrc;
                double leafsum = 0, leafsum2 = 0;
                for (int b0 = nd.s; b0 < nd.e; b0 += 64) {
                    const int cnt = min(64, nd.e - b0);
                    if (lane < cnt)
                        s_src[wid][lane] = src[b0 + lane];
                    tc_wave_sync();
                    if (act) {
#pragma clang fp contract(fast)
                        int j = 0;
                        for (; j + 2 <= cnt; j += 2) {
                            const double4 q0 = s_src[wid][j], q1 = s_src[wid][j + 1];
                            const double ax = x - q0.x, ay = y - q0.y;
                            const double bx = x - q1.x, by = y - q1.y;
                            leafsum += log64(ax * ax + ay * ay + EPS_REG) * q0.z;
                            leafsum2 += log64(bx * bx + by * by + EPS_REG) * q1.z;
                        }
                        if (j < cnt) {
                            const double4 q = s_src[wid][j];
                            const double ax = x - q.x, ay = y - q.y;
                            leafsum += log64(ax * ax + ay * ay + EPS_REG) * q.z;
                        }
                    }
                    tc_wave_sync();
                }
                acc += (leafsum + leafsum2) * 0.5;
            }
            if constexpr (CNT)
                cp2p += tc_clock() - tc0;
        } else {
            if constexpr (F32) {
                if (omask == live) {
                    // the whole wave opens this node: children that qualify for the
                    // wave's local expansion are queued, the others pushed (order
                    // 0..3 kept), empty ones dropped (they contribute nothing)
                    bool m2l = false, push = false;
                    float2 z0 = make_float2(0.f, 0.f);
                    tc_wave_sync();
                    if (lane < 4) {
                        const double4 wc = s_wc[wid];
                        const TcNode *cn = nodes + nd.child + lane;
                        const int2 cse = *reinterpret_cast<const int2 *>(&cn->s);
                        const double4 cg = *reinterpret_cast<const double4 *>(cn);
                        if (cse.x < cse.y) {
                            // |z0| > R + r / theta' and |z0| > r + R / rho, compared squared
                            // with a relative margin (wc.w = 1 / theta')
                            const double ddx = cg.x - wc.x, ddy = cg.y - wc.y;
                            const double d2 = ddx * ddx + ddy * ddy;
                            const double a = fmax(wc.z + cg.z * wc.w, cg.z + wc.z * (1.0 / M2L_RHO));
                            m2l = th2 > 0 && d2 > a * a * M2L_SLACK;  // theta 0 accepts nothing
                            push = !m2l;
                            z0 = make_float2((float)ddx, (float)ddy);
                        }
                    }
                    // a full queue sends the rest down the per-lane path
                    if (m2l && nq + 4 > M2L_Q) {
                        m2l = false;
                        push = true;
                    }
                    const unsigned long long mm = tc_ballot(m2l), pm = tc_ballot(push);
                    if (m2l) {
                        const int qi = nq + __popcll(mm & tc_lanemask_lt());
                        s_q[wid][qi] = nd.child + lane;
                        s_qz[wid][qi] = z0;
                    }
                    if (push) {
                        const int pos = sp + __popcll(pm & tc_lanemask_lt());
                        s_stk[wid][pos] = TcStackEntry{omask, nd.child + lane, 0};
                    }
                    sp += __popcll(pm);
                    const int nadd = __popcll(mm);
                    nq += nadd;
                    if constexpr (CNT) {
                        nm2l += nadd;
                        ne2pw += (unsigned long long)nadd * __popcll(live);
                        ne2p += (unsigned long long)nadd * __popcll(live);
                    }
                    continue;
                }
            }
            // children pushed in order 0..3, popped 3..0 (treecode-potential.cpp:76-79)
            if (lane < 4)
                s_stk[wid][sp + lane] = TcStackEntry{omask, nd.child + lane, 0};
            sp += 4;
        }
    }
    if constexpr (F32) {
        if (nq > 0) {
            // the queued translations, four nodes per pass (the source tile's
            // LDS stages their coefficients)
            tc_wave_sync();
            float2 *s_coef = reinterpret_cast<float2 *>(&s_src[wid][0]);
            const double4 wc = s_wc[wid];
            const double cwx = wc.x, cwy = wc.y, rwave = wc.z;
            const int ll = lane % M2L_G;
            const float qs = ll ? (1.f / LN2_HALF) / (float)ll : 0.f;
            double lre = 0, lim = 0;
            // coefficients of the next pass are fetched while the current one is computed
            constexpr int NST = (4 * (P + 1) + 63) / 64;
            float2 st[NST];
            auto fetch = [&](int q0) {
                const int nqq = min(4, nq - q0);
#pragma unroll
                for (int j = 0; j < NST; ++j) {
                    const int i = lane + 64 * j, g = i / (P + 1), kk = i - g * (P + 1);
                    st[j] = g < nqq ? exps32[(size_t)(P + 1) * s_q[wid][q0 + g] + kk] : make_float2(0.f, 0.f);
                }
            };
            if constexpr (CNT)
                tc0 = tc_clock();
            fetch(0);
            for (int q0 = 0; q0 < nq; q0 += 4) {
#pragma unroll
                for (int j = 0; j < NST; ++j)
                    if (lane + 64 * j < 4 * (P + 1))
                        s_coef[lane + 64 * j] = st[j];
                tc_wave_sync();
                if (q0 + 4 < nq)
                    fetch(q0 + 4);
                m2l_wave_pass<P>(min(4, nq - q0), s_qz[wid] + q0, s_coef, (float)rwave, qs, s_beta, lre, lim);
                tc_wave_sync();
            }
            if constexpr (CNT) {
                const unsigned long long t1 = tc_clock();
                cloc += t1 - tc0;
                tc0 = t1;
            }
            // sum the four node groups' partial coefficients, then every lane
            // evaluates sum_l L_l zeta^l at its own target (Horner, FP64)
            lre += __shfl_xor(lre, 16, 64);
            lre += __shfl_xor(lre, 32, 64);
            lim += __shfl_xor(lim, 16, 64);
            lim += __shfl_xor(lim, 32, 64);
            const double sc = rwave > 0 ? 1.0 / rwave : 0.0;
            const double zr = (x - cwx) * sc, zi = (y - cwy) * sc;
            double hr = 0, hi = 0;
#pragma unroll 1
            for (int l = M2L_G - 1; l >= 0; --l) {
                const double cr = tc_lane_value(lre, l), ci = tc_lane_value(lim, l);
                const double t = hr * zr - hi * zi + cr;
                hi = hr * zi + hi * zr + ci;
                hr = t;
            }
            acc += hr;
            if constexpr (CNT)
                cl2p += tc_clock() - tc0;
        }
    }
    if (valid)
        out[ti] = acc + (double)acc_e2p;
    if (CNT && lane == 0) {
        atomicAdd(&counters[TC_CNT_P2P], np2p);
        atomicAdd(&counters[TC_CNT_E2P], ne2p);
        atomicAdd(&counters[TC_CNT_P2P_SLOTS], nslots);
        atomicAdd(&counters[TC_CNT_E2P_SLOTS], neslots);
        if (F32) {
            atomicAdd(&counters[TC_CNT_M2L_WAVE], nm2l);
            atomicAdd(&counters[TC_CNT_E2P_WAVE], ne2pw);
        }
        atomicAdd(&counters[TC_CNT_CYC_P2P], cp2p);
        atomicAdd(&counters[TC_CNT_CYC_E2P], ce2p);
        atomicAdd(&counters[TC_CNT_CYC_LOCAL], cloc);
        atomicAdd(&counters[TC_CNT_CYC_L2P], cl2p);
        atomicAdd(&counters[TC_CNT_CYC_TOTAL], tc_clock() - cyc0);
    }
}

// (separate kernel arguments rather than a TcPotReq: a by-value struct
// argument costs this kernel registers and occupancy)
template <int P, bool F32, bool CNT>
__global__ __launch_bounds__(256) void k_potential(const TcNode *__restrict__ nodes, const double *__restrict__ exps,
                                                   const float2 *__restrict__ exps32,
                                                   const double4 *__restrict__ src,
                                                   const float4 *__restrict__ src32, const double *__restrict__ xt,
                                                   const double *__restrict__ yt, const uint32_t *__restrict__ tperm,
                                                   int m, double th2, double *__restrict__ out,
                                                   unsigned long long *counters, int xcd_lc)
{
    __shared__ TcPotShared<P, F32> sh;
    potential_shared_init<P, F32>(sh);
    potential_wave<P, F32, CNT>(sh, nodes, exps, exps32, src, src32, xt, yt, tperm, out, th2, m,
                                (tc_xcd_chunk_block(xcd_lc) * 4 + (threadIdx.x >> 6)) * 64, counters);
}

// Cross-request traversal: a persistent grid drains the target waves of all
// meta[0] requests of a batch (meta[1] waves in total).  Each wave takes the
// next wave index g from the work counter meta[2] (zero in the descriptor
// block that is copied in before every launch), finds g's request by binary
// search over the requests' first waves and does the same per-wave work as
// k_potential, so each result equals the one-request launch's whichever wave
// of the grid computes it.  Every wave leaves once the counter has passed
// meta[1].  The descriptors, the batch size and the counter live in device
// memory and the grid depends only on the size bucket (its waves or the
// resident capacity, whichever is smaller), so a launch has fixed arguments:
// one graph per (order, precision, size bucket) replays it.
template <int P, bool F32, bool CNT>
__global__ __launch_bounds__(256) void k_potential_batch(const TcPotReq *__restrict__ reqs, int *meta,
                                                         unsigned long long *counters)
{
    __shared__ TcPotShared<P, F32> sh;
    potential_shared_init<P, F32>(sh);
    const int nreq = meta[0], nwaves = meta[1];
    for (;;) {
        int g = 0;
        if (tc_lane() == 0)
            g = atomicAdd(&meta[2], 1);
        g = __builtin_amdgcn_readfirstlane(g);
        if (g >= nwaves)
            break;
        int lo = 0, hi = nreq - 1;
        while (lo < hi) {
            const int mid = (lo + hi + 1) >> 1;
            if (reqs[mid].wave0 <= g)
                lo = mid;
            else
                hi = mid - 1;
        }
        // the descriptor is wave-uniform (scalar registers) and its pointers
        // point to global memory: telling the compiler so lets node records and
        // coefficients travel through scalar loads as in k_potential
        const TcPotReq &rq = reqs[lo];
        auto up = [](const void *p) {
            typedef const __attribute__((address_space(1))) char *gptr;
            const gptr q = (gptr)(tc_uniform64((unsigned long long)p));
            return (const void *)q;
        };
        potential_wave<P, F32, CNT>(
            sh, static_cast<const TcNode *>(up(rq.nodes)), static_cast<const double *>(up(rq.exps)),
            static_cast<const float2 *>(up(rq.exps32)), static_cast<const double4 *>(up(rq.src)),
            static_cast<const float4 *>(up(rq.src32)), static_cast<const double *>(up(rq.xt)),
            static_cast<const double *>(up(rq.yt)), static_cast<const uint32_t *>(up(rq.tperm)),
            const_cast<double *>(static_cast<const double *>(up(rq.out))), tc_uniform(rq.th2), tc_uniform(rq.m),
            (g - tc_uniform(rq.wave0)) * 64, counters);
    }
}

// ------------------------------------------------------------------ force kernels
// Multipole-to-local (downward_e2l, force-kernels.m4:102-155) for up to
// E2L_NQ nodes at once, from coefficients and node records staged in LDS: the
// wave splits into groups of E2L_G lanes (8 for p <= 8, 16 for p <= 16, else
// 32), group g takes node g and lane l of a group owns local coefficient l+1.
// z0 = com - brick centre, u = 1/z0; each lane forms u^(l+1) by binary
// powering, publishes (-1)^k a_k u^k, then sums C(l+k, k-1) (-1)^k a_k u^k
// over k and multiplies by u^(l+1).
template <int P> constexpr int E2L_G = P <= 8 ? 8 : (P <= 16 ? 16 : 32);
template <int P> constexpr int E2L_NQ = 64 / E2L_G<P>;

// rec[g] = {xcom, ycom, r, mass}; co[g * 2P + j] = Re a_(j+1), co[g * 2P + P + j] = Im a_(j+1)
template <int P>
TC_DEV void e2l_group(const double4 *rec, const double *co, int nq, double x0b, double y0b, const double *s_binom,
                      const double *s_ninv, double2 *s_c, double &lre, double &lim)
{
#pragma clang fp contract(fast)
    constexpr int G = E2L_G<P>;
    const int lane = tc_lane();
    const int h = lane / G, l = lane % G;
    const bool on = h < nq && l < P;
    double ur = 0, ui = 0, pr = 1, pi = 0, mass = 0;
    if (h < nq) {
        const double4 nd = rec[h];
        mass = nd.w;
        const double zx = nd.x - x0b, zy = nd.y - y0b;
        const double inv = rcp64(zx * zx + zy * zy);
        ur = zx * inv;
        ui = -zy * inv;
        double br = ur, bi = ui;
        int e = l + 1;  // <= G: 4 bits for G = 8, 5 for G = 16, 6 for G = 32
#pragma unroll
        for (int bit = 0; bit < (G == 8 ? 4 : (G == 16 ? 5 : 6)); ++bit) {
            if (e & 1) {
                const double tr = pr * br - pi * bi;
                pi = pr * bi + pi * br;
                pr = tr;
            }
            e >>= 1;
            const double sr = br * br - bi * bi;
            bi = 2.0 * br * bi;
            br = sr;
        }
        if (on) {
            const double ar = co[h * 2 * P + l], ai = co[h * 2 * P + P + l];
            const double sgn = (l & 1) ? 1.0 : -1.0;  // (-1)^k, k = l + 1
            s_c[h * (G + 1) + l + 1] = make_double2(sgn * (ar * pr - ai * pi), sgn * (ar * pi + ai * pr));
        }
    }
    tc_wave_sync();
    if (on) {
        const int L = l + 1;
        double rt = mass * s_ninv[L];  // -Q / L
        double it = 0;
        const double2 *c = s_c + h * (G + 1);
#pragma unroll 2
        for (int kk = 1; kk <= P; ++kk) {
            const double bb = s_binom[(kk - 1) * P + l];  // C(L + kk - 1, kk - 1)
            const double2 ck = c[kk];
            rt += bb * ck.x;
            it += bb * ck.y;
        }
        lre += rt * pr - it * pi;
        lim += rt * pi + it * pr;
    }
    tc_wave_sync();
}

// Node record of the force traversal: the first 48 bytes of a TcNode (centre,
// radius, mass, source range, first child).
struct alignas(16) TcForceRec {
    double xcom, ycom, r, mass;
    int s, e, child, level;
};
static_assert(sizeof(TcForceRec) == 48 && offsetof(TcNode, level) == 44, "force record is TcNode's first 48 B");

enum TcAction { ACT_SKIP = 0, ACT_E2L = 1, ACT_E2P = 2, ACT_P2P = 3, ACT_OPEN = 4 };

struct TcBrick {
    double x0b, y0b, rbrick, xlo, xhi, ylo, yhi, invtheta, th2;
};

TC_DEV TcForceRec force_load(const TcNode *__restrict__ nodes, int id)
{
    const double2 *g = reinterpret_cast<const double2 *>(nodes + id);
    const double2 a = g[0], b = g[1], c = g[2];
    TcForceRec nd;
    nd.xcom = a.x;
    nd.ycom = a.y;
    nd.r = b.x;
    nd.mass = b.y;
    nd.s = __double2loint(c.x);
    nd.e = __double2hiint(c.x);
    nd.child = __double2loint(c.y);
    nd.level = __double2hiint(c.y);
    return nd;
}

// The per-brick decisions of treecode-force.cpp:101-199 for one node: empty
// nodes are skipped; E2L when (|c - brick centre| / r - 1) > 1/theta and the
// brick radius <= r; E2P when r^2 < theta^2 d^2, d the distance from the node
// centre to the brick's closest point; else P2P at a leaf, else open.
TC_DEV int force_action(const TcForceRec &nd, const TcBrick &bk)
{
    if (nd.s == nd.e)
        return ACT_SKIP;
    const double ddx = bk.x0b - nd.xcom, ddy = bk.y0b - nd.ycom;
    const double distance = sqrt(ddx * ddx + ddy * ddy);
    if ((distance / nd.r - 1) > bk.invtheta && bk.rbrick <= nd.r)
        return ACT_E2L;
    const double xm = nd.xcom < bk.xhi ? nd.xcom : bk.xhi, ym = nd.ycom < bk.yhi ? nd.ycom : bk.yhi;
    const double cxt = bk.xlo < xm ? xm : bk.xlo;
    const double cyt = bk.ylo < ym ? ym : bk.ylo;
    const double r2 = (cxt - nd.xcom) * (cxt - nd.xcom) + (cyt - nd.ycom) * (cyt - nd.ycom);
    if (nd.r * nd.r < bk.th2 * r2)
        return ACT_E2P;
    return nd.child == 0 ? ACT_P2P : ACT_OPEN;
}

// Per-wave LDS of the force traversal.  The stack holds node ids; each step
// pops up to 64 of them, one per lane, so one round trip to memory classifies
// a whole frontier.  The E2L / E2P / P2P lists collect node ids and are
// flushed in batches whose records, coefficients or sources are copied into
// `buf` by global_load_lds (memory to LDS without passing through VGPRs).
constexpr int FS_CAP = 256;                  // stack entries
constexpr int FS_RESERVE = 3 * TC_LMAX + 4;  // growth bound of a one-node-at-a-time DFS
constexpr int FQ_CAP = 128;                  // entries per work list
constexpr int FBUF_BYTES = 2048;             // staging buffer

template <int P, bool F32> struct ForceCfg {
    // E2P batch: node records (32 B) + coefficients per node
    static constexpr int E2P_BYTES = 32 + (F32 ? 8 * (P + 1) : 16 * P);
    static constexpr int E2P_QB = FBUF_BYTES / E2P_BYTES;
    // E2L batch: node records (32 B) + 2p coefficients per node, whole groups
    static constexpr int E2L_BYTES = 32 + 16 * P;
    static constexpr int E2L_QB = FBUF_BYTES / E2L_BYTES / E2L_NQ<P> * E2L_NQ<P>;
    static_assert(E2P_QB >= 1 && E2L_QB >= E2L_NQ<P>, "staging buffer too small for this order");
};

// P2P tiles: two 1 KiB halves of the staging buffer, 64 FP32 sources (float4)
// or 32 FP64 sources (x,y and v,- halves of the double4 in separate 512 B rows)
template <bool F32> constexpr int P2P_TILE_N = F32 ? 64 : 32;
// Near field of the force traversal for one range of bricks (brick indices
// relative to the range).  The traversal sums a brick's first FP2P_INLINE
// P2P leaves itself; further leaves go to chunks of up to 64 leaves from a
// shared pool, linked per brick in traversal order, each summed by its own
// wave (k_force_chunks) and added to the brick in chain order
// (k_force_combine), so that a brick with thousands of near leaves does not
// leave one wave running long after the rest.  A brick whose chunks do not fit
// in the pool is marked in pmark and listed in ovf; the NEAR instantiation of
// k_force then repeats its traversal and sums all its P2P leaves.
constexpr int FP2P_INLINE = 64;  // default of the inline count (TcNearLists::inl)
enum { FN_OVF = 0, FN_CHUNKS = 1, FN_OBRICKS = 2, FN_NCTR = 4 };
struct TcNearLists {
    uint32_t *pmark;    // [nb] 1: the brick is listed in ovf
    int *ohead;         // [nb] first chunk of a brick listed in obricks
    uint32_t *obricks;  // [nb] bricks with chunks
    uint32_t *ovf;      // [nb] bricks marked FP2P_OVERFLOW (absolute brick ids)
    int4 *cdesc;        // [maxc] chunk: brick, leaf count, next chunk of the brick (-1 ends)
    uint32_t *cpool;    // [maxc * 64] leaf ids
    double2 *cpart;     // [maxc * 64] chunk sums per target
    uint32_t *ctr;      // [FN_NCTR] counters
    int maxc;
    int inl;            // P2P leaves a brick sums in the traversal before chunks take the rest
};

typedef __attribute__((address_space(3))) void tc_lds_void;
typedef __attribute__((address_space(1))) const void tc_glb_void;

// global_load_lds: lane i copies BYTES (4 or 16) from gsrc to lds + i * BYTES.
// Lanes that are switched off copy nothing.  The copy is counted by vmcnt.
template <int BYTES> TC_DEV void lds_dma(const void *gsrc, void *lds)
{
    static_assert(BYTES == 4 || BYTES == 16, "global_load_lds copies 4 or 16 bytes per lane here");
    if constexpr (BYTES == 16)
        __builtin_amdgcn_global_load_lds((tc_glb_void *)gsrc, (tc_lds_void *)lds, 16, 0, 0);
    else
        __builtin_amdgcn_global_load_lds((tc_glb_void *)gsrc, (tc_lds_void *)lds, 4, 0, 0);
}

// s_waitcnt vmcnt(0): every vector memory operation of this wave, the LDS
// copies included, has completed.
TC_DEV void vm_wait_all() { __builtin_amdgcn_s_waitcnt(0x0F70); }

// Node records {xcom, ycom, r, mass} of list[b0 .. b0 + cnt) into rec[0 .. cnt).
TC_DEV void stage_records(const int *list, int b0, int cnt, const TcNode *__restrict__ nodes, double4 *rec)
{
    const int lane = tc_lane();
    if (lane < 2 * cnt)
        lds_dma<16>(reinterpret_cast<const char *>(nodes + list[b0 + (lane >> 1)]) + 16 * (lane & 1), rec);
    if (64 + lane < 2 * cnt)
        lds_dma<16>(reinterpret_cast<const char *>(nodes + list[b0 + 32 + (lane >> 1)]) + 16 * (lane & 1),
                    rec + 32);
}

// Rows of ROW bytes (a multiple of CH, CH = 4 or 16) at base + ROW * list[b0 + q]
// for q < cnt, packed into dst.
template <int ROW, int CH, int QB>
TC_DEV void stage_rows(const int *list, int b0, int cnt, const void *base, void *dst)
{
    constexpr int NCH = ROW / CH;
    constexpr int IT = (QB * NCH + 63) / 64;
    static_assert(ROW % CH == 0, "row is not a whole number of chunks");
    const int lane = tc_lane();
#pragma unroll
    for (int u = 0; u < IT; ++u) {
        const int i = lane + 64 * u, q = i / NCH, c = i - q * NCH;
        if (i < cnt * NCH)
            lds_dma<CH>(static_cast<const char *>(base) + (size_t)ROW * list[b0 + q] + CH * c,
                        static_cast<char *>(dst) + 64 * CH * u);
    }
}

// E2P of the listed nodes at the brick's 64 targets (force_e2p_8x8,
// force-kernels.m4:57-98), E2P_QB nodes per staged batch.
template <int P, bool F32>
TC_DEV void force_flush_e2p(const int *list, int n, const TcNode *__restrict__ nodes,
                            const double *__restrict__ exps, const float2 *__restrict__ exps32, double4 *buf,
                            double xb, double yb, int ix, int iy, double h, double &fx, double &fy)
{
    constexpr int QB = ForceCfg<P, F32>::E2P_QB;
    double4 *rec = buf;
    void *co = rec + QB;
    for (int b0 = 0; b0 < n; b0 += QB) {
        const int cnt = min(QB, n - b0);
        stage_records(list, b0, cnt, nodes, rec);
        if (F32)
            stage_rows<8 * (P + 1), 4, QB>(list, b0, cnt, exps32, co);
        else
            stage_rows<16 * P, 16, QB>(list, b0, cnt, exps, co);
        vm_wait_all();
        tc_wave_sync();
        for (int t = 0; t < cnt; ++t) {
            const double4 r = rec[t];
            const double rz = (xb - r.x) + ix * h, iz = (yb - r.y) + iy * h;
            if (F32) {
                float gx, gy;
                e2p_force32<P>((float)rz, (float)iz, static_cast<const float2 *>(co) + t * (P + 1), gx, gy);
                fx += (double)gx;
                fy += (double)gy;
            } else {
                e2p_force64<P>(rz, iz, r.w, static_cast<const double *>(co) + t * 2 * P, fx, fy);
            }
        }
        tc_wave_sync();
    }
}

// E2L of the listed nodes about the brick centre (downward_e2l), E2L_QB
// nodes per staged batch, E2L_NQ per e2l_group call.
template <int P, bool F32>
TC_DEV void force_flush_e2l(const int *list, int n, const TcNode *__restrict__ nodes,
                            const double *__restrict__ exps, double4 *buf, double x0b, double y0b,
                            const double *s_binom, const double *s_ninv, double2 *s_c, double &lre, double &lim)
{
    constexpr int QB = ForceCfg<P, F32>::E2L_QB;
    double4 *rec = buf;
    double *co = reinterpret_cast<double *>(rec + QB);
    for (int b0 = 0; b0 < n; b0 += QB) {
        const int cnt = min(QB, n - b0);
        stage_records(list, b0, cnt, nodes, rec);
        stage_rows<16 * P, 16, QB>(list, b0, cnt, exps, co);
        vm_wait_all();
        tc_wave_sync();
        for (int g0 = 0; g0 < cnt; g0 += E2L_NQ<P>)
            e2l_group<P>(rec + g0, co + g0 * 2 * P, min(E2L_NQ<P>, cnt - g0), x0b, y0b, s_binom, s_ninv, s_c, lre,
                         lim);
    }
}

TC_DEV double readlane_f64(double v, int l)
{
    const long long b = __double_as_longlong(v);
    const int lo = __builtin_amdgcn_readlane((int)(b & 0xffffffff), l);
    const int hi = __builtin_amdgcn_readlane((int)(b >> 32), l);
    return __longlong_as_double(((long long)hi << 32) | (unsigned int)lo);
}

// Copy the P2P tile [b0, min(b0 + TILE_N, ne)) into an LDS half (see ForceCfg).
template <bool F32>
TC_DEV void p2p_tile_load(const double4 *__restrict__ src, const float4 *__restrict__ src32, int b0, int ne,
                          void *dst)
{
    const int lane = tc_lane();
    if (F32) {
        if (b0 + lane < ne)
            lds_dma<16>(src32 + b0 + lane, dst);
    } else {
        if (b0 + (lane & 31) < ne)
            lds_dma<16>(reinterpret_cast<const char *>(src + b0 + (lane & 31)) + 16 * (lane >> 5), dst);
    }
}

// One P2P tile of cnt sources from `cur` while the next tile (if `more`) is
// copied into `nxt`.  The two halves never alias, which the restrict
// qualifiers tell the compiler, so it does not wait for the copy before
// reading `cur`.
template <bool F32>
TC_DEV void p2p_tile_sum(const void *__restrict__ cur, void *__restrict__ nxt, bool more, int nb0, int nne,
                         const double4 *__restrict__ src, const float4 *__restrict__ src32, int cnt, double xt,
                         double yt, double xcom, double ycom, double &fx, double &fy)
{
    if (more)
        p2p_tile_load<F32>(src, src32, nb0, nne, nxt);
    if (F32) {
        // Lanes l and l^32 split the tile (even / odd sources) and each
        // evaluates both targets, its own and its partner's: one broadcast LDS
        // read feeds two pair evaluations.  Partner sums are exchanged at the
        // end of the tile.
        const float tx = (float)(xt - xcom), ty = (float)(yt - ycom);
        const float ux = __shfl_xor(tx, 32, 64), uy = __shfl_xor(ty, 32, 64);
        const int half = tc_lane() >> 5;
        const float4 zero4 = make_float4(0.f, 0.f, 0.f, 0.f);
        const float4 *tile = static_cast<const float4 *>(cur);
        float sx0 = 0.f, sy0 = 0.f, sx1 = 0.f, sy1 = 0.f;  // own target
        float ox0 = 0.f, oy0 = 0.f, ox1 = 0.f, oy1 = 0.f;  // partner's target
        int j = 0;
        for (; 2 * j + 3 < cnt; j += 2) {
#pragma clang fp contract(fast)
            const float4 qa = tile[2 * j + half], qb = tile[2 * j + 2 + half];
            float dx = tx - qa.x, dy = ty - qa.y;
            float f = qa.z * fast_rcp(dx * dx + (dy * dy + EPS_REG32));
            sx0 += dx * f;
            sy0 += dy * f;
            dx = ux - qa.x;
            dy = uy - qa.y;
            f = qa.z * fast_rcp(dx * dx + (dy * dy + EPS_REG32));
            ox0 += dx * f;
            oy0 += dy * f;
            dx = tx - qb.x;
            dy = ty - qb.y;
            f = qb.z * fast_rcp(dx * dx + (dy * dy + EPS_REG32));
            sx1 += dx * f;
            sy1 += dy * f;
            dx = ux - qb.x;
            dy = uy - qb.y;
            f = qb.z * fast_rcp(dx * dx + (dy * dy + EPS_REG32));
            ox1 += dx * f;
            oy1 += dy * f;
        }
        for (; 2 * j < cnt; ++j) {
#pragma clang fp contract(fast)
            const int idx = 2 * j + half;
            const float4 qa = idx < cnt ? tile[idx] : zero4;
            float dx = tx - qa.x, dy = ty - qa.y;
            float f = qa.z * fast_rcp(dx * dx + (dy * dy + EPS_REG32));
            sx0 += dx * f;
            sy0 += dy * f;
            dx = ux - qa.x;
            dy = uy - qa.y;
            f = qa.z * fast_rcp(dx * dx + (dy * dy + EPS_REG32));
            ox0 += dx * f;
            oy0 += dy * f;
        }
        fx += (double)(sx0 + sx1) + (double)__shfl_xor(ox0 + ox1, 32, 64);
        fy += (double)(sy0 + sy1) + (double)__shfl_xor(oy0 + oy1, 32, 64);
    } else {
        const double2 *xy = static_cast<const double2 *>(cur);
        const double2 *vv = xy + 32;
        double gx = 0, gy = 0;
        int j = 0;
        for (; j + 2 <= cnt; j += 2) {
#pragma clang fp contract(fast)
            const double2 p0 = xy[j], p1 = xy[j + 1];
            const double v0 = vv[j].x, v1 = vv[j + 1].x;
            const double dx0 = xt - p0.x, dy0 = yt - p0.y;
            const double dx1 = xt - p1.x, dy1 = yt - p1.y;
            const double f0 = v0 * rcp64(dx0 * dx0 + (dy0 * dy0 + EPS_REG));
            const double f1 = v1 * rcp64(dx1 * dx1 + (dy1 * dy1 + EPS_REG));
            fx += dx0 * f0;
            fy += dy0 * f0;
            gx += dx1 * f1;
            gy += dy1 * f1;
        }
        if (j < cnt) {
#pragma clang fp contract(fast)
            const double2 p = xy[j];
            const double v = vv[j].x;
            const double dx = xt - p.x, dy = yt - p.y;
            const double f = v * rcp64(dx * dx + (dy * dy + EPS_REG));
            fx += dx * f;
            fy += dy * f;
        }
        fx += gx;
        fy += gy;
    }
}

// P2P of the listed leaves at the brick's 64 targets (force_p2p_8x8,
// force-kernels.m4:15-55).  Sources pass through the two LDS halves of buf in
// tiles of TILE_N: the copy of the next tile (of this leaf or the next one)
// runs while the current one is summed.
template <bool F32>
TC_DEV void force_flush_p2p(const int *list, int n, const TcNode *__restrict__ nodes,
                            const double4 *__restrict__ src, const float4 *__restrict__ src32, double4 *buf,
                            double xt, double yt, double &fx, double &fy, unsigned long long &np2p)
{
    constexpr int T = P2P_TILE_N<F32>;
    const int lane = tc_lane();
    char *const base = reinterpret_cast<char *>(buf);
    for (int c0 = 0; c0 < n; c0 += 64) {
        const int cn = min(64, n - c0);
        double lx = 0, ly = 0;
        int ls = 0, le = 0;
        if (lane < cn) {
            const TcForceRec nd = force_load(nodes, list[c0 + lane]);
            lx = nd.xcom;
            ly = nd.ycom;
            ls = nd.s;
            le = nd.e;
        }
        int t = 0;
        int ns = __builtin_amdgcn_readlane(ls, 0), ne = __builtin_amdgcn_readlane(le, 0);
        double xc = readlane_f64(lx, 0), yc = readlane_f64(ly, 0);
        int b0 = ns, ph = 0;
        np2p += 64ull * (unsigned long long)(ne - ns);
        p2p_tile_load<F32>(src, src32, b0, ne, base);
        while (t < cn) {
            const int cnt = min(T, ne - b0);
            const double xcc = xc, ycc = yc;
            b0 += T;
            if (b0 >= ne) {
                ++t;
                if (t < cn) {
                    ns = __builtin_amdgcn_readlane(ls, t);
                    ne = __builtin_amdgcn_readlane(le, t);
                    xc = readlane_f64(lx, t);
                    yc = readlane_f64(ly, t);
                    b0 = ns;
                    np2p += 64ull * (unsigned long long)(ne - ns);
                }
            }
            vm_wait_all();  // the current tile has landed
            tc_wave_sync();
            p2p_tile_sum<F32>(base + ph * (FBUF_BYTES / 2), base + (ph ^ 1) * (FBUF_BYTES / 2), t < cn, b0, ne, src,
                              src32, cnt, xt, yt, xcc, ycc, fx, fy);
            tc_wave_sync();
            ph ^= 1;
        }
    }
}

// P2P of the listed leaves of brick `wave`: the brick's first FP2P_INLINE
// leaves (ninl counts them) are summed here, further ones go to chunks (see
// TcNearLists); prev is the brick's last chunk.
template <bool F32>
TC_DEV void force_near_p2p(const int *list, int n, const TcNearLists &nl, int wave, int &ninl, int &prev,
                           bool &ovfl, const TcNode *__restrict__ nodes, const double4 *__restrict__ src,
                           const float4 *__restrict__ src32, double4 *buf, double xt, double yt, double &fx,
                           double &fy, unsigned long long &np2p)
{
    if (ovfl)
        return;
    const int lane = tc_lane();
    const int take = min(n, nl.inl - ninl);
    if (take > 0)
        force_flush_p2p<F32>(list, take, nodes, src, src32, buf, xt, yt, fx, fy, np2p);
    ninl += take;
    for (int off = take; off < n; off += 64) {
        const int cnt = min(64, n - off);
        int c = 0;
        if (lane == 0)
            c = (int)atomicAdd(&nl.ctr[FN_CHUNKS], 1u);
        c = __builtin_amdgcn_readfirstlane(c);
        if (c >= nl.maxc) {
            ovfl = true;
            return;
        }
        if (lane < cnt)
            nl.cpool[(size_t)c * 64 + lane] = (uint32_t)list[off + lane];
        if (lane == 0) {
            nl.cdesc[c] = make_int4(wave, cnt, -1, 0);
            if (prev >= 0) {
                nl.cdesc[prev].z = c;
            } else {
                nl.ohead[wave] = c;
                nl.obricks[atomicAdd(&nl.ctr[FN_OBRICKS], 1u)] = (uint32_t)wave;
            }
        }
        prev = c;
    }
}

// P2P of the listed leaves after the first `skip` (which count down).
template <bool F32>
TC_DEV void force_skip_p2p(const int *list, int n, int &skip, const TcNode *__restrict__ nodes,
                           const double4 *__restrict__ src, const float4 *__restrict__ src32, double4 *buf,
                           double xt, double yt, double &fx, double &fy, unsigned long long &np2p)
{
    const int s0 = min(n, skip);
    skip -= s0;
    if (n > s0)
        force_flush_p2p<F32>(list + s0, n - s0, nodes, src, src32, buf, xt, yt, fx, fy, np2p);
}

// EvaluateForce::evaluate (treecode-force.cpp:94-199) with one wavefront per
// 8x8 brick (lane = target).  Bricks are dealt out one at a time from a device
// counter and each wave runs until the queue is empty: brick costs differ by
// orders of magnitude across block sizes, and a workgroup's LDS is only
// released when all four of its waves are done, so fixed brick-per-wave
// launches left SIMDs idle.  The brick's interaction lists are the
// reference's (same decisions per node); they are evaluated in list order
// rather than depth-first order.  With NEAR = false the pass computes the
// far field (E2L, E2P, L2P) and the first FP2P_INLINE P2P leaves, writes the
// outputs and hands further P2P leaves to k_force_chunks (see TcNearLists).
// With NEAR = true the waves take the bricks whose leaves did not fit there,
// repeat the same traversal and add all their P2P sums.
// CNT: the instance launched with counters, which also times the categories
template <int P, bool F32, bool NEAR, bool CNT>
__global__ __launch_bounds__(256, 4) void k_force(const TcNode *__restrict__ nodes, const double *__restrict__ exps,
                                               const float2 *__restrict__ exps32, const double4 *__restrict__ src,
                                               const float4 *__restrict__ src32, const double *__restrict__ x0s,
                                               const double *__restrict__ y0s, const double *__restrict__ hs,
                                               int first, int nbricks, double theta,
                                               double *__restrict__ xres, double *__restrict__ yres,
                                               unsigned long long *counters, uint32_t *__restrict__ work,
                                               TcNearLists nlist)
{
    __shared__ int s_stack[4][FS_CAP];
    __shared__ int s_list[4][3][FQ_CAP];  // E2L, E2P, P2P node ids
    __shared__ double4 s_buf[4][FBUF_BYTES / 32];
    __shared__ double2 s_coef[4][72];  // E2L_NQ groups of E2L_G + 1 entries (at most 8 x 9); L2P uses [1..P]
    __shared__ double s_binom[P * P];  // [k - 1][L - 1] = C(L + k - 1, k - 1)
    __shared__ double s_ninv[TC_MAXORDER + 1];  // -1/L, nearest double

    if constexpr (!NEAR) {
        for (int i = threadIdx.x; i < P * P; i += 256)
            s_binom[i] = c_binomT[(i / P) * 64 + i % P + i / P + 1];
        if (threadIdx.x <= TC_MAXORDER)
            s_ninv[threadIdx.x] = threadIdx.x ? -(1.0 / (double)threadIdx.x) : 0.0;
        __syncthreads();
    }
    // NEAR: the bricks listed in nlist.ovf, dealt out by position
    const uint32_t nwork = NEAR ? nlist.ctr[FN_OVF] : (uint32_t)nbricks;

    const int wid = threadIdx.x >> 6, lane = tc_lane();
    int *stack = s_stack[wid];
    int *lst_l = s_list[wid][0], *lst_e = s_list[wid][1], *lst_p = s_list[wid][2];
    unsigned long long np2p = 0, ne2p = 0, ne2l = 0, nl2p = 0;
    // with counters: wave clock cycles per category and in total
    constexpr bool timing = CNT;
    unsigned long long cp2p = 0, ce2p = 0, cloc = 0, cl2p = 0, tc0 = 0;
    const unsigned long long cyc0 = timing ? tc_clock() : 0ull;
    for (uint32_t iter = 0;; ++iter) {
        uint32_t grab = 0;
        if (NEAR)
            grab = (blockIdx.x + iter * gridDim.x) * 4 + wid;
        else if (lane == 0)
            grab = atomicAdd(work, 1u);
        const uint32_t wave = (uint32_t)__builtin_amdgcn_readfirstlane((int)grab);
        if (wave >= nwork)
            break;
        const int gbrick = NEAR ? (int)nlist.ovf[wave] : first + (int)wave;
        const int b = gbrick >> 4, brick = gbrick & 15;
        const int by = (brick >> 2) * TILE, bx = (brick & 3) * TILE;
        const double x0 = x0s[b], y0 = y0s[b], h = hs[b];
        const int ix = lane & 7, iy = lane >> 3;
        int ninl = 0, prev = -1;
        bool ovfl = false;
        // NEAR: the far-field pass summed the brick's first FP2P_INLINE P2P
        // leaves; the repeated traversal lists the leaves in the same order
        int skip = nlist.inl;
        // targets as in force_p2p_8x8 (force-kernels.m4:28-32); the brick-wide
        // values are moved to scalar registers
        const double xb = tc_uniform(x0 + (bx + 0) * h), yb = tc_uniform(y0 + (by + 0) * h);
        const double xt = xb + ix * h, yt = yb + iy * h;
        // brick geometry (treecode-force.cpp:108,113-114,146-147)
        TcBrick bk;
        bk.x0b = tc_uniform(x0 + h * (bx + 0.5 * (TILE - 1)));
        bk.y0b = tc_uniform(y0 + h * (by + 0.5 * (TILE - 1)));
        bk.rbrick = tc_uniform(1.4142135623730951 * h * (TILE - 1) * 0.5);
        bk.xlo = tc_uniform(x0 + bx * h);
        bk.xhi = tc_uniform(x0 + (bx + TILE - 1) * h);
        bk.ylo = tc_uniform(y0 + by * h);
        bk.yhi = tc_uniform(y0 + (by + TILE - 1) * h);
        bk.invtheta = tc_uniform(1 / theta);
        bk.th2 = tc_uniform(theta * theta);

        double fx = 0, fy = 0;
        double lre = 0, lim = 0;
        int sp = 1, nl = 0, ne = 0, np = 0;
        if (lane == 0)
            stack[0] = 0;
        tc_wave_sync();
        while (sp > 0) {
            // Pop k ids.  Opening all of them pushes at most 3k more, so k is
            // capped to keep FS_RESERVE entries free; with one pop per step the
            // stack grows by at most 3 per level below the popped node, which
            // the reserve covers, so the stack never overflows.
            int k = (FS_CAP - FS_RESERVE - sp) / 3;
            k = k < 1 ? 1 : (k > 64 ? 64 : k);
            k = k < sp ? k : sp;
            sp -= k;
            int act = ACT_SKIP, id = 0, child = 0;
            if (lane < k) {
                id = stack[sp + lane];
                const TcForceRec nd = force_load(nodes, id);
                act = force_action(nd, bk);
                child = nd.child;
            }
            const unsigned long long lt = tc_lanemask_lt();
            const unsigned long long mo = tc_ballot(act == ACT_OPEN), ml = tc_ballot(act == ACT_E2L);
            const unsigned long long me = tc_ballot(act == ACT_E2P), mp = tc_ballot(act == ACT_P2P);
            tc_wave_sync();
            if (act == ACT_OPEN) {
                const int pos = sp + 4 * __popcll(mo & lt);
#pragma unroll
                for (int c = 0; c < 4; ++c)
                    stack[pos + c] = child + c;
            } else if (act == ACT_E2L && !NEAR) {
                lst_l[nl + __popcll(ml & lt)] = id;
            } else if (act == ACT_E2P && !NEAR) {
                lst_e[ne + __popcll(me & lt)] = id;
            } else if (act == ACT_P2P) {
                lst_p[np + __popcll(mp & lt)] = id;
            }
            sp += 4 * __popcll(mo);
            if constexpr (!NEAR) {
                nl += __popcll(ml);
                ne += __popcll(me);
            }
            np += __popcll(mp);
            tc_wave_sync();
            // a list that might not take another 64 ids is flushed now
            if (nl > FQ_CAP - 64) {
                ne2l += nl;
                if (timing)
                    tc0 = tc_clock();
                force_flush_e2l<P, F32>(lst_l, nl, nodes, exps, s_buf[wid], bk.x0b, bk.y0b, s_binom, s_ninv,
                                        s_coef[wid], lre, lim);
                if (timing)
                    cloc += tc_clock() - tc0;
                nl = 0;
            }
            if (ne > FQ_CAP - 64) {
                ne2p += 64ull * ne;
                if (timing)
                    tc0 = tc_clock();
                force_flush_e2p<P, F32>(lst_e, ne, nodes, exps, exps32, s_buf[wid], xb, yb, ix, iy, h, fx, fy);
                if (timing)
                    ce2p += tc_clock() - tc0;
                ne = 0;
            }
            if (np > FQ_CAP - 64) {
                if (timing)
                    tc0 = tc_clock();
                if constexpr (NEAR)
                    force_skip_p2p<F32>(lst_p, np, skip, nodes, src, src32, s_buf[wid], xt, yt, fx, fy, np2p);
                else
                    force_near_p2p<F32>(lst_p, np, nlist, (int)wave, ninl, prev, ovfl, nodes, src, src32,
                                        s_buf[wid], xt, yt, fx, fy, np2p);
                if (timing)
                    cp2p += tc_clock() - tc0;
                np = 0;
            }
        }
        if constexpr (NEAR) {
            if (timing)
                tc0 = tc_clock();
            if (np)
                force_skip_p2p<F32>(lst_p, np, skip, nodes, src, src32, s_buf[wid], xt, yt, fx, fy, np2p);
            if (timing)
                cp2p += tc_clock() - tc0;
            const size_t o = (size_t)b * 1024 + (size_t)(bx + ix) + 32 * (size_t)(by + iy);
            xres[o] += fx;
            yres[o] += fy;
            tc_wave_sync();
            continue;
        }
        if (timing)
            tc0 = tc_clock();
        if (np)
            force_near_p2p<F32>(lst_p, np, nlist, (int)wave, ninl, prev, ovfl, nodes, src, src32, s_buf[wid], xt,
                                yt, fx, fy, np2p);
        if (timing) {
            const unsigned long long t1 = tc_clock();
            cp2p += t1 - tc0;
            tc0 = t1;
        }
        if (lane == 0) {
            nlist.pmark[wave] = ovfl ? 1u : 0u;
            if (ovfl)
                nlist.ovf[atomicAdd(&nlist.ctr[FN_OVF], 1u)] = (uint32_t)gbrick;
        }
        if (ne) {
            ne2p += 64ull * ne;
            if (timing)
                tc0 = tc_clock();
            force_flush_e2p<P, F32>(lst_e, ne, nodes, exps, exps32, s_buf[wid], xb, yb, ix, iy, h, fx, fy);
            if (timing)
                ce2p += tc_clock() - tc0;
        }
        if (timing)
            tc0 = tc_clock();
        if (nl) {
            ne2l += nl;
            force_flush_e2l<P, F32>(lst_l, nl, nodes, exps, s_buf[wid], bk.x0b, bk.y0b, s_binom, s_ninv,
                                    s_coef[wid], lre, lim);
        }
        // fold the groups' partial local expansions into lanes 0..E2L_G-1
#pragma unroll
        for (int off = 32; off >= E2L_G<P>; off >>= 1) {
            lre += __shfl_xor(lre, off, 64);
            lim += __shfl_xor(lim, off, 64);
        }
        if (timing) {
            const unsigned long long t1 = tc_clock();
            cloc += t1 - tc0;
            tc0 = t1;
        }
        // downward_l2p_8x8 (force-kernels.m4:157-191) about the brick centre
        if (lane < P)
            s_coef[wid][lane + 1] = make_double2(lre, lim);
        tc_wave_sync();
        {
#pragma clang fp contract(fast)
            const double zl = h * (0 - 0.5 * (TILE - 1));
            const double zr1 = zl + ix * h, zi1 = zl + iy * h;
            const double2 b1 = s_coef[wid][1];
            double rres = b1.x, ires = b1.y;
            double zr = zr1, zi = zi1;
            for (int l = 2; l <= P; ++l) {
                const double2 bl = s_coef[wid][l];
                rres += l * (bl.x * zr - bl.y * zi);
                ires += l * (bl.x * zi + bl.y * zr);
                const double rn = zr * zr1 - zi * zi1;
                const double in = zr * zi1 + zi * zr1;
                zr = rn;
                zi = in;
            }
            fx += rres;
            fy -= ires;
        }
        const size_t o = (size_t)b * 1024 + (size_t)(bx + ix) + 32 * (size_t)(by + iy);
        xres[o] = fx;
        yres[o] = fy;
        nl2p += 64;
        if (timing)
            cl2p += tc_clock() - tc0;
        tc_wave_sync();
    }
    if (counters && lane == 0) {
        atomicAdd(&counters[TC_CNT_P2P], np2p);
        atomicAdd(&counters[TC_CNT_E2P], ne2p);
        atomicAdd(&counters[TC_CNT_E2L], ne2l);
        atomicAdd(&counters[TC_CNT_L2P], nl2p);
        atomicAdd(&counters[TC_CNT_P2P_SLOTS], np2p);
        atomicAdd(&counters[TC_CNT_E2P_SLOTS], ne2p);
        atomicAdd(&counters[TC_CNT_CYC_P2P], cp2p);
        atomicAdd(&counters[TC_CNT_CYC_E2P], ce2p);
        atomicAdd(&counters[TC_CNT_CYC_LOCAL], cloc);
        atomicAdd(&counters[TC_CNT_CYC_L2P], cl2p);
        atomicAdd(&counters[TC_CNT_CYC_TOTAL], tc_clock() - cyc0);
    }
}

// P2P sums of the near-leaf chunks (see TcNearLists), one wave per chunk,
// into cpart; k_force_combine adds them to the outputs.
template <bool F32>
__global__ __launch_bounds__(256) void k_force_chunks(const TcNode *__restrict__ nodes,
                                                      const double4 *__restrict__ src,
                                                      const float4 *__restrict__ src32,
                                                      const double *__restrict__ x0s, const double *__restrict__ y0s,
                                                      const double *__restrict__ hs, int first, TcNearLists nlist,
                                                      unsigned long long *counters)
{
    __shared__ double4 s_buf[4][FBUF_BYTES / 32];
    const int wid = threadIdx.x >> 6, lane = tc_lane();
    const int nch = min((int)nlist.ctr[FN_CHUNKS], nlist.maxc);
    unsigned long long np2p = 0;
    const unsigned long long cyc0 = counters ? tc_clock() : 0ull;
    for (int c = blockIdx.x * 4 + wid; c < nch; c += gridDim.x * 4) {
        const int4 d = nlist.cdesc[c];
        if (nlist.pmark[__builtin_amdgcn_readfirstlane(d.x)])
            continue;  // summed by k_force<P, F32, true, *>
        const int gbrick = first + __builtin_amdgcn_readfirstlane(d.x);
        const int b = gbrick >> 4, brick = gbrick & 15;
        const int by = (brick >> 2) * TILE, bx = (brick & 3) * TILE;
        const double x0 = x0s[b], y0 = y0s[b], h = hs[b];
        const int ix = lane & 7, iy = lane >> 3;
        const double xb = tc_uniform(x0 + (bx + 0) * h), yb = tc_uniform(y0 + (by + 0) * h);
        const double xt = xb + ix * h, yt = yb + iy * h;
        double fx = 0, fy = 0;
        force_flush_p2p<F32>(reinterpret_cast<const int *>(nlist.cpool + (size_t)c * 64),
                             __builtin_amdgcn_readfirstlane(d.y), nodes, src, src32, s_buf[wid], xt, yt, fx, fy,
                             np2p);
        nlist.cpart[(size_t)c * 64 + lane] = make_double2(fx, fy);
    }
    if (counters && lane == 0) {
        // a chunk wave does direct sums only
        const unsigned long long cyc = tc_clock() - cyc0;
        atomicAdd(&counters[TC_CNT_P2P], np2p);
        atomicAdd(&counters[TC_CNT_P2P_SLOTS], np2p);
        atomicAdd(&counters[TC_CNT_CYC_P2P], cyc);
        atomicAdd(&counters[TC_CNT_CYC_TOTAL], cyc);
    }
}

// Adds each brick's chunk sums to its outputs, in the order the traversal
// produced the chunks.
__global__ __launch_bounds__(256) void k_force_combine(const double *__restrict__ x0s, const double *__restrict__ y0s,
                                                       const double *__restrict__ hs, int first, TcNearLists nlist,
                                                       double *__restrict__ xres, double *__restrict__ yres)
{
    const int lane = tc_lane();
    const int nob = (int)nlist.ctr[FN_OBRICKS];
    for (int i = blockIdx.x * 4 + (int)(threadIdx.x >> 6); i < nob; i += gridDim.x * 4) {
        const int wave = (int)nlist.obricks[i];
        if (nlist.pmark[wave])
            continue;
        double sx = 0, sy = 0;
        for (int c = nlist.ohead[wave]; c >= 0; c = nlist.cdesc[c].z) {
            const double2 pc = nlist.cpart[(size_t)c * 64 + lane];
            sx += pc.x;
            sy += pc.y;
        }
        const int gbrick = first + wave;
        const int b = gbrick >> 4, brick = gbrick & 15;
        const int by = (brick >> 2) * TILE, bx = (brick & 3) * TILE;
        const int ix = lane & 7, iy = lane >> 3;
        const size_t o = (size_t)b * 1024 + (size_t)(bx + ix) + 32 * (size_t)(by + iy);
        xres[o] += sx;
        yres[o] += sy;
    }
}

// Force at arbitrary targets (no counterpart in the reference: main.cpp:191-195
// rejects non-block targets for velocities).  Same masked wave traversal as
// k_potential -- per-target accept test r^2 < theta^2 d^2 -- with the force
// E2P (force-kernels.m4:57-98 per target) and the force P2P pair term
// (force-kernels.m4:46-52).
template <int P, bool F32>
__global__ __launch_bounds__(256) void k_force_targets(const TcNode *__restrict__ nodes,
                                                       const double *__restrict__ exps,
                                                       const float2 *__restrict__ exps32,
                                                       const double4 *__restrict__ src,
                                                       const float4 *__restrict__ src32,
                                                       const double *__restrict__ xt, const double *__restrict__ yt,
                                                       const uint32_t *__restrict__ tperm, int m, double th2,
                                                       double *__restrict__ xres, double *__restrict__ yres,
                                                       unsigned long long *counters)
{
    __shared__ int s_node[4][TC_STACK];
    __shared__ unsigned long long s_mask[4][TC_STACK];
    __shared__ double4 s_src[4][64];
    const int wid = threadIdx.x >> 6, lane = tc_lane();
    const int k = blockIdx.x * 256 + threadIdx.x;
    const bool valid = k < m;
    const uint32_t ti = valid ? tperm[k] : 0u;
    const double x = valid ? xt[ti] : 0.0, y = valid ? yt[ti] : 0.0;
    const unsigned long long live = tc_ballot(valid);
    if (!live)
        return;
    double fx = 0, fy = 0;
    unsigned long long np2p = 0, ne2p = 0, nslots = 0, neslots = 0;
    s_node[wid][0] = 0;
    s_mask[wid][0] = live;
    int sp = 1;
    while (sp > 0) {
        --sp;
        tc_wave_sync();
        const int id = tc_uniform(s_node[wid][sp]);
        const unsigned long long mask = tc_uniform64(s_mask[wid][sp]);
        // no early exit for empty nodes: all fields of the record are needed
        // before any branch, so they arrive in one round of scalar loads (an
        // empty node leaves no lane accepting or opening it)
        const TcNode nd = nodes[id];
        const bool nonempty = nd.s != nd.e;
        const bool in = ((mask >> lane) & 1ull) && nonempty;
        const double dx = x - nd.xcom, dy = y - nd.ycom;
        const double r2 = dx * dx + dy * dy;
        const bool acc_ok = in && (nd.r * nd.r < th2 * r2);
        const unsigned long long amask = tc_ballot(acc_ok);
        if (acc_ok) {
            if (F32) {
                float gx, gy;
                e2p_force32<P>((float)dx, (float)dy, exps32 + (size_t)(P + 1) * id, gx, gy);
                fx += (double)gx;
                fy += (double)gy;
            } else {
                e2p_force64<P>(dx, dy, nd.mass, exps + (size_t)2 * P * id, fx, fy);
            }
        }
        ne2p += __popcll(amask);
        neslots += amask ? 64ull : 0ull;
        const unsigned long long omask = nonempty ? mask & ~amask : 0ull;
        if (!omask)
            continue;
        if (nd.child == 0) {
            const bool act = (omask >> lane) & 1ull;
            np2p += (unsigned long long)__popcll(omask) * (unsigned long long)(nd.e - nd.s);
            nslots += 64ull * (unsigned long long)(nd.e - nd.s);
            if (F32) {
                const float tx = (float)(x - nd.xcom), ty = (float)(y - nd.ycom);
                float4 *tile = reinterpret_cast<float4 *>(&s_src[wid][0]);
                for (int b0 = nd.s; b0 < nd.e; b0 += 64) {
                    const int cnt = min(64, nd.e - b0);
                    if (lane < cnt)
                        tile[lane] = src32[b0 + lane];
                    tc_wave_sync();
                    if (act) {
#pragma clang fp contract(fast)
                        float ax32 = 0.f, ay32 = 0.f;
                        for (int j = 0; j < cnt; ++j) {
                            const float4 q = tile[j];
                            const float ddx = tx - q.x, ddy = ty - q.y;
                            const float f = q.z * fast_rcp(ddx * ddx + (ddy * ddy + EPS_REG32));
                            ax32 += ddx * f;
                            ay32 += ddy * f;
                        }
                        fx += (double)ax32;
                        fy += (double)ay32;
                    }
                    tc_wave_sync();
                }
            } else {
                for (int b0 = nd.s; b0 < nd.e; b0 += 64) {
                    const int cnt = min(64, nd.e - b0);
                    if (lane < cnt)
                        s_src[wid][lane] = src[b0 + lane];
                    tc_wave_sync();
                    if (act) {
#pragma clang fp contract(fast)
                        for (int j = 0; j < cnt; ++j) {
                            const double4 q = s_src[wid][j];
                            const double ddx = x - q.x, ddy = y - q.y;
                            const double f = q.z * rcp64(ddx * ddx + (ddy * ddy + EPS_REG));
                            fx += ddx * f;
                            fy += ddy * f;
                        }
                    }
                    tc_wave_sync();
                }
            }
        } else {
            if (lane < 4) {
                s_node[wid][sp + lane] = nd.child + lane;
                s_mask[wid][sp + lane] = omask;
            }
            sp += 4;
        }
    }
    if (valid) {
        xres[ti] = fx;
        yres[ti] = fy;
    }
    if (counters && lane == 0) {
        atomicAdd(&counters[TC_CNT_P2P], np2p);
        atomicAdd(&counters[TC_CNT_E2P], ne2p);
        atomicAdd(&counters[TC_CNT_P2P_SLOTS], nslots);
        atomicAdd(&counters[TC_CNT_E2P_SLOTS], neslots);
    }
}

// ------------------------------------------------------------------ direct sums
// The FP64 O(N) sums of main.cpp:206-245 at selected targets, one block per
// target: potential sum_j v_j log(sqrt(r^2 + 10 eps)), force sum_j v_j d/(r^2 + 10 eps).
__global__ __launch_bounds__(256) void k_direct(const double *__restrict__ xs, const double *__restrict__ ys,
                                                const double *__restrict__ vs, int n,
                                                const double *__restrict__ xt, const double *__restrict__ yt,
                                                const int *__restrict__ idx, int force, double *__restrict__ out0,
                                                double *__restrict__ out1)
{
    __shared__ double s0[4], s1[4];
    const int k = blockIdx.x;
    const int t = idx[k];
    const double xd = xt[t], yd = yt[t];
    double a = 0, b = 0;
    if (force) {
        for (int j = threadIdx.x; j < n; j += 256) {
            const double xr = xd - xs[j], yr = yd - ys[j];
            const double f = vs[j] / (xr * xr + yr * yr + EPS_REG);
            a += xr * f;
            b += yr * f;
        }
    } else {
        for (int j = threadIdx.x; j < n; j += 256) {
            const double xr = xd - xs[j], yr = yd - ys[j];
            a += log(sqrt(xr * xr + yr * yr + EPS_REG)) * vs[j];
        }
    }
    a = tc_wave_tree_sum(a);
    b = tc_wave_tree_sum(b);
    if (tc_lane() == 0) {
        s0[threadIdx.x >> 6] = a;
        s1[threadIdx.x >> 6] = b;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        out0[k] = (s0[0] + s0[1]) + (s0[2] + s0[3]);
        if (force)
            out1[k] = (s1[0] + s1[1]) + (s1[2] + s1[3]);
    }
}

// Plain 16-byte-per-thread copy: the reference kernel for scaling HBM counters
// (FETCH_SIZE / WRITE_SIZE of a copy of known size).
__global__ __launch_bounds__(256) void k_copy16(const uint4 *__restrict__ src, uint4 *__restrict__ dst, size_t n)
{
    for (size_t i = blockIdx.x * (size_t)256 + threadIdx.x; i < n; i += (size_t)gridDim.x * 256)
        dst[i] = src[i];
}

int tiles_of(int n, int t) { return (n + t - 1) / t; }

TcPotReq potential_req(const TcTree &t, double th2, const double *xt, const double *yt, const uint32_t *tperm,
                       int m, double *out, int precision, int wave0)
{
    TcPotReq rq;
    rq.nodes = t.nodes.as<TcNode>();
    rq.exps = t.exps.as<double>();
    rq.exps32 = precision == 1 ? t.exps32.as<float2>() : nullptr;
    rq.src = t.src.as<double4>();
    rq.src32 = precision == 1 ? t.src32.as<float4>() : nullptr;
    rq.xt = xt;
    rq.yt = yt;
    rq.tperm = tperm;
    rq.out = out;
    rq.th2 = th2;
    rq.m = m;
    rq.wave0 = wave0;
    return rq;
}

template <int P>
void launch_potential(const TcTree &t, double th2, const double *xt, const double *yt, const uint32_t *tperm,
                      int m, double *out, int precision, unsigned long long *counters, hipStream_t st,
                      int xcd_lc = TC_XCD_CHUNK_LOG2)
{
    const int grid = tiles_of(m, 256);
    const TcPotReq q = potential_req(t, th2, xt, yt, tperm, m, out, precision, 0);
    auto go = [&](auto kern) {
        kern<<<grid, 256, 0, st>>>(q.nodes, q.exps, q.exps32, q.src, q.src32, q.xt, q.yt, q.tperm, q.m, q.th2, q.out,
                                   counters, xcd_lc);
    };
    if (precision == 1 && counters)
        go(k_potential<P, true, true>);
    else if (precision == 1)
        go(k_potential<P, true, false>);
    else if (counters)
        go(k_potential<P, false, true>);
    else
        go(k_potential<P, false, false>);
    TC_HIP(hipGetLastError());
}

template <int P>
void launch_force_targets(const TcTree &t, double th2, const double *xt, const double *yt, const uint32_t *tperm,
                          int m, double *xr, double *yr, int precision, unsigned long long *counters,
                          hipStream_t st)
{
    const int grid = tiles_of(m, 256);
    if (precision == 1)
        k_force_targets<P, true><<<grid, 256, 0, st>>>(t.nodes.as<TcNode>(), t.exps.as<double>(),
                                                       t.exps32.as<float2>(), t.src.as<double4>(),
                                                       t.src32.as<float4>(), xt, yt, tperm, m, th2, xr, yr,
                                                       counters);
    else
        k_force_targets<P, false><<<grid, 256, 0, st>>>(t.nodes.as<TcNode>(), t.exps.as<double>(), nullptr,
                                                        t.src.as<double4>(), nullptr, xt, yt, tperm, m, th2, xr,
                                                        yr, counters);
    TC_HIP(hipGetLastError());
}

// Workgroups of a persistent kernel that fit the device at once (occupancy x
// CUs), cached per kernel and device.
template <auto Kernel> int resident_blocks(int device)
{
    static std::atomic<int> cache[64];
    if (device < 0 || device >= 64)
        throw TcError(-3, "device id out of range");
    int v = cache[device].load(std::memory_order_relaxed);
    if (!v) {
        int per_cu = 0, cus = 0;
        TC_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, Kernel, 256, 0));
        TC_HIP(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, device));
        if (per_cu < 1 || cus < 1)
            throw TcError(-3, "kernel does not fit on the device");
        v = per_cu * cus;
        cache[device].store(v, std::memory_order_relaxed);
    }
    return v;
}

template <int P, bool F32>
void launch_force_pass(const TcTree &t, double theta, const double *x0s, const double *y0s, const double *hs,
                       int first, int nb, double *xr, double *yr, unsigned long long *counters, uint32_t *work,
                       const TcNearLists &nl, int device, hipStream_t st)
{
    const TcNode *nodes = t.nodes.as<TcNode>();
    const double *exps = t.exps.as<double>();
    const float2 *exps32 = F32 ? t.exps32.as<float2>() : nullptr;
    const double4 *src = t.src.as<double4>();
    const float4 *src32 = F32 ? t.src32.as<float4>() : nullptr;
    tc_fill_u32(work, 0u, 1, st);
    tc_fill_u32(nl.ctr, 0u, FN_NCTR, st);
    if (counters) {
        const int grid = std::min(tiles_of(nb, 4), resident_blocks<k_force<P, F32, false, true>>(device));
        k_force<P, F32, false, true><<<grid, 256, 0, st>>>(nodes, exps, exps32, src, src32, x0s, y0s, hs, first, nb,
                                                           theta, xr, yr, counters, work, nl);
    } else {
        const int grid = std::min(tiles_of(nb, 4), resident_blocks<k_force<P, F32, false, false>>(device));
        k_force<P, F32, false, false><<<grid, 256, 0, st>>>(nodes, exps, exps32, src, src32, x0s, y0s, hs, first,
                                                            nb, theta, xr, yr, counters, work, nl);
    }
    // the chunk and overflow counts are on the device: persistent grids
    k_force_chunks<F32><<<512, 256, 0, st>>>(nodes, src, src32, x0s, y0s, hs, first, nl, counters);
    k_force_combine<<<64, 256, 0, st>>>(x0s, y0s, hs, first, nl, xr, yr);
    if (counters)
        k_force<P, F32, true, true><<<256, 256, 0, st>>>(nodes, exps, exps32, src, src32, x0s, y0s, hs, first, nb,
                                                         theta, xr, yr, counters, work, nl);
    else
        k_force<P, F32, true, false><<<256, 256, 0, st>>>(nodes, exps, exps32, src, src32, x0s, y0s, hs, first, nb,
                                                          theta, xr, yr, counters, work, nl);
    TC_HIP(hipGetLastError());
}

template <int P>
void launch_force(const TcTree &t, double theta, const double *x0s, const double *y0s, const double *hs,
                  int nblocks, double *xr, double *yr, int precision, unsigned long long *counters, TcScratch &ws,
                  int max_chunks, int inline_leaves, hipStream_t st)
{
    if (nblocks > (1 << 26))
        throw TcError(-2, "tc_eval_force: too many blocks");
    const int nbricks = nblocks * 16;
    int device = 0;
    TC_HIP(hipGetDevice(&device));
    // bricks per pass, which bounds the near-field hand-off buffers
    constexpr int CHUNK = 1 << 20;
    const int chunk = std::min(nbricks, CHUNK);
    const int maxc = max_chunks > 0 ? max_chunks : std::max(4096, chunk / 4);
    const size_t part_bytes = sizeof(double2) * 64 * (size_t)maxc;
    const size_t u32s = (size_t)chunk * 4 + (size_t)maxc * (64 + 4) + FN_NCTR;
    ws.work.reserve(256);
    ws.flist.reserve(part_bytes + sizeof(uint32_t) * u32s);
    TcNearLists nl;
    nl.cpart = ws.flist.as<double2>();
    uint32_t *u = reinterpret_cast<uint32_t *>(ws.flist.as<char>() + part_bytes);
    nl.cdesc = reinterpret_cast<int4 *>(u);  // 16-byte aligned: part_bytes is a multiple of 1024
    u += (size_t)maxc * 4;
    nl.cpool = u;
    u += (size_t)maxc * 64;
    nl.pmark = u;
    u += chunk;
    nl.ohead = reinterpret_cast<int *>(u);
    u += chunk;
    nl.obricks = u;
    u += chunk;
    nl.ovf = u;
    u += chunk;
    nl.ctr = u;
    nl.maxc = maxc;
    nl.inl = inline_leaves > 0 ? inline_leaves : FP2P_INLINE;
    uint32_t *work = ws.work.as<uint32_t>();
    for (int first = 0; first < nbricks; first += chunk) {
        const int nb = std::min(chunk, nbricks - first);
        if (precision == 1)
            launch_force_pass<P, true>(t, theta, x0s, y0s, hs, first, nb, xr, yr, counters, work, nl, device, st);
        else
            launch_force_pass<P, false>(t, theta, x0s, y0s, hs, first, nb, xr, yr, counters, work, nl, device, st);
    }
}

}  // namespace

void tc_upload_binomials_eval(hipStream_t st)
{
    std::vector<double> tab(TC_MAXORDER * 64, 0.0);
    for (int k = 0; k < TC_MAXORDER; ++k)
        for (int n = k; n < 64; ++n) {
            unsigned __int128 r = 1;
            const int kk = k < n - k ? k : n - k;
            for (int i = 1; i <= kk; ++i)
                r = r * (unsigned __int128)(n - kk + i) / (unsigned __int128)i;
            tab[k * 64 + n] = (double)r;
        }
    TC_HIP(hipMemcpyToSymbolAsync(HIP_SYMBOL(c_binomT), tab.data(), sizeof(double) * tab.size(), 0,
                                  hipMemcpyHostToDevice, st));
    TC_HIP(hipStreamSynchronize(st));
}

void tc_copy16(const void *src, void *dst, size_t bytes, hipStream_t st)
{
    const size_t n = bytes / 16;
    if (!n)
        return;
    k_copy16<<<2048, 256, 0, st>>>(static_cast<const uint4 *>(src), static_cast<uint4 *>(dst), n);
    TC_HIP(hipGetLastError());
}

void tc_direct_sums(const double *xs, const double *ys, const double *vs, int n, const double *xt,
                    const double *yt, const int *idx, int nidx, int force, double *out0, double *out1,
                    hipStream_t st)
{
    if (nidx <= 0)
        return;
    k_direct<<<nidx, 256, 0, st>>>(xs, ys, vs, n, xt, yt, idx, force, out0, out1);
    TC_HIP(hipGetLastError());
}

// Device layout of a batch (one descriptor block): TcPotReq[MAXREQ], meta {k, waves, work counter}.
constexpr size_t BATCH_DESC = sizeof(TcPotReq) * TcBatchWs::MAXREQ;
constexpr size_t BATCH_HOST = BATCH_DESC + 256;  // one block, a multiple of 256 bytes

TcPotReq *batch_desc(TcBatchWs &ws, int slot) { return reinterpret_cast<TcPotReq *>(ws.dev.as<char>() + (size_t)slot * BATCH_HOST); }
int *batch_meta(TcBatchWs &ws, int slot) { return reinterpret_cast<int *>(ws.dev.as<char>() + (size_t)slot * BATCH_HOST + BATCH_DESC); }

// persistent grid of a batch launch: the bucket's waves in workgroups of
// four, at most as many workgroups as are resident at once
template <int P, bool F32, bool CNT> int batch_grid(int bucket)
{
    int device = 0;
    TC_HIP(hipGetDevice(&device));
    return std::min((1 << bucket) / 4, resident_blocks<k_potential_batch<P, F32, CNT>>(device));
}

template <int P, bool F32, bool CNT>
void batch_kernel(TcBatchWs &ws, int slot, unsigned long long *counters, int bucket, hipStream_t st)
{
    k_potential_batch<P, F32, CNT><<<batch_grid<P, F32, CNT>(bucket), 256, 0, st>>>(batch_desc(ws, slot),
                                                                                 batch_meta(ws, slot), counters);
    TC_HIP(hipGetLastError());
}

// The graph of one (order, precision, size bucket, descriptor block): the
// copy of that block from pinned host memory, then the launch over the
// bucket's grid.  Built node by node rather than by stream capture, so
// building it never restricts what other threads may call meanwhile.
template <int P, bool F32>
hipGraphExec_t batch_graph(TcBatchWs &ws, int slot, int bucket)
{
    const int grid = batch_grid<P, F32, false>(bucket);
    TcPotReq *desc = batch_desc(ws, slot);
    int *meta = batch_meta(ws, slot);
    unsigned long long *counters = nullptr;
    const char *host = static_cast<const char *>(ws.host) + (size_t)slot * BATCH_HOST;
    hipGraph_t g = nullptr;
    TC_HIP(hipGraphCreate(&g, 0));
    hipGraphExec_t ex = nullptr;
    try {
        hipGraphNode_t cp = nullptr, kn = nullptr;
        TC_HIP(hipGraphAddMemcpyNode1D(&cp, g, nullptr, 0, desc, host, BATCH_HOST, hipMemcpyHostToDevice));
        void *args[] = {&desc, &meta, &counters};
        hipKernelNodeParams kp{};
        kp.func = reinterpret_cast<void *>(k_potential_batch<P, F32, false>);
        kp.gridDim = dim3(grid);
        kp.blockDim = dim3(256);
        kp.sharedMemBytes = 0;
        kp.kernelParams = args;
        kp.extra = nullptr;
        TC_HIP(hipGraphAddKernelNode(&kn, g, &cp, 1, &kp));
        TC_HIP(hipGraphInstantiate(&ex, g, nullptr, nullptr, 0));
    } catch (...) {
        (void)hipGraphDestroy(g);
        throw;
    }
    (void)hipGraphDestroy(g);
    return ex;
}

// size bucket of a batch: its waves rounded up to a power of two (at least 64 waves)
int batch_bucket(int waves)
{
    int b = 6;
    while ((1 << b) < waves)
        ++b;
    return b;
}

template <int P>
void batch_launch(TcBatchWs &ws, int slot, int precision, int waves, unsigned long long *counters, hipStream_t st,
                  bool graph)
{
    const bool f32 = precision == 1;
    const int bucket = batch_bucket(waves);
    if (counters || !graph) {
        TC_HIP(hipMemcpyAsync(batch_desc(ws, slot), static_cast<char *>(ws.host) + (size_t)slot * BATCH_HOST,
                              BATCH_HOST, hipMemcpyHostToDevice, st));
        if (f32 && counters)
            batch_kernel<P, true, true>(ws, slot, counters, bucket, st);
        else if (f32)
            batch_kernel<P, true, false>(ws, slot, nullptr, bucket, st);
        else if (counters)
            batch_kernel<P, false, true>(ws, slot, counters, bucket, st);
        else
            batch_kernel<P, false, false>(ws, slot, nullptr, bucket, st);
        return;
    }
    hipGraphExec_t &ex = ws.graph[(2 * P + (f32 ? 1 : 0)) * TcBatchWs::NBUCKET + bucket - 6][slot];
    if (!ex)
        ex = f32 ? batch_graph<P, true>(ws, slot, bucket) : batch_graph<P, false>(ws, slot, bucket);
    TC_HIP(hipGraphLaunch(ex, st));
}

void TcBatchWs::release()
{
    for (auto &row : graph)
        for (auto &g : row)
            if (g) {
                (void)hipGraphExecDestroy(g);
                g = nullptr;
            }
    for (int i = 0; i < RING; ++i) {
        if (done[i]) {
            if (pending[i])
                (void)hipEventSynchronize(done[i]);
            (void)hipEventDestroy(done[i]);
            done[i] = nullptr;
        }
        pending[i] = false;
    }
    if (host)
        (void)hipHostFree(host);
    host = nullptr;
    dev.release();
}

void tc_potential_batch(const TcPotBatchItem *items, int k, int order, int precision, TcBatchWs &ws,
                        unsigned long long *counters, hipStream_t st, bool graph)
{
    if (k <= 0)
        return;
    if (k > TcBatchWs::MAXREQ)
        throw TcError(-3, "tc_potential_batch: at most 64 requests per batch");
    if (!tc_order_supported(order))
        throw TcError(-3, "tc_potential_batch: unsupported order");
    for (int i = 0; i < k; ++i) {
        const TcTree &t = *items[i].tree;
        if (t.order != order || (precision == 1 && !t.has_fp32))
            throw TcError(-3, "tc_potential_batch: a tree of another order or without FP32 mirrors");
    }
    // the largest size bucket bounds a batch's waves (2^25 waves, 2^31 targets)
    long long total = 0;
    for (int i = 0; i < k; ++i)
        total += (items[i].m + 63) / 64;
    if (total > (1ll << (TcBatchWs::NBUCKET + 5)))
        throw TcError(-3, "tc_potential_batch: more targets than one launch takes");
    if (!ws.host) {
        TC_HIP(hipHostMalloc(&ws.host, BATCH_HOST * TcBatchWs::RING, hipHostMallocDefault));
        for (auto &e : ws.done)
            TC_HIP(hipEventCreateWithFlags(&e, hipEventDisableTiming));
        ws.dev.reserve(BATCH_HOST * TcBatchWs::RING);
    }
    // a descriptor block is written again only after the launch that read it
    // has finished (launches may run on different streams)
    const int slot = ws.next;
    ws.next = (ws.next + 1) % TcBatchWs::RING;
    if (ws.pending[slot]) {
        TC_HIP(hipEventSynchronize(ws.done[slot]));
        ws.pending[slot] = false;
    }
    char *host = static_cast<char *>(ws.host) + (size_t)slot * BATCH_HOST;
    TcPotReq *desc = reinterpret_cast<TcPotReq *>(host);
    int waves = 0;
    for (int i = 0; i < k; ++i) {
        const TcPotBatchItem &it = items[i];
        desc[i] = potential_req(*it.tree, it.theta * it.theta, it.xt, it.yt, it.tperm, it.m, it.out, precision,
                                waves);
        waves += (it.m + 63) / 64;
    }
    int *meta = reinterpret_cast<int *>(host + BATCH_DESC);
    meta[0] = k;
    meta[1] = waves;
    meta[2] = 0;  // work counter of the persistent grid
    switch (order) {
#define TC_CASE(P) \
    case P: batch_launch<P>(ws, slot, precision, waves, counters, st, graph); break;
        TC_ORDERS(TC_CASE)
#undef TC_CASE
    }
    TC_HIP(hipEventRecord(ws.done[slot], st));
    ws.pending[slot] = true;
}

void tc_potential_item(const TcPotBatchItem &it, int precision, unsigned long long *counters, hipStream_t st)
{
    if (it.m <= 0)
        return;
    const TcTree &t = *it.tree;
    if (precision == 1 && !t.has_fp32)
        throw TcError(-3, "tc_potential_item: tree was built without FP32 mirrors");
    switch (t.order) {
#define TC_CASE(P) \
    case P: launch_potential<P>(t, it.theta * it.theta, it.xt, it.yt, it.tperm, it.m, it.out, precision, counters, st); break;
        TC_ORDERS(TC_CASE)
#undef TC_CASE
    default: throw TcError(-3, "tc_potential_item: unsupported order");
    }
}

void tc_eval_potential(const TcTree &t, double theta, const double *xt, const double *yt, int m, double *out,
                       int precision, unsigned long long *counters, TcScratch &ws, hipStream_t st,
                       hipEvent_t ev0, hipEvent_t ev1, bool sort_targets, int xcd_chunk_log2)
{
    if (m <= 0)
        return;
    if (precision == 1 && !t.has_fp32)
        throw TcError(-3, "tc_eval_potential: tree was built without FP32 mirrors");
    if (xcd_chunk_log2 < 0 || xcd_chunk_log2 > 8)
        throw TcError(-3, "tc_eval_potential: XCD chunk of 1 .. 256 workgroups");
    if (sort_targets)
        tc_sort_targets(xt, yt, m, ws, st);
    const double th2 = theta * theta;
    const uint32_t *tperm = ws.tperm.as<uint32_t>();
    if (ev0)
        TC_HIP(hipEventRecord(ev0, st));
    switch (t.order) {
#define TC_CASE(P) \
    case P: launch_potential<P>(t, th2, xt, yt, tperm, m, out, precision, counters, st, xcd_chunk_log2); break;
        TC_ORDERS(TC_CASE)
#undef TC_CASE
    default: throw TcError(-3, "tc_eval_potential: unsupported order");
    }
    if (ev1)
        TC_HIP(hipEventRecord(ev1, st));
}

void tc_eval_force(const TcTree &t, double theta, const double *x0s, const double *y0s, const double *hs,
                   int nblocks, double *xr, double *yr, int precision, unsigned long long *counters,
                   TcScratch &ws, int max_chunks, int inline_leaves, hipStream_t st, hipEvent_t ev0, hipEvent_t ev1)
{
    if (nblocks <= 0)
        return;
    if (precision == 1 && !t.has_fp32)
        throw TcError(-3, "tc_eval_force: tree was built without FP32 mirrors");
    if (ev0)
        TC_HIP(hipEventRecord(ev0, st));
    switch (t.order) {
#define TC_CASE(P)                                                                                         \
    case P:                                                                                                \
        launch_force<P>(t, theta, x0s, y0s, hs, nblocks, xr, yr, precision, counters, ws, max_chunks, \
                        inline_leaves, st);                                                                \
        break;
        TC_ORDERS(TC_CASE)
#undef TC_CASE
    default: throw TcError(-3, "tc_eval_force: unsupported order");
    }
    if (ev1)
        TC_HIP(hipEventRecord(ev1, st));
}

void tc_eval_force_targets(const TcTree &t, double theta, const double *xt, const double *yt, int m, double *xr,
                           double *yr, int precision, unsigned long long *counters, TcScratch &ws,
                           hipStream_t st, hipEvent_t ev0, hipEvent_t ev1)
{
    if (m <= 0)
        return;
    if (precision == 1 && !t.has_fp32)
        throw TcError(-3, "tc_eval_force_targets: tree was built without FP32 mirrors");
    tc_sort_targets(xt, yt, m, ws, st);
    const double th2 = theta * theta;
    const uint32_t *tperm = ws.tperm.as<uint32_t>();
    if (ev0)
        TC_HIP(hipEventRecord(ev0, st));
    switch (t.order) {
#define TC_CASE(P) \
    case P: launch_force_targets<P>(t, th2, xt, yt, tperm, m, xr, yr, precision, counters, st); break;
        TC_ORDERS(TC_CASE)
#undef TC_CASE
    default: throw TcError(-3, "tc_eval_force_targets: unsupported order");
    }
    if (ev1)
        TC_HIP(hipEventRecord(ev1, st));
}
