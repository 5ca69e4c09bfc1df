/*
 * tc_oracle.c -- CPU oracle (checker) for the 2-D treecode hot path.
 *
 * TEST INFRASTRUCTURE ONLY: loaded by tests/, the smoke check and bench.py's
 * CPU-baseline leg.  The product path never calls it.
 *
 * Plain C restatement of Virtual-X/2d-treecodes (file:line cited per
 * function).  Where the reference leaves a result unspecified (OpenMP task
 * timing, ISPC lane order under --opt=fast-math, undefined behaviour on
 * degenerate input) this file fixes ONE canonical choice, and the HIP
 * library follows the same choice:
 *
 *  C1 quantisation uses the form the reference's release flags compile
 *     sort-sources.cpp:117 into: q = floor((x - xmin) * ((1/ext) * 2^15)),
 *     clamped to [0, 2^15 - 1];
 *  C2 bbox accumulators start at min=1e13, max=0 (upward-common.ispc:132),
 *     so the box always contains the origin's positive quadrant edge; a
 *     zero extent (all sources coincident) is replaced by 1.0;
 *  C3 leaf sums (upward-common.ispc:13-57, upward-kernels.m4:52-92) use 64
 *     lane-strided partial sums, lane j taking sources j, j+64, ..., reduced
 *     by the pairwise tree p[j] += p[j+off], off = 32,16,...,1;
 *  C4 the P2P log is FP64 (the reference's bulk float log,
 *     potential-kernels.m4:42, is not emulated);
 *  C5 the node pool is dynamic; node ids are breadth-first: root 0, the
 *     children of the k-th internal node in (level, Morton) order at
 *     1 + 4k + c (the reference's ids depend on OpenMP task timing);
 *  C6 a parent whose children all have w = 0 gets the cell centre as com;
 *     children with w == 0 are skipped in the M2M sum (their terms are
 *     exactly zero); a child with z0 == 0 is translated by the identity;
 *     the zcount==4 side effect on the triggering child
 *     (upward.cpp:178-179) is not emulated;
 *  C7 traversals skip empty nodes (s == e), whose contributions are exactly
 *     zero;
 *  C8 the constants 1/k and -1/l are the nearest doubles and binomials are
 *     exact (upward-kernels.m4:72,128,130; force-kernels.m4:138,141).
 *
 * Build with -ffp-contract=off: every decision quantity (keys, com, r,
 * accept tests) must be rounded exactly as written.
 */
#include "tc_oracle.h"

#include <float.h>
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#define LMAX ORC_LMAX
#define TILE 8
#define BLOCKSIZE 32
#define MAXORDER 64

static const double EPS_REG = 10.0 * DBL_EPSILON; /* potential-kernels.m4:15, force-kernels.m4:26 */

int orc_max_threads(void)
{
#ifdef _OPENMP
    return omp_get_max_threads();
#else
    return 1;
#endif
}

static void set_threads(int nthreads)
{
#ifdef _OPENMP
    if (nthreads > 0)
        omp_set_num_threads(nthreads);
#else
    (void)nthreads;
#endif
}

/* Threads of the following parallel regions (tree builds included). */
void orc_set_threads(int nthreads) { set_threads(nthreads); }

/* std::max / std::min semantics: max(a,b) = a<b ? b : a, min(a,b) = b<a ? b : a */
static inline double smax(double a, double b) { return a < b ? b : a; }
static inline double smin(double a, double b) { return b < a ? b : a; }

/* ---------------------------------------------------------------- bbox, keys, sort */

/* sort-sources.cpp:35-82, upward-common.ispc:126-150 */
int orc_bbox(const double *x, const double *y, int n, double *out)
{
    double xmi = 1e13, ymi = 1e13, xma = 0, yma = 0;
    for (int i = 0; i < n; ++i) {
        xmi = smin(xmi, x[i]);
        xma = smax(xma, x[i]);
        ymi = smin(ymi, y[i]);
        yma = smax(yma, y[i]);
    }
    const double ext0 = xma - xmi, ext1 = yma - ymi;
    const double eps = 10000 * DBL_EPSILON;
    double ext = smax(ext0, ext1) * (1 + 2 * eps);
    if (!(ext > 0))
        ext = 1.0; /* C2 */
    out[0] = xmi - eps * ext;
    out[1] = ymi - eps * ext;
    out[2] = ext;
    return 0;
}

static inline int spread15(int x)
{
    x = (x | (x << 8)) & 0x00FF00FF;
    x = (x | (x << 4)) & 0x0F0F0F0F;
    x = (x | (x << 2)) & 0x33333333;
    x = (x | (x << 1)) & 0x55555555;
    return x;
}

static inline int quantise(double v, double vmin, double scale)
{
    double q = floor((v - vmin) * scale);
    if (!(q >= 0))
        q = 0;
    if (q > (1 << LMAX) - 1)
        q = (1 << LMAX) - 1;
    return (int)q;
}

/* sort-sources.cpp:87-110 in the compiled form C1 */
void orc_keys(const double *x, const double *y, int n, double xmin, double ymin, double ext, int *keys)
{
    const double scale = (1.0 / ext) * (double)(1 << LMAX);
#pragma omp parallel for schedule(static)
    for (int i = 0; i < n; ++i) {
        const int qx = quantise(x[i], xmin, scale);
        const int qy = quantise(y[i], ymin, scale);
        keys[i] = spread15(qx) | (spread15(qy) << 1);
    }
}

/* Stable LSD counting sort on 30-bit keys: the unique order of the
 * lexicographic pair sort at sort-sources.cpp:112 (ties by original index).
 * Each pass is parallel and stays stable: thread t counts, then scatters, its
 * own contiguous chunk, and the output runs of digit d are laid out in thread
 * order. */
void orc_sort(const int *keys, int n, int *perm, int *keys_sorted)
{
    int *pa = (int *)malloc(sizeof(int) * (size_t)(n > 0 ? n : 1));
    int *pb = (int *)malloc(sizeof(int) * (size_t)(n > 0 ? n : 1));
    const int nt = n < 65536 ? 1 : orc_max_threads();
    int *count = (int *)calloc((size_t)nt * 1024, sizeof(int));
#pragma omp parallel for schedule(static) num_threads(nt)
    for (int i = 0; i < n; ++i)
        pa[i] = i;
    for (int shift = 0; shift < 30; shift += 10) {
#pragma omp parallel num_threads(nt)
        {
#ifdef _OPENMP
            const int t = omp_get_thread_num();
#else
            const int t = 0;
#endif
            const int lo = (int)((long long)n * t / nt), hi = (int)((long long)n * (t + 1) / nt);
            int *c = count + (size_t)t * 1024;
            memset(c, 0, sizeof(int) * 1024);
            for (int i = lo; i < hi; ++i)
                c[(keys[pa[i]] >> shift) & 1023]++;
#pragma omp barrier
#pragma omp single
            {
                int run = 0;
                for (int d = 0; d < 1024; ++d)
                    for (int tt = 0; tt < nt; ++tt) {
                        const int v = count[(size_t)tt * 1024 + d];
                        count[(size_t)tt * 1024 + d] = run;
                        run += v;
                    }
            }
            for (int i = lo; i < hi; ++i)
                pb[c[(keys[pa[i]] >> shift) & 1023]++] = pa[i];
        }
        int *tmp = pa;
        pa = pb;
        pb = tmp;
    }
    free(count);
#pragma omp parallel for schedule(static) num_threads(nt)
    for (int i = 0; i < n; ++i) {
        if (perm)
            perm[i] = pa[i];
        if (keys_sorted)
            keys_sorted[i] = keys[pa[i]];
    }
    free(pa);
    free(pb);
}

/* ---------------------------------------------------------------- tree */

struct orc_tree {
    int n, order, leafcap, nnodes, cap_nodes;
    double xmin, ymin, ext;
    int *keys, *perm;
    double *xs, *ys, *vs;
    int *s, *e, *child, *level, *cx, *cy;
    double *xcom, *ycom, *r, *mass, *w, *wx, *wy;
    double *exp; /* 2*order per node */
};

static void tree_grow(orc_tree *t, int need)
{
    if (need <= t->cap_nodes)
        return;
    int cap = t->cap_nodes ? t->cap_nodes : 64;
    while (cap < need)
        cap *= 2;
#define GROW(p, T) t->p = (T *)realloc(t->p, sizeof(T) * (size_t)cap)
    GROW(s, int); GROW(e, int); GROW(child, int); GROW(level, int); GROW(cx, int); GROW(cy, int);
    GROW(xcom, double); GROW(ycom, double); GROW(r, double); GROW(mass, double);
    GROW(w, double); GROW(wx, double); GROW(wy, double);
#undef GROW
    t->exp = (double *)realloc(t->exp, sizeof(double) * 2 * (size_t)t->order * (size_t)cap);
    t->cap_nodes = cap;
}

static int lower_bound_i(const int *k, int s, int e, int val)
{
    while (s < e) {
        int mid = s + (e - s) / 2;
        if (k[mid] < val) s = mid + 1; else e = mid;
    }
    return s;
}

static int upper_bound_i(const int *k, int s, int e, int val)
{
    while (s < e) {
        int mid = s + (e - s) / 2;
        if (k[mid] <= val) s = mid + 1; else e = mid;
    }
    return s;
}

/* Canonical leaf reduction C3 of a per-lane partial array p[64]. */
static inline double tree64(double *p)
{
    for (int off = 32; off >= 1; off >>= 1)
        for (int l = 0; l < off; ++l)
            p[l] = p[l] + p[l + off];
    return p[0];
}

/* upward-common.ispc:13-57 (canonical C3 order): mass, w, wx, wy */
static void leaf_sums(const double *x, const double *y, const double *v, int n,
                      double *m, double *w, double *wx, double *wy)
{
    double pm[64], pw[64], px[64], py[64];
    for (int l = 0; l < 64; ++l) {
        pm[l] = pw[l] = px[l] = py[l] = 0;
        for (int i = l; i < n; i += 64) {
            const double a = fabs(v[i]);
            pm[l] += v[i];
            pw[l] += a;
            px[l] += x[i] * a;
            py[l] += y[i] * a;
        }
    }
    *m = tree64(pm);
    *w = tree64(pw);
    *wx = tree64(px);
    *wy = tree64(py);
}

/* upward-kernels.m4:52-92 (canonical C3 order) */
void orc_p2e(const double *x, const double *y, const double *v, int n, double xc, double yc,
             int order, double *re, double *im)
{
    static const int L = 64;
    double *pr = (double *)calloc((size_t)L * order, sizeof(double));
    double *pi = (double *)calloc((size_t)L * order, sizeof(double));
    double inv[MAXORDER + 2];
    for (int k = 1; k <= order; ++k)
        inv[k] = 1.0 / (double)k; /* C8 */
    for (int l = 0; l < L; ++l) {
        double *rx = pr + (size_t)l * order, *ix = pi + (size_t)l * order;
        for (int i = l; i < n; i += L) {
            const double r0 = x[i] - xc, i0 = y[i] - yc, src = v[i];
            double rtmp = r0 * src, itmp = i0 * src;
            rx[0] -= rtmp;
            ix[0] -= itmp;
            double rprod = r0, iprod = i0;
            for (int k = 1; k < order; ++k) {
                rtmp = rprod * r0 - iprod * i0;
                itmp = rprod * i0 + iprod * r0;
                const double term = src * inv[k + 1];
                rprod = rtmp;
                iprod = itmp;
                rtmp = rprod * term;
                itmp = iprod * term;
                rx[k] -= rtmp;
                ix[k] -= itmp;
            }
        }
    }
    double tr[64], ti[64];
    for (int k = 0; k < order; ++k) {
        for (int l = 0; l < L; ++l) {
            tr[l] = pr[(size_t)l * order + k];
            ti[l] = pi[(size_t)l * order + k];
        }
        re[k] = tree64(tr);
        im[k] = tree64(ti);
    }
    free(pr);
    free(pi);
}

double orc_binomial(int n, int k)
{
    if (k < 0 || k > n)
        return 0;
    if (k > n - k)
        k = n - k;
    /* exact while the result < 2^63 (n <= 66 is plenty: n <= 2*MAXORDER-1 used with small k) */
    unsigned __int128 r = 1;
    for (int i = 1; i <= k; ++i)
        r = r * (unsigned __int128)(n - k + i) / (unsigned __int128)i;
    return (double)r;
}

/* upward-kernels.m4:95-149 for nchild children, summed in child order (C6). */
void orc_e2e(const double *x0s, const double *y0s, const double *masses, const double *const *rxps,
             const double *const *ixps, int nchild, int order, double *dre, double *dim)
{
    const int p = order;
    for (int l = 0; l < p; ++l)
        dre[l] = dim[l] = 0;
    for (int c = 0; c < nchild; ++c) {
        const double x0 = x0s[c], y0 = y0s[c], mass = masses[c];
        const double *ra = rxps[c], *ia = ixps[c];
        if (x0 == 0 && y0 == 0) { /* C6 identity */
            for (int l = 0; l < p; ++l) {
                dre[l] += ra[l];
                dim[l] += ia[l];
            }
            continue;
        }
        double rinv[MAXORDER + 1], iinv[MAXORDER + 1], rc[MAXORDER + 1], ic[MAXORDER + 1];
        const double r2z0 = x0 * x0 + y0 * y0;
        rinv[1] = x0 / r2z0;
        iinv[1] = -y0 / r2z0;
        for (int j = 2; j <= p; ++j) {
            rinv[j] = rinv[j - 1] * rinv[1] - iinv[j - 1] * iinv[1];
            iinv[j] = rinv[j - 1] * iinv[1] + iinv[j - 1] * rinv[1];
        }
        for (int j = 1; j <= p; ++j) {
            rc[j] = ra[j - 1] * rinv[j] - ia[j - 1] * iinv[j];
            ic[j] = ra[j - 1] * iinv[j] + ia[j - 1] * rinv[j];
        }
        for (int l = 1; l <= p; ++l) {
            double rt = (l == 1) ? -mass : mass * (-(1.0 / (double)l));
            double it = ic[1];
            rt = rt + rc[1];
            for (int k = 2; k <= l; ++k) {
                const double b = orc_binomial(l - 1, k - 1);
                rt = rt + rc[k] * b;
                it = it + ic[k] * b;
            }
            const double invz2 = rinv[l] * rinv[l] + iinv[l] * iinv[l];
            const double invinv = invz2 != 0 ? 1 / invz2 : 0;
            const double rz = rinv[l] * invinv;
            const double iz = -iinv[l] * invinv;
            const double rpart = rt * rz - it * iz;
            const double ipart = rt * iz + it * rz;
            dre[l - 1] += rpart;
            dim[l - 1] += ipart;
        }
    }
}

/* Tree::build (upward.cpp:280-314) with BFS numbering (C5); build_tree
 * (upward.cpp:223-277) one level at a time; then the upward pass
 * (build_leaf upward.cpp:61-219) leaves first, parents in decreasing id. */
orc_tree *orc_tree_build(const double *xsrc, const double *ysrc, const double *vsrc, int n,
                         int leafcap, int order)
{
    orc_tree *t = (orc_tree *)calloc(1, sizeof(orc_tree));
    t->n = n;
    t->order = order;
    t->leafcap = leafcap;
    const size_t nn = (size_t)(n > 0 ? n : 1);
    t->keys = (int *)malloc(sizeof(int) * nn);
    t->perm = (int *)malloc(sizeof(int) * nn);
    t->xs = (double *)malloc(sizeof(double) * nn);
    t->ys = (double *)malloc(sizeof(double) * nn);
    t->vs = (double *)malloc(sizeof(double) * nn);

    double bb[3];
    orc_bbox(xsrc, ysrc, n, bb);
    t->xmin = bb[0];
    t->ymin = bb[1];
    t->ext = bb[2];
    int *raw = (int *)malloc(sizeof(int) * nn);
    orc_keys(xsrc, ysrc, n, t->xmin, t->ymin, t->ext, raw);
    orc_sort(raw, n, t->perm, t->keys);
    free(raw);
#pragma omp parallel for schedule(static)
    for (int i = 0; i < n; ++i) {
        t->xs[i] = xsrc[t->perm[i]];
        t->ys[i] = ysrc[t->perm[i]];
        t->vs[i] = vsrc[t->perm[i]];
    }

    /* topology, breadth first */
    tree_grow(t, 1);
    t->s[0] = 0; t->e[0] = n; t->child[0] = 0; t->level[0] = 0; t->cx[0] = 0; t->cy[0] = 0;
    int nnodes = 1, lvl_begin = 0, lvl_end = 1;
    int lvl_start[LMAX + 2];
    int nlevels = 0;
    for (int l = 0; lvl_begin < lvl_end; ++l) {
        lvl_start[nlevels++] = lvl_begin;
        /* child bases in id order (serial, one add per node), then the four
         * children of every internal node of the level in parallel */
        for (int id = lvl_begin; id < lvl_end; ++id) {
            const int leaf = t->e[id] - t->s[id] <= leafcap || l + 1 > LMAX;
            t->child[id] = leaf ? 0 : nnodes;
            nnodes += leaf ? 0 : 4;
        }
        tree_grow(t, nnodes);
#pragma omp parallel for schedule(dynamic, 64)
        for (int id = lvl_begin; id < lvl_end; ++id) {
            const int s = t->s[id], e = t->e[id];
            const int cb = t->child[id];
            if (!cb)
                continue;
            /* mask = Morton prefix of the cell, rebuilt from its coordinates */
            const int shift_l = 2 * (LMAX - l);
            const int mask = (shift_l >= 30) ? 0
                             : ((spread15(t->cx[id]) | (spread15(t->cy[id]) << 1)) << shift_l);
            for (int c = 0; c < 4; ++c) {
                const int shift = 2 * (LMAX - l - 1);
                const int key1 = mask | (c << shift);
                const int key2 = key1 + (1 << shift) - 1;
                const int imin = c == 0 ? s : lower_bound_i(t->keys, s, e, key1);
                const int isup = c == 3 ? e : upper_bound_i(t->keys, s, e, key2);
                const int k = cb + c;
                t->s[k] = imin;
                t->e[k] = isup;
                t->child[k] = 0;
                t->level[k] = l + 1;
                t->cx[k] = (t->cx[id] << 1) + (c & 1);
                t->cy[k] = (t->cy[id] << 1) + (c >> 1);
            }
        }
        lvl_begin = lvl_end;
        lvl_end = nnodes;
    }
    lvl_start[nlevels] = nnodes;
    t->nnodes = nnodes;

    /* upward: leaves (independent) */
    const int p = order;
#pragma omp parallel for schedule(dynamic, 16)
    for (int id = 0; id < nnodes; ++id) {
        if (t->child[id])
            continue;
        const int s = t->s[id], e = t->e[id], l = t->level[id];
        const double h = t->ext / (double)(1 << l);
        const double x0 = t->xmin + h * t->cx[id], y0 = t->ymin + h * t->cy[id];
        double m, w, wx, wy;
        leaf_sums(t->xs + s, t->ys + s, t->vs + s, e - s, &m, &w, &wx, &wy);
        t->mass[id] = m;
        t->w[id] = w;
        t->wx[id] = wx;
        t->wy[id] = wy;
        t->xcom[id] = w != 0 ? wx / w : (x0 + 0.5 * h);
        t->ycom[id] = w != 0 ? wy / w : (y0 + 0.5 * h);
        double r2 = 0;
        for (int i = s; i < e; ++i) {
            const double xr = t->xs[i] - t->xcom[id], yr = t->ys[i] - t->ycom[id];
            r2 = smax(r2, xr * xr + yr * yr);
        }
        t->r[id] = sqrt(r2);
        orc_p2e(t->xs + s, t->ys + s, t->vs + s, e - s, t->xcom[id], t->ycom[id], p,
                t->exp + (size_t)2 * p * id, t->exp + (size_t)2 * p * id + p);
    }
    /* upward: parents (upward.cpp:95-218), deepest level first; the parents of
     * one level are independent */
    for (int lv = nlevels - 1; lv >= 0; --lv)
#pragma omp parallel for schedule(dynamic, 16)
    for (int id = lvl_start[lv]; id < lvl_start[lv + 1]; ++id) {
        if (!t->child[id])
            continue;
        const int cb = t->child[id], l = t->level[id];
        double msum = 0, wsum = 0, wxsum = 0, wysum = 0;
        for (int c = 0; c < 4; ++c) {
            msum += t->mass[cb + c];
            wsum += t->w[cb + c];
            wxsum += t->wx[cb + c];
            wysum += t->wy[cb + c];
        }
        t->mass[id] = msum;
        t->w[id] = wsum;
        t->wx[id] = wxsum;
        t->wy[id] = wysum;
        const double h = t->ext / (double)(1 << l);
        const double xcp = wsum != 0 ? wxsum / wsum : (t->xmin + h * t->cx[id]) + 0.5 * h; /* C6 */
        const double ycp = wsum != 0 ? wysum / wsum : (t->ymin + h * t->cy[id]) + 0.5 * h;
        double rr = 0;
        for (int c = 0; c < 4; ++c) {
            const int k = cb + c;
            if (t->w[k] != 0) {
                const double rx = xcp - t->xcom[k], ry = ycp - t->ycom[k];
                rr = smax(rr, t->r[k] + sqrt(rx * rx + ry * ry));
            }
        }
        t->r[id] = smin(rr, (double)1.4143f * t->ext / (double)(1 << l));
        t->xcom[id] = xcp;
        t->ycom[id] = ycp;

        int zcount = 0, nz = 0;
        for (int c = 0; c < 4; ++c) {
            const int neg = fabs(t->w[cb + c]) < 1.1e-16;
            zcount += neg;
            if (!neg)
                nz = cb + c;
        }
        double *dst = t->exp + (size_t)2 * p * id;
        if (zcount == 4) {
            for (int i = 0; i < 2 * p; ++i)
                dst[i] = 0;
        } else if (zcount == 3) {
            memcpy(dst, t->exp + (size_t)2 * p * nz, sizeof(double) * 2 * (size_t)p);
        } else {
            double x0s[4], y0s[4], ms[4];
            const double *rp[4], *ip[4];
            int nc = 0;
            for (int c = 0; c < 4; ++c) {
                const int k = cb + c;
                if (t->w[k] == 0)
                    continue; /* C6 */
                x0s[nc] = t->xcom[k] - xcp;
                y0s[nc] = t->ycom[k] - ycp;
                ms[nc] = t->mass[k];
                rp[nc] = t->exp + (size_t)2 * p * k;
                ip[nc] = t->exp + (size_t)2 * p * k + p;
                ++nc;
            }
            orc_e2e(x0s, y0s, ms, rp, ip, nc, p, dst, dst + p);
        }
    }
    return t;
}

void orc_tree_free(orc_tree *t)
{
    if (!t)
        return;
    free(t->keys); free(t->perm); free(t->xs); free(t->ys); free(t->vs);
    free(t->s); free(t->e); free(t->child); free(t->level); free(t->cx); free(t->cy);
    free(t->xcom); free(t->ycom); free(t->r); free(t->mass); free(t->w); free(t->wx); free(t->wy);
    free(t->exp);
    free(t);
}

int orc_tree_nnodes(const orc_tree *t) { return t->nnodes; }
int orc_tree_order(const orc_tree *t) { return t->order; }

void orc_tree_export(const orc_tree *t, double *bbox, int *keys, int *perm, double *xyv,
                     int *node_se, int *node_child, int *node_level, int *node_cell,
                     double *node_geom, double *node_w, double *expansions)
{
    const int n = t->n, nn = t->nnodes;
    if (bbox) { bbox[0] = t->xmin; bbox[1] = t->ymin; bbox[2] = t->ext; }
    if (keys) memcpy(keys, t->keys, sizeof(int) * (size_t)n);
    if (perm) memcpy(perm, t->perm, sizeof(int) * (size_t)n);
    if (xyv) {
        memcpy(xyv, t->xs, sizeof(double) * (size_t)n);
        memcpy(xyv + n, t->ys, sizeof(double) * (size_t)n);
        memcpy(xyv + 2 * (size_t)n, t->vs, sizeof(double) * (size_t)n);
    }
    for (int i = 0; i < nn; ++i) {
        if (node_se) { node_se[2 * i] = t->s[i]; node_se[2 * i + 1] = t->e[i]; }
        if (node_child) node_child[i] = t->child[i];
        if (node_level) node_level[i] = t->level[i];
        if (node_cell) { node_cell[2 * i] = t->cx[i]; node_cell[2 * i + 1] = t->cy[i]; }
        if (node_geom) {
            node_geom[4 * i + 0] = t->xcom[i];
            node_geom[4 * i + 1] = t->ycom[i];
            node_geom[4 * i + 2] = t->r[i];
            node_geom[4 * i + 3] = t->mass[i];
        }
        if (node_w) { node_w[3 * i] = t->w[i]; node_w[3 * i + 1] = t->wx[i]; node_w[3 * i + 2] = t->wy[i]; }
    }
    if (expansions)
        memcpy(expansions, t->exp, sizeof(double) * 2 * (size_t)t->order * (size_t)nn);
}

/* ---------------------------------------------------------------- potential */

/* potential-kernels.m4:59-89 for one expansion */
double orc_e2p_potential(double rz, double iz, double mass, const double *ra, const double *ia, int order)
{
    const double r2 = rz * rz + iz * iz;
    const double rinv1 = rz / r2, iinv1 = -iz / r2;
    double rinv = rinv1, iinv = iinv1;
    double sum = ra[0] * rinv - ia[0] * iinv;
    for (int j = 2; j <= order; ++j) {
        const double rn = rinv * rinv1 - iinv * iinv1;
        const double in = rinv * iinv1 + iinv * rinv1;
        rinv = rn;
        iinv = in;
        sum += ra[j - 1] * rinv - ia[j - 1] * iinv;
    }
    return mass * log(r2) / 2 + sum;
}

/* potential-kernels.m4:19-57 with FP64 log (C4) */
static double p2p_potential(const double *xs, const double *ys, const double *vs, int n, double xt, double yt)
{
    double s = 0;
    for (int i = 0; i < n; ++i) {
        const double xr = xt - xs[i], yr = yt - ys[i];
        s += log(xr * xr + yr * yr + EPS_REG) * vs[i];
    }
    return s * 0.5;
}

/* EvaluatePotential::evaluate, treecode-potential.cpp:25-86 */
static double eval_potential(const orc_tree *t, double th2, double xt, double yt, long long *cnt)
{
    int stack[4 * LMAX + 8];
    int sp = 0;
    stack[0] = 0;
    double res = 0;
    const int p = t->order;
    long long np2p = 0, ne2p = 0;
    while (sp > -1) {
        const int id = stack[sp--];
        if (t->s[id] == t->e[id])
            continue; /* C7 */
        const double dx = xt - t->xcom[id], dy = yt - t->ycom[id];
        const double r2 = dx * dx + dy * dy;
        if (t->r[id] * t->r[id] < th2 * r2) {
            const double *ex = t->exp + (size_t)2 * p * id;
            res += orc_e2p_potential(dx, dy, t->mass[id], ex, ex + p, p);
            ++ne2p;
        } else if (!t->child[id]) {
            const int s = t->s[id], e = t->e[id];
            res += p2p_potential(t->xs + s, t->ys + s, t->vs + s, e - s, xt, yt);
            np2p += e - s;
        } else {
            for (int c = 0; c < 4; ++c)
                stack[++sp] = t->child[id] + c;
        }
    }
    if (cnt) {
        cnt[ORC_CNT_P2P_PAIRS] += np2p;
        cnt[ORC_CNT_E2P] += ne2p;
    }
    return res;
}

void orc_potential_tree(const orc_tree *t, double theta, const double *xt, const double *yt,
                        int t0, int t1, double *out, long long *counters, int nthreads)
{
    set_threads(nthreads);
    const double th2 = theta * theta;
    long long c0 = 0, c1 = 0;
#pragma omp parallel for schedule(static, 1) reduction(+ : c0, c1)
    for (int i = t0; i < t1; ++i) {
        long long cnt[ORC_NCOUNTERS] = {0, 0, 0, 0};
        out[i] = eval_potential(t, th2, xt[i], yt[i], cnt);
        c0 += cnt[0];
        c1 += cnt[1];
    }
    if (counters) {
        counters[ORC_CNT_P2P_PAIRS] += c0;
        counters[ORC_CNT_E2P] += c1;
    }
}

/* treecode_potential, treecode-potential.cpp:91-106 */
void orc_potential(double theta, const double *xs, const double *ys, const double *vs, int n,
                   const double *xt, const double *yt, int m, double *out,
                   int order, int leafcap, long long *counters, int nthreads)
{
    if (order <= 0) order = 12;
    if (leafcap <= 0) leafcap = 32 * 16;
    if (n <= 0) {
        for (int i = 0; i < m; ++i)
            out[i] = 0;
        return;
    }
    orc_tree *t = orc_tree_build(xs, ys, vs, n, leafcap, order);
    orc_potential_tree(t, theta, xt, yt, 0, m, out, counters, nthreads);
    orc_tree_free(t);
}

/* ---------------------------------------------------------------- force */

/* force-kernels.m4:57-98 for one target */
void orc_e2p_force(double mass, double rz, double iz, const double *ra, const double *ia, int order,
                   double *fx, double *fy)
{
    const double r2 = rz * rz + iz * iz;
    const double rinv1 = rz / r2, iinv1 = -iz / r2;
    double rsum = mass * rinv1, isum = mass * iinv1;
    double rprod = rinv1, iprod = iinv1;
    for (int j = 0; j < order; ++j) {
        const double rt = rprod * rinv1 - iprod * iinv1;
        const double it = rprod * iinv1 + iprod * rinv1;
        rprod = rt;
        iprod = it;
        rsum -= (j + 1) * (ra[j] * rprod - ia[j] * iprod);
        isum -= (j + 1) * (ra[j] * iprod + ia[j] * rprod);
    }
    *fx += rsum;
    *fy -= isum;
}

/* force-kernels.m4:102-155 for one expansion: lre/lim indexed 0..order */
void orc_e2l(double x0, double y0, double mass, const double *ra, const double *ia, int order,
             double *lre, double *lim)
{
    const int p = order;
    double rinv[MAXORDER + 1], iinv[MAXORDER + 1], rc[MAXORDER + 1], ic[MAXORDER + 1];
    const double r2z0 = x0 * x0 + y0 * y0;
    rinv[1] = x0 / r2z0;
    iinv[1] = -y0 / r2z0;
    for (int j = 2; j <= p; ++j) {
        rinv[j] = rinv[j - 1] * rinv[1] - iinv[j - 1] * iinv[1];
        iinv[j] = rinv[j - 1] * iinv[1] + iinv[j - 1] * rinv[1];
    }
    for (int j = 1; j <= p; ++j) {
        rc[j] = ra[j - 1] * rinv[j] - ia[j - 1] * iinv[j];
        ic[j] = ra[j - 1] * iinv[j] + ia[j - 1] * rinv[j];
    }
    for (int l = 1; l <= p; ++l) {
        double rt = (l == 1) ? -mass : mass * (-(1.0 / (double)l));
        double it = 0;
        for (int k = 1; k <= p; ++k) {
            const double b = orc_binomial(l + k - 1, k - 1);
            if (k & 1) {
                rt -= b * rc[k];
                it -= b * ic[k];
            } else {
                rt += b * rc[k];
                it += b * ic[k];
            }
        }
        lre[l] += rt * rinv[l] - it * iinv[l];
        lim[l] += rt * iinv[l] + it * rinv[l];
    }
}

/* force-kernels.m4:157-191 for one target */
void orc_l2p(double rz1, double iz1, const double *lre, const double *lim, int order, double *fx, double *fy)
{
    double rres = lre[1], ires = lim[1];
    double rz = rz1, iz = iz1; /* z^(l-1) */
    for (int l = 2; l <= order; ++l) {
        rres = rres + l * (lre[l] * rz - lim[l] * iz);
        ires = ires + l * (lre[l] * iz + lim[l] * rz);
        const double rn = rz * rz1 - iz * iz1;
        const double in = rz * iz1 + iz * rz1;
        rz = rn;
        iz = in;
    }
    *fx += rres;
    *fy -= ires;
}

/* force-kernels.m4:15-55: 64 targets x = x0 + ix*h, y = y0 + iy*h */
static void p2p_force_8x8(const double *xs, const double *ys, const double *vs, int n,
                          double x0, double y0, double h, double *xres, double *yres)
{
    double xt[TILE], yt[TILE];
    for (int i = 0; i < TILE; ++i) {
        xt[i] = x0 + i * h;
        yt[i] = y0 + i * h;
    }
    for (int s = 0; s < n; ++s) {
        const double xsrc = xs[s], ysrc = ys[s], vsrc = vs[s];
        for (int iy = 0; iy < TILE; ++iy) {
            const double yr = yt[iy] - ysrc;
            const double yr2 = yr * yr + EPS_REG;
            for (int ix = 0; ix < TILE; ++ix) {
                const double xr = xt[ix] - xsrc;
                const double xr2 = xr * xr;
                const double f = vsrc / (xr2 + yr2);
                xres[ix + TILE * iy] += xr * f;
                yres[ix + TILE * iy] += yr * f;
            }
        }
    }
}

/* EvaluateForce::evaluate, treecode-force.cpp:94-199 */
static void eval_force_block(const orc_tree *t, double theta, double x0, double y0, double h,
                             double *xrb, double *yrb, long long *cnt)
{
    const int p = t->order;
    int stack[4 * LMAX + 8];
    double result[2 * TILE * TILE];
    double lre[MAXORDER + 1], lim[MAXORDER + 1];
    const double rbrick = 1.4142135623730951 * h * (TILE - 1) * 0.5;
    const double invtheta = 1 / theta;
    const double th2 = theta * theta;
    for (int by = 0; by < BLOCKSIZE; by += TILE)
        for (int bx = 0; bx < BLOCKSIZE; bx += TILE) {
            const double x0b = x0 + h * (bx + 0.5 * (TILE - 1));
            const double y0b = y0 + h * (by + 0.5 * (TILE - 1));
            for (int i = 0; i <= p; ++i)
                lre[i] = lim[i] = 0;
            for (int i = 0; i < 2 * TILE * TILE; ++i)
                result[i] = 0;
            int ne2l = 0;
            int sp = 0;
            stack[0] = 0;
            while (sp > -1) {
                const int id = stack[sp--];
                if (t->s[id] == t->e[id])
                    continue; /* C7 */
                const double xcom = t->xcom[id], ycom = t->ycom[id], r = t->r[id];
                const double ddx = x0b - xcom, ddy = y0b - ycom;
                const double distance = sqrt(ddx * ddx + ddy * ddy);
                const int e2l = (distance / r - 1) > invtheta && rbrick <= r;
                const double *ex = t->exp + (size_t)2 * p * id;
                if (e2l) {
                    orc_e2l(xcom - x0b, ycom - y0b, t->mass[id], ex, ex + p, p, lre, lim);
                    ++ne2l;
                    continue;
                }
                const double xt = smax(x0 + bx * h, smin(x0 + (bx + TILE - 1) * h, xcom));
                const double yt = smax(y0 + by * h, smin(y0 + (by + TILE - 1) * h, ycom));
                const double r2 = (xt - xcom) * (xt - xcom) + (yt - ycom) * (yt - ycom);
                if (r * r < theta * theta * r2) {
                    (void)th2;
                    const double rz0 = x0 + (bx + 0) * h - xcom, iz0 = y0 + (by + 0) * h - ycom;
                    for (int iy = 0; iy < TILE; ++iy)
                        for (int ix = 0; ix < TILE; ++ix) {
                            const int d = ix + TILE * iy;
                            orc_e2p_force(t->mass[id], rz0 + ix * h, iz0 + iy * h, ex, ex + p, p,
                                          &result[d], &result[TILE * TILE + d]);
                        }
                    cnt[ORC_CNT_E2P] += TILE * TILE;
                } else if (!t->child[id]) {
                    const int s = t->s[id];
                    p2p_force_8x8(t->xs + s, t->ys + s, t->vs + s, t->e[id] - s, x0 + (bx + 0) * h,
                                  y0 + (by + 0) * h, h, result, result + TILE * TILE);
                    cnt[ORC_CNT_P2P_PAIRS] += (long long)TILE * TILE * (t->e[id] - s);
                } else {
                    for (int c = 0; c < 4; ++c)
                        stack[++sp] = t->child[id] + c;
                }
            }
            cnt[ORC_CNT_E2L] += ne2l;
            cnt[ORC_CNT_L2P] += TILE * TILE;
            const double zl = h * (0 - 0.5 * (TILE - 1));
            for (int iy = 0; iy < TILE; ++iy)
                for (int ix = 0; ix < TILE; ++ix) {
                    const int d = ix + TILE * iy;
                    orc_l2p(zl + ix * h, zl + iy * h, lre, lim, p, &result[d], &result[TILE * TILE + d]);
                }
            for (int iy = 0; iy < TILE; ++iy)
                for (int ix = 0; ix < TILE; ++ix) {
                    xrb[bx + ix + BLOCKSIZE * (by + iy)] = result[ix + TILE * iy];
                    yrb[bx + ix + BLOCKSIZE * (by + iy)] = result[TILE * TILE + ix + TILE * iy];
                }
        }
}

void orc_force_tree(const orc_tree *t, double theta, const double *x0s, const double *y0s,
                    const double *hs, int b0, int b1, double *xr, double *yr,
                    long long *counters, int nthreads)
{
    set_threads(nthreads);
    long long c0 = 0, c1 = 0, c2 = 0, c3 = 0;
#pragma omp parallel for schedule(dynamic, 1) reduction(+ : c0, c1, c2, c3)
    for (int b = b0; b < b1; ++b) {
        long long cnt[ORC_NCOUNTERS] = {0, 0, 0, 0};
        eval_force_block(t, theta, x0s[b], y0s[b], hs[b], xr + (size_t)b * BLOCKSIZE * BLOCKSIZE,
                         yr + (size_t)b * BLOCKSIZE * BLOCKSIZE, cnt);
        c0 += cnt[0]; c1 += cnt[1]; c2 += cnt[2]; c3 += cnt[3];
    }
    if (counters) {
        counters[0] += c0; counters[1] += c1; counters[2] += c2; counters[3] += c3;
    }
}

/* treecode_force_mrag, treecode-force.cpp:201-236 */
void orc_force_mrag(double theta, const double *xs, const double *ys, const double *vs, int n,
                    const double *x0s, const double *y0s, const double *hs, int nblocks,
                    double *xr, double *yr, int order, int leafcap, long long *counters, int nthreads)
{
    if (order <= 0) order = 24;
    if (leafcap <= 0) leafcap = 192;
    if (n <= 0) {
        memset(xr, 0, sizeof(double) * (size_t)nblocks * BLOCKSIZE * BLOCKSIZE);
        memset(yr, 0, sizeof(double) * (size_t)nblocks * BLOCKSIZE * BLOCKSIZE);
        return;
    }
    orc_tree *t = orc_tree_build(xs, ys, vs, n, leafcap, order);
    orc_force_tree(t, theta, x0s, y0s, hs, 0, nblocks, xr, yr, counters, nthreads);
    orc_tree_free(t);
}

/* ---------------------------------------------------------------- force at arbitrary targets */

/* The gradient at arbitrary targets (no counterpart in the reference, whose
 * velocities exist only on 32x32 blocks, main.cpp:191-195): the per-target DFS
 * and opening test of treecode-potential.cpp:25-86 with the force E2P of
 * force-kernels.m4:57-98 and the pair term of force-kernels.m4:46-52. */
static void eval_force_target(const orc_tree *t, double th2, double xt, double yt, double *fx, double *fy,
                              long long *cnt)
{
    int stack[4 * LMAX + 8];
    int sp = 0;
    stack[0] = 0;
    const int p = t->order;
    double ax = 0, ay = 0;
    while (sp > -1) {
        const int id = stack[sp--];
        if (t->s[id] == t->e[id])
            continue; /* C7 */
        const double dx = xt - t->xcom[id], dy = yt - t->ycom[id];
        const double r2 = dx * dx + dy * dy;
        if (t->r[id] * t->r[id] < th2 * r2) {
            const double *ex = t->exp + (size_t)2 * p * id;
            orc_e2p_force(t->mass[id], dx, dy, ex, ex + p, p, &ax, &ay);
            cnt[ORC_CNT_E2P] += 1;
        } else if (!t->child[id]) {
            for (int i = t->s[id]; i < t->e[id]; ++i) {
                const double xr = xt - t->xs[i], yr = yt - t->ys[i];
                const double f = t->vs[i] / (xr * xr + (yr * yr + EPS_REG));
                ax += xr * f;
                ay += yr * f;
            }
            cnt[ORC_CNT_P2P_PAIRS] += t->e[id] - t->s[id];
        } else {
            for (int c = 0; c < 4; ++c)
                stack[++sp] = t->child[id] + c;
        }
    }
    *fx = ax;
    *fy = ay;
}

void orc_force_targets_tree(const orc_tree *t, double theta, const double *xt, const double *yt, int t0, int t1,
                            double *xr, double *yr, long long *counters, int nthreads)
{
    set_threads(nthreads);
    const double th2 = theta * theta;
    long long c0 = 0, c1 = 0;
#pragma omp parallel for schedule(static, 1) reduction(+ : c0, c1)
    for (int i = t0; i < t1; ++i) {
        long long cnt[ORC_NCOUNTERS] = {0, 0, 0, 0};
        eval_force_target(t, th2, xt[i], yt[i], &xr[i], &yr[i], cnt);
        c0 += cnt[0];
        c1 += cnt[1];
    }
    if (counters) {
        counters[ORC_CNT_P2P_PAIRS] += c0;
        counters[ORC_CNT_E2P] += c1;
    }
}

void orc_force_targets(double theta, const double *xs, const double *ys, const double *vs, int n,
                       const double *xt, const double *yt, int m, double *xr, double *yr, int order, int leafcap,
                       long long *counters, int nthreads)
{
    if (order <= 0) order = 24;
    if (leafcap <= 0) leafcap = 192;
    if (n <= 0) {
        memset(xr, 0, sizeof(double) * (size_t)m);
        memset(yr, 0, sizeof(double) * (size_t)m);
        return;
    }
    orc_tree *t = orc_tree_build(xs, ys, vs, n, leafcap, order);
    orc_force_targets_tree(t, theta, xt, yt, 0, m, xr, yr, counters, nthreads);
    orc_tree_free(t);
}

/* ---------------------------------------------------------------- direct sums */

/* main.cpp:206-224 */
void orc_direct_potential(const double *xs, const double *ys, const double *vs, int n,
                          const double *xt, const double *yt, const int *idx, int nidx,
                          double *out, int nthreads)
{
    set_threads(nthreads);
    const double eps = DBL_EPSILON * 10;
#pragma omp parallel for schedule(dynamic, 4)
    for (int k = 0; k < nidx; ++k) {
        const double xd = xt[idx[k]], yd = yt[idx[k]];
        double s = 0;
        for (int j = 0; j < n; ++j) {
            const double xr = xd - xs[j], yr = yd - ys[j];
            const double r = sqrt(xr * xr + yr * yr + eps);
            s += log(r) * vs[j];
        }
        out[k] = s;
    }
}

/* main.cpp:225-245 */
void orc_direct_force(const double *xs, const double *ys, const double *vs, int n,
                      const double *xt, const double *yt, const int *idx, int nidx,
                      double *outx, double *outy, int nthreads)
{
    set_threads(nthreads);
    const double eps = DBL_EPSILON * 10;
#pragma omp parallel for schedule(dynamic, 4)
    for (int k = 0; k < nidx; ++k) {
        const double xd = xt[idx[k]], yd = yt[idx[k]];
        double sx = 0, sy = 0;
        for (int j = 0; j < n; ++j) {
            const double xr = xd - xs[j], yr = yd - ys[j];
            const double f = vs[j] / (xr * xr + yr * yr + eps);
            sx += xr * f;
            sy += yr * f;
        }
        outx[k] = sx;
        outy[k] = sy;
    }
}
