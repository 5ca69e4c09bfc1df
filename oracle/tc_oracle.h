/*
 * tc_oracle.h -- CPU oracle for the 2-D treecode hot path.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the checker that the HIP
 * product path is compared against; only tests/, the smoke check and the
 * CPU-baseline leg of bench.py load it.  Nothing in the product links it.
 *
 * It restates, in plain C, the algorithm of Virtual-X/2d-treecodes:
 *   drivers/sort-sources.cpp:35-131   bbox, quantise, interleave, sort, gather
 *   drivers/upward.cpp:61-314         build_tree, build_leaf, parent update, E2E
 *   kernels/upward-common.ispc:13-150 upward_setup, minmax_vec
 *   kernels/upward-kernels.m4:18-149  upward_p2e, upward_e2e
 *   kernels/potential-kernels.m4      potential_p2p, potential_e2p
 *   kernels/force-kernels.m4          force_p2p_8x8, force_e2p_8x8,
 *                                     downward_e2l, downward_l2p_8x8
 *   drivers/treecode-potential.cpp    EvaluatePotential::evaluate
 *   drivers/treecode-force.cpp        EvaluateForce::evaluate
 *   main.cpp:175-245                  FP64 direct sums
 * with the canonical choices listed in tc_oracle.c's header comment.
 */
#ifndef TC_ORACLE_H
#define TC_ORACLE_H

#ifdef __cplusplus
extern "C" {
#endif

#define ORC_LMAX 15

/* counters[] layout shared with the HIP library (see treecodes2d.h) */
enum {
    ORC_CNT_P2P_PAIRS = 0, /* target-source pairs evaluated directly     */
    ORC_CNT_E2P = 1,       /* (target, node) multipole evaluations       */
    ORC_CNT_E2L = 2,       /* (brick, node) multipole-to-local shifts    */
    ORC_CNT_L2P = 3,       /* targets receiving a local expansion        */
    ORC_NCOUNTERS = 4
};

typedef struct orc_tree orc_tree;

/* bbox: out[0]=xmin, out[1]=ymin, out[2]=ext (padded as sort-sources.cpp:71-82) */
int orc_bbox(const double *x, const double *y, int n, double *out);

/* Morton keys, 15 bits per axis, x in even bits (sort-sources.cpp:87-110) */
void orc_keys(const double *x, const double *y, int n, double xmin, double ymin, double ext, int *keys);

/* stable sort by key: perm[i] = original index of the i-th smallest (sort-sources.cpp:112) */
void orc_sort(const int *keys, int n, int *perm, int *keys_sorted);

/* Full Tree::build (upward.cpp:280-314): sort, BFS-numbered topology, upward pass. */
orc_tree *orc_tree_build(const double *xs, const double *ys, const double *vs, int n, int leafcap, int order);
void orc_tree_free(orc_tree *t);
int orc_tree_nnodes(const orc_tree *t);
int orc_tree_order(const orc_tree *t);

/* Export arrays (any pointer may be NULL):
 *   bbox[3], keys[n] sorted, perm[n], xyv[3n] sorted sources (x then y then v),
 *   node_se[2*nn], node_child[nn] (0 = leaf, else first child id), node_level[nn],
 *   node_cell[2*nn] (integer cell coordinates at the node's level),
 *   node_geom[4*nn] (xcom, ycom, r, mass), node_w[3*nn] (w, wx, wy),
 *   expansions[2*order*nn] ([Re a_1..a_p][Im a_1..a_p] per node) */
void orc_tree_export(const orc_tree *t, double *bbox, int *keys, int *perm, double *xyv,
                     int *node_se, int *node_child, int *node_level, int *node_cell,
                     double *node_geom, double *node_w, double *expansions);

/* treecode_potential restated (treecode-potential.cpp:91-106), order 12, leaf 512
 * unless overridden.  counters may be NULL.  nthreads<=0 -> OpenMP default. */
void orc_potential(double theta, const double *xs, const double *ys, const double *vs, int n,
                   const double *xt, const double *yt, int m, double *out,
                   int order, int leafcap, long long *counters, int nthreads);

/* Evaluate targets [t0, t1) against a prebuilt tree (same traversal). */
void orc_potential_tree(const orc_tree *t, double theta, const double *xt, const double *yt,
                        int t0, int t1, double *out, long long *counters, int nthreads);

/* treecode_force_mrag restated (treecode-force.cpp:201-236), order 24, leaf 192, TILE 8,
 * 32x32 blocks. */
void orc_force_mrag(double theta, const double *xs, const double *ys, const double *vs, int n,
                    const double *x0s, const double *y0s, const double *hs, int nblocks,
                    double *xr, double *yr, int order, int leafcap, long long *counters, int nthreads);

/* Evaluate blocks [b0, b1) against a prebuilt tree. Outputs are written for those blocks only
 * (xr/yr point at the full 32*32*nblocks arrays). */
void orc_force_tree(const orc_tree *t, double theta, const double *x0s, const double *y0s,
                    const double *hs, int b0, int b1, double *xr, double *yr,
                    long long *counters, int nthreads);

/* Gradient at arbitrary targets: per-target DFS and opening test of the potential
 * traversal with the force E2P / P2P terms (order 24, leaf 192 by default). */
void orc_force_targets(double theta, const double *xs, const double *ys, const double *vs, int n,
                       const double *xt, const double *yt, int m, double *xr, double *yr, int order, int leafcap,
                       long long *counters, int nthreads);
void orc_force_targets_tree(const orc_tree *t, double theta, const double *xt, const double *yt, int t0, int t1,
                            double *xr, double *yr, long long *counters, int nthreads);

/* FP64 direct sums at selected targets (main.cpp:206-245): out[k] for target idx[k]. */
void orc_direct_potential(const double *xs, const double *ys, const double *vs, int n,
                          const double *xt, const double *yt, const int *idx, int nidx,
                          double *out, int nthreads);
void orc_direct_force(const double *xs, const double *ys, const double *vs, int n,
                      const double *xt, const double *yt, const int *idx, int nidx,
                      double *outx, double *outy, int nthreads);

/* Building blocks exposed for closed-form tests. */
void orc_p2e(const double *x, const double *y, const double *v, int n, double xc, double yc,
             int order, double *re, double *im);
void orc_e2e(const double *x0, const double *y0, const double *mass, const double *const *re,
             const double *const *im, int nchild, int order, double *dre, double *dim);
double orc_e2p_potential(double rz, double iz, double mass, const double *re, const double *im, int order);
void orc_e2p_force(double mass, double rz, double iz, const double *re, const double *im, int order,
                   double *fx, double *fy);
void orc_e2l(double x0, double y0, double mass, const double *re, const double *im, int order,
             double *lre, double *lim);
void orc_l2p(double rz, double iz, const double *lre, const double *lim, int order, double *fx, double *fy);
double orc_binomial(int n, int k);

int orc_max_threads(void);
void orc_set_threads(int nthreads);

#ifdef __cplusplus
}
#endif
#endif
