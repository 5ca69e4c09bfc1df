"""Pins the CPU oracle (the checker) without trusting its own formulas.

Each test checks an oracle function against something the mathematics or an
independent implementation fixes: closed forms, brute force, numpy's stable
sort, a reference-style recursive tree build, the translation identities of
multipole expansions, and the FP64 direct sum.
"""
import bisect
import math

import numpy as np
import pytest

from conftest import mrag_targets, rel_l2
from golden_util import load_fixture

LMAX = 15


def np_bbox(x, y):
    # sort-sources.cpp:44-82 + upward-common.ispc:132 (max starts at 0, min at 1e13)
    xmi = min(1e13, float(np.min(x)))
    ymi = min(1e13, float(np.min(y)))
    xma = max(0.0, float(np.max(x)))
    yma = max(0.0, float(np.max(y)))
    eps = 10000 * np.finfo(float).eps
    ext = max(xma - xmi, yma - ymi) * (1 + 2 * eps)
    return np.array([xmi - eps * ext, ymi - eps * ext, ext])


def np_keys(x, y, box):
    s = (1.0 / box[2]) * 32768.0
    qx = np.clip(np.floor((x - box[0]) * s), 0, 32767).astype(np.int64)
    qy = np.clip(np.floor((y - box[1]) * s), 0, 32767).astype(np.int64)
    k = np.zeros_like(qx)
    for b in range(15):
        k |= ((qx >> b) & 1) << (2 * b)
        k |= ((qy >> b) & 1) << (2 * b + 1)
    return k.astype(np.int32)


def test_bbox_closed_form(oracle, wl):
    st = wl.Stream(1)
    for lo, hi in [(0, 1), (-3, -1), (2, 5), (-1, 1)]:
        x, y = st.uniform(1000, lo, hi), st.uniform(1000, lo * 0.5, hi * 2)
        assert np.array_equal(oracle.bbox(x, y), np_bbox(x, y))
    # all-negative coordinates: the box still reaches 0 (max accumulators start at 0)
    b = oracle.bbox(np.array([-2.0, -1.0]), np.array([-1.0, -3.0]))
    assert b[0] + b[2] > 0


def test_bbox_degenerate(oracle):
    b = oracle.bbox(np.full(5, 0.3), np.full(5, 0.3))
    assert b[2] == 1.0  # all coincident: extent replaced by 1 (reference: undefined)


def test_keys_closed_form(oracle, wl):
    st = wl.Stream(2)
    x, y = st.uniform(20000), st.uniform(20000)
    box = oracle.bbox(x, y)
    assert np.array_equal(oracle.keys(x, y, box), np_keys(x, y, box))


def test_keys_cell_edge_fixture(oracle):
    fx = load_fixture("morton_edges")
    assert np.array_equal(oracle.bbox(fx["x"], fx["y"]), fx["box"])
    k = oracle.keys(fx["x"], fx["y"], fx["box"])
    assert np.array_equal(k, fx["keys"])
    # the fixture separates the compiled quantisation form from the source form
    assert int(fx["n_differ"]) > 0
    assert not np.array_equal(k, fx["keys_source_form"])


def test_sort_is_stable_argsort(oracle, wl):
    st = wl.Stream(3)
    k = (st.uniform(50000) * 2000).astype(np.int32) * 1000003 % (1 << 30)
    k = k.astype(np.int32)
    perm, ks = oracle.sort(k)
    ref = np.argsort(k, kind="stable")
    assert np.array_equal(perm, ref)
    assert np.array_equal(ks, k[ref])


def ref_style_tree(keys, n, cap):
    """build_tree (upward.cpp:223-277) recursion with sequential fetch_and_add numbering,
    std::lower_bound / std::upper_bound, children c=0..3 in order."""
    nodes = {0: dict(s=0, e=n, l=0, mask=0, child=0)}
    counter = [1]

    def rec(nid):
        nd = nodes[nid]
        s, e, l, mask = nd["s"], nd["e"], nd["l"], nd["mask"]
        if e - s <= cap or l + 1 > LMAX:
            return
        cb = counter[0]
        counter[0] += 4
        nd["child"] = cb
        for c in range(4):
            shift = 2 * (LMAX - l - 1)
            key1 = mask | (c << shift)
            key2 = key1 + (1 << shift) - 1
            imin = s if c == 0 else bisect.bisect_left(keys, key1, s, e)
            isup = e if c == 3 else bisect.bisect_right(keys, key2, s, e)
            nodes[cb + c] = dict(s=imin, e=isup, l=l + 1, mask=key1, child=0)
        for c in range(4):
            rec(cb + c)

    rec(0)
    return nodes


def canonical_bfs(nodes):
    """Renumber breadth first: children of the k-th internal node (level order) at 1+4k+c."""
    out, queue, nxt = [], [0], 1
    while queue:
        level_next = []
        for old in queue:
            nd = dict(nodes[old])
            if nd["child"]:
                kids = [nd["child"] + c for c in range(4)]
                nd["child"] = nxt
                nxt += 4
                level_next += kids
            out.append(nd)
        queue = level_next
    return out


@pytest.mark.parametrize("gen,cap", [("uniform", 16), ("gauss", 8), ("coincident", 4)])
def test_topology_vs_recursive_build(oracle, wl, gen, cap):
    if gen == "uniform":
        xs, ys, vs, _, _ = wl.uniform_potential(3000, seed=5)
    elif gen == "gauss":
        xs, ys, vs, *_ = wl.gaussian4_force(3000, seed=6)
    else:
        xs = np.concatenate([np.full(40, 0.3), wl.Stream(7).uniform(200)])
        ys = np.concatenate([np.full(40, 0.6), wl.Stream(8).uniform(200)])
        vs = np.ones(len(xs))
    t = oracle.Tree(xs, ys, vs, cap, 12)
    keys = [int(k) for k in t.keys]
    ref = canonical_bfs(ref_style_tree(keys, len(keys), cap))
    assert len(ref) == t.nnodes
    assert [d["s"] for d in ref] == list(t.se[:, 0])
    assert [d["e"] for d in ref] == list(t.se[:, 1])
    assert [d["child"] for d in ref] == list(t.child)
    assert [d["l"] for d in ref] == list(t.level)
    if gen == "coincident":
        assert t.level.max() == LMAX  # 40 coincident sources force level-15 leaves


def test_tree_invariants_and_geometry(oracle, wl):
    xs, ys, vs, *_ = wl.gaussian4_force(20000, seed=9)
    cap = 64
    t = oracle.Tree(xs, ys, vs, cap, 12)
    xsrt, ysrt, vsrt = t.xyv[: t.n], t.xyv[t.n: 2 * t.n], t.xyv[2 * t.n:]
    assert np.all(np.diff(t.keys) >= 0)
    assert np.array_equal(xsrt, xs[t.perm])
    xmin, ymin, ext = t.bbox
    for i in range(t.nnodes):
        s, e = t.se[i]
        lvl = t.level[i]
        h = ext / (1 << lvl)
        x0, y0 = xmin + h * t.cell[i, 0], ymin + h * t.cell[i, 1]
        if e > s:  # sources inside the cell (upward.cpp:240-243)
            assert np.all((xsrt[s:e] >= x0) & (xsrt[s:e] < x0 + h))
            assert np.all((ysrt[s:e] >= y0) & (ysrt[s:e] < y0 + h))
        xc, yc, r, m = t.geom[i]
        w = np.abs(vsrt[s:e])
        if t.child[i]:
            cb = t.child[i]
            assert t.se[cb, 0] == s and t.se[cb + 3, 1] == e
            assert all(t.se[cb + c, 1] == t.se[cb + c + 1, 0] for c in range(3))
            assert e - s > cap and lvl < LMAX
        else:
            assert e - s <= cap or lvl == LMAX
        if e > s:
            com_x = math.fsum(xsrt[s:e] * w) / math.fsum(w)
            com_y = math.fsum(ysrt[s:e] * w) / math.fsum(w)
            assert abs(xc - com_x) <= 1e-13 * max(1.0, abs(com_x))
            assert abs(yc - com_y) <= 1e-13 * max(1.0, abs(com_y))
            assert abs(m - math.fsum(vsrt[s:e])) <= 1e-12 * max(1.0, math.fsum(w))
            dist = np.sqrt((xsrt[s:e] - xc) ** 2 + (ysrt[s:e] - yc) ** 2).max()
            if t.child[i]:
                assert r >= dist * (1 - 1e-15)  # parent radius encloses its sources (upward.cpp:149-159)
                assert r <= float(np.float32(1.4143)) * h * (1 + 1e-15)
            else:
                assert r == np.sqrt(((xsrt[s:e] - xc) ** 2 + (ysrt[s:e] - yc) ** 2).max())
                assert r < 1.5 * h  # upward.cpp:87
                assert x0 <= xc < x0 + h and y0 <= yc < y0 + h  # upward.cpp:89-93


def test_leaf_sum_order_canonical(oracle, wl):
    # leaf sums use 64 lane-strided partials + a pairwise tree; check bit-exactly with numpy
    xs, ys, vs, _, _ = wl.uniform_potential(3000, seed=10)
    t = oracle.Tree(xs, ys, vs, 600, 12)
    n = t.n
    xsrt, vsrt = t.xyv[:n], t.xyv[2 * n:]
    for i in range(t.nnodes):
        if t.child[i]:
            continue
        s, e = t.se[i]
        p = np.zeros(64)
        for lane in range(64):
            acc = 0.0
            for j in range(s + lane, e, 64):
                acc = acc + vsrt[j]
            p[lane] = acc
        off = 32
        while off >= 1:
            p[:off] = p[:off] + p[off:2 * off]
            off //= 2
        assert t.geom[i, 3] == p[0]
        pw = np.zeros(64)
        px = np.zeros(64)
        for lane in range(64):
            aw = ax = 0.0
            for j in range(s + lane, e, 64):
                a = abs(vsrt[j])
                aw = aw + a
                ax = ax + xsrt[j] * a
            pw[lane], px[lane] = aw, ax
        off = 32
        while off >= 1:
            pw[:off] = pw[:off] + pw[off:2 * off]
            px[:off] = px[:off] + px[off:2 * off]
            off //= 2
        assert t.w[i, 0] == pw[0] and t.w[i, 1] == px[0]
        assert t.geom[i, 0] == px[0] / pw[0]


def test_p2e_single_source_closed_form(oracle):
    for d in [0.3 + 0.1j, -0.05 + 0.2j, 1e-3 - 2e-3j]:
        v = -0.7
        xs, ys = np.array([0.5 + d.real]), np.array([0.25 + d.imag])
        a = oracle.p2e(xs, ys, np.array([v]), 0.5, 0.25, 24)
        d = complex(xs[0] - 0.5, ys[0] - 0.25)  # the offset as the kernel sees it
        k = np.arange(1, 25)
        ref = -v * d ** k / k
        # complex relative error: the power recursion rounds ~k times
        assert np.all(np.abs(a - ref) <= 4e-16 * k * np.abs(ref))


def test_e2e_equals_p2e_about_parent(oracle, wl):
    xs, ys, vs, *_ = wl.gaussian4_force(8000, seed=11)
    for p in (12, 24):
        t = oracle.Tree(xs, ys, vs, 32, p)
        n = t.n
        xsrt, ysrt, vsrt = t.xyv[:n], t.xyv[n:2 * n], t.xyv[2 * n:]
        checked = 0
        for i in range(t.nnodes):
            if not t.child[i]:
                continue
            s, e = t.se[i]
            a_ref = oracle.p2e(xsrt[s:e], ysrt[s:e], vsrt[s:e], t.geom[i, 0], t.geom[i, 1], p)
            a = t.exp[i, 0] + 1j * t.exp[i, 1]
            # M2M is exact for finite p (compare coefficient-wise, scaled by r^k)
            r = t.geom[i, 2]
            sc = r ** np.arange(1, p + 1)
            err = np.abs((a - a_ref) / sc).max() / max(np.abs(a_ref / sc).max(), 1e-300)
            assert err <= 1e-11, (i, err)
            checked += 1
        assert checked > 10


def test_e2p_converges_to_direct(oracle, wl):
    st = wl.Stream(12)
    m = 500
    ang = st.uniform(m) * 2 * np.pi
    rad = 0.1 * np.sqrt(st.uniform(m))
    xs, ys = rad * np.cos(ang), rad * np.sin(ang)
    vs = st.uniform(m, -1, 1)
    xc = float(np.sum(xs * np.abs(vs)) / np.sum(np.abs(vs)))
    yc = float(np.sum(ys * np.abs(vs)) / np.sum(np.abs(vs)))
    r = float(np.sqrt(((xs - xc) ** 2 + (ys - yc) ** 2).max()))
    z = complex(xc + 4 * r * 0.6, yc + 4 * r * 0.8)
    direct = oracle.direct_potential(xs, ys, vs, np.array([z.real]), np.array([z.imag]), np.array([0], np.int32))[0]
    errs = []
    for p in (4, 8, 12, 16, 24, 32):
        a = oracle.p2e(xs, ys, vs, xc, yc, p)
        val = oracle.e2p_potential(complex(z.real - xc, z.imag - yc), float(np.sum(vs)), a)
        errs.append(abs(val - direct))
        bound = np.sum(np.abs(vs)) * (0.25 ** (p + 1)) / (p + 1) / (1 - 0.25) + 1e-12
        assert errs[-1] <= bound, (p, errs[-1], bound)
    assert errs[-1] < errs[0] * 1e-9  # down to the regularisation + rounding floor


def test_force_e2p_e2l_l2p_vs_direct(oracle, wl):
    st = wl.Stream(13)
    m = 400
    xs = 0.02 * st.normal(m) + 0.8
    ys = 0.02 * st.normal(m) + 0.7
    vs = st.uniform(m, -1, 1)
    w = np.abs(vs)
    xc, yc = float(np.sum(xs * w) / w.sum()), float(np.sum(ys * w) / w.sum())
    p = 24
    a = oracle.p2e(xs, ys, vs, xc, yc, p)
    mass = float(vs.sum())
    # local expansion about a brick centre far away, evaluated at points around it
    bc = complex(0.2, 0.25)
    b = oracle.e2l(complex(xc - bc.real, yc - bc.imag), mass, a)
    pts = [bc + complex(dx, dy) for dx in (-0.01, 0.0, 0.007) for dy in (-0.006, 0.01)]
    tx = np.array([q.real for q in pts])
    ty = np.array([q.imag for q in pts])
    dx, dy = oracle.direct_force(xs, ys, vs, tx, ty, np.arange(len(pts), dtype=np.int32))
    for k, q in enumerate(pts):
        fx, fy = oracle.e2p_force(complex(q.real - xc, q.imag - yc), mass, a)
        gx, gy = oracle.l2p(complex(q.real - bc.real, q.imag - bc.imag), b)
        scale = math.hypot(dx[k], dy[k])
        assert abs(fx - dx[k]) <= 1e-12 * scale and abs(fy - dy[k]) <= 1e-12 * scale
        assert abs(gx - dx[k]) <= 1e-10 * scale and abs(gy - dy[k]) <= 1e-10 * scale


def test_l2p_is_polynomial_derivative(oracle, wl):
    st = wl.Stream(14)
    p = 12
    b = st.uniform(p + 1, -1, 1) + 1j * st.uniform(p + 1, -1, 1)
    z = complex(0.3, -0.2)
    fx, fy = oracle.l2p(z, b)
    deriv = sum(l * b[l] * z ** (l - 1) for l in range(1, p + 1))
    assert abs(fx - deriv.real) <= 1e-14 and abs(fy + deriv.imag) <= 1e-14


def test_potential_theta0_is_direct_sum(oracle, wl):
    xs, ys, vs, xt, yt = wl.uniform_potential(2000, seed=15)
    out, cnt = oracle.potential(0.0, xs, ys, vs, xt, yt)
    ref = oracle.direct_potential(xs, ys, vs, xt, yt, np.arange(len(xt), dtype=np.int32))
    assert rel_l2(out, ref) <= 1e-13
    assert cnt[0] == len(xs) * len(xt) and cnt[1] == 0


def test_potential_accuracy_vs_direct(oracle, wl):
    xs, ys, vs, xt, yt = wl.uniform_potential(20000, seed=16)
    idx = np.arange(7, len(xt), 433, dtype=np.int32)  # main.cpp:203-204 sampling
    ref = oracle.direct_potential(xs, ys, vs, xt, yt, idx)
    for theta, tol in [(0.3, 1e-9), (0.5, 1e-6)]:
        out, _ = oracle.potential(theta, xs, ys, vs, xt, yt)
        assert rel_l2(out[idx], ref) <= tol


def test_force_theta0_and_accuracy(oracle, wl):
    xs, ys, vs, x0s, y0s, hs = wl.gaussian4_force(5000, nbx=2, h=1.0 / 64, seed=17)
    tx, ty = mrag_targets(x0s, y0s, hs)
    idx = np.arange(7, len(tx), 13, dtype=np.int32)
    dx, dy = oracle.direct_force(xs, ys, vs, tx, ty, idx)
    fx, fy, cnt = oracle.force_mrag(0.0, xs, ys, vs, x0s, y0s, hs)
    assert rel_l2(fx[idx], dx) <= 1e-13 and rel_l2(fy[idx], dy) <= 1e-13
    assert cnt[0] == len(xs) * len(tx) and cnt[1] == 0 and cnt[2] == 0
    fx, fy, cnt = oracle.force_mrag(0.5, xs, ys, vs, x0s, y0s, hs)
    assert rel_l2(fx[idx], dx) <= 1e-10 and rel_l2(fy[idx], dy) <= 1e-10
    assert cnt[2] > 0 and cnt[3] == len(tx)


def test_output_layout_mrag(oracle, wl):
    # point (ix, iy) of block b is written at ix + 32*(iy + 32*b) (README.md:36-37)
    xs, ys, vs, x0s, y0s, hs = wl.gaussian4_force(3000, nbx=2, h=1.0 / 64, seed=18)
    fx, fy, _ = oracle.force_mrag(0.5, xs, ys, vs, x0s, y0s, hs)
    b, ix, iy = 3, 5, 17
    t = np.array([x0s[b] + ix * hs[b]]), np.array([y0s[b] + iy * hs[b]])
    dx, dy = oracle.direct_force(xs, ys, vs, t[0], t[1], np.array([0], np.int32))
    k = ix + 32 * (iy + 32 * b)
    assert abs(fx[k] - dx[0]) <= 1e-9 * abs(dx[0]) and abs(fy[k] - dy[0]) <= 1e-9 * abs(dy[0])


def test_binomials_exact(oracle):
    fx = load_fixture("binomials")
    for n in range(64):
        for k in range(n + 1):
            exact = (int(fx["hi"][n, k]) << 32) | int(fx["lo"][n, k])
            assert exact == math.comb(n, k)
            assert oracle.binomial(n, k) == float(exact)


def test_bc_constants_round_to_nearest(oracle):
    # -1/l printed by `bc -l` with 20 digits parses to the nearest double of -1/l
    fx = load_fixture("neg_inv")
    assert np.array_equal(fx["bc20"], fx["nearest"])


def test_force_targets_vs_direct(oracle, wl):
    xs, ys, vs, xt, yt = wl.uniform_potential(4000, seed=19)
    st = wl.Stream(20)
    xt, yt = st.uniform(700, -0.2, 1.2), st.uniform(700, -0.2, 1.2)
    idx = np.arange(len(xt), dtype=np.int32)
    dx, dy = oracle.direct_force(xs, ys, vs, xt, yt, idx)
    fx, fy, cnt = oracle.force_targets(0.0, xs, ys, vs, xt, yt)
    assert rel_l2(fx, dx) <= 1e-13 and rel_l2(fy, dy) <= 1e-13
    assert cnt[0] == len(xs) * len(xt)
    fx, fy, cnt = oracle.force_targets(0.5, xs, ys, vs, xt, yt)
    assert rel_l2(fx, dx) <= 1e-9 and rel_l2(fy, dy) <= 1e-9 and cnt[1] > 0


# ------------------------------------------------------------------ reference-style traversals
# The interaction lists (and so the counters the GPU must match) are pinned by a
# literal restatement of the reference's own loops on a tree built the reference's
# way: the recursive build above (reference numbering, not breadth-first), node
# geometry from the leaf setup and parent update of upward.cpp:60-160 in numpy,
# expansions by direct P2E of each node's sources about its centre of mass (equal
# to the M2M translations for a finite order), then the depth-first loops of
# treecode-potential.cpp:25-86 and treecode-force.cpp:94-199 with their
# accept / E2L / E2P tests and push order.  Empty nodes contribute exactly zero
# and are not counted (the oracle's convention).
EPS_REG = 10 * np.finfo(float).eps


def ref_style_geometry(keys, xs, ys, vs, cap, p, ext):
    nodes = ref_style_tree(keys, len(keys), cap)
    for nid in sorted(nodes, reverse=True):  # children carry larger ids than their parent
        nd = nodes[nid]
        s, e = nd["s"], nd["e"]
        if e == s:
            nd.update(w=0.0, wx=0.0, wy=0.0, mass=0.0)
            continue
        if not nd["child"]:
            a = np.abs(vs[s:e])
            nd.update(mass=float(np.sum(vs[s:e])), w=float(np.sum(a)), wx=float(np.sum(xs[s:e] * a)),
                      wy=float(np.sum(ys[s:e] * a)))
            nd["xcom"], nd["ycom"] = nd["wx"] / nd["w"], nd["wy"] / nd["w"]
            nd["r"] = math.sqrt(float(np.max((xs[s:e] - nd["xcom"]) ** 2 + (ys[s:e] - nd["ycom"]) ** 2)))
        else:
            kids = [nodes[nd["child"] + c] for c in range(4)]
            for k in ("mass", "w", "wx", "wy"):
                nd[k] = sum(kid[k] for kid in kids)
            nd["xcom"], nd["ycom"] = nd["wx"] / nd["w"], nd["wy"] / nd["w"]
            rr = max(kid["r"] + math.hypot(nd["xcom"] - kid["xcom"], nd["ycom"] - kid["ycom"])
                     for kid in kids if kid["w"])
            nd["r"] = min(rr, float(np.float32(1.4143)) * ext / (1 << nd["l"]))
        d = (xs[s:e] - nd["xcom"]) + 1j * (ys[s:e] - nd["ycom"])
        k = np.arange(1, p + 1)
        nd["a"] = -np.sum(vs[s:e, None] * d[:, None] ** k[None, :], axis=0) / k
    return nodes


def ref_style_potential(nodes, xs, ys, vs, theta, xt, yt):
    out = np.zeros(len(xt))
    npairs = ne2p = 0
    for i, (x, y) in enumerate(zip(xt, yt)):
        res = 0.0
        stack = [0]
        while stack:
            nd = nodes[stack.pop()]
            if nd["e"] == nd["s"]:
                continue
            r2 = (x - nd["xcom"]) ** 2 + (y - nd["ycom"]) ** 2
            if nd["r"] * nd["r"] < theta * theta * r2:
                z = complex(x - nd["xcom"], y - nd["ycom"])
                u = (1 / z) ** np.arange(1, len(nd["a"]) + 1)
                res += nd["mass"] * math.log(r2) / 2 + float(np.sum(nd["a"] * u).real)
                ne2p += 1
            elif not nd["child"]:
                s, e = nd["s"], nd["e"]
                dx, dy = x - xs[s:e], y - ys[s:e]
                res += 0.5 * float(np.sum(np.log(dx * dx + dy * dy + EPS_REG) * vs[s:e]))
                npairs += e - s
            else:
                stack += [nd["child"] + c for c in range(4)]
        out[i] = res
    return out, npairs, ne2p


def ref_style_force_block(nodes, xs, ys, vs, theta, x0, y0, h, p):
    tile = 8
    fx, fy = np.zeros(1024), np.zeros(1024)
    counts = np.zeros(4, np.int64)
    rbrick = 1.4142135623730951 * h * (tile - 1) * 0.5
    ix, iy = np.arange(64) % tile, np.arange(64) // tile
    l = np.arange(1, p + 1)
    binom = np.array([[math.comb(ll + k - 1, k - 1) for k in range(1, p + 1)] for ll in range(1, p + 1)], float)
    for by in range(0, 32, tile):
        for bx in range(0, 32, tile):
            xb, yb = x0 + h * (bx + 0.5 * (tile - 1)), y0 + h * (by + 0.5 * (tile - 1))
            rx, ry = np.zeros(64), np.zeros(64)
            loc = np.zeros(p, complex)  # b_1..b_p
            tx, ty = x0 + (bx + ix) * h, y0 + (by + iy) * h
            stack = [0]
            while stack:
                nd = nodes[stack.pop()]
                if nd["e"] == nd["s"]:
                    continue
                xc, yc, r = nd["xcom"], nd["ycom"], nd["r"]
                dist = math.sqrt((xb - xc) ** 2 + (yb - yc) ** 2)
                with np.errstate(divide="ignore", invalid="ignore"):
                    ratio = np.float64(dist) / np.float64(r)  # IEEE: a one-source leaf has r = 0
                if (ratio - 1) > (1 / theta) and rbrick <= r:
                    z0 = complex(xc - xb, yc - yb)
                    sgn = (-1.0) ** l
                    inner = -nd["mass"] / l + binom @ (sgn * nd["a"] * z0 ** (-l.astype(float)))
                    loc += z0 ** (-l.astype(float)) * inner
                    counts[2] += 1
                    continue
                xt = max(x0 + bx * h, min(x0 + (bx + tile - 1) * h, xc))
                yt = max(y0 + by * h, min(y0 + (by + tile - 1) * h, yc))
                if r * r < theta * theta * ((xt - xc) ** 2 + (yt - yc) ** 2):
                    z = (tx - xc) + 1j * (ty - yc)
                    dphi = nd["mass"] / z - np.sum(l * nd["a"] * z[:, None] ** (-(l + 1.0)), axis=1)
                    rx += dphi.real
                    ry -= dphi.imag
                    counts[1] += 64
                elif not nd["child"]:
                    s, e = nd["s"], nd["e"]
                    dx, dy = tx[:, None] - xs[None, s:e], ty[:, None] - ys[None, s:e]
                    den = dx * dx + (dy * dy + EPS_REG)
                    rx += np.sum(vs[None, s:e] * dx / den, axis=1)
                    ry += np.sum(vs[None, s:e] * dy / den, axis=1)
                    counts[0] += 64 * (e - s)
                else:
                    stack += [nd["child"] + c for c in range(4)]
            zl = (tx - xb) + 1j * (ty - yb)
            d = np.sum(l * loc * zl[:, None] ** (l - 1.0), axis=1)
            rx += d.real
            ry -= d.imag
            counts[3] += 64
            k = bx + ix + 32 * (by + iy)
            fx[k], fy[k] = rx, ry
    return fx, fy, counts


@pytest.mark.parametrize("theta", [0.3, 0.5, 0.7])
def test_potential_reference_style_traversal(oracle, wl, theta):
    xs, ys, vs, xt, yt = wl.mixed_plummer_potential(3000, seed=23)
    p, cap = 12, 24
    t = oracle.Tree(xs, ys, vs, cap, p)
    n = t.n
    xsrt, ysrt, vsrt = t.xyv[:n], t.xyv[n:2 * n], t.xyv[2 * n:]
    nodes = ref_style_geometry([int(k) for k in t.keys], xsrt, ysrt, vsrt, cap, p, t.bbox[2])
    sel = np.arange(0, len(xt), 11)
    ref, npairs, ne2p = ref_style_potential(nodes, xsrt, ysrt, vsrt, theta, xt[sel], yt[sel])
    out, cnt = t.potential(theta, xt[sel], yt[sel])
    assert (int(cnt[0]), int(cnt[1])) == (npairs, ne2p)
    assert rel_l2(out, ref) <= 1e-12


@pytest.mark.parametrize("theta", [0.3, 0.5, 0.7])
def test_force_reference_style_traversal(oracle, wl, theta):
    xs, ys, vs, x0s, y0s, hs = wl.gaussian16_force(4000, nblocks=6, seed=24)
    p, cap = 12, 16
    t = oracle.Tree(xs, ys, vs, cap, p)
    n = t.n
    xsrt, ysrt, vsrt = t.xyv[:n], t.xyv[n:2 * n], t.xyv[2 * n:]
    nodes = ref_style_geometry([int(k) for k in t.keys], xsrt, ysrt, vsrt, cap, p, t.bbox[2])
    fx, fy, cnt = t.force(theta, x0s, y0s, hs)
    total = np.zeros(4, np.int64)
    for b in range(len(x0s)):
        rx, ry, c = ref_style_force_block(nodes, xsrt, ysrt, vsrt, theta, x0s[b], y0s[b], hs[b], p)
        total += c
        sl = slice(1024 * b, 1024 * (b + 1))
        assert rel_l2(fx[sl], rx) <= 1e-12 and rel_l2(fy[sl], ry) <= 1e-12, b
    assert tuple(int(c) for c in cnt) == tuple(int(c) for c in total)
    assert total[2] > 0 and total[1] > 0 and total[0] > 0


def test_parallel_build_equals_serial(oracle, wl):
    # the tree build runs its sort, topology levels, leaves and parent levels in
    # parallel; every step is per element or per node, so any thread count gives
    # the serial build bit for bit
    xs, ys, vs, _, _ = wl.mixed_plummer_potential(300000, seed=25)
    n0 = oracle.max_threads()
    trees = []
    for nt in (1, 4):
        oracle.set_threads(nt)
        trees.append(oracle.Tree(xs, ys, vs, 48, 12))
    oracle.set_threads(n0)
    a, b = trees
    for k in ("keys", "perm", "se", "child", "level", "geom", "w", "exp"):
        assert np.array_equal(getattr(a, k), getattr(b, k)), k
