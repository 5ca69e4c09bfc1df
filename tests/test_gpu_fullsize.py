"""Full-size checks at the benchmark shapes (the 2^24 potential and the
16,384-block force), through properties that do not need a CPU solve of that
size: error against the FP64 direct sum on sampled targets, identical
interaction counts in FP32 and FP64, keys and permutation against the oracle's
key function, and the tree invariants of the reference build (children
partition their parent's range in quadrant order, leaf iff at most `cap`
sources or level 15)."""
import numpy as np
import pytest

from conftest import rel_l2

pytestmark = pytest.mark.gpu

N24 = 1 << 24


@pytest.fixture(scope="module")
def c3(wl):
    return wl.mixed_plummer_potential(N24)


def _solve_potential(pkg, data, precision, order, leaf):
    xs, ys, vs, xt, yt = data
    ctx = pkg.Context(precision=precision, pot_order=order, pot_leaf=leaf, counters=True)
    d = {k: ctx.to_device(v) for k, v in dict(xs=xs, ys=ys, vs=vs, xt=xt, yt=yt).items()}
    out = ctx.empty(len(xt))
    ctx.potential(0.5, d["xs"], d["ys"], d["vs"], d["xt"], d["yt"], out=out, device=True)
    ctx.synchronize()
    return ctx, d, out.numpy()


def test_potential_2e24_accuracy_and_counts(pkg, c3):
    idx = np.sort(np.random.default_rng(7).choice(N24, 4096, replace=False)).astype(np.int32)
    counts = {}
    for prec in (pkg.FP32, pkg.FP64):
        ctx, d, out = _solve_potential(pkg, c3, prec, 10, 256)
        counts[prec] = ctx.stats()["counters"]
        ref = ctx.direct_potential(d["xs"], d["ys"], d["vs"], d["xt"], d["yt"], ctx.to_device(idx, np.int32))
        assert rel_l2(out[idx], ref) <= 1e-6
        ctx.close()
    # the FP32 path takes exactly the FP64 path's accept/open decisions
    for k in ("p2p", "e2p"):
        assert counts[pkg.FP32][k] == counts[pkg.FP64][k]


def test_tree_2e24_invariants(pkg, oracle, c3):
    xs, ys, vs, xt, yt = c3
    cap = 256
    ctx, _, _ = _solve_potential(pkg, (xs, ys, vs, xt[:4096], yt[:4096]), pkg.FP64, 10, cap)
    t = ctx.tree()
    ctx.close()
    keys, perm = t["keys"], t["perm"]
    n = len(xs)
    # keys: the oracle's key function on the permuted points; stable sort order
    assert np.array_equal(t["box"], oracle.bbox(xs, ys))
    assert np.array_equal(np.sort(perm), np.arange(n))
    okeys = oracle.keys(xs, ys, t["box"])
    assert np.array_equal(keys, okeys[perm])
    assert np.all(np.diff(keys.astype(np.int64)) >= 0)
    tie = np.diff(keys) == 0
    assert np.all(np.diff(perm)[tie] > 0)
    # topology
    s, e, child, level = (t["topo"][:, i] for i in range(4))
    assert s[0] == 0 and e[0] == n and level[0] == 0
    inner = np.nonzero(child)[0]
    leaves = np.nonzero(child == 0)[0]
    c0 = child[inner]
    assert np.array_equal(s[c0], s[inner]) and np.array_equal(e[c0 + 3], e[inner])
    for c in range(3):
        assert np.array_equal(e[c0 + c], s[c0 + c + 1])
    for c in range(4):
        assert np.array_equal(level[c0 + c], level[inner] + 1)
    assert np.all(e[inner] - s[inner] > cap)
    assert np.all((e[leaves] - s[leaves] <= cap) | (level[leaves] == 15))
    # every source sits in exactly one leaf
    assert np.sum(e[leaves] - s[leaves]) == n
    # geometry: radius non-negative, centres of non-empty nodes inside the box
    g = t["geom"]
    ne = e > s
    xmin, ymin, ext = t["box"]
    assert np.all(g[:, 2] >= 0)
    assert np.all((g[ne, 0] >= xmin) & (g[ne, 0] <= xmin + ext) & (g[ne, 1] >= ymin) & (g[ne, 1] <= ymin + ext))


def test_force_16384_blocks_accuracy_and_counts(pkg, wl):
    xs, ys, vs, x0s, y0s, hs = wl.gaussian16_force(1 << 23, nblocks=16384)
    m = len(x0s) * 1024
    idx = np.sort(np.random.default_rng(8).choice(m, 4096, replace=False)).astype(np.int32)
    b = idx // 1024
    tx = x0s[b] + (idx % 32) * hs[b]
    ty = y0s[b] + ((idx // 32) % 32) * hs[b]
    counts = {}
    for prec in (pkg.FP32, pkg.FP64):
        ctx = pkg.Context(precision=prec, force_order=10, force_leaf=192, counters=True)
        d = {k: ctx.to_device(v) for k, v in dict(xs=xs, ys=ys, vs=vs, x0s=x0s, y0s=y0s, hs=hs).items()}
        xr, yr = ctx.empty(m), ctx.empty(m)
        ctx.force(0.5, d["xs"], d["ys"], d["vs"], d["x0s"], d["y0s"], d["hs"], xr=xr, yr=yr, device=True)
        ctx.synchronize()
        counts[prec] = ctx.stats()["counters"]
        rx, ry = ctx.direct_force(d["xs"], d["ys"], d["vs"], ctx.to_device(tx), ctx.to_device(ty),
                                  ctx.to_device(np.arange(len(idx), dtype=np.int32), np.int32))
        fx, fy = xr.numpy()[idx], yr.numpy()[idx]
        err = np.sqrt(np.sum((fx - rx) ** 2) + np.sum((fy - ry) ** 2)) / np.sqrt(np.sum(rx ** 2) + np.sum(ry ** 2))
        assert err <= 1e-6
        ctx.close()
    assert counts[pkg.FP32]["l2p"] == m
    for k in ("p2p", "e2p", "e2l", "l2p"):
        assert counts[pkg.FP32][k] == counts[pkg.FP64][k]
