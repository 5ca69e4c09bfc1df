"""HIP path vs the CPU oracle, through the C ABI (libtreecodes2d.so).

Bars: keys, permutation, topology and node geometry bit-exact; expansion
coefficients within 1e-13 relative; FP64 potentials and velocities within
1e-12 relative L2 with identical interaction counts; the FP32-bulk path within
1e-6 relative L2 of the FP64 direct sum and of the FP64 treecode.
"""
import os

import numpy as np
import pytest

from conftest import compare_tree, mrag_targets, rel_l2

pytestmark = pytest.mark.gpu

FP64_TOL = 1e-12
FP32_TOL = 1e-6


def _device_sort(pkg, ctx, x, y):
    xd, yd = ctx.to_device(x), ctx.to_device(y)
    n = len(x)
    box = ctx.empty(4)
    keys = ctx.empty(n, np.int32)
    perm = ctx.empty(n, np.int32)
    ctx.morton_sort_device(xd, yd, box, keys, perm)
    ctx.synchronize()
    return box.numpy(), keys.numpy(), perm.numpy()


# up to 16384 points the sort runs in one workgroup, above it in tiles of 5120
# whose passes look back over the counts earlier tiles publish; 881 tiles are
# more than the GPU holds at once (768), so the last tiles start after the first finish
@pytest.mark.parametrize("n", [1, 63, 64, 4095, 4096, 4097, 16384, 16385, 20480, 20481, 100003, 262144, 262145,
                               1000003, 1100 * 4096 + 9])
def test_morton_sort_bit_exact(pkg, oracle, wl, ctx64, n):
    st = wl.Stream(100 + n)
    x, y = st.uniform(n), st.uniform(n)
    box, keys, perm = _device_sort(pkg, ctx64, x, y)
    obox = oracle.bbox(x, y)
    assert np.array_equal(box[:3], obox)
    okeys = oracle.keys(x, y, obox)
    operm, oks = oracle.sort(okeys)
    assert np.array_equal(keys, oks)
    assert np.array_equal(perm, operm)


def test_morton_sort_duplicates_stable(pkg, oracle, wl, ctx64):
    # many equal keys: the permutation must keep input order within a key (stable sort)
    st = wl.Stream(5)
    n = 50000
    x = np.floor(st.uniform(n) * 8) / 8 + 0.01
    y = np.floor(st.uniform(n) * 8) / 8 + 0.01
    box, keys, perm = _device_sort(pkg, ctx64, x, y)
    okeys = oracle.keys(x, y, oracle.bbox(x, y))
    operm, oks = oracle.sort(okeys)
    assert np.array_equal(perm, operm)
    assert np.array_equal(keys, oks)


@pytest.mark.parametrize("n", [3 * 4096 * 16 + 77, 80 * 4096 + 5])
def test_morton_sort_two_keys_stable(pkg, oracle, wl, ctx64, n):
    # two distinct keys over many tiles: each 1024-key run of the sort holds
    # hundreds of equal digits, so ranks within a run reach their maximum
    # (39 and 65 tiles, all resident at once, so every tile looks back)
    st = wl.Stream(6)
    pick = st.uniform(n) < 0.5
    x = np.where(pick, 0.1, 0.9)
    y = np.where(pick, 0.2, 0.8)
    box, keys, perm = _device_sort(pkg, ctx64, x, y)
    okeys = oracle.keys(x, y, oracle.bbox(x, y))
    operm, oks = oracle.sort(okeys)
    assert np.array_equal(perm, operm)
    assert np.array_equal(keys, oks)


def test_morton_sort_cell_edges(pkg, oracle, ctx64):
    from golden_util import load_fixture
    fx = load_fixture("morton_edges")
    box, keys, perm = _device_sort(pkg, ctx64, fx["x"], fx["y"])
    assert np.array_equal(box[:3], fx["box"])
    assert np.array_equal(np.sort(keys), np.sort(fx["keys"]))
    inv = np.empty_like(perm)
    inv[perm] = np.arange(len(perm))
    assert np.array_equal(keys[inv], fx["keys"])


def _compare_tree(ctx, ot, order):
    compare_tree(ctx.tree(), ot)


@pytest.mark.parametrize("case", ["uniform", "gauss", "small", "tiny", "plummer"])
def test_tree_potential_config(pkg, oracle, wl, ctx64, case):
    if case == "uniform":
        xs, ys, vs, xt, yt = wl.uniform_potential(65536)
    elif case == "gauss":
        xs, ys, vs, *_ = wl.gaussian4_force(50000)
        xt, yt = xs.copy(), ys.copy()
    elif case == "small":
        xs, ys, vs, xt, yt = wl.uniform_potential(300, seed=3)
    elif case == "tiny":
        xs, ys, vs, xt, yt = wl.uniform_potential(1, seed=4)
    else:
        xs, ys, vs, xt, yt = wl.mixed_plummer_potential(40000, seed=9)
    out = ctx64.potential(0.5, xs, ys, vs, xt, yt)
    ot = oracle.Tree(xs, ys, vs, 512, 12)
    _compare_tree(ctx64, ot, 12)
    ref, cnt = ot.potential(0.5, xt, yt)
    assert rel_l2(out, ref) <= FP64_TOL
    st = ctx64.stats()["counters"]
    assert st["p2p"] == cnt[0] and st["e2p"] == cnt[1]


@pytest.mark.parametrize("theta", [0.3, 0.5, 0.7])
def test_potential_fp64_vs_oracle(pkg, oracle, wl, ctx64, theta):
    xs, ys, vs, xt, yt = wl.uniform_potential(65536)
    out = ctx64.potential(theta, xs, ys, vs, xt, yt)
    ref, cnt = oracle.potential(theta, xs, ys, vs, xt, yt)
    assert rel_l2(out, ref) <= FP64_TOL
    st = ctx64.stats()["counters"]
    assert (st["p2p"], st["e2p"]) == (cnt[0], cnt[1])


@pytest.mark.parametrize("order,leaf", [(6, 256), (9, 192), (10, 256), (14, 128), (20, 64), (32, 512)])
def test_potential_orders_vs_oracle(pkg, oracle, wl, order, leaf):
    # every compiled expansion order (tuned configurations use 10 / 256 and 10 / 64)
    ctx = pkg.Context(counters=True, pot_order=order, pot_leaf=leaf)
    xs, ys, vs, xt, yt = wl.mixed_plummer_potential(30000, seed=40 + order)
    out = ctx.potential(0.5, xs, ys, vs, xt, yt)
    ref, cnt = oracle.potential(0.5, xs, ys, vs, xt, yt, order=order, leafcap=leaf)
    assert rel_l2(out, ref) <= FP64_TOL
    st = ctx.stats()["counters"]
    assert (st["p2p"], st["e2p"]) == (cnt[0], cnt[1])
    ctx.close()


@pytest.mark.parametrize("order,leaf", [(8, 64), (9, 96), (12, 128), (12, 192), (16, 192), (20, 128)])
def test_force_orders_vs_oracle(pkg, oracle, wl, order, leaf):
    # E2L runs four nodes per wave for orders <= 16 and two above
    ctx = pkg.Context(counters=True, force_order=order, force_leaf=leaf)
    xs, ys, vs, x0s, y0s, hs = wl.gaussian16_force(40000, nblocks=64, seed=60 + order)
    fx, fy = ctx.force(0.5, xs, ys, vs, x0s, y0s, hs)
    rx, ry, cnt = oracle.force_mrag(0.5, xs, ys, vs, x0s, y0s, hs, order=order, leafcap=leaf)
    assert rel_l2(fx, rx) <= FP64_TOL
    assert rel_l2(fy, ry) <= FP64_TOL
    st = ctx.stats()["counters"]
    assert (st["p2p"], st["e2p"], st["e2l"], st["l2p"]) == tuple(int(c) for c in cnt)
    ctx.close()


@pytest.mark.parametrize("precision,chunks", [(0, 0), (1, 0), (0, 2), (1, 2), (0, 1)])
def test_force_many_near_leaves(pkg, oracle, wl, precision, chunks):
    # leaves of at most 8 sources under coarse blocks (bricks a quarter of the
    # box wide): every brick of the coarse block has hundreds of P2P leaves,
    # more than the traversal sums itself, so the rest goes to 64-leaf chunks
    # summed by other waves; the fine blocks need no chunks.  A pool of one or
    # two chunks (TC_OPT_FORCE_CHUNKS) fills up, and the bricks left over take
    # the repeated-traversal path, which must skip the leaves already summed.
    st = wl.Stream(31)
    n = 30000
    xs, ys, vs = st.uniform(n), st.uniform(n), st.uniform(n, -1, 1)
    x0s = np.array([0.0, 0.31, 0.62, 0.0])
    y0s = np.array([0.0, 0.27, 0.55, 0.5])
    hs = np.array([1.0 / 32, 1.0 / 1024, 1.0 / 2048, 1.0 / 64])
    ctx = pkg.Context(precision=precision, counters=True, force_order=12, force_leaf=8)
    ctx.set(pkg.OPT_FORCE_CHUNKS, chunks)
    fx, fy = ctx.force(0.5, xs, ys, vs, x0s, y0s, hs)
    rx, ry, cnt = oracle.force_mrag(0.5, xs, ys, vs, x0s, y0s, hs, order=12, leafcap=8)
    tol = FP64_TOL if precision == 0 else FP32_TOL
    assert rel_l2(fx, rx) <= tol
    assert rel_l2(fy, ry) <= tol
    c = ctx.stats()["counters"]
    assert (c["p2p"], c["e2p"], c["e2l"], c["l2p"]) == tuple(int(v) for v in cnt)
    ctx.close()


@pytest.mark.parametrize("precision", [0, 1])
@pytest.mark.parametrize("inline", [1, 16, 4096])
def test_force_inline_tiling(pkg, oracle, wl, precision, inline):
    # the inline P2P count (a tiling option) only moves leaves between the
    # traversal and the chunk waves: counts equal the oracle's, results agree
    # with it and with the default tiling up to summation order
    st = wl.Stream(32)
    n = 30000
    xs, ys, vs = st.uniform(n), st.uniform(n), st.uniform(n, -1, 1)
    x0s = np.array([0.0, 0.31, 0.62, 0.0])
    y0s = np.array([0.0, 0.27, 0.55, 0.5])
    hs = np.array([1.0 / 32, 1.0 / 1024, 1.0 / 2048, 1.0 / 64])
    ctx = pkg.Context(precision=precision, counters=True, force_order=12, force_leaf=8)
    base = ctx.force(0.5, xs, ys, vs, x0s, y0s, hs)
    ctx.set(pkg.OPT_FORCE_INLINE, inline)
    fx, fy = ctx.force(0.5, xs, ys, vs, x0s, y0s, hs)
    rx, ry, cnt = oracle.force_mrag(0.5, xs, ys, vs, x0s, y0s, hs, order=12, leafcap=8)
    tol = FP64_TOL if precision == 0 else FP32_TOL
    assert rel_l2(fx, rx) <= tol and rel_l2(fy, ry) <= tol
    same = 1e-14 if precision == 0 else 1e-7
    assert rel_l2(fx, base[0]) <= same and rel_l2(fy, base[1]) <= same
    c = ctx.stats()["counters"]
    assert (c["p2p"], c["e2p"], c["e2l"], c["l2p"]) == tuple(int(v) for v in cnt)
    for bad in (-1, 4097):
        with pytest.raises(pkg.TreecodeError, match="invalid option"):
            ctx.set(pkg.OPT_FORCE_INLINE, bad)
    ctx.close()


@pytest.mark.parametrize("precision", [0, 1])
def test_potential_xcd_chunk_tiling(pkg, wl, precision):
    # the traversal's workgroup-to-XCD chunking (a tiling option) only changes
    # which workgroup runs which targets: results are bit-identical
    xs, ys, vs, xt, yt = wl.mixed_plummer_potential(200000, seed=33)
    ctx = pkg.Context(precision=precision, pot_order=10, pot_leaf=128)
    ref = ctx.potential(0.5, xs, ys, vs, xt, yt)
    for chunk in (1, 2, 64, 256):
        ctx.set(pkg.OPT_POT_XCD_CHUNK, chunk)
        assert np.array_equal(ctx.potential(0.5, xs, ys, vs, xt, yt), ref), chunk
    for bad in (-1, 3, 512):
        with pytest.raises(pkg.TreecodeError, match="invalid option"):
            ctx.set(pkg.OPT_POT_XCD_CHUNK, bad)
    ctx.close()


@pytest.mark.parametrize("n", [(1 << 19) + 3, 1 << 21])
@pytest.mark.parametrize("precision", [0, 1])
def test_host_targets_uploaded_beside_build(pkg, wl, precision, n):
    # host target arrays of 4 MiB and more (2^19 points and more) are uploaded by
    # a helper thread on the second stream while the sources build: the results
    # equal the device-pointer solve's, and a non-finite target is still refused
    xs, ys, vs, xt, yt = wl.mixed_plummer_potential(n, seed=35)
    ctx = pkg.Context(precision=precision, pot_order=10, pot_leaf=192)
    host = ctx.potential(0.5, xs, ys, vs, xt, yt)
    d = {k: ctx.to_device(v) for k, v in dict(xs=xs, ys=ys, vs=vs, xt=xt, yt=yt).items()}
    out = ctx.empty(n)
    ctx.potential(0.5, d["xs"], d["ys"], d["vs"], d["xt"], d["yt"], out=out, device=True)
    ctx.synchronize()
    assert np.array_equal(host, out.numpy())
    bad = yt.copy()
    bad[n - 5] = np.inf
    with pytest.raises(pkg.TreecodeError, match="non-finite"):
        ctx.potential(0.5, xs, ys, vs, xt, bad)
    again = ctx.potential(0.5, xs, ys, vs, xt, yt)
    assert np.array_equal(again, host)
    ctx.close()


def test_potential_targets_elsewhere(pkg, oracle, wl, ctx64):
    # targets outside the source box, in random order, fewer than sources
    st = wl.Stream(77)
    xs, ys, vs = st.uniform(30000), st.uniform(30000), st.uniform(30000, -1, 1)
    xt, yt = st.uniform(7777, -1.0, 2.0), st.uniform(7777, -0.5, 1.5)
    out = ctx64.potential(0.5, xs, ys, vs, xt, yt)
    ref, _ = oracle.potential(0.5, xs, ys, vs, xt, yt)
    assert rel_l2(out, ref) <= FP64_TOL


def test_potential_direct_sum_and_theta0(pkg, oracle, wl, ctx64):
    xs, ys, vs, xt, yt = wl.uniform_potential(3000, seed=11)
    idx = np.arange(len(xt), dtype=np.int32)
    direct = oracle.direct_potential(xs, ys, vs, xt, yt, idx)
    # theta = 0 never accepts: the treecode becomes the direct sum (P2P only)
    out0 = ctx64.potential(0.0, xs, ys, vs, xt, yt)
    assert rel_l2(out0, direct) <= 1e-13
    out = ctx64.potential(0.5, xs, ys, vs, xt, yt)
    assert rel_l2(out, direct) <= 1e-6


@pytest.mark.parametrize("theta", [0.3, 0.5, 0.7])
def test_force_fp64_vs_oracle(pkg, oracle, wl, ctx64, theta):
    xs, ys, vs, x0s, y0s, hs = wl.gaussian4_force(65536, nbx=16, h=1.0 / 512)
    fx, fy = ctx64.force(theta, xs, ys, vs, x0s, y0s, hs)
    rx, ry, cnt = oracle.force_mrag(theta, xs, ys, vs, x0s, y0s, hs)
    assert rel_l2(fx, rx) <= FP64_TOL
    assert rel_l2(fy, ry) <= FP64_TOL
    st = ctx64.stats()["counters"]
    assert (st["p2p"], st["e2p"], st["e2l"], st["l2p"]) == tuple(int(c) for c in cnt)


def test_force_tree_and_mixed_blocks(pkg, oracle, wl, ctx64):
    xs, ys, vs, x0s, y0s, hs = wl.gaussian16_force(60000, nblocks=96, seed=21)
    fx, fy = ctx64.force(0.5, xs, ys, vs, x0s, y0s, hs)
    ot = oracle.Tree(xs, ys, vs, 192, 24)
    _compare_tree(ctx64, ot, 24)
    rx, ry, cnt = ot.force(0.5, x0s, y0s, hs)
    assert rel_l2(fx, rx) <= FP64_TOL
    assert rel_l2(fy, ry) <= FP64_TOL


def test_force_direct_sum(pkg, oracle, wl, ctx64):
    xs, ys, vs, x0s, y0s, hs = wl.gaussian4_force(20000, nbx=4, h=1.0 / 128, seed=31)
    fx, fy = ctx64.force(0.5, xs, ys, vs, x0s, y0s, hs)
    tx, ty = mrag_targets(x0s, y0s, hs)
    idx = np.arange(7, len(tx), 7, dtype=np.int32)
    dx, dy = oracle.direct_force(xs, ys, vs, tx, ty, idx)
    assert rel_l2(fx[idx], dx) <= 1e-10
    assert rel_l2(fy[idx], dy) <= 1e-10
    f0x, f0y = ctx64.force(0.0, xs, ys, vs, x0s, y0s, hs)
    assert rel_l2(f0x[idx], dx) <= 1e-13
    assert rel_l2(f0y[idx], dy) <= 1e-13


def test_dropin_abi(pkg, oracle, wl):
    xs, ys, vs, xt, yt = wl.uniform_potential(20000, seed=12)
    out = pkg.treecode_potential(0.5, xs, ys, vs, xt, yt)
    ref, _ = oracle.potential(0.5, xs, ys, vs, xt, yt)
    assert rel_l2(out, ref) <= FP64_TOL
    xs, ys, vs, x0s, y0s, hs = wl.gaussian4_force(20000, nbx=4, h=1.0 / 128, seed=13)
    fx, fy = pkg.treecode_force_mrag(0.5, xs, ys, vs, x0s, y0s, hs)
    rx, ry, _ = oracle.force_mrag(0.5, xs, ys, vs, x0s, y0s, hs)
    assert rel_l2(fx, rx) <= FP64_TOL and rel_l2(fy, ry) <= FP64_TOL


def test_host_arrays_staged_copies(pkg, wl, ctx32):
    # host arrays of 4 MiB and more travel through the pinned staging buffers in
    # 16 MiB chunks (2.1M doubles = 2 chunks): results equal the device-array path
    xs, ys, vs, xt, yt = wl.uniform_potential(2_100_000, seed=71)
    host = ctx32.potential(0.5, xs, ys, vs, xt, yt)
    d = {k: ctx32.to_device(v) for k, v in dict(xs=xs, ys=ys, vs=vs, xt=xt, yt=yt).items()}
    out = ctx32.empty(len(xt))
    ctx32.potential(0.5, d["xs"], d["ys"], d["vs"], d["xt"], d["yt"], out=out, device=True)
    ctx32.synchronize()
    assert np.array_equal(host, out.numpy())
    xs, ys, vs, x0s, y0s, hs = wl.gaussian4_force(600_000, nbx=50, h=1.0 / 2048, seed=72)
    x0s, y0s, hs = x0s[:2100].copy(), y0s[:2100].copy(), hs[:2100].copy()  # 2.15M targets
    fx, fy = ctx32.force(0.5, xs, ys, vs, x0s, y0s, hs)
    d = {k: ctx32.to_device(v) for k, v in dict(xs=xs, ys=ys, vs=vs, x0s=x0s, y0s=y0s, hs=hs).items()}
    xr, yr = ctx32.empty(len(x0s) * 1024), ctx32.empty(len(x0s) * 1024)
    ctx32.force(0.5, d["xs"], d["ys"], d["vs"], d["x0s"], d["y0s"], d["hs"], xr=xr, yr=yr, device=True)
    ctx32.synchronize()
    assert np.array_equal(fx, xr.numpy()) and np.array_equal(fy, yr.numpy())


def test_host_arrays_packed_copies(pkg, wl, ctx32):
    # a call whose host inputs fit one 16 MiB staging buffer together sends them in one DMA
    # (and its outputs come back in one): results equal the device-array path, for small
    # calls and for ones just under the buffer (5 x 400,000 doubles = 16.0 MB in; force
    # outputs 2 x 921,600 doubles = 14.7 MB out)
    for n in (3000, 400_000):
        xs, ys, vs, xt, yt = wl.uniform_potential(n, seed=73)
        host = ctx32.potential(0.5, xs, ys, vs, xt, yt)
        d = {k: ctx32.to_device(v) for k, v in dict(xs=xs, ys=ys, vs=vs, xt=xt, yt=yt).items()}
        out = ctx32.empty(len(xt))
        ctx32.potential(0.5, d["xs"], d["ys"], d["vs"], d["xt"], d["yt"], out=out, device=True)
        ctx32.synchronize()
        assert np.array_equal(host, out.numpy()), n
        res = ctx32.potential_resident(0.5, xt, yt)
        assert np.array_equal(host, res), n
        fx, fy = ctx32.force_targets(0.5, xs, ys, vs, xt, yt)
        gx, gy = ctx32.empty(len(xt)), ctx32.empty(len(xt))
        ctx32.force_targets(0.5, d["xs"], d["ys"], d["vs"], d["xt"], d["yt"], xr=gx, yr=gy, device=True)
        ctx32.synchronize()
        assert np.array_equal(fx, gx.numpy()) and np.array_equal(fy, gy.numpy()), n
    xs, ys, vs, x0s, y0s, hs = wl.gaussian4_force(200_000, nbx=32, h=1.0 / 2048, seed=74)
    for nb in (7, 900):
        bx, by, bh = x0s[:nb].copy(), y0s[:nb].copy(), hs[:nb].copy()
        fx, fy = ctx32.force(0.5, xs, ys, vs, bx, by, bh)
        d = {k: ctx32.to_device(v) for k, v in dict(xs=xs, ys=ys, vs=vs, x0s=bx, y0s=by, hs=bh).items()}
        xr, yr = ctx32.empty(nb * 1024), ctx32.empty(nb * 1024)
        ctx32.force(0.5, d["xs"], d["ys"], d["vs"], d["x0s"], d["y0s"], d["hs"], xr=xr, yr=yr, device=True)
        ctx32.synchronize()
        assert np.array_equal(fx, xr.numpy()) and np.array_equal(fy, yr.numpy()), nb


def test_deterministic_and_resident(pkg, wl, ctx64):
    xs, ys, vs, xt, yt = wl.uniform_potential(65536, seed=14)
    a = ctx64.potential(0.5, xs, ys, vs, xt, yt)
    b = ctx64.potential(0.5, xs, ys, vs, xt, yt)
    assert np.array_equal(a, b)
    c = ctx64.potential_resident(0.5, xt, yt)
    assert np.array_equal(a, c)
    # new targets against the resident tree == full solve with those targets
    xt2, yt2 = xt[::3].copy(), yt[::3].copy()
    d = ctx64.potential_resident(0.5, xt2, yt2)
    e = ctx64.potential(0.5, xs, ys, vs, xt2, yt2)
    assert np.array_equal(d, e)


def test_empty_and_error_paths(pkg, wl, ctx64):
    out = ctx64.potential(0.5, np.zeros(0), np.zeros(0), np.zeros(0), np.ones(5), np.ones(5))
    assert np.array_equal(out, np.zeros(5))
    # a failed build (the target sort already queued on the context's second stream) leaves the
    # context usable, and the next solve equals one made before the failure; 40,000 points use
    # the tiled sorts, 1000 the one-workgroup ones
    for n in (40_000, 1000):
        xs, ys, vs, xt, yt = wl.uniform_potential(n, seed=15)
        before = ctx64.potential(0.5, xs, ys, vs, xt, yt)
        x10 = xs[10]
        xs[10] = np.nan
        with pytest.raises(pkg.TreecodeError):
            ctx64.potential(0.5, xs, ys, vs, xt, yt)
        xs[10] = x10
        assert np.array_equal(ctx64.potential(0.5, xs, ys, vs, xt, yt), before), n
    xs[10] = 0.5
    ctx64.potential(0.5, xs, ys, vs, xt, yt)
    # non-finite targets are reported too (flagged on the device by the target sort:
    # the one-workgroup sort for 1000 targets, the one-launch passes for 40,000)
    for n in (1000, 40_000):
        xs, ys, vs, xt, yt = wl.uniform_potential(n, seed=16)
        bad = yt.copy()
        bad[777] = np.inf
        with pytest.raises(pkg.TreecodeError, match="non-finite"):
            ctx64.potential(0.5, xs, ys, vs, xt, bad)
        with pytest.raises(pkg.TreecodeError, match="non-finite"):
            ctx64.force_targets(0.5, xs, ys, vs, xt, bad)
        good = ctx64.potential(0.5, xs, ys, vs, xt, yt)
        assert np.all(np.isfinite(good))


def test_failed_build_leaves_no_resident_tree(pkg, wl):
    # an FP64 tree is resident; a rebuild in FP32 fails at its census (non-finite source)
    # after it has overwritten the tree's sizes and source arrays: resident calls and the
    # export must refuse (no kernel runs on the half-built tree), and a new build works
    ctx = pkg.Context(precision=pkg.FP64)
    xs, ys, vs, xt, yt = wl.uniform_potential(20000, seed=18)
    ctx.potential(0.5, xs, ys, vs, xt, yt)
    ctx.set(pkg.OPT_PRECISION, pkg.FP32)
    bad = xs.copy()
    bad[3] = np.nan
    with pytest.raises(pkg.TreecodeError, match="non-finite"):
        ctx.potential(0.5, bad, ys, vs, xt, yt)
    x0s, y0s, hs = np.array([0.2]), np.array([0.3]), np.array([1.0 / 256])
    for call in (lambda: ctx.potential_resident(0.5, xt, yt), lambda: ctx.force_resident(0.5, x0s, y0s, hs),
                 ctx.tree):
        with pytest.raises(pkg.TreecodeError, match="no resident tree"):
            call()
    out = ctx.potential(0.5, xs, ys, vs, xt, yt)
    assert np.array_equal(ctx.potential_resident(0.5, xt, yt), out)
    ctx.close()


def test_leaf_capacity_limit(pkg):
    # source index + leaf capacity must stay below 2^31: capacities above 2^30 are refused
    ctx = pkg.Context()
    for key in (pkg.OPT_POT_LEAF, pkg.OPT_FORCE_LEAF):
        ctx.set(key, 1 << 30)
        for bad in ((1 << 30) + 1, (1 << 31) - 1, 0, -5):
            with pytest.raises(pkg.TreecodeError, match="invalid option"):
                ctx.set(key, bad)
    ctx.close()


def test_coincident_sources(pkg, oracle, ctx64, ctx32):
    # zero-radius nodes (all sources at one point) next to ordinary ones
    n = 2000
    xs = np.concatenate([np.full(n, 0.25), np.linspace(0.0, 1.0, 3000)])
    ys = np.concatenate([np.full(n, 0.75), np.linspace(1.0, 0.0, 3000) ** 2])
    vs = np.concatenate([np.linspace(-1, 1, n) + 0.3, np.full(3000, -0.2)])
    xt = np.array([0.0, 0.25, 1.0, 0.3, 0.6])
    yt = np.array([0.0, 0.75, 1.0, 0.7, 0.2])
    out = ctx64.potential(0.5, xs, ys, vs, xt, yt)
    ref, _ = oracle.potential(0.5, xs, ys, vs, xt, yt)
    assert rel_l2(out, ref) <= FP64_TOL
    out32 = ctx32.potential(0.5, xs, ys, vs, xt, yt)
    assert np.all(np.isfinite(out32)) and rel_l2(out32, ref) <= FP32_TOL
    x0s, y0s, hs = np.array([0.2, 0.5]), np.array([0.7, 0.1]), np.array([1 / 512, 1 / 256])
    fx, fy = ctx32.force(0.5, xs, ys, vs, x0s, y0s, hs)
    rx, ry, _ = oracle.force_mrag(0.5, xs, ys, vs, x0s, y0s, hs)
    assert np.all(np.isfinite(fx)) and rel_l2(fx, rx) <= FP32_TOL and rel_l2(fy, ry) <= FP32_TOL


# ------------------------------------------------------------------ FP32 bulk
def test_fp32_potential_accuracy(pkg, oracle, wl, ctx32, ctx64):
    xs, ys, vs, xt, yt = wl.mixed_plummer_potential(1 << 18, seed=16)
    out32 = ctx32.potential(0.5, xs, ys, vs, xt, yt)
    c32 = ctx32.stats()["counters"]
    out64 = ctx64.potential(0.5, xs, ys, vs, xt, yt)
    c64 = ctx64.stats()["counters"]
    assert (c32["p2p"], c32["e2p"]) == (c64["p2p"], c64["e2p"])  # same interaction lists
    rng = np.random.default_rng(0)
    idx = np.sort(rng.choice(len(xt), 4096, replace=False)).astype(np.int32)
    direct = oracle.direct_potential(xs, ys, vs, xt, yt, idx)
    assert rel_l2(out32[idx], direct) <= FP32_TOL
    assert rel_l2(out32, out64) <= FP32_TOL


def test_fp32_wave_local_expansions(pkg, oracle, wl, ctx32, ctx64):
    # FP32: far nodes that a whole wave accepts are evaluated through one local expansion
    # about the wave's target centre; interaction counts stay those of the per-target walk
    xs, ys, vs, xt, yt = wl.mixed_plummer_potential(1 << 18, seed=21)
    out32 = ctx32.potential(0.5, xs, ys, vs, xt, yt)
    c32 = ctx32.stats()["counters"]
    out64 = ctx64.potential(0.5, xs, ys, vs, xt, yt)
    c64 = ctx64.stats()["counters"]
    assert (c32["p2p"], c32["e2p"]) == (c64["p2p"], c64["e2p"])
    assert c32["m2l_wave"] > 0 and 4 * c32["e2p_wave"] > c32["e2p"]
    assert c64["m2l_wave"] == 0 and c64["e2p_wave"] == 0
    assert rel_l2(out32, out64) <= FP32_TOL
    # degenerate waves: a single target, 64 copies of one point (zero radius), a wave of
    # targets outside the box, a ring of targets around the sources
    st = wl.Stream(22)
    ang = 2 * np.pi * st.uniform(3000)
    cases = [(xt[:1], yt[:1]), (np.full(64, 0.3), np.full(64, 0.7)),
             (1.6 + 0.01 * st.uniform(500), -0.4 + 0.01 * st.uniform(500)),
             (0.5 + 0.9 * np.cos(ang), 0.5 + 0.9 * np.sin(ang))]
    for i, (tx, ty) in enumerate(cases):
        a = ctx32.potential(0.5, xs, ys, vs, tx, ty)
        b = ctx64.potential(0.5, xs, ys, vs, tx, ty)
        assert rel_l2(a, b) <= FP32_TOL, i
        d = oracle.direct_potential(xs, ys, vs, tx, ty, np.arange(len(tx), dtype=np.int32))
        assert rel_l2(a, d) <= FP32_TOL, i


def test_fp32_force_accuracy(pkg, oracle, wl, ctx32, ctx64):
    xs, ys, vs, x0s, y0s, hs = wl.gaussian4_force(1 << 18, nbx=16, h=1.0 / 512, seed=17)
    fx32, fy32 = ctx32.force(0.5, xs, ys, vs, x0s, y0s, hs)
    fx, fy = ctx64.force(0.5, xs, ys, vs, x0s, y0s, hs)
    assert rel_l2(fx32, fx) <= FP32_TOL
    assert rel_l2(fy32, fy) <= FP32_TOL
    tx, ty = mrag_targets(x0s, y0s, hs)
    idx = np.arange(7, len(tx), 433, dtype=np.int32)
    dx, dy = oracle.direct_force(xs, ys, vs, tx, ty, idx)
    assert rel_l2(fx32[idx], dx) <= FP32_TOL
    assert rel_l2(fy32[idx], dy) <= FP32_TOL


def test_gpu_direct_sums_vs_oracle(pkg, oracle, wl, ctx64):
    # the GPU direct-sum kernels used by bench.py to measure error at full size
    xs, ys, vs, xt, yt = wl.uniform_potential(5000, seed=19)
    idx = np.arange(0, 5000, 7, dtype=np.int32)
    d = [ctx64.to_device(a) for a in (xs, ys, vs, xt, yt)]
    di = ctx64.to_device(idx, np.int32)
    g = ctx64.direct_potential(*d, di)
    assert rel_l2(g, oracle.direct_potential(xs, ys, vs, xt, yt, idx)) <= 1e-14
    gx, gy = ctx64.direct_force(*d, di)
    rx, ry = oracle.direct_force(xs, ys, vs, xt, yt, idx)
    assert rel_l2(gx, rx) <= 1e-14 and rel_l2(gy, ry) <= 1e-14


def _float_dropin(pkg):
    """The float build's drop-in pair under the reference's names (float/lib2d-treecodes-*.so)."""
    import ctypes
    libdir = os.path.join(os.path.dirname(pkg.LIB_PATH), "float")
    F = ctypes.POINTER(ctypes.c_float)
    fl, it = ctypes.c_float, ctypes.c_int
    pot = ctypes.CDLL(os.path.join(libdir, "lib2d-treecodes-potential.so")).treecode_potential
    pot.argtypes = [fl, F, F, F, it, F, F, it, F]
    pot.restype = None
    frc = ctypes.CDLL(os.path.join(libdir, "lib2d-treecodes-force.so")).treecode_force_mrag
    frc.argtypes = [fl, F, F, F, it, F, F, F, it, F, F]
    frc.restype = None
    return pot, frc


def test_float_abi(pkg, oracle, wl):
    # the reference's real=float build: treecode_potential / treecode_force_mrag over float
    # arrays, called by the reference names in the float library, and the _float-suffixed
    # pair of the double library; both run the FP32-bulk path and give the same results
    import ctypes
    pot, frc = _float_dropin(pkg)
    fp = lambda a: a.ctypes.data_as(ctypes.POINTER(ctypes.c_float))  # noqa: E731
    xs, ys, vs, xt, yt = (a.astype(np.float32) for a in wl.uniform_potential(20000, seed=61))
    out = np.empty(len(xt), np.float32)
    pot(0.5, fp(xs), fp(ys), fp(vs), len(xs), fp(xt), fp(yt), len(xt), fp(out))
    ref, _ = oracle.potential(0.5, *(a.astype(np.float64) for a in (xs, ys, vs, xt, yt)))
    assert rel_l2(out, ref) <= FP32_TOL
    assert np.array_equal(out, pkg.treecode_potential_float(0.5, xs, ys, vs, xt, yt))
    xs, ys, vs, x0s, y0s, hs = (a.astype(np.float32) for a in wl.gaussian4_force(20000, nbx=4, h=1.0 / 128, seed=62))
    fx, fy = np.empty(len(x0s) * 1024, np.float32), np.empty(len(x0s) * 1024, np.float32)
    frc(0.5, fp(xs), fp(ys), fp(vs), len(xs), fp(x0s), fp(y0s), fp(hs), len(x0s), fp(fx), fp(fy))
    rx, ry, _ = oracle.force_mrag(0.5, *(a.astype(np.float64) for a in (xs, ys, vs, x0s, y0s, hs)))
    assert rel_l2(fx, rx) <= FP32_TOL and rel_l2(fy, ry) <= FP32_TOL
    gx, gy = pkg.treecode_force_mrag_float(0.5, xs, ys, vs, x0s, y0s, hs)
    assert np.array_equal(fx, gx) and np.array_equal(fy, gy)


REF_MAIN = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "oracle", "_ref",
                        "main_dropin")


@pytest.mark.skipif(not os.path.exists(REF_MAIN), reason="oracle/_ref/main_dropin not built (needs the reference)")
def test_reference_main_runs_unchanged(pkg, tmp_path):
    # the reference's main.cpp (built unchanged against this library by oracle/Makefile)
    # on synthetic files at its eleven dataset paths: it calls treecode_potential /
    # treecode_force_mrag and checks sampled results against its own FP64 direct sum
    # with its own rule (tol 1e-8, velocities 1e-6; asserts abort on a failure)
    import subprocess
    from treecodes2d_amd import harness
    harness.write_synthetic_datasets(str(tmp_path))
    r = subprocess.run([REF_MAIN, "0.2", "1e-8"], cwd=str(tmp_path), capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stdout[-2000:] + r.stderr[-2000:]
    assert r.stdout.count("TEST PASSED.") == 11, r.stdout[-2000:]
    assert "not found" not in r.stdout


def test_velocity_wrapper(pkg, oracle, wl):
    # main.cpp:136-143 turns stored velocities into the library's convention by
    # xref = -2 pi v, yref = 2 pi u; the wrapper inverts it: u = grad_y / (2 pi),
    # v = -grad_x / (2 pi) of the oracle's treecode_force_mrag, within the FP64 bar
    xs, ys, vs, x0s, y0s, hs = wl.gaussian4_force(20000, nbx=4, h=1.0 / 128, seed=63)
    u, v = pkg.treecode_velocity_mrag(0.5, xs, ys, vs, x0s, y0s, hs)
    rx, ry, _ = oracle.force_mrag(0.5, xs, ys, vs, x0s, y0s, hs)
    assert rel_l2(u, ry / (2 * np.pi)) <= FP64_TOL and rel_l2(v, -rx / (2 * np.pi)) <= FP64_TOL
    # and the harness's reading of a stored velocity file maps (u, v) back onto the force
    assert rel_l2(-2 * np.pi * v, rx) <= FP64_TOL and rel_l2(2 * np.pi * u, ry) <= FP64_TOL


def test_log64_accuracy(pkg, ctx64):
    # theta = 0 with one source of value 2 at the origin: out = log(r^2 + 10 eps) exactly as
    # computed by the P2P loop, whose log is the library's own FP64 routine
    r2 = np.concatenate([np.logspace(-14, 3, 4000), 1 + np.linspace(-0.3, 0.4, 3000)])
    xt = np.sqrt(r2)
    yt = np.zeros_like(xt)
    out = ctx64.potential(0.0, np.zeros(1), np.zeros(1), np.full(1, 2.0), xt, yt)
    ref = np.log(xt * xt + 10 * np.finfo(float).eps)
    ulp = np.abs(out - ref) / np.spacing(np.abs(ref))
    assert ulp.max() <= 4, ulp.max()


@pytest.mark.parametrize("precision", [0, 1])
def test_force_targets(pkg, oracle, wl, ctx64, ctx32, precision):
    xs, ys, vs, *_ = wl.gaussian4_force(60000, seed=64)
    st = wl.Stream(65)
    xt, yt = st.uniform(9000, 0.1, 0.9), st.uniform(9000, 0.1, 0.9)
    ctx = ctx32 if precision else ctx64
    fx, fy = ctx.force_targets(0.5, xs, ys, vs, xt, yt)
    rx, ry, cnt = oracle.force_targets(0.5, xs, ys, vs, xt, yt)
    tol = FP32_TOL if precision else FP64_TOL
    assert rel_l2(fx, rx) <= tol and rel_l2(fy, ry) <= tol
    c = ctx.stats()["counters"]
    assert (c["p2p"], c["e2p"]) == (cnt[0], cnt[1])
    if not precision:
        gx, gy = pkg.treecode_force(0.5, xs, ys, vs, xt, yt)
        assert np.array_equal(gx, fx) and np.array_equal(gy, fy)


def test_target_order_invariance(pkg, wl, ctx64, ctx32):
    # FP64: each lane accumulates in its own target's DFS order, so results do not depend on
    # how targets are grouped into waves: a permutation of the inputs permutes the outputs
    # bit for bit.  FP32: a leaf that few lanes of a wave need is summed source-parallel, and
    # far nodes a whole wave accepts go through the wave's local expansion; both depend on
    # the grouping, so a permutation (which regroups targets whose sort keys tie) changes the
    # rounding of those terms: bounded by FP32 rounding of the far field, 1e-7 (a tenth of the
    # FP32 path's error bar).
    xs, ys, vs, xt, yt = wl.mixed_plummer_potential(50000, seed=66)
    perm = np.random.default_rng(1).permutation(len(xt))
    a = ctx64.potential(0.5, xs, ys, vs, xt, yt)
    b = ctx64.potential(0.5, xs, ys, vs, xt[perm], yt[perm])
    assert np.array_equal(a[perm], b)
    a = ctx32.potential(0.5, xs, ys, vs, xt, yt)
    b = ctx32.potential(0.5, xs, ys, vs, xt[perm], yt[perm])
    print("FP32 permutation difference %.3e" % rel_l2(b, a[perm]))
    assert rel_l2(b, a[perm]) < 1e-7


def test_profile_phases_and_cycles(pkg, wl):
    # tc_get_profile: the build's phases (ms) after a call that built a tree, and the
    # traversal's clock cycles per category; the parts never exceed the whole
    xs, ys, vs, xt, yt = wl.mixed_plummer_potential(200000, seed=71)
    for prec in (pkg.FP64, pkg.FP32):
        ctx = pkg.Context(precision=prec, counters=True, profile=True, pot_order=10, pot_leaf=128)
        ctx.potential(0.5, xs, ys, vs, xt, yt)
        pr, st = ctx.profile(), ctx.stats()
        b = pr["build_ms"]
        assert all(v > 0 for v in b.values()), b
        assert sum(b.values()) <= st["times_ms"]["build"] + 1e-3
        c = pr["cycles"]
        assert c["p2p"] > 0 and c["e2p"] > 0 and c["total"] > 0
        assert c["p2p"] + c["e2p"] + c["local"] + c["l2p"] <= c["total"]
        if prec == pkg.FP32:
            assert c["local"] > 0 and c["l2p"] > 0  # wave-level local expansions
        else:
            assert c["local"] == 0 and c["l2p"] == 0
        # a resident solve builds nothing: no build phases
        ctx.potential_resident(0.5, xt[:5000], yt[:5000])
        assert all(v == 0 for v in ctx.profile()["build_ms"].values())
        ctx.close()
    xs, ys, vs, x0s, y0s, hs = wl.gaussian4_force(100000, nbx=8, h=1.0 / 256, seed=72)
    ctx = pkg.Context(counters=True, profile=True, force_order=12, force_leaf=64)
    ctx.force(0.5, xs, ys, vs, x0s, y0s, hs)
    c = ctx.profile()["cycles"]
    assert min(c["p2p"], c["e2p"], c["local"], c["l2p"]) > 0
    assert c["p2p"] + c["e2p"] + c["local"] + c["l2p"] <= c["total"]
    # counting off: no cycles
    ctx.set(pkg.OPT_COUNTERS, 0)
    ctx.force(0.5, xs, ys, vs, x0s, y0s, hs)
    assert ctx.profile()["cycles"]["total"] == 0
    ctx.close()


@pytest.mark.parametrize("precision", [0, 1])
def test_force_host_ranges_equal_device(pkg, wl, precision):
    # a host-array force solve of 4096 blocks is evaluated in block ranges whose
    # outputs are copied back while the next range runs; the result equals the
    # one-launch device-pointer solve bit for bit
    xs, ys, vs, x0s, y0s, hs = wl.gaussian4_force(1 << 18)  # 64 x 64 = 4096 blocks
    assert len(x0s) == 4096
    ctx = pkg.Context(precision=precision, force_order=8, force_leaf=64)
    hx, hy = ctx.force(0.5, xs, ys, vs, x0s, y0s, hs)
    d = {k: ctx.to_device(v) for k, v in dict(xs=xs, ys=ys, vs=vs, x0s=x0s, y0s=y0s, hs=hs).items()}
    xr, yr = ctx.empty(len(x0s) * 1024), ctx.empty(len(x0s) * 1024)
    ctx.force(0.5, d["xs"], d["ys"], d["vs"], d["x0s"], d["y0s"], d["hs"], xr=xr, yr=yr, device=True)
    ctx.synchronize()
    assert np.array_equal(hx, xr.numpy()) and np.array_equal(hy, yr.numpy())
    assert np.all(np.isfinite(hx)) and np.any(hx != 0)
    ctx.close()
