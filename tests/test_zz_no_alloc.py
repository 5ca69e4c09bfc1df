"""tc_potential_noalloc, the potential solve that allocates nothing: every
buffer lives in the library's static device scratch (64 MiB per GPU).

Checked through the C ABI against tc_potential with device pointers on the
same inputs (the same kernels in the same order, so bit-identical results and
equal interaction counts), on two calls in a row with the second smaller, in
both precisions; plus the limits (oversized calls are refused before any work,
a tree that outgrows the scratch is refused at the census and leaves the next
call correct), zero sources, and no change of free device memory across calls.
"""
import ctypes

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _hip_runtime():
    """The libamdhip64 the treecode library runs on (already loaded in this process)."""
    with open("/proc/self/maps") as f:
        paths = {l.split()[-1] for l in f if "libamdhip64.so" in l}
    assert paths, "HIP runtime not loaded"
    L = ctypes.CDLL(sorted(paths)[0])
    L.hipMemGetInfo.argtypes = [ctypes.POINTER(ctypes.c_size_t), ctypes.POINTER(ctypes.c_size_t)]
    return L


def _free_bytes(hip):
    free, total = ctypes.c_size_t(), ctypes.c_size_t()
    assert hip.hipMemGetInfo(ctypes.byref(free), ctypes.byref(total)) == 0
    return free.value


def _inputs(ctx, wl, n, seed):
    xs, ys, vs, xt, yt = wl.mixed_plummer_potential(n, seed=seed)
    d = {k: ctx.to_device(v) for k, v in dict(xs=xs, ys=ys, vs=vs, xt=xt, yt=yt).items()}
    return d


def _solve_both(pkg, ref, na, d, theta):
    m = len(d["xt"])
    out_ref = ref.empty(m)
    ref.potential(theta, d["xs"], d["ys"], d["vs"], d["xt"], d["yt"], out=out_ref, device=True)
    ref.synchronize()
    out_na = na.empty(m)
    na.potential_noalloc(theta, d["xs"], d["ys"], d["vs"], d["xt"], d["yt"], out_na)
    return out_ref.numpy(), out_na.numpy(), ref.stats()["counters"], na.stats()["counters"]


# two calls in a row, the second smaller (the multi-kernel sort path, then the
# one-workgroup path), in both precisions with their tuned-style and reference settings
@pytest.mark.parametrize("precision,order,leaf", [(0, 12, 512), (1, 10, 192), (0, 10, 64)])
def test_noalloc_equals_device_call(pkg, wl, precision, order, leaf):
    kw = dict(precision=precision, pot_order=order, pot_leaf=leaf, counters=True)
    ref = pkg.Context(**kw)
    na = pkg.Context(**kw)
    for n, theta, seed in ((150000, 0.5, 11), (9000, 0.7, 12)):
        print("noalloc parity: precision %d order %d leaf %d n=m=%d theta %.1f" % (precision, order, leaf, n, theta),
              flush=True)
        d = _inputs(ref, wl, n, seed)
        a, b, ca, cb = _solve_both(pkg, ref, na, d, theta)
        assert np.all(np.isfinite(b))
        assert np.array_equal(a, b), (n, np.max(np.abs(a - b)))
        assert ca == cb, n
    ref.close()
    na.close()


def test_noalloc_allocates_nothing(pkg, wl):
    hip = _hip_runtime()
    ctx = pkg.Context(precision=pkg.FP32, pot_order=10, pot_leaf=192, counters=True)
    small = _inputs(ctx, wl, 3000, 21)
    large = _inputs(ctx, wl, pkg.NOALLOC_MAX_POINTS, 22)
    out_s, out_l = ctx.empty(3000), ctx.empty(pkg.NOALLOC_MAX_POINTS)
    # warm-up: both sort paths once (the runtime may set up kernel resources on a first launch)
    print("noalloc memory: warm-up calls n=m=%d and 3000" % pkg.NOALLOC_MAX_POINTS, flush=True)
    ctx.potential_noalloc(0.5, large["xs"], large["ys"], large["vs"], large["xt"], large["yt"], out_l)
    ctx.potential_noalloc(0.5, small["xs"], small["ys"], small["vs"], small["xt"], small["yt"], out_s)
    before = _free_bytes(hip)
    print("noalloc memory: n=m=%d then 3000" % pkg.NOALLOC_MAX_POINTS, flush=True)
    ctx.potential_noalloc(0.5, large["xs"], large["ys"], large["vs"], large["xt"], large["yt"], out_l)
    ctx.potential_noalloc(0.5, small["xs"], small["ys"], small["vs"], small["xt"], small["yt"], out_s)
    after = _free_bytes(hip)
    assert after == before
    assert np.all(np.isfinite(out_l.numpy()))
    ctx.close()


def test_noalloc_limits(pkg, wl):
    ctx = pkg.Context(precision=pkg.FP64, pot_order=12, pot_leaf=512)
    big = pkg.NOALLOC_MAX_POINTS + 1
    d = _inputs(ctx, wl, big, 31)
    out = ctx.empty(big)
    # more points than the scratch is sized for: refused before any work is queued
    print("noalloc limits: n=m=%d refused" % big, flush=True)
    with pytest.raises(pkg.TreecodeError, match="at most 262144"):
        ctx.potential_noalloc(0.5, d["xs"], d["ys"], d["vs"], d["xt"], d["yt"], out)
    with pytest.raises(pkg.TreecodeError, match="at most 262144"):
        ctx.potential_noalloc(0.5, d["xs"], d["ys"], d["vs"], d["xt"], d["yt"], out, n=1000)
    # a tree larger than what the points leave of the scratch: refused at the census
    n = pkg.NOALLOC_MAX_POINTS
    print("noalloc limits: leaf 1 tree of %d sources refused at the census" % n, flush=True)
    ctx.set(pkg.OPT_POT_LEAF, 1)
    with pytest.raises(pkg.TreecodeError, match="no-allocation scratch"):
        ctx.potential_noalloc(0.5, d["xs"], d["ys"], d["vs"], d["xt"], d["yt"], out, n=n, m=n)
    # the next call is whole again
    ctx.set(pkg.OPT_POT_LEAF, 512)
    ref = pkg.Context(precision=pkg.FP64, pot_order=12, pot_leaf=512)
    print("noalloc limits: call after the refusal, n=m=20000", flush=True)
    ctx.potential_noalloc(0.5, d["xs"], d["ys"], d["vs"], d["xt"], d["yt"], out, n=20000, m=20000)
    s = {k: ref.to_device(v.numpy()[:20000]) for k, v in d.items()}
    r = ref.empty(20000)
    ref.potential(0.5, s["xs"], s["ys"], s["vs"], s["xt"], s["yt"], out=r, device=True)
    ref.synchronize()
    assert np.array_equal(out.numpy()[:20000], r.numpy())
    ref.close()
    ctx.close()


def test_noalloc_zero_sources(pkg, wl):
    ctx = pkg.Context()
    d = _inputs(ctx, wl, 5000, 41)
    out = ctx.to_device(np.full(5000, 7.0))
    ctx.potential_noalloc(0.5, d["xs"], d["ys"], d["vs"], d["xt"], d["yt"], out, n=0)
    assert np.array_equal(out.numpy(), np.zeros(5000))
    ctx.close()
