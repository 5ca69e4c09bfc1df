"""The request server: results are identical alone and in a concurrent mix,
reused trees give the same answers as rebuilt ones, and admission respects
the HBM budget."""
import numpy as np
import pytest

from conftest import rel_l2

pytestmark = pytest.mark.gpu


def _mix(wl, count=24):
    reqs = wl.request_stream(count, rate_hz=1e6, seed=99, pot_n=(1e3, 6e4), force_n=(1e4, 6e4), force_nb=(4, 48))
    items = []
    for r in reqs:
        xs, ys, vs = wl.request_sources(r)
        tg = wl.request_targets(r, xs, ys)
        items.append((r, xs, ys, vs, tg))
    return items


def _alone(pkg, r, xs, ys, vs, tg, precision):
    c = pkg.Context(precision=precision)
    try:
        if r["kind"] == "potential":
            return c.potential(r["theta"], xs, ys, vs, *tg)
        return c.force(r["theta"], xs, ys, vs, *tg)
    finally:
        c.close()


@pytest.mark.parametrize("precision,batch", [(0, False), (1, False), (0, True), (1, True)])
def test_mix_equals_alone(pkg, wl, precision, batch):
    items = _mix(wl)
    srv = pkg.Server(slots=4, hbm_budget_gib=8.0, batch=batch)
    try:
        tickets = []
        for r, xs, ys, vs, tg in items:
            sid = r["src_seed"]
            if r["kind"] == "potential":
                t = srv.submit_potential(r["theta"], xs, ys, vs, *tg, precision=precision, source_id=sid)
            else:
                t = srv.submit_force(r["theta"], xs, ys, vs, *tg, precision=precision, source_id=sid)
            tickets.append(t)
        outs = [srv.wait(t) for t in tickets]
        st = srv.stats()
    finally:
        srv.close()
    assert st["completed"] == len(items) and st["failed"] == 0
    assert st["max_inflight"] >= 2  # requests overlapped
    npot = sum(r["kind"] == "potential" for r, *_ in items)
    if batch:
        # every potential traversal went through the batcher, in at most as many launches
        assert st["batched"] == npot and 1 <= st["batches"] <= npot
    else:
        assert st["batched"] == 0
    for (r, xs, ys, vs, tg), (res, info) in zip(items, outs):
        ref = _alone(pkg, r, xs, ys, vs, tg, precision)
        if r["kind"] == "potential":
            assert np.array_equal(res, ref)
        else:
            assert np.array_equal(res[0], ref[0]) and np.array_equal(res[1], ref[1])
        assert info["latency_ms"] >= info["service_ms"] >= 0


def test_reused_tree_matches_rebuild(pkg, wl, oracle):
    xs, ys, vs, xt, yt = wl.uniform_potential(40000, seed=41)
    srv = pkg.Server(slots=1, hbm_budget_gib=4.0)
    try:
        t1 = srv.submit_potential(0.5, xs, ys, vs, xt, yt, source_id=7)
        t2 = srv.submit_potential(0.5, xs, ys, vs, xt[::5], yt[::5], source_id=7)
        a, ia = srv.wait(t1)
        b, ib = srv.wait(t2)
    finally:
        srv.close()
    assert not ia["reused"] and ib["reused"]
    assert np.array_equal(a[::5], b)
    ref, _ = oracle.potential(0.5, xs, ys, vs, xt, yt)
    assert rel_l2(a, ref) <= 1e-12


def test_budget_rejects_oversized(pkg, wl):
    srv = pkg.Server(slots=1, hbm_budget_gib=0.01)
    try:
        xs, ys, vs, xt, yt = wl.uniform_potential(200000, seed=42)
        with pytest.raises(pkg.TreecodeError):
            srv.submit_potential(0.5, xs, ys, vs, xt, yt)
        small = wl.uniform_potential(2000, seed=43)
        t = srv.submit_potential(0.5, *small)
        out, _ = srv.wait(t)
        assert np.all(np.isfinite(out))
    finally:
        srv.close()


def test_request_order_and_leaf(pkg, wl, oracle):
    # per-request expansion order and leaf capacity (as a server sets them from
    # the autotuner): results equal a context solve with the same options, and
    # a resident tree built with other parameters is not reused
    xs, ys, vs, xt, yt = wl.uniform_potential(30000, seed=44)
    srv = pkg.Server(slots=1, hbm_budget_gib=4.0)
    try:
        t1 = srv.submit_potential(0.5, xs, ys, vs, xt, yt, source_id=9, order=8, leaf=64)
        t2 = srv.submit_potential(0.5, xs, ys, vs, xt, yt, source_id=9)
        t3 = srv.submit_potential(0.5, xs, ys, vs, xt[::3], yt[::3], source_id=9)
        a, ia = srv.wait(t1)
        b, ib = srv.wait(t2)
        c, ic = srv.wait(t3)
    finally:
        srv.close()
    assert not ia["reused"] and not ib["reused"] and ic["reused"]
    ctx = pkg.Context(pot_order=8, pot_leaf=64)
    assert np.array_equal(a, ctx.potential(0.5, xs, ys, vs, xt, yt))
    ctx.close()
    ref, _ = oracle.potential(0.5, xs, ys, vs, xt, yt)
    assert rel_l2(b, ref) <= 1e-12
    assert np.array_equal(b[::3], c)
    ref8, _ = oracle.potential(0.5, xs, ys, vs, xt, yt, order=8, leafcap=64)
    assert rel_l2(a, ref8) <= 1e-12


def test_tree_cache_shared_by_slots(pkg, wl):
    # a tree built by one slot serves requests in every slot
    xs, ys, vs, xt, yt = wl.uniform_potential(60000, seed=45)
    srv = pkg.Server(slots=4, hbm_budget_gib=4.0)
    try:
        first, info0 = srv.wait(srv.submit_potential(0.5, xs, ys, vs, xt, yt, source_id=11))
        tickets = [srv.submit_potential(0.5, xs, ys, vs, xt[k::8], yt[k::8], source_id=11) for k in range(8)]
        outs = [srv.wait(t) for t in tickets]
        st = srv.stats()
    finally:
        srv.close()
    assert not info0["reused"]
    assert all(info["reused"] for _, info in outs)
    assert len({info["slot"] for _, info in outs}) >= 2
    assert st["reused"] == 8 and st["cached_trees"] == 1
    for k, (res, _) in enumerate(outs):
        assert np.array_equal(res, first[k::8])


def test_waits_for_tree_being_built(pkg, wl):
    # requests submitted together with the same sources: one builds, the others
    # wait for its tree rather than building their own
    xs, ys, vs, xt, yt = wl.uniform_potential(80000, seed=46)
    srv = pkg.Server(slots=4, hbm_budget_gib=4.0)
    try:
        tickets = [srv.submit_potential(0.5, xs, ys, vs, xt[k::4], yt[k::4], source_id=12) for k in range(4)]
        outs = [srv.wait(t) for t in tickets]
        st = srv.stats()
    finally:
        srv.close()
    assert st["reused"] == 3
    ctx = pkg.Context()
    ref = ctx.potential(0.5, xs, ys, vs, xt, yt)
    ctx.close()
    for k, (res, _) in enumerate(outs):
        assert np.array_equal(res, ref[k::4])


@pytest.mark.parametrize("precision", [0, 1])
def test_batched_graph_replays(pkg, wl, precision):
    # the batcher replays graphs captured per size bucket: rounds of concurrent
    # requests of varied sizes, with eager solves in between, each result equal
    # to the request solved alone
    rng = np.random.default_rng(47)
    rounds = []
    for _ in range(4):
        items = []
        for j in range(10):
            n = int(rng.integers(2000, 40000))
            m = int(rng.integers(500, 50000))
            xs, ys, vs, _, _ = wl.uniform_potential(n, seed=int(rng.integers(1 << 30)))
            xt, yt = rng.uniform(-0.1, 1.1, m), rng.uniform(-0.1, 1.1, m)
            items.append((float(rng.choice([0.3, 0.5, 0.7])), xs, ys, vs, xt, yt))
        rounds.append(items)
    ctx = pkg.Context(precision=precision)
    srv = pkg.Server(slots=4, hbm_budget_gib=4.0, batch="always")
    try:
        for items in rounds:
            tickets = [srv.submit_potential(*it, precision=precision) for it in items]
            outs = [srv.wait(t)[0] for t in tickets]
            for it, res in zip(items, outs):
                assert np.array_equal(res, ctx.potential(*it))
        st = srv.stats()
    finally:
        srv.close()
        ctx.close()
    assert st["batched"] == 40 and st["failed"] == 0
    assert st["batches"] <= 40


def _hip():
    import ctypes
    import glob
    import os
    paths = [p for p in glob.glob("/opt/rocm/lib/libamdhip64.so*") if os.path.isfile(p)]
    L = ctypes.CDLL(sorted(paths)[0])
    L.hipMemGetInfo.argtypes = [ctypes.POINTER(ctypes.c_size_t), ctypes.POINTER(ctypes.c_size_t)]
    return L


def test_peak_device_memory_within_budget(pkg, wl):
    # requests whose trees together exceed the budget: the server evicts idle trees
    # and trims idle slots, and the device memory it holds (sampled while the
    # requests run) never exceeds the budget
    import ctypes
    import threading
    hip = _hip()

    def free_bytes():
        f, t = ctypes.c_size_t(), ctypes.c_size_t()
        assert hip.hipMemGetInfo(ctypes.byref(f), ctypes.byref(t)) == 0
        return f.value

    budget = 1.0
    rng = np.random.default_rng(48)
    items = []
    for j in range(24):
        n = int(rng.integers(200000, 1000000))
        xs, ys, vs, xt, yt = wl.uniform_potential(n, seed=1000 + j)
        items.append(("potential", j, xs, ys, vs, xt, yt))
        if j % 3 == 0:
            x0s, y0s, hs = wl.mrag_grid_blocks(8, 8, 1.0 / 256)
            items.append(("force", j, xs, ys, vs, x0s, y0s, hs))
    srv = pkg.Server(slots=4, hbm_budget_gib=budget, batch=True)
    low = [None]
    stop = threading.Event()

    def sample():
        while not stop.is_set():
            f = free_bytes()
            low[0] = f if low[0] is None else min(low[0], f)
            stop.wait(0.0005)

    try:
        base = free_bytes()
        th = threading.Thread(target=sample, daemon=True)
        th.start()
        tickets = []
        for kind, j, *a in items:
            if kind == "potential":
                tickets.append(srv.submit_potential(0.5, *a, precision=j % 2, source_id=1 + j))
            else:
                tickets.append(srv.submit_force(0.5, *a, precision=j % 2, source_id=1 + j))
        for t in tickets:
            srv.wait(t)
        stop.set()
        th.join()
        st = srv.stats()
    finally:
        stop.set()
        srv.close()
    used = base - low[0]
    msg = "sampled %.1f MiB, admitted peak %.1f MiB, budget %.0f MiB, stats %s" % (
        used / 2 ** 20, st["peak_bytes"] / 2 ** 20, budget * 1024, st)
    print(msg)
    assert st["completed"] == len(items) and st["failed"] == 0
    assert st["evicted"] > 0
    assert st["estimate_misses"] == 0, msg
    assert st["peak_bytes"] <= budget * 2 ** 30, msg
    assert used <= budget * 2 ** 30, msg


@pytest.mark.parametrize("precision", [0, 1])
def test_batch_resident_equals_single(pkg, wl, precision):
    # tc_potential_batch_resident: several contexts' resident trees evaluated in one
    # launch of the persistent cross-request kernel, each result bit-identical to the
    # context's own tc_potential_resident; repeated (graph replay) after an eager solve
    rng = np.random.default_rng(71)
    leaves = (64, 128, 192, 512, 96)
    thetas = (0.3, 0.5, 0.7, 0.5, 0.4)
    sizes = (1, 63, 1000, 20000, 70000)
    ctxs, reqs = [], []
    try:
        for i, (leaf, th, m) in enumerate(zip(leaves, thetas, sizes)):
            c = pkg.Context(precision=precision, pot_order=10, pot_leaf=leaf)
            ctxs.append(c)
            n = int(rng.integers(3000, 60000))
            if i % 2:
                xs, ys, vs, _, _ = wl.uniform_potential(n, seed=100 + i)
            else:
                xs, ys, vs, _, _ = wl.mixed_plummer_potential(n, seed=100 + i)
            c.potential(th, xs, ys, vs, xs[:5], ys[:5])  # leaves the tree resident
            xt, yt = rng.uniform(-0.2, 1.2, m), rng.uniform(-0.2, 1.2, m)
            reqs.append((c, th, c.to_device(xt), c.to_device(yt), c.empty(m)))
        alone = []
        for c, th, dxt, dyt, _ in reqs:
            o = c.empty(len(dxt))
            c.potential_resident(th, dxt, dyt, out=o, device=True)
            c.synchronize()
            alone.append(o.numpy())
        for rep in range(2):
            for c, th, dxt, dyt, out in reqs:
                out.upload(np.full(len(dxt), np.nan))
            pkg.potential_batch_resident(reqs)
            for (c, th, dxt, dyt, out), ref in zip(reqs, alone):
                assert np.array_equal(out.numpy(), ref)
            ctxs[0].potential_resident(0.5, reqs[0][2], reqs[0][3], out=ctxs[0].empty(1), device=True)
        # refused: a context twice, more than 64 requests, trees of different orders
        with pytest.raises(pkg.TreecodeError, match="twice"):
            pkg.potential_batch_resident([reqs[0], reqs[0]])
        with pytest.raises(pkg.TreecodeError, match="at most 64"):
            pkg.potential_batch_resident([reqs[0]] * 65)
        c6 = pkg.Context(precision=precision, pot_order=12, pot_leaf=64)
        ctxs.append(c6)
        xs, ys, vs, _, _ = wl.uniform_potential(5000, seed=9)
        c6.potential(0.5, xs, ys, vs, xs[:5], ys[:5])
        with pytest.raises(pkg.TreecodeError, match="orders or precisions"):
            pkg.potential_batch_resident([reqs[1], (c6, 0.5, reqs[1][2], reqs[1][3], reqs[1][4])])
        c7 = pkg.Context(precision=precision, pot_order=10, pot_leaf=64)
        ctxs.append(c7)
        with pytest.raises(pkg.TreecodeError, match="no resident tree"):
            pkg.potential_batch_resident([reqs[1], (c7, 0.5, reqs[1][2], reqs[1][3], reqs[1][4])])
    finally:
        for c in ctxs:
            c.close()
