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


@pytest.mark.parametrize("precision", [0, 1])
def test_mix_equals_alone(pkg, wl, precision):
    items = _mix(wl)
    srv = pkg.Server(slots=4, hbm_budget_gib=8.0)
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
