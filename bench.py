#!/usr/bin/env python3
"""Benchmark of the treecode hot path on one MI355X.

Default workload (the precision form of treecode_potential): N = M = 2^24,
half uniform in [0,1]^2 and half Plummer (a = 0.02) about (0.5, 0.5)
truncated to [0,1]^2, values uniform in [-1, 1), theta = 0.5, each precision
at the order, leaf capacity and tiling autotuned for it
(2d-treecodes_amd/tuned.json).  One step = one whole solve (bbox, Morton sort, tree
build, upward pass, target sort, traversal) with inputs already resident
in HBM.  `value` is the FP32-bulk rate in Mtargets/s; the plain FP64 GPU
path is timed beside it, and both are checked against the FP64 direct sum
on 4096 random targets (GPU direct-sum kernel).

Other workloads (--config): c1 (potential, 65,536 uniform), c2 (force_mrag,
1,048,576 sources in 4 Gaussians, 4,096 blocks), c5 (force_mrag, 8,388,608
sources in 16 Gaussians, 16,384 blocks of mixed spacing).  Unless --no-legs,
the default run also times C3 at theta 0.3 and 0.7, C5, C2 and C1 (`legs`)
and replays the served stream (`serving`).

The JSON line also carries `roofline` (the traversal kernel's algorithmic
FLOP/s from deterministic interaction counts over its HIP-event duration)
and `cpu_baseline` (the C restatement timed on this host's cores on a
bounded sample of the same workload).
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

import __graft_entry__  # noqa: E402

pkg = __graft_entry__.load_package()
from treecodes2d_amd import workloads  # noqa: E402

PEAK_TFLOPS = {"f32": 157.3, "f64": 78.6}  # MI355X vector FP32 / FP64 (non-packed FP32 issues at the FP64 rate)


def flops_potential(cnt, p):
    # per-interaction model: P2P 9 flops per pair (log = 1), E2P 10p + 6 per (target, node)
    return 9.0 * cnt["p2p"] + (10.0 * p + 6.0) * cnt["e2p"]


def flops_force(cnt, p):
    # P2P 11 per pair, E2P 16p + 11 per (target, node), E2L 4p^2 + 21p - 1 per (brick, node), L2P 16p - 12
    return (11.0 * cnt["p2p"] + (16.0 * p + 11.0) * cnt["e2p"] + (4.0 * p * p + 21.0 * p - 1.0) * cnt["e2l"]
            + (16.0 * p - 12.0) * cnt["l2p"])


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def make_inputs(cfg, n):
    if cfg == "c1":
        n = n or 65536
        xs, ys, vs, xt, yt = workloads.uniform_potential(n)
        return "potential", dict(xs=xs, ys=ys, vs=vs, xt=xt, yt=yt), "potential uniform N=M=%d" % n
    if cfg == "c3":
        n = n or (1 << 24)
        xs, ys, vs, xt, yt = workloads.mixed_plummer_potential(n)
        return "potential", dict(xs=xs, ys=ys, vs=vs, xt=xt, yt=yt), \
            "potential half-uniform half-Plummer(a=0.02) N=M=%d" % n
    if cfg == "c2":
        n = n or (1 << 20)
        xs, ys, vs, x0s, y0s, hs = workloads.gaussian4_force(n)
        return "force", dict(xs=xs, ys=ys, vs=vs, x0s=x0s, y0s=y0s, hs=hs), \
            "force_mrag 4-Gaussian N=%d, 4096 blocks of 32x32 (h=1/2048)" % n
    if cfg == "c5":
        n = n or (1 << 23)
        xs, ys, vs, x0s, y0s, hs = workloads.gaussian16_force(n, nblocks=16384)
        return "force", dict(xs=xs, ys=ys, vs=vs, x0s=x0s, y0s=y0s, hs=hs), \
            "force_mrag 16-Gaussian N=%d, 16384 blocks of 32x32, h in {1/1024,1/2048,1/4096}" % n
    if cfg == "near":
        # near-field probe: theta = 0 never accepts a node, so the traversal is P2P
        # only; 2^20 targets give every SIMD its full complement of waves
        n = n or (1 << 16)
        xs, ys, vs, _, _ = workloads.mixed_plummer_potential(n)
        _, _, _, xt, yt = workloads.mixed_plummer_potential(1 << 20, seed=1457)
        return "potential", dict(xs=xs, ys=ys, vs=vs, xt=xt, yt=yt), \
            "near-field probe: potential half-uniform half-Plummer(a=0.02) N=%d, M=%d, theta 0 (P2P only)" % (
                n, len(xt))
    if cfg == "far":
        # far-field probe: targets in a frame around the source box, at least 1/20 of
        # its side away, are served by multipoles only (no leaf is ever opened)
        n = n or (1 << 24)
        xs, ys, vs, _, _ = workloads.mixed_plummer_potential(n)
        xt, yt = workloads.frame_targets(n, inner=0.05, outer=0.5, seed=1456)
        return "potential", dict(xs=xs, ys=ys, vs=vs, xt=xt, yt=yt), \
            "far-field probe: potential N=M=%d, Plummer-mix sources in [0,1]^2, targets in the frame " \
            "0.05-0.5 around it (E2P only)" % n
    raise SystemExit("unknown config " + cfg)


class GpuCase:
    def __init__(self, kind, data, precision, theta, order, leaf, tile=None):
        self.kind, self.theta = kind, theta
        kw = dict(precision=precision, profile=True)
        if kind == "potential":
            kw.update(pot_order=order or 12, pot_leaf=leaf or 512)
        else:
            kw.update(force_order=order or 24, force_leaf=leaf or 192)
        self.ctx = pkg.Context(**kw)
        if tile:
            self.ctx.set(*tile)  # (tiling option, value)
        self.order = order or (12 if kind == "potential" else 24)
        self.leaf = leaf or (512 if kind == "potential" else 192)
        self.d = {k: self.ctx.to_device(v) for k, v in data.items()}
        if kind == "potential":
            self.m = len(data["xt"])
            self.out = self.ctx.empty(self.m)
        else:
            self.m = len(data["x0s"]) * 1024
            self.xr = self.ctx.empty(self.m)
            self.yr = self.ctx.empty(self.m)

    def step(self):
        d = self.d
        if self.kind == "potential":
            self.ctx.potential(self.theta, d["xs"], d["ys"], d["vs"], d["xt"], d["yt"], out=self.out, device=True)
        else:
            self.ctx.force(self.theta, d["xs"], d["ys"], d["vs"], d["x0s"], d["y0s"], d["hs"], xr=self.xr,
                           yr=self.yr, device=True)
        self.ctx.synchronize()

    def timed(self, steps, warmup):
        for _ in range(warmup):
            self.step()
        phases = []
        t0 = time.perf_counter()
        for _ in range(steps):
            self.step()
            phases.append(self.ctx.stats()["times_ms"])
        dt = (time.perf_counter() - t0) / steps
        avg = {k: float(np.mean([p[k] for p in phases])) for k in phases[0]}
        return dt, avg

    def host_timed(self, data, reps=3):
        """Best of `reps` solves through the host-pointer path (inputs and outputs in pageable host
        memory, so each solve includes the H2D and D2H copies).  The output arrays are allocated and
        touched once, outside the timing, and reused, as a solver calling the library every step does."""
        best = None
        if self.kind == "potential":
            out = np.zeros(self.m)
        else:
            xr, yr = np.zeros(self.m), np.zeros(self.m)
        for _ in range(reps):
            t0 = time.perf_counter()
            if self.kind == "potential":
                self.ctx.potential(self.theta, data["xs"], data["ys"], data["vs"], data["xt"], data["yt"], out=out)
            else:
                self.ctx.force(self.theta, data["xs"], data["ys"], data["vs"], data["x0s"], data["y0s"], data["hs"],
                               xr=xr, yr=yr)
            dt = time.perf_counter() - t0
            best = dt if best is None else min(best, dt)
        return best

    def counters(self):
        """A solve with interaction counters (and per-category cycles) on: the counts, and the
        build's phases and the traversal's cycle shares (tc_get_profile)."""
        self.ctx.set(pkg.OPT_COUNTERS, 1)
        self.step()
        st = self.ctx.stats()
        st["profile"] = self.ctx.profile()
        self.ctx.set(pkg.OPT_COUNTERS, 0)
        return st

    def result(self):
        if self.kind == "potential":
            return self.out.numpy()
        return self.xr.numpy(), self.yr.numpy()


def sampled_error(case32, case64, kind, data, nsample=4096, seed=2024):
    rng = np.random.default_rng(seed)
    ctx = case64.ctx
    if kind == "potential":
        m = len(data["xt"])
        idx = np.sort(rng.choice(m, min(nsample, m), replace=False)).astype(np.int32)
        didx = ctx.to_device(idx, np.int32)
        d = case64.d
        ref = ctx.direct_potential(d["xs"], d["ys"], d["vs"], d["xt"], d["yt"], didx)
        out = {}
        for name, c in (("f32", case32), ("f64", case64)):
            r = c.result()[idx]
            out[name] = float(np.linalg.norm(r - ref) / np.linalg.norm(ref))
        return out
    x0s, y0s, hs = data["x0s"], data["y0s"], data["hs"]
    m = len(x0s) * 1024
    idx = np.sort(rng.choice(m, min(nsample, m), replace=False)).astype(np.int32)
    b = idx // 1024
    tx = x0s[b] + (idx % 32) * hs[b]
    ty = y0s[b] + ((idx // 32) % 32) * hs[b]
    dtx, dty = ctx.to_device(tx), ctx.to_device(ty)
    d = case64.d
    rx, ry = ctx.direct_force(d["xs"], d["ys"], d["vs"], dtx, dty, ctx.to_device(np.arange(len(idx), dtype=np.int32),
                                                                                  np.int32))
    out = {}
    for name, c in (("f32", case32), ("f64", case64)):
        fx, fy = c.result()
        num = np.sqrt(np.sum((fx[idx] - rx) ** 2) + np.sum((fy[idx] - ry) ** 2))
        out[name] = float(num / np.sqrt(np.sum(rx ** 2) + np.sum(ry ** 2)))
    return out


def cpu_baseline(kind, data, theta, order, leaf, budget_s):
    """The C restatement (oracle) on this host: full tree build + a timed sample of targets,
    extrapolated to the whole workload."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle
    threads = oracle.max_threads()
    t0 = time.perf_counter()
    if kind == "potential":
        tree = oracle.Tree(data["xs"], data["ys"], data["vs"], leaf or 512, order or 12)
    else:
        tree = oracle.Tree(data["xs"], data["ys"], data["vs"], leaf or 192, order or 24)
    t_build = time.perf_counter() - t0
    rng = np.random.default_rng(7)
    if kind == "potential":
        m = len(data["xt"])
        k = 2048
        while True:
            idx = rng.choice(m, min(k, m), replace=False)
            t1 = time.perf_counter()
            tree.potential(theta, data["xt"][idx], data["yt"][idx], nthreads=threads)
            te = time.perf_counter() - t1
            if te > 0.25 * budget_s or k >= m:
                break
            k = int(min(m, k * max(2.0, 0.5 * budget_s / max(te, 1e-3))))
        per_target = te / len(idx)
        sample = "oracle tree build on all %d sources (%.1f s, %d threads) + %d random targets (%.1f s, %d threads)" \
            % (len(data["xs"]), t_build, threads, len(idx), te, threads)
    else:
        nb = len(data["x0s"])
        k = 16
        while True:
            bidx = rng.choice(nb, min(k, nb), replace=False)
            t1 = time.perf_counter()
            tree.force(theta, data["x0s"][bidx], data["y0s"][bidx], data["hs"][bidx], nthreads=threads)
            te = time.perf_counter() - t1
            if te > 0.25 * budget_s or k >= nb:
                break
            k = int(min(nb, k * max(2.0, 0.5 * budget_s / max(te, 1e-3))))
        m = nb * 1024
        per_target = te / (len(bidx) * 1024)
        sample = "oracle tree build on all %d sources (%.1f s, %d threads) + %d random blocks (%.1f s, %d threads)" \
            % (len(data["xs"]), t_build, threads, len(bidx), te, threads)
    total = t_build + per_target * m
    return {"value": m / total / 1e6, "unit": "Mtargets/s", "cores": threads, "kind": "port",
            "sample": sample + "; extrapolated to %d targets" % m}


# Size buckets tuned for the served stream (FP32): potential N = M; force
# (N, blocks).  With the nearest-bucket lookup (at most 3 in summed log2
# distance) they cover the stream's request sizes.
SERVE_POT_N = (1 << 11, 1 << 14, 1 << 17, 1 << 20)
SERVE_FORCE = [(n, nb) for n in (1 << 14, 1 << 17, 1 << 20) for nb in (32, 256, 2048)]
SERVE_THETAS = (0.3, 0.5, 0.7)
SERVE_ERR = 2.5e-7  # a 4x margin under the metric's 1e-6 for sizes between buckets


def tune_serving(cache_path):
    """Autotune order and leaf capacity (FP32) for the served stream's size buckets.
    Each bucket is measured on both of the stream's source generators (uniform and
    4-Gaussian).  A cell qualifies if, on each generator, its error is within the
    looser of SERVE_ERR and the error of the reference's own parameters (12 / 512
    potential, 24 / 192 force, both in the grid) plus 1%, so a tuned choice is no
    less accurate than the reference's on either kind of input; the fastest
    qualifying cell (by the longer of its two times) is recorded."""
    from treecodes2d_amd import autotune
    tuner = autotune.Tuner(pkg, cache_path=cache_path or autotune.CACHE)
    done = []
    cases = [("potential", n, 0) for n in SERVE_POT_N] + [("force", n, nb) for n, nb in SERVE_FORCE]
    for kind, n, nb in cases:
        datas = []
        for gen in ("uniform", "gauss4"):
            r = {"kind": kind, "n": n, "nblocks": nb, "gen": gen, "src_seed": 4242 + n + nb, "tgt_seed": 77 + nb}
            xs, ys, vs = workloads.request_sources(r)
            tg = workloads.request_targets(r, xs, ys)
            if kind == "potential":
                datas.append(dict(xs=xs, ys=ys, vs=vs, xt=tg[0], yt=tg[1]))
            else:
                datas.append(dict(xs=xs, ys=ys, vs=vs, x0s=tg[0], y0s=tg[1], hs=tg[2]))
        leaves = (64, 128, 256, 512) if kind == "potential" else (64, 128, 192, 384)
        m = n if kind == "potential" else nb * 1024
        for theta in SERVE_THETAS:
            tables = [tuner.tune(kind, d, theta=theta, precision=1, err_target=1.0, reps=2, save=False,
                                 orders=(6, 8, 10, 12, 16, 24), leaves=leaves)["table"] for d in datas]
            p0, l0 = autotune.DEFAULT[kind]
            targets = [max(SERVE_ERR, 1.01 * [c for c in t if c["order"] == p0 and c["leaf"] == l0][0]["err"])
                       for t in tables]
            cells = [dict(a, ms=max(a["ms"], b["ms"]), err=max(a["err"], b["err"]),
                          ok=a["err"] <= targets[0] and b["err"] <= targets[1]) for a, b in zip(*tables)]
            base = [c for c in cells if c["order"] == p0 and c["leaf"] == l0][0]
            # the plain FP64 error at the reference's parameters, for comparison
            ref64 = max(tuner.tune(kind, d, theta=theta, precision=0, err_target=1.0, reps=1, save=False,
                                   orders=(p0,), leaves=(l0,))["err"] for d in datas)
            target = max(targets)
            best = min((c for c in cells if c["ok"]), key=lambda c: c["ms"])
            tuner.cache[autotune.bucket_key(kind, 1, theta, n, m)] = {
                "order": best["order"], "leaf": best["leaf"], "ms": best["ms"], "err": best["err"],
                "err_target": target}
            tuner._save()
            log("  %s N=%d blocks=%d theta=%.1f: order %d leaf %d (%.3f ms, err %.1e; FP32 at the reference's %d/%d: "
                "%.3f ms, err %.1e; FP64 there: err %.1e)" % (kind, n, nb, theta, best["order"], best["leaf"], best["ms"],
                                                              best["err"], p0, l0, base["ms"], base["err"], ref64))
            done.append({"kind": kind, "n": n, "blocks": nb, "theta": theta, "order": best["order"],
                         "leaf": best["leaf"], "ms": best["ms"], "err": best["err"], "ref_ms": base["ms"],
                         "ref_err": base["err"], "ref_err_fp64": ref64})
    return done


def batch_resident_case(k=16, n=1 << 14, m=1 << 11, reps=20, order=10, leaf=128):
    """k small potential requests against resident trees (one context each, FP32): one at a
    time (each call synchronised), all issued on their own streams then synchronised, and as one
    tc_potential_batch_resident launch of the persistent cross-request kernel.  Median wall times
    over reps; the batched results are checked against the one-at-a-time ones."""
    rng = np.random.default_rng(77)
    ctxs, reqs = [], []
    try:
        for i in range(k):
            c = pkg.Context(precision=1, pot_order=order, pot_leaf=leaf)
            ctxs.append(c)
            xs, ys, vs, _, _ = workloads.uniform_potential(n, seed=500 + i)
            c.potential(0.5, xs, ys, vs, xs[:4], ys[:4])
            xt, yt = rng.uniform(0, 1, m), rng.uniform(0, 1, m)
            reqs.append((c, 0.5, c.to_device(xt), c.to_device(yt), c.empty(m)))
        outs = [c.empty(m) for c in ctxs]

        def one_by_one():
            for (c, th, dxt, dyt, _), o in zip(reqs, outs):
                c.potential_resident(th, dxt, dyt, out=o, device=True)
                c.synchronize()

        def overlapped():
            for (c, th, dxt, dyt, _), o in zip(reqs, outs):
                c.potential_resident(th, dxt, dyt, out=o, device=True)
            for c in ctxs:
                c.synchronize()

        def batched():
            pkg.potential_batch_resident(reqs)

        res = {}
        for name, f in (("one_by_one", one_by_one), ("own_streams", overlapped), ("batched", batched)):
            f()
            ts = []
            for _ in range(reps):
                t0 = time.perf_counter()
                f()
                ts.append(time.perf_counter() - t0)
            res[name + "_ms"] = 1e3 * float(np.median(ts))
        one_by_one()
        batched()
        same = all(np.array_equal(r[4].numpy(), o.numpy()) for r, o in zip(reqs, outs))
        res.update(requests=k, sources=n, targets_each=m, order=order, leaf=leaf, precision="f32",
                   results_identical=bool(same))
        return res
    finally:
        for c in ctxs:
            c.close()


def serve_stream(nreq, rate_hz, slots, budget_gib, precision=1, seed=1454, tuned=True, batch=False, **mix):
    """Replay the served mixed-size stream: Poisson arrivals, 70% potential / 30% force,
    25% of requests reuse the previous request's sources (the server keeps the tree).
    With `tuned`, each request carries the order and leaf capacity of the nearest
    autotuned size bucket (FP32); others use the reference's.  With `batch`, the
    potential traversals of concurrent requests are launched together (graph replay
    per size bucket).  Latency is measured by the server from submission to completion."""
    import threading
    reqs = workloads.request_stream(nreq, rate_hz=rate_hz, seed=seed, **mix)
    pools = workloads.request_pools(seed=seed)
    tuner = None
    if tuned and precision == 1:
        from treecodes2d_amd import autotune
        tuner = autotune.Tuner(pkg)
    prepared = []
    ntuned = 0
    last = {}  # (kind, source seed) -> (order, leaf) of its latest request
    for r in reqs:
        xs, ys, vs = workloads.pool_sources(pools, r)
        # potential targets are the sources' own coordinates: passed as the same
        # (read-only) host arrays rather than 10,000 copies held at once
        tg = (xs, ys) if r["kind"] == "potential" else workloads.request_targets(r, xs, ys)
        ol = None
        if tuner:
            m = len(tg[0]) * (1 if r["kind"] == "potential" else 1024)
            ol = tuner.nearest(r["kind"], 1, r["theta"], len(xs), m)
            # a tree already built for these sources at an order no lower than
            # the tuned one is kept (no rebuild, no loss of accuracy)
            prev = last.get((r["kind"], r["src_seed"]))
            if ol and prev and (prev[0] or autotune.DEFAULT[r["kind"]][0]) >= ol[0]:
                ol = prev
            last[(r["kind"], r["src_seed"])] = ol or (0, 0)
        ntuned += ol is not None
        prepared.append((r, xs, ys, vs, tg, ol or (0, 0)))
    srv = pkg.Server(slots=slots, hbm_budget_gib=budget_gib, batch=batch)
    info = [None] * nreq
    kept = {}  # outputs of the requests whose accuracy is checked (every 40th)
    tickets = [None] * nreq
    done = threading.Event()

    def waiter():
        for i in range(nreq):
            while tickets[i] is None:
                time.sleep(0.0002)
            out_i, info[i] = srv.wait(tickets[i])
            if i % 40 == 0:
                kept[i] = out_i
        done.set()

    th = threading.Thread(target=waiter, daemon=True)
    th.start()
    t0 = time.perf_counter()
    for i, (r, xs, ys, vs, tg, (order, leaf)) in enumerate(prepared):
        now = time.perf_counter() - t0
        if r["arrival"] > now:
            time.sleep(r["arrival"] - now)
        sid = r["src_seed"]
        if r["kind"] == "potential":
            tickets[i] = srv.submit_potential(r["theta"], xs, ys, vs, *tg, precision=precision, source_id=sid,
                                              order=order, leaf=leaf)
        else:
            tickets[i] = srv.submit_force(r["theta"], xs, ys, vs, *tg, precision=precision, source_id=sid,
                                          order=order, leaf=leaf)
    done.wait()
    makespan = time.perf_counter() - t0
    st = srv.stats()
    srv.close()
    # accuracy of served results: every 40th request against the FP64 direct sum
    # on 256 of its targets (GPU direct-sum kernels)
    errs, errs64 = [], []
    ectx = pkg.Context()
    rctx = pkg.Context()  # plain FP64 at the reference's order / leaf, on the same sampled targets
    rng = np.random.default_rng(5)
    dev = ectx.to_device
    for i in sorted(kept):
        r, xs, ys, vs, tg, _ = prepared[i]
        res = kept[i]
        src = (dev(np.ascontiguousarray(xs)), dev(np.ascontiguousarray(ys)), dev(np.ascontiguousarray(vs)))
        if r["kind"] == "potential":
            idx = np.sort(rng.choice(len(tg[0]), min(256, len(tg[0])), replace=False)).astype(np.int32)
            ref = ectx.direct_potential(*src, dev(tg[0]), dev(tg[1]), dev(idx, np.int32))
            errs.append(float(np.linalg.norm(res[idx] - ref) / np.linalg.norm(ref)))
            r64 = rctx.potential(r["theta"], xs, ys, vs, *tg)
            errs64.append(float(np.linalg.norm(r64[idx] - ref) / np.linalg.norm(ref)))
        else:
            m = len(tg[0]) * 1024
            idx = np.sort(rng.choice(m, min(256, m), replace=False)).astype(np.int32)
            b = idx // 1024
            tx = tg[0][b] + (idx % 32) * tg[2][b]
            ty = tg[1][b] + ((idx // 32) % 32) * tg[2][b]
            rx, ry = ectx.direct_force(*src, dev(tx), dev(ty), dev(np.arange(len(idx), dtype=np.int32), np.int32))
            num = np.sum((res[0][idx] - rx) ** 2) + np.sum((res[1][idx] - ry) ** 2)
            errs.append(float(np.sqrt(num / (np.sum(rx ** 2) + np.sum(ry ** 2)))))
            f64 = rctx.force(r["theta"], xs, ys, vs, *tg)
            num = np.sum((f64[0][idx] - rx) ** 2) + np.sum((f64[1][idx] - ry) ** 2)
            errs64.append(float(np.sqrt(num / (np.sum(rx ** 2) + np.sum(ry ** 2)))))
    ectx.close()
    rctx.close()
    lat = np.array([x["latency_ms"] for x in info])
    svc = np.array([x["service_ms"] for x in info])
    targets = sum(len(p[4][0]) * (1 if p[0]["kind"] == "potential" else 1024) for p in prepared)
    return {"requests": nreq, "rate_hz": rate_hz, "slots": slots, "precision": "f32" if precision else "f64",
            "p50_ms": float(np.percentile(lat, 50)), "p99_ms": float(np.percentile(lat, 99)),
            "mean_service_ms": float(svc.mean()), "makespan_s": makespan,
            "targets_per_s": targets / makespan, "reused_trees": int(st["reused"]),
            "reusing_requests": int(sum(p[0]["reuses_sources"] for p in prepared)),
            "max_inflight": int(st["max_inflight"]), "batched": bool(batch),
            "batch_launches": int(st["batches"]), "batched_requests": int(st["batched"]),
            "merged_requests": int(st["merged"]),
            "budget_gib": budget_gib, "peak_admitted_gib": st["peak_bytes"] / 2 ** 30,
            "evicted_trees": int(st["evicted"]), "trimmed_slots": int(st["trimmed"]),
            "rel_err_sampled": {"requests": len(errs), "max": max(errs), "median": float(np.median(errs))},
            "rel_err_sampled_fp64_reference_params": {"requests": len(errs64), "max": max(errs64),
                                                      "median": float(np.median(errs64)),
                                                      "note": "the same sampled requests solved in plain FP64 at the "
                                                              "reference's order / leaf (12 / 512, 24 / 192)"},
            "params": ("order / leaf of the nearest autotuned size bucket (2d-treecodes_amd/tuned.json) for %d of %d "
                       "requests, the reference's for the rest" % (ntuned, nreq)) if tuner else
                      "the reference's order / leaf (12 / 512 potential, 24 / 192 force)"}


def executed_flops_potential(cnt, p):
    # what the kernel actually evaluates: per-lane E2P for the interactions not taken by a
    # wavefront's local expansion, 16 local coefficients per wave-level translation (the
    # Horner over p terms, 8 flops each, plus about 24 for the powers and the log term)
    # and 16 complex multiply-adds per target for the local evaluation
    e2p_lane = cnt["e2p"] - cnt.get("e2p_wave", 0)
    m2l = cnt.get("m2l_wave", 0)
    l2p = 128.0 * (cnt.get("targets", 0) if m2l else 0)
    return 9.0 * cnt["p2p"] + (10.0 * p + 6.0) * e2p_lane + 16.0 * (8.0 * p + 24.0) * m2l + l2p


def measure(cfg, theta, args, cpu_budget, host=False, n=0):
    """Time one workload in both precisions (each at its tuned order / leaf unless given),
    check both against the FP64 direct sum, and build the result record."""
    log("generating workload", cfg, "theta", theta)
    kind, data, desc = make_inputs(cfg, n)
    cases, timings, tuned = {}, {}, {}
    for prec in (["f32", "f64"] if args.only is None else [args.only]):
        order, leaf, tile = args.order, args.leaf, None
        from treecodes2d_amd import autotune
        if not (args.tune or args.reference_params or order or leaf):
            m_t = len(data["xt"]) if kind == "potential" else len(data["x0s"]) * 1024
            tuner = autotune.Tuner(pkg)
            hit = tuner.lookup(kind, 1 if prec == "f32" else 0, theta, len(data["xs"]), m_t)
            if hit:
                order, leaf = hit
                tile = tuner.lookup_tile(kind, 1 if prec == "f32" else 0, theta, len(data["xs"]), m_t)
                tuned[prec] = {"order": order, "leaf": leaf, "source": "2d-treecodes_amd/tuned.json"}
                if tile:
                    tuned[prec]["tile"] = {autotune.TILING[kind][0]: tile[1]}
                log("  tuned parameters for %s: order %d leaf %d%s" % (
                    prec, order, leaf, " tile %s=%d" % (autotune.TILING[kind][0], tile[1]) if tile else ""))
        if args.tune_tiling and order and leaf:
            # the tiling sweep alone, at this precision's tuned (or given) order and leaf
            log("tiling sweep", prec)
            tt = autotune.Tuner(pkg, cache_path=args.tune_cache or autotune.CACHE)
            res = tt.tune(kind, data, theta=theta, precision=1 if prec == "f32" else 0, orders=(order,),
                          leaves=(leaf,), err_target=args.err_target, reps=3, tilings=autotune.TILING[kind][1])
            tile = (getattr(pkg, autotune.TILING[kind][0]), res["tile"])
            tuned[prec] = {"order": order, "leaf": leaf, "tile": {autotune.TILING[kind][0]: res["tile"]},
                           "tile_table": [(t["tile"], round(t["ms"], 3), t["err"]) for t in res["tile_table"]]}
            log("  %s: %s" % (autotune.TILING[kind][0], ", ".join("%d: %.3f ms" % (t["tile"], t["ms"])
                                                                   for t in res["tile_table"])))
        if args.tune:
            log("autotuning", prec)
            grid = dict(orders=(8, 9, 10, 12, 14, 16, 20, 24), leaves=(64, 128, 192, 256, 384, 512)) \
                if kind == "potential" else dict(orders=(8, 10, 12, 16), leaves=(32, 64, 128, 192, 384))
            tuner = autotune.Tuner(pkg, cache_path=args.tune_cache or autotune.CACHE)
            res = tuner.tune(kind, data, theta=theta, precision=1 if prec == "f32" else 0,
                             err_target=args.err_target, reps=2, tilings=autotune.TILING[kind][1], **grid)
            order, leaf = res["order"], res["leaf"]
            tile = (getattr(pkg, autotune.TILING[kind][0]), res["tile"])
            tuned[prec] = {"order": order, "leaf": leaf, "err_sampled": res["err"],
                           "tile": {autotune.TILING[kind][0]: res["tile"]},
                           "table": [(t["order"], t["leaf"], round(t["ms"], 3), t["err"]) for t in res["table"]],
                           "tile_table": [(t["tile"], round(t["ms"], 3), t["err"]) for t in res["tile_table"]]}
            log("  chose order %d leaf %d %s=%d (%.3f ms, err %.2e)" % (order, leaf, autotune.TILING[kind][0],
                                                                       res["tile"], res["ms"], res["err"]))
        log("timing", prec)
        c = GpuCase(kind, data, 1 if prec == "f32" else 0, theta, order, leaf, tile)
        dt, phases = c.timed(args.steps, args.warmup)
        cases[prec] = c
        timings[prec] = (dt, phases)
        log("  %s: %.3f ms/step, phases %s" % (prec, dt * 1e3, {k: round(v, 3) for k, v in phases.items()}))

    head = "f32" if "f32" in cases else "f64"
    p = cases[head].order
    dt, phases = timings[head]
    m = cases[head].m
    st = cases[head].counters()
    cnt = dict(st["counters"])
    fl = flops_potential(cnt, p) if kind == "potential" else flops_force(cnt, p)
    eval_ms = phases["eval"]
    achieved = fl / (eval_ms * 1e-3) / 1e12
    peak = PEAK_TFLOPS[head]
    roof = {"bound": "valu", "achieved": achieved, "peak": peak, "unit": "TFLOP/s", "frac": achieved / peak,
            "traffic": None, "traffic_source": None,
            "kernel": "k_potential (traversal)" if kind == "potential" else "k_force + k_force_chunks (traversal)",
            "flops_model": "per interaction: P2P %d, E2P %s" % ((9, "10p+6") if kind == "potential" else
                                                                (11, "16p+11, E2L 4p^2+21p-1, L2P 16p-12")),
            "note": "VALU-bound; peak is the vector %s rate (no MFMA in this path)" % head.upper()}
    if kind == "potential" and cnt.get("m2l_wave"):
        cnt_t = dict(cnt, targets=m)
        ex = executed_flops_potential(cnt_t, p) / (eval_ms * 1e-3) / 1e12
        roof["executed_tflops"] = ex
        roof["note"] += ("; %.0f%% of the (target, node) multipole interactions are evaluated through "
                         "wavefront-level local expansions, so the algorithmic rate (achieved) exceeds the "
                         "executed one (executed_tflops)" % (100.0 * cnt["e2p_wave"] / max(cnt["e2p"], 1)))
    tfile = os.path.join(ROOT, "profiles", "potential_2e24_traffic.json")
    if kind == "potential" and head == "f32" and os.path.exists(tfile):
        tj = json.load(open(tfile))
        if tj["workload"] == desc and tj["order"] == p and tj["theta"] == theta and tj["leaf"] == cases[head].leaf:
            roof["traffic"] = tj["hbm_bytes_per_launch"]
            roof["traffic_source"] = ("copy-calibrated rocprofv3 FETCH_SIZE + WRITE_SIZE of this kernel on this "
                                      "workload, from profiles/potential_2e24_hbm_traffic.md (not measured in "
                                      "this run)")
    errs = None
    if not args.no_error and "f32" in cases and "f64" in cases:
        log("direct-sum error check")
        errs = sampled_error(cases["f32"], cases["f64"], kind, data)
    hostr = None
    if host:
        log("host-array C ABI timing")
        hdt = cases[head].host_timed(data)
        hostr = {"ms": hdt * 1e3, "mtargets_per_s": m / hdt / 1e6, "dtype": head,
                 "headline_form": "end-to-end: the metric as a drop-in caller sees it",
                 "path": "end-to-end C ABI with host arrays in and out (pageable memory): H2D copies, build, "
                         "solve, D2H copy; PCIe-inclusive (value is the HBM-resident rate)"}
        log("  %.3f ms -> %.1f Mtargets/s" % (hdt * 1e3, hostr["mtargets_per_s"]))
    for c in cases.values():
        c.ctx.close()
    cpu = None
    if cpu_budget > 0:
        log("cpu baseline")
        cpu = cpu_baseline(kind, data, theta, args.order, args.leaf, cpu_budget)
    res = {"value": m / dt / 1e6, "ms_per_step": dt * 1e3, "dtype": head,
           "config": {"workload": desc, "theta": theta, "order": p, "leaf": cases[head].leaf,
                      "precision": "FP32 P2P/E2P on leaf-relative offsets, FP64 decisions/P2E/E2E/E2L"
                      if head == "f32" else "plain FP64", "inputs": "resident in HBM"},
           "phases_ms": {k: round(v, 4) for k, v in phases.items()},
           "counters": {k: int(v) for k, v in cnt.items()},
           "profile": {"build_ms": {k: round(v, 4) for k, v in st["profile"]["build_ms"].items()},
                       "traversal_cycle_shares": {k: round(v, 4) for k, v in
                                                  st["profile"]["cycle_shares"].items()},
                       "note": "from the solve with counters on (build phases on the context stream, the "
                               "target sort beside them; cycle shares of the traversal's wavefront time)"},
           "roofline": roof, "cpu_baseline": cpu, "host_abi": hostr}
    if "f32" in timings and "f64" in timings:
        res["fp64_value"] = m / timings["f64"][0] / 1e6
        res["fp64_ms_per_step"] = timings["f64"][0] * 1e3
        res["speedup_fp32_vs_fp64"] = timings["f64"][0] / timings["f32"][0]
        res["fp64_phases_ms"] = {k: round(v, 4) for k, v in timings["f64"][1].items()}
    if errs:
        res["rel_err_vs_direct"] = errs
    if tuned:
        res["autotune"] = tuned
    if "f64" in cases:
        res["fp64_config"] = {"order": cases["f64"].order, "leaf": cases["f64"].leaf}
    return res


# the other benchmark workloads, measured after the headline one (theta, CPU-baseline seconds)
LEGS = [("c3", 0.3, 0.0), ("c3", 0.7, 0.0), ("c5", 0.5, 8.0), ("c2", 0.5, 0.0), ("c1", 0.5, 4.0)]


def main():
    ap = argparse.ArgumentParser(description=__doc__, formatter_class=argparse.RawDescriptionHelpFormatter)
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--config", default="c3", choices=["c1", "c2", "c3", "c5", "near", "far"])
    ap.add_argument("--n", type=int, default=0, help="override the number of sources")
    ap.add_argument("--theta", type=float, default=None, help="default 0.5 (0 for the near-field probe)")
    ap.add_argument("--order", type=int, default=0)
    ap.add_argument("--leaf", type=int, default=0)
    ap.add_argument("--cpu-budget", type=float, default=20.0, help="seconds of CPU-baseline work")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-host", action="store_true", help="skip the host-array (PCIe-inclusive) timing")
    ap.add_argument("--no-error", action="store_true")
    ap.add_argument("--no-legs", action="store_true",
                    help="only the selected workload (default: also C1, C2, C5 and C3 at theta 0.3 / 0.7)")
    ap.add_argument("--only", choices=["f32", "f64"], default=None, help="time one precision only (profiling)")
    ap.add_argument("--serve-requests", type=int, default=10000,
                    help="requests of the replayed served stream (0 = skip)")
    ap.add_argument("--serve-rate", type=float, default=500.0, help="Poisson arrival rate of the stream (1/s)")
    ap.add_argument("--serve-slots", type=int, default=4)
    ap.add_argument("--serve-budget", type=float, default=16.0, help="the server's HBM budget (GiB)")
    ap.add_argument("--no-serve-batch", action="store_true",
                    help="replay the stream only without cross-request batching")
    ap.add_argument("--tune", action="store_true",
                    help="re-run the autotuner for this workload (grid of order x leaf) before timing")
    ap.add_argument("--err-target", type=float, default=8e-7,
                    help="autotuner error target (default 8e-7, a 20%% margin under the metric's 1e-6)")
    ap.add_argument("--tune-cache", default="", metavar="PATH",
                    help="autotuner cache to write (default 2d-treecodes_amd/tuned.json)")
    ap.add_argument("--tune-tiling", action="store_true",
                    help="sweep the tiling option at the tuned order / leaf of each precision (see autotune.TILING)")
    ap.add_argument("--tune-serving", action="store_true",
                    help="autotune the served stream's size buckets and exit")
    ap.add_argument("--serve-reference-params", action="store_true",
                    help="serve every request at the reference's order / leaf instead of the tuned buckets")
    ap.add_argument("--reference-params", action="store_true",
                    help="time both precisions at the reference's order / leaf (no autotuned parameters)")
    ap.add_argument("--copy-calibration", type=int, default=0, metavar="MIB",
                    help="also run a plain copy kernel of this size (MiB), the reference for HBM counters")
    args = ap.parse_args()
    if args.theta is None:
        args.theta = 0.0 if args.config == "near" else 0.5
    if args.gpus != 1:
        print("bench.py: this build measures one GPU; multi-GPU runs are not supported (--gpus %d)" % args.gpus,
              file=sys.stderr)
        sys.exit(2)

    if args.tune_serving:
        done = tune_serving(args.tune_cache)
        print(json.dumps({"tuned_serving_buckets": done}))
        return

    if args.copy_calibration:
        cctx = pkg.Context()
        nb = args.copy_calibration << 20
        a, b = cctx.empty(nb // 8), cctx.empty(nb // 8)
        for _ in range(3):
            cctx.copy_kernel(b, a)
        log("copy kernel: 3 x %d MiB" % args.copy_calibration)
        a.free()
        b.free()
        cctx.close()

    head = measure(args.config, args.theta, args, 0.0 if args.no_cpu else args.cpu_budget, host=not args.no_host,
                   n=args.n)
    legs = {}
    if not args.no_legs and args.only is None and not args.tune:
        for cfg, theta, budget in LEGS:
            if (cfg, theta) == (args.config, args.theta):
                continue
            key = "%s_theta%.1f" % (cfg, theta)
            # the force leg also times the drop-in call with host arrays (how an MRAG solver calls it)
            legs[key] = measure(cfg, theta, args, 0.0 if args.no_cpu else budget,
                                host=cfg == "c5" and not args.no_host)
            log("  leg %s: %.1f Mtargets/s (%.3f ms), FP32/FP64 %.2fx" % (
                key, legs[key]["value"], legs[key]["ms_per_step"], legs[key].get("speedup_fp32_vs_fp64", 0)))
    serving = None
    if args.serve_requests > 0:
        log("served stream: %d requests at %.0f/s" % (args.serve_requests, args.serve_rate))
        serving = serve_stream(args.serve_requests, args.serve_rate, args.serve_slots, args.serve_budget,
                               tuned=not args.serve_reference_params)
        log("  p50 %.2f ms p99 %.2f ms, %d trees reused" % (serving["p50_ms"], serving["p99_ms"],
                                                             serving["reused_trees"]))
        if not args.no_serve_batch:
            log("served stream with batched traversals")
            b = serve_stream(args.serve_requests, args.serve_rate, args.serve_slots, args.serve_budget,
                             tuned=not args.serve_reference_params, batch=True)
            log("  p50 %.2f ms p99 %.2f ms, %d launches for %d traversals" % (
                b["p50_ms"], b["p99_ms"], b["batch_launches"], b["batched_requests"]))
            # a burst: 3000 small potential requests at once, where traversals of
            # different slots do meet
            burst = {}
            for mode in (False, True):
                r = serve_stream(3000, 1e6, args.serve_slots, args.serve_budget, batch=mode, pot_frac=1.0,
                                 pot_n=(1e3, 3e4))
                burst["batched" if mode else "unbatched"] = {k: r[k] for k in (
                    "p50_ms", "p99_ms", "makespan_s", "mean_service_ms", "batch_launches", "batched_requests",
                    "merged_requests")}
            log("  burst of 3000 small requests: makespan %.3f s unbatched, %.3f s batched" % (
                burst["unbatched"]["makespan_s"], burst["batched"]["makespan_s"]))
            serving = {"unbatched": serving, "batched": b, "burst_3000_small": burst}
            br = batch_resident_case()
            log("  16 small requests on resident trees: %.3f ms one by one, %.3f ms on own streams, %.3f ms "
                "in one batched launch" % (br["one_by_one_ms"], br["own_streams_ms"], br["batched_ms"]))
            serving["batch_resident_16"] = br

    res = {
        "metric": "Mtargets/s at rel err <=1e-6, FP32-bulk vs FP64 plain",
        "value": head["value"],
        "unit": "Mtargets/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": head["ms_per_step"],
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": head["dtype"],
        "data": "synthetic",
    }
    res.update({k: v for k, v in head.items() if k not in ("value", "ms_per_step", "dtype")})
    res["serving"] = serving
    if legs:
        res["legs"] = legs
    print(json.dumps(res), flush=True)


if __name__ == "__main__":
    main()
