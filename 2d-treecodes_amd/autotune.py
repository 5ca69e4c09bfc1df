"""Autotuner for leaf capacity, expansion order and tiling.

For one workload (sources, targets or blocks, theta, precision) it measures,
on the GPU, every (leaf capacity, order) pair of a grid: the whole-solve time
with inputs resident in HBM, and the relative L2 error against the FP64
direct sum on a random sample of targets.  It returns the fastest pair whose
error is within the target (default: no worse than the reference defaults,
order 12 / leaf 512 for the potential and 24 / 192 for the force, and at
most 1e-6), and caches the choice per (kind, precision, theta, log2 N,
log2 targets) bucket in a JSON file so a server can look it up per request
size.  With `tilings`, it then times the chosen pair under each tiling value
(TILING: the potential traversal's workgroups per XCD chunk, or the force
traversal's inline P2P leaves per brick) and records the fastest as "tile".
"""
import json
import math
import os
import time

import numpy as np

DEFAULT = {"potential": (12, 512), "force": (24, 192)}
CACHE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "tuned.json")
ORDERS = (6, 8, 9, 10, 12, 14, 16, 20, 24, 32)
# tiling option per kind, the values a sweep tries and the kernel's default
TILING = {"potential": ("OPT_POT_XCD_CHUNK", (1, 4, 16, 64)), "force": ("OPT_FORCE_INLINE", (16, 32, 64, 128))}
TILING_DEFAULT = {"potential": 16, "force": 64}
TILE_MARGIN = 0.01  # another tiling replaces the default only when at least this much faster
LEAVES = (64, 128, 192, 256, 384, 512, 768, 1024)


def bucket_key(kind, precision, theta, n, m):
    return "%s/p%d/t%.2f/n%d/m%d" % (kind, precision, theta, int(round(math.log2(max(n, 1)))),
                                     int(round(math.log2(max(m, 1)))))


class Tuner:
    def __init__(self, pkg, cache_path=CACHE, device=0):
        self.pkg = pkg
        self.device = device
        self.cache_path = cache_path
        self.cache = {}
        if cache_path and os.path.exists(cache_path):
            with open(cache_path) as f:
                self.cache = json.load(f)

    def lookup(self, kind, precision, theta, n, m):
        """Tuned (order, leaf) for this size bucket, or None."""
        e = self.cache.get(bucket_key(kind, precision, theta, n, m))
        return (e["order"], e["leaf"]) if e else None

    def lookup_tile(self, kind, precision, theta, n, m):
        """(option key, value) of the tuned tiling for this size bucket, or None."""
        e = self.cache.get(bucket_key(kind, precision, theta, n, m))
        if not e or not e.get("tile"):
            return None
        return getattr(self.pkg, TILING[kind][0]), e["tile"]

    def nearest(self, kind, precision, theta, n, m, max_dist=3):
        """(order, leaf) of the tuned bucket of the same kind, precision and theta
        whose log2 N and log2 M are closest (sum of distances at most max_dist),
        or None.  For a server: requests of every size, tuned on a few buckets."""
        prefix = "%s/p%d/t%.2f/" % (kind, precision, theta)
        ln, lm = math.log2(max(n, 1)), math.log2(max(m, 1))
        best, best_d = None, None
        for key, e in self.cache.items():
            if not key.startswith(prefix):
                continue
            bn, bm = key[len(prefix):].split("/")
            d = abs(ln - int(bn[1:])) + abs(lm - int(bm[1:]))
            if d <= max_dist and (best_d is None or d < best_d):
                best, best_d = (e["order"], e["leaf"]), d
        return best

    def _save(self):
        if self.cache_path:
            tmp = self.cache_path + ".tmp"
            with open(tmp, "w") as f:
                json.dump(self.cache, f, indent=1, sort_keys=True)
            os.replace(tmp, self.cache_path)

    def tune(self, kind, data, theta=0.5, precision=1, orders=ORDERS, leaves=LEAVES, err_target=None,
             nsample=1024, reps=3, seed=0, log=None, save=True, tilings=None):
        """data: potential {xs, ys, vs, xt, yt}; force {xs, ys, vs, x0s, y0s, hs}.
        Returns {"order", "leaf", "ms", "err", "table": [...]} (with tilings also "tile" and
        "tile_table") and records it in the cache."""
        pkg = self.pkg
        ctx = pkg.Context(device=self.device, precision=precision)
        d = {k: ctx.to_device(v) for k, v in data.items()}
        rng = np.random.default_rng(seed)
        if kind == "potential":
            m = len(data["xt"])
            idx = np.sort(rng.choice(m, min(nsample, m), replace=False)).astype(np.int32)
            ref = ctx.direct_potential(d["xs"], d["ys"], d["vs"], d["xt"], d["yt"], ctx.to_device(idx, np.int32))
            out = ctx.empty(m)
        else:
            m = len(data["x0s"]) * 1024
            idx = np.sort(rng.choice(m, min(nsample, m), replace=False)).astype(np.int32)
            b = idx // 1024
            tx = data["x0s"][b] + (idx % 32) * data["hs"][b]
            ty = data["y0s"][b] + ((idx // 32) % 32) * data["hs"][b]
            rx, ry = ctx.direct_force(d["xs"], d["ys"], d["vs"], ctx.to_device(tx), ctx.to_device(ty),
                                      ctx.to_device(np.arange(len(idx), dtype=np.int32), np.int32))
            xr, yr = ctx.empty(m), ctx.empty(m)

        def run():
            if kind == "potential":
                ctx.potential(theta, d["xs"], d["ys"], d["vs"], d["xt"], d["yt"], out=out, device=True)
            else:
                ctx.force(theta, d["xs"], d["ys"], d["vs"], d["x0s"], d["y0s"], d["hs"], xr=xr, yr=yr, device=True)
            ctx.synchronize()

        def error():
            if kind == "potential":
                r = out.numpy()[idx]
                return float(np.linalg.norm(r - ref) / np.linalg.norm(ref))
            fx, fy = xr.numpy()[idx], yr.numpy()[idx]
            return float(np.sqrt(np.sum((fx - rx) ** 2) + np.sum((fy - ry) ** 2)) /
                         np.sqrt(np.sum(rx ** 2) + np.sum(ry ** 2)))

        okey, lkey = (pkg.OPT_POT_ORDER, pkg.OPT_POT_LEAF) if kind == "potential" else \
            (pkg.OPT_FORCE_ORDER, pkg.OPT_FORCE_LEAF)
        table = []
        for p in orders:
            for leaf in leaves:
                ctx.set(okey, p)
                ctx.set(lkey, leaf)
                run()
                err = error()
                ts = []
                for _ in range(reps):
                    t0 = time.perf_counter()
                    run()
                    ts.append(time.perf_counter() - t0)
                table.append({"order": p, "leaf": leaf, "ms": 1e3 * min(ts), "err": err})
                if log:
                    log("tune %s p=%d leaf=%d: %.3f ms err %.2e" % (kind, p, leaf, table[-1]["ms"], err))
        p0, l0 = DEFAULT[kind]
        base = [t for t in table if t["order"] == p0 and t["leaf"] == l0]
        if err_target is None:
            err_target = min(1e-6, max(base[0]["err"], 1e-15) * 1.01) if base else 1e-6
        ok = [t for t in table if t["err"] <= err_target]
        met = bool(ok)
        if not met:
            # nothing in the grid meets the target: keep the reference's parameters (if
            # measured, else the most accurate cell) and say so, rather than the fastest
            ok = base or [min(table, key=lambda t: t["err"])]
            if log:
                log("tune %s: no cell within %.2e; keeping order %d leaf %d" % (kind, err_target, ok[0]["order"],
                                                                              ok[0]["leaf"]))
        best = min(ok, key=lambda t: t["ms"])
        res = {"order": best["order"], "leaf": best["leaf"], "ms": best["ms"], "err": best["err"],
               "err_target": err_target, "met_target": met, "table": table}
        if tilings:
            # the chosen pair under each tiling value; the tiling moves results only
            # by rounding (summation order) or not at all, so the error is rechecked
            # but the choice is by time
            tkey = getattr(pkg, TILING[kind][0])
            ctx.set(okey, best["order"])
            ctx.set(lkey, best["leaf"])
            ttable = []
            for tv in tilings:
                ctx.set(tkey, tv)
                run()
                err = error()
                ts = []
                for _ in range(reps):
                    t0 = time.perf_counter()
                    run()
                    ts.append(time.perf_counter() - t0)
                ttable.append({"tile": tv, "ms": 1e3 * min(ts), "err": err})
                if log:
                    log("tune %s tile %s=%d: %.3f ms err %.2e" % (kind, TILING[kind][0], tv, ttable[-1]["ms"], err))
            ctx.set(tkey, 0)
            okt = [t for t in ttable if t["err"] <= max(err_target, best["err"] * 1.01)] or ttable
            bt = min(okt, key=lambda t: t["ms"])
            dflt = [t for t in okt if t["tile"] == TILING_DEFAULT[kind]]
            if dflt and bt["ms"] > dflt[0]["ms"] * (1.0 - TILE_MARGIN):
                bt = dflt[0]  # within timing noise of the default
            res.update(tile=bt["tile"], ms=bt["ms"], err=bt["err"], tile_table=ttable)
        n = len(data["xs"])
        if save:
            self.cache[bucket_key(kind, precision, theta, n, m)] = {
                k: res[k] for k in ("order", "leaf", "ms", "err", "err_target", "met_target", "tile") if k in res}
            self._save()
        ctx.close()
        return res
