"""Deterministic synthetic inputs for the treecode workloads.

Every array comes from a counter-based splitmix64 stream, so the same
(config, seed) gives bit-identical inputs on any host; the GPU path and the
CPU oracle always read the same numpy arrays.

Workloads (sizes as named):
  uniform_potential(n)        N = M sources/targets uniform in [0,1)^2, v in [-1,1);
                              targets are a separate copy of the source coordinates
  gaussian4_force(n)          4 isotropic Gaussians (sigma 0.05) at (0.3|0.7, 0.3|0.7),
                              strengths s_g*exp(-r^2/(2 sigma^2)), random sign per Gaussian;
                              64x64 blocks of 32x32 points, h = 1/2048, tiling [0,1)^2
  mixed_plummer_potential(n)  N/2 uniform + N/2 Plummer (a = 0.02) about (0.5, 0.5),
                              rejection-truncated to [0,1]^2, v in [-1,1)
  gaussian16_force(n, nb)     16 Gaussians, sigma log-uniform in [0.005, 0.05], centres in
                              [0.1,0.9]^2; nb blocks at h in {1/1024, 1/2048, 1/4096}
                              placed on the densest cells of each resolution
  request_stream(...)         the served mixed potential/force request stream
"""
import math

import numpy as np

_GOLDEN = np.uint64(0x9E3779B97F4A7C15)
_M1 = np.uint64(0xBF58476D1CE4E5B9)
_M2 = np.uint64(0x94D049BB133111EB)


def _splitmix64(seed, start, n):
    with np.errstate(over="ignore"):
        z = (np.arange(start + 1, start + n + 1, dtype=np.uint64) * _GOLDEN) + np.uint64(seed) * _M2
        z = (z ^ (z >> np.uint64(30))) * _M1
        z = (z ^ (z >> np.uint64(27))) * _M2
        z = z ^ (z >> np.uint64(31))
    return z


class Stream:
    """Counter-based stream of uniform doubles in [0, 1)."""

    def __init__(self, seed):
        self.seed = int(seed) & 0xFFFFFFFFFFFFFFFF
        self.pos = 0

    def uniform(self, n, lo=0.0, hi=1.0):
        u = (_splitmix64(self.seed, self.pos, n) >> np.uint64(11)).astype(np.float64) * (1.0 / 9007199254740992.0)
        self.pos += n
        if lo == 0.0 and hi == 1.0:
            return u
        return lo + (hi - lo) * u

    def normal(self, n):
        u1 = self.uniform(n)
        u2 = self.uniform(n)
        return np.sqrt(-2.0 * np.log1p(-u1)) * np.cos(2.0 * math.pi * u2)

    def integers(self, n, k):
        return np.minimum((self.uniform(n) * k).astype(np.int64), k - 1)


def uniform_potential(n, seed=1451):
    st = Stream(seed)
    xs = st.uniform(n)
    ys = st.uniform(n)
    vs = st.uniform(n, -1.0, 1.0)
    return xs, ys, vs, xs.copy(), ys.copy()


def _gaussian_mixture(st, n, centres, sigmas):
    g = st.integers(n, len(centres))
    signs = np.where(st.uniform(len(centres)) < 0.5, -1.0, 1.0)
    dx = st.normal(n) * sigmas[g]
    dy = st.normal(n) * sigmas[g]
    xs = centres[g, 0] + dx
    ys = centres[g, 1] + dy
    vs = signs[g] * np.exp(-(dx * dx + dy * dy) / (2.0 * sigmas[g] ** 2))
    return xs, ys, vs


def mrag_grid_blocks(nbx, nby, h, x0=0.0, y0=0.0):
    """nbx*nby blocks of 32x32 points with spacing h tiling from (x0, y0)."""
    i, j = np.meshgrid(np.arange(nbx), np.arange(nby), indexing="xy")
    bx = x0 + i.ravel() * (32 * h)
    by = y0 + j.ravel() * (32 * h)
    return bx.astype(np.float64), by.astype(np.float64), np.full(nbx * nby, h, dtype=np.float64)


def gaussian4_force(n, seed=1452, nbx=64, h=1.0 / 2048):
    st = Stream(seed)
    centres = np.array([[0.3, 0.3], [0.7, 0.3], [0.3, 0.7], [0.7, 0.7]])
    xs, ys, vs = _gaussian_mixture(st, n, centres, np.full(4, 0.05))
    x0s, y0s, hs = mrag_grid_blocks(nbx, nbx, h)
    return xs, ys, vs, x0s, y0s, hs


def mixed_plummer_potential(n, seed=1453, a=0.02):
    st = Stream(seed)
    nu = n // 2
    npl = n - nu
    xs = np.empty(n)
    ys = np.empty(n)
    xs[:nu] = st.uniform(nu)
    ys[:nu] = st.uniform(nu)
    todo = np.arange(nu, n)
    while len(todo):
        k = len(todo)
        u = st.uniform(k)
        rad = a * np.sqrt(u / (1.0 - u))
        ang = 2.0 * math.pi * st.uniform(k)
        px = 0.5 + rad * np.cos(ang)
        py = 0.5 + rad * np.sin(ang)
        ok = (px >= 0) & (px <= 1) & (py >= 0) & (py <= 1)
        xs[todo[ok]] = px[ok]
        ys[todo[ok]] = py[ok]
        todo = todo[~ok]
    del npl
    vs = st.uniform(n, -1.0, 1.0)
    return xs, ys, vs, xs.copy(), ys.copy()


def gaussian16_force(n, nblocks=16384, seed=1455):
    st = Stream(seed)
    centres = 0.1 + 0.8 * st.uniform(32).reshape(16, 2)
    sig = np.exp(np.log(0.005) + (np.log(0.05) - np.log(0.005)) * st.uniform(16))
    xs, ys, vs = _gaussian_mixture(st, n, centres, sig)
    # blocks on the densest cells of each spacing: a quarter at h=1/1024 and half at
    # 1/2048 (at most all 32^2 and 64^2 cells of [0,1]^2), the rest at h=1/4096 (at
    # most 128^2), so that nblocks <= 21504 blocks exist and finer blocks sit where
    # the particles are densest
    coarse = min(nblocks // 4, 32 * 32)
    mid = min(nblocks // 2, 64 * 64)
    fine = nblocks - coarse - mid
    if fine > 128 * 128:
        raise ValueError("gaussian16_force: at most 21504 blocks")
    counts = [fine, mid, coarse]
    x0l, y0l, hl = [], [], []
    for h, cnt in zip([1.0 / 4096, 1.0 / 2048, 1.0 / 1024], counts):
        side = 32 * h
        nc = int(round(1.0 / side))
        ix = np.clip((xs / side).astype(np.int64), 0, nc - 1)
        iy = np.clip((ys / side).astype(np.int64), 0, nc - 1)
        dens = np.bincount(iy * nc + ix, minlength=nc * nc)
        order = np.argsort(-dens, kind="stable")[:cnt]
        x0l.append((order % nc) * side)
        y0l.append((order // nc) * side)
        hl.append(np.full(cnt, h))
    return xs, ys, vs, np.concatenate(x0l), np.concatenate(y0l), np.concatenate(hl)


def frame_targets(m, inner=0.25, outer=1.0, seed=1456):
    """m points uniform in the square frame [-outer, 1+outer]^2 minus [-inner, 1+inner]^2."""
    st = Stream(seed)
    x = np.empty(m)
    y = np.empty(m)
    todo = np.arange(m)
    while len(todo):
        k = len(todo)
        px = -outer + (1.0 + 2.0 * outer) * st.uniform(k)
        py = -outer + (1.0 + 2.0 * outer) * st.uniform(k)
        ok = ~((px > -inner) & (px < 1.0 + inner) & (py > -inner) & (py < 1.0 + inner))
        x[todo[ok]] = px[ok]
        y[todo[ok]] = py[ok]
        todo = todo[~ok]
    return x, y


def request_stream(nreq=10000, rate_hz=2000.0, seed=1454, pot_frac=0.7, reuse_frac=0.25,
                   pot_n=(1e3, 1e6), force_n=(1e4, 2e6), force_nb=(16, 2048), thetas=(0.3, 0.5, 0.7)):
    """Request descriptors of the served stream (inputs are generated lazily from the seeds)."""
    st = Stream(seed)
    gaps = -np.log1p(-st.uniform(nreq)) / rate_hz
    arrival = np.cumsum(gaps)
    kind = np.where(st.uniform(nreq) < pot_frac, 0, 1)
    ul = st.uniform(nreq)
    ub = st.uniform(nreq)
    th = st.integers(nreq, len(thetas))
    reuse = st.uniform(nreq) < reuse_frac
    gen = st.integers(nreq, 2)
    reqs = []
    prev = None
    for i in range(nreq):
        if kind[i] == 0:
            n = int(round(math.exp(math.log(pot_n[0]) + ul[i] * (math.log(pot_n[1]) - math.log(pot_n[0])))))
            nb = 0
        else:
            n = int(round(math.exp(math.log(force_n[0]) + ul[i] * (math.log(force_n[1]) - math.log(force_n[0])))))
            nb = int(round(math.exp(math.log(force_nb[0]) + ub[i] * (math.log(force_nb[1]) - math.log(force_nb[0])))))
        src_seed = 100000 + i
        g = "uniform" if gen[i] == 0 else "gauss4"
        if reuse[i] and prev is not None:
            n, src_seed, g = prev["n"], prev["src_seed"], prev["gen"]
        r = {"id": i, "arrival": float(arrival[i]), "kind": "potential" if kind[i] == 0 else "force",
             "n": n, "nblocks": nb, "theta": float(thetas[th[i]]), "src_seed": src_seed,
             "gen": g, "tgt_seed": 200000 + i,
             "reuses_sources": bool(reuse[i] and prev is not None)}
        reqs.append(r)
        prev = r
    return reqs


def request_pools(seed=1454, size=2_000_000):
    """Pre-generated source pools (uniform and 4-Gaussian) of the served stream."""
    st = Stream(seed + 7)
    uni = (st.uniform(size), st.uniform(size), st.uniform(size, -1.0, 1.0))
    centres = np.array([[0.3, 0.3], [0.7, 0.3], [0.3, 0.7], [0.7, 0.7]])
    gau = _gaussian_mixture(st, size, centres, np.full(4, 0.05))
    return {"uniform": uni, "gauss4": gau, "size": size}


def pool_sources(pools, r):
    """Sources of request r as a window of the pool chosen by its source seed."""
    size = pools["size"]
    n = min(r["n"], size)
    off = (r["src_seed"] * 2654435761) % (size - n + 1)
    xs, ys, vs = pools[r["gen"]]
    return xs[off:off + n], ys[off:off + n], vs[off:off + n]


def request_sources(r):
    st = Stream(r["src_seed"])
    n = r["n"]
    if r["gen"] == "uniform":
        xs, ys, vs = st.uniform(n), st.uniform(n), st.uniform(n, -1.0, 1.0)
    else:
        centres = np.array([[0.3, 0.3], [0.7, 0.3], [0.3, 0.7], [0.7, 0.7]])
        xs, ys, vs = _gaussian_mixture(st, n, centres, np.full(4, 0.05))
    return xs, ys, vs


def request_targets(r, xs, ys):
    """Potential: a copy of the sources; force: nblocks blocks at h=1/2048 on random grid cells."""
    if r["kind"] == "potential":
        return xs.copy(), ys.copy()
    st = Stream(r["tgt_seed"])
    nb = r["nblocks"]
    cells = st.integers(nb, 64 * 64)
    h = 1.0 / 2048
    return (cells % 64) * (32 * h), (cells // 64) * (32 * h), np.full(nb, h)
