"""Writes the small fixtures under tests/golden/ (run: python tests/golden/make_fixtures.py).

morton_edges.npz
    Points within a few ulps of Morton cell edges, where the quantisation of
    drivers/sort-sources.cpp:117, floor((x - xmin) / ext * 2^15), and the form
    the reference's release flags (-O3 --fast-math) compile it into,
    floor((x - xmin) * ((1/ext) * 2^15)), give different cells.  Expected keys
    follow the compiled form; `keys_source_form` holds the other one, so the
    fixture shows which of the two an implementation follows.  Keys are
    interleaved bit by bit with plain Python integers (x in even bits).
binomials.npz
    C(n, k) for n < 64 as exact integers (Python math.comb) and the nearest
    doubles; the E2E/E2L constants of kernels/upward-kernels.m4:128-130 and
    kernels/force-kernels.m4:138-141.
neg_inv.npz
    -1/l for l = 1..64 as the nearest double and as the 20-digit decimal that
    `bc -l` prints (upward-kernels.m4:128, force-kernels.m4:138), parsed to double.
"""
import math
import os
import sys

import numpy as np

OUT = os.path.dirname(os.path.abspath(__file__))
DBL_EPS = 2.220446049250313e-16


def interleave(qx, qy):
    k = 0
    for b in range(15):
        k |= ((qx >> b) & 1) << (2 * b)
        k |= ((qy >> b) & 1) << (2 * b + 1)
    return k


def bbox(xs, ys):
    xmi = min([1e13] + list(xs))
    ymi = min([1e13] + list(ys))
    xma = max([0.0] + list(xs))
    yma = max([0.0] + list(ys))
    eps = 10000 * DBL_EPS
    ext = max(xma - xmi, yma - ymi) * (1 + 2 * eps)
    return xmi - eps * ext, ymi - eps * ext, ext


def q_compiled(v, vmin, ext):
    s = (1.0 / ext) * 32768.0
    return min(max(int(math.floor((v - vmin) * s)), 0), 32767)


def q_source(v, vmin, ext):
    return min(max(int(math.floor((v - vmin) / ext * 32768.0)), 0), 32767)


def morton_edges():
    rng = np.random.default_rng(1451)
    anchors_x = [0.0, 0.999, 0.5]
    anchors_y = [0.0, 0.999, 0.25]
    xmin, ymin, ext = bbox(anchors_x, anchors_y)
    xs, ys = list(anchors_x), list(anchors_y)
    differ = 0
    tries = 0
    while differ < 64 and tries < 200000:
        tries += 1
        k = int(rng.integers(40, 32700))
        base = xmin + k * ext / 32768.0
        v = base
        for step in range(-4, 5):
            v = base
            for _ in range(abs(step)):
                v = math.nextafter(v, math.inf if step > 0 else -math.inf)
            if q_compiled(v, xmin, ext) != q_source(v, xmin, ext):
                xs.append(v)
                ys.append(float(rng.uniform(0.01, 0.98)))
                differ += 1
                break
    # points exactly at representable cell edges and random points
    for k in rng.integers(1, 32767, 64):
        xs.append(xmin + float(k) * (ext / 32768.0))
        ys.append(ymin + float(rng.integers(1, 32767)) * (ext / 32768.0))
    xs += list(rng.uniform(0.001, 0.998, 128))
    ys += list(rng.uniform(0.001, 0.998, 128))
    assert bbox(xs, ys) == (xmin, ymin, ext)
    keys = [interleave(q_compiled(x, xmin, ext), q_compiled(y, ymin, ext)) for x, y in zip(xs, ys)]
    keys_src = [interleave(q_source(x, xmin, ext), q_source(y, ymin, ext)) for x, y in zip(xs, ys)]
    np.savez(os.path.join(OUT, "morton_edges.npz"), x=np.array(xs), y=np.array(ys),
             box=np.array([xmin, ymin, ext]), keys=np.array(keys, dtype=np.int32),
             keys_source_form=np.array(keys_src, dtype=np.int32), n_differ=np.int64(differ))
    print("morton_edges: %d points, %d where the two forms differ" % (len(xs), differ))


def binomials():
    exact = np.zeros((64, 64), dtype=object)
    dbl = np.zeros((64, 64))
    for n in range(64):
        for k in range(n + 1):
            exact[n, k] = math.comb(n, k)
            dbl[n, k] = float(math.comb(n, k))
    hi = np.array([[int(exact[n, k]) >> 32 if k <= n else 0 for k in range(64)] for n in range(64)], np.int64)
    lo = np.array([[int(exact[n, k]) & 0xFFFFFFFF if k <= n else 0 for k in range(64)] for n in range(64)],
                  np.int64)
    np.savez(os.path.join(OUT, "binomials.npz"), hi=hi, lo=lo, dbl=dbl)
    print("binomials: C(n,k), n < 64")


def neg_inv():
    from decimal import Decimal, getcontext
    getcontext().prec = 60
    near = np.array([-(1.0 / l) for l in range(1, 65)])
    bc = []
    for l in range(1, 65):
        q = Decimal(1) / Decimal(l)
        s = "-" + ("%.20f" % 0)[:0] + str(q.quantize(Decimal("1e-20"), rounding="ROUND_DOWN"))
        bc.append(float(s))
    np.savez(os.path.join(OUT, "neg_inv.npz"), nearest=near, bc20=np.array(bc))
    print("neg_inv: %d of 64 bc constants equal the nearest double" % int(np.sum(near == np.array(bc))))


if __name__ == "__main__":
    morton_edges()
    binomials()
    neg_inv()
    sys.exit(0)
