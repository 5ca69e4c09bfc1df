"""ctypes binding of the CPU oracle (tc_oracle.c).

TEST INFRASTRUCTURE ONLY: tests/, the smoke check and bench.py's CPU-baseline
leg import this module as the checker.  The product path never does.
"""
import ctypes
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liboracle_tc.so")

CNT_P2P, CNT_E2P, CNT_E2L, CNT_L2P = 0, 1, 2, 3

_lib = None

_D = ctypes.POINTER(ctypes.c_double)
_I = ctypes.POINTER(ctypes.c_int)
_LL = ctypes.POINTER(ctypes.c_longlong)


def build():
    subprocess.run(["make", "-s", "-C", HERE], check=True)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        L = ctypes.CDLL(LIB_PATH)
        L.orc_bbox.argtypes = [_D, _D, ctypes.c_int, _D]
        L.orc_keys.argtypes = [_D, _D, ctypes.c_int, ctypes.c_double, ctypes.c_double, ctypes.c_double, _I]
        L.orc_sort.argtypes = [_I, ctypes.c_int, _I, _I]
        L.orc_tree_build.argtypes = [_D, _D, _D, ctypes.c_int, ctypes.c_int, ctypes.c_int]
        L.orc_tree_build.restype = ctypes.c_void_p
        L.orc_tree_free.argtypes = [ctypes.c_void_p]
        L.orc_tree_nnodes.argtypes = [ctypes.c_void_p]
        L.orc_tree_export.argtypes = [ctypes.c_void_p, _D, _I, _I, _D, _I, _I, _I, _I, _D, _D, _D]
        L.orc_potential.argtypes = [ctypes.c_double, _D, _D, _D, ctypes.c_int, _D, _D, ctypes.c_int, _D,
                                    ctypes.c_int, ctypes.c_int, _LL, ctypes.c_int]
        L.orc_potential_tree.argtypes = [ctypes.c_void_p, ctypes.c_double, _D, _D, ctypes.c_int, ctypes.c_int,
                                         _D, _LL, ctypes.c_int]
        L.orc_force_mrag.argtypes = [ctypes.c_double, _D, _D, _D, ctypes.c_int, _D, _D, _D, ctypes.c_int,
                                     _D, _D, ctypes.c_int, ctypes.c_int, _LL, ctypes.c_int]
        L.orc_force_tree.argtypes = [ctypes.c_void_p, ctypes.c_double, _D, _D, _D, ctypes.c_int, ctypes.c_int,
                                     _D, _D, _LL, ctypes.c_int]
        L.orc_force_targets.argtypes = [ctypes.c_double, _D, _D, _D, ctypes.c_int, _D, _D, ctypes.c_int, _D, _D,
                                        ctypes.c_int, ctypes.c_int, _LL, ctypes.c_int]
        L.orc_direct_potential.argtypes = [_D, _D, _D, ctypes.c_int, _D, _D, _I, ctypes.c_int, _D, ctypes.c_int]
        L.orc_direct_force.argtypes = [_D, _D, _D, ctypes.c_int, _D, _D, _I, ctypes.c_int, _D, _D, ctypes.c_int]
        L.orc_p2e.argtypes = [_D, _D, _D, ctypes.c_int, ctypes.c_double, ctypes.c_double, ctypes.c_int, _D, _D]
        L.orc_e2p_potential.argtypes = [ctypes.c_double, ctypes.c_double, ctypes.c_double, _D, _D, ctypes.c_int]
        L.orc_e2p_potential.restype = ctypes.c_double
        L.orc_e2p_force.argtypes = [ctypes.c_double, ctypes.c_double, ctypes.c_double, _D, _D, ctypes.c_int, _D, _D]
        L.orc_e2l.argtypes = [ctypes.c_double, ctypes.c_double, ctypes.c_double, _D, _D, ctypes.c_int, _D, _D]
        L.orc_l2p.argtypes = [ctypes.c_double, ctypes.c_double, _D, _D, ctypes.c_int, _D, _D]
        L.orc_binomial.argtypes = [ctypes.c_int, ctypes.c_int]
        L.orc_binomial.restype = ctypes.c_double
        L.orc_max_threads.restype = ctypes.c_int
        _lib = L
    return _lib


def _d(a):
    return a.ctypes.data_as(_D)


def _i(a):
    return a.ctypes.data_as(_I)


def _f64(a):
    return np.ascontiguousarray(a, dtype=np.float64)


def max_threads():
    return lib().orc_max_threads()


def set_threads(n):
    """Threads of the following OpenMP regions of the oracle (tree builds included)."""
    lib().orc_set_threads(int(n))


def bbox(x, y):
    x, y = _f64(x), _f64(y)
    out = np.zeros(3)
    lib().orc_bbox(_d(x), _d(y), len(x), _d(out))
    return out


def keys(x, y, box):
    x, y = _f64(x), _f64(y)
    k = np.zeros(len(x), dtype=np.int32)
    lib().orc_keys(_d(x), _d(y), len(x), float(box[0]), float(box[1]), float(box[2]), _i(k))
    return k


def sort(k):
    k = np.ascontiguousarray(k, dtype=np.int32)
    perm = np.zeros(len(k), dtype=np.int32)
    ks = np.zeros(len(k), dtype=np.int32)
    lib().orc_sort(_i(k), len(k), _i(perm), _i(ks))
    return perm, ks


class Tree:
    """Oracle tree (Tree::build restated) with its arrays exported as numpy."""

    def __init__(self, xs, ys, vs, leafcap, order):
        self.xs, self.ys, self.vs = _f64(xs), _f64(ys), _f64(vs)
        self.n, self.order, self.leafcap = len(self.xs), order, leafcap
        self.h = lib().orc_tree_build(_d(self.xs), _d(self.ys), _d(self.vs), self.n, leafcap, order)
        nn = lib().orc_tree_nnodes(self.h)
        n = self.n
        self.nnodes = nn
        self.bbox = np.zeros(3)
        self.keys = np.zeros(n, np.int32)
        self.perm = np.zeros(n, np.int32)
        self.xyv = np.zeros(3 * n)
        self.se = np.zeros(2 * nn, np.int32)
        self.child = np.zeros(nn, np.int32)
        self.level = np.zeros(nn, np.int32)
        self.cell = np.zeros(2 * nn, np.int32)
        self.geom = np.zeros(4 * nn)
        self.w = np.zeros(3 * nn)
        self.exp = np.zeros(2 * order * nn)
        lib().orc_tree_export(self.h, _d(self.bbox), _i(self.keys), _i(self.perm), _d(self.xyv), _i(self.se),
                              _i(self.child), _i(self.level), _i(self.cell), _d(self.geom), _d(self.w),
                              _d(self.exp))
        self.se = self.se.reshape(nn, 2)
        self.cell = self.cell.reshape(nn, 2)
        self.geom = self.geom.reshape(nn, 4)
        self.w = self.w.reshape(nn, 3)
        self.exp = self.exp.reshape(nn, 2, order)

    def potential(self, theta, xt, yt, nthreads=0):
        xt, yt = _f64(xt), _f64(yt)
        out = np.zeros(len(xt))
        cnt = np.zeros(4, np.int64)
        lib().orc_potential_tree(self.h, theta, _d(xt), _d(yt), 0, len(xt), _d(out),
                                 cnt.ctypes.data_as(_LL), nthreads)
        return out, cnt

    def force(self, theta, x0s, y0s, hs, nthreads=0):
        x0s, y0s, hs = _f64(x0s), _f64(y0s), _f64(hs)
        nb = len(x0s)
        xr = np.zeros(nb * 1024)
        yr = np.zeros(nb * 1024)
        cnt = np.zeros(4, np.int64)
        lib().orc_force_tree(self.h, theta, _d(x0s), _d(y0s), _d(hs), 0, nb, _d(xr), _d(yr),
                             cnt.ctypes.data_as(_LL), nthreads)
        return xr, yr, cnt

    def __del__(self):
        if getattr(self, "h", None):
            lib().orc_tree_free(self.h)
            self.h = None


def potential(theta, xs, ys, vs, xt, yt, order=12, leafcap=512, nthreads=0):
    xs, ys, vs, xt, yt = map(_f64, (xs, ys, vs, xt, yt))
    out = np.zeros(len(xt))
    cnt = np.zeros(4, np.int64)
    lib().orc_potential(theta, _d(xs), _d(ys), _d(vs), len(xs), _d(xt), _d(yt), len(xt), _d(out),
                        order, leafcap, cnt.ctypes.data_as(_LL), nthreads)
    return out, cnt


def force_mrag(theta, xs, ys, vs, x0s, y0s, hs, order=24, leafcap=192, nthreads=0):
    xs, ys, vs, x0s, y0s, hs = map(_f64, (xs, ys, vs, x0s, y0s, hs))
    nb = len(x0s)
    xr = np.zeros(nb * 1024)
    yr = np.zeros(nb * 1024)
    cnt = np.zeros(4, np.int64)
    lib().orc_force_mrag(theta, _d(xs), _d(ys), _d(vs), len(xs), _d(x0s), _d(y0s), _d(hs), nb, _d(xr), _d(yr),
                         order, leafcap, cnt.ctypes.data_as(_LL), nthreads)
    return xr, yr, cnt


def force_targets(theta, xs, ys, vs, xt, yt, order=24, leafcap=192, nthreads=0):
    xs, ys, vs, xt, yt = map(_f64, (xs, ys, vs, xt, yt))
    xr = np.zeros(len(xt))
    yr = np.zeros(len(xt))
    cnt = np.zeros(4, np.int64)
    lib().orc_force_targets(theta, _d(xs), _d(ys), _d(vs), len(xs), _d(xt), _d(yt), len(xt), _d(xr), _d(yr),
                            order, leafcap, cnt.ctypes.data_as(_LL), nthreads)
    return xr, yr, cnt


def direct_potential(xs, ys, vs, xt, yt, idx, nthreads=0):
    xs, ys, vs, xt, yt = map(_f64, (xs, ys, vs, xt, yt))
    idx = np.ascontiguousarray(idx, dtype=np.int32)
    out = np.zeros(len(idx))
    lib().orc_direct_potential(_d(xs), _d(ys), _d(vs), len(xs), _d(xt), _d(yt), _i(idx), len(idx), _d(out),
                               nthreads)
    return out


def direct_force(xs, ys, vs, xt, yt, idx, nthreads=0):
    xs, ys, vs, xt, yt = map(_f64, (xs, ys, vs, xt, yt))
    idx = np.ascontiguousarray(idx, dtype=np.int32)
    ox = np.zeros(len(idx))
    oy = np.zeros(len(idx))
    lib().orc_direct_force(_d(xs), _d(ys), _d(vs), len(xs), _d(xt), _d(yt), _i(idx), len(idx), _d(ox), _d(oy),
                           nthreads)
    return ox, oy


def p2e(x, y, v, xc, yc, order):
    x, y, v = map(_f64, (x, y, v))
    re = np.zeros(order)
    im = np.zeros(order)
    lib().orc_p2e(_d(x), _d(y), _d(v), len(x), xc, yc, order, _d(re), _d(im))
    return re + 1j * im


def e2p_potential(z, mass, a):
    re, im = _f64(a.real), _f64(a.imag)
    return lib().orc_e2p_potential(z.real, z.imag, mass, _d(re), _d(im), len(re))


def e2p_force(z, mass, a):
    re, im = _f64(a.real), _f64(a.imag)
    fx = np.zeros(1)
    fy = np.zeros(1)
    lib().orc_e2p_force(mass, z.real, z.imag, _d(re), _d(im), len(re), _d(fx), _d(fy))
    return fx[0], fy[0]


def e2l(z0, mass, a):
    re, im = _f64(a.real), _f64(a.imag)
    p = len(re)
    lre = np.zeros(p + 1)
    lim = np.zeros(p + 1)
    lib().orc_e2l(z0.real, z0.imag, mass, _d(re), _d(im), p, _d(lre), _d(lim))
    return lre + 1j * lim


def l2p(z, b):
    lre, lim = _f64(b.real), _f64(b.imag)
    fx = np.zeros(1)
    fy = np.zeros(1)
    lib().orc_l2p(z.real, z.imag, _d(lre), _d(lim), len(lre) - 1, _d(fx), _d(fy))
    return fx[0], fy[0]


def binomial(n, k):
    return lib().orc_binomial(n, k)
