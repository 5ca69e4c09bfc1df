"""Python host mirror of the 2d-treecodes interface over libtreecodes2d.so.

The reference exposes two C functions (README.md:5-25):
    treecode_potential(theta, xs, ys, vs, n, xt, yt, m, out)
    treecode_force_mrag(theta, xs, ys, vs, n, x0s, y0s, hs, nblocks, xr, yr)
`treecode_potential` / `treecode_force_mrag` below call exactly those C
symbols (host numpy arrays in, results out).  `Context` exposes the
extensions: precision, order and leaf options, device-resident buffers
(DeviceArray buffers in HBM), phase timers, interaction counters and the
resident tree.

There is no CPU fallback: importing works without a GPU, but every call
raises TreecodeError when the HIP library or a device is missing.
"""
import ctypes
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libtreecodes2d.so")

(OPT_PRECISION, OPT_POT_ORDER, OPT_POT_LEAF, OPT_FORCE_ORDER, OPT_FORCE_LEAF, OPT_PROFILE, OPT_COUNTERS,
 OPT_FORCE_CHUNKS, OPT_POT_XCD_CHUNK, OPT_FORCE_INLINE) = range(10)
PH_BUILD, PH_TSORT, PH_EVAL, PH_TOTAL = range(4)
CNT_P2P, CNT_E2P, CNT_E2L, CNT_L2P = range(4)
FP64, FP32 = 0, 1

EXPORTED = [
    "treecode_potential", "treecode_force_mrag", "treecode_last_error", "tc_context_create",
    "tc_context_destroy", "tc_set_option", "tc_context_stream", "tc_potential", "tc_force_mrag",
    "tc_synchronize", "tc_potential_resident", "tc_force_mrag_resident", "tc_get_stats", "tc_tree_export",
    "tc_morton_sort_device", "tc_device_alloc", "tc_device_free", "tc_copy_to_device", "tc_copy_to_host",
    "tc_direct_potential", "tc_direct_force", "tc_server_create", "tc_server_destroy", "tc_server_submit",
    "tc_server_wait", "tc_server_stats", "treecode_potential_float", "treecode_force_mrag_float",
    "treecode_velocity_mrag", "tc_copy_kernel", "treecode_force", "tc_force_targets", "tc_potential_noalloc",
    "tc_get_profile", "tc_server_create_ex", "tc_server_stats_ex", "tc_potential_batch_resident",
]
SERVER_BATCH = 1  # tc_server_create_ex flag: batch concurrent potential traversals
SERVER_BATCH_ALWAYS = 2  # also a lone traversal goes through the batch kernel (tests its graphs)
NOALLOC_MAX_POINTS = 1 << 18  # sources and targets per tc_potential_noalloc call
NOALLOC_SCRATCH_BYTES = 64 << 20  # its static device scratch


class TreecodeError(RuntimeError):
    pass


class TcRequest(ctypes.Structure):
    """struct tc_request of treecodes2d.h"""
    _fields_ = [("kind", ctypes.c_int), ("precision", ctypes.c_int), ("theta", ctypes.c_double),
                ("xs", ctypes.c_void_p), ("ys", ctypes.c_void_p), ("vs", ctypes.c_void_p), ("n", ctypes.c_int),
                ("xt", ctypes.c_void_p), ("yt", ctypes.c_void_p), ("m", ctypes.c_int),
                ("x0s", ctypes.c_void_p), ("y0s", ctypes.c_void_p), ("hs", ctypes.c_void_p),
                ("nblocks", ctypes.c_int), ("out0", ctypes.c_void_p), ("out1", ctypes.c_void_p),
                ("source_id", ctypes.c_longlong), ("order", ctypes.c_int), ("leaf", ctypes.c_int)]


_lib = None
_D = ctypes.POINTER(ctypes.c_double)
_I = ctypes.POINTER(ctypes.c_int)
_P = ctypes.c_void_p


def lib():
    """Load libtreecodes2d.so (built by build_lib.build()); fail loudly if absent."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise TreecodeError("libtreecodes2d.so is not built (run __graft_entry__.build()); "
                            "there is no CPU fallback")
    L = ctypes.CDLL(LIB_PATH)
    dbl, it = ctypes.c_double, ctypes.c_int
    L.treecode_potential.argtypes = [dbl, _D, _D, _D, it, _D, _D, it, _D]
    L.treecode_potential.restype = None
    L.treecode_force_mrag.argtypes = [dbl, _D, _D, _D, it, _D, _D, _D, it, _D, _D]
    L.treecode_force_mrag.restype = None
    L.treecode_last_error.restype = ctypes.c_char_p
    fl = ctypes.c_float
    _F = ctypes.POINTER(ctypes.c_float)
    if hasattr(L, 'treecode_potential_float'):
        L.treecode_potential_float.argtypes = [fl, _F, _F, _F, it, _F, _F, it, _F]
    if hasattr(L, 'treecode_potential_float'):
        L.treecode_potential_float.restype = None
    if hasattr(L, 'treecode_force_mrag_float'):
        L.treecode_force_mrag_float.argtypes = [fl, _F, _F, _F, it, _F, _F, _F, it, _F, _F]
    if hasattr(L, 'treecode_force_mrag_float'):
        L.treecode_force_mrag_float.restype = None
    if hasattr(L, 'treecode_velocity_mrag'):
        L.treecode_velocity_mrag.argtypes = [dbl, _D, _D, _D, it, _D, _D, _D, it, _D, _D]
    if hasattr(L, 'treecode_velocity_mrag'):
        L.treecode_velocity_mrag.restype = None
    if hasattr(L, 'tc_context_create'):
        L.tc_context_create.argtypes = [it]
    if hasattr(L, 'tc_context_create'):
        L.tc_context_create.restype = _P
    if hasattr(L, 'tc_context_destroy'):
        L.tc_context_destroy.argtypes = [_P]
    if hasattr(L, 'tc_set_option'):
        L.tc_set_option.argtypes = [_P, it, it]
    if hasattr(L, 'tc_context_stream'):
        L.tc_context_stream.argtypes = [_P]
    if hasattr(L, 'tc_context_stream'):
        L.tc_context_stream.restype = _P
    if hasattr(L, 'tc_potential'):
        L.tc_potential.argtypes = [_P, dbl, _P, _P, _P, it, _P, _P, it, _P, it]
    if hasattr(L, 'tc_force_mrag'):
        L.tc_force_mrag.argtypes = [_P, dbl, _P, _P, _P, it, _P, _P, _P, it, _P, _P, it]
    if hasattr(L, 'tc_potential_resident'):
        L.tc_potential_resident.argtypes = [_P, dbl, _P, _P, it, _P, it]
    if hasattr(L, 'tc_force_mrag_resident'):
        L.tc_force_mrag_resident.argtypes = [_P, dbl, _P, _P, _P, it, _P, _P, it]
    if hasattr(L, 'tc_potential_batch_resident'):
        L.tc_potential_batch_resident.argtypes = [_P, it, _D, _P, _P, _I, _P]
    if hasattr(L, 'tc_synchronize'):
        L.tc_synchronize.argtypes = [_P]
    if hasattr(L, 'tc_get_stats'):
        L.tc_get_stats.argtypes = [_P, _D, ctypes.POINTER(ctypes.c_ulonglong), _I]
    if hasattr(L, 'tc_tree_export'):
        L.tc_tree_export.argtypes = [_P, _I, _I, _D, _D, _I, _D]
    if hasattr(L, 'tc_morton_sort_device'):
        L.tc_morton_sort_device.argtypes = [_P, _P, _P, it, _P, _P, _P]
    if hasattr(L, 'tc_device_alloc'):
        L.tc_device_alloc.argtypes = [_P, ctypes.c_size_t]
    if hasattr(L, 'tc_device_alloc'):
        L.tc_device_alloc.restype = _P
    if hasattr(L, 'tc_device_free'):
        L.tc_device_free.argtypes = [_P, _P]
    if hasattr(L, 'tc_copy_to_device'):
        L.tc_copy_to_device.argtypes = [_P, _P, _P, ctypes.c_size_t]
    if hasattr(L, 'tc_copy_to_host'):
        L.tc_copy_to_host.argtypes = [_P, _P, _P, ctypes.c_size_t]
    if hasattr(L, 'tc_server_create'):
        L.tc_server_create.argtypes = [it, it, dbl]
    if hasattr(L, 'tc_server_create'):
        L.tc_server_create.restype = _P
    if hasattr(L, 'tc_server_destroy'):
        L.tc_server_destroy.argtypes = [_P]
    if hasattr(L, 'tc_server_submit'):
        L.tc_server_submit.argtypes = [_P, ctypes.POINTER(TcRequest)]
    if hasattr(L, 'tc_server_submit'):
        L.tc_server_submit.restype = ctypes.c_longlong
    if hasattr(L, 'tc_server_wait'):
        L.tc_server_wait.argtypes = [_P, ctypes.c_longlong, _D]
    if hasattr(L, 'tc_server_stats'):
        L.tc_server_stats.argtypes = [_P, ctypes.POINTER(ctypes.c_longlong)]
    if hasattr(L, 'tc_server_create_ex'):
        L.tc_server_create_ex.argtypes = [it, it, dbl, it]
        L.tc_server_create_ex.restype = _P
    if hasattr(L, 'tc_server_stats_ex'):
        L.tc_server_stats_ex.argtypes = [_P, ctypes.POINTER(ctypes.c_longlong), it]
    if hasattr(L, 'treecode_force'):
        L.treecode_force.argtypes = [dbl, _D, _D, _D, it, _D, _D, it, _D, _D]
        L.treecode_force.restype = None
    if hasattr(L, 'tc_force_targets'):
        L.tc_force_targets.argtypes = [_P, dbl, _P, _P, _P, it, _P, _P, it, _P, _P, it]
    if hasattr(L, 'tc_copy_kernel'):
        L.tc_copy_kernel.argtypes = [_P, _P, _P, ctypes.c_size_t]
    if hasattr(L, 'tc_get_profile'):
        L.tc_get_profile.argtypes = [_P, _D, ctypes.POINTER(ctypes.c_ulonglong)]
    if hasattr(L, 'tc_potential_noalloc'):
        L.tc_potential_noalloc.argtypes = [_P, dbl, _P, _P, _P, it, _P, _P, it, _P]
    if hasattr(L, 'tc_direct_potential'):
        L.tc_direct_potential.argtypes = [_P, _P, _P, _P, it, _P, _P, _P, it, _P]
    if hasattr(L, 'tc_direct_force'):
        L.tc_direct_force.argtypes = [_P, _P, _P, _P, it, _P, _P, _P, it, _P, _P]
    _lib = L
    return L


def last_error():
    return lib().treecode_last_error().decode()


def _f64(a):
    return np.ascontiguousarray(a, dtype=np.float64)


def _dp(a):
    return a.ctypes.data_as(_D)


def treecode_potential(theta, xs, ys, vs, xt, yt):
    """The reference's treecode_potential (drivers/treecode-potential.h:21-29) on host arrays."""
    xs, ys, vs, xt, yt = map(_f64, (xs, ys, vs, xt, yt))
    out = np.empty(len(xt))
    lib().treecode_potential(float(theta), _dp(xs), _dp(ys), _dp(vs), len(xs), _dp(xt), _dp(yt), len(xt), _dp(out))
    if len(out) and np.isnan(out[0]) and last_error():
        raise TreecodeError(last_error())
    return out


def treecode_force_mrag(theta, xs, ys, vs, x0s, y0s, hs):
    """The reference's treecode_force_mrag (drivers/treecode-force.h:21-31) on host arrays;
    returns (xresults, yresults) with point (ix, iy) of block b at ix + 32*(iy + 32*b)."""
    xs, ys, vs, x0s, y0s, hs = map(_f64, (xs, ys, vs, x0s, y0s, hs))
    nb = len(x0s)
    xr = np.empty(nb * 1024)
    yr = np.empty(nb * 1024)
    lib().treecode_force_mrag(float(theta), _dp(xs), _dp(ys), _dp(vs), len(xs), _dp(x0s), _dp(y0s), _dp(hs), nb,
                              _dp(xr), _dp(yr))
    if nb and np.isnan(xr[0]) and last_error():
        raise TreecodeError(last_error())
    return xr, yr


def _f32(a):
    return np.ascontiguousarray(a, dtype=np.float32)


def _fp(a):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_float))


def treecode_potential_float(theta, xs, ys, vs, xt, yt):
    """The reference's real=float build of treecode_potential (libraries.Makefile:13)."""
    xs, ys, vs, xt, yt = map(_f32, (xs, ys, vs, xt, yt))
    out = np.empty(len(xt), dtype=np.float32)
    lib().treecode_potential_float(float(theta), _fp(xs), _fp(ys), _fp(vs), len(xs), _fp(xt), _fp(yt), len(xt),
                                   _fp(out))
    if len(out) and np.isnan(out[0]) and last_error():
        raise TreecodeError(last_error())
    return out


def treecode_force_mrag_float(theta, xs, ys, vs, x0s, y0s, hs):
    xs, ys, vs, x0s, y0s, hs = map(_f32, (xs, ys, vs, x0s, y0s, hs))
    nb = len(x0s)
    xr = np.empty(nb * 1024, dtype=np.float32)
    yr = np.empty(nb * 1024, dtype=np.float32)
    lib().treecode_force_mrag_float(float(theta), _fp(xs), _fp(ys), _fp(vs), len(xs), _fp(x0s), _fp(y0s), _fp(hs),
                                    nb, _fp(xr), _fp(yr))
    if nb and np.isnan(xr[0]) and last_error():
        raise TreecodeError(last_error())
    return xr, yr


def treecode_force(theta, xs, ys, vs, xt, yt):
    """Gradient at arbitrary targets (extension; the reference only has the MRAG layout)."""
    xs, ys, vs, xt, yt = map(_f64, (xs, ys, vs, xt, yt))
    xr = np.empty(len(xt))
    yr = np.empty(len(xt))
    lib().treecode_force(float(theta), _dp(xs), _dp(ys), _dp(vs), len(xs), _dp(xt), _dp(yt), len(xt), _dp(xr),
                         _dp(yr))
    if len(xr) and np.isnan(xr[0]) and last_error():
        raise TreecodeError(last_error())
    return xr, yr


def treecode_velocity_mrag(theta, xs, ys, vorticity, x0s, y0s, hs):
    """(u, v) = (grad_y, -grad_x) / (2 pi) of treecode_force_mrag (main.cpp:136-143)."""
    xs, ys, vs, x0s, y0s, hs = map(_f64, (xs, ys, vorticity, x0s, y0s, hs))
    nb = len(x0s)
    u = np.empty(nb * 1024)
    v = np.empty(nb * 1024)
    lib().treecode_velocity_mrag(float(theta), _dp(xs), _dp(ys), _dp(vs), len(xs), _dp(x0s), _dp(y0s), _dp(hs), nb,
                                 _dp(u), _dp(v))
    if nb and np.isnan(u[0]) and last_error():
        raise TreecodeError(last_error())
    return u, v


def _ptr(a):
    """Pointer of a numpy array (host) or a DeviceArray (device)."""
    if isinstance(a, np.ndarray):
        return a.ctypes.data
    return a.ptr


class DeviceArray:
    """A 1-D buffer in the HBM of a context's GPU (tc_device_alloc)."""

    def __init__(self, ctx, n, dtype):
        self.ctx, self.n, self.dtype = ctx, int(n), np.dtype(dtype)
        self.nbytes = self.n * self.dtype.itemsize
        self.ptr = lib().tc_device_alloc(ctx.h, self.nbytes)
        if not self.ptr:
            raise TreecodeError(last_error())

    def __len__(self):
        return self.n

    def upload(self, a):
        a = np.ascontiguousarray(a, dtype=self.dtype)
        assert a.size == self.n
        self.ctx._check(lib().tc_copy_to_device(self.ctx.h, self.ptr, a.ctypes.data, self.nbytes))
        return self

    def numpy(self):
        out = np.empty(self.n, dtype=self.dtype)
        self.ctx._check(lib().tc_copy_to_host(self.ctx.h, out.ctypes.data, self.ptr, self.nbytes))
        return out

    def free(self):
        if self.ptr and self.ctx.h:
            lib().tc_device_free(self.ctx.h, self.ptr)
        self.ptr = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


class Context:
    """One HIP stream + resident tree + grow-only HBM buffers (tc_context)."""

    def __init__(self, device=0, precision=FP64, pot_order=12, pot_leaf=512, force_order=24, force_leaf=192,
                 profile=False, counters=False):
        L = lib()
        self.h = L.tc_context_create(device)
        if not self.h:
            raise TreecodeError(last_error())
        self.set(OPT_PRECISION, precision)
        self.set(OPT_POT_ORDER, pot_order)
        self.set(OPT_POT_LEAF, pot_leaf)
        self.set(OPT_FORCE_ORDER, force_order)
        self.set(OPT_FORCE_LEAF, force_leaf)
        self.set(OPT_PROFILE, int(profile))
        self.set(OPT_COUNTERS, int(counters))

    def set(self, key, value):
        if lib().tc_set_option(self.h, key, int(value)) != 0:
            raise TreecodeError(last_error())

    def _check(self, rc):
        if rc != 0:
            raise TreecodeError(last_error())

    @property
    def stream(self):
        return lib().tc_context_stream(self.h)

    def potential(self, theta, xs, ys, vs, xt, yt, out=None, device=False):
        if not device:
            xs, ys, vs, xt, yt = map(_f64, (xs, ys, vs, xt, yt))
            out = np.empty(len(xt)) if out is None else out
        self._check(lib().tc_potential(self.h, float(theta), _ptr(xs), _ptr(ys), _ptr(vs), len(xs), _ptr(xt),
                                       _ptr(yt), len(xt), _ptr(out), int(device)))
        return out

    def potential_noalloc(self, theta, xs, ys, vs, xt, yt, out, n=None, m=None):
        """tc_potential_noalloc: DeviceArrays only; the first n sources and m targets
        (default: all); synchronous, allocates nothing on the device."""
        n = len(xs) if n is None else n
        m = len(xt) if m is None else m
        self._check(lib().tc_potential_noalloc(self.h, float(theta), xs.ptr, ys.ptr, vs.ptr, int(n), xt.ptr, yt.ptr,
                                               int(m), out.ptr))
        return out

    def potential_resident(self, theta, xt, yt, out=None, device=False):
        if not device:
            xt, yt = _f64(xt), _f64(yt)
            out = np.empty(len(xt)) if out is None else out
        self._check(lib().tc_potential_resident(self.h, float(theta), _ptr(xt), _ptr(yt), len(xt), _ptr(out),
                                                int(device)))
        return out

    def force(self, theta, xs, ys, vs, x0s, y0s, hs, xr=None, yr=None, device=False):
        if not device:
            xs, ys, vs, x0s, y0s, hs = map(_f64, (xs, ys, vs, x0s, y0s, hs))
            xr = np.empty(len(x0s) * 1024) if xr is None else xr
            yr = np.empty(len(x0s) * 1024) if yr is None else yr
        self._check(lib().tc_force_mrag(self.h, float(theta), _ptr(xs), _ptr(ys), _ptr(vs), len(xs), _ptr(x0s),
                                        _ptr(y0s), _ptr(hs), len(x0s), _ptr(xr), _ptr(yr), int(device)))
        return xr, yr

    def force_resident(self, theta, x0s, y0s, hs, xr=None, yr=None, device=False):
        if not device:
            x0s, y0s, hs = map(_f64, (x0s, y0s, hs))
            xr = np.empty(len(x0s) * 1024) if xr is None else xr
            yr = np.empty(len(x0s) * 1024) if yr is None else yr
        self._check(lib().tc_force_mrag_resident(self.h, float(theta), _ptr(x0s), _ptr(y0s), _ptr(hs), len(x0s),
                                                 _ptr(xr), _ptr(yr), int(device)))
        return xr, yr

    def force_targets(self, theta, xs, ys, vs, xt, yt, xr=None, yr=None, device=False):
        if not device:
            xs, ys, vs, xt, yt = map(_f64, (xs, ys, vs, xt, yt))
            xr = np.empty(len(xt)) if xr is None else xr
            yr = np.empty(len(xt)) if yr is None else yr
        self._check(lib().tc_force_targets(self.h, float(theta), _ptr(xs), _ptr(ys), _ptr(vs), len(xs), _ptr(xt),
                                           _ptr(yt), len(xt), _ptr(xr), _ptr(yr), int(device)))
        return xr, yr

    def synchronize(self):
        self._check(lib().tc_synchronize(self.h))

    def stats(self):
        t = np.zeros(4)
        c = (ctypes.c_ulonglong * 8)()
        info = np.zeros(4, np.int32)
        self._check(lib().tc_get_stats(self.h, _dp(t), c, info.ctypes.data_as(_I)))
        return {"times_ms": {"build": t[0], "tsort": t[1], "eval": t[2], "total": t[3]},
                "counters": {"p2p": c[0], "e2p": c[1], "e2l": c[2], "l2p": c[3], "p2p_slots": c[4], "e2p_slots": c[5],
                             "m2l_wave": c[6], "e2p_wave": c[7]},
                "n": int(info[0]), "nnodes": int(info[1]), "ninternal": int(info[2]), "order": int(info[3])}

    def profile(self):
        """Build phases (ms) and the traversal's clock cycles per category (tc_get_profile)."""
        b = np.zeros(4)
        c = (ctypes.c_ulonglong * 5)()
        self._check(lib().tc_get_profile(self.h, _dp(b), c))
        cyc = dict(zip(("p2p", "e2p", "local", "l2p", "total"), (int(v) for v in c)))
        walk = cyc["total"] - cyc["p2p"] - cyc["e2p"] - cyc["local"] - cyc["l2p"]
        shares = {k: (v / cyc["total"] if cyc["total"] else 0.0) for k, v in cyc.items() if k != "total"}
        shares["walk"] = walk / cyc["total"] if cyc["total"] else 0.0
        return {"build_ms": dict(zip(("keys", "sort", "topology", "upward"), (float(v) for v in b))),
                "cycles": cyc, "cycle_shares": shares}

    def tree(self):
        st = self.stats()
        n, nn, p = st["n"], st["nnodes"], st["order"]
        keys = np.zeros(n, np.int32)
        perm = np.zeros(n, np.int32)
        box = np.zeros(3)
        geom = np.zeros(4 * nn)
        topo = np.zeros(4 * nn, np.int32)
        exps = np.zeros(2 * p * nn)
        self._check(lib().tc_tree_export(self.h, keys.ctypes.data_as(_I), perm.ctypes.data_as(_I), _dp(box),
                                         _dp(geom), topo.ctypes.data_as(_I), _dp(exps)))
        return {"keys": keys, "perm": perm, "box": box, "geom": geom.reshape(nn, 4), "topo": topo.reshape(nn, 4),
                "exps": exps.reshape(nn, 2, p)}

    def empty(self, n, dtype=np.float64):
        return DeviceArray(self, n, dtype)

    def to_device(self, a, dtype=np.float64):
        a = np.ascontiguousarray(a, dtype=dtype)
        return DeviceArray(self, a.size, dtype).upload(a)

    def direct_potential(self, xs, ys, vs, xt, yt, idx):
        """FP64 direct sum at targets idx on the GPU (arguments: DeviceArrays)."""
        out = DeviceArray(self, len(idx), np.float64)
        self._check(lib().tc_direct_potential(self.h, xs.ptr, ys.ptr, vs.ptr, len(xs), xt.ptr, yt.ptr, idx.ptr,
                                              len(idx), out.ptr))
        return out.numpy()

    def direct_force(self, xs, ys, vs, xt, yt, idx):
        ox = DeviceArray(self, len(idx), np.float64)
        oy = DeviceArray(self, len(idx), np.float64)
        self._check(lib().tc_direct_force(self.h, xs.ptr, ys.ptr, vs.ptr, len(xs), xt.ptr, yt.ptr, idx.ptr,
                                          len(idx), ox.ptr, oy.ptr))
        return ox.numpy(), oy.numpy()

    def copy_kernel(self, dst, src):
        self._check(lib().tc_copy_kernel(self.h, dst.ptr, src.ptr, min(dst.nbytes, src.nbytes)))

    def morton_sort_device(self, x, y, box, keys, perm):
        self._check(lib().tc_morton_sort_device(self.h, _ptr(x), _ptr(y), len(x), _ptr(box), _ptr(keys),
                                                _ptr(perm)))

    def close(self):
        if getattr(self, "h", None):
            lib().tc_context_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def potential_batch_resident(requests):
    """tc_potential_batch_resident: requests = [(ctx, theta, xt, yt, out), ...] with DeviceArrays,
    each context holding a resident potential tree (same order and precision for all); one
    launch of the cross-request traversal writes every out.  Synchronous."""
    k = len(requests)
    if k == 0:
        raise TreecodeError("potential_batch_resident: no request")
    P = ctypes.c_void_p
    ctxs = (P * k)(*[r[0].h for r in requests])
    th = (ctypes.c_double * k)(*[float(r[1]) for r in requests])
    xt = (P * k)(*[r[2].ptr for r in requests])
    yt = (P * k)(*[r[3].ptr for r in requests])
    m = (ctypes.c_int * k)(*[len(r[2]) for r in requests])
    out = (P * k)(*[r[4].ptr for r in requests])
    if lib().tc_potential_batch_resident(ctxs, k, th, xt, yt, m, out) != 0:
        raise TreecodeError(last_error())
    return [r[4] for r in requests]


class Server:
    """Continuous-batching request server (tc_server): slots of (streams + host thread), one
    source-keyed tree cache that every slot reads, HBM admission against a budget; with
    batch=True the potential traversals of concurrent requests run as one kernel launch."""

    STATS = ["completed", "reused", "failed", "max_inflight", "peak_bytes", "budget_bytes", "queued", "batches",
             "batched", "evicted", "trimmed", "cached_trees", "held_bytes", "estimate_misses", "estimate_excess", "merged"]

    def __init__(self, device=0, slots=4, hbm_budget_gib=64.0, batch=False):
        """batch: False, True, or "always" (a lone traversal also goes through the batch kernel)."""
        flags = SERVER_BATCH_ALWAYS if batch == "always" else SERVER_BATCH if batch else 0
        self.h = lib().tc_server_create_ex(device, slots, float(hbm_budget_gib), flags)
        if not self.h:
            raise TreecodeError("tc_server_create failed: " + last_error())
        self._keep = {}

    def submit_potential(self, theta, xs, ys, vs, xt, yt, precision=FP64, source_id=0, order=0, leaf=0):
        xs, ys, vs, xt, yt = map(_f64, (xs, ys, vs, xt, yt))
        out = np.empty(len(xt))
        r = TcRequest(kind=0, precision=precision, theta=theta, xs=xs.ctypes.data, ys=ys.ctypes.data,
                      vs=vs.ctypes.data, n=len(xs), xt=xt.ctypes.data, yt=yt.ctypes.data, m=len(xt),
                      out0=out.ctypes.data, source_id=source_id, order=order, leaf=leaf)
        return self._submit(r, (xs, ys, vs, xt, yt), (out,))

    def submit_force(self, theta, xs, ys, vs, x0s, y0s, hs, precision=FP64, source_id=0, order=0, leaf=0):
        xs, ys, vs, x0s, y0s, hs = map(_f64, (xs, ys, vs, x0s, y0s, hs))
        xr = np.empty(len(x0s) * 1024)
        yr = np.empty(len(x0s) * 1024)
        r = TcRequest(kind=1, precision=precision, theta=theta, xs=xs.ctypes.data, ys=ys.ctypes.data,
                      vs=vs.ctypes.data, n=len(xs), x0s=x0s.ctypes.data, y0s=y0s.ctypes.data, hs=hs.ctypes.data,
                      nblocks=len(x0s), out0=xr.ctypes.data, out1=yr.ctypes.data, source_id=source_id,
                      order=order, leaf=leaf)
        return self._submit(r, (xs, ys, vs, x0s, y0s, hs), (xr, yr))

    def _submit(self, req, inputs, outputs):
        t = lib().tc_server_submit(self.h, ctypes.byref(req))
        if t < 0:
            raise TreecodeError("request does not fit the server's HBM budget")
        self._keep[t] = (inputs, outputs)
        return t

    def wait(self, ticket):
        """-> (outputs, {queue_ms, service_ms, latency_ms, reused, slot})"""
        times = np.zeros(5)
        st = lib().tc_server_wait(self.h, ticket, _dp(times))
        inputs, outputs = self._keep.pop(ticket)
        if st != 0:
            raise TreecodeError("request %d failed (status %d): %s" % (ticket, st, last_error()))
        info = {"queue_ms": times[0], "service_ms": times[1], "latency_ms": times[2], "reused": bool(times[3]),
                "slot": int(times[4])}
        return (outputs[0] if len(outputs) == 1 else outputs), info

    def stats(self):
        out = (ctypes.c_longlong * len(self.STATS))()
        n = lib().tc_server_stats_ex(self.h, out, len(self.STATS))
        return dict(zip(self.STATS[:n], list(out)[:n]))

    def close(self):
        if getattr(self, "h", None):
            lib().tc_server_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
