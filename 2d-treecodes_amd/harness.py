"""Dataset harness with the reference benchmark's file format and pass rule.

Restates main.cpp's `test` / `file2test` / `check` (main.cpp:31-58, 60-287,
307-329) on top of this library:

File layout (native-endian, main.cpp:62-135):
    int NSRC; double x[NSRC], y[NSRC], v[NSRC]; int NDST;
    non-MRAG file:  double xdst[NDST], ydst[NDST], xref[NDST] (+ yref[NDST] for velocity files)
    MRAG file:      NB := NDST; double x0s[NB], y0s[NB], hs[NB] (+ yref[NB*1024])
Velocity references are rotated and scaled before comparison (main.cpp:136-143):
    xref <- -2 pi yref,  yref <- 2 pi xref.
A non-MRAG velocity file whose NDST is a multiple of 1024 is read as blocks whose
origin/spacing come from the first points of each block (main.cpp:155-173).

Verification (main.cpp:201-263): the FP64 direct sum at every 433rd target from
index 7, regularised with r^2 + 10 DBL_EPSILON, computed on the GPU
(tc_direct_potential / tc_direct_force); pass iff for every sampled target
|err| < tol or |err| / max(1e-6, max(|ref|, |res|)) < tol; velocities use
100 * tol (main.cpp:327).

CLI (main.cpp:289-353): python harness.py [theta=1] [tol=1e-8] [profile] [files...]
With no files, the reference's eleven dataset paths are tried and skipped
with a warning when absent (main.cpp:309-312, 331-343).
"""
import math
import os
import struct
import sys
import time

import numpy as np

OFFSET, JUMP = 7, 433  # main.cpp:203-204
BS = 32
BS2 = BS * BS

REFERENCE_FILES = [
    ("testDiego/diegoBinaryN400", False, "P"), ("testDiego/diegoBinaryN2000", False, "P"),
    ("testDiego/diegoBinaryN12000", False, "P"),
    ("diegoVel/velocityPoissonFishLmax6", False, "V"), ("diegoVel/velocityPoissonCylUnif2048", False, "V"),
    ("diegoVel/velocityPoissonFishLmax8Early", False, "V"), ("diegoVel/velocityPoissonFishLmax8Late", False, "V"),
    ("testSid/diegoSolverCylUniform", True, "V"), ("testSid/diegoSolverAdaptiveGrid", True, "V"),
    ("testSid/diegoVelTestsDec10", True, "V"), ("testSid/diegoVelTestsDec14", True, "V"),
]


class _Reader:
    def __init__(self, f):
        self.f = f

    def int(self):
        b = self.f.read(4)
        if len(b) != 4:
            raise ValueError("truncated dataset (int)")
        return struct.unpack("=i", b)[0]

    def doubles(self, n):
        a = np.fromfile(self.f, dtype=np.float64, count=n)
        if len(a) != n:
            raise ValueError("truncated dataset (%d of %d doubles)" % (len(a), n))
        return a


def read_dataset(path, potential, mragfile):
    """Parse one test case as main.cpp:60-173 does."""
    with open(path, "rb") as f:
        r = _Reader(f)
        nsrc = r.int()
        xs, ys, vs = r.doubles(nsrc), r.doubles(nsrc), r.doubles(nsrc)
        ndst = r.int()
        case = {"xs": xs, "ys": ys, "vs": vs}
        if mragfile:
            nb = ndst
            ndst = nb * BS2
            x0s, y0s, hs = r.doubles(nb), r.doubles(nb), r.doubles(nb)
            case.update(x0s=x0s, y0s=y0s, hs=hs)
            xref = np.zeros(ndst)  # never read for MRAG files (main.cpp:110-123)
        else:
            xdst, ydst, xref = r.doubles(ndst), r.doubles(ndst), r.doubles(ndst)
            case.update(xdst=xdst, ydst=ydst)
        yref = None
        if not potential:
            yref = r.doubles(ndst)
            xref, yref = -yref * (2 * math.pi), xref * (2 * math.pi)
        case.update(xref=xref, yref=yref, ndst=ndst)
    if not mragfile and not potential and ndst % BS2 == 0:
        nb = ndst // BS2
        xd, yd = case["xdst"], case["ydst"]
        case.update(x0s=xd[::BS2].copy(), y0s=yd[::BS2].copy(), hs=(xd[1::BS2] - xd[::BS2]).copy())
    return case


def write_dataset(path, xs, ys, vs, xdst=None, ydst=None, xref=None, yref=None, x0s=None, y0s=None, hs=None):
    """Write a case in the same layout (for tests and for converting other data)."""
    with open(path, "wb") as f:
        f.write(struct.pack("=i", len(xs)))
        for a in (xs, ys, vs):
            np.ascontiguousarray(a, np.float64).tofile(f)
        if x0s is not None:
            f.write(struct.pack("=i", len(x0s)))
            for a in (x0s, y0s, hs):
                np.ascontiguousarray(a, np.float64).tofile(f)
            if yref is not None:
                np.ascontiguousarray(yref, np.float64).tofile(f)
        else:
            f.write(struct.pack("=i", len(xdst)))
            for a in (xdst, ydst, xref if xref is not None else np.zeros(len(xdst))):
                np.ascontiguousarray(a, np.float64).tofile(f)
            if yref is not None:
                np.ascontiguousarray(yref, np.float64).tofile(f)


def write_synthetic_datasets(root, seed=1457):
    """Synthetic cases under the eleven dataset paths the reference's main.cpp reads
    (main.cpp:331-343; the originals are remote and not available offline), so that
    an unmodified main.cpp linked against this library runs every test it has:
    three potential files (N = 400, 2000, 12000 uniform sources, targets a copy of
    them), four velocity files whose targets form 32x32 blocks (main.cpp:155-173)
    and four MRAG files, from the 4-Gaussian vortex generator.  Stored reference
    values are zero: main.cpp replaces the sampled ones with its own direct sum."""
    here = os.path.dirname(os.path.abspath(__file__))
    sys.path.insert(0, here)
    import workloads
    written = []
    for k, (rel, mragfile, tt) in enumerate(REFERENCE_FILES):
        path = os.path.join(root, rel)
        os.makedirs(os.path.dirname(path), exist_ok=True)
        if tt == "P":
            n = int(rel.rsplit("N", 1)[1])
            xs, ys, vs, xt, yt = workloads.uniform_potential(n, seed=seed + k)
            write_dataset(path, xs, ys, vs, xdst=xt, ydst=yt)
        else:
            xs, ys, vs, x0s, y0s, hs = workloads.gaussian4_force(6000 + 3000 * k, seed=seed + k, nbx=3 + k % 3,
                                                                 h=1.0 / (64 + 32 * (k % 4)))
            if mragfile:
                write_dataset(path, xs, ys, vs, x0s=x0s, y0s=y0s, hs=hs, yref=np.zeros(len(x0s) * BS2))
            else:
                tx, ty = mrag_points(x0s, y0s, hs)
                write_dataset(path, xs, ys, vs, xdst=tx, ydst=ty, yref=np.zeros(len(tx)))
        written.append(path)
    return written


def check(ref, res, tol):
    """main.cpp:31-58: -> (passed, linf_abs, linf_rel, l1_abs, l1_rel, failures)."""
    ref = np.asarray(ref, np.float64)
    res = np.asarray(res, np.float64)
    if np.any(np.isnan(ref)) or np.any(np.isnan(res)):
        return False, math.nan, math.nan, math.nan, math.nan, [-1]
    err = ref - res
    maxval = np.maximum(np.abs(res), np.abs(ref))
    rel = err / np.maximum(1e-6, maxval)
    bad = (np.abs(rel) >= tol) & (np.abs(err) >= tol)
    return (not bool(np.any(bad)), float(np.max(np.abs(err), initial=0)), float(np.max(np.abs(rel), initial=0)),
            float(np.sum(np.abs(err))), float(np.sum(np.abs(rel))), list(np.nonzero(bad)[0]))


def mrag_points(x0s, y0s, hs):
    ix = np.arange(BS2) % BS
    iy = np.arange(BS2) // BS
    tx = (x0s[:, None] + ix[None, :] * hs[:, None]).ravel()
    ty = (y0s[:, None] + iy[None, :] * hs[:, None]).ravel()
    return tx, ty


def run_case(pkg, ctx, case, potential, theta, tol, verify=True, out=print):
    """One `test` call (main.cpp:60-287).  Returns the check results (or None without verify)."""
    xs, ys, vs = case["xs"], case["ys"], case["vs"]
    mrag = "x0s" in case
    out("Testing %s with %d sources and %d targets (theta %.3e)..." % (
        "POTENTIAL" if potential else "FORCE", len(xs), case["ndst"], theta))
    t0 = time.perf_counter()
    if potential:
        xd, yd = case["xdst"], case["ydst"]
        res = pkg.treecode_potential(theta, xs, ys, vs, xd, yd)
    else:
        if not mrag:
            raise ValueError("This file format is not supported for velocities (main.cpp:191-195).")
        resx, resy = pkg.treecode_force_mrag(theta, xs, ys, vs, case["x0s"], case["y0s"], case["hs"])
        xd, yd = mrag_points(case["x0s"], case["y0s"], case["hs"])
    out("solved in %.2f ms" % ((time.perf_counter() - t0) * 1e3))
    if not verify:
        return None
    idx = np.arange(OFFSET, case["ndst"], JUMP, dtype=np.int32)
    d = {k: ctx.to_device(a) for k, a in dict(xs=xs, ys=ys, vs=vs, xd=xd, yd=yd).items()}
    didx = ctx.to_device(idx, np.int32)
    results = []
    if potential:
        ref = ctx.direct_potential(d["xs"], d["ys"], d["vs"], d["xd"], d["yd"], didx)
        results.append(check(ref, res[idx], tol))
    else:
        rx, ry = ctx.direct_force(d["xs"], d["ys"], d["vs"], d["xd"], d["yd"], didx)
        results.append(check(rx, resx[idx], tol))
        results.append(check(ry, resy[idx], tol))
    for ok, linf, linf_rel, l1, l1_rel, _ in results:
        out("l-infinity errors: %.03e (absolute) %.03e (relative)" % (linf, linf_rel))
        out("       l-1 errors: %.03e (absolute) %.03e (relative)" % (l1, l1_rel))
    if all(r[0] for r in results):
        out("TEST PASSED.")
    else:
        out("TEST FAILED.")
    return results


def file2test(pkg, ctx, path, mragfile, testtype, theta, tol, verify=True, out=print):
    """main.cpp:307-329"""
    if not os.access(path, os.R_OK):
        out("WARNING: reference file <%s> not found, i skip this test." % path)
        return None
    out("reading from <%s> ..." % path)
    results = []
    if not mragfile and "P" in testtype:
        results += run_case(pkg, ctx, read_dataset(path, True, False), True, theta, tol, verify, out) or []
    if "V" in testtype:
        results += run_case(pkg, ctx, read_dataset(path, False, mragfile), False, theta, tol * 100, verify,
                            out) or []
    return results


def main(argv):
    here = os.path.dirname(os.path.abspath(__file__))
    sys.path.insert(0, os.path.dirname(here))
    import __graft_entry__
    pkg = __graft_entry__.load_package()
    args = [a for a in argv[1:]]
    theta = float(args[0]) if len(args) > 0 else 1.0
    tol = float(args[1]) if len(args) > 1 else 1e-8
    verify = not (len(args) > 2 and args[2] == "profile")
    files = [(p, p.endswith("_mrag"), "PV") for p in args[3:]] or REFERENCE_FILES
    ctx = pkg.Context()
    ok = True
    for path, mragfile, tt in files:
        res = file2test(pkg, ctx, path, mragfile, tt, theta, tol, verify)
        ok &= all(r[0] for r in (res or []))
    return 0 if ok else 1


if __name__ == "__main__":
    sys.exit(main(sys.argv))
