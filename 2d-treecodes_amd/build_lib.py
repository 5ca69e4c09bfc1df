"""Build libtreecodes2d.so for gfx950 with hipcc (in-tree, no JIT cache).

Flags:
  -ffp-contract=off   keys, centres of mass, radii and every accept/open test
                      are rounded exactly as written (hot loops opt back in to
                      contraction with `#pragma clang fp contract(fast)`)
  -Xclang -target-feature -Xclang -packed-fp32-ops
                      no packed FP32 instructions are generated
"""
import concurrent.futures
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
OBJ = os.path.join(HERE, "_build")
LIB = os.path.join(HERE, "libtreecodes2d.so")
ALIASES = ["lib2d-treecodes-potential.so", "lib2d-treecodes-force.so"]
SOURCES = ["tc_build.hip", "tc_eval.hip", "tc_api.cpp", "tc_serve.cpp"]
ARCH = os.environ.get("TC_OFFLOAD_ARCH", "gfx950")

CFLAGS = ["--offload-arch=" + ARCH, "-O3", "-std=c++17", "-fPIC", "-ffp-contract=off",
          "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops", "-fvisibility=hidden",
          "-Wall", "-Wno-unused-function", "-Wno-unused-variable"]


def _hipcc():
    for c in ("/opt/rocm/bin/hipcc", "hipcc"):
        if os.path.sep not in c or os.path.exists(c):
            return c
    return "hipcc"


def _compile(src):
    out = os.path.join(OBJ, src + ".o")
    cmd = [_hipcc()] + CFLAGS + ["-c", os.path.join(CSRC, src), "-o", out]
    r = subprocess.run(cmd, capture_output=True, text=True)
    msgs = "\n".join(l for l in (r.stdout + r.stderr).splitlines() if "packed-fp32-ops" not in l)
    if r.returncode != 0:
        raise RuntimeError("hipcc failed for %s:\n%s" % (src, msgs))
    return out


def build(verbose=False):
    os.makedirs(OBJ, exist_ok=True)
    with concurrent.futures.ThreadPoolExecutor(len(SOURCES)) as ex:
        objs = list(ex.map(_compile, SOURCES))
    tmp = LIB + ".tmp"
    cmd = [_hipcc(), "--offload-arch=" + ARCH, "-shared", "-fPIC", "-o", tmp] + objs
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("link failed:\n" + r.stdout + r.stderr)
    os.replace(tmp, LIB)
    for a in ALIASES:
        p = os.path.join(HERE, a)
        if os.path.lexists(p):
            os.remove(p)
        os.symlink(os.path.basename(LIB), p)
    if verbose:
        print("built", LIB)
    return LIB


if __name__ == "__main__":
    build(verbose=True)
    sys.exit(0)
