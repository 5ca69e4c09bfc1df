"""Build libtreecodes2d.so and libtreecodes2d_float.so for gfx950 with hipcc
(in-tree, no JIT cache).

libtreecodes2d.so        the FP64 drop-in pair (treecode_potential / treecode_force_mrag
                         over double arrays) and the extension API; also reachable
                         under the reference's library names lib2d-treecodes-{potential,
                         force}.so (symlinks beside it)
libtreecodes2d_float.so  the reference's real=float build: the same two names over
                         float arrays (tc_api.cpp compiled with TREECODES2D_FLOAT_DROPIN,
                         the kernels shared); it exports only the drop-in pair and
                         treecode_last_error, and float/lib2d-treecodes-{potential,
                         force}.so link to it

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
LIB_FLOAT = os.path.join(HERE, "libtreecodes2d_float.so")
ALIASES = ["lib2d-treecodes-potential.so", "lib2d-treecodes-force.so"]
# (source, object name, extra flags)
UNITS = [("tc_build.hip", "tc_build", []), ("tc_eval.hip", "tc_eval", []), ("tc_api.cpp", "tc_api", []),
         ("tc_serve.cpp", "tc_serve", []), ("tc_noalloc.hip", "tc_noalloc", []),
         ("tc_api.cpp", "tc_api_float", ["-DTREECODES2D_FLOAT_DROPIN"])]
DOUBLE_OBJS = ["tc_build", "tc_eval", "tc_api", "tc_serve", "tc_noalloc"]
FLOAT_OBJS = ["tc_build", "tc_eval", "tc_api_float"]
FLOAT_EXPORTS = ["treecode_potential", "treecode_force_mrag", "treecode_last_error"]
ARCH = os.environ.get("TC_OFFLOAD_ARCH", "gfx950")

CFLAGS = ["--offload-arch=" + ARCH, "-O3", "-std=c++17", "-fPIC", "-ffp-contract=off",
          "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops", "-fvisibility=hidden",
          "-Wall", "-Wno-unused-function", "-Wno-unused-variable"]


def _hipcc():
    for c in ("/opt/rocm/bin/hipcc", "hipcc"):
        if os.path.sep not in c or os.path.exists(c):
            return c
    return "hipcc"


def _obj(name):
    return os.path.join(OBJ, name + ".o")


def _compile(unit):
    src, name, extra = unit
    cmd = [_hipcc()] + CFLAGS + extra + ["-c", os.path.join(CSRC, src), "-o", _obj(name)]
    r = subprocess.run(cmd, capture_output=True, text=True)
    msgs = "\n".join(l for l in (r.stdout + r.stderr).splitlines() if "packed-fp32-ops" not in l)
    if r.returncode != 0:
        raise RuntimeError("hipcc failed for %s:\n%s" % (src, msgs))
    return _obj(name)


def _link(out, objs, extra=()):
    tmp = out + ".tmp"
    cmd = [_hipcc(), "--offload-arch=" + ARCH, "-shared", "-fPIC", "-o", tmp] + [_obj(o) for o in objs] + list(extra)
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("link failed:\n" + r.stdout + r.stderr)
    os.replace(tmp, out)


def _alias(directory, target):
    os.makedirs(directory, exist_ok=True)
    for a in ALIASES:
        p = os.path.join(directory, a)
        if os.path.lexists(p):
            os.remove(p)
        os.symlink(os.path.relpath(target, directory), p)


def build(verbose=False):
    os.makedirs(OBJ, exist_ok=True)
    with concurrent.futures.ThreadPoolExecutor(len(UNITS)) as ex:
        list(ex.map(_compile, UNITS))
    _link(LIB, DOUBLE_OBJS)
    _alias(HERE, LIB)
    vscript = os.path.join(OBJ, "float_exports.map")
    with open(vscript, "w") as f:
        f.write("{\n  global: %s;\n  local: *;\n};\n" % "; ".join(FLOAT_EXPORTS))
    _link(LIB_FLOAT, FLOAT_OBJS, ["-Wl,--version-script=" + vscript])
    _alias(os.path.join(HERE, "float"), LIB_FLOAT)
    if verbose:
        print("built", LIB, "and", LIB_FLOAT)
    return LIB


if __name__ == "__main__":
    build(verbose=True)
    sys.exit(0)
