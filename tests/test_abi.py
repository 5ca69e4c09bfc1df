"""CPU checks of the C ABI: the library loads, exports every declared symbol,
the headers are valid C, and with no GPU every entry point fails loudly
(there is no CPU fallback)."""
import ctypes
import os
import re
import subprocess

import numpy as np
import pytest

from conftest import ROOT

HEADERS = ["treecodes2d.h", "2d-treecodes-potential.h", "2d-treecodes-force.h"]
HAVE_GPU = os.path.exists("/dev/kfd")


def declared_symbols():
    names = set()
    for h in HEADERS:
        text = open(os.path.join(ROOT, "include", h)).read()
        text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
        for m in re.finditer(r"\b(tc_\w+|treecode_\w+)\s*\(", text):
            names.add(m.group(1))
    return sorted(names)


def test_headers_declare_the_dropin_pair():
    names = declared_symbols()
    assert "treecode_potential" in names and "treecode_force_mrag" in names


def test_library_exports_every_declared_symbol(pkg):
    L = ctypes.CDLL(pkg.LIB_PATH)
    missing = [s for s in declared_symbols() if not hasattr(L, s)]
    assert not missing, missing
    assert sorted(pkg.EXPORTED) == sorted(s for s in declared_symbols())


def test_compat_library_names(pkg):
    for alias in ["lib2d-treecodes-potential.so", "lib2d-treecodes-force.so"]:
        p = os.path.join(os.path.dirname(pkg.LIB_PATH), alias)
        L = ctypes.CDLL(p)
        assert hasattr(L, "treecode_potential") and hasattr(L, "treecode_force_mrag")


def test_only_abi_symbols_are_exported(pkg):
    out = subprocess.run(["nm", "-D", "--defined-only", pkg.LIB_PATH], capture_output=True, text=True).stdout
    text_syms = {l.split()[-1] for l in out.splitlines() if " T " in l}
    assert text_syms == set(declared_symbols())


@pytest.mark.parametrize("lang", ["c", "c++"])
def test_headers_compile(tmp_path, lang):
    src = tmp_path / ("t.c" if lang == "c" else "t.cpp")
    src.write_text("\n".join('#include "%s"' % h for h in HEADERS) + "\nint main(void){return 0;}\n")
    cc = "gcc" if lang == "c" else "g++"
    r = subprocess.run([cc, "-fsyntax-only", "-Wall", "-Werror", "-I", os.path.join(ROOT, "include"), str(src)],
                       capture_output=True, text=True)
    assert r.returncode == 0, r.stderr


@pytest.mark.skipif(HAVE_GPU, reason="checks the no-GPU failure path")
def test_no_gpu_fails_loudly(pkg):
    with pytest.raises(pkg.TreecodeError):
        pkg.Context()
    x = np.linspace(0, 1, 10)
    with pytest.raises(pkg.TreecodeError):
        pkg.treecode_potential(0.5, x, x, x, x, x)
    with pytest.raises(pkg.TreecodeError):
        pkg.treecode_force_mrag(0.5, x, x, x, x[:1], x[:1], x[:1])
    assert pkg.last_error()


def test_missing_library_is_an_error(pkg, monkeypatch):
    monkeypatch.setattr(pkg, "_lib", None)
    monkeypatch.setattr(pkg, "LIB_PATH", os.path.join(ROOT, "does-not-exist.so"))
    with pytest.raises(pkg.TreecodeError):
        pkg.lib()


def test_workloads_deterministic(wl):
    a = wl.uniform_potential(1000)
    b = wl.uniform_potential(1000)
    for u, v in zip(a, b):
        assert np.array_equal(u, v)
    xs, ys, vs, xt, yt = a
    assert xs.min() >= 0 and xs.max() < 1 and vs.min() >= -1 and vs.max() < 1
    assert xt is not xs and np.array_equal(xt, xs)
    xs, ys, vs, x0s, y0s, hs = wl.gaussian4_force(5000)
    assert len(x0s) == 4096 and np.all(hs == 1.0 / 2048)
    assert x0s.max() + 31 * hs[0] < 1.0
    xs, ys, vs, _, _ = wl.mixed_plummer_potential(10000)
    assert xs.min() >= 0 and xs.max() <= 1 and ys.min() >= 0 and ys.max() <= 1
    xs, ys, vs, x0s, y0s, hs = wl.gaussian16_force(10000, nblocks=64)
    assert len(x0s) == 64 and set(np.unique(hs)) <= {1 / 1024, 1 / 2048, 1 / 4096}
    reqs = wl.request_stream(200)
    assert len(reqs) == 200 and all(r["arrival"] > 0 for r in reqs)
    assert any(r["reuses_sources"] for r in reqs)


def test_request_stream_reuse_is_consistent(wl):
    reqs = wl.request_stream(500, seed=3)
    pools = wl.request_pools(seed=3, size=100000)
    by_id = {}
    for r in reqs:
        key = r["src_seed"]
        src = wl.pool_sources(pools, dict(r, n=min(r["n"], 100000)))
        if key in by_id:
            # same source id -> identical sources (the server may reuse the tree)
            assert all(np.array_equal(a, b) for a, b in zip(by_id[key], src))
        by_id[key] = src
    assert sum(r["reuses_sources"] for r in reqs) > 50


# ------------------------------------------------------------------ float build and reference clients
REF = "/root/reference"
HAVE_REF = os.path.exists(os.path.join(REF, "main.cpp"))
FLOAT_HEADERS = ["2d-treecodes-potential.h", "2d-treecodes-force.h"]


def _float_lib(pkg):
    return os.path.join(os.path.dirname(pkg.LIB_PATH), "libtreecodes2d_float.so")


def test_float_library_exports_reference_names(pkg):
    # the real=float build exports the two reference names (float signatures) and nothing
    # of the double library's extension API
    out = subprocess.run(["nm", "-D", "--defined-only", _float_lib(pkg)], capture_output=True, text=True).stdout
    text_syms = {l.split()[-1] for l in out.splitlines() if " T " in l}
    assert text_syms == {"treecode_potential", "treecode_force_mrag", "treecode_last_error"}
    for alias in FLOAT_HEADERS:
        p = os.path.join(os.path.dirname(pkg.LIB_PATH), "float", "lib" + alias[:-2] + ".so")
        assert os.path.realpath(p) == os.path.realpath(_float_lib(pkg))


def _decl_tokens(text):
    text = re.sub(r"/\*.*?\*/", " ", text, flags=re.S)
    text = re.sub(r"//[^\n]*", " ", text)
    return re.findall(r"[A-Za-z_]\w*|\S", text)


@pytest.mark.skipif(not HAVE_REF, reason="reference sources absent")
@pytest.mark.parametrize("real", ["double", "float"])
def test_headers_equal_reference_generation(real):
    # libraries.Makefile:30-31,35-36: m4 -D realtype=<real> drivers/treecode-*.h, minus the
    # typedef and attribute lines; the shipped headers must declare exactly the same tokens
    for h, src in zip(FLOAT_HEADERS, ["treecode-potential.h", "treecode-force.h"]):
        lines = open(os.path.join(REF, "drivers", src)).read().splitlines()
        gen = "\n".join(l for l in lines if "typedef" not in l and "attribute" not in l)
        gen = re.sub(r"\brealtype\b", real, gen)
        mine = os.path.join(ROOT, "include", h) if real == "double" else os.path.join(ROOT, "include", "float", h)
        assert _decl_tokens(open(mine).read()) == _decl_tokens(gen), (real, h)


@pytest.mark.parametrize("real", ["double", "float"])
def test_client_links_both_builds(pkg, tmp_path, real):
    # a C client of the reference headers links against the reference library names of
    # either build (the float client only resolves against the float library)
    inc = os.path.join(ROOT, "include") if real == "double" else os.path.join(ROOT, "include", "float")
    libdir = os.path.dirname(pkg.LIB_PATH) if real == "double" else os.path.join(os.path.dirname(pkg.LIB_PATH),
                                                                                   "float")
    src = tmp_path / "client.c"
    src.write_text('#include "2d-treecodes-potential.h"\n#include "2d-treecodes-force.h"\n'
                   "int main(void) { %s a[1] = {0}; treecode_potential(0.5f, a, a, a, 0, a, a, 0, a);\n"
                   "  treecode_force_mrag(0.5f, a, a, a, 0, a, a, a, 0, a, a); return 0; }\n" % real)
    exe = tmp_path / "client"
    r = subprocess.run(["gcc", "-Wall", "-Werror", "-I", inc, str(src), "-L", libdir, "-l2d-treecodes-potential",
                        "-l2d-treecodes-force", "-Wl,-rpath," + libdir, "-o", str(exe)], capture_output=True,
                       text=True)
    assert r.returncode == 0, r.stderr


@pytest.mark.skipif(not HAVE_REF, reason="reference sources absent")
def test_reference_main_links_unchanged(pkg, tmp_path):
    # the reference's main.cpp, read in place, compiled with its Makefile's flags against
    # include/ and linked against lib2d-treecodes-{potential,force}.so (Makefile:17-18,25-26);
    # with none of its dataset files present it skips every case without calling the library
    exe = tmp_path / "test"
    libdir = os.path.dirname(pkg.LIB_PATH)
    r = subprocess.run(["g++", "-std=c++11", "-fopenmp", "-DBLOCKSIZE=32", "-g", "-I", os.path.join(ROOT, "include"),
                        os.path.join(REF, "main.cpp"), "-L", libdir, "-l2d-treecodes-potential",
                        "-l2d-treecodes-force", "-Wl,-rpath," + libdir, "-o", str(exe)], capture_output=True,
                       text=True)
    assert r.returncode == 0, r.stderr
    run = subprocess.run([str(exe)], cwd=str(tmp_path), capture_output=True, text=True, timeout=60)
    assert run.returncode == 0
    assert run.stdout.count("not found, i skip this test") == 11
