"""The dataset harness (main.cpp's file format and check rule)."""
import math

import numpy as np
import pytest

from conftest import mrag_targets


@pytest.fixture(scope="module")
def harness(pkg):
    from treecodes2d_amd import harness as h
    return h


def test_roundtrip_potential_file(harness, wl, tmp_path):
    xs, ys, vs, xt, yt = wl.uniform_potential(500, seed=51)
    ref = np.arange(500.0)
    p = tmp_path / "pot"
    harness.write_dataset(str(p), xs, ys, vs, xdst=xt, ydst=yt, xref=ref)
    c = harness.read_dataset(str(p), potential=True, mragfile=False)
    assert np.array_equal(c["xs"], xs) and np.array_equal(c["xdst"], xt) and np.array_equal(c["xref"], ref)
    assert c["ndst"] == 500 and c["yref"] is None


def test_roundtrip_velocity_files(harness, wl, tmp_path):
    xs, ys, vs, x0s, y0s, hs = wl.gaussian4_force(300, nbx=2, h=1.0 / 64, seed=52)
    yref = np.linspace(0, 1, len(x0s) * 1024)
    p = tmp_path / "mrag"
    harness.write_dataset(str(p), xs, ys, vs, x0s=x0s, y0s=y0s, hs=hs, yref=yref)
    c = harness.read_dataset(str(p), potential=False, mragfile=True)
    assert np.array_equal(c["x0s"], x0s) and c["ndst"] == len(x0s) * 1024
    # rotation and 2 pi scaling (main.cpp:136-143); xref is zero for MRAG files
    assert np.array_equal(c["xref"], -yref * (2 * math.pi)) and np.all(c["yref"] == 0)
    # a non-MRAG velocity file whose targets form 32x32 blocks is read as blocks (main.cpp:155-173)
    tx, ty = mrag_targets(x0s, y0s, hs)
    p2 = tmp_path / "vel"
    harness.write_dataset(str(p2), xs, ys, vs, xdst=tx, ydst=ty, xref=np.ones(len(tx)), yref=np.zeros(len(tx)))
    c2 = harness.read_dataset(str(p2), potential=False, mragfile=False)
    assert np.allclose(c2["x0s"], x0s) and np.allclose(c2["hs"], hs)
    assert np.all(c2["yref"] == 2 * math.pi)


def test_truncated_file_is_an_error(harness, tmp_path):
    p = tmp_path / "bad"
    p.write_bytes(b"\x05\x00\x00\x00" + b"\x00" * 16)
    with pytest.raises(ValueError):
        harness.read_dataset(str(p), True, False)


def test_check_rule(harness):
    # pass iff |err| < tol or |err|/max(1e-6, max|.|) < tol (main.cpp:40-47)
    ok, *_ = harness.check([1.0, 2.0], [1.0 + 5e-9, 2.0], 1e-8)
    assert ok
    ok, *_ = harness.check([1e3], [1e3 + 5e-6], 1e-8)  # |err| > tol but relative 5e-9 < tol
    assert ok
    ok, *_, bad = harness.check([1.0, 2.0], [1.0, 2.0 + 1e-6], 1e-8)
    assert not ok and bad == [1]
    ok, *_ = harness.check([1e-9], [2e-9], 1e-8)  # tiny values: absolute test passes
    assert ok
    ok, *_ = harness.check([np.nan], [1.0], 1e-8)
    assert not ok


def test_missing_file_skips(harness, pkg):
    msgs = []
    assert harness.file2test(pkg, None, "/nonexistent/file", False, "P", 1.0, 1e-8, out=msgs.append) is None
    assert "not found" in msgs[0]


@pytest.mark.gpu
def test_harness_runs_files(harness, pkg, wl, tmp_path):
    ctx = pkg.Context()
    xs, ys, vs, xt, yt = wl.uniform_potential(20000, seed=53)
    p = tmp_path / "pot"
    harness.write_dataset(str(p), xs, ys, vs, xdst=xt, ydst=yt)
    res = harness.file2test(pkg, ctx, str(p), False, "P", 0.5, 1e-4, out=lambda *_: None)
    assert res and all(r[0] for r in res)
    xs, ys, vs, x0s, y0s, hs = wl.gaussian4_force(20000, nbx=4, h=1.0 / 128, seed=54)
    p2 = tmp_path / "mrag"
    harness.write_dataset(str(p2), xs, ys, vs, x0s=x0s, y0s=y0s, hs=hs, yref=np.zeros(len(x0s) * 1024))
    res = harness.file2test(pkg, ctx, str(p2), True, "V", 0.5, 1e-8, out=lambda *_: None)
    assert res and all(r[0] for r in res)
