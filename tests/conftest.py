import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.join(ROOT, "oracle"))

import __graft_entry__  # noqa: E402

pkg_mod = __graft_entry__.load_package()


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs through the HIP library)")


@pytest.fixture(scope="session")
def pkg():
    return pkg_mod


@pytest.fixture(scope="session")
def oracle():
    import oracle as o
    o.lib()
    return o


@pytest.fixture(scope="session")
def wl():
    from treecodes2d_amd import workloads
    return workloads


@pytest.fixture(scope="session")
def ctx64(pkg):
    c = pkg.Context(counters=True, profile=True)
    yield c
    c.close()


@pytest.fixture(scope="session")
def ctx32(pkg):
    c = pkg.Context(precision=pkg.FP32, counters=True, profile=True)
    yield c
    c.close()


def rel_l2(a, b):
    a = np.asarray(a, dtype=np.float64)
    b = np.asarray(b, dtype=np.float64)
    d = np.linalg.norm(b)
    return float(np.linalg.norm(a - b) / (d if d > 0 else 1.0))


def mrag_targets(x0s, y0s, hs):
    """Target coordinates of the 32x32 blocks in output order ix + 32*(iy + 32*b)."""
    ix = np.arange(1024) % 32
    iy = np.arange(1024) // 32
    tx = (np.asarray(x0s)[:, None] + ix[None, :] * np.asarray(hs)[:, None]).ravel()
    ty = (np.asarray(y0s)[:, None] + iy[None, :] * np.asarray(hs)[:, None]).ravel()
    return tx, ty


def compare_tree(t, ot):
    """Exported GPU tree (Context.tree()) against an oracle.Tree: box, keys, permutation,
    topology and node geometry bit-exact (geometry decides every accept/open test),
    expansion coefficients within 1e-13 of each node's largest."""
    assert np.array_equal(t["box"], ot.bbox)
    assert np.array_equal(t["keys"], ot.keys)
    assert np.array_equal(t["perm"], ot.perm)
    nn = ot.nnodes
    assert t["topo"].shape[0] == nn
    assert np.array_equal(t["topo"][:, 0], ot.se[:, 0])
    assert np.array_equal(t["topo"][:, 1], ot.se[:, 1])
    assert np.array_equal(t["topo"][:, 2], ot.child)
    assert np.array_equal(t["topo"][:, 3], ot.level)
    assert np.array_equal(t["geom"], ot.geom)
    scale = np.maximum(np.abs(ot.exp).max(axis=(1, 2), keepdims=True), 1e-300)
    assert np.max(np.abs(t["exps"] - ot.exp) / scale) <= 1e-13
