"""The autotuner's cache: bucket keys and the nearest-bucket lookup a server
uses for request sizes that were not tuned (host logic, no GPU)."""
import json


def _tuner(pkg, tmp_path, entries):
    from treecodes2d_amd import autotune
    path = tmp_path / "tuned.json"
    path.write_text(json.dumps(entries))
    return autotune.Tuner(pkg, cache_path=str(path)), autotune


def test_bucket_key_rounds_log2(pkg, tmp_path):
    _, at = _tuner(pkg, tmp_path, {})
    assert at.bucket_key("potential", 1, 0.5, 1 << 20, 1 << 20) == "potential/p1/t0.50/n20/m20"
    assert at.bucket_key("force", 0, 0.3, 3000, 70000) == "force/p0/t0.30/n12/m16"


def test_nearest_bucket(pkg, tmp_path):
    e = lambda o, l: {"order": o, "leaf": l, "ms": 1.0, "err": 1e-7, "err_target": 2.5e-7}
    tuner, _ = _tuner(pkg, tmp_path, {
        "potential/p1/t0.50/n11/m11": e(8, 64), "potential/p1/t0.50/n17/m17": e(10, 128),
        "potential/p1/t0.70/n17/m17": e(16, 128), "force/p1/t0.50/n17/m18": e(12, 192),
        "potential/p0/t0.50/n17/m17": e(12, 64)})
    # exact bucket, then nearest of the same kind / precision / theta
    assert tuner.lookup("potential", 1, 0.5, 1 << 17, 1 << 17) == (10, 128)
    assert tuner.nearest("potential", 1, 0.5, 1 << 16, 1 << 16) == (10, 128)
    assert tuner.nearest("potential", 1, 0.5, 1 << 12, 1 << 12) == (8, 64)
    assert tuner.nearest("potential", 1, 0.7, 1 << 17, 1 << 17) == (16, 128)
    assert tuner.nearest("potential", 0, 0.5, 1 << 17, 1 << 17) == (12, 64)
    assert tuner.nearest("force", 1, 0.5, 1 << 16, 1 << 18) == (12, 192)
    # too far from every bucket, or no bucket for this theta: None (reference defaults)
    assert tuner.nearest("potential", 1, 0.5, 1 << 14, 1 << 14) is None  # 3 + 3 from both buckets
    assert tuner.nearest("potential", 1, 0.5, 1 << 24, 1 << 24) is None
    assert tuner.nearest("potential", 1, 0.3, 1 << 17, 1 << 17) is None
    assert tuner.nearest("force", 1, 0.5, 1 << 21, 1 << 14) is None


# ------------------------------------------------------------------ selection rule on the GPU
import numpy as np  # noqa: E402
import pytest  # noqa: E402


@pytest.mark.gpu
def test_tuner_selection_rule(pkg, wl, tmp_path):
    # the chosen cell meets the error target and is the fastest cell that does (so it is
    # no slower than the reference's own order / leaf when that cell qualifies); with a
    # target no cell can meet, the reference's cell is kept and flagged
    from treecodes2d_amd import autotune
    tuner = autotune.Tuner(pkg, cache_path=str(tmp_path / "t.json"))
    xs, ys, vs, xt, yt = wl.uniform_potential(1 << 14, seed=31)
    data = dict(xs=xs, ys=ys, vs=vs, xt=xt, yt=yt)
    res = tuner.tune("potential", data, theta=0.5, precision=1, orders=(8, 12), leaves=(128, 512),
                     err_target=1e-6, reps=2)
    ok = [t for t in res["table"] if t["err"] <= 1e-6]
    assert res["met_target"] and res["err"] <= 1e-6
    assert res["ms"] == min(t["ms"] for t in ok)
    ref = [t for t in res["table"] if (t["order"], t["leaf"]) == (12, 512)][0]
    if ref["err"] <= 1e-6:
        assert res["ms"] <= ref["ms"]
    cached = tuner.lookup("potential", 1, 0.5, 1 << 14, 1 << 14)
    assert cached == (res["order"], res["leaf"])
    res = tuner.tune("potential", data, theta=0.5, precision=1, orders=(8, 12), leaves=(128, 512),
                     err_target=1e-30, reps=1, save=False)
    assert not res["met_target"] and (res["order"], res["leaf"]) == (12, 512)


@pytest.mark.gpu
@pytest.mark.parametrize("kind", ["potential", "force"])
def test_tuner_tiling_sweep(pkg, wl, tmp_path, kind):
    # with tilings, the chosen (order, leaf) is timed under each tiling value; the
    # fastest is recorded unless the default is within the margin of it, it is
    # looked up, and the chosen cell's error is unchanged
    from treecodes2d_amd import autotune
    tuner = autotune.Tuner(pkg, cache_path=str(tmp_path / "t.json"))
    if kind == "potential":
        xs, ys, vs, xt, yt = wl.uniform_potential(1 << 15, seed=34)
        data, m = dict(xs=xs, ys=ys, vs=vs, xt=xt, yt=yt), 1 << 15
        grid = dict(orders=(10, 12), leaves=(128,))
    else:
        xs, ys, vs, x0s, y0s, hs = wl.gaussian4_force(1 << 15)
        data, m = dict(xs=xs, ys=ys, vs=vs, x0s=x0s[:64], y0s=y0s[:64], hs=hs[:64]), 64 * 1024
        grid = dict(orders=(8, 12), leaves=(32,))
    values = autotune.TILING[kind][1]
    res = tuner.tune(kind, data, theta=0.5, precision=1, err_target=1.0, reps=1, tilings=values, **grid)
    assert [t["tile"] for t in res["tile_table"]] == list(values)
    assert res["tile"] in values
    fastest = min(t["ms"] for t in res["tile_table"])
    dflt = [t for t in res["tile_table"] if t["tile"] == autotune.TILING_DEFAULT[kind]][0]
    if dflt["ms"] * (1.0 - autotune.TILE_MARGIN) < fastest:
        assert res["tile"] == dflt["tile"] and res["ms"] == dflt["ms"]
    else:
        assert res["ms"] == fastest
    chosen = [t for t in res["table"] if (t["order"], t["leaf"]) == (res["order"], res["leaf"])][0]
    for t in res["tile_table"]:
        assert abs(t["err"] - chosen["err"]) <= 1e-3 * chosen["err"] + 1e-12
    key, val = tuner.lookup_tile(kind, 1, 0.5, len(xs), m)
    assert key == getattr(pkg, autotune.TILING[kind][0]) and val == res["tile"]
