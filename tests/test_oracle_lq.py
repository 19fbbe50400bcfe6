"""Pins the oracle's bin mapping and neighbour sets against the reference's own
src/lq.c: the committed fixture tests/golden/lq_golden.npz was written by
tests/golden/gen_lq_golden.py running lq.c compiled from the reference
sources.  When those sources are present, the same checks also run live."""
from pathlib import Path

import numpy as np
import pytest

HERE = Path(__file__).resolve().parent
FIX = HERE / "golden" / "lq_golden.npz"


@pytest.fixture(scope="module")
def fx():
    return dict(np.load(FIX))


def test_bin_index_matches_lq(orc, fx):
    g = orc.grid(fx["grid_origin"], fx["grid_size"], fx["grid_div"])
    pts = fx["points"]
    lq_bin = fx["bin"]
    ours = np.array([orc.cell_of(g, *p) for p in pts], np.int64)
    inside = lq_bin >= 0
    assert inside.sum() > 1000 and (~inside).sum() > 100
    # In the lattice: the same bin, lq.c's (int)(((x - ox) / sx) * dx) mapping
    # and its ix*divy*divz + iy*divz + iz linear index.
    np.testing.assert_array_equal(ours[inside], lq_bin[inside])
    # lq.c's "other" bin holds exactly the points outside the super-brick;
    # the restatement clamps them into the boundary cells instead.
    o, s = fx["grid_origin"], fx["grid_size"]
    out_box = np.any((pts < o) | (pts >= o + s), axis=1)
    np.testing.assert_array_equal(~inside, out_box)
    d = fx["grid_div"]
    clamped = np.clip(np.floor((pts[~inside] - o) / s * d), 0, d - 1).astype(np.int64)
    lin = clamped[:, 0] * d[1] * d[2] + clamped[:, 1] * d[2] + clamped[:, 2]
    np.testing.assert_array_equal(ours[~inside], lin)


def _d2(q, pts):
    dx = (q[0] - pts[:, 0]).astype(np.float32)
    dy = (q[1] - pts[:, 1]).astype(np.float32)
    dz = (q[2] - pts[:, 2]).astype(np.float32)
    return (dx * dx + dy * dy) + dz * dz


@pytest.mark.parametrize("binned", [False, True])
def test_knn_agrees_with_lq_radius_queries(orc, fx, binned):
    """For every lq locality query (all objects with d2 < r^2), the oracle's
    k nearest with k = |lq set| + 3 begin with exactly that set."""
    pts = fx["points"]
    g = orc.grid(fx["grid_origin"], fx["grid_size"], fx["grid_div"])
    q = fx["queries"]
    offs, mem = fx["nbr_offsets"], fx["nbr_members"]
    checked = 0
    for i in range(q.shape[0]):
        lq_set = set(mem[offs[i]:offs[i + 1]].tolist())
        k = min(len(lq_set) + 3, 32)
        if len(lq_set) + 1 > 32:
            continue
        fn = (lambda a, b: orc.knn_binned(g, a, b, k, False)) if binned else (
            lambda a, b: orc.knn_brute(a, b, k, False))
        idx, dist = fn(q[i:i + 1], pts)
        idx = idx[:, 0]
        d2 = _d2(q[i], pts[idx])
        r2 = np.float32(fx["radii"][i]) * np.float32(fx["radii"][i])
        ours = set(idx[d2 < r2].tolist())
        assert ours == lq_set, f"query {i}"
        checked += 1
    assert checked > 60


def test_live_lq_if_reference_present(orc):
    import sys
    sys.path.insert(0, str(HERE / "golden"))
    import lqref

    if not lqref.available():
        pytest.skip("reference lq.c not built here (oracle/_ref is built only where the reference sources exist)")
    g0 = np.random.default_rng(99)
    origin, size, div = (-20.0, -5.0, -20.0), (40.0, 10.0, 40.0), (8, 1, 8)
    db = lqref.LqDB(origin, size, div)
    pts = g0.uniform(-25, 25, size=(500, 3)).astype(np.float32)
    pts[:, 1] = g0.uniform(-4, 4, size=500)
    g = orc.grid(origin, size, div)
    for p in pts:
        b = db.bin_index(*p)
        if b >= 0:
            assert orc.cell_of(g, *p) == b
    db.close()
