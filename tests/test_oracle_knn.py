"""Pins the oracle's bin build and KNN.

* bins: the counting sort equals a stable sort by cell (numpy lexsort), start
  is the exclusive scan of the cell counts (numpy cumsum of bincount);
* KNN: binned equals brute force on random, tied-lattice, out-of-grid, 2-D
  and two-group inputs; brute force equals a pure-Python sort by
  (fp32 d2, index) on small inputs (the order KNNBruteForceCUDAKernelV2's
  strict '<' scan produces, include/OpenSteer/CUDA/KNNBruteForceCudaKernel.cu:245-270).
"""
import numpy as np
import pytest

from gpusteer_amd import scenarios
from helpers import lattice_points


def test_bins_build_is_stable_counting_sort(orc):
    st, p = scenarios.boids_sphere(n=3000)
    pos = st.pos()
    g = orc.grid(p.grid_origin, p.grid_size, p.grid_cells)
    key, start, perm, sp = orc.bins_build(g, pos)
    C = int(np.prod(p.grid_cells))
    np.testing.assert_array_equal(perm, np.lexsort((np.arange(pos.shape[0]), key)))
    counts = np.bincount(key, minlength=C)
    np.testing.assert_array_equal(start, np.concatenate([[0], np.cumsum(counts)]))
    np.testing.assert_array_equal(sp, pos[perm])


def _py_knn(a, b, k, self_group):
    """Pure-Python reference: every pair's fp32 d2, sorted by (d2, index)."""
    out = np.full((k, a.shape[0]), 0xFFFFFFFF, np.uint32)
    for q in range(a.shape[0]):
        dx = (a[q, 0] - b[:, 0]).astype(np.float32)
        dy = (a[q, 1] - b[:, 1]).astype(np.float32)
        dz = (a[q, 2] - b[:, 2]).astype(np.float32)
        d2 = (dx * dx + dy * dy) + dz * dz
        cand = [(float(d2[j]), j) for j in range(b.shape[0]) if not (self_group and j == q)]
        cand.sort()
        for s, (_, j) in enumerate(cand[:k]):
            out[s, q] = j
    return out


@pytest.mark.parametrize("k", [1, 3, 7, 12])
def test_brute_equals_python_sort(orc, k):
    g = np.random.default_rng(k)
    pts = lattice_points(5, 1.0)  # many exact ties
    pts = np.ascontiguousarray(pts[g.permutation(pts.shape[0])][:90])
    idx, dist = orc.knn_brute(pts, pts, k, True)
    np.testing.assert_array_equal(idx, _py_knn(pts, pts, k, True))
    assert np.all(np.diff(dist, axis=0) >= 0)


def test_ties_resolve_to_lower_index(orc):
    # Query at the origin, 6 points at distance 1 in shuffled index order.
    pts = np.array([[0, 0, 0], [0, 0, 1], [1, 0, 0], [0, -1, 0], [-1, 0, 0], [0, 1, 0], [0, 0, -1]], np.float32)
    idx, dist = orc.knn_brute(pts, pts, 6, True)
    np.testing.assert_array_equal(idx[:, 0], [1, 2, 3, 4, 5, 6])
    assert np.all(dist[:, 0] == 1.0)


CASES = [
    ("sphere-16", lambda: scenarios.boids_sphere(n=2500)[0].pos(), (-55,) * 3, (110,) * 3, (16, 16, 16), 7),
    ("lattice-ties", lambda: lattice_points(11, 1.0), (-0.5,) * 3, (11,) * 3, (4, 4, 4), 10),
    ("out-of-grid", lambda: np.random.default_rng(2).uniform(-40, 40, (2000, 3)).astype(np.float32), (-10,) * 3,
     (20,) * 3, (4, 4, 4), 7),
    ("2d-dense", lambda: scenarios.choke(n_total=3000).groups[0].pos(), (-100, -5, -100), (200, 10, 200),
     (40, 1, 40), 7),
    ("k32", lambda: scenarios.boids_sphere(n=1500)[0].pos(), (-55,) * 3, (110,) * 3, (8, 8, 8), 32),
    ("tiny-padded", lambda: np.random.default_rng(3).uniform(-5, 5, (5, 3)).astype(np.float32), (-5,) * 3,
     (10,) * 3, (2, 2, 2), 7),
]


@pytest.mark.parametrize("name,mk,origin,size,cells,k", CASES)
def test_binned_equals_brute(orc, name, mk, origin, size, cells, k):
    pos = mk()
    g = orc.grid(origin, size, cells)
    a = orc.knn_brute(pos, pos, k, True)
    b = orc.knn_binned(g, pos, pos, k, True)
    np.testing.assert_array_equal(a[0], b[0])
    np.testing.assert_array_equal(a[1], b[1])
    if pos.shape[0] - 1 < k:
        assert np.all(a[0][pos.shape[0] - 1:] == 0xFFFFFFFF)
        assert np.all(a[1][pos.shape[0] - 1:] == np.finfo(np.float32).max)


def test_two_groups(orc):
    g0 = np.random.default_rng(4)
    a = g0.uniform(-50, 50, (1200, 3)).astype(np.float32)
    b = g0.uniform(-50, 50, (700, 3)).astype(np.float32)
    g = orc.grid((-55,) * 3, (110,) * 3, (8, 8, 8))
    x = orc.knn_brute(a, b, 9, False)
    y = orc.knn_binned(g, a, b, 9, False)
    np.testing.assert_array_equal(x[0], y[0])
    np.testing.assert_array_equal(x[1], y[1])
    np.testing.assert_array_equal(x[0][:, :50], _py_knn(a[:50], b, 9, False))
