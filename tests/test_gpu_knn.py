"""GPU parity: bin database and exact KNN (binned and brute force) against the
CPU oracle, bit-exact (indices and fp32 distances)."""
import numpy as np
import pytest

from gpusteer_amd import scenarios
from helpers import lattice_points

pytestmark = pytest.mark.gpu


def _group(gs, ctx, pos):
    st = scenarios.make_agents(pos, np.tile([[0.0, 0.0, 1.0]], (pos.shape[0], 1)), np.zeros(pos.shape[0]), 0.5,
                               1.0, 1.0)
    return gs.AgentGroup(ctx, st)


# The binned search unless a test asks for the default dispatch (which takes
# an exhaustive scan for small query groups and small or sparse targets).
BINNED = 2  # GS_KNN_BINNED (gpusteer_amd.KNN_BINNED)


def _run_binned(gs, ctx, a_pos, b_pos, k, origin, size, cells, self_group, flags=BINNED, repeat=1, sr=1):
    bins = gs.KNNBinData(ctx, cells, size, origin=origin)
    A = _group(gs, ctx, a_pos)
    B = A if self_group else _group(gs, ctx, b_pos)
    gs.updateKNNDatabase(A, bins)
    if not self_group:
        gs.updateKNNDatabase(B, bins)
    knn = gs.KNNData(ctx, a_pos.shape[0], k)
    for _ in range(repeat):
        gs.findKNearestNeighbors(A, knn, bins, B, sr, flags)
    out = knn.download()
    knn.close()
    A.close()
    if not self_group:
        B.close()
    bins.close()
    return out


CASES = [
    # name, n, k, cells, search radius (fast-path column radius)
    ("c0-boids-k7", 4096, 7, (16, 16, 16), 1),
    ("c0-boids-k1", 4096, 1, (16, 16, 16), 1),
    # k = 7 has lists of exactly 7 entries; other k <= 8 use 8-entry lists
    ("c0-boids-k8", 4096, 8, (16, 16, 16), 1),
    ("c0-boids-k5-r2", 4096, 5, (16, 16, 16), 2),
    ("c0-boids-k16", 4096, 16, (16, 16, 16), 1),
    ("c0-boids-k32", 4096, 32, (16, 16, 16), 1),
    ("coarse-1cell", 500, 7, (1, 1, 1), 1),
    ("fine-64", 3000, 9, (64, 64, 64), 1),
    ("c0-boids-k7-r2", 4096, 7, (16, 16, 16), 2),
    ("fine-32-k16-r2", 4096, 16, (32, 32, 32), 2),
    ("fine-64-r3", 3000, 7, (64, 64, 64), 3),
    # finer z cells: boxes of R cells in x, y and R * zr in z
    ("aniso-z2", 4096, 7, (16, 16, 32), 1),
    ("aniso-z3-k16", 4096, 16, (16, 16, 48), 1),
    ("aniso-fine-z2", 3000, 9, (48, 48, 96), 1),
    ("aniso-z2-r2", 4096, 7, (32, 32, 64), 2),
]


@pytest.mark.parametrize("name,n,k,cells,sr", CASES)
def test_binned_equals_oracle_brute(gs, ctx, orc, name, n, k, cells, sr):
    print(f"case {name}", flush=True)
    st, p = scenarios.boids_sphere(n=n)
    pos = st.pos()
    r_idx, r_dist = orc.knn_brute(pos, pos, k, True)
    idx, dist = _run_binned(gs, ctx, pos, pos, k, p.grid_origin, p.grid_size, cells, True, sr=sr)
    np.testing.assert_array_equal(idx, r_idx)
    np.testing.assert_array_equal(dist, r_dist)


def test_brute_equals_oracle(gs, ctx, orc):
    st, _ = scenarios.boids_sphere(n=3000)
    pos = st.pos()
    r_idx, r_dist = orc.knn_brute(pos, pos, 7, True)
    A = _group(gs, ctx, pos)
    knn = gs.KNNData(ctx, pos.shape[0], 7)
    gs.findKNearestNeighborsBruteForce(A, knn, A)
    idx, dist = knn.download()
    np.testing.assert_array_equal(idx, r_idx)
    np.testing.assert_array_equal(dist, r_dist)


@pytest.mark.parametrize("dims", [2, 3])
def test_lattice_ties_lower_index_wins(gs, ctx, orc, dims):
    # Integer lattice, shuffled so index order differs from spatial order.
    pts = lattice_points(12 if dims == 3 else 40, 1.0, dims=dims)
    perm = np.random.default_rng(3).permutation(pts.shape[0])
    pts = np.ascontiguousarray(pts[perm])
    k = 12
    r_idx, r_dist = orc.knn_brute(pts, pts, k, True)
    cells = (6, 6, 6) if dims == 3 else (10, 1, 10)
    size = (12.0, 12.0, 12.0) if dims == 3 else (40.0, 2.0, 40.0)
    origin = (-0.5, -0.5, -0.5) if dims == 3 else (-0.5, -1.0, -0.5)
    idx, dist = _run_binned(gs, ctx, pts, pts, k, origin, size, cells, True)
    np.testing.assert_array_equal(idx, r_idx)
    np.testing.assert_array_equal(dist, r_dist)
    # Ties exist and the lower index comes first among equal distances.
    d = r_dist
    ties = (d[1:] == d[:-1])
    assert ties.any()
    assert np.all(r_idx[1:][ties] > r_idx[:-1][ties])


def test_two_groups_and_padding(gs, ctx, orc):
    g = np.random.default_rng(5)
    a = g.uniform(-50, 50, size=(2000, 3)).astype(np.float32)
    b = g.uniform(-50, 50, size=(5, 3)).astype(np.float32)  # fewer than k
    k = 8
    r_idx, r_dist = orc.knn_brute(a, b, k, False)
    idx, dist = _run_binned(gs, ctx, a, b, k, (-55, -55, -55), (110, 110, 110), (8, 8, 8), False)
    np.testing.assert_array_equal(idx, r_idx)
    np.testing.assert_array_equal(dist, r_dist)
    assert np.all(idx[5:] == 0xFFFFFFFF)
    assert np.all(dist[5:] == np.finfo(np.float32).max)


def test_out_of_grid_points(gs, ctx, orc):
    # Lattice covers [-10, 10]^3, points spread over [-40, 40]^3: clamped
    # boundary cells must still give exact answers.
    g = np.random.default_rng(7)
    pos = g.uniform(-40, 40, size=(3000, 3)).astype(np.float32)
    r_idx, r_dist = orc.knn_brute(pos, pos, 7, True)
    idx, dist = _run_binned(gs, ctx, pos, pos, 7, (-10, -10, -10), (20, 20, 20), (4, 4, 4), True)
    np.testing.assert_array_equal(idx, r_idx)
    np.testing.assert_array_equal(dist, r_dist)


@pytest.mark.parametrize("sr", [1, 2])
def test_seeded_call_equals_unseeded(gs, ctx, orc, sr):
    st, p = scenarios.boids_sphere(n=4096)
    pos = st.pos()
    r_idx, r_dist = orc.knn_brute(pos, pos, 7, True)
    # Second call on the same container uses the first result as its bound.
    idx, dist = _run_binned(gs, ctx, pos, pos, 7, p.grid_origin, p.grid_size, p.grid_cells, True, repeat=3,
                            sr=sr)
    np.testing.assert_array_equal(idx, r_idx)
    np.testing.assert_array_equal(dist, r_dist)


def test_bin_database_matches_oracle(gs, ctx, orc):
    for name, pos, origin, size, cells in [
        ("c0", scenarios.boids_sphere(n=4096)[0].pos(), (-55,) * 3, (110,) * 3, (16, 16, 16)),
        ("dense-2d", scenarios.choke(n_total=4000).groups[0].pos(), (-100, -5, -100), (200, 10, 200), (40, 1, 40)),
        ("out-of-grid", np.random.default_rng(1).uniform(-30, 30, size=(2000, 3)).astype(np.float32),
         (-10,) * 3, (20,) * 3, (5, 5, 5)),
        ("one-point", np.zeros((1, 3), np.float32), (-1,) * 3, (2,) * 3, (2, 2, 2)),
        # 5,000 agents in each of two cells: members above the old 4,096 cap
        # still come out in ascending agent index.
        ("crowded-cells", np.random.default_rng(2).uniform(-10, 10, size=(10000, 3)).astype(np.float32),
         (-10,) * 3, (20,) * 3, (2, 1, 1)),
    ]:
        print(f"case {name}", flush=True)
        g = orc.grid(origin, size, cells)
        r_key, r_start, r_perm, r_sp = orc.bins_build(g, pos)
        bins = gs.KNNBinData(ctx, cells, size, origin=origin)
        A = _group(gs, ctx, pos)
        gs.updateKNNDatabase(A, bins)
        key, start, perm, sp = A.download_bins(bins.num_cells)
        np.testing.assert_array_equal(key, r_key, err_msg=name)
        np.testing.assert_array_equal(start, r_start, err_msg=name)
        np.testing.assert_array_equal(perm, r_perm, err_msg=name)
        np.testing.assert_array_equal(sp, r_sp, err_msg=name)
        # Rebuild on the same group (counts must have been cleared).
        gs.updateKNNDatabase(A, bins)
        key2, start2, perm2, _ = A.download_bins(bins.num_cells)
        np.testing.assert_array_equal(start2, r_start, err_msg=name)
        np.testing.assert_array_equal(perm2, r_perm, err_msg=name)
        A.close()
        bins.close()


def test_bin_database_across_lattices(gs, ctx, orc):
    """One group rebuilt on lattices of 2 tiles of 2,048 cells, 2,560,000
    cells (more than 1,024 tiles: the two-pass scan) and 256 tiles, back and
    forth, two builds each: the single-pass scan's tile sums (kept between
    builds) never carry counts over from another lattice or build."""
    pos = scenarios.boids_sphere(n=4096)[0].pos()
    origin, size = (-55,) * 3, (110,) * 3
    A = _group(gs, ctx, pos)
    for cells in [(16, 16, 16), (160, 160, 100), (16, 16, 16), (64, 64, 128), (160, 160, 100), (64, 64, 128)]:
        print(f"lattice {cells}", flush=True)
        g = orc.grid(origin, size, cells)
        r_key, r_start, r_perm, _ = orc.bins_build(g, pos)
        bins = gs.KNNBinData(ctx, cells, size, origin=origin)
        for rep in range(2):
            gs.updateKNNDatabase(A, bins)
            key, start, perm, _ = A.download_bins(bins.num_cells)
            np.testing.assert_array_equal(start, r_start, err_msg=f"{cells} build {rep}")
            np.testing.assert_array_equal(perm, r_perm, err_msg=f"{cells} build {rep}")
        bins.close()
    A.close()


@pytest.mark.parametrize("k,cells", [(7, (64, 64, 64)), (7, (64, 64, 192)), (16, (48, 48, 96)), (16, (48, 48, 144))])
def test_large_flock_subset_against_brute(gs, ctx, orc, k, cells):
    """2^20 agents in a 1000^3 box at k = 7 (64^3, and 64 x 64 x 192, the
    autotuner's lattice) and k = 16 (48 x 48 x 96 and 48 x 48 x 144, its
    lattices for k = 16):
    binned KNN for every agent, checked against the oracle's brute force on
    16,384 random queries (each against all 2^20 agents)."""
    print(f"flock 2^20 k={k} cells={cells}", flush=True)
    st, p = scenarios.boids_box(n=1 << 20, k=k)
    pos = st.pos()
    idx, dist = _run_binned(gs, ctx, pos, pos, k, p.grid_origin, p.grid_size, cells, True, repeat=2)
    q = np.random.default_rng(11).choice(pos.shape[0], size=16384, replace=False).astype(np.uint32)
    r_idx, r_dist = orc.knn_brute(pos, pos, k, True, queries=q)
    np.testing.assert_array_equal(idx[:, q], r_idx[:, q])
    np.testing.assert_array_equal(dist[:, q], r_dist[:, q])
    # Size-independent properties for all agents: no self, ascending keys.
    n = pos.shape[0]
    assert not np.any(idx == np.arange(n, dtype=np.uint32)[None, :])
    assert np.all(dist[1:] >= dist[:-1])


@pytest.mark.parametrize("cap", [2.0, 7.5])
@pytest.mark.parametrize("self_group", [True, False])
def test_capped_knn_equals_oracle(gs, ctx, orc, cap, self_group):
    """gs_find_knn_within: the k nearest among points closer than the cap
    (strict, lq.c's locality test), padded when fewer qualify."""
    g = np.random.default_rng(21)
    a = g.uniform(-50, 50, size=(3000, 3)).astype(np.float32)
    b = a if self_group else g.uniform(-50, 50, size=(2500, 3)).astype(np.float32)
    k = 7
    r_idx, r_dist = orc.knn_brute(a, b, k, self_group, cap=cap)
    bins = gs.KNNBinData(ctx, (10, 10, 10), (110.0,) * 3, origin=(-55.0,) * 3)
    A = _group(gs, ctx, a)
    B = A if self_group else _group(gs, ctx, b)
    gs.updateKNNDatabase(A, bins)
    if not self_group:
        gs.updateKNNDatabase(B, bins)
    knn = gs.KNNData(ctx, a.shape[0], k)
    for _ in range(2):  # the second call is seeded by the first
        gs.findKNearestNeighborsWithin(A, knn, bins, B, cap)
        idx, dist = knn.download()
        np.testing.assert_array_equal(idx, r_idx)
        np.testing.assert_array_equal(dist, r_dist)
    assert np.any(idx == 0xFFFFFFFF) and np.all(dist[idx != 0xFFFFFFFF] < cap)
    for h in (knn, A, bins) + (() if self_group else (B,)):
        h.close()


@pytest.mark.parametrize("flags", [BINNED, 0])
def test_capped_knn_across_moving_groups(gs, ctx, orc, flags):
    """A capped search of group A in group B while B moves from beyond the cap
    (every list empty; from the second build on, the search skips queries by
    B's bounding box) through partial overlap to full overlap and back, one
    rebuild of B per stage, each search done twice (the second seeded): equal
    to the oracle at every stage, with and without the forced binned search."""
    g = np.random.default_rng(31)
    a = g.uniform(-20, 20, size=(4000, 3)).astype(np.float32)
    b0 = g.uniform(-10, 10, size=(3000, 3)).astype(np.float32)
    cap, k = 3.0, 7
    bins = gs.KNNBinData(ctx, (24, 24, 24), (120.0,) * 3, origin=(-60.0,) * 3)
    A = _group(gs, ctx, a)
    gs.updateKNNDatabase(A, bins)
    B = _group(gs, ctx, b0)
    knn = gs.KNNData(ctx, a.shape[0], k)
    found = []
    for shift in (45.0, 40.0, 33.0, 28.0, 20.0, 8.0, 0.0, 30.0, 45.0):
        print(f"B shifted by {shift}", flush=True)
        b = b0 + np.float32(shift)
        B.upload(scenarios.make_agents(b, np.tile([[0.0, 0.0, 1.0]], (b.shape[0], 1)), np.zeros(b.shape[0]), 0.5,
                                       1.0, 1.0))
        gs.updateKNNDatabase(B, bins)
        r_idx, r_dist = orc.knn_brute(a, b, k, False, cap=cap)
        for rep in range(2):
            gs.findKNearestNeighborsWithin(A, knn, bins, B, cap, 1, flags)
            idx, dist = knn.download()
            np.testing.assert_array_equal(idx, r_idx, err_msg=f"shift {shift} call {rep}")
            np.testing.assert_array_equal(dist, r_dist, err_msg=f"shift {shift} call {rep}")
        found.append(int(np.sum(idx != 0xFFFFFFFF)))
    # The stages cover empty, partial and full overlap.
    assert found[0] == 0 and found[-1] == 0 and 0 < found[4] < found[6]
    for h in (knn, A, B, bins):
        h.close()


@pytest.mark.parametrize("nb", [1, 6, 50, 256])
def test_small_target_set(gs, ctx, orc, nb):
    """Target sets of at most 256 points are scanned exhaustively (walls,
    obstacles; 20,000 queries, above the split scan's query limit); the
    result is the same exact KNN."""
    g = np.random.default_rng(nb)
    a = g.uniform(-50, 50, size=(20000, 3)).astype(np.float32)
    b = g.uniform(-50, 50, size=(nb, 3)).astype(np.float32)
    k = 3
    r_idx, r_dist = orc.knn_brute(a, b, k, False)
    idx, dist = _run_binned(gs, ctx, a, b, k, (-55,) * 3, (110,) * 3, (40, 40, 40), False, flags=0)
    np.testing.assert_array_equal(idx, r_idx)
    np.testing.assert_array_equal(dist, r_dist)


@pytest.mark.parametrize("cap", [float("inf"), 40.0, 1.5], ids=["uncapped", "cap40", "cap1.5"])
def test_sparse_target_set(gs, ctx, orc, cap):
    """3000 targets over 65,536 cells (sparse: the split walls' case): with
    no cap or a wide one the search scans them exhaustively, with a narrow
    cap it walks the bins; the result is the same exact KNN."""
    g = np.random.default_rng(77)
    a = np.stack([g.uniform(-128, 128, 20000), g.uniform(-4, 4, 20000), g.uniform(-128, 128, 20000)],
                 axis=1).astype(np.float32)
    b = np.stack([g.uniform(-128, 128, 3000), np.zeros(3000), g.uniform(-128, 128, 3000)], axis=1).astype(np.float32)
    k = 3
    r_idx, r_dist = orc.knn_brute(a, b, k, False, cap=cap)
    bins = gs.KNNBinData(ctx, (256, 1, 256), (260.0, 10.0, 260.0))
    A, B = _group(gs, ctx, a), _group(gs, ctx, b)
    gs.updateKNNDatabase(A, bins)
    gs.updateKNNDatabase(B, bins)
    knn = gs.KNNData(ctx, a.shape[0], k)
    gs.findKNearestNeighborsWithin(A, knn, bins, B, cap)
    idx, dist = knn.download()
    np.testing.assert_array_equal(idx, r_idx)
    np.testing.assert_array_equal(dist, r_dist)
    for h in (knn, A, B, bins):
        h.close()


SPLIT_CASES = [
    # name, nA, nB, k, cap (the 12^3 lattice's cells are 10.8 wide: every
    # case reaches more than 64 cells, so the default dispatch splits)
    ("5000x5000-k7", 5000, 5000, 7, float("inf")),
    ("5000x3000-k7-cap12", 5000, 3000, 7, 12.0),
    ("1000x129-k16", 1000, 129, 16, float("inf")),
    ("4096x16384-k8", 4096, 16384, 8, float("inf")),
    ("777x900-k32", 777, 900, 32, float("inf")),
    ("300x200-k7-cap20-padded", 300, 200, 7, 20.0),
    ("2048x2048-k1", 2048, 2048, 1, float("inf")),
]


@pytest.mark.parametrize("name,na,nb,k,cap", SPLIT_CASES)
def test_split_exhaustive_equals_oracle(gs, ctx, orc, name, na, nb, k, cap):
    """A small query group searching another small group over a wide reach
    takes the split exhaustive scan (B cut into up to 64 ranges, partial
    lists merged by a wavefront per query): KNN sets and distances equal the
    oracle's brute force, including exact duplicates among B's positions and
    queries sitting on B's points (ties broken by agent index), a cap that
    leaves rows padded, ranges of 65 points and ragged query counts,
    unseeded and seeded by the previous call; the binned search gives the
    same result."""
    print(f"case {name}", flush=True)
    g = np.random.default_rng(na + nb + k)
    a = g.uniform(-60, 60, size=(na, 3)).astype(np.float32)
    b = g.uniform(-60, 60, size=(nb, 3)).astype(np.float32)
    b[1::17] = b[::17][: len(b[1::17])]  # duplicated positions
    m = min(na, nb) // 5
    a[:m] = b[:m]  # queries on B's points
    r_idx, r_dist = orc.knn_brute(a, b, k, False, cap=cap)
    bins = gs.KNNBinData(ctx, (12, 12, 12), (130.0,) * 3, origin=(-65.0,) * 3)
    A, B = _group(gs, ctx, a), _group(gs, ctx, b)
    gs.updateKNNDatabase(A, bins)
    gs.updateKNNDatabase(B, bins)
    for flags in (0, BINNED):
        knn = gs.KNNData(ctx, na, k)
        for call in range(2):  # the second call is seeded by the first
            gs.findKNearestNeighborsWithin(A, knn, bins, B, cap, flags=flags)
            idx, dist = knn.download()
            np.testing.assert_array_equal(idx, r_idx, err_msg=f"flags {flags} call {call}")
            np.testing.assert_array_equal(dist, r_dist, err_msg=f"flags {flags} call {call}")
        knn.close()
    if "padded" in name:
        assert np.any(r_idx == 0xFFFFFFFF)
    for h in (A, B, bins):
        h.close()


@pytest.mark.parametrize("flags", [0, BINNED], ids=["default", "binned"])
def test_empty_groups(gs, ctx, orc, flags):
    """An empty query group gets an empty result and an empty target group
    all pads, through every search the dispatch can take (uncapped and over
    a wide cap: exhaustive and split scans, bins); behaviours and the
    integration on an empty group do nothing."""
    g = np.random.default_rng(3)
    b = g.uniform(-40, 40, size=(300, 3)).astype(np.float32)
    empty = np.zeros((0, 3), np.float32)
    bins = gs.KNNBinData(ctx, (8, 8, 8), (100.0,) * 3, origin=(-50.0,) * 3)
    E, B = _group(gs, ctx, empty), _group(gs, ctx, b)
    gs.updateKNNDatabase(E, bins)
    gs.updateKNNDatabase(B, bins)
    for cap in (float("inf"), 60.0):
        print(f"empty A, cap {cap}, flags {flags}", flush=True)
        knn = gs.KNNData(ctx, 0, 7)
        gs.findKNearestNeighborsWithin(E, knn, bins, B, cap, flags=flags)
        idx, dist = knn.download()
        assert idx.shape == (7, 0) and dist.shape == (7, 0)
        gs.steerForSeparation(E, knn, B, 0.0, 5.0, -0.7, 1.0)
        knn.close()
        print(f"empty B, cap {cap}, flags {flags}", flush=True)
        knn = gs.KNNData(ctx, 300, 7)
        gs.findKNearestNeighborsWithin(B, knn, bins, E, cap, flags=flags)
        idx, dist = knn.download()
        assert np.all(idx == 0xFFFFFFFF) and np.all(dist == np.finfo(np.float32).max)
        knn.close()
    gs.updateGroup(E, 0.1)
    gs.wrapWorld(E, (100.0, 100.0, 100.0))
    assert E.download().n == 0
    for h in (E, B, bins):
        h.close()


def test_lattice_above_2_pow_24_cells(gs, ctx, orc):
    """A 512 x 256 x 256 lattice (2^25 cells of one unit) with 8,000 agents
    packed into a 20-unit cube near its far x end, so their cell indices lie
    above 2^24: the bin database (keys, starts, order) and the binned KNN
    equal the oracle's."""
    g = np.random.default_rng(25)
    pos = np.stack([g.uniform(200, 220, 8000), g.uniform(-10, 10, 8000), g.uniform(-10, 10, 8000)],
                   axis=1).astype(np.float32)
    origin, size, cells = (-256.0, -128.0, -128.0), (512.0, 256.0, 256.0), (512, 256, 256)
    og = orc.grid(origin, size, cells)
    r_key, r_start, r_perm, r_sp = orc.bins_build(og, pos)
    assert r_key.min() > (1 << 24)
    bins = gs.KNNBinData(ctx, cells, size, origin=origin)
    A = _group(gs, ctx, pos)
    print("lattice 2^25: bins", flush=True)
    gs.updateKNNDatabase(A, bins)
    key, start, perm, sp = A.download_bins(bins.num_cells)
    np.testing.assert_array_equal(key, r_key)
    np.testing.assert_array_equal(start, r_start)
    np.testing.assert_array_equal(perm, r_perm)
    np.testing.assert_array_equal(sp, r_sp)
    print("lattice 2^25: knn", flush=True)
    r_idx, r_dist = orc.knn_brute(pos, pos, 7, True)
    knn = gs.KNNData(ctx, pos.shape[0], 7)
    gs.findKNearestNeighbors(A, knn, bins, A, 1, BINNED)
    idx, dist = knn.download()
    np.testing.assert_array_equal(idx, r_idx)
    np.testing.assert_array_equal(dist, r_dist)
    for h in (knn, A, bins):
        h.close()


@pytest.mark.parametrize("fused", [False, True], ids=["binned", "fused-step"])
def test_isolated_agents_on_a_wide_lattice(gs, ctx, orc, fused):
    """Five agents (fewer than k + 1) on a 1400 x 1400 x 64 lattice over a
    1000^3 box: every query's rings would cross the lattice before finding
    k neighbours; once a ring's box holds more cells than the lattice holds
    points the search scans the points instead.  The forced binned search
    (GS_KNN_BINNED) and the fused Boids step equal the oracle's brute force,
    padded rows included."""
    from gpusteer_amd import scenarios as sc

    st, p = sc.boids_box(n=5, k=7)
    p.grid_cells = (1400, 1400, 64)
    pos = st.pos()
    print(f"isolated agents, fused={fused}", flush=True)
    bins = gs.KNNBinData(ctx, p.grid_cells, p.grid_size, origin=p.grid_origin)
    A = gs.AgentGroup(ctx, st)
    knn = gs.KNNData(ctx, st.n, p.k)
    if fused:
        ref = st.copy()
        gs.stepBoids(A, knn, bins, p, 2)
        for _ in range(2):
            r_idx, r_dist = orc.boids_step(ref, p, knn_fn=lambda pp: orc.knn_brute(pp, pp, p.k, True))
        from helpers import assert_state_equal
        assert_state_equal(A.download(), ref)
    else:
        gs.updateKNNDatabase(A, bins)
        gs.findKNearestNeighbors(A, knn, bins, A, 1, BINNED)
        r_idx, r_dist = orc.knn_brute(pos, pos, p.k, True)
    idx, dist = knn.download()
    np.testing.assert_array_equal(idx, r_idx)
    np.testing.assert_array_equal(dist, r_dist)
    assert (idx[4:] == 0xFFFFFFFF).all()
    for h in (knn, A, bins):
        h.close()
