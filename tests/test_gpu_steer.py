"""GPU parity of the behaviours, integration and wrap against the CPU oracle.

The HIP kernels and the oracle perform the same sequence of correctly rounded
fp32 operations (both built without FMA contraction), so the results are
compared for exact equality; a 1e-5 relative tolerance on the steering
vectors is also asserted separately as a second, weaker bar."""
import numpy as np
import pytest

from gpusteer_amd import scenarios
from helpers import assert_state_equal

pytestmark = pytest.mark.gpu

MODES = [0, 1]  # OPENSTEER, GPUREF


def _knn(orc, st, other=None, k=7):
    pos = st.pos()
    opos = pos if other is None else other.pos()
    return orc.knn_brute(pos, opos, k, other is None)


def _upload_knn_via_brute(gs, ctx, A, B, k):
    knn = gs.KNNData(ctx, A.Size(), k)
    gs.findKNearestNeighborsBruteForce(A, knn, B)
    return knn


def _state_with_bits(n=4096, seed=1):
    st, p = scenarios.boids_sphere(n=n)
    g = np.random.default_rng(seed)
    # Random pre-existing steering and bits exercise accumulate + masks.
    for f in ("gx", "gy", "gz"):
        st.arrays[f][:] = g.normal(size=n).astype(np.float32)
    st.bits[:] = g.choice([0, 16, 32, 64, 256, 4096], size=n).astype(np.uint32)
    return st, p


@pytest.mark.parametrize("mode", MODES)
@pytest.mark.parametrize("which", ["separation", "alignment", "cohesion"])
def test_flock_term(gs, ctx, orc, mode, which):
    print(f"case {which} mode {mode}", flush=True)
    st, p = _state_with_bits()
    ref = st.copy()
    A = gs.AgentGroup(ctx, st)
    knn = _upload_knn_via_brute(gs, ctx, A, A, 7)
    idx, _ = _knn(orc, ref)
    t = getattr(p, which)
    min_d = t.min_d if mode == 0 else 1.5
    mask = 4096 | (16 if which != "separation" else 0)
    fn = {"separation": gs.steerForSeparation, "alignment": gs.steerForAlignment,
          "cohesion": gs.steerForCohesion}[which]
    fn(A, knn, A, min_d, t.max_d, t.cos_max, t.weight, mask, mode)
    w = {"separation": 0, "alignment": 1, "cohesion": 2}[which]
    orc.steer_flock(w, ref, idx, ref, min_d, t.max_d, t.cos_max, t.weight, mask, mode)
    out = A.download()
    assert_state_equal(out, ref)
    g = np.stack([out.arrays[f] for f in ("gx", "gy", "gz")])
    r = np.stack([ref.arrays[f] for f in ("gx", "gy", "gz")])
    np.testing.assert_allclose(g, r, rtol=1e-5, atol=1e-6)
    A.close()
    knn.close()


@pytest.mark.parametrize("mode", MODES)
def test_fused_flocking_equals_three_calls(gs, ctx, orc, mode):
    st, p = _state_with_bits(seed=2)
    A1 = gs.AgentGroup(ctx, st)
    A2 = gs.AgentGroup(ctx, st)
    knn1 = _upload_knn_via_brute(gs, ctx, A1, A1, 7)
    knn2 = _upload_knn_via_brute(gs, ctx, A2, A2, 7)
    terms = [p.separation, p.alignment, p.cohesion]
    for t in terms:
        t.mask = 272  # AVOID_WALLS | SEPARATION: alignment/cohesion see separation's bit
    p.separation.mask = 4096
    md = (lambda t: t.min_d) if mode == 0 else (lambda t: 1.5)
    gs.steerForSeparation(A1, knn1, A1, md(p.separation), p.separation.max_d, p.separation.cos_max,
                          p.separation.weight, p.separation.mask, mode)
    gs.steerForAlignment(A1, knn1, A1, md(p.alignment), p.alignment.max_d, p.alignment.cos_max,
                         p.alignment.weight, p.alignment.mask, mode)
    gs.steerForCohesion(A1, knn1, A1, md(p.cohesion), p.cohesion.max_d, p.cohesion.cos_max, p.cohesion.weight,
                        p.cohesion.mask, mode)
    if mode == 1:
        for t in terms:
            t.min_d = 1.5
    gs.steerForFlocking(A2, knn2, A2, p.separation, p.alignment, p.cohesion, mode)
    assert_state_equal(A1.download(), A2.download())
    for h in (A1, A2, knn1, knn2):
        h.close()


@pytest.mark.parametrize("mode", MODES)
def test_seek(gs, ctx, orc, mode):
    st, _ = _state_with_bits(seed=3)
    ref = st.copy()
    A = gs.AgentGroup(ctx, st)
    tgt = (3.0, -7.0, 11.0)
    gs.steerForSeek(A, tgt, 1.5, 4480, mode)
    orc.steer_seek(ref, tgt, 1.5, 4480, mode)
    assert_state_equal(A.download(), ref)
    A.close()


def test_avoid_obstacles(gs, ctx, orc):
    sc = scenarios.choke(n_total=6000)
    st = sc.groups[0]
    g = np.random.default_rng(4)
    st.arrays["speed"][:] = g.uniform(0, 3, size=st.n).astype(np.float32)
    ref = st.copy()
    A = gs.AgentGroup(ctx, st)
    O = gs.ObstacleGroup(ctx, *sc.obstacles)
    knn = gs.KNNData(ctx, st.n, 2)
    gs.findKNearestNeighborsBruteForce(A, knn, O)
    kidx, _ = orc.knn_brute(st.pos(), np.stack(sc.obstacles[:3], axis=1), 2, False)
    np.testing.assert_array_equal(knn.download()[0], kidx)
    gs.steerToAvoidObstacles(A, O, knn, 5.0, 1.0, 0)
    orc.avoid_obstacles(ref, kidx, sc.obstacles, 5.0, 1.0, 0)
    out = A.download()
    assert_state_equal(out, ref)
    assert np.any(out.bits & 128), "no agent avoided an obstacle: the case is too weak"
    for h in (A, O, knn):
        h.close()


def test_avoid_walls(gs, ctx, orc):
    sc = scenarios.choke(n_total=6000, offset=0.2)  # many agents start at the wall
    st = sc.groups[0]
    g = np.random.default_rng(6)
    st.arrays["speed"][:] = g.uniform(0, 3, size=st.n).astype(np.float32)
    ref = st.copy()
    A = gs.AgentGroup(ctx, st)
    W = gs.WallGroup(ctx, sc.walls)
    knn = gs.KNNData(ctx, st.n, 3)
    gs.findKNearestNeighborsBruteForce(A, knn, W)
    kidx, _ = orc.knn_brute(st.pos(), sc.walls.mid(), 3, False)
    np.testing.assert_array_equal(knn.download()[0], kidx)
    gs.steerToAvoidWalls(A, knn, W, 10.0, 1.0, 0)
    orc.avoid_walls(ref, kidx, sc.walls, 10.0, 1.0, 0)
    out = A.download()
    assert_state_equal(out, ref)
    assert np.any(out.bits & 256) and np.any(out.bits & 4096), "case must exercise both wall branches"
    for h in (A, W, knn):
        h.close()


@pytest.mark.parametrize("mode", MODES)
def test_update(gs, ctx, orc, mode):
    st, p = _state_with_bits(seed=7)
    # Include slow agents (steering adjustment branch) and zero force.
    st.arrays["speed"][:64] = np.float32(0.5)
    st.arrays["gx"][64:80] = 0
    st.arrays["gy"][64:80] = 0
    st.arrays["gz"][64:80] = 0
    # Unit and non-unit masses interleaved (a unit mass skips the divisions).
    st.arrays["mass"][1::2] = np.random.default_rng(8).uniform(0.5, 2.0, st.n // 2).astype(np.float32)
    ref = st.copy()
    A = gs.AgentGroup(ctx, st)
    gs.updateGroup(A, p.dt, mode)
    orc.update(ref, p.dt, mode)
    assert_state_equal(A.download(), ref)
    A.close()


@pytest.mark.parametrize("shape", [0, 1])
def test_wrap(gs, ctx, orc, shape):
    st, _ = scenarios.boids_sphere(n=4096)
    g = np.random.default_rng(8)
    for f in ("px", "py", "pz"):
        st.arrays[f][:] = g.uniform(-70, 70, size=st.n).astype(np.float32)
    ref = st.copy()
    A = gs.AgentGroup(ctx, st)
    if shape == 1:
        gs.wrapWorld(A, shape=1, center=(0, 0, 0), radius=50.0)
        orc.wrap_world(ref, 1, (0, 0, 0), (50.0, 0, 0))
    else:
        gs.wrapWorld(A, (100.0, 80.0, 120.0), shape=0, center=(1.0, 2.0, 3.0))
        orc.wrap_world(ref, 0, (1.0, 2.0, 3.0), (100.0, 80.0, 120.0))
    assert_state_equal(A.download(), ref)
    A.close()


def test_check_finite(gs, ctx):
    st, _ = scenarios.boids_sphere(n=1000)
    st.arrays["speed"][417] = np.nan
    A = gs.AgentGroup(ctx, st)
    status, bad = A.check_finite()
    assert status == 8 and bad == 417
    A.close()


def _cone_lattice(n_side=16, spacing=0.7, origin=(100.3, -57.1, 33.9), jitter=2e-5):
    """Agents on a lattice far from the origin, moved by a few ulps of their
    coordinates, all with one forward vector: the cone inputs
    dot(fwd, off / d) of each lattice direction form clusters about 1e-4
    wide, many times the estimate's margin."""
    ax = np.arange(n_side, dtype=np.float32) * np.float32(spacing)
    g = np.stack(np.meshgrid(ax, ax, ax, indexing="ij"), axis=-1).reshape(-1, 3)
    r = np.random.default_rng(5)
    pos = (g + np.asarray(origin, np.float32) + r.uniform(-jitter, jitter, size=g.shape)).astype(np.float32)
    f = np.array([1.0, 2.0, 3.0], np.float32)
    f = (f / np.float32(np.sqrt(np.float32(14.0)))).astype(np.float32)
    fwd = np.tile(f, (len(pos), 1))
    return scenarios.make_agents(pos, fwd, np.full(len(pos), 1.0, np.float32), 0.5, 9.0, 27.0)


def _cone_inputs(st, idx):
    """dot3(fwd, off / d) in the kernels' fp32 order for every KNN pair."""
    pos = st.pos()
    f32 = np.float32
    out = np.empty(idx.shape, np.float32)
    for j in range(idx.shape[0]):
        off = (pos[idx[j]] - pos).astype(f32)
        d2 = ((off[:, 0] * off[:, 0]) + (off[:, 1] * off[:, 1])) + (off[:, 2] * off[:, 2])
        d = np.sqrt(d2).astype(f32)
        q = (off / d[:, None]).astype(f32)
        fw = np.stack([st.arrays["fx"], st.arrays["fy"], st.arrays["fz"]], axis=1)
        out[j] = ((fw[:, 0] * q[:, 0]) + (fw[:, 1] * q[:, 1])) + (fw[:, 2] * q[:, 2])
    return out


@pytest.mark.parametrize("mode", MODES)
def test_cone_thresholds_at_the_cone_inputs(gs, ctx, orc, mode):
    """The kernels decide the cone test `dot(fwd, off / d) > cos_max` from an
    estimate unless it is near the threshold.  Thresholds placed exactly on
    attained cone inputs (the middle of a cluster about 1e-4 wide, so the
    test flips inside it, with pairs both within and beyond the estimate's
    margin of it) give the oracle's results bit for bit for each term alone
    and for the fused flocking."""
    st = _cone_lattice()
    idx, _ = _knn(orc, st)
    c = _cone_inputs(st, idx)
    # Three clusters: the inputs of the first, fourth and last KNN slots'
    # most common rounded values.
    cos = []
    for j in (0, 3, 6):
        vals, counts = np.unique(np.round(c[j].astype(np.float64), 3), return_counts=True)
        centre = vals[np.argmax(counts)]
        cl = np.sort(c[j][np.abs(c[j] - centre) < 1e-3])
        cos.append(float(cl[len(cl) // 2]))
        below, at, above = (np.count_nonzero(cl < cl[len(cl) // 2]), np.count_nonzero(cl == cl[len(cl) // 2]),
                            np.count_nonzero(cl > cl[len(cl) // 2]))
        print(f"slot {j}: threshold {cos[-1]!r} cluster below {below} at {at} above {above}", flush=True)
        assert at > 0 and above > 0
    terms = [scenarios.FlockTerm(0.1, 5.0, cos[0], 1.5, 0), scenarios.FlockTerm(0.1, 5.0, cos[1], 0.75, 0),
             scenarios.FlockTerm(0.1, 5.0, cos[2], 2.0, 0)]
    for w, (name, t) in enumerate(zip(("separation", "alignment", "cohesion"), terms)):
        ref = st.copy()
        A = gs.AgentGroup(ctx, st)
        knn = _upload_knn_via_brute(gs, ctx, A, A, 7)
        fn = {"separation": gs.steerForSeparation, "alignment": gs.steerForAlignment,
              "cohesion": gs.steerForCohesion}[name]
        fn(A, knn, A, t.min_d, t.max_d, t.cos_max, t.weight, 0, mode)
        orc.steer_flock(w, ref, idx, ref, t.min_d, t.max_d, t.cos_max, t.weight, 0, mode)
        assert_state_equal(A.download(), ref, what=f"{name}: ")
        A.close()
        knn.close()
    ref = st.copy()
    A = gs.AgentGroup(ctx, st)
    knn = _upload_knn_via_brute(gs, ctx, A, A, 7)
    gs.steerForFlocking(A, knn, A, *terms, mode)
    for w, t in enumerate(terms):
        orc.steer_flock(w, ref, idx, ref, t.min_d, t.max_d, t.cos_max, t.weight, 0, mode)
    assert_state_equal(A.download(), ref, what="fused: ")
    A.close()
    knn.close()
