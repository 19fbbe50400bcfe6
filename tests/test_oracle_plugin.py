"""Pins the oracle's plugin-compat restatements (oracle/oracle.cpp:
update_anti_penetration, anti_penetration_agents, path_init, follow_path) by
what geometry fixes: an agent overlapping or crossing a wall, two
overlapping agents, paths whose lengths, directions and points are known,
and agents inside / outside a straight tube.  Without walls the
anti-penetration integrate must equal the pinned GPU-reference integrate.
Each case is chosen so that a dropped term, a flipped sign or a swapped
operand fails it."""
import math

import numpy as np
import pytest

from gpusteer_amd.state import WallState
from test_oracle_math import F64, one_agent

PAD = 0xFFFFFFFF


def walls(*segs, dtype=np.float64):
    """Walls from (start, end, normal) triples."""
    cols = {f: [] for f in ("sx", "sy", "sz", "ex", "ey", "ez", "nx", "ny", "nz")}
    for s, e, n in segs:
        for f, v in zip(("sx", "sy", "sz"), s):
            cols[f].append(v)
        for f, v in zip(("ex", "ey", "ez"), e):
            cols[f].append(v)
        for f, v in zip(("nx", "ny", "nz"), n):
            cols[f].append(v)
    return WallState({f: np.array(v, dtype) for f, v in cols.items()})


def pos(s, i=0):
    a = s.arrays
    return np.array([a["px"][i], a["py"][i], a["pz"][i]])


def fwd(s, i=0):
    a = s.arrays
    return np.array([a["fx"][i], a["fy"][i], a["fz"][i]])


WALL_X = ((-5.0, 0.0, 0.0), (5.0, 0.0, 0.0), (0.0, 0.0, 1.0))  # along x at z = 0, facing +z


# ---------------------------------------------------------------------------
# Integrate with wall anti-penetration.


def test_anti_penetration_update_without_walls_is_the_gpu_integrate(orc):
    """With only pads in the wall KNN the result is UpdateCUDAKernel's, except
    that the bits are kept."""
    g = np.random.default_rng(3)
    for trial in range(20):
        s = one_agent(pos=tuple(g.normal(size=3)), fwd=(0.6, 0.0, 0.8), side=(-0.8, 0, 0.6), speed=float(g.uniform(0, 1)),
                      max_speed=1.5, max_force=2.0, mass=1.3)
        s.arrays["gx"][0], s.arrays["gy"][0], s.arrays["gz"][0] = g.normal(size=3)
        s.bits[0] = 0x30
        a, b = s.copy(), s.copy()
        orc.update_anti_penetration(a, np.full((3, 1), PAD, np.uint32), walls(WALL_X), 0.1)
        orc.update(b, 0.1, orc.MODE_GPUREF)
        for f in ("px", "py", "pz", "fx", "fy", "fz", "speed", "sx", "sy", "sz", "ux", "uy", "uz", "gx", "gy", "gz"):
            assert a.arrays[f][0] == b.arrays[f][0], f
        assert a.bits[0] == 0x30 and b.bits[0] == 0


@pytest.mark.parametrize("closest", [0, 1])
def test_overlapping_wall_pushes_out_to_the_radius(orc, closest):
    """Agent at rest 0.3 above a wall it overlaps (radius 0.5): with the
    segment's closest point (the origin) it is moved along +normal by
    r - d = 0.2; the reference's closest point is the wall's end (5, 0, 0),
    25.09 away squared, so it stays."""
    s = one_agent(pos=(0.0, 0.0, 0.3), radius=0.5)
    orc.update_anti_penetration(s, np.zeros((1, 1), np.uint32), walls(WALL_X), 0.1, closest)
    if closest == orc.CLOSEST_CLAMPED:
        assert pos(s) == pytest.approx((0.0, 0.0, 0.5), abs=1e-15)
        assert s.bits[0] == 1 << 12
    else:
        assert pos(s).tolist() == [0.0, 0.0, 0.3] and s.bits[0] == 0


def test_reference_closest_point_is_the_end(orc):
    """Near the wall's end both rules see the same closest point: an agent at
    (5, 0, 0.2) overlapping the end is pushed by 0.3 in both."""
    for closest in (0, 1):
        s = one_agent(pos=(5.0, 0.0, 0.2), radius=0.5)
        orc.update_anti_penetration(s, np.zeros((1, 1), np.uint32), walls(WALL_X), 0.1, closest)
        assert pos(s) == pytest.approx((5.0, 0.0, 0.5), abs=1e-15)


def test_crossing_keeps_the_tangential_velocity(orc):
    """Agent at (0, 0, 1) moving along (0.6, 0, -0.8) at speed 2 with dt 1
    would cross the wall: the velocity loses its normal (z) part, leaving
    (1.2, 0, 0); forward becomes +x, side = normalize(forward x up)."""
    s = one_agent(pos=(0.0, 0.0, 1.0), fwd=(0.6, 0.0, -0.8), speed=2.0, max_speed=5.0, radius=0.25)
    orc.update_anti_penetration(s, np.zeros((1, 1), np.uint32), walls(WALL_X), 1.0)
    assert s.arrays["speed"][0] == pytest.approx(1.2, abs=1e-15)
    assert fwd(s) == pytest.approx((1.0, 0.0, 0.0), abs=1e-15)
    assert (s.arrays["sx"][0], s.arrays["sz"][0]) == pytest.approx((0.0, 1.0), abs=1e-15)  # (1,0,0) x (0,1,0)
    assert pos(s) == pytest.approx((1.2, 0.0, 1.0), abs=1e-15)
    assert s.bits[0] == 1 << 12


def test_nearest_crossing_and_last_overlap(orc):
    """Two parallel walls ahead (z = 0 and z = -0.5, normals +z and +x): the
    crossing taken is the nearer one (z = 0, normal +z, removing vz).  For the
    overlap, the last wall in KNN order within the radius is used, not the
    nearest: the agent 0.2 from wall A and 0.4 from wall B is pushed by B."""
    s = one_agent(pos=(0.0, 0.0, 1.0), fwd=(0.0, 0.0, -1.0), speed=3.0, max_speed=5.0, radius=0.1)
    W = walls(WALL_X, ((-5.0, 0.0, -0.5), (5.0, 0.0, -0.5), (1.0, 0.0, 0.0)))
    orc.update_anti_penetration(s, np.array([[1], [0]], np.uint32), W, 1.0)
    assert s.arrays["speed"][0] == 0.0 and pos(s).tolist() == [0.0, 0.0, 1.0]
    # overlap: A at z = 0 (normal +z) 0.2 away, B at x = 0.4 (normal -x) 0.4 away
    s = one_agent(pos=(0.0, 0.0, 0.2), radius=0.5)
    W = walls(WALL_X, ((0.4, 0.0, -5.0), (0.4, 0.0, 5.0), (-1.0, 0.0, 0.0)))
    orc.update_anti_penetration(s, np.array([[0], [1]], np.uint32), W, 0.1)
    assert pos(s) == pytest.approx((-0.1, 0.0, 0.2), abs=1e-15)
    s = one_agent(pos=(0.0, 0.0, 0.2), radius=0.5)
    orc.update_anti_penetration(s, np.array([[1], [0]], np.uint32), W, 0.1)
    assert pos(s) == pytest.approx((0.0, 0.0, 0.5), abs=1e-15)


# ---------------------------------------------------------------------------
# Agent anti-penetration.


def test_overlapping_pair_moves_apart_by_half_each(orc):
    """Two agents 0.6 apart on x with radius 0.5 (overlap 0.4): each moves
    0.2 away from the other, both read before either moves."""
    s = one_agent(pos=(0.0, 0.0, 0.0), extra=[(0.6, 0.0, 0.0)])
    idx = np.array([[1, 0]], np.uint32)
    dist = np.array([[0.6, 0.6]], np.float32)
    orc.anti_penetration_agents(s, idx, dist, s)
    assert pos(s, 0) == pytest.approx((-0.2, 0.0, 0.0), abs=1e-7)
    assert pos(s, 1) == pytest.approx((0.8, 0.0, 0.0), abs=1e-7)
    assert s.bits[0] == s.bits[1] == 1 << 13


def test_masked_other_takes_the_whole_overlap(orc):
    """B's agent has a masked bit: A's agent moves by the whole overlap; a
    masked A agent stays; a pad or a non-overlap leaves it; the KNN's
    distance is used, not the current one."""
    a = one_agent(pos=(0.0, 0.0, 0.0))
    b = one_agent(pos=(0.0, 0.0, 0.6))
    b.bits[0] = 1 << 12
    orc.anti_penetration_agents(a, np.array([[0]], np.uint32), np.array([[0.6]], np.float32), b, mask=1 << 12)
    assert pos(a) == pytest.approx((0.0, 0.0, -0.4), abs=1e-7)
    a = one_agent(pos=(0.0, 0.0, 0.0))
    a.bits[0] = 1 << 12
    orc.anti_penetration_agents(a, np.array([[0]], np.uint32), np.array([[0.6]], np.float32), b, mask=1 << 12)
    assert pos(a).tolist() == [0.0, 0.0, 0.0]
    a = one_agent(pos=(0.0, 0.0, 0.0))
    orc.anti_penetration_agents(a, np.array([[PAD]], np.uint32), np.array([[0.6]], np.float32), b)
    assert pos(a).tolist() == [0.0, 0.0, 0.0] and a.bits[0] == 0
    a = one_agent(pos=(0.0, 0.0, 0.0))
    orc.anti_penetration_agents(a, np.array([[0]], np.uint32), np.array([[1.5]], np.float32), b)
    assert pos(a).tolist() == [0.0, 0.0, 0.0] and a.bits[0] == 0
    # stale distance 0.9 (overlap 0.1) with the agents actually 0.6 apart
    a = one_agent(pos=(0.0, 0.0, 0.0))
    orc.anti_penetration_agents(a, np.array([[0]], np.uint32), np.array([[0.9]], np.float32), b)
    assert pos(a) == pytest.approx((0.0, 0.0, -0.05), abs=1e-7)


# ---------------------------------------------------------------------------
# Paths.


def test_path_init_closed_forms(orc):
    """A 3-4-5 triangle: segment lengths 3, 4 and (closing, cyclic) 5; unit
    directions; total 12 (7 when open); entry 0 zero."""
    pts = [(0.0, 0.0, 0.0), (3.0, 0.0, 0.0), (3.0, 0.0, 4.0)]
    p = orc.Pathway(pts, 1.0, cyclic=True, dtype=F64)
    assert p.m == 4 and p.total == 12.0
    assert p.lengths.tolist() == [0.0, 3.0, 4.0, 5.0]
    assert p.points[3].tolist() == [0.0, 0.0, 0.0]
    np.testing.assert_allclose(p.normals, [[0, 0, 0], [1, 0, 0], [0, 0, 1], [-0.6, 0, -0.8]], atol=1e-15)
    q = orc.Pathway(pts, 1.0, cyclic=False, dtype=F64)
    assert q.m == 3 and q.total == 7.0


def test_path_point_at(orc):
    p = orc.Pathway([(0.0, 0.0, 0.0), (3.0, 0.0, 0.0), (3.0, 0.0, 4.0)], 1.0, cyclic=False, dtype=F64)
    assert p.point_at(-1.0).tolist() == [0.0, 0.0, 0.0]
    assert p.point_at(7.0).tolist() == [3.0, 0.0, 4.0]
    assert p.point_at(1.5).tolist() == [1.5, 0.0, 0.0]
    assert p.point_at(5.0) == pytest.approx((3.0, 0.0, 2.0), abs=1e-15)
    c = orc.Pathway([(0.0, 0.0, 0.0), (3.0, 0.0, 0.0), (3.0, 0.0, 4.0)], 1.0, cyclic=True, dtype=F64)
    # 12 + 9.5 wraps to 9.5: 2.5 along the closing segment from (3, 0, 4)
    assert c.point_at(21.5) == pytest.approx((1.5, 0.0, 2.0), abs=1e-12)
    # a negative distance wraps to a negative remainder: extrapolated before the first point
    assert c.point_at(-1.0) == pytest.approx((-1.0, 0.0, 0.0), abs=1e-15)


def test_path_distance_of(orc):
    p = orc.Pathway([(0.0, 0.0, 0.0), (3.0, 0.0, 0.0), (3.0, 0.0, 4.0)], 1.0, cyclic=False, dtype=F64)
    assert p.distance_of((1.0, 0.0, -0.5)) == 1.0
    assert p.distance_of((4.0, 0.0, 1.0)) == 4.0
    assert p.distance_of((-2.0, 0.0, 0.0)) == 0.0     # before the start: projection clamped to 0
    assert p.distance_of((3.5, 0.0, 9.0)) == 7.0      # beyond the end: the last segment's length


def _path_agent(p, f, speed):
    return one_agent(pos=p, fwd=f, side=(-f[2], 0.0, f[0]), speed=speed, max_speed=10.0)


def test_follow_path_inside_the_tube_and_moving_forward_is_free(orc):
    P = orc.Pathway([(0.0, 0.0, 0.0), (0.0, 0.0, 100.0)], 2.0, cyclic=False, dtype=F64)
    s = _path_agent((0.5, 0.0, 10.0), (0.0, 0.0, 1.0), 2.0)
    orc.follow_path(s, P, 3.0, 1.0)
    assert s.arrays["gz"][0] == 0.0 and s.arrays["gx"][0] == 0.0 and s.bits[0] == 0


def test_follow_path_outside_the_tube_seeks_ahead(orc):
    """Agent at (5, 0, 10) moving +z at speed 2, prediction 3 s: the future
    position (5, 0, 16) is 3 outside the radius-2 tube, so it seeks the path
    point 6 ahead of its own path distance 10: target (0, 0, 16);
    steering = normalize((0, 0, 16) - (5, 0, 10) - (0, 0, 1)) * w."""
    P = orc.Pathway([(0.0, 0.0, 0.0), (0.0, 0.0, 100.0)], 2.0, cyclic=False, dtype=F64)
    s = _path_agent((5.0, 0.0, 10.0), (0.0, 0.0, 1.0), 2.0)
    orc.follow_path(s, P, 3.0, 2.0)
    want = np.array([-5.0, 0.0, 5.0]) / math.sqrt(50.0) * 2.0
    np.testing.assert_allclose([s.arrays[f][0] for f in ("gx", "gy", "gz")], want, atol=1e-15)
    assert s.bits[0] == 1 << 11


def test_follow_path_moving_backward_turns(orc):
    """Inside the tube but moving -z (path distance decreasing with a
    positive offset): seeks the point 6 ahead, (0, 0, 16), from (0.5, 0, 10)
    with forward (0, 0, -1)."""
    P = orc.Pathway([(0.0, 0.0, 0.0), (0.0, 0.0, 100.0)], 2.0, cyclic=False, dtype=F64)
    s = _path_agent((0.5, 0.0, 10.0), (0.0, 0.0, -1.0), 2.0)
    orc.follow_path(s, P, 3.0, 1.0)
    d = np.array([-0.5, 0.0, 7.0])
    np.testing.assert_allclose([s.arrays[f][0] for f in ("gx", "gy", "gz")], d / np.linalg.norm(d), atol=1e-15)


def test_follow_path_masked_and_accumulating(orc):
    P = orc.Pathway([(0.0, 0.0, 0.0), (0.0, 0.0, 100.0)], 2.0, cyclic=False, dtype=F64)
    s = _path_agent((5.0, 0.0, 10.0), (0.0, 0.0, 1.0), 2.0)
    s.bits[0] = 1 << 8
    orc.follow_path(s, P, 3.0, 1.0, mask=1 << 8)
    assert s.arrays["gx"][0] == 0.0
    s = _path_agent((5.0, 0.0, 10.0), (0.0, 0.0, 1.0), 2.0)
    s.arrays["gx"][0] = 1.0
    orc.follow_path(s, P, 3.0, 1.0)
    assert s.arrays["gx"][0] == pytest.approx(1.0 - 5.0 / math.sqrt(50.0), abs=1e-15)


# ---------------------------------------------------------------------------
# Neighbour avoidance.


def _pair(a_pos, a_fwd, a_speed, others):
    """Agent 0 (side (-1, 0, 0) for forward +z) and neighbours given as
    (position, forward, speed); KNN rows list the neighbours in order."""
    s = one_agent(pos=a_pos, fwd=a_fwd, side=(-1.0, 0.0, 0.0), speed=a_speed, extra=[o[0] for o in others])
    for i, (_, f, v) in enumerate(others, start=1):
        s.arrays["fx"][i], s.arrays["fy"][i], s.arrays["fz"][i] = f
        s.arrays["speed"][i] = v
    idx = knn_rows_f(len(others))
    pos = np.stack([s.arrays["px"], s.arrays["py"], s.arrays["pz"]], axis=1)
    dist = np.full(idx.shape, np.finfo(np.float32).max, np.float32)
    for j in range(len(others)):
        dist[j, 0] = np.linalg.norm(pos[j + 1] - pos[0])
    return s, idx, dist


def knn_rows_f(n_nb):
    out = np.full((n_nb, 1 + n_nb), PAD, np.uint32)
    out[:, 0] = np.arange(1, n_nb + 1)
    return out


def _steer(s):
    return [s.arrays[f][0] for f in ("gx", "gy", "gz")]


BIT_AN, BIT_CLOSE = 1 << 9, 1 << 10


@pytest.mark.parametrize("mode", [0, 1])
def test_avoid_head_on(orc, mode):
    """Head-on: the other agent 0.3 to the +x side, 10 ahead, both at speed
    1: nearest approach at t = 5, 0.3 apart (< 2 r): steer along +side, away
    from +x: (-1, 0, 0) x w."""
    s, idx, dist = _pair((0.0, 0.0, 0.0), (0.0, 0.0, 1.0), 1.0, [((0.3, 0.0, 10.0), (0.0, 0.0, -1.0), 1.0)])
    orc.avoid_neighbors(s, idx, dist, s, 10.0, 0.5, False, 2.0, mode=mode)
    assert _steer(s) == pytest.approx([-2.0, 0.0, 0.0], abs=1e-15)
    assert s.bits[0] == BIT_AN


@pytest.mark.parametrize("mode", [0, 1])
def test_avoid_parallel_overtaking(orc, mode):
    """Parallel: a faster agent behind, 0.3 to +x, catching up at t = 5:
    steer away from its current position (+x): (-1, 0, 0)."""
    s, idx, dist = _pair((0.0, 0.0, 0.0), (0.0, 0.0, 1.0), 1.0, [((0.3, 0.0, -5.0), (0.0, 0.0, 1.0), 2.0)])
    orc.avoid_neighbors(s, idx, dist, s, 10.0, 0.5, False, 1.0, mode=mode)
    assert _steer(s) == pytest.approx([-1.0, 0.0, 0.0], abs=1e-15)


def test_avoid_perpendicular_slower_steers_behind(orc):
    """Crossing paths meeting at (0, 0, 5) at t = 5 (the other moves -x at
    speed 1): the GPU rule lets the agent not faster than the threat steer
    behind it, sign of side . threat velocity = +1 -> steer -1 -> (1, 0, 0);
    the CPU rule lets the agent steer when the threat is not faster (equal
    speeds: both rules steer).  With the agent at speed 2 (meeting arranged at
    t = 2.5) only the CPU rule steers; the GPU rule adds zero with the bit
    unset."""
    s, idx, dist = _pair((0.0, 0.0, 0.0), (0.0, 0.0, 1.0), 1.0, [((5.0, 0.0, 5.0), (-1.0, 0.0, 0.0), 1.0)])
    orc.avoid_neighbors(s, idx, dist, s, 10.0, 0.5, False, 1.0, mode=1)
    assert _steer(s) == pytest.approx([1.0, 0.0, 0.0], abs=1e-15)
    for mode, want, bit in ((1, [0.0, 0.0, 0.0], 0), (0, [1.0, 0.0, 0.0], BIT_AN)):
        s, idx, dist = _pair((0.0, 0.0, 0.0), (0.0, 0.0, 1.0), 2.0, [((2.5, 0.0, 5.0), (-1.0, 0.0, 0.0), 1.0)])
        orc.avoid_neighbors(s, idx, dist, s, 10.0, 0.5, False, 1.0, mode=mode)
        assert _steer(s) == pytest.approx(want, abs=1e-15), mode
        assert s.bits[0] == bit


def test_avoid_no_threat_adds_nothing(orc):
    """Moving apart (nearest approach in the past) or meeting after
    min_time: no threat, steering unchanged, no bit."""
    for other in (((0.3, 0.0, -10.0), (0.0, 0.0, -1.0), 1.0), ((0.3, 0.0, 30.0), (0.0, 0.0, -1.0), 1.0)):
        s, idx, dist = _pair((0.0, 0.0, 0.0), (0.0, 0.0, 1.0), 1.0, [other])
        s.arrays["gx"][0] = 0.25
        orc.avoid_neighbors(s, idx, dist, s, 10.0, 0.5, False, 1.0, mode=1)
        assert _steer(s) == [0.25, 0.0, 0.0] and s.bits[0] == 0


def test_avoid_head_on_uses_the_last_approach_position(orc):
    """GPU rule: the head-on offset is the approach position the kernel's
    helper wrote last.  Threat T1 (0.2 to +x, head-on, t = 3) is chosen; T2
    (5 to -x, head-on, t = 2 < 3, never closer than 5) is evaluated after it
    and overwrites the approach position with (-5, 0, 2): the agent steers
    toward +x.  The CPU rule keeps T1's approach point and steers to -x."""
    others = [((0.2, 0.0, 6.0), (0.0, 0.0, -1.0), 1.0), ((-5.0, 0.0, 4.0), (0.0, 0.0, -1.0), 1.0)]
    s, idx, dist = _pair((0.0, 0.0, 0.0), (0.0, 0.0, 1.0), 1.0, others)
    orc.avoid_neighbors(s, idx, dist, s, 10.0, 0.5, False, 1.0, mode=1)
    assert _steer(s) == pytest.approx([1.0, 0.0, 0.0], abs=1e-15)
    s, idx, dist = _pair((0.0, 0.0, 0.0), (0.0, 0.0, 1.0), 1.0, others)
    orc.avoid_neighbors(s, idx, dist, s, 10.0, 0.5, False, 1.0, mode=0)
    assert _steer(s) == pytest.approx([-1.0, 0.0, 0.0], abs=1e-15)


def test_avoid_close_neighbours(orc):
    """A neighbour at (0.8, 0, 0.3) (KNN distance 0.854 < 0.5 + 1): steer
    along the component of -offset perpendicular to forward, (-0.8, 0, 0),
    normalized and weighted (GPU), with the close bit; the collision pass is
    skipped.  Without avoid_close the GPU rule only looks for collisions.
    The CPU rule always checks close neighbours (min separation 0: 0.854 <
    1) and adds the unnormalized (-0.8, 0, 0)."""
    other = [((0.8, 0.0, 0.3), (0.0, 0.0, -1.0), 1.0)]
    s, idx, dist = _pair((0.0, 0.0, 0.0), (0.0, 0.0, 1.0), 1.0, other)
    orc.avoid_neighbors(s, idx, dist, s, 10.0, 0.5, True, 3.0, mode=1)
    assert _steer(s) == pytest.approx([-3.0, 0.0, 0.0], abs=1e-15) and s.bits[0] == BIT_CLOSE
    s, idx, dist = _pair((0.0, 0.0, 0.0), (0.0, 0.0, 1.0), 1.0, other)
    orc.avoid_neighbors(s, idx, dist, s, 10.0, 0.5, False, 3.0, mode=1)
    assert s.bits[0] == BIT_AN  # head-on collision pass instead
    s, idx, dist = _pair((0.0, 0.0, 0.0), (0.0, 0.0, 1.0), 1.0, other)
    orc.avoid_neighbors(s, idx, dist, s, 10.0, 0.5, False, 3.0, mode=0)
    assert _steer(s) == pytest.approx([-2.4, 0.0, 0.0], abs=1e-12) and s.bits[0] == BIT_CLOSE
    # masked: nothing
    s, idx, dist = _pair((0.0, 0.0, 0.0), (0.0, 0.0, 1.0), 1.0, other)
    s.bits[0] = 1 << 4
    orc.avoid_neighbors(s, idx, dist, s, 10.0, 0.5, True, 3.0, mask=1 << 4, mode=1)
    assert _steer(s) == [0.0, 0.0, 0.0] and s.bits[0] == 1 << 4
