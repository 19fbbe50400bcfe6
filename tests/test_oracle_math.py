"""Pins the oracle's force, integration and wrap restatements by what the
mathematics fixes: constructed single-agent cases whose answers follow from
geometry (a single neighbour, symmetric neighbours, a ray against a sphere, a
feeler against a wall), closed-form integration in float64, std::pow for the
x^20 restatement, and exact antipode / face mappings.  Each case is chosen so
that a dropped term, a flipped sign or a swapped operand in the oracle fails
it."""
import math

import numpy as np
import pytest

from gpusteer_amd import _capi as K
from gpusteer_amd.state import AgentState, WallState

F64 = np.float64
# The integrator's literal float constants (0.2f, 0.15f, 0.4f of
# src/SimpleVehicle.cpp:115,185) carried into float64 arithmetic.
F32_02, F32_015, F32_04 = (float(np.float32(v)) for v in (0.2, 0.15, 0.4))


def one_agent(pos=(0, 0, 0), fwd=(0, 0, 1), side=(-1, 0, 0), up=(0, 1, 0), speed=0.0, radius=0.5, max_speed=1.0,
              max_force=1.0, mass=1.0, dtype=F64, extra=()):
    """Agent 0 as given, followed by `extra` neighbour positions (unit forward +z)."""
    n = 1 + len(extra)
    s = AgentState.empty(n, dtype)
    a = s.arrays
    for i in range(n):
        p = pos if i == 0 else extra[i - 1]
        a["px"][i], a["py"][i], a["pz"][i] = p
        f = fwd if i == 0 else (0, 0, 1)
        a["fx"][i], a["fy"][i], a["fz"][i] = f
        a["sx"][i], a["sy"][i], a["sz"][i] = side
        a["ux"][i], a["uy"][i], a["uz"][i] = up
    a["speed"][0] = speed
    a["radius"][:] = radius
    a["max_speed"][:] = max_speed
    a["max_force"][:] = max_force
    a["mass"][:] = mass
    return s


def knn_rows(n_nb):
    """KNN [k][n] (n = 1 + n_nb agents): agent 0 lists agents 1..n_nb, every
    other agent's row is padding."""
    out = np.full((n_nb, 1 + n_nb), 0xFFFFFFFF, np.uint32)
    out[:, 0] = np.arange(1, n_nb + 1)
    return out


def steering(s):
    return np.array([s.arrays["gx"][0], s.arrays["gy"][0], s.arrays["gz"][0]])


# ---------------------------------------------------------------------------
# Vector utilities.


def test_truncate_length(orc):
    out = np.zeros(3, np.float32)
    orc.lib().or_truncate_length(orc._f(np.array([3, 4, 0], np.float32)), 2.5, orc._f(out))
    np.testing.assert_allclose(out, [1.5, 2.0, 0.0], rtol=1e-7)
    orc.lib().or_truncate_length(orc._f(np.array([0.3, 0.4, 0], np.float32)), 2.5, orc._f(out))
    np.testing.assert_array_equal(out, np.array([0.3, 0.4, 0], np.float32))


@pytest.mark.parametrize("cosc", [0.5, 0.0, -0.5, 0.9])
def test_limit_max_deviation_angle_cone(orc, cosc):
    """Outside the cone: the result keeps the source's length, lies on the
    cone's surface and in the plane of source and basis.  Inside: unchanged."""
    basis = np.array([0, 0, 1], np.float32)
    src = np.array([2.0, 0.5, -1.0], np.float32)  # cos = -1/sqrt(5.25) = -0.436
    out = np.zeros(3, np.float32)
    orc.lib().or_limit_max_deviation_angle(orc._f(src), cosc, orc._f(basis), orc._f(out))
    cos_src = float(src @ basis) / np.linalg.norm(src)
    if cos_src >= cosc:
        np.testing.assert_array_equal(out, src)
        return
    assert np.linalg.norm(out) == pytest.approx(np.linalg.norm(src), rel=1e-6)
    assert float(out @ basis) / np.linalg.norm(out) == pytest.approx(cosc, abs=1e-6)
    # Coplanar with source and basis; on the source's side of the basis.
    assert abs(np.dot(np.cross(src, basis), out)) < 1e-5
    assert np.dot(out[:2], src[:2]) > 0


def test_pow20_matches_std_pow(orc):
    """The repeated-squaring restatement of pow(range, 20) agrees with the
    correctly rounded float of the exact power (itself within an ulp of
    std::pow's) for every range the integrator can produce (0..1]."""
    r = np.concatenate([np.linspace(0, 1, 20001, dtype=np.float32),
                        np.random.default_rng(0).uniform(0, 1, 20000).astype(np.float32)])
    ours = np.array([orc.lib().or_pow20(float(x)) for x in r], np.float32)
    exact = np.array([float(math.pow(float(x), 20)) for x in r], np.float64).astype(np.float32)
    ulp = np.abs(ours.view(np.int32).astype(np.int64) - exact.view(np.int32).astype(np.int64))
    assert ulp.max() <= 1
    assert (ulp == 0).mean() > 0.999
    assert orc.lib().or_pow20(1.0) == 1.0 and orc.lib().or_pow20(0.0) == 0.0
    assert orc.lib().or_pow20(0.5) == np.float32(0.5 ** 20)


def _li(orc, sa, ea, sb, eb):
    ip = np.zeros(3, np.float32)
    args = [orc._f(np.array(v, np.float32)) for v in (sa, ea, sb, eb)]
    r = orc.lib().or_lines_intersect(*args, orc._f(ip))
    return r, ip


def test_lines_intersect(orc):
    r, ip = _li(orc, (0, 7, 0), (2, 7, 2), (0, 0, 2), (2, 0, 0))  # an X in the XZ plane; y ignored
    assert r == 1
    np.testing.assert_allclose(ip, [1, 0, 1], atol=1e-7)
    assert _li(orc, (0, 0, 0), (1, 0, 0), (0, 0, 1), (1, 0, 1))[0] == 0  # parallel
    r, ip = _li(orc, (0, 0, 0), (2, 0, 0), (1, 0, 0), (3, 0, 0))  # collinear: coincident -> start of A
    assert r == 1 and np.array_equal(ip, [0, 0, 0])
    assert _li(orc, (0, 0, 0), (1, 0, 1), (3, 0, 0), (0, 0, 3))[0] == 0  # lines cross beyond A's end
    r, ip = _li(orc, (0, 0, 0), (4, 0, 0), (4, 0, -1), (4, 0, 1))  # touching at A's end point (ua = 1)
    assert r == 1 and np.allclose(ip, [4, 0, 0])


# ---------------------------------------------------------------------------
# Flocking forces.


@pytest.mark.parametrize("mode", [K.GS_MODE_OPENSTEER, K.GS_MODE_GPUREF])
def test_separation_single_neighbour(orc, mode):
    # Neighbour at +x distance 2: separation points away from it, unit length.
    s = one_agent(extra=[(2.0, 0.0, 0.0)])
    orc.steer_flock(orc.FLOCK_SEPARATION, s, knn_rows(1), s, 1.0, 10.0, -1.0, 3.0, 0, mode)
    np.testing.assert_allclose(steering(s), [-3.0, 0, 0])
    assert s.bits[0] == K.BIT_SEPARATION


def test_separation_inverse_square_weighting(orc):
    # Neighbours at +x (d=1) and -x (d=2): each contributes -offset/d^2, i.e.
    # (-1,0,0) and (+0.5,0,0); the normalized mean points away from the nearer.
    s = one_agent(extra=[(1.0, 0, 0), (-2.0, 0, 0)])
    orc.steer_flock(orc.FLOCK_SEPARATION, s, knn_rows(2), s, 5.0, 10.0, -1.0, 1.0)
    np.testing.assert_allclose(steering(s), [-1.0, 0, 0])
    # Symmetric neighbours cancel: zero force, bit not set.
    s = one_agent(extra=[(1.0, 0, 0), (-1.0, 0, 0)])
    orc.steer_flock(orc.FLOCK_SEPARATION, s, knn_rows(2), s, 5.0, 10.0, -1.0, 1.0)
    np.testing.assert_array_equal(steering(s), 0)
    assert s.bits[0] == 0


def test_alignment(orc):
    # Same heading as every neighbour -> zero; neighbour heading +x -> direction
    # normalize(mean_fwd - fwd) = normalize((1,0,-1)).
    s = one_agent(extra=[(1.0, 0, 0), (0, 1.0, 0)])
    orc.steer_flock(orc.FLOCK_ALIGNMENT, s, knn_rows(2), s, 5.0, 10.0, -1.0, 1.0)
    np.testing.assert_array_equal(steering(s), 0)
    s = one_agent(extra=[(1.0, 0, 0)])
    s.arrays["fx"][1], s.arrays["fz"][1] = 1.0, 0.0
    orc.steer_flock(orc.FLOCK_ALIGNMENT, s, knn_rows(1), s, 5.0, 10.0, -1.0, 2.0)
    np.testing.assert_allclose(steering(s), [math.sqrt(2), 0, -math.sqrt(2)], rtol=1e-12)
    assert s.bits[0] == K.BIT_ALIGNMENT


def test_cohesion(orc):
    s = one_agent(extra=[(3.0, 0, 0), (-3.0, 0, 0)])
    orc.steer_flock(orc.FLOCK_COHESION, s, knn_rows(2), s, 5.0, 10.0, -1.0, 1.0)
    np.testing.assert_array_equal(steering(s), 0)
    s = one_agent(extra=[(3.0, 4.0, 0), (3.0, -4.0, 0), (0, 0, 6.0)])
    orc.steer_flock(orc.FLOCK_COHESION, s, knn_rows(3), s, 10.0, 10.0, -1.0, 1.0)
    np.testing.assert_allclose(steering(s), np.array([2, 0, 2]) / math.sqrt(8), rtol=1e-12)
    assert s.bits[0] == K.BIT_COHESION


def test_neighbourhood_rules(orc):
    """A neighbour behind the agent counts only inside the min distance.
    OPENSTEER with min_d < 0 uses 3 * radius; GPUREF uses min_d as given."""
    behind = [(0, 0, -1.2)]  # radius 0.5 -> 3r = 1.5 > 1.2
    for mode, min_d, counted in [(K.GS_MODE_OPENSTEER, -1.0, True), (K.GS_MODE_OPENSTEER, 1.0, False),
                                 (K.GS_MODE_GPUREF, 1.0, False), (K.GS_MODE_GPUREF, 1.3, True)]:
        s = one_agent(extra=behind)
        orc.steer_flock(orc.FLOCK_COHESION, s, knn_rows(1), s, min_d, 10.0, 0.0, 1.0, 0, mode)
        assert (s.bits[0] != 0) == counted, (mode, min_d)
    # In front, beyond min distance but inside max distance and the cone.
    s = one_agent(extra=[(0, 0, 4.0)])
    orc.steer_flock(orc.FLOCK_COHESION, s, knn_rows(1), s, 1.0, 5.0, 0.5, 1.0)
    np.testing.assert_allclose(steering(s), [0, 0, 1])
    # Beyond max distance.
    s = one_agent(extra=[(0, 0, 6.0)])
    orc.steer_flock(orc.FLOCK_COHESION, s, knn_rows(1), s, 1.0, 5.0, 0.5, 1.0)
    assert s.bits[0] == 0
    # Padding (index >= n) stops the scan: a real neighbour after it is ignored.
    s = one_agent(extra=[(0, 0, 4.0)])
    knn = np.array([[0xFFFFFFFF, 0xFFFFFFFF], [1, 0xFFFFFFFF]], np.uint32)
    orc.steer_flock(orc.FLOCK_COHESION, s, knn, s, 1.0, 5.0, 0.5, 1.0)
    assert s.bits[0] == 0


def test_mask_skips_and_accumulates(orc):
    s = one_agent(extra=[(2.0, 0, 0)])
    s.arrays["gx"][0] = 0.25
    orc.steer_flock(orc.FLOCK_SEPARATION, s, knn_rows(1), s, 5.0, 10.0, -1.0, 1.0)
    np.testing.assert_allclose(steering(s), [-0.75, 0, 0])  # (f*w) + previous
    s.bits[0] = K.BIT_SEEK
    orc.steer_flock(orc.FLOCK_SEPARATION, s, knn_rows(1), s, 5.0, 10.0, -1.0, 1.0, K.BIT_SEEK)
    np.testing.assert_allclose(steering(s), [-0.75, 0, 0])  # masked: unchanged


@pytest.mark.parametrize("mode", [K.GS_MODE_OPENSTEER, K.GS_MODE_GPUREF])
def test_seek(orc, mode):
    s = one_agent(pos=(1, 0, 0), speed=2.0)
    orc.steer_seek(s, (1, 0, 10), 1.0, 0, mode)
    if mode == K.GS_MODE_OPENSTEER:
        np.testing.assert_allclose(steering(s), [0, 0, 8])  # desired (0,0,10) minus velocity (0,0,2)
    else:
        np.testing.assert_allclose(steering(s), [0, 0, 1])  # normalize(desired - forward)
    assert s.bits[0] == K.BIT_SEEK


# ---------------------------------------------------------------------------
# Obstacles and walls.


def _obstacles(c, r):
    return [np.array([c[0]]), np.array([c[1]]), np.array([c[2]]), np.array([r])]


@pytest.mark.parametrize("center,min_time,expect", [
    ((1.0, 0.0, 10.0), 2.0, [-1.0, 0, 0]),   # ray hits the inflated sphere at z = 10 - sqrt(9 - 1)
    ((1.0, 0.0, 10.0), 1.0, [0, 0, 0]),      # hit beyond min_time * speed
    ((3.5, 0.0, 10.0), 2.0, [0, 0, 0]),      # lateral 3.5 > radius sum 3: miss
    ((1.0, 0.0, -10.0), 2.0, [0, 0, 0]),     # behind: both roots negative
    ((0.0, -2.0, 1.0), 2.0, [0, 2.0, 0]),    # agent inside the inflated sphere: one root positive
])
def test_avoid_obstacle_sphere(orc, center, min_time, expect):
    s = one_agent(speed=5.0, radius=1.0)
    orc.avoid_obstacles(s, np.zeros((1, 1), np.uint32), _obstacles(center, 2.0), min_time, 1.0)
    np.testing.assert_allclose(steering(s), expect, atol=1e-12)
    assert (s.bits[0] == K.BIT_AVOID_OBSTACLES) == bool(np.any(expect))


def _wall(sx, sz, ex, ez, nx, nz):
    ar = {k: np.zeros(1) for k in ("sx", "sy", "sz", "ex", "ey", "ez", "nx", "ny", "nz")}
    ar["sx"][0], ar["sz"][0], ar["ex"][0], ar["ez"][0], ar["nx"][0], ar["nz"][0] = sx, sz, ex, ez, nx, nz
    return WallState(ar)


def test_avoid_wall_feeler(orc):
    # Heading +x at speed 2 with min_time 5: the forward feeler reaches x = 10,
    # crossing the wall x = 5 at distance 5 (side feelers reach only x = 4).
    s = one_agent(fwd=(1, 0, 0), side=(0, 0, 1), speed=2.0, radius=1.0)
    orc.avoid_walls(s, np.zeros((1, 1), np.uint32), _wall(5, -10, 5, 10, -1, 0), 5.0, 0.5)
    np.testing.assert_allclose(steering(s), [-2.5, 0, 0])  # normal * |feeler tip - hit| * w
    assert s.bits[0] == K.BIT_AVOID_WALLS
    # Far wall: nothing.
    s = one_agent(fwd=(1, 0, 0), side=(0, 0, 1), speed=2.0, radius=1.0)
    orc.avoid_walls(s, np.zeros((1, 1), np.uint32), _wall(50, -10, 50, 10, -1, 0), 5.0, 0.5)
    assert s.bits[0] == 0 and not steering(s).any()


def test_avoid_wall_touching_pushes_out(orc):
    # Radius 1 agent at x = 4.5, wall at x = 5 facing -x: penetration 0.5, the
    # agent moves back to x = 4 and its forward loses the normal component.
    r2 = 1 / math.sqrt(2)
    s = one_agent(pos=(4.5, 0, 0), fwd=(r2, 0, r2), side=(r2, 0, -r2), speed=1.0, radius=1.0)
    orc.avoid_walls(s, np.zeros((1, 1), np.uint32), _wall(5, -10, 5, 10, -1, 0), 5.0, 1.0)
    a = s.arrays
    np.testing.assert_allclose([a["px"][0], a["py"][0], a["pz"][0]], [4.0, 0, 0], atol=1e-12)
    np.testing.assert_allclose([a["fx"][0], a["fy"][0], a["fz"][0]], [0, 0, 1], atol=1e-12)
    assert s.bits[0] == K.BIT_ANTI_PENETRATION_WALL
    assert not steering(s).any()


# ---------------------------------------------------------------------------
# Integration.


def test_update_gpuref_closed_form(orc):
    s = one_agent(fwd=(0, 0, 1), side=(-1, 0, 0), speed=0.5, max_speed=2.0, max_force=3.0, mass=2.0)
    s.arrays["gx"][0], s.arrays["gz"][0] = 4.0, 0.0  # |force| 4 > max_force 3 -> (3,0,0)
    orc.update(s, 0.1, K.GS_MODE_GPUREF)
    v = np.array([0, 0, 0.5]) + np.array([3.0, 0, 0]) / 2.0 * 0.1
    a = s.arrays
    np.testing.assert_allclose([a["px"][0], a["py"][0], a["pz"][0]], v * 0.1, rtol=1e-14)
    assert a["speed"][0] == pytest.approx(np.linalg.norm(v), rel=1e-14)
    f = v / np.linalg.norm(v)
    np.testing.assert_allclose([a["fx"][0], a["fy"][0], a["fz"][0]], f, rtol=1e-14)
    side = np.cross(f, [0, 1, 0])
    np.testing.assert_allclose([a["sx"][0], a["sy"][0], a["sz"][0]], side / np.linalg.norm(side), rtol=1e-14)
    assert not steering(s).any() and s.bits[0] == 0


def test_update_opensteer_smoothing_closed_form(orc):
    # speed 1.0 > 0.2 * max_speed: no cone adjustment.  rate = clip(9 dt, .15, .4).
    dt = 0.02
    s = one_agent(fwd=(0, 0, 1), speed=1.0, max_speed=2.0, max_force=10.0, mass=2.0)
    s.arrays["ax"][0] = 1.0
    s.arrays["gy"][0] = 4.0
    orc.update(s, dt, K.GS_MODE_OPENSTEER)
    rate = min(max(9 * dt, F32_015), F32_04)
    acc0, acc_new = np.array([1.0, 0, 0]), np.array([0, 2.0, 0])
    sacc = acc0 + (acc_new - acc0) * rate
    v = np.array([0, 0, 1.0]) + sacc * dt
    a = s.arrays
    np.testing.assert_allclose([a["ax"][0], a["ay"][0], a["az"][0]], sacc, rtol=1e-14)
    np.testing.assert_allclose([a["px"][0], a["py"][0], a["pz"][0]], v * dt, rtol=1e-14)
    f = v / np.linalg.norm(v)
    np.testing.assert_allclose([a["fx"][0], a["fy"][0], a["fz"][0]], f, rtol=1e-14)
    # side = forward x up (un-normalized), up = side x forward.
    side = np.cross(f, [0, 1.0, 0])
    np.testing.assert_allclose([a["sx"][0], a["sy"][0], a["sz"][0]], side, rtol=1e-14)
    np.testing.assert_allclose([a["ux"][0], a["uy"][0], a["uz"][0]], np.cross(side, f), rtol=1e-14)


def test_update_opensteer_cone_at_rest(orc):
    """At speed 0 the deviation cone is closed (cos = 1 - 2*0^20 = 1): a force
    perpendicular to forward is turned onto forward, keeping its length."""
    dt = 0.05
    s = one_agent(fwd=(0, 0, 1), speed=0.0, max_speed=1.0, max_force=10.0)
    s.arrays["gx"][0] = 2.0
    orc.update(s, dt, K.GS_MODE_OPENSTEER)
    rate = min(max(9 * dt, F32_015), F32_04)
    np.testing.assert_allclose([s.arrays["ax"][0], s.arrays["ay"][0], s.arrays["az"][0]], [0, 0, 2.0 * rate],
                               atol=1e-15)
    # Speed 0.1 (range 0.5): cos = 1 - 2 * 0.5^20, the cone is almost closed.
    s = one_agent(fwd=(0, 0, 1), speed=0.1, max_speed=1.0, max_force=10.0)
    s.arrays["gx"][0] = 2.0
    orc.update(s, dt, K.GS_MODE_OPENSTEER)
    acc = np.array([s.arrays["ax"][0], s.arrays["ay"][0], s.arrays["az"][0]]) / rate
    cos = 1 - 2 * (0.1 / F32_02) ** 20
    assert acc[2] / np.linalg.norm(acc) == pytest.approx(cos, abs=1e-12)
    assert np.linalg.norm(acc) == pytest.approx(2.0, rel=1e-12)


def test_update_truncates_speed(orc):
    for mode in (K.GS_MODE_OPENSTEER, K.GS_MODE_GPUREF):
        s = one_agent(fwd=(0, 0, 1), speed=0.9, max_speed=1.0, max_force=100.0)
        s.arrays["gz"][0] = 100.0
        orc.update(s, 0.5, mode)
        assert s.arrays["speed"][0] == pytest.approx(1.0, rel=1e-14)


# ---------------------------------------------------------------------------
# World wrap.


def test_wrap_sphere_antipode(orc):
    s = one_agent(pos=(0, 0, 60.0), extra=[(30.0, 40.0, 0.0), (10.0, 0, 0)])
    orc.wrap_world(s, orc.WRAP_SPHERE, (0, 0, 0), (50.0,))
    np.testing.assert_allclose(s.pos()[0], [0, 0, -40.0])        # 60 - 2*50 along the ray
    np.testing.assert_allclose(s.pos()[1], [30.0, 40.0, 0.0])    # exactly on the sphere: r > R fails
    np.testing.assert_allclose(s.pos()[2], [10.0, 0, 0])


def test_wrap_box_faces(orc):
    s = one_agent(pos=(511.0, -1.0, 1000.5), extra=[(500.0, 500.0, 500.0)])
    orc.wrap_world(s, orc.WRAP_BOX, (500.0, 500.0, 500.0), (1000.0, 1000.0, 1000.0))
    np.testing.assert_array_equal(s.pos()[0], [511.0, 1000.0, 0.0])
    np.testing.assert_array_equal(s.pos()[1], [500.0, 500.0, 500.0])
