"""Pins the oracle's calibration model (the fused crowd step the adjoint
differentiates) to the oracle's single-operator functions, which the closed
forms of test_oracle_math.py pin: one model step must equal, bit for bit in
fp32, the operator sequence

    per group: clear -> avoid walls -> avoid obstacles -> separation (own,
    other) -> alignment -> cohesion -> seek;  then integrate both groups

(GPU-kernel formulas, plugins/ChokePointGroup.cpp:683-728 order plus the
alignment and cohesion terms), with KNN lists from the brute-force oracle
(own group, other group within the cap, wall midpoints, obstacle centres)
and every neighbour read from the step-start state (the agent's own
position and heading are the ones the wall pass left: an agent within its
radius of a wall is moved out and its heading turned along the wall).  One
crowd stays clear of the walls; one is spawned against them, so the
touching branch fires on some agents.
The loss is recomputed in numpy float64 from the composed trajectory:
mean_i |p_T - goal|^2 + pen_w * sum_t sum_i sum_{j in lists} max(0, r_i + r_j - d_ij)^2.
"""
import numpy as np
import pytest

from gpusteer_amd import _capi as K
from gpusteer_amd import scenarios
from gpusteer_amd.state import AgentState

GPUREF = 1


def _theta():
    return scenarios.calib_params(w_sep=(9.0, 11.0), w_ali=(1.5, 2.5), w_coh=(0.8, 1.2), w_obs=(1.1, 0.9),
                                  w_wall=(1.3, 0.7), radius=(0.45, 0.55), max_speed=(2.7, 3.3))


def _groups(sc, th, cfg):
    out = []
    for g, st in enumerate(sc.groups):
        a = {k: v.astype(np.float32).copy() for k, v in st.arrays.items()}
        n = st.n
        a["radius"] = np.full(n, th[10 + g], np.float32)
        a["max_speed"] = np.full(n, th[12 + g], np.float32)
        a["max_force"] = np.full(n, cfg.max_force, np.float32)
        a["mass"] = np.full(n, cfg.mass, np.float32)
        for f in ("gx", "gy", "gz", "ax", "ay", "az"):
            a[f][:] = 0
        out.append(AgentState(a, np.zeros(n, np.uint32)))
    return out


def _crowd(workload, offset=1.5, region=12.0, n_total=600):
    """A small choke-point crowd close to the wall line, moving (speed
    U(1, 2.5)) so the obstacle look-ahead (min time x speed) is not empty."""
    sc = scenarios.choke(n_total=n_total, offset=offset, region=region, workload=workload)
    sc.mode = GPUREF
    g = np.random.default_rng(workload)
    for st in sc.groups:
        st.arrays["speed"][:] = g.uniform(1.0, 2.5, st.n).astype(np.float32)
    return sc


def _pos(st):
    return np.stack([st.arrays["px"], st.arrays["py"], st.arrays["pz"]], axis=1).astype(np.float32)


def _composed(orc, sc, cfg, th, steps):
    G = _groups(sc, th, cfg)
    wmid = sc.walls.mid()
    opos = np.stack(sc.obstacles[:3], axis=1).astype(np.float32)
    k = sc.k
    r = (np.float32(th[10]), np.float32(th[11]))
    pen = 0.0
    moved = 0
    for _ in range(steps):
        pos = [_pos(s) for s in G]
        snap = [s.subset(np.arange(s.n)) for s in G]
        lists = []
        for gi in range(2):
            ks = orc.knn_brute(pos[gi], pos[gi], k, True)[0]
            ko = orc.knn_brute(pos[gi], pos[1 - gi], k, False, cap=sc.other_radius)[0]
            kw = orc.knn_brute(pos[gi], wmid, sc.knw, False)[0]
            kob = orc.knn_brute(pos[gi], opos, sc.kno, False)[0]
            lists.append((ks, ko, kw, kob))
            for which, lst, R in ((0, ks, r[gi] + r[gi]), (1, ko, r[gi] + r[1 - gi])):
                other = pos[gi] if which == 0 else pos[1 - gi]
                for j in range(k):
                    b = lst[j]
                    ok = b < other.shape[0]
                    d = np.linalg.norm(other[b[ok]].astype(np.float64) - pos[gi][ok].astype(np.float64), axis=1)
                    e = float(R) - d
                    pen += float((e[e > 0] ** 2).sum())
        for gi, st in enumerate(G):
            ks, ko, kw, kob = lists[gi]
            st.bits[:] = 0
            orc.avoid_walls(st, kw, sc.walls, cfg.walls_min_time, th[8 + gi], 0)
            orc.avoid_obstacles(st, kob, sc.obstacles, cfg.obstacles_min_time, th[6 + gi], 0)
            s, a, c = cfg.separation, cfg.alignment, cfg.cohesion
            moved += int(np.any(_pos(st) != pos[gi], axis=1).sum())
            orc.steer_flock(orc.FLOCK_SEPARATION, st, ks, snap[gi], s.min_d, s.max_d, s.cos_max, th[0 + gi], s.mask,
                            GPUREF)
            orc.steer_flock(orc.FLOCK_SEPARATION, st, ko, snap[1 - gi], s.min_d, s.max_d, s.cos_max, th[0 + gi],
                            s.mask, GPUREF)
            orc.steer_flock(orc.FLOCK_ALIGNMENT, st, ks, snap[gi], a.min_d, a.max_d, a.cos_max, th[2 + gi], a.mask,
                            GPUREF)
            orc.steer_flock(orc.FLOCK_COHESION, st, ks, snap[gi], c.min_d, c.max_d, c.cos_max, th[4 + gi], c.mask,
                            GPUREF)
            orc.steer_seek(st, sc.goals[gi], cfg.seek_weight, cfg.seek_mask, GPUREF)
        for st in G:
            orc.update(st, cfg.dt, GPUREF)
    term = 0.0
    for gi, st in enumerate(G):
        term += float(((_pos(st).astype(np.float64) - np.array(sc.goals[gi])) ** 2).sum())
    loss = cfg.penetration_weight * pen + term / (G[0].n + G[1].n)
    return G, loss, moved


@pytest.mark.parametrize("steps,workload,offset", [(1, 3, 1.5), (4, 5, 1.5), (1, 7, 0.3), (4, 9, 0.3)],
                         ids=["clear-1", "clear-4", "touching-1", "touching-4"])
def test_calibration_model_equals_operator_sequence(orc, steps, workload, offset):
    # Crowds spawned from 1.5 units off the wall line: the feelers (4 r ahead
    # and to the sides) reach the walls, but nobody comes within r of them in
    # four steps of at most 3.3 * 0.05.  Spawned from 0.3 over x in
    # [-15, 15] (the gap is x in [-7.5, 7.5], the wall 0.5 thick, agents of
    # radius 0.45 / 0.55): the first rows beside the gap are within their
    # radius of the wall faces and are moved out.
    sc = _crowd(workload, offset, 12.0, 600) if offset > 1.0 else _crowd(workload, offset, 30.0, 4000)
    cfg = scenarios.CalibConfig(steps=steps, checkpoint_every=steps)
    th = _theta()
    G, loss, moved = _composed(orc, sc, cfg, th, steps)
    print(f"agents moved out of a wall: {moved}")
    assert (moved > 0) == (offset < 1.0), moved
    c = orc.C4(sc, cfg)
    loss32, fs = c.run(th, 1, want_state=True, precision=32)
    c.close()
    n0, n1 = sc.groups[0].n, sc.groups[1].n
    for g, (off, n) in enumerate(((0, n0), (13 * n0, n1))):
        for j, f in enumerate(orc.C4_STATE):
            np.testing.assert_array_equal(fs[off + j * n: off + (j + 1) * n].astype(np.float32), G[g].arrays[f],
                                          err_msg=f"group {g} field {f}")
    assert abs(loss32 - loss) <= 1e-5 * abs(loss)


def test_operator_sequence_exercises_every_term(orc):
    """The composition above is only a pin if the terms fire: walls,
    obstacles, separation, alignment, cohesion and seek all set their bits on
    some agent of the test crowd."""
    sc = _crowd(3)
    cfg = scenarios.CalibConfig(steps=1)
    th = _theta()
    G = _groups(sc, th, cfg)
    wmid = sc.walls.mid()
    opos = np.stack(sc.obstacles[:3], axis=1).astype(np.float32)
    seen = 0
    for gi, st in enumerate(G):
        p = _pos(st)
        ks = orc.knn_brute(p, p, sc.k, True)[0]
        kw = orc.knn_brute(p, wmid, sc.knw, False)[0]
        kob = orc.knn_brute(p, opos, sc.kno, False)[0]
        orc.avoid_walls(st, kw, sc.walls, cfg.walls_min_time, 1.0, 0)
        orc.avoid_obstacles(st, kob, sc.obstacles, cfg.obstacles_min_time, 1.0, 0)
        s, a, c = cfg.separation, cfg.alignment, cfg.cohesion
        orc.steer_flock(orc.FLOCK_SEPARATION, st, ks, st, s.min_d, s.max_d, s.cos_max, 1.0, 0, GPUREF)
        orc.steer_flock(orc.FLOCK_ALIGNMENT, st, ks, st, a.min_d, a.max_d, a.cos_max, 1.0, 0, GPUREF)
        orc.steer_flock(orc.FLOCK_COHESION, st, ks, st, c.min_d, c.max_d, c.cos_max, 1.0, 0, GPUREF)
        orc.steer_seek(st, sc.goals[gi], 1.0, 0, GPUREF)
        seen |= int(np.bitwise_or.reduce(st.bits))
    for name in ("AVOID_WALLS", "AVOID_OBSTACLES", "SEPARATION", "ALIGNMENT", "COHESION", "SEEK"):
        assert seen & getattr(K, "BIT_" + name), name
