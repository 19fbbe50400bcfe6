"""GPU parity of the plugin-compat kernels (gs_update_group_anti_penetration,
gs_anti_penetration_agents, gs_steer_follow_path over gs_path) and of the
choke-point plugins' per-step sequences that use them, bit for bit against
the oracle, on the reference's own ChokePoint.params and its split map."""
from pathlib import Path

import numpy as np
import pytest

from gpusteer_amd import formats, scenarios
from gpusteer_amd.plugins import ChokePointNew
from gpusteer_amd.state import WallState
from helpers import assert_state_equal

pytestmark = pytest.mark.gpu

FILES = formats.DATA_DIR


@pytest.fixture(scope="module")
def choke_files():
    p = formats.load_params(FILES / "ChokePoint.params")
    ws = formats.WallSet.load_map(FILES / p.map_filename).split(p.dimensions, p.cells)
    walls, mid = ws.get()
    ws.close()
    return p, walls, mid


def _grid(orc, p):
    return orc.grid(tuple(-0.5 * np.float32(v) for v in p.dimensions), p.dimensions, p.cells)


def _near_walls(walls: WallState, n, seed, spread=1.0):
    """Agents on random points of random wall pieces, offset along the normal
    by U(-spread, spread), random XZ headings, speeds U(0, 3), random
    steering; radius 0.5."""
    g = scenarios.rng(seed)
    a = walls.arrays
    w = g.integers(0, walls.n, size=n)
    t = g.uniform(0, 1, size=n)
    off = g.uniform(-spread, spread, size=n)
    s = np.stack([a["sx"][w], a["sy"][w], a["sz"][w]], axis=1).astype(np.float64)
    e = np.stack([a["ex"][w], a["ey"][w], a["ez"][w]], axis=1).astype(np.float64)
    nr = np.stack([a["nx"][w], a["ny"][w], a["nz"][w]], axis=1).astype(np.float64)
    pos = (s + (e - s) * t[:, None] + nr * off[:, None]).astype(np.float32)
    hd = g.uniform(0, 2 * np.pi, size=n)
    fwd = np.stack([np.cos(hd), np.zeros(n), np.sin(hd)], axis=1)
    st = scenarios.make_agents(pos, fwd, g.uniform(0, 3, size=n).astype(np.float32), 0.5, 3.0, 3.0)
    st.arrays["gx"][:] = g.normal(scale=2.0, size=n)
    st.arrays["gz"][:] = g.normal(scale=2.0, size=n)
    return st


@pytest.mark.parametrize("closest", [0, 1], ids=["clamped", "as_reference"])
def test_update_anti_penetration_equals_oracle(gs, ctx, orc, choke_files, closest):
    """20,000 agents straddling the 1254 split walls, dt 0.5: the wall KNN
    (capped at 40) and the integrate with anti-penetration equal the
    oracle's; a good share of agents is pushed out or stopped at a wall."""
    p, walls, mid = choke_files
    st = _near_walls(walls, 20000, 41)
    st.bits[:] = np.where(np.arange(st.n) % 3 == 0, 0x110, 0).astype(np.uint32)
    ref = st.copy()
    print(f"update anti-penetration closest={closest} n={st.n}", flush=True)
    bins = gs.KNNBinData(ctx, p.cells, p.dimensions)
    W = gs.WallGroup(ctx)
    W.upload(walls, mid)
    A = gs.AgentGroup(ctx, st)
    gs.updateKNNDatabase(W, bins)
    gs.updateKNNDatabase(A, bins)
    kw = gs.KNNData(ctx, st.n, p.knw)
    gs.findKNearestNeighborsWithin(A, kw, bins, W, 40.0, p.search_radius)
    gs.updateGroupWithAntiPenetration(A, kw, W, 0.5, closest)
    pos = np.stack([ref.arrays[f] for f in ("px", "py", "pz")], axis=1)
    r_kw, _ = orc.knn_binned(_grid(orc, p), pos, mid, p.knw, False, cap=40.0)
    orc.update_anti_penetration(ref, r_kw, walls, 0.5, closest)
    np.testing.assert_array_equal(kw.download()[0], r_kw)
    out = A.download()
    assert_state_equal(out, ref)
    hit = np.count_nonzero(out.bits & (1 << 12))
    print(f"agents with the anti-penetration bit: {hit}", flush=True)
    assert hit > st.n // 20
    assert np.array_equal(out.bits & 0x110, st.bits)  # bits kept
    for h in (A, W, bins, kw):
        h.close()


def _crowd(n, half, seed, mass_bits=True):
    g = scenarios.rng(seed)
    pos = np.stack([g.uniform(-half, half, size=n), np.zeros(n), g.uniform(-half, half, size=n)],
                   axis=1).astype(np.float32)
    hd = g.uniform(0, 2 * np.pi, size=n)
    fwd = np.stack([np.cos(hd), np.zeros(n), np.sin(hd)], axis=1)
    st = scenarios.make_agents(pos, fwd, g.uniform(0, 3, size=n).astype(np.float32), 0.5, 3.0, 3.0)
    st.arrays["radius"][:] = g.uniform(0.3, 0.7, size=n)
    if mass_bits:
        st.bits[:] = np.where(g.uniform(size=n) < 0.2, 1 << 12, 0).astype(np.uint32)
    return st


def test_anti_penetration_agents_self_equals_oracle(gs, ctx, orc):
    """30,000 agents at ~8 per unit area (most overlap their nearest), a
    fifth with the masked bit: two calls in a row (each reads the state the
    last one left), then a seek on the result, equal the oracle."""
    st = _crowd(30000, 30.0, 43)
    ref = st.copy()
    mask = 1 << 12
    cells, size = (16, 1, 16), (64.0, 10.0, 64.0)
    print(f"anti-penetration agents self n={st.n}", flush=True)
    bins = gs.KNNBinData(ctx, cells, size)
    A = gs.AgentGroup(ctx, st)
    knn = gs.KNNData(ctx, st.n, 6)
    g = orc.grid((-32.0, -5.0, -32.0), size, cells)
    for call in range(2):
        gs.updateKNNDatabase(A, bins)
        gs.findKNearestNeighbors(A, knn, bins, A, 1)
        gs.antiPenetrationAgents(A, knn, A, mask)
        pos = np.stack([ref.arrays[f] for f in ("px", "py", "pz")], axis=1)
        ks, ds = orc.knn_binned(g, pos, pos, 6, True)
        orc.anti_penetration_agents(ref, ks, ds, ref, mask)
        np.testing.assert_array_equal(knn.download()[0], ks)
        assert_state_equal(A.download(), ref, what=f"call {call}: ")
    moved = np.count_nonzero(ref.bits & (1 << 13))
    print(f"agents moved: {moved}", flush=True)
    assert moved > st.n // 2
    gs.steerForSeek(A, (5.0, 0.0, 5.0), 1.0, 0, gs.MODE_GPUREF)
    orc.steer_seek(ref, (5.0, 0.0, 5.0), 1.0, 0, orc.MODE_GPUREF)
    assert_state_equal(A.download(), ref, what="seek after: ")
    for h in (A, bins, knn):
        h.close()


def test_anti_penetration_agents_two_groups_equals_oracle(gs, ctx, orc):
    """A pushed by B's agents: B's masked bits take the whole overlap; B is
    unchanged."""
    a, b = _crowd(12000, 25.0, 44, mass_bits=False), _crowd(9000, 25.0, 45)
    ra, rb = a.copy(), b.copy()
    cells, size = (16, 1, 16), (64.0, 10.0, 64.0)
    bins = gs.KNNBinData(ctx, cells, size)
    A, B = gs.AgentGroup(ctx, a), gs.AgentGroup(ctx, b)
    gs.updateKNNDatabase(A, bins)
    gs.updateKNNDatabase(B, bins)
    knn = gs.KNNData(ctx, a.n, 4)
    gs.findKNearestNeighbors(A, knn, bins, B, 1)
    gs.antiPenetrationAgents(A, knn, B, 1 << 12)
    pa = np.stack([ra.arrays[f] for f in ("px", "py", "pz")], axis=1)
    pb = np.stack([rb.arrays[f] for f in ("px", "py", "pz")], axis=1)
    ks, ds = orc.knn_binned(orc.grid((-32.0, -5.0, -32.0), size, cells), pa, pb, 4, False)
    orc.anti_penetration_agents(ra, ks, ds, rb, 1 << 12)
    assert_state_equal(A.download(), ra)
    assert_state_equal(B.download(), b)
    for h in (A, B, bins, knn):
        h.close()


PATHS = [
    ("choke", [(0.0, 0.0, 13.5), (0.0, 0.0, -1650.0)], 20.0, False),
    ("cyclic", [(0.0, 0.0, 0.0), (60.0, 0.0, 10.0), (70.0, 0.0, 80.0), (-20.0, 0.0, 90.0), (-40.0, 0.0, 30.0)],
     6.0, True),
    ("open3d", [(-50.0, -5.0, -50.0), (0.0, 5.0, 0.0), (50.0, 0.0, -10.0)], 4.0, False),
]


@pytest.mark.parametrize("name,pts,radius,cyclic", PATHS, ids=[p[0] for p in PATHS])
def test_follow_path_equals_oracle(gs, ctx, orc, name, pts, radius, cyclic):
    """Agents inside and outside the tube, moving along, against and across
    the path, a tenth at rest, a fifth masked: steering and bits equal the
    oracle's; the host copies of the path equal the oracle's
    PolylinePathwayCUDA arrays."""
    g = scenarios.rng(46)
    n = 50000
    c = np.mean(np.asarray(pts, np.float64), axis=0)
    pos = (c + g.normal(scale=60.0, size=(n, 3)) * np.array([1.0, 0.05, 1.0])).astype(np.float32)
    fwd = scenarios.unit_vectors(g, n)
    speed = np.where(g.uniform(size=n) < 0.1, 0.0, g.uniform(0, 3, size=n)).astype(np.float32)
    st = scenarios.make_agents(pos, fwd, speed, 0.5, 3.0, 3.0)
    st.arrays["gx"][:] = g.normal(size=n)
    st.bits[:] = np.where(g.uniform(size=n) < 0.2, 16, 0).astype(np.uint32)
    ref = st.copy()
    print(f"follow path {name} n={n}", flush=True)
    P = gs.PolylinePathway(ctx, pts, radius, cyclic)
    R = orc.Pathway(pts, radius, cyclic)
    hp, hd, hl = P.host_arrays()
    np.testing.assert_array_equal(hp, R.points)
    np.testing.assert_array_equal(hd, R.normals)
    np.testing.assert_array_equal(hl, R.lengths)
    assert np.float32(P.total) == R.total
    A = gs.AgentGroup(ctx, st)
    gs.steerToFollowPath(A, P, 5.0, 1.5, 16)
    orc.follow_path(ref, R, 5.0, 1.5, 16)
    out = A.download()
    assert_state_equal(out, ref)
    on = np.count_nonzero(out.bits & (1 << 11))
    assert 0 < on < n
    for h in (A, P):
        h.close()


@pytest.mark.parametrize("variant,closest", [("new", 0), ("group", 0), ("group", 1)])
def test_choke_point_plugin_steps_bit_exact(gs, ctx, orc, choke_files, variant, closest):
    """The choke-point plugins' per-step sequences on ChokePoint.params with
    its split map (4096 agents started around (0, 0, 60) within 50, dt 0.25,
    80 steps): KNN lists and the whole state equal the oracle's."""
    p, walls, mid = choke_files
    sc = scenarios.choke_point_new(p, walls, mid, n_agents=4096, dt=0.25, start=(0.0, 0.0, 60.0),
                                   max_start_radius=50.0)
    ref = sc.state.copy()
    R = orc.Pathway(sc.path_points, sc.path_radius, sc.path_cyclic)
    sim = ChokePointNew(ctx, sc, closest)
    print(f"choke-point plugin {variant} closest={closest}", flush=True)
    for step in range(80):
        if variant == "new":
            sim.step()
        else:
            sim.step_with_wall_update()
        ks, kw, _ = orc.choke_new_step(ref, sc, R, variant, closest)
        if step % 20 == 0 or step == 79:
            np.testing.assert_array_equal(sim.knn_self.download()[0], ks, err_msg=f"step {step}")
            np.testing.assert_array_equal(sim.knn_walls.download()[0], kw, err_msg=f"step {step}")
            assert_state_equal(sim.group.download(), ref, what=f"step {step}: ")
    z = ref.arrays["pz"]
    print(f"agents past z = 13.5: {np.count_nonzero(z < 13.5)}", flush=True)
    sim.close()


def _crossing(n, seed, heading_z):
    g = scenarios.rng(seed)
    pos = np.stack([g.uniform(-40, 40, n), np.zeros(n), g.uniform(-40, 40, n)], axis=1).astype(np.float32)
    hd = np.where(g.uniform(size=n) < 0.7, np.arctan2(heading_z, 0.0), g.uniform(0, 2 * np.pi, n))
    hd = hd + g.normal(scale=0.3, size=n)
    fwd = np.stack([np.cos(hd), np.zeros(n), np.sin(hd)], axis=1)
    st = scenarios.make_agents(pos, fwd, g.uniform(0, 3, n).astype(np.float32), 0.5, 3.0, 3.0)
    st.arrays["radius"][:] = g.uniform(0.3, 0.7, n)
    st.arrays["gx"][:] = g.normal(size=n)
    st.bits[:] = np.where(g.uniform(size=n) < 0.1, 16, 0).astype(np.uint32)
    return st


@pytest.mark.parametrize("mode", [0, 1], ids=["openSteer", "gpuRef"])
@pytest.mark.parametrize("close", [0, 1])
@pytest.mark.parametrize("self_group", [False, True])
def test_avoid_neighbors_equals_oracle(gs, ctx, orc, mode, close, self_group):
    """Two crowds crossing head-on (70 % heading +z / -z, the rest random),
    k = 7: steering and bits of steerToAvoidNeighbors equal the oracle's; the
    close-neighbour rule fires where it applies (the crowd is dense, so it
    takes most agents) and the collision rule on the rest."""
    a = _crossing(12000, 51, 1.0)
    b = a if self_group else _crossing(10000, 52, -1.0)
    ra = a.copy()
    rb = ra if self_group else b.copy()
    cells, size = (32, 1, 32), (84.0, 10.0, 84.0)
    bins = gs.KNNBinData(ctx, cells, size)
    A = gs.AgentGroup(ctx, a)
    B = A if self_group else gs.AgentGroup(ctx, b)
    gs.updateKNNDatabase(A, bins)
    if not self_group:
        gs.updateKNNDatabase(B, bins)
    knn = gs.KNNData(ctx, a.n, 7)
    gs.findKNearestNeighbors(A, knn, bins, B, 1)
    print(f"avoid neighbours mode={mode} close={close} self={self_group}", flush=True)
    gs.steerToAvoidNeighbors(A, knn, B, 10.0, 0.5, close, 1.5, 16, mode)
    pa = np.stack([ra.arrays[f] for f in ("px", "py", "pz")], axis=1)
    pb = np.stack([rb.arrays[f] for f in ("px", "py", "pz")], axis=1)
    g = orc.grid((-42.0, -5.0, -42.0), size, cells)
    ki, kd = orc.knn_binned(g, pa, pb, 7, self_group)
    orc.avoid_neighbors(ra, ki, kd, rb, 10.0, 0.5, close, 1.5, 16, mode)
    np.testing.assert_array_equal(knn.download()[0], ki)
    out = A.download()
    assert_state_equal(out, ra)
    n_an = np.count_nonzero(out.bits & (1 << 9))
    n_close = np.count_nonzero(out.bits & (1 << 10))
    print(f"avoid {n_an} close {n_close}", flush=True)
    if close or mode == 0:
        assert n_close > 100 and n_an > 0
    else:
        assert n_close == 0 and n_an > 100
    for h in (A, bins, knn) + (() if self_group else (B,)):
        h.close()


@pytest.mark.parametrize("name,close", [("MultiGroup.params", 0), ("MultiGroup4.params", 1)])
def test_multi_group_steps_bit_exact(gs, ctx, orc, name, close):
    """plugins/MultiGroup.cpp's frame (preUpdate of every group, then each
    group's update: KNN to every group, avoid the others, separation from
    all, alignment and cohesion with its own, path and seek, integrate,
    anti-penetration against all) on the reference's MultiGroup files, 1500
    agents per group started 0.15 x as far out within 30, dt 0.5, 30 frames,
    group searches capped at the behaviours' reach: every group's KNN lists
    and state equal the oracle's."""
    from gpusteer_amd.plugins import MultiGroup

    p = formats.load_params(FILES / name)
    p.values = dict(p.values, avoid_close_neighbors=close)
    sc = scenarios.multi_group(p, n_per_group=1500, dt=0.5, max_start_radius=30.0, start_scale=0.15)
    refs = [s.copy() for s in sc.states]
    paths = [orc.Pathway(pts, r, c) if len(pts) >= 2 else None for pts, r, c in sc.paths]
    sim = MultiGroup(ctx, sc)
    print(f"multi-group {name} groups={len(refs)}", flush=True)
    for step in range(30):
        sim.step()
        knn = orc.multi_group_step(refs, sc, paths)
        if step % 10 == 0 or step == 29:
            for gi in range(len(refs)):
                for j in range(len(refs)):
                    np.testing.assert_array_equal(sim.knn[gi][j].download()[0], knn[gi][j][0],
                                                  err_msg=f"step {step} knn {gi}->{j}")
                assert_state_equal(sim.groups[gi].download(), refs[gi], what=f"step {step} group {gi}: ")
    sim.close()
