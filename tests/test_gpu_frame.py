"""Recorded frames (gs_frame_begin / end / launch): a call sequence recorded
once and replayed gives the same results, bit for bit, as issuing the calls
(the fused Boids step against the oracle, the choke-point plugins' and the
multi-group plugin's frames and the fused crowd step against eager runs of
the same calls); recording runs nothing and leaves every handle as it was;
calls that copy, allocate or synchronize are refused while recording; a
replay on handles in another state is refused before it launches."""

import numpy as np
import pytest

from gpusteer_amd import _capi as capi
from gpusteer_amd import formats, scenarios
from gpusteer_amd.plugins import ChokePointNew, MultiGroup
from helpers import assert_state_equal

pytestmark = pytest.mark.gpu

FILES = formats.DATA_DIR


def _knn_equal(a, b, what):
    ia, da = a.download()
    ib, db = b.download()
    np.testing.assert_array_equal(ia, ib, err_msg=what)
    np.testing.assert_array_equal(da, db, err_msg=what)


def test_frame_boids_steps_equal_oracle(gs, ctx, orc):
    """The 4,096-agent sphere flock: 2 unrecorded steps, then a frame of
    gs_step_boids(2) replayed 40 times (80 steps): KNN lists and state equal
    the oracle's after 82 steps.  Recording itself runs nothing."""
    st, p = scenarios.boids_sphere()
    ref = st.copy()
    A = gs.AgentGroup(ctx, st)
    bins = gs.KNNBinData(ctx, p.grid_cells, p.grid_size, origin=p.grid_origin)
    knn = gs.KNNData(ctx, st.n, p.k)
    gs.stepBoids(A, knn, bins, p, 2)
    for _ in range(2):
        orc.boids_step(ref, p)
    before = A.download()
    print("frame: recording gs_step_boids(2)", flush=True)
    f = ctx.record(lambda: gs.stepBoids(A, knn, bins, p, 2))
    assert_state_equal(A.download(), before, what="after recording: ")
    assert_state_equal(before, ref, what="after recording: ")
    f.launch(40)
    for _ in range(80):
        r_idx, r_dist = orc.boids_step(ref, p)
    idx, dist = knn.download()
    np.testing.assert_array_equal(idx, r_idx)
    np.testing.assert_array_equal(dist, r_dist)
    assert_state_equal(A.download(), ref, what="after 82 steps: ")
    # An unrecorded step after the replays continues from the replayed state.
    gs.stepBoids(A, knn, bins, p, 1)
    r_idx, _ = orc.boids_step(ref, p)
    np.testing.assert_array_equal(knn.download()[0], r_idx)
    assert_state_equal(A.download(), ref, what="after 83 steps: ")
    # The other double-buffer state: the frame refuses to replay.
    with pytest.raises(capi.GpuSteerError) as e:
        f.launch(1)
    assert e.value.status == capi.GS_ERR_FRAME
    assert_state_equal(A.download(), ref, what="after the refused replay: ")
    f.close()
    for h in (A, bins, knn):
        h.close()


def test_frame_boids_large_flock_equals_eager(gs, ctx):
    """The 2^20-agent box flock at k = 7 on 64 x 64 x 192 cells: a frame of
    two fused steps replayed 5 times equals 10 unrecorded steps of a twin
    group, bit for bit."""
    st, p = scenarios.boids_box(n=1 << 20, k=7, cells=64)
    p.grid_cells = (64, 64, 192)
    A, B = gs.AgentGroup(ctx, st), gs.AgentGroup(ctx, st)
    bins = gs.KNNBinData(ctx, p.grid_cells, p.grid_size, origin=p.grid_origin)
    ka, kb = gs.KNNData(ctx, st.n, p.k), gs.KNNData(ctx, st.n, p.k)
    gs.stepBoids(A, ka, bins, p, 2)
    gs.stepBoids(B, kb, bins, p, 12)
    print("frame: 2^20 flock", flush=True)
    with ctx.record(lambda: gs.stepBoids(A, ka, bins, p, 2)) as f:
        f.launch(5)
    _knn_equal(ka, kb, "2^20 flock lists")
    assert_state_equal(A.download(), B.download(), what="2^20 flock: ")
    for h in (A, B, bins, ka, kb):
        h.close()


@pytest.fixture(scope="module")
def choke_files():
    p = formats.load_params(FILES / "ChokePoint.params")
    ws = formats.WallSet.load_map(FILES / p.map_filename).split(p.dimensions, p.cells)
    walls, mid = ws.get()
    ws.close()
    return p, walls, mid


@pytest.mark.parametrize("variant", ["step", "step_with_wall_update"])
def test_frame_choke_point_plugin_equals_eager(gs, ctx, choke_files, variant):
    """The choke-point plugins' frames on ChokePoint.params (4096 agents, its
    split map): 25 steps from recorded frames (odd counts, so both
    double-buffer states get a frame) equal 25 issued steps bit for bit."""
    p, walls, mid = choke_files
    sc = scenarios.choke_point_new(p, walls, mid, n_agents=4096, dt=0.25, start=(0.0, 0.0, 60.0),
                                   max_start_radius=50.0)
    a, b = ChokePointNew(ctx, sc), ChokePointNew(ctx, sc)
    print(f"frame: choke-point plugin {variant}", flush=True)
    for n in (3, 7, 10, 5):
        a.recorded(variant, n)
        getattr(b, variant)(n)
    assert len(a._frames[variant].frames) == 2
    _knn_equal(a.knn_self, b.knn_self, "own lists")
    _knn_equal(a.knn_walls, b.knn_walls, "wall lists")
    assert_state_equal(a.group.download(), b.group.download(), what=f"{variant}: ")
    a.close()
    b.close()


def test_frame_multi_group_equals_eager(gs, ctx):
    """plugins/MultiGroup.cpp's frame on MultiGroup.params (two groups of
    1500): 20 steps from recorded frames equal 20 issued steps."""
    p = formats.load_params(FILES / "MultiGroup.params")
    sc = scenarios.multi_group(p, n_per_group=1500, dt=0.5, max_start_radius=30.0, start_scale=0.15)
    a, b = MultiGroup(ctx, sc), MultiGroup(ctx, sc)
    print("frame: multi-group", flush=True)
    a.recorded(20)
    b.step(20)
    for ga, gb in zip(a.groups, b.groups):
        assert_state_equal(ga.download(), gb.download(), what="multi-group: ")
    for ra, rb in zip(a.knn, b.knn):
        for ka, kb in zip(ra, rb):
            _knn_equal(ka, kb, "multi-group lists")
    a.close()
    b.close()


def test_frame_choke_crowd_equals_eager(gs, ctx):
    """gs_step_choke on the 10,000-agent crowd: a frame of 2 steps replayed
    50 times equals 100 issued steps (both groups' state and all 8 lists)."""
    from gpusteer_amd.ensemble import step_choke

    sc = scenarios.choke()
    sims = []
    for _ in range(2):
        bins = gs.KNNBinData(ctx, sc.grid_cells, sc.grid_size, origin=sc.grid_origin)
        G = [gs.AgentGroup(ctx, g) for g in sc.groups]
        W = gs.WallGroup(ctx, sc.walls)
        O = gs.ObstacleGroup(ctx, *sc.obstacles)
        gs.updateKNNDatabase(W, bins)
        gs.updateKNNDatabase(O, bins)
        knns = [[gs.KNNData(ctx, g.n, k) for k in (sc.k, sc.k, sc.knw, sc.kno)] for g in sc.groups]
        sims.append((G, knns, W, O, bins))
    (Ga, ka, Wa, Oa, ba), (Gb, kb, Wb, Ob, bb) = sims
    step_choke(Ga, ka, Wa, Oa, ba, sc, 2)
    step_choke(Gb, kb, Wb, Ob, bb, sc, 102)
    print("frame: choke crowd", flush=True)
    with ctx.record(lambda: step_choke(Ga, ka, Wa, Oa, ba, sc, 2)) as f:
        f.launch(50)
    for gi in range(2):
        assert_state_equal(Ga[gi].download(), Gb[gi].download(), what=f"group {gi}: ")
        for j in range(4):
            _knn_equal(ka[gi][j], kb[gi][j], f"group {gi} list {j}")
    for G, knns, W, O, bins in sims:
        for h in G + [W, O, bins] + [k for ks in knns for k in ks]:
            h.close()


def test_frame_refusals_leave_handles_intact(gs, ctx, orc):
    """While recording, an upload, a download, a synchronize and a search
    whose buffers would grow fail with GS_ERR_FRAME; gs_frame_end then fails
    too and records nothing; the handles are as before (the next issued
    steps equal the oracle's).  A frame recorded before an upload of the
    group refuses to replay after it."""
    st, p = scenarios.boids_sphere()
    ref = st.copy()
    A = gs.AgentGroup(ctx, st)
    bins = gs.KNNBinData(ctx, p.grid_cells, p.grid_size, origin=p.grid_origin)
    knn = gs.KNNData(ctx, st.n, p.k)
    gs.stepBoids(A, knn, bins, p, 2)
    for _ in range(2):
        orc.boids_step(ref, p)
    L = capi.lib
    fresh = gs.AgentGroup(ctx, st)  # never binned: its bin buffers would grow
    for name, call in (("upload", lambda: L.gs_group_upload_agents(A.h, 0, None)),
                       ("synchronize", lambda: L.gs_context_synchronize(ctx.h)),
                       ("download", lambda: A.download()),
                       ("grow", lambda: gs.updateKNNDatabase(fresh, bins))):
        print(f"frame: refusal {name}", flush=True)
        assert L.gs_frame_begin(ctx.h) == 0
        gs.stepBoids(A, knn, bins, p, 1)
        try:
            rc = call()
        except capi.GpuSteerError as e:
            rc = e.status
        assert rc == capi.GS_ERR_FRAME, name
        h = capi.C.c_void_p()
        assert L.gs_frame_end(ctx.h, capi.C.byref(h)) == capi.GS_ERR_FRAME
        assert not h.value
    gs.stepBoids(A, knn, bins, p, 3)
    for _ in range(3):
        r_idx, _ = orc.boids_step(ref, p)
    np.testing.assert_array_equal(knn.download()[0], r_idx)
    assert_state_equal(A.download(), ref, what="after the refused recordings: ")
    gs.stepBoids(A, knn, bins, p, 1)
    f = ctx.record(lambda: gs.stepBoids(A, knn, bins, p, 2))
    A.upload(st)
    with pytest.raises(capi.GpuSteerError) as e:
        f.launch(1)
    assert e.value.status == capi.GS_ERR_FRAME
    f.close()
    for h in (A, bins, knn, fresh):
        h.close()
