"""Argument checking of the plugin-compat, neighbour-avoidance and
extraction entry points on a live context: every bad call returns its
status (raised as GpuSteerError by the Python layer) and leaves the
library usable."""
import ctypes as C

import numpy as np
import pytest

from gpusteer_amd import _capi as capi
from gpusteer_amd import scenarios
from gpusteer_amd._capi import GpuSteerError

pytestmark = pytest.mark.gpu


def _agents(n=256, seed=71):
    g = scenarios.rng(seed)
    pos = g.uniform(-10, 10, size=(n, 3)).astype(np.float32)
    return scenarios.make_agents(pos, scenarios.unit_vectors(g, n), np.ones(n, np.float32), 0.5, 3.0, 3.0)


def test_path_arguments(gs, ctx):
    with pytest.raises(GpuSteerError, match="at least 2 points"):
        gs.PolylinePathway(ctx, [(0.0, 0.0, 0.0)], 1.0)
    h = C.c_void_p()
    pts = np.zeros(6, np.float32)
    assert capi.lib.gs_path_create(ctx.h, 2, capi.fptr(pts), 1.0, 0, None) == capi.GS_ERR_INVALID_ARGUMENT
    assert capi.lib.gs_path_create(ctx.h, 2, None, 1.0, 0, C.byref(h)) == capi.GS_ERR_INVALID_ARGUMENT
    P = gs.PolylinePathway(ctx, [(0.0, 0.0, 0.0), (1.0, 0.0, 0.0)], 1.0, cyclic=True)
    assert P.m == 3 and P.total == 2.0
    P.close()


def test_knn_must_match_the_groups(gs, ctx):
    """A KNN container computed for another pair of groups is refused by the
    anti-penetration and avoidance calls."""
    a, b = _agents(256, 72), _agents(256, 73)
    bins = gs.KNNBinData(ctx, (4, 4, 4), (24.0, 24.0, 24.0))
    A, B = gs.AgentGroup(ctx, a), gs.AgentGroup(ctx, b)
    gs.updateKNNDatabase(A, bins)
    gs.updateKNNDatabase(B, bins)
    knn = gs.KNNData(ctx, a.n, 4)
    gs.findKNearestNeighbors(A, knn, bins, A, 1)
    with pytest.raises(GpuSteerError, match="does not hold a result"):
        gs.antiPenetrationAgents(A, knn, B, 0)
    with pytest.raises(GpuSteerError, match="does not hold a result"):
        gs.steerToAvoidNeighbors(A, knn, B, 5.0, 0.5, True, 1.0, 0, gs.MODE_GPUREF)
    with pytest.raises(GpuSteerError, match="unknown mode"):
        gs.steerToAvoidNeighbors(A, knn, A, 5.0, 0.5, True, 1.0, 0, 7)
    W = gs.WallGroup(ctx)
    with pytest.raises(GpuSteerError):
        gs.updateGroupWithAntiPenetration(A, knn, W, 0.1)
    with pytest.raises(GpuSteerError, match="closest-point"):
        capi.check(capi.lib.gs_update_group_anti_penetration(A.h, knn.h, W.h, 0.1, 5), "x")
    # still usable afterwards
    gs.antiPenetrationAgents(A, knn, A, 0)
    for h in (A, B, W, bins, knn):
        h.close()


def test_extraction_arguments(gs, ctx, tmp_path):
    A = gs.AgentGroup(ctx, _agents(128, 74))
    with pytest.raises(GpuSteerError, match="no bin database"):
        A.cell_density(8)
    bins = gs.KNNBinData(ctx, (2, 2, 2), (24.0, 24.0, 24.0))
    gs.updateKNNDatabase(A, bins)
    assert A.cell_density(8).sum() == 128
    with pytest.raises(GpuSteerError, match="selected cell out of range"):
        A.write_cell_density(tmp_path / "d", [8], tmp_path / "s")
    sel = np.zeros(1, np.uint32)
    assert capi.lib.gs_write_cell_density(A.h, b"x", capi.uptr(sel), 1, None, 0) == capi.GS_ERR_INVALID_ARGUMENT
    assert capi.lib.gs_write_cell_velocity(A.h, None, None, 0, None, 0) == capi.GS_ERR_INVALID_ARGUMENT
    with pytest.raises(GpuSteerError, match="cannot open"):
        A.write_cell_velocity(tmp_path / "missing_dir" / "v")
    for h in (A, bins):
        h.close()
