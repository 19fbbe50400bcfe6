"""Launch shapes of the k = 7 search and flocking kernels
(gs_context_set_shape): every shape the autotuner may pick gives the same
results, bit for bit, as the default, on the box flock (the lists of the
k = 7 searches and the whole state after fused steps, including a group
whose size is not a multiple of any block), and the shape setters check
their arguments."""
import numpy as np
import pytest

from gpusteer_amd import _capi as capi
from gpusteer_amd import autotune, scenarios
from helpers import assert_state_equal

pytestmark = pytest.mark.gpu


def _run(gs, ctx, st, p, steps):
    A = gs.AgentGroup(ctx, st)
    bins = gs.KNNBinData(ctx, p.grid_cells, p.grid_size, origin=p.grid_origin)
    knn = gs.KNNData(ctx, st.n, 7)
    gs.stepBoids(A, knn, bins, p, steps)
    # A stand-alone search of the final state (the plugin-compat call).
    gs.updateKNNDatabase(A, bins)
    k2 = gs.KNNData(ctx, st.n, 7)
    gs.findKNearestNeighbors(A, k2, bins, A, 1)
    out = (A.download(), knn.download(), k2.download())
    for h in (A, bins, knn, k2):
        h.close()
    return out


@pytest.mark.parametrize("n", [65536, 100003])
def test_every_launch_shape_gives_the_same_results(gs, n):
    st, p = scenarios.boids_box(n=n, k=7, cells=16)
    p.grid_cells = (16, 16, 48)
    with gs.Context(0) as ctx:
        ref = _run(gs, ctx, st, p, 6)
        for shape in autotune.SHAPES:
            print(f"shape {shape} n={n}", flush=True)
            ctx.set_shape(capi.GS_SHAPE_KNN, *shape)
            ctx.set_shape(capi.GS_SHAPE_FLOCK, *shape)
            assert ctx.shape(capi.GS_SHAPE_KNN) == shape
            got = _run(gs, ctx, st, p, 6)
            assert_state_equal(got[0], ref[0], what=f"shape {shape}: ")
            for j in (1, 2):
                np.testing.assert_array_equal(got[j][0], ref[j][0], err_msg=f"shape {shape} lists {j}")
                np.testing.assert_array_equal(got[j][1], ref[j][1], err_msg=f"shape {shape} distances {j}")


def test_shape_arguments_checked(gs, ctx):
    L = capi.lib
    assert L.gs_context_set_shape(ctx.h, capi.GS_SHAPE_KNN, 64, 1) == capi.GS_ERR_UNSUPPORTED
    assert L.gs_context_set_shape(ctx.h, capi.GS_SHAPE_FLOCK, 256, 3) == capi.GS_ERR_UNSUPPORTED
    assert L.gs_context_set_shape(ctx.h, 7, 256, 1) == capi.GS_ERR_INVALID_ARGUMENT
    assert L.gs_context_set_shape(None, capi.GS_SHAPE_KNN, 256, 1) == capi.GS_ERR_INVALID_HANDLE
    assert ctx.shape(capi.GS_SHAPE_KNN) == (256, 1)
