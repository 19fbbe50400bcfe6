"""gs_find_knn_noalloc (the exact KNN of points in device memory, with its
bin database in the library's fixed device array) against the existing
entry points on a group holding the same points (gs_update_knn_database +
gs_find_knn of the group with itself, then gs_knn_download): indices and
fp32 distances bit for bit, on two calls in a row, the second with a smaller
input (the array keeps the first call's contents); sizes beyond the limits
refused with GS_ERR_UNSUPPORTED."""
import ctypes as C

import numpy as np
import pytest

from gpusteer_amd import _capi as capi
from gpusteer_amd import scenarios

pytestmark = pytest.mark.gpu


class _Hip:
    """Device buffers for the test through the HIP runtime the library links
    (libamdhip64.so.7 of the system ROCm), loaded by ctypes: the caller's side
    of the entry point allocates its inputs and outputs itself."""

    def __init__(self):
        self.lib = C.CDLL("libamdhip64.so.7")
        self.lib.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
        self.lib.hipFree.argtypes = [C.c_void_p]
        self.lib.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
        self.lib.hipDeviceSynchronize.argtypes = []
        self.bufs = []

    def alloc(self, nbytes):
        p = C.c_void_p()
        assert self.lib.hipMalloc(C.byref(p), max(4, nbytes)) == 0
        self.bufs.append(p)
        return p.value

    def upload(self, a):
        a = np.ascontiguousarray(a)
        p = self.alloc(a.nbytes)
        assert self.lib.hipMemcpy(p, a.ctypes.data, a.nbytes, 1) == 0  # host to device
        return p

    def download(self, p, shape, dtype):
        out = np.empty(shape, dtype)
        assert self.lib.hipMemcpy(out.ctypes.data, p, out.nbytes, 2) == 0  # device to host
        return out

    def fill(self, a):
        return self.upload(a)

    def close(self):
        assert self.lib.hipDeviceSynchronize() == 0
        for p in self.bufs:
            assert self.lib.hipFree(p) == 0
        self.bufs = []


def _points(n, seed, extent):
    g = np.random.default_rng(seed)
    return g.uniform(0.0, extent, size=(n, 3)).astype(np.float32)


def _with_entry_points(gs, ctx, pos, k, cells, extent):
    n = pos.shape[0]
    st = scenarios.make_agents(pos, np.tile([[0.0, 0.0, 1.0]], (n, 1)), np.zeros(n), 0.5, 1.0, 1.0)
    A = gs.AgentGroup(ctx, st)
    bins = gs.KNNBinData(ctx, cells, (extent,) * 3, origin=(0.0, 0.0, 0.0))
    knn = gs.KNNData(ctx, n, k)
    gs.updateKNNDatabase(A, bins)
    gs.findKNearestNeighbors(A, knn, bins, A, 1, 0)
    out = knn.download()
    for h in (knn, bins, A):
        h.close()
    return out


def _noalloc(gs, ctx, pos, k, cells, extent):
    n = pos.shape[0]
    hip = _Hip()
    p = [hip.upload(np.ascontiguousarray(pos[:, a])) for a in range(3)]
    # Outputs start as garbage, so every entry must be written.
    idx = hip.fill(np.full((k, n), 12345, np.uint32))
    dist = hip.fill(np.full((k, n), -1.0, np.float32))
    bins = gs.KNNBinData(ctx, cells, (extent,) * 3, origin=(0.0, 0.0, 0.0))
    rc = capi.lib.gs_find_knn_noalloc(ctx.h, bins.h, n, p[0], p[1], p[2], k, idx, dist)
    assert rc == capi.GS_OK, capi.lib.gs_last_error()
    ctx.synchronize()
    out = hip.download(idx, (k, n), np.uint32), hip.download(dist, (k, n), np.float32)
    bins.close()
    hip.close()
    return out


def test_noalloc_equals_find_knn_twice(gs, ctx):
    # First call: 40,000 points, k = 7, 24^3 cells; second: 5,000 points,
    # k = 16, 8^3 cells (fewer points, fewer cells, longer lists).
    for n, k, cells, seed in ((40000, 7, (24, 24, 24), 1), (5000, 16, (8, 8, 8), 2)):
        print(f"noalloc case n={n} k={k} cells={cells}", flush=True)
        pos = _points(n, seed, 100.0)
        want_idx, want_dist = _with_entry_points(gs, ctx, pos, k, cells, 100.0)
        idx, dist = _noalloc(gs, ctx, pos, k, cells, 100.0)
        np.testing.assert_array_equal(idx, want_idx)
        np.testing.assert_array_equal(dist, want_dist)


def test_noalloc_pads_and_limits(gs, ctx):
    # Fewer points than k + 1: every row ends in pads.
    pos = _points(5, 3, 10.0)
    want_idx, want_dist = _with_entry_points(gs, ctx, pos, 8, (2, 2, 2), 10.0)
    idx, dist = _noalloc(gs, ctx, pos, 8, (2, 2, 2), 10.0)
    assert (idx[4:] == 0xFFFFFFFF).all()
    np.testing.assert_array_equal(idx, want_idx)
    np.testing.assert_array_equal(dist, want_dist)
    # Limits are refused before any device work.
    hip = _Hip()
    ptr = hip.alloc(64)
    small = gs.KNNBinData(ctx, (4, 4, 4), (10.0,) * 3, origin=(0.0, 0.0, 0.0))
    big = gs.KNNBinData(ctx, (64, 64, 64), (10.0,) * 3, origin=(0.0, 0.0, 0.0))
    L = capi.lib
    assert L.gs_find_knn_noalloc(ctx.h, small.h, capi.GS_NOALLOC_MAX_POINTS + 1, ptr, ptr, ptr, 7, ptr,
                                 ptr) == capi.GS_ERR_UNSUPPORTED
    assert L.gs_find_knn_noalloc(ctx.h, big.h, 4, ptr, ptr, ptr, 7, ptr, ptr) == capi.GS_ERR_UNSUPPORTED
    assert L.gs_find_knn_noalloc(ctx.h, small.h, 4, ptr, ptr, ptr, capi.GS_NOALLOC_MAX_K + 1, ptr,
                                 ptr) == capi.GS_ERR_UNSUPPORTED
    assert L.gs_find_knn_noalloc(ctx.h, small.h, 4, ptr, ptr, ptr, 0, ptr, ptr) == capi.GS_ERR_UNSUPPORTED
    assert L.gs_find_knn_noalloc(ctx.h, small.h, 4, None, ptr, ptr, 7, ptr, ptr) == capi.GS_ERR_INVALID_ARGUMENT
    assert L.gs_find_knn_noalloc(None, small.h, 4, ptr, ptr, ptr, 7, ptr, ptr) == capi.GS_ERR_INVALID_HANDLE
    # n = 0 is a no-op.
    assert L.gs_find_knn_noalloc(ctx.h, small.h, 0, None, None, None, 7, None, None) == capi.GS_OK
    small.close()
    big.close()
    hip.close()
