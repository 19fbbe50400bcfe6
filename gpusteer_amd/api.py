"""Python host mirror of the reference's operator / group API on top of the
C-ABI.

Names and argument order follow include/OpenSteer/CUDA/GroupSteerLibrary.cuh
(the static free functions plugins call every frame) and the group classes
(include/OpenSteer/AgentGroup.h, ObstacleGroup.h, WallGroup.h,
CUDA/KNNData.cuh, CUDA/KNNBinDataV2.cuh).  Errors raise GpuSteerError with
the library's status code; nothing falls back to a CPU path.
"""
from __future__ import annotations

import ctypes as C
from typing import Optional, Sequence, Tuple

import numpy as np

from . import _capi as capi
from ._capi import check, f3, fptr, u3, uptr
from .state import AGENT_FIELDS, WALL_FIELDS, AgentState, WallState

MODE_OPENSTEER = capi.GS_MODE_OPENSTEER
MODE_GPUREF = capi.GS_MODE_GPUREF
# findKNearestNeighbors flags: no seeding from the container's previous
# result; always search through the bins (never an exhaustive scan).
KNN_DEFAULT = capi.GS_KNN_DEFAULT
KNN_NO_SEED = capi.GS_KNN_NO_SEED
KNN_BINNED = capi.GS_KNN_BINNED


class Context:
    """Device + stream.  stream: a hipStream_t (int) or None for an owned one."""

    def __init__(self, device: int = 0, stream: Optional[int] = None):
        h = C.c_void_p()
        check(capi.lib.gs_context_create(device, C.c_void_p(stream) if stream else None, C.byref(h)),
              "gs_context_create")
        self.h = h

    def synchronize(self) -> None:
        check(capi.lib.gs_context_synchronize(self.h), "gs_context_synchronize")

    def stream(self) -> int:
        s = C.c_void_p()
        check(capi.lib.gs_context_stream(self.h, C.byref(s)), "gs_context_stream")
        return s.value or 0

    def set_shape(self, kernel: int, block_threads: int, agents_per_lane: int) -> None:
        """Launch shape of the k = 7 search (capi.GS_SHAPE_KNN) or flocking
        kernel (capi.GS_SHAPE_FLOCK): 128 / 256 / 512 threads per block,
        1 or 2 agents per lane."""
        check(capi.lib.gs_context_set_shape(self.h, kernel, block_threads, agents_per_lane), "gs_context_set_shape")

    def shape(self, kernel: int) -> Tuple[int, int]:
        b, a = C.c_uint32(), C.c_uint32()
        check(capi.lib.gs_context_get_shape(self.h, kernel, C.byref(b), C.byref(a)), "gs_context_get_shape")
        return b.value, a.value

    def record(self, calls) -> "Frame":
        """Records the library calls that calls() issues on this context into
        a frame (gs_frame_begin / gs_frame_end): nothing runs, and every
        handle they touch is left as it was.  Frame.launch(times) replays
        them."""
        check(capi.lib.gs_frame_begin(self.h), "gs_frame_begin")
        try:
            calls()
        except BaseException:
            f = C.c_void_p()
            if capi.lib.gs_frame_end(self.h, C.byref(f)) == 0:
                capi.lib.gs_frame_destroy(f)
            raise
        f = C.c_void_p()
        check(capi.lib.gs_frame_end(self.h, C.byref(f)), "gs_frame_end")
        return Frame(f)

    def close(self) -> None:
        if self.h:
            check(capi.lib.gs_context_destroy(self.h), "gs_context_destroy")
            self.h = None

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()


class FrameCache:
    """Runs a step function (library calls only) from recorded frames.  The
    first per_frame steps run unrecorded, so every device buffer has its
    size; then per_frame steps are recorded as one frame and replayed.  A
    step leaves double buffers swapped, so after an odd number of unrecorded
    steps the handles are in the other state: a frame is recorded for each
    state met (GS_ERR_FRAME from a replay means "not this state")."""

    def __init__(self, ctx: "Context", step, per_frame: int = 2):
        self.ctx, self.step, self.per_frame = ctx, step, int(per_frame)
        self.frames: list = []
        self.warm = 0

    def _launch(self, times: int) -> bool:
        for f in self.frames:
            try:
                f.launch(1)
            except capi.GpuSteerError as e:
                if e.status != capi.GS_ERR_FRAME:
                    raise
                continue
            if times > 1:
                f.launch(times - 1)
            return True
        return False

    def run(self, steps: int) -> None:
        n = int(steps)
        while n > 0 and self.warm < self.per_frame:
            self.step()
            self.warm += 1
            n -= 1
        if n >= self.per_frame:
            k = n // self.per_frame
            if not self._launch(k):
                self.frames.append(self.ctx.record(lambda: [self.step() for _ in range(self.per_frame)]))
                if not self._launch(k):
                    raise RuntimeError("a frame recorded in the current state does not replay in it")
            n -= k * self.per_frame
        for _ in range(n):
            self.step()

    def close(self) -> None:
        for f in self.frames:
            f.close()
        self.frames = []


class Frame:
    """A recorded call sequence (gs_frame_*): launch(times) replays it on the
    context's stream, with the same results as issuing the calls."""

    def __init__(self, h):
        self.h = h

    def launch(self, times: int = 1) -> None:
        check(capi.lib.gs_frame_launch(self.h, int(times)), "gs_frame_launch")

    def close(self) -> None:
        if self.h:
            check(capi.lib.gs_frame_destroy(self.h), "gs_frame_destroy")
            self.h = None

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()


class KNNBinData:
    """Bin lattice (KNNBinDataV2, CUDA/KNNBinDataV2.cuh:43-61).  With
    origin=None the world is centred at 0 as in the reference's constructor."""

    def __init__(self, ctx: Context, world_cells: Sequence[int], world_size: Sequence[float],
                 origin: Optional[Sequence[float]] = None):
        h = C.c_void_p()
        if origin is None:
            check(capi.lib.gs_bins_create_centered(ctx.h, u3(world_cells), f3(world_size), C.byref(h)),
                  "gs_bins_create_centered")
        else:
            check(capi.lib.gs_bins_create(ctx.h, f3(origin), f3(world_size), u3(world_cells), C.byref(h)),
                  "gs_bins_create")
        self.h = h
        self.ctx = ctx
        self.cells = tuple(int(c) for c in world_cells)

    @property
    def num_cells(self) -> int:
        return self.cells[0] * self.cells[1] * self.cells[2]

    def close(self):
        if self.h:
            check(capi.lib.gs_bins_destroy(self.h), "gs_bins_destroy")
            self.h = None


class _Group:
    kind = -1

    def __init__(self, ctx: Context):
        h = C.c_void_p()
        check(capi.lib.gs_group_create(ctx.h, self.kind, C.byref(h)), "gs_group_create")
        self.h = h
        self.ctx = ctx

    def Size(self) -> int:
        n = C.c_uint32()
        check(capi.lib.gs_group_size(self.h, C.byref(n)), "gs_group_size")
        return n.value

    def download_bins(self, num_cells: int):
        """(key[n], start[C+1], perm[n], sorted positions [n,3]) of the bin
        database (KNNDatabase, CUDA/KNNDatabase.cuh:30-38)."""
        n = self.Size()
        key = np.zeros(n, np.uint32)
        start = np.zeros(num_cells + 1, np.uint32)
        perm = np.zeros(n, np.uint32)
        sp = [np.zeros(n, np.float32) for _ in range(3)]
        check(capi.lib.gs_group_download_bins(self.h, uptr(key), uptr(start), uptr(perm), fptr(sp[0]),
                                              fptr(sp[1]), fptr(sp[2])), "gs_group_download_bins")
        return key, start, perm, np.stack(sp, axis=1)

    def close(self):
        if self.h:
            check(capi.lib.gs_group_destroy(self.h), "gs_group_destroy")
            self.h = None


def _agents_struct(state: AgentState) -> capi.GsAgents:
    s = capi.GsAgents()
    for f in AGENT_FIELDS:
        a = state.arrays[f]
        setattr(s, f, fptr(a))
    s.bits = uptr(state.bits)
    return s


class AgentGroup(_Group):
    """AgentGroup (include/OpenSteer/AgentGroup.h:12-56) as device SoA arrays.
    upload() replaces AddAgent* + SyncDevice(); download() is SyncHost()."""

    kind = capi.GS_KIND_AGENTS

    def __init__(self, ctx: Context, state: Optional[AgentState] = None):
        super().__init__(ctx)
        if state is not None:
            self.upload(state)

    def upload(self, state: AgentState) -> None:
        st = state if state.arrays["px"].dtype == np.float32 else state.copy(np.float32)
        for f in AGENT_FIELDS:
            st.arrays[f] = np.ascontiguousarray(st.arrays[f], dtype=np.float32)
        st.bits = np.ascontiguousarray(st.bits, dtype=np.uint32)
        s = _agents_struct(st)
        check(capi.lib.gs_group_upload_agents(self.h, st.n, C.byref(s)), "gs_group_upload_agents")

    def download(self) -> AgentState:
        out = AgentState.empty(self.Size())
        s = _agents_struct(out)
        check(capi.lib.gs_group_download_agents(self.h, C.byref(s)), "gs_group_download_agents")
        return out

    def device_arrays(self) -> dict:
        s = capi.GsAgents()
        check(capi.lib.gs_group_device_agents(self.h, C.byref(s)), "gs_group_device_agents")
        d = {f: C.cast(getattr(s, f), C.c_void_p).value for f in AGENT_FIELDS}
        d["bits"] = C.cast(s.bits, C.c_void_p).value
        return d

    def cell_density(self, num_cells: int) -> np.ndarray:
        """Agents per cell of the last bin build (ComputeCellDensity,
        CUDA/ProcessData.cu:52-80), KNNBinDataV2 cell numbering."""
        out = np.zeros(num_cells, np.uint32)
        check(capi.lib.gs_group_cell_density(self.h, uptr(out)), "gs_group_cell_density")
        return out

    def cell_velocity(self, num_cells: int) -> np.ndarray:
        """Mean velocity per cell ([C][3], ComputeAvgCellVelocity,
        CUDA/ProcessData.cu:83-129)."""
        out = np.zeros((num_cells, 3), np.float32)
        check(capi.lib.gs_group_cell_velocity(self.h, fptr(out)), "gs_group_cell_velocity")
        return out

    def write_cell_density(self, prefix, selected=(), selected_prefix="Frames/cell_density",
                           all_rows: bool = False) -> None:
        """WriteCellDensity (src/ExtractData.cpp:19-88)."""
        sel = np.ascontiguousarray(selected, dtype=np.uint32)
        check(capi.lib.gs_write_cell_density(self.h, str(prefix).encode(), uptr(sel), sel.shape[0],
                                             str(selected_prefix).encode(), int(all_rows)), "gs_write_cell_density")

    def write_cell_velocity(self, prefix, selected=(), selected_prefix="Frames/cell_avg_velocity",
                            all_rows: bool = False) -> None:
        """WriteAvgCellVelocity (src/ExtractData.cpp:90-140)."""
        sel = np.ascontiguousarray(selected, dtype=np.uint32)
        check(capi.lib.gs_write_cell_velocity(self.h, str(prefix).encode(), uptr(sel), sel.shape[0],
                                              str(selected_prefix).encode(), int(all_rows)), "gs_write_cell_velocity")

    def check_finite(self) -> Tuple[int, int]:
        bad = C.c_uint32()
        st = capi.lib.gs_group_check_finite(self.h, C.byref(bad))
        return st, bad.value


class ObstacleGroup(_Group):
    """ObstacleGroup (include/OpenSteer/ObstacleGroup.h:10-43): spheres."""

    kind = capi.GS_KIND_OBSTACLES

    def __init__(self, ctx: Context, cx=None, cy=None, cz=None, radius=None):
        super().__init__(ctx)
        if cx is not None:
            self.upload(cx, cy, cz, radius)

    def upload(self, cx, cy, cz, radius) -> None:
        arrs = [np.ascontiguousarray(a, dtype=np.float32) for a in (cx, cy, cz, radius)]
        check(capi.lib.gs_group_upload_obstacles(self.h, arrs[0].shape[0], *[fptr(a) for a in arrs]),
              "gs_group_upload_obstacles")


class WallGroup(_Group):
    """WallGroup (include/OpenSteer/WallGroup.h:12-37): line segments, KNN on
    their midpoints."""

    kind = capi.GS_KIND_WALLS

    def __init__(self, ctx: Context, walls: Optional[WallState] = None):
        super().__init__(ctx)
        if walls is not None:
            self.upload(walls)

    def upload(self, walls: WallState, mid: Optional[np.ndarray] = None) -> None:
        """Walls with their KNN positions: 0.5f * (start + end) unless `mid`
        ([n][3]) gives them, as split walls carry (WallGroupData.cu:185-190)."""
        arrs = {f: np.ascontiguousarray(walls.arrays[f], dtype=np.float32) for f in WALL_FIELDS}
        s = capi.GsWalls(**{f: fptr(arrs[f]) for f in WALL_FIELDS})
        if mid is None:
            check(capi.lib.gs_group_upload_walls(self.h, walls.n, C.byref(s)), "gs_group_upload_walls")
            return
        m = np.asarray(mid, dtype=np.float32).reshape(walls.n, 3)
        cols = [np.ascontiguousarray(m[:, a]) for a in range(3)]
        check(capi.lib.gs_group_upload_walls_mid(self.h, walls.n, C.byref(s), *[fptr(c) for c in cols]),
              "gs_group_upload_walls_mid")

    def upload_wall_set(self, ws) -> None:
        """The walls of a formats.WallSet (WallGroup::SyncDevice after
        LoadFromFile / SplitWalls)."""
        check(capi.lib.gs_group_upload_wall_set(self.h, ws.h), "gs_group_upload_wall_set")


class KNNData:
    """KNN container (CUDA/KNNData.cuh:10-55): slot-major [k][n]."""

    def __init__(self, ctx: Context, n: int, k: int):
        h = C.c_void_p()
        check(capi.lib.gs_knn_create(ctx.h, n, k, C.byref(h)), "gs_knn_create")
        self.h, self.n, self.k = h, n, k

    def download(self) -> Tuple[np.ndarray, np.ndarray]:
        idx = np.zeros((self.k, self.n), np.uint32)
        dist = np.zeros((self.k, self.n), np.float32)
        check(capi.lib.gs_knn_download(self.h, uptr(idx), fptr(dist)), "gs_knn_download")
        return idx, dist

    def getAgentData(self, i: int):
        idx, dist = self.download()
        return idx[:, i].copy(), dist[:, i].copy()

    def close(self):
        if self.h:
            check(capi.lib.gs_knn_destroy(self.h), "gs_knn_destroy")
            self.h = None


# ---------------------------------------------------------------------------
# Operator API (GroupSteerLibrary.cuh).


def updateKNNDatabase(group: _Group, bins: KNNBinData) -> None:
    check(capi.lib.gs_update_knn_database(group.h, bins.h), "gs_update_knn_database")


def findKNearestNeighbors(A: _Group, knn: KNNData, bins: KNNBinData, B: _Group, searchRadius: int = 1,
                          flags: int = capi.GS_KNN_DEFAULT) -> None:
    check(capi.lib.gs_find_knn(A.h, knn.h, bins.h, B.h, searchRadius, flags), "gs_find_knn")


def findKNearestNeighborsWithin(A: _Group, knn: KNNData, bins: KNNBinData, B: _Group, maxDistance: float,
                                searchRadius: int = 1, flags: int = capi.GS_KNN_DEFAULT) -> None:
    """findKNearestNeighbors keeping only points closer than maxDistance."""
    check(capi.lib.gs_find_knn_within(A.h, knn.h, bins.h, B.h, searchRadius, flags, maxDistance),
          "gs_find_knn_within")


def findKNearestNeighborsBruteForce(A: _Group, knn: KNNData, B: _Group) -> None:
    check(capi.lib.gs_find_knn_brute(A.h, knn.h, B.h), "gs_find_knn_brute")


def steerForSeparation(A, knn, B, minDistance, maxDistance, cosMaxAngle, fWeight, doNotApplyWith=0,
                       mode=MODE_OPENSTEER):
    check(capi.lib.gs_steer_separation(A.h, knn.h, B.h, minDistance, maxDistance, cosMaxAngle, fWeight,
                                       doNotApplyWith, mode), "gs_steer_separation")


def steerForAlignment(A, knn, B, minDistance, maxDistance, cosMaxAngle, fWeight, doNotApplyWith=0,
                      mode=MODE_OPENSTEER):
    check(capi.lib.gs_steer_alignment(A.h, knn.h, B.h, minDistance, maxDistance, cosMaxAngle, fWeight,
                                      doNotApplyWith, mode), "gs_steer_alignment")


def steerForCohesion(A, knn, B, minDistance, maxDistance, cosMaxAngle, fWeight, doNotApplyWith=0,
                     mode=MODE_OPENSTEER):
    check(capi.lib.gs_steer_cohesion(A.h, knn.h, B.h, minDistance, maxDistance, cosMaxAngle, fWeight,
                                     doNotApplyWith, mode), "gs_steer_cohesion")


def _term(t) -> capi.GsFlockTerm:
    return capi.GsFlockTerm(t.min_d, t.max_d, t.cos_max, t.weight, t.mask)


def steerForFlocking(A, knn, B, separation, alignment, cohesion, mode=MODE_OPENSTEER):
    """The three terms fused (order separation, alignment, cohesion); any term
    may be None."""
    ts = [None if t is None else C.byref(_term(t)) for t in (separation, alignment, cohesion)]
    check(capi.lib.gs_steer_flocking(A.h, knn.h, B.h, *ts, mode), "gs_steer_flocking")


def steerForSeek(A, target, fWeight, doNotApplyWith=0, mode=MODE_OPENSTEER):
    check(capi.lib.gs_steer_seek(A.h, f3(target), fWeight, doNotApplyWith, mode), "gs_steer_seek")


def steerToAvoidObstacles(A, obstacles, knn, fMinTimeToCollision, fWeight, doNotApplyWith=0):
    check(capi.lib.gs_steer_avoid_obstacles(A.h, obstacles.h, knn.h, fMinTimeToCollision, fWeight,
                                            doNotApplyWith), "gs_steer_avoid_obstacles")


def steerToAvoidWalls(A, knn, walls, fMinTimeToCollision, fWeight, doNotApplyWith=0):
    check(capi.lib.gs_steer_avoid_walls(A.h, knn.h, walls.h, fMinTimeToCollision, fWeight, doNotApplyWith),
          "gs_steer_avoid_walls")


def updateGroup(A, elapsedTime, mode=MODE_OPENSTEER):
    check(capi.lib.gs_update_group(A.h, elapsedTime, mode), "gs_update_group")


def updateGroupWithAntiPenetration(A, knn, walls, elapsedTime, closest=capi.GS_CLOSEST_CLAMPED):
    """updateGroupWithAntiPenetration (GroupSteerLibrary.cuh:198): the GPU
    integrate with wall anti-penetration over A's KNN to the walls; bits kept.
    closest=GS_CLOSEST_AS_REFERENCE reproduces the reference's closest point
    (always the wall's end, UpdateCUDAKernelNew.cu:79)."""
    check(capi.lib.gs_update_group_anti_penetration(A.h, knn.h, walls.h, elapsedTime, closest),
          "gs_update_group_anti_penetration")


def antiPenetrationAgents(A, knn, other, doNotApplyWith=0):
    """antiPenetrationAgents (GroupSteerLibrary.cuh:216): push apart from the
    nearest agent of `other` (the first entry of A's KNN to it)."""
    check(capi.lib.gs_anti_penetration_agents(A.h, knn.h, other.h, doNotApplyWith), "gs_anti_penetration_agents")


class PolylinePathway:
    """PolylinePathwayCUDA (CUDA/PolylinePathwayCUDA.cuh:11-68): points,
    tube radius, cyclic flag; host copies of the derived arrays."""

    def __init__(self, ctx: Context, points, radius: float, cyclic: bool = False):
        pts = np.ascontiguousarray(np.asarray(points, dtype=np.float32).reshape(-1, 3))
        h = C.c_void_p()
        check(capi.lib.gs_path_create(ctx.h, pts.shape[0], fptr(pts), radius, int(bool(cyclic)), C.byref(h)),
              "gs_path_create")
        self.h = h
        m, total = C.c_uint32(), C.c_float()
        check(capi.lib.gs_path_info(h, C.byref(m), C.byref(total)), "gs_path_info")
        self.m, self.total, self.radius, self.cyclic = m.value, total.value, np.float32(radius), bool(cyclic)

    def host_arrays(self) -> Tuple[np.ndarray, np.ndarray, np.ndarray]:
        """(points[m][3], unit segment directions[m][3], lengths[m]) =
        hvPoints / hvNormals / hvLengths."""
        p = np.zeros((self.m, 3), np.float32)
        d = np.zeros((self.m, 3), np.float32)
        L = np.zeros(self.m, np.float32)
        check(capi.lib.gs_path_get(self.h, fptr(p), fptr(d), fptr(L)), "gs_path_get")
        return p, d, L

    def close(self) -> None:
        if self.h:
            capi.lib.gs_path_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def steerToFollowPath(A, path: PolylinePathway, predictionTime, fWeight, doNotApplyWith=0):
    check(capi.lib.gs_steer_follow_path(A.h, path.h, predictionTime, fWeight, doNotApplyWith),
          "gs_steer_follow_path")


def steerToAvoidNeighbors(A, knn, other, fMinTimeToCollision, fMinSeparationDistance, bAvoidCloseNeighbors,
                          fWeight, doNotApplyWith=0, mode=MODE_OPENSTEER):
    """steerToAvoidNeighbors (GroupSteerLibrary.cuh:111) over A's KNN in
    `other`."""
    check(capi.lib.gs_steer_avoid_neighbors(A.h, knn.h, other.h, fMinTimeToCollision, fMinSeparationDistance,
                                            int(bool(bAvoidCloseNeighbors)), fWeight, doNotApplyWith, mode),
          "gs_steer_avoid_neighbors")


def wrapWorld(A, worldSize=None, shape=capi.GS_WRAP_BOX, center=(0.0, 0.0, 0.0), radius=None):
    """Box: wrapWorld(A, worldSize) of the reference (centred box).  Sphere:
    shape=GS_WRAP_SPHERE, radius=R."""
    size = (radius, 0.0, 0.0) if shape == capi.GS_WRAP_SPHERE else worldSize
    check(capi.lib.gs_wrap_world(A.h, shape, f3(center), f3(size)), "gs_wrap_world")


def clearAppliedKernels(A):
    check(capi.lib.gs_clear_bits(A.h), "gs_clear_bits")


def boids_params_struct(p) -> capi.GsBoidsParams:
    s = capi.GsBoidsParams()
    s.wrap_shape = p.wrap_shape
    s.wrap_center = (C.c_float * 3)(*p.wrap_center)
    s.wrap_size = (C.c_float * 3)(*p.wrap_size)
    s.search_radius = p.search_radius
    s.knn_flags = 0
    s.separation, s.alignment, s.cohesion = _term(p.separation), _term(p.alignment), _term(p.cohesion)
    s.dt = p.dt
    s.mode = p.mode
    return s


def stepBoids(A: AgentGroup, knn: KNNData, bins: KNNBinData, params, steps: int = 1, knn_flags: int = 0) -> None:
    """Fused wrap -> bin -> KNN -> flock -> integrate, `steps` times."""
    s = boids_params_struct(params)
    s.knn_flags = knn_flags
    check(capi.lib.gs_step_boids(A.h, knn.h, bins.h, C.byref(s), steps), "gs_step_boids")


def stepBoidsTimed(A: AgentGroup, knn: KNNData, bins: KNNBinData, params, steps: int = 1, knn_flags: int = 0):
    """stepBoids with HIP-event timing on the library's stream; returns
    (bin_build_ms, knn_ms, steer_integrate_ms, step_ms) averaged per step."""
    s = boids_params_struct(params)
    s.knn_flags = knn_flags
    t = capi.GsBoidsTiming()
    check(capi.lib.gs_step_boids_timed(A.h, knn.h, bins.h, C.byref(s), steps, C.byref(t)), "gs_step_boids_timed")
    return t.bin_build_ms, t.knn_ms, t.steer_integrate_ms, t.step_ms
