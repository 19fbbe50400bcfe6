"""The choke-point plugin's per-step sequence (CtfGroup::update,
plugins/ChokePointNew.cpp:587-630) written against the operator API, as a
plugin would call it.  Device work only; the host never touches agent state
between steps."""
from __future__ import annotations

from . import api
from .api import MODE_GPUREF
from .scenarios import ChokeNewScenario, MultiGroupScenario

BIT_ANTI_PENETRATION_WALL = 1 << 12


class ChokePointNew:
    """World (bin lattice + split walls, CtfWorld :101-118) and one group
    (CtfGroup :212-227: KNN to itself and to the walls, the group's path)."""

    def __init__(self, ctx: api.Context, sc: ChokeNewScenario, closest=api.capi.GS_CLOSEST_CLAMPED):
        self.sc = sc
        self.closest = closest
        p = sc.params
        self.bins = api.KNNBinData(ctx, sc.world_cells, sc.world_size)
        self.walls = api.WallGroup(ctx)
        self.walls.upload(sc.walls, sc.wall_mid)
        api.updateKNNDatabase(self.walls, self.bins)
        self.group = api.AgentGroup(ctx, sc.state)
        self.knn_self = api.KNNData(ctx, sc.state.n, p.knn)
        self.knn_walls = api.KNNData(ctx, sc.state.n, p.knw)
        self.path = api.PolylinePathway(ctx, sc.path_points, sc.path_radius, sc.path_cyclic)
        self.ctx = ctx
        self._frames = {}

    def recorded(self, variant: str, steps: int) -> None:
        """`steps` steps of step / step_with_wall_update replayed from
        recorded frames (the same calls; see api.FrameCache)."""
        if variant not in self._frames:
            self._frames[variant] = api.FrameCache(self.ctx, lambda: getattr(self, variant)(1))
        self._frames[variant].run(steps)

    def step(self, steps: int = 1) -> None:
        sc, p, A = self.sc, self.sc.params, self.group
        for _ in range(steps):
            api.wrapWorld(A, sc.world_size)
            api.updateKNNDatabase(A, self.bins)
            api.clearAppliedKernels(A)
            api.findKNearestNeighbors(A, self.knn_self, self.bins, A, p.search_radius)
            api.findKNearestNeighborsWithin(A, self.knn_walls, self.bins, self.walls, sc.wall_cap, p.search_radius)
            api.steerToAvoidWalls(A, self.knn_walls, self.walls, p.min_time_to_wall, p.weight_avoid_walls,
                                  p.mask_avoid_walls)
            # separation's mask is KERNEL_ANTI_PENETRATION_WALL in the plugin (:615)
            api.steerForSeparation(A, self.knn_self, A, p.min_separation_distance, p.max_separation_distance,
                                   p.cos_max_flocking_angle, p.weight_separation, BIT_ANTI_PENETRATION_WALL,
                                   MODE_GPUREF)
            api.steerForAlignment(A, self.knn_self, A, p.min_flocking_distance, p.max_flocking_distance,
                                  p.cos_max_flocking_angle, p.weight_alignment, p.mask_alignment, MODE_GPUREF)
            api.steerForCohesion(A, self.knn_self, A, p.min_flocking_distance, p.max_flocking_distance,
                                 p.cos_max_flocking_angle, p.weight_cohesion, p.mask_cohesion, MODE_GPUREF)
            api.steerToFollowPath(A, self.path, p.path_prediction_time, p.weight_follow_path, p.mask_follow_path)
            api.steerForSeek(A, sc.seek_target, p.weight_seek, p.mask_seek, MODE_GPUREF)
            api.updateGroup(A, sc.dt, MODE_GPUREF)
            api.antiPenetrationAgents(A, self.knn_self, A, p.mask_anti_penetration_agents)

    def step_with_wall_update(self, steps: int = 1) -> None:
        """The ChokePointGroup variant (CtfEnemyGroup::update,
        plugins/ChokePointGroup.cpp:683-728): walls, path, seek, then the
        three flocking terms, the integrate with wall anti-penetration, and
        agent anti-penetration masked by AVOID_WALLS | ANTI_PENETRATION_WALL
        (:203)."""
        sc, p, A = self.sc, self.sc.params, self.group
        for _ in range(steps):
            api.wrapWorld(A, sc.world_size)
            api.updateKNNDatabase(A, self.bins)
            api.clearAppliedKernels(A)
            api.findKNearestNeighbors(A, self.knn_self, self.bins, A, p.search_radius)
            api.findKNearestNeighborsWithin(A, self.knn_walls, self.bins, self.walls, sc.wall_cap, p.search_radius)
            api.steerToAvoidWalls(A, self.knn_walls, self.walls, p.min_time_to_wall, p.weight_avoid_walls,
                                  p.mask_avoid_walls)
            api.steerToFollowPath(A, self.path, p.path_prediction_time, p.weight_follow_path, p.mask_follow_path)
            api.steerForSeek(A, sc.seek_target, p.weight_seek, p.mask_seek, MODE_GPUREF)
            api.steerForSeparation(A, self.knn_self, A, p.min_separation_distance, p.max_separation_distance,
                                   p.cos_max_flocking_angle, p.weight_separation, p.mask_separation, MODE_GPUREF)
            api.steerForAlignment(A, self.knn_self, A, p.min_flocking_distance, p.max_flocking_distance,
                                  p.cos_max_flocking_angle, p.weight_alignment, p.mask_alignment, MODE_GPUREF)
            api.steerForCohesion(A, self.knn_self, A, p.min_flocking_distance, p.max_flocking_distance,
                                 p.cos_max_flocking_angle, p.weight_cohesion, p.mask_cohesion, MODE_GPUREF)
            api.updateGroupWithAntiPenetration(A, self.knn_walls, self.walls, sc.dt, self.closest)
            api.antiPenetrationAgents(A, self.knn_self, A, (1 << 8) | BIT_ANTI_PENETRATION_WALL)

    def close(self) -> None:
        for f in self._frames.values():
            f.close()
        for h in (self.group, self.walls, self.bins, self.knn_self, self.knn_walls, self.path):
            h.close()


class MultiGroup:
    """plugins/MultiGroup.cpp: groups sharing one bin lattice, each with a KNN
    to every group and an optional path (Group :180-230).  A frame runs
    Group::preUpdate for every group (wrap, bins, clear bits, :546-556), then
    Group::update for each group in turn (:558-611): KNN to every group,
    avoid the other groups, separation from every group, alignment and
    cohesion with its own, path and seek, integrate, anti-penetration against
    every group.  A later group sees the earlier groups' new state.  The
    group searches are capped at the scenario's knn_cap (the reach of every
    behaviour that reads them)."""

    def __init__(self, ctx: api.Context, sc: MultiGroupScenario, mode=MODE_GPUREF):
        self.sc = sc
        self.mode = mode
        p = sc.params
        self.bins = api.KNNBinData(ctx, sc.world_cells, sc.world_size)
        self.groups = [api.AgentGroup(ctx, st) for st in sc.states]
        self.knn = [[api.KNNData(ctx, st.n, p.knn) for _ in sc.states] for st in sc.states]
        self.paths = [api.PolylinePathway(ctx, pts, r, c) if len(pts) >= 2 else None for pts, r, c in sc.paths]
        self.frames = api.FrameCache(ctx, lambda: self.step(1))

    def recorded(self, steps: int) -> None:
        """`steps` steps replayed from recorded frames (api.FrameCache)."""
        self.frames.run(steps)

    def step(self, steps: int = 1) -> None:
        sc, p, G, m = self.sc, self.sc.params, self.groups, self.mode
        for _ in range(steps):
            for A in G:
                api.wrapWorld(A, sc.world_size)
                api.updateKNNDatabase(A, self.bins)
                api.clearAppliedKernels(A)
            for gi, A in enumerate(G):
                K = self.knn[gi]
                for j, B in enumerate(G):
                    api.findKNearestNeighborsWithin(A, K[j], self.bins, B, sc.knn_cap, p.search_radius)
                for j, B in enumerate(G):
                    if j != gi:
                        api.steerToAvoidNeighbors(A, K[j], B, p.min_time_to_collision, p.min_separation_distance,
                                                  p.avoid_close_neighbors, p.weight_avoid_neighbors,
                                                  p.mask_avoid_neighbors, m)
                for j, B in enumerate(G):
                    api.steerForSeparation(A, K[j], B, p.min_separation_distance, p.max_separation_distance,
                                           p.cos_max_flocking_angle, p.weight_separation, p.mask_separation, m)
                api.steerForAlignment(A, K[gi], A, p.min_flocking_distance, p.max_flocking_distance,
                                      p.cos_max_flocking_angle, p.weight_alignment, p.mask_alignment, m)
                api.steerForCohesion(A, K[gi], A, p.min_flocking_distance, p.max_flocking_distance,
                                     p.cos_max_flocking_angle, p.weight_cohesion, p.mask_cohesion, m)
                P = self.paths[gi]
                if P is not None:
                    api.steerToFollowPath(A, P, p.path_prediction_time, p.weight_follow_path, p.mask_follow_path)
                    api.steerForSeek(A, tuple(float(v) for v in sc.paths[gi][0][1]), p.weight_seek, p.mask_seek, m)
                api.updateGroup(A, sc.dt, m)
                for j, B in enumerate(G):
                    api.antiPenetrationAgents(A, K[j], B, p.mask_anti_penetration_agents)

    def close(self) -> None:
        self.frames.close()
        for A in self.groups:
            A.close()
        for row in self.knn:
            for k in row:
                k.close()
        for P in self.paths:
            if P is not None:
                P.close()
        self.bins.close()
