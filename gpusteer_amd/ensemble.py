"""Batched ensemble (gs_ensemble_*): many independent Boids flocks and
choke-point crowds stepped together, one launch per phase over all scenarios
and steps replayed from a captured hipGraph.  Also the sequential runner that
steps the same scenarios one at a time through the single-instance calls,
which is what the ensemble must equal and what it is measured against."""
from __future__ import annotations

import ctypes as C
from typing import List

import numpy as np

from . import _capi as capi
from . import api
from ._capi import check, fptr
from .state import AGENT_FIELDS, WALL_FIELDS, AgentState


def _term(t) -> capi.GsFlockTerm:
    return capi.GsFlockTerm(t.min_d, t.max_d, t.cos_max, t.weight, t.mask)


def scenario_struct(spec) -> capi.GsScenario:
    s = capi.GsScenario()
    if spec[0] == "boids":
        _, st, p = spec
        s.kind = capi.GS_SCENARIO_BOIDS
        s.n[0], s.n[1] = st.n, 0
        s.origin, s.size = (C.c_float * 3)(*p.grid_origin), (C.c_float * 3)(*p.grid_size)
        s.cells = (C.c_uint32 * 3)(*p.grid_cells)
        s.dt, s.mode = p.dt, p.mode
        s.wrap_shape = p.wrap_shape
        s.wrap_center, s.wrap_size = (C.c_float * 3)(*p.wrap_center), (C.c_float * 3)(*p.wrap_size)
        s.separation, s.alignment, s.cohesion = _term(p.separation), _term(p.alignment), _term(p.cohesion)
    else:
        _, sc = spec
        s.kind = capi.GS_SCENARIO_CHOKE
        s.n[0], s.n[1] = sc.groups[0].n, sc.groups[1].n
        s.origin, s.size = (C.c_float * 3)(*sc.grid_origin), (C.c_float * 3)(*sc.grid_size)
        s.cells = (C.c_uint32 * 3)(*sc.grid_cells)
        s.dt, s.mode = sc.dt, sc.mode
        s.wrap_shape = -1
        s.choke_separation = _term(sc.sep)
        for gi in range(2):
            s.goal[gi] = (C.c_float * 3)(*sc.goals[gi])
        s.seek_weight, s.seek_mask = sc.w_seek, sc.mask_seek
        s.walls_min_time, s.walls_weight = sc.min_time_walls, sc.w_walls
        s.obstacles_min_time, s.obstacles_weight = sc.min_time_obstacles, sc.w_obstacles
        s.k_walls, s.k_obstacles = sc.knw, sc.kno
        s.other_max_distance = sc.other_radius
    return s


def _f32_state(st: AgentState) -> AgentState:
    out = st.copy(np.float32)
    for f in AGENT_FIELDS:
        out.arrays[f] = np.ascontiguousarray(out.arrays[f], dtype=np.float32)
    out.bits = np.ascontiguousarray(out.bits, dtype=np.uint32)
    return out


class Ensemble:
    """All scenarios of `specs` in one ensemble (k neighbours per agent)."""

    def __init__(self, ctx: api.Context, specs, k: int = 7):
        self.specs = specs
        self.k = k
        arr = (capi.GsScenario * max(1, len(specs)))(*[scenario_struct(s) for s in specs])
        h = C.c_void_p()
        check(capi.lib.gs_ensemble_create(ctx.h, len(specs), arr, k, C.byref(h)), "gs_ensemble_create")
        self.h = h
        for i, spec in enumerate(specs):
            groups = [spec[1]] if spec[0] == "boids" else list(spec[1].groups)
            for gi, st in enumerate(groups):
                self.upload(i, gi, st)
            if spec[0] == "choke":
                sc = spec[1]
                w = {f: np.ascontiguousarray(sc.walls.arrays[f], dtype=np.float32) for f in WALL_FIELDS}
                ws = capi.GsWalls(*[fptr(w[f]) for f in WALL_FIELDS])
                if sc.wall_mid is None:
                    check(capi.lib.gs_ensemble_upload_walls(self.h, i, sc.walls.n, C.byref(ws)),
                          "gs_ensemble_upload_walls")
                else:
                    m = np.asarray(sc.wall_mid, dtype=np.float32).reshape(sc.walls.n, 3)
                    cols = [np.ascontiguousarray(m[:, a]) for a in range(3)]
                    check(capi.lib.gs_ensemble_upload_walls_mid(self.h, i, sc.walls.n, C.byref(ws),
                                                                *[fptr(c) for c in cols]),
                          "gs_ensemble_upload_walls_mid")
                o = [np.ascontiguousarray(a, dtype=np.float32) for a in sc.obstacles]
                check(capi.lib.gs_ensemble_upload_obstacles(self.h, i, o[0].shape[0], *[fptr(a) for a in o]),
                      "gs_ensemble_upload_obstacles")

    @property
    def slots(self) -> int:
        v = C.c_uint32()
        check(capi.lib.gs_ensemble_slots(self.h, C.byref(v)), "gs_ensemble_slots")
        return v.value

    def upload(self, scenario: int, group: int, st: AgentState) -> None:
        s = _f32_state(st)
        check(capi.lib.gs_ensemble_upload_agents(self.h, scenario, group, C.byref(api._agents_struct(s))),
              "gs_ensemble_upload_agents")

    def download(self, scenario: int, group: int = 0) -> AgentState:
        spec = self.specs[scenario]
        n = spec[1].n if spec[0] == "boids" else spec[1].groups[group].n
        out = AgentState.empty(n)
        check(capi.lib.gs_ensemble_download_agents(self.h, scenario, group, C.byref(api._agents_struct(out))),
              "gs_ensemble_download_agents")
        return out

    def download_knn(self, scenario: int, group: int = 0) -> np.ndarray:
        spec = self.specs[scenario]
        n = spec[1].n if spec[0] == "boids" else spec[1].groups[group].n
        idx = np.zeros((self.k, n), np.uint32)
        check(capi.lib.gs_ensemble_download_knn(self.h, scenario, group, capi.uptr(idx)), "gs_ensemble_download_knn")
        return idx

    def step(self, steps: int = 1, eager: bool = False) -> None:
        check(capi.lib.gs_ensemble_step(self.h, steps, capi.GS_ENSEMBLE_EAGER if eager else capi.GS_ENSEMBLE_GRAPH),
              "gs_ensemble_step")

    def agents(self) -> int:
        return sum(s[1].n if s[0] == "boids" else s[1].groups[0].n + s[1].groups[1].n for s in self.specs)

    def close(self) -> None:
        if self.h:
            check(capi.lib.gs_ensemble_destroy(self.h), "gs_ensemble_destroy")
            self.h = None


def choke_step_calls(G, knns, W, O, bins, sc) -> None:
    """One choke-point step through the single-instance calls, in the order of
    plugins/ChokePointGroup.cpp:683-728: bin both groups; per group clear
    bits, KNN to itself / the other group (within sc.other_radius) / the
    walls / the obstacles, avoid
    walls, avoid obstacles, separation from both groups, seek; integrate both."""
    for grp in G:
        api.updateKNNDatabase(grp, bins)
    for gi, grp in enumerate(G):
        other = G[1 - gi]
        ks, ko, kw, kob = knns[gi]
        api.clearAppliedKernels(grp)
        api.findKNearestNeighbors(grp, ks, bins, grp, 1)
        api.findKNearestNeighborsWithin(grp, ko, bins, other, sc.other_radius, 1)
        api.findKNearestNeighbors(grp, kw, bins, W, 1)
        api.findKNearestNeighbors(grp, kob, bins, O, 1)
        api.steerToAvoidWalls(grp, kw, W, sc.min_time_walls, sc.w_walls, 0)
        api.steerToAvoidObstacles(grp, O, kob, sc.min_time_obstacles, sc.w_obstacles, 0)
        t = sc.sep
        api.steerForSeparation(grp, ks, grp, t.min_d, t.max_d, t.cos_max, t.weight, t.mask, sc.mode)
        api.steerForSeparation(grp, ko, other, t.min_d, t.max_d, t.cos_max, t.weight, t.mask, sc.mode)
        api.steerForSeek(grp, sc.goals[gi], sc.w_seek, sc.mask_seek, sc.mode)
    for grp in G:
        api.updateGroup(grp, sc.dt, sc.mode)


def step_choke(G, knns, W, O, bins, sc, steps: int = 1) -> None:
    """choke_step_calls `steps` times, issued by the library
    (gs_step_choke): one call instead of about thirty per step."""
    hs = (C.c_void_p * 8)(*[k.h.value if isinstance(k.h, C.c_void_p) else k.h for ks in knns for k in ks])
    s = scenario_struct(("choke", sc))
    check(capi.lib.gs_step_choke(G[0].h, G[1].h, W.h, O.h, bins.h, hs, C.byref(s), steps), "gs_step_choke")


class Sequential:
    """The same scenarios, each stepped alone through the single-instance
    calls (gs_step_boids for a flock, gs_step_choke -- the choke call
    sequence issued by the library -- for a crowd), one scenario after
    another."""

    def __init__(self, ctx: api.Context, specs, k: int = 7):
        self.specs = specs
        self.items = []
        for spec in specs:
            if spec[0] == "boids":
                _, st, p = spec
                A = api.AgentGroup(ctx, st)
                bins = api.KNNBinData(ctx, p.grid_cells, p.grid_size, origin=p.grid_origin)
                knn = api.KNNData(ctx, st.n, k)
                self.items.append(("boids", A, bins, knn, p))
            else:
                sc = spec[1]
                bins = api.KNNBinData(ctx, sc.grid_cells, sc.grid_size, origin=sc.grid_origin)
                G = [api.AgentGroup(ctx, g) for g in sc.groups]
                W = api.WallGroup(ctx)
                W.upload(sc.walls, sc.wall_mid)
                O = api.ObstacleGroup(ctx, *sc.obstacles)
                api.updateKNNDatabase(W, bins)
                api.updateKNNDatabase(O, bins)
                knns = [[api.KNNData(ctx, g.n, kk) for kk in (k, k, sc.knw, sc.kno)] for g in sc.groups]
                self.items.append(("choke", G, knns, W, O, bins, sc))

    def step(self, steps: int = 1) -> None:
        for it in self.items:
            if it[0] == "boids":
                _, A, bins, knn, p = it
                api.stepBoids(A, knn, bins, p, steps)
            else:
                _, G, knns, W, O, bins, sc = it
                step_choke(G, knns, W, O, bins, sc, steps)

    def upload(self, scenario: int, group: int, st: AgentState) -> None:
        it = self.items[scenario]
        (it[1] if it[0] == "boids" else it[1][group]).upload(st)

    def download(self, scenario: int, group: int = 0) -> AgentState:
        it = self.items[scenario]
        return it[1].download() if it[0] == "boids" else it[1][group].download()

    def download_knn(self, scenario: int, group: int = 0) -> np.ndarray:
        it = self.items[scenario]
        return (it[3] if it[0] == "boids" else it[2][group][0]).download()[0]

    def close(self) -> None:
        for it in self.items:
            if it[0] == "boids":
                for h in it[1:4]:
                    h.close()
            else:
                _, G, knns, W, O, bins, sc = it
                for h in G + [W, O, bins] + [x for ks in knns for x in ks]:
                    h.close()
        self.items = []


class Calibration:
    """gs_calib_*: loss and its gradient with respect to the 14 parameters
    (scenarios.PARAM_NAMES x 2 groups) over the crowd scenarios of an
    ensemble, from the ensemble's current state."""

    def __init__(self, ens: Ensemble, config):
        c = capi.GsCalibConfig()
        c.separation, c.alignment, c.cohesion = _term(config.separation), _term(config.alignment), _term(
            config.cohesion)
        c.seek_weight, c.seek_mask = config.seek_weight, config.seek_mask
        c.walls_min_time, c.obstacles_min_time = config.walls_min_time, config.obstacles_min_time
        c.max_force, c.mass, c.dt = config.max_force, config.mass, config.dt
        c.penetration_weight = config.penetration_weight
        c.steps, c.checkpoint_every = config.steps, config.checkpoint_every
        h = C.c_void_p()
        check(capi.lib.gs_calib_create(ens.h, C.byref(c), C.byref(h)), "gs_calib_create")
        self.h = h
        self.ens = ens
        self.config = config

    @staticmethod
    def _params(theta) -> capi.GsCalibParams:
        p = capi.GsCalibParams()
        v = np.asarray(theta, dtype=np.float32)
        assert v.shape == (14,)
        C.memmove(C.byref(p), v.ctypes.data, 14 * 4)
        return p

    def forward(self, theta) -> float:
        loss = C.c_double()
        check(capi.lib.gs_calib_forward(self.h, C.byref(self._params(theta)), C.byref(loss)), "gs_calib_forward")
        return loss.value

    def backward(self) -> np.ndarray:
        g = capi.GsCalibParams()
        check(capi.lib.gs_calib_backward(self.h, C.byref(g)), "gs_calib_backward")
        out = np.zeros(14, np.float32)
        C.memmove(out.ctypes.data, C.byref(g), 14 * 4)
        return out.astype(np.float64)

    def timing(self):
        f, b = C.c_float(), C.c_float()
        check(capi.lib.gs_calib_timing(self.h, C.byref(f), C.byref(b)), "gs_calib_timing")
        return f.value, b.value

    def final_state(self, scenario: int, group: int) -> AgentState:
        sc = self.ens.specs[scenario][1]
        out = AgentState.empty(sc.groups[group].n)
        check(capi.lib.gs_calib_download_state(self.h, scenario, group, C.byref(api._agents_struct(out))),
              "gs_calib_download_state")
        return out

    def close(self) -> None:
        if self.h:
            check(capi.lib.gs_calib_destroy(self.h), "gs_calib_destroy")
            self.h = None
