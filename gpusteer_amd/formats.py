"""Scenario formats through the library's host-side C-ABI: wall maps
(WallGroupData::LoadFromFile, include/OpenSteer/CUDA/WallGroupData.cu:73-103),
wall splitting at bin-cell borders (WallGroupData::SplitWalls, :105-206) and
simulation parameter files (SimulationParams::load, src/Simulation.cpp:7-250).
No device work: these run wherever the library loads."""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass, field
from pathlib import Path
from typing import List, Optional, Tuple

import numpy as np

from . import _capi as capi
from ._capi import check
from .state import WALL_FIELDS, WallState

# The reference's scenario files (data/*.params, data/*.map), as its plugins
# load them.
DATA_DIR = Path(__file__).resolve().parent / "data"


def _f(a: np.ndarray):
    return a.ctypes.data_as(C.POINTER(C.c_float))


class WallSet:
    """A list of walls with their KNN positions (midpoints), as a WallGroup
    holds it on the host before SyncDevice."""

    def __init__(self, handle):
        self.h = handle

    @classmethod
    def load_map(cls, path) -> "WallSet":
        h = C.c_void_p()
        check(capi.lib.gs_wall_set_load_map(str(path).encode(), C.byref(h)), "gs_wall_set_load_map")
        return cls(h)

    @classmethod
    def from_walls(cls, walls: WallState) -> "WallSet":
        arrs = [np.ascontiguousarray(walls.arrays[f], dtype=np.float32) for f in WALL_FIELDS]
        w = capi.GsWalls(*[_f(a) for a in arrs])
        h = C.c_void_p()
        check(capi.lib.gs_wall_set_create(walls.n, C.byref(w), C.byref(h)), "gs_wall_set_create")
        return cls(h)

    def split(self, world_size, cells) -> "WallSet":
        """SplitWalls over the cells of a lattice of `cells` over a box of
        `world_size` centred at the origin (KNNBinDataV2::CreateCells)."""
        s = (C.c_float * 3)(*world_size)
        c = (C.c_uint32 * 3)(*cells)
        check(capi.lib.gs_wall_set_split(self.h, s, c), "gs_wall_set_split")
        return self

    def split_cells(self, min_bounds: np.ndarray, max_bounds: np.ndarray) -> "WallSet":
        """SplitWalls over an explicit cell list (rows of min / max bounds)."""
        lo = np.ascontiguousarray(min_bounds, dtype=np.float32).reshape(-1, 3)
        hi = np.ascontiguousarray(max_bounds, dtype=np.float32).reshape(-1, 3)
        check(capi.lib.gs_wall_set_split_cells(self.h, lo.shape[0], _f(lo), _f(hi)), "gs_wall_set_split_cells")
        return self

    def __len__(self) -> int:
        n = C.c_uint32()
        check(capi.lib.gs_wall_set_size(self.h, C.byref(n)), "gs_wall_set_size")
        return n.value

    def get(self) -> Tuple[WallState, np.ndarray]:
        """The walls and their midpoints ([n][3])."""
        n = len(self)
        arrs = {f: np.zeros(n, np.float32) for f in WALL_FIELDS}
        mid = [np.zeros(n, np.float32) for _ in range(3)]
        w = capi.GsWalls(*[_f(arrs[f]) for f in WALL_FIELDS])
        check(capi.lib.gs_wall_set_get(self.h, C.byref(w), _f(mid[0]), _f(mid[1]), _f(mid[2])), "gs_wall_set_get")
        return WallState(arrs), np.stack(mid, axis=1)

    def close(self) -> None:
        if self.h:
            capi.lib.gs_wall_set_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


@dataclass
class GroupParams:
    """GroupParams (include/OpenSteer/Simulation.h:16-48)."""
    num_agents: int = 0
    max_speed: float = 0.0
    max_force: float = 0.0
    start_position: Tuple[float, float, float] = (0.0, 0.0, 0.0)
    min_start_radius: float = 0.0
    max_start_radius: float = 0.0
    path: List[Tuple[float, float, float]] = field(default_factory=list)
    path_radius: float = 0.0
    path_cyclic: bool = False
    body_color: Tuple[float, float, float] = (0.0, 0.0, 0.0)


@dataclass
class SimParams:
    """SimulationParams (Simulation.h:66-120): world, groups and the scalar
    keys by their C names (seed, knn, ..., weight_*, mask_*)."""
    dimensions: Tuple[float, float, float]
    cells: Tuple[int, int, int]
    obstacle_count: int
    map_filename: str
    groups: List[GroupParams]
    values: dict

    def __getattr__(self, name):
        v = self.__dict__.get("values", {})
        if name in v:
            return v[name]
        raise AttributeError(name)


def load_params(path) -> SimParams:
    p = capi.GsSimParams()
    check(capi.lib.gs_sim_params_load(str(path).encode(), C.byref(p)), "gs_sim_params_load")
    groups = []
    for i in range(p.n_groups):
        g = p.groups[i]
        groups.append(GroupParams(
            num_agents=g.num_agents, max_speed=g.max_speed, max_force=g.max_force,
            start_position=tuple(g.start_position), min_start_radius=g.min_start_radius,
            max_start_radius=g.max_start_radius, path=[tuple(g.path[j]) for j in range(g.path_points)],
            path_radius=g.path_radius, path_cyclic=bool(g.path_cyclic), body_color=tuple(g.body_color)))
    values = {f: getattr(p, f) for f in ("seed", "knn", "kno", "knw", "search_radius", "avoid_close_neighbors")}
    values.update({f: getattr(p, f) for f in capi.SIM_FLOAT_FIELDS + capi.SIM_WEIGHT_FIELDS + capi.SIM_MASK_FIELDS})
    w = p.world
    return SimParams(dimensions=tuple(w.dimensions), cells=tuple(w.cells), obstacle_count=w.obstacle_count,
                     map_filename=w.map_filename.decode(), groups=groups, values=values)


def lattice_cells(world_size, cells) -> Tuple[np.ndarray, np.ndarray]:
    """Min / max bounds of KNNBinDataV2::CreateCells' cells (a lattice over a
    box centred at the origin) in its order: y outer, z, x inner."""
    step = np.float32(world_size[0]) / np.float32(cells[0]), np.float32(world_size[1]) / np.float32(cells[1]), \
        np.float32(world_size[2]) / np.float32(cells[2])
    lo, hi = [], []
    for iy in range(cells[1]):
        for iz in range(cells[2]):
            for ix in range(cells[0]):
                m = [np.float32(i) * step[a] - np.float32(0.5) * np.float32(world_size[a])
                     for a, i in enumerate((ix, iy, iz))]
                lo.append(m)
                hi.append([np.float32(m[a] + step[a]) for a in range(3)])
    return np.array(lo, np.float32), np.array(hi, np.float32)
