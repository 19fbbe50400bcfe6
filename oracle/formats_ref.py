"""Oracle (test infrastructure only; never imported by the product): a plain
Python / numpy-float32 restatement of the scenario formats, written from the
reference's behaviour, against which the library's host-side C-ABI is
checked.

* load_map: WallGroupData::LoadFromFile (include/OpenSteer/CUDA/WallGroupData.cu:73-103):
  whitespace-separated records of 9 floats (start, end, normal); a record is
  kept only if all 9 values were read and the last one did not end the file
  (the reference's istream loop keeps a record while the stream is good);
  mid = 0.5f * (start + end).
* lattice_cells: KNNBinDataV2::CreateCells (CUDA/KNNBinDataV2.cu:73-136):
  step = size / cells; min = i * step - 0.5 * size; max = min + step, in the
  order y (outer), z, x (inner).
* split_walls: WallGroupData::SplitWalls (WallGroupData.cu:105-206) with the
  slab test WallGroupData::intersects (:28-71) and EPSILON = 0.001
  (include/OpenSteer/VectorUtils.cuh:8).
* load_params: SimulationParams::load (src/Simulation.cpp:7-250).

Every float operation is a numpy float32 scalar operation (IEEE single
rounding, no fused multiply-add), as the reference's host code.
"""
from __future__ import annotations

import re
from typing import Dict, List, Tuple

import numpy as np

F = np.float32
EPS = F(0.001)


def load_map(path) -> List[Tuple[Tuple, Tuple, Tuple, Tuple]]:
    text = open(path, "rb").read().decode()
    out = []
    pos = 0
    num = re.compile(r"\s*([-+]?(?:\d+\.?\d*|\.\d+)(?:[eE][-+]?\d+)?)")
    while True:
        vals = []
        at_eof = False
        for j in range(9):
            m = num.match(text, pos)
            if not m:
                break
            vals.append(F(float(m.group(1))))
            pos = m.end()
            if j == 8:
                at_eof = pos >= len(text)
        if len(vals) < 9 or at_eof:
            break
        s, e, n = tuple(vals[0:3]), tuple(vals[3:6]), tuple(vals[6:9])
        mid = tuple(F(0.5) * (s[a] + e[a]) for a in range(3))
        out.append((s, e, n, mid))
    return out


def lattice_cells(size, cells):
    step = [F(size[a]) / F(cells[a]) for a in range(3)]
    out = []
    for iy in range(cells[1]):
        for iz in range(cells[2]):
            for ix in range(cells[0]):
                lo = tuple(F(F(i) * step[a]) - F(0.5) * F(size[a]) for a, i in enumerate((ix, iy, iz)))
                hi = tuple(lo[a] + step[a] for a in range(3))
                out.append((lo, hi))
    return out


def _intersects(s, e, lo, hi):
    fst, fet = F(0), F(1)
    for a in range(3):
        si, ei, bmin, bmax = s[a], e[a], lo[a], hi[a]
        if si < ei:
            if si > bmax or ei < bmin:
                return None
            di = ei - si
            st = (bmin - si) / di if si < bmin else F(0)
            et = (bmax - si) / di if ei > bmax else F(1)
        else:
            if ei > bmax or si < bmin:
                return None
            di = ei - si
            st = (bmax - si) / di if si > bmax else F(0)
            et = (bmin - si) / di if ei < bmin else F(1)
        if st > fst:
            fst = st
        if et < fet:
            fet = et
        if fet < fst:
            return None
    return tuple(s[a] + (e[a] - s[a]) * fst for a in range(3))


def _dist(a, b):
    d = [a[i] - b[i] for i in range(3)]
    return np.sqrt(F((d[0] * d[0] + d[1] * d[1]) + d[2] * d[2]), dtype=F)


def _first_split(s, e, LO, HI):
    """The first cell (in list order) whose entry point splits s->e into two
    parts longer than EPS: _intersects evaluated for every cell at once in
    float32 arrays (each element the same IEEE single operations), cells
    rejected on an earlier axis staying rejected.  Returns the entry point or
    None."""
    nc = LO.shape[0]
    alive = np.ones(nc, bool)
    t0 = np.zeros(nc, F)
    t1 = np.ones(nc, F)
    with np.errstate(divide="ignore", invalid="ignore"):
        for a in range(3):
            si, ei = F(s[a]), F(e[a])
            bmin, bmax = LO[:, a], HI[:, a]
            d = F(ei - si)
            if si < ei:
                alive &= ~((si > bmax) | (ei < bmin))
                st = np.where(si < bmin, (bmin - si) / d, F(0)).astype(F)
                et = np.where(ei > bmax, (bmax - si) / d, F(1)).astype(F)
            else:
                alive &= ~((ei > bmax) | (si < bmin))
                st = np.where(si > bmax, (bmax - si) / d, F(0)).astype(F)
                et = np.where(ei < bmin, (bmin - si) / d, F(1)).astype(F)
            t0 = np.where(st > t0, st, t0)
            t1 = np.where(et < t1, et, t1)
            alive &= ~(t1 < t0)
    idx = np.nonzero(alive)[0]
    for i in idx:  # few cells survive: the exact scalar test on them, in order
        ip = tuple(F(s[a] + F(F(e[a] - s[a]) * t0[i])) for a in range(3))
        if _dist(s, ip) > EPS and _dist(ip, e) > EPS:
            return ip
    return None


def split_walls(walls, cells_list):
    LO = np.array([c[0] for c in cells_list], F).reshape(-1, 3)
    HI = np.array([c[1] for c in cells_list], F).reshape(-1, 3)
    L = list(walls)
    i = 0
    while i < len(L):
        s, e, n, _ = L[i]
        ip = _first_split(s, e, LO, HI)
        if ip is None:
            i += 1
            continue
        m1 = tuple(s[a] + (ip[a] - s[a]) * F(0.5) for a in range(3))
        m2 = tuple(ip[a] + (e[a] - ip[a]) * F(0.5) for a in range(3))
        L.append((s, ip, n, m1))
        L.append((ip, e, n, m2))
        del L[i]
    return L


# ---- parameter files -------------------------------------------------------

_SCALAR_F = {"maxPursuitPredictionTime": "max_pursuit_prediction_time", "minTimeToCollision": "min_time_to_collision",
             "minTimeToObstacle": "min_time_to_obstacle", "minTimeToWall": "min_time_to_wall",
             "pathPredictionTime": "path_prediction_time", "minSeparationDistance": "min_separation_distance",
             "maxSeparationDistance": "max_separation_distance", "minFlockingDistance": "min_flocking_distance",
             "maxFlockingDistance": "max_flocking_distance", "cosMaxFlockingAngle": "cos_max_flocking_angle",
             "weightAlignment": "weight_alignment", "weightCohesion": "weight_cohesion",
             "weightSeparation": "weight_separation", "weightPursuit": "weight_pursuit", "weightSeek": "weight_seek",
             "weightFollowPath": "weight_follow_path", "weightAvoidObstacles": "weight_avoid_obstacles",
             "weightAvoidWalls": "weight_avoid_walls", "weightAvoidNeighbors": "weight_avoid_neighbors"}
_SCALAR_U = {"seed": "seed", "knn": "knn", "kno": "kno", "knw": "knw", "searchRadius": "search_radius",
             "maskAlignment": "mask_alignment", "maskAntiPenetrationAgents": "mask_anti_penetration_agents",
             "maskCohesion": "mask_cohesion", "maskSeparation": "mask_separation", "maskSeek": "mask_seek",
             "maskFlee": "mask_flee", "maskPursuit": "mask_pursuit", "maskEvade": "mask_evade",
             "maskFollowPath": "mask_follow_path", "maskAvoidObstacles": "mask_avoid_obstacles",
             "maskAvoidWalls": "mask_avoid_walls", "maskAvoidNeighbors": "mask_avoid_neighbors",
             "avoidCloseNeighbors": "avoid_close_neighbors"}


class ParamError(ValueError):
    pass


def _tokens(text):
    for line in text.splitlines():
        for tok in line.split():
            if tok.startswith("#"):
                break  # the rest of the line is a comment
            yield tok


def load_params(path) -> Dict:
    toks = _tokens(open(path, "rb").read().decode())

    def nxt():
        try:
            return next(toks)
        except StopIteration:
            raise ParamError("unexpected end of file")

    P = {"world": {"dimensions": (0.0, 0.0, 0.0), "cells": (0, 0, 0), "obstacle_count": 0, "map_filename": ""},
         "groups": []}
    key = nxt()
    while key != "</simulation>":
        if key == "<world>":
            k = nxt()
            while k != "</world>":
                nxt()
                if k == "dimensions":
                    P["world"]["dimensions"] = tuple(F(nxt()) for _ in range(3))
                elif k == "cells":
                    P["world"]["cells"] = tuple(int(nxt()) for _ in range(3))
                elif k == "obstacleCount":
                    P["world"]["obstacle_count"] = int(nxt())
                elif k == "mapFilename":
                    P["world"]["map_filename"] = nxt()
                else:
                    raise ParamError(k)
                k = nxt()
        elif key == "<group>":
            g = {"num_agents": 0, "max_speed": 0.0, "max_force": 0.0, "start_position": (0.0, 0.0, 0.0),
                 "min_start_radius": 0.0, "max_start_radius": 0.0, "path": [], "path_radius": 0.0,
                 "path_cyclic": 0, "body_color": (0.0, 0.0, 0.0)}
            k = nxt()
            while k != "</group>":
                if k == "<path>":
                    pk = nxt()
                    while pk != "</path>":
                        nxt()
                        if pk == "point":
                            g["path"].append(tuple(F(nxt()) for _ in range(3)))
                        elif pk == "cyclic":
                            g["path_cyclic"] = int(nxt())
                        elif pk == "radius":
                            g["path_radius"] = F(nxt())
                        else:
                            raise ParamError(pk)
                        pk = nxt()
                else:
                    nxt()
                    if k == "numAgents":
                        g["num_agents"] = int(nxt())
                    elif k in ("maxSpeed", "maxForce", "minStartRadius", "maxStartRadius"):
                        g[{"maxSpeed": "max_speed", "maxForce": "max_force", "minStartRadius": "min_start_radius",
                           "maxStartRadius": "max_start_radius"}[k]] = F(nxt())
                    elif k in ("startPosition", "bodyColor"):
                        g["start_position" if k == "startPosition" else "body_color"] = tuple(
                            F(nxt()) for _ in range(3))
                    else:
                        raise ParamError(k)
                k = nxt()
            P["groups"].append(g)
        else:
            nxt()
            if key in _SCALAR_F:
                P[_SCALAR_F[key]] = F(nxt())
            elif key in _SCALAR_U:
                P[_SCALAR_U[key]] = int(nxt())
            else:
                raise ParamError(key)
        key = nxt()
    return P
