"""ctypes wrapper of the CPU oracle (oracle/liboracle.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and
bench.py's cpu_baseline leg, to check and time the HIP path.  The product
never imports it.

The composite steps below issue the oracle's per-behaviour passes in the same
order as the corresponding sequence of C-ABI calls, so a state advanced here
and on the GPU must agree (bit for bit for KNN and forces).
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
# GS_ORACLE_LIB selects another build of the same source (the sanitizer
# build, oracle/_san/liboracle_san.so).
LIB_PATH = Path(os.environ["GS_ORACLE_LIB"]) if os.environ.get("GS_ORACLE_LIB") else HERE / "liboracle.so"
REF_LQ = HERE / "_ref" / "liblq.so"

AGENT_FIELDS = (
    "px", "py", "pz", "fx", "fy", "fz", "speed", "ax", "ay", "az",
    "sx", "sy", "sz", "ux", "uy", "uz", "gx", "gy", "gz",
    "radius", "max_speed", "max_force", "mass",
)

MODE_OPENSTEER, MODE_GPUREF = 0, 1
WRAP_BOX, WRAP_SPHERE = 0, 1
FLOCK_SEPARATION, FLOCK_ALIGNMENT, FLOCK_COHESION = 0, 1, 2


def build() -> Path:
    """Compile liboracle.so (and the reference's lq.c into _ref/ when the
    reference sources are present)."""
    subprocess.run(["make", "-s", "-C", str(HERE), "liboracle.so"], check=True)
    if Path("/root/reference/src/lq.c").exists():
        subprocess.run(["make", "-s", "-C", str(HERE), "ref"], check=True)
    return LIB_PATH


class OrGrid(C.Structure):
    _fields_ = [("ox", C.c_float), ("oy", C.c_float), ("oz", C.c_float), ("sx", C.c_float),
                ("sy", C.c_float), ("sz", C.c_float), ("dx", C.c_int), ("dy", C.c_int), ("dz", C.c_int)]


def _agents_type(ctype):
    class OrAgents(C.Structure):
        _fields_ = [(f, C.POINTER(ctype)) for f in AGENT_FIELDS] + [("bits", C.POINTER(C.c_uint32))]
    return OrAgents


OrAgents32 = _agents_type(C.c_float)
OrAgents64 = _agents_type(C.c_double)

_lib = None


def lib() -> C.CDLL:
    global _lib
    if _lib is None:
        if not LIB_PATH.exists():
            build()
        L = C.CDLL(str(LIB_PATH))
        f, u, i = C.POINTER(C.c_float), C.POINTER(C.c_uint32), C.c_int
        L.or_cell_of.restype = C.c_uint32
        L.or_cell_of.argtypes = [C.POINTER(OrGrid), C.c_float, C.c_float, C.c_float]
        L.or_bins_build.argtypes = [C.POINTER(OrGrid), i, f, f, f, u, u, u, f, f, f]
        L.or_knn_brute.argtypes = [i, f, f, f, i, f, f, f, i, i, u, f, u, i, C.c_float]
        L.or_knn_binned.argtypes = [C.POINTER(OrGrid), i, f, f, f, i, f, f, f, i, i, u, f, C.c_float]
        L.or_pow20.restype = C.c_float
        L.or_pow20.argtypes = [C.c_float]
        L.or_num_threads.restype = C.c_int
        L.or_set_num_threads.argtypes = [C.c_int]
        for suf, T, A in (("f32", C.c_float, OrAgents32), ("f64", C.c_double, OrAgents64)):
            t = C.POINTER(T)
            getattr(L, f"or_steer_flock_{suf}").argtypes = [i, i, C.POINTER(A), i, u, i, C.POINTER(A), i, T, T, T,
                                                             T, C.c_uint32]
            getattr(L, f"or_steer_seek_{suf}").argtypes = [i, C.POINTER(A), i, t, t, t, i, T, C.c_uint32]
            getattr(L, f"or_avoid_obstacles_{suf}").argtypes = [C.POINTER(A), i, u, i, i, t, t, t, t, T, T,
                                                                 C.c_uint32]
            getattr(L, f"or_avoid_walls_{suf}").argtypes = [C.POINTER(A), i, u, i, i] + [t] * 9 + [T, T, C.c_uint32]
            getattr(L, f"or_update_{suf}").argtypes = [i, C.POINTER(A), i, T]
            getattr(L, f"or_wrap_world_{suf}").argtypes = [i, C.POINTER(A), i, t, t]
            getattr(L, f"or_update_anti_penetration_{suf}").argtypes = [C.POINTER(A), i, u, i, i] + [t] * 9 + [T, i]
            getattr(L, f"or_anti_penetration_agents_{suf}").argtypes = [C.POINTER(A), i, u, f, C.POINTER(A), i,
                                                                         C.c_uint32]
            getattr(L, f"or_path_init_{suf}").restype = T
            getattr(L, f"or_path_init_{suf}").argtypes = [i, t, i, t, t, t]
            getattr(L, f"or_follow_path_{suf}").argtypes = [C.POINTER(A), i, i, t, t, t, T, T, i, T, T, C.c_uint32]
            getattr(L, f"or_path_point_at_{suf}").argtypes = [i, t, t, t, T, i, T, t]
            getattr(L, f"or_avoid_neighbors_{suf}").argtypes = [i, C.POINTER(A), i, u, f, i, C.POINTER(A), i, T, T, i,
                                                                 T, C.c_uint32]
            getattr(L, f"or_path_distance_of_{suf}").restype = T
            getattr(L, f"or_path_distance_of_{suf}").argtypes = [i, t, t, t, t]
        L.or_limit_max_deviation_angle.argtypes = [f, C.c_float, f, f]
        L.or_truncate_length.argtypes = [f, C.c_float, f]
        L.or_lines_intersect.restype = C.c_int
        L.or_lines_intersect.argtypes = [f, f, f, f, f]
        _lib = L
    return _lib


def set_threads(n: int) -> None:
    lib().or_set_num_threads(int(n))


def num_threads() -> int:
    return lib().or_num_threads()


def _f(a):
    assert a.dtype == np.float32 and a.flags.c_contiguous, (a.dtype, a.flags)
    return a.ctypes.data_as(C.POINTER(C.c_float))


def _d(a):
    assert a.dtype == np.float64 and a.flags.c_contiguous
    return a.ctypes.data_as(C.POINTER(C.c_double))


def _u(a):
    assert a.dtype == np.uint32 and a.flags.c_contiguous
    return a.ctypes.data_as(C.POINTER(C.c_uint32))


def grid(origin, size, cells) -> OrGrid:
    return OrGrid(*[float(x) for x in origin], *[float(x) for x in size], *[int(c) for c in cells])


def _c32(x):
    return np.ascontiguousarray(x, dtype=np.float32)


# ---------------------------------------------------------------------------
# Bins and KNN.


def cell_of(g: OrGrid, x, y, z) -> int:
    return int(lib().or_cell_of(C.byref(g), float(x), float(y), float(z)))


def bins_build(g: OrGrid, pos: np.ndarray):
    n = pos.shape[0]
    C_ = g.dx * g.dy * g.dz
    px, py, pz = (_c32(pos[:, j]) for j in range(3))
    key = np.zeros(n, np.uint32)
    start = np.zeros(C_ + 1, np.uint32)
    perm = np.zeros(n, np.uint32)
    sp = [np.zeros(n, np.float32) for _ in range(3)]
    lib().or_bins_build(C.byref(g), n, _f(px), _f(py), _f(pz), _u(key), _u(start), _u(perm), _f(sp[0]),
                        _f(sp[1]), _f(sp[2]))
    return key, start, perm, np.stack(sp, axis=1)


def knn_brute(a_pos: np.ndarray, b_pos: np.ndarray, k: int, self_group: bool, queries=None, cap=float("inf")):
    """Exact KNN (key (d2, index)) among points with d2 < cap^2; returns idx,
    dist as [k][nA]."""
    nA, nB = a_pos.shape[0], b_pos.shape[0]
    A = [_c32(a_pos[:, j]) for j in range(3)]
    B = A if self_group else [_c32(b_pos[:, j]) for j in range(3)]
    idx = np.full((k, nA), 0xFFFFFFFF, np.uint32)
    dist = np.full((k, nA), np.finfo(np.float32).max, np.float32)
    q = None if queries is None else np.ascontiguousarray(queries, dtype=np.uint32)
    lib().or_knn_brute(nA, _f(A[0]), _f(A[1]), _f(A[2]), nB, _f(B[0]), _f(B[1]), _f(B[2]), int(self_group), k,
                       _u(idx), _f(dist), None if q is None else _u(q), 0 if q is None else q.shape[0], cap)
    return idx, dist


def knn_binned(g: OrGrid, a_pos: np.ndarray, b_pos: np.ndarray, k: int, self_group: bool, cap=float("inf")):
    nA, nB = a_pos.shape[0], b_pos.shape[0]
    A = [_c32(a_pos[:, j]) for j in range(3)]
    B = A if self_group else [_c32(b_pos[:, j]) for j in range(3)]
    idx = np.zeros((k, nA), np.uint32)
    dist = np.zeros((k, nA), np.float32)
    lib().or_knn_binned(C.byref(g), nA, _f(A[0]), _f(A[1]), _f(A[2]), nB, _f(B[0]), _f(B[1]), _f(B[2]),
                        int(self_group), k, _u(idx), _f(dist), cap)
    return idx, dist


# ---------------------------------------------------------------------------
# Agent passes (operate in place on an AgentState of float32 or float64).


class _Bound:
    """Keeps the numpy arrays of an AgentState alive while a struct points at them."""

    def __init__(self, state):
        dt = state.arrays["px"].dtype
        self.f64 = dt == np.float64
        for k in AGENT_FIELDS:
            a = state.arrays[k]
            assert a.dtype == dt and a.flags.c_contiguous
        conv = _d if self.f64 else _f
        S = OrAgents64 if self.f64 else OrAgents32
        self.s = S(*[conv(state.arrays[k]) for k in AGENT_FIELDS], _u(state.bits))
        self.suf = "f64" if self.f64 else "f32"
        self.T = np.float64 if self.f64 else np.float32
        self.conv = conv
        self.state = state


def steer_flock(which, state, knn_idx, other, min_d, max_d, cos_max, w, mask=0, mode=MODE_OPENSTEER):
    a = _Bound(state)
    b = a if other is state else _Bound(other)
    k = knn_idx.shape[0]
    knn = np.ascontiguousarray(knn_idx, dtype=np.uint32)
    getattr(lib(), f"or_steer_flock_{a.suf}")(which, mode, C.byref(a.s), state.n, _u(knn), k, C.byref(b.s),
                                               other.n, min_d, max_d, cos_max, w, mask)


def steer_seek(state, target, w, mask=0, mode=MODE_OPENSTEER):
    a = _Bound(state)
    t = [np.array([target[j]], dtype=a.T) for j in range(3)]
    getattr(lib(), f"or_steer_seek_{a.suf}")(mode, C.byref(a.s), state.n, a.conv(t[0]), a.conv(t[1]),
                                              a.conv(t[2]), 0, w, mask)


def avoid_obstacles(state, knn_idx, obstacles, min_time, w, mask=0):
    a = _Bound(state)
    o = [np.ascontiguousarray(x, dtype=a.T) for x in obstacles]
    knn = np.ascontiguousarray(knn_idx, dtype=np.uint32)
    getattr(lib(), f"or_avoid_obstacles_{a.suf}")(C.byref(a.s), state.n, _u(knn), knn.shape[0], o[0].shape[0],
                                                   *[a.conv(x) for x in o], min_time, w, mask)


def avoid_walls(state, knn_idx, walls, min_time, w, mask=0):
    a = _Bound(state)
    ws = [np.ascontiguousarray(walls.arrays[f], dtype=a.T) for f in ("sx", "sy", "sz", "ex", "ey", "ez", "nx",
                                                                      "ny", "nz")]
    knn = np.ascontiguousarray(knn_idx, dtype=np.uint32)
    getattr(lib(), f"or_avoid_walls_{a.suf}")(C.byref(a.s), state.n, _u(knn), knn.shape[0], walls.n,
                                               *[a.conv(x) for x in ws], min_time, w, mask)


def update(state, dt, mode=MODE_OPENSTEER):
    a = _Bound(state)
    getattr(lib(), f"or_update_{a.suf}")(mode, C.byref(a.s), state.n, dt)


CLOSEST_CLAMPED, CLOSEST_AS_REFERENCE = 0, 1


def update_anti_penetration(state, knn_idx, walls, dt, closest=CLOSEST_CLAMPED):
    """UpdateCUDAKernelNew: GPU-reference integrate with wall anti-penetration
    over the k nearest walls; bits are kept (ANTI_PENETRATION_WALL added)."""
    a = _Bound(state)
    ws = [np.ascontiguousarray(walls.arrays[f], dtype=a.T) for f in ("sx", "sy", "sz", "ex", "ey", "ez", "nx",
                                                                      "ny", "nz")]
    knn = np.ascontiguousarray(knn_idx, dtype=np.uint32)
    getattr(lib(), f"or_update_anti_penetration_{a.suf}")(C.byref(a.s), state.n, _u(knn), knn.shape[0], walls.n,
                                                           *[a.conv(x) for x in ws], dt, closest)


def anti_penetration_agents(state, knn_idx, knn_dist, other, mask=0):
    """AntiPenetrationAgentsCUDAKernel over the first KNN entry of each agent
    (all reads of the state before the call)."""
    a = _Bound(state)
    b = a if other is state else _Bound(other)
    knn = np.ascontiguousarray(knn_idx[0], dtype=np.uint32)
    dist = np.ascontiguousarray(knn_dist[0], dtype=np.float32)
    getattr(lib(), f"or_anti_penetration_agents_{a.suf}")(C.byref(a.s), state.n, _u(knn), _f(dist), C.byref(b.s),
                                                           other.n, mask)


def avoid_neighbors(state, knn_idx, knn_dist, other, min_time, min_sep, avoid_close, w, mask=0, mode=MODE_GPUREF):
    """steerToAvoidNeighbors over the KNN of `state` in `other` (close
    neighbours first when avoid_close, GPU-reference rules; or the CPU
    SteerLibrary rules in MODE_OPENSTEER)."""
    a = _Bound(state)
    b = a if other is state else _Bound(other)
    knn = np.ascontiguousarray(knn_idx, dtype=np.uint32)
    dist = np.ascontiguousarray(knn_dist, dtype=np.float32)
    getattr(lib(), f"or_avoid_neighbors_{a.suf}")(mode, C.byref(a.s), state.n, _u(knn), _f(dist), knn.shape[0],
                                                   C.byref(b.s), other.n, min_time, min_sep, int(bool(avoid_close)), w,
                                                   mask)


class Pathway:
    """PolylinePathwayCUDA's arrays: points[m][3], normals[m][3], lengths[m],
    total length (m = n, + 1 when cyclic)."""

    def __init__(self, points, radius, cyclic, dtype=np.float32):
        pin = np.ascontiguousarray(np.asarray(points, dtype=dtype).reshape(-1, 3))
        n = pin.shape[0]
        m = n + (1 if cyclic else 0)
        self.points = np.zeros((m, 3), dtype)
        self.normals = np.zeros((m, 3), dtype)
        self.lengths = np.zeros(m, dtype)
        suf = "f64" if dtype == np.float64 else "f32"
        conv = _d if dtype == np.float64 else _f
        self.total = dtype(getattr(lib(), f"or_path_init_{suf}")(n, conv(pin), int(bool(cyclic)), conv(self.points),
                                                                  conv(self.normals), conv(self.lengths)))
        self.m, self.radius, self.cyclic, self.dtype, self.suf, self.conv = m, dtype(radius), bool(cyclic), dtype, suf, conv

    def point_at(self, dist):
        out = np.zeros(3, self.dtype)
        getattr(lib(), f"or_path_point_at_{self.suf}")(self.m, self.conv(self.points), self.conv(self.normals),
                                                        self.conv(self.lengths), self.total, int(self.cyclic),
                                                        dist, self.conv(out))
        return out

    def distance_of(self, p):
        q = np.ascontiguousarray(np.asarray(p, self.dtype))
        return self.dtype(getattr(lib(), f"or_path_distance_of_{self.suf}")(
            self.m, self.conv(self.points), self.conv(self.normals), self.conv(self.lengths), self.conv(q)))


def follow_path(state, path: Pathway, prediction_time, w, mask=0):
    a = _Bound(state)
    getattr(lib(), f"or_follow_path_{a.suf}")(C.byref(a.s), state.n, path.m, path.conv(path.points),
                                               path.conv(path.normals), path.conv(path.lengths), path.radius,
                                               path.total, int(path.cyclic), prediction_time, w, mask)


def wrap_world(state, shape, center, size_or_radius):
    a = _Bound(state)
    c = np.array(center, dtype=a.T)
    s = np.array(size_or_radius, dtype=a.T)
    getattr(lib(), f"or_wrap_world_{a.suf}")(shape, C.byref(a.s), state.n, a.conv(c), a.conv(s))


# ---------------------------------------------------------------------------
# Composite steps.


def boids_step(state, p, knn_fn=None):
    """One Boids step in the order of plugins/BoidsGroup.cpp:606-637 with the
    force composition of plugins/Boids.cpp:157-178: wrap -> KNN (self) ->
    separation, alignment, cohesion -> integrate.  Returns the KNN (idx, dist)
    computed from the post-wrap positions."""
    if p.wrap_shape in (WRAP_BOX, WRAP_SPHERE):
        wrap_world(state, p.wrap_shape, p.wrap_center, p.wrap_size)
    pos = np.stack([state.arrays["px"], state.arrays["py"], state.arrays["pz"]], axis=1).astype(np.float32)
    g = grid(p.grid_origin, p.grid_size, p.grid_cells)
    idx, dist = (knn_fn or (lambda pp: knn_binned(g, pp, pp, p.k, True)))(pos)
    for which, t in ((FLOCK_SEPARATION, p.separation), (FLOCK_ALIGNMENT, p.alignment),
                     (FLOCK_COHESION, p.cohesion)):
        steer_flock(which, state, idx, state, t.min_d, t.max_d, t.cos_max, t.weight, t.mask, p.mode)
    update(state, p.dt, p.mode)
    return idx, dist


def choke_step(groups, sc):
    """One choke-point step for both groups (Jacobi across groups): per group,
    clear bits -> KNN to itself, to the other group (within sc.other_radius),
    to walls and obstacles ->
    avoid walls -> avoid obstacles -> separation (self, other) -> seek; then
    integrate both.  Returns the KNN results per group."""
    g = grid(sc.grid_origin, sc.grid_size, sc.grid_cells)
    wmid = sc.walls.mid() if getattr(sc, "wall_mid", None) is None else np.asarray(sc.wall_mid, np.float32)
    opos = np.stack(sc.obstacles[:3], axis=1).astype(np.float32)
    pos = [np.stack([s.arrays["px"], s.arrays["py"], s.arrays["pz"]], axis=1).astype(np.float32) for s in groups]
    out = []
    for gi, st in enumerate(groups):
        other = groups[1 - gi]
        st.bits[:] = 0
        ks = knn_binned(g, pos[gi], pos[gi], sc.k, True)[0]
        ko = knn_binned(g, pos[gi], pos[1 - gi], sc.k, False, cap=sc.other_radius)[0]
        # Walls and obstacles are a few dozen points: brute force gives the
        # same exact lists (binned == brute is pinned) without expanding
        # rings across the lattice's empty cells.
        kw = knn_brute(pos[gi], wmid, sc.knw, False)[0]
        kob = knn_brute(pos[gi], opos, sc.kno, False)[0]
        avoid_walls(st, kw, sc.walls, sc.min_time_walls, sc.w_walls, 0)
        avoid_obstacles(st, kob, sc.obstacles, sc.min_time_obstacles, sc.w_obstacles, 0)
        steer_flock(FLOCK_SEPARATION, st, ks, st, sc.sep.min_d, sc.sep.max_d, sc.sep.cos_max, sc.sep.weight,
                    sc.sep.mask, sc.mode)
        steer_flock(FLOCK_SEPARATION, st, ko, other, sc.sep.min_d, sc.sep.max_d, sc.sep.cos_max, sc.sep.weight,
                    sc.sep.mask, sc.mode)
        steer_seek(st, sc.goals[gi], sc.w_seek, sc.mask_seek, sc.mode)
        out.append((ks, ko, kw, kob))
    for st in groups:
        update(st, sc.dt, sc.mode)
    return out


def choke_new_step(state, sc, path: "Pathway", variant="new", closest=CLOSEST_CLAMPED):
    """One step of the choke-point plugin in the GPU-reference formulas.
    variant "new": CtfGroup::update (plugins/ChokePointNew.cpp:587-630) --
    wrap, clear bits, KNN to itself and to the walls (capped at sc.wall_cap),
    avoid walls, separation (mask ANTI_PENETRATION_WALL), alignment,
    cohesion, follow path, seek, integrate, agent anti-penetration.
    variant "group": CtfEnemyGroup::update (plugins/ChokePointGroup.cpp:683-728)
    -- walls, path, seek, separation, alignment, cohesion, the integrate with
    wall anti-penetration, agent anti-penetration masked by AVOID_WALLS |
    ANTI_PENETRATION_WALL.  Returns (self KNN, wall KNN, the positions the
    step binned)."""
    p = sc.params
    size = sc.world_size
    wrap_world(state, WRAP_BOX, (0.0, 0.0, 0.0), size)
    g = grid(tuple(-0.5 * np.float32(v) for v in size), size, sc.world_cells)
    state.bits[:] = 0
    pos = np.stack([state.arrays["px"], state.arrays["py"], state.arrays["pz"]], axis=1).astype(np.float32)
    ks, ds = knn_binned(g, pos, pos, p.knn, True)
    kw, _ = knn_binned(g, pos, np.asarray(sc.wall_mid, np.float32), p.knw, False, cap=sc.wall_cap)
    avoid_walls(state, kw, sc.walls, p.min_time_to_wall, p.weight_avoid_walls, p.mask_avoid_walls)
    flock = [(FLOCK_SEPARATION, p.min_separation_distance, p.max_separation_distance, p.weight_separation,
              1 << 12 if variant == "new" else p.mask_separation),
             (FLOCK_ALIGNMENT, p.min_flocking_distance, p.max_flocking_distance, p.weight_alignment,
              p.mask_alignment),
             (FLOCK_COHESION, p.min_flocking_distance, p.max_flocking_distance, p.weight_cohesion, p.mask_cohesion)]
    path_terms = [lambda: follow_path(state, path, p.path_prediction_time, p.weight_follow_path, p.mask_follow_path),
                  lambda: steer_seek(state, sc.seek_target, p.weight_seek, p.mask_seek, MODE_GPUREF)]
    if variant == "new":
        for which, lo, hi, w, m in flock:
            steer_flock(which, state, ks, state, lo, hi, p.cos_max_flocking_angle, w, m, MODE_GPUREF)
        for f in path_terms:
            f()
        update(state, sc.dt, MODE_GPUREF)
        anti_penetration_agents(state, ks, ds, state, p.mask_anti_penetration_agents)
    else:
        for f in path_terms:
            f()
        for which, lo, hi, w, m in flock:
            steer_flock(which, state, ks, state, lo, hi, p.cos_max_flocking_angle, w, m, MODE_GPUREF)
        update_anti_penetration(state, kw, sc.walls, sc.dt, closest)
        anti_penetration_agents(state, ks, ds, state, (1 << 8) | (1 << 12))
    return ks, kw, pos


def multi_group_step(states, sc, paths, mode=MODE_GPUREF):
    """One frame of plugins/MultiGroup.cpp: Group::preUpdate of every group
    (:546-556: wrap, clear bits; the bins are the positions after the wrap),
    then Group::update of each group in turn (:558-611).  KNN to group j uses
    group j's positions at its preUpdate (its bin database); the forces read
    the groups' current state.  Returns the KNN results of the last frame
    [group][other] = (idx, dist)."""
    p = sc.params
    size = sc.world_size
    g = grid(tuple(-0.5 * np.float32(v) for v in size), size, sc.world_cells)
    snap = []
    for st in states:
        wrap_world(st, WRAP_BOX, (0.0, 0.0, 0.0), size)
        st.bits[:] = 0
        snap.append(np.stack([st.arrays["px"], st.arrays["py"], st.arrays["pz"]], axis=1).astype(np.float32))
    out = []
    for gi, A in enumerate(states):
        knn = [knn_binned(g, snap[gi], snap[j], p.knn, j == gi, cap=sc.knn_cap) for j in range(len(states))]
        for j, B in enumerate(states):
            if j != gi:
                avoid_neighbors(A, knn[j][0], knn[j][1], B, p.min_time_to_collision, p.min_separation_distance,
                                p.avoid_close_neighbors, p.weight_avoid_neighbors, p.mask_avoid_neighbors, mode)
        for j, B in enumerate(states):
            steer_flock(FLOCK_SEPARATION, A, knn[j][0], B, p.min_separation_distance, p.max_separation_distance,
                        p.cos_max_flocking_angle, p.weight_separation, p.mask_separation, mode)
        for which, w, m in ((FLOCK_ALIGNMENT, p.weight_alignment, p.mask_alignment),
                            (FLOCK_COHESION, p.weight_cohesion, p.mask_cohesion)):
            steer_flock(which, A, knn[gi][0], A, p.min_flocking_distance, p.max_flocking_distance,
                        p.cos_max_flocking_angle, w, m, mode)
        if paths[gi] is not None:
            follow_path(A, paths[gi], p.path_prediction_time, p.weight_follow_path, p.mask_follow_path)
            steer_seek(A, tuple(float(v) for v in sc.paths[gi][0][1]), p.weight_seek, p.mask_seek, mode)
        update(A, sc.dt, mode)
        for j, B in enumerate(states):
            anti_penetration_agents(A, knn[j][0], knn[j][1], B, p.mask_anti_penetration_agents)
        out.append(knn)
    return out


# ---------------------------------------------------------------------------
# Calibration model (FP64), for gradient checks of the adjoint.

C4_STATE = ("px", "py", "pz", "fx", "fy", "fz", "speed", "sx", "sy", "sz", "ux", "uy", "uz")


class C4:
    """One crowd scenario in the oracle's FP64 calibration model.  run(theta,
    mode=1) records the decision tape and KNN lists; mode=2 replays them, so
    central differences around the recorded run are the derivative of its
    branch."""

    def __init__(self, sc, cfg, k=7):
        L = lib()
        d = C.POINTER(C.c_double)
        L.or_c4_create.restype = C.c_void_p
        L.or_c4_create.argtypes = [C.c_int, C.c_int, d, d, d, C.c_int, d, C.c_int, d]
        L.or_c4_destroy.argtypes = [C.c_void_p]
        L.or_c4_run.restype = C.c_double
        L.or_c4_run.argtypes = [C.c_void_p, d, d, C.POINTER(C.c_int), C.c_int, d, C.c_int]
        self.n = [g.n for g in sc.groups]
        st = [np.ascontiguousarray(np.stack([g.arrays[f].astype(np.float64) for f in C4_STATE])) for g in sc.groups]
        goals = np.ascontiguousarray(np.array(sc.goals, np.float64))
        w = np.ascontiguousarray(np.stack([sc.walls.arrays[f].astype(np.float64) for f in
                                           ("sx", "sy", "sz", "ex", "ey", "ez", "nx", "ny", "nz")]))
        o = np.ascontiguousarray(np.stack([np.asarray(a, np.float64) for a in sc.obstacles]))
        self._keep = (st, goals, w, o)
        self.h = L.or_c4_create(self.n[0], self.n[1], _d(st[0]), _d(st[1]), _d(goals), w.shape[1], _d(w),
                                o.shape[1], _d(o))
        t = cfg
        self.cfgd = np.array([t.separation.min_d, t.separation.max_d, t.separation.cos_max, t.alignment.min_d,
                              t.alignment.max_d, t.alignment.cos_max, t.cohesion.min_d, t.cohesion.max_d,
                              t.cohesion.cos_max, t.seek_weight, t.walls_min_time, t.obstacles_min_time,
                              t.max_force, t.mass, t.dt, t.penetration_weight, sc.other_radius], np.float64)
        self.cfgi = (C.c_int * 8)(t.separation.mask, t.alignment.mask, t.cohesion.mask, t.seek_mask, k, sc.knw,
                                  sc.kno, t.steps)

    def run(self, theta, mode: int, want_state: bool = False, precision: int = 64):
        """mode 1 records the tape, 2 replays it; precision 32 runs the model
        in fp32 (the GPU's arithmetic), 64 in FP64."""
        th = np.ascontiguousarray(theta, dtype=np.float64)
        fs = np.zeros(13 * (self.n[0] + self.n[1]), np.float64) if want_state else None
        loss = lib().or_c4_run(self.h, _d(th), _d(self.cfgd), self.cfgi, mode, None if fs is None else _d(fs),
                               precision)
        return (loss, fs) if want_state else loss

    def fd_grad(self, theta, h: float = 1e-3, record_precision: int = 64) -> np.ndarray:
        """Central differences of the FP64 loss at theta with the branch
        decisions and KNN lists frozen at a run recorded in the given
        precision (32: the branches the GPU's fp32 run takes)."""
        self.run(theta, 1, precision=record_precision)
        base = self.run(theta, 2)
        g = np.zeros(14)
        for j in range(14):
            e = np.zeros(14)
            e[j] = h
            g[j] = (self.run(theta + e, 2) - self.run(theta - e, 2)) / (2 * h)
        return base, g

    def tangent_grad(self, theta, primal_precision: int = 64, want_rms: bool = False):
        """The analytic gradient of the taped model by forward-mode tangents
        (replays the tape and KNN lists of the last recording run): returns
        (loss, grad[14]) and, with want_rms, the RMS over agents of the
        position tangent after every step ([steps][14]).  primal_precision
        32 evaluates it along the fp32 trajectory (the GPU's), 64 along the
        FP64 one (the finite differences' function)."""
        L = lib()
        d = C.POINTER(C.c_double)
        L.or_c4_tangent.restype = C.c_double
        L.or_c4_tangent.argtypes = [C.c_void_p, d, d, C.POINTER(C.c_int), C.c_int, d, d]
        th = np.ascontiguousarray(theta, dtype=np.float64)
        g = np.zeros(14, np.float64)
        rms = np.zeros((self.cfgi[7], 14), np.float64) if want_rms else None
        loss = L.or_c4_tangent(self.h, _d(th), _d(self.cfgd), self.cfgi, primal_precision, _d(g),
                               None if rms is None else _d(rms))
        return (loss, g, rms) if want_rms else (loss, g)

    def close(self):
        if self.h:
            lib().or_c4_destroy(self.h)
            self.h = None
