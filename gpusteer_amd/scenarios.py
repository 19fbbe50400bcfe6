"""Scenario generators: seeded synthetic states for the benchmark workloads.

Every generator is deterministic (numpy Philox, seed 12345 + workload index,
the seed of data/ChokePoint.params:3) and returns float32 struct-of-arrays
state, so the HIP path and the CPU oracle run on identical inputs.

Workloads
---------
boids_sphere   4,096-agent Boids flock in a sphere of radius 50 with antipode
               wrap (plugins/Boids.cpp:89-209 parameters; k = 7, 16^3 lattice
               over the 1.1 x R margin of plugins/Boids.cpp:375-380).
choke          two 5,000-agent groups in 40 x 40 regions either side of a
               wall with one gap (full width 15), 50 spherical obstacles,
               seek goal in the opposite region (data/ChokePoint.params style).
boids_box      2^20-agent flock in a 1000^3 periodic box (k = 7 or 16).
"""
from __future__ import annotations

from dataclasses import dataclass
from typing import Optional, Tuple

import numpy as np

from .state import AgentState, WallState, basis_from_forward

SEED = 12345


def rng(workload: int, extra: int = 0) -> np.random.Generator:
    return np.random.Generator(np.random.Philox(key=SEED + workload + 1000 * extra))


def unit_vectors(g: np.random.Generator, n: int) -> np.ndarray:
    v = g.normal(size=(n, 3))
    v /= np.linalg.norm(v, axis=1, keepdims=True)
    return v


def make_agents(pos: np.ndarray, fwd: np.ndarray, speed: np.ndarray, radius: float, max_speed: float,
                max_force: float, mass: float = 1.0) -> AgentState:
    n = pos.shape[0]
    s = AgentState.empty(n)
    a = s.arrays
    a["px"][:], a["py"][:], a["pz"][:] = pos[:, 0], pos[:, 1], pos[:, 2]
    f = fwd.astype(np.float32)
    a["fx"][:], a["fy"][:], a["fz"][:] = f[:, 0], f[:, 1], f[:, 2]
    side, up = basis_from_forward(a["fx"], a["fy"], a["fz"])
    a["sx"][:], a["sy"][:], a["sz"][:] = side[:, 0], side[:, 1], side[:, 2]
    a["ux"][:], a["uy"][:], a["uz"][:] = up[:, 0], up[:, 1], up[:, 2]
    a["speed"][:] = speed
    a["radius"][:] = radius
    a["max_speed"][:] = max_speed
    a["max_force"][:] = max_force
    a["mass"][:] = mass
    return s


@dataclass
class FlockTerm:
    min_d: float
    max_d: float
    cos_max: float
    weight: float
    mask: int = 0


@dataclass
class BoidsParams:
    """Flocking composition of plugins/Boids.cpp:136-178: separation r=5,
    cos -0.707, w 12; alignment r=7.5, cos 0.7, w 8; cohesion r=9, cos -0.15,
    w 8.  min_d < 0 selects SteerLibrary's 3 x radius minimum distance."""

    separation: FlockTerm = None
    alignment: FlockTerm = None
    cohesion: FlockTerm = None
    dt: float = 1.0 / 60.0
    k: int = 7
    wrap_shape: int = 1  # GS_WRAP_SPHERE
    wrap_center: Tuple[float, float, float] = (0.0, 0.0, 0.0)
    wrap_size: Tuple[float, float, float] = (50.0, 0.0, 0.0)
    grid_origin: Tuple[float, float, float] = (-55.0, -55.0, -55.0)
    grid_size: Tuple[float, float, float] = (110.0, 110.0, 110.0)
    grid_cells: Tuple[int, int, int] = (16, 16, 16)
    mode: int = 0  # GS_MODE_OPENSTEER
    search_radius: int = 1

    def __post_init__(self):
        if self.separation is None:
            self.separation = FlockTerm(-1.0, 5.0, -0.707, 12.0)
        if self.alignment is None:
            self.alignment = FlockTerm(-1.0, 7.5, 0.7, 8.0)
        if self.cohesion is None:
            self.cohesion = FlockTerm(-1.0, 9.0, -0.15, 8.0)


def boids_sphere(n: int = 4096, radius: float = 50.0, workload: int = 0) -> Tuple[AgentState, BoidsParams]:
    """Positions uniform in the ball (rejection sampling), forward uniform on
    the sphere, speed U(0, 9); radius 0.5, maxSpeed 9, maxForce 27, mass 1
    (plugins/Boids.cpp:95-101, include/OpenSteer/SimpleVehicle.h:107-110)."""
    g = rng(workload)
    pts = []
    count = 0
    while count < n:
        c = g.uniform(-1.0, 1.0, size=(2 * n, 3))
        c = c[np.einsum("ij,ij->i", c, c) < 1.0]
        pts.append(c)
        count += c.shape[0]
    pos = (np.concatenate(pts)[:n] * radius).astype(np.float32)
    fwd = unit_vectors(g, n)
    speed = g.uniform(0.0, 9.0, size=n).astype(np.float32)
    st = make_agents(pos, fwd, speed, 0.5, 9.0, 27.0)
    margin = radius * 1.1
    p = BoidsParams(wrap_shape=1, wrap_center=(0.0, 0.0, 0.0), wrap_size=(radius, 0.0, 0.0),
                    grid_origin=(-margin,) * 3, grid_size=(2 * margin,) * 3, grid_cells=(16, 16, 16))
    return st, p


def boids_box(n: int = 1 << 20, size: float = 1000.0, k: int = 7, cells: int = 64,
              workload: int = 2) -> Tuple[AgentState, BoidsParams]:
    """Positions uniform in [0, size)^3 with face wrap, velocities as in
    boids_sphere."""
    g = rng(workload)
    pos = g.uniform(0.0, size, size=(n, 3)).astype(np.float32)
    fwd = unit_vectors(g, n)
    speed = g.uniform(0.0, 9.0, size=n).astype(np.float32)
    st = make_agents(pos, fwd, speed, 0.5, 9.0, 27.0)
    h = size / 2
    p = BoidsParams(k=k, wrap_shape=0, wrap_center=(h, h, h), wrap_size=(size, size, size),
                    grid_origin=(0.0, 0.0, 0.0), grid_size=(size, size, size), grid_cells=(cells, cells, cells))
    return st, p


# ---------------------------------------------------------------------------
# Choke point.


@dataclass
class ChokeScenario:
    groups: Tuple[AgentState, AgentState]
    goals: Tuple[Tuple[float, float, float], Tuple[float, float, float]]
    walls: WallState
    obstacles: Tuple[np.ndarray, np.ndarray, np.ndarray, np.ndarray]  # cx, cy, cz, radius
    grid_origin: Tuple[float, float, float]
    grid_size: Tuple[float, float, float]
    grid_cells: Tuple[int, int, int]
    k: int = 7
    kno: int = 2
    knw: int = 3
    dt: float = 0.05
    min_time_walls: float = 10.0
    min_time_obstacles: float = 5.0
    sep: FlockTerm = None
    w_seek: float = 1.5
    mask_seek: int = 4480  # ANTI_PENETRATION_WALL | AVOID_WALLS | AVOID_OBSTACLES (ChokePoint.params)
    w_walls: float = 1.0
    w_obstacles: float = 1.0
    mode: int = 0
    # KNN to the other group keeps only agents closer than this: beyond the
    # separation term's max distance (3), so the forces equal the uncapped
    # search's.
    other_radius: float = 3.5
    # KNN positions of the walls ([n][3]); None means 0.5f * (start + end).
    # Walls split at bin-cell borders carry their own (start + 0.5f * (end -
    # start) of each piece, WallGroupData.cu:185-190).
    wall_mid: Optional[np.ndarray] = None


def choke_walls(gap: float, half_len: float = 100.0, thickness: float = 0.5) -> WallState:
    """A wall along x at z = 0 of the given thickness with one gap of full
    width `gap` centred at x = 0, as one-sided segments in the .map layout of
    data/ChokePoint15.map (start, end, normal): the +z face (normal +z), the
    -z face (normal -z) and the end cap at the gap (normal toward the gap),
    for each side.  (The reference's ChokePointN.map gap spans x in [-N, N];
    here the number is the full width.)"""
    h = gap / 2.0
    t = thickness / 2.0
    segs = [
        # left piece x in [-half_len, -h]
        ((-half_len, 0, t), (-h, 0, t), (0, 0, 1)),
        ((-half_len, 0, -t), (-h, 0, -t), (0, 0, -1)),
        ((-h, 0, -t), (-h, 0, t), (1, 0, 0)),
        # right piece x in [h, half_len]
        ((h, 0, t), (half_len, 0, t), (0, 0, 1)),
        ((h, 0, -t), (half_len, 0, -t), (0, 0, -1)),
        ((h, 0, t), (h, 0, -t), (-1, 0, 0)),
    ]
    arr = np.array([[*s, *e, *nrm] for s, e, nrm in segs], dtype=np.float32)
    return WallState({f: np.ascontiguousarray(arr[:, j]) for j, f in enumerate(
        ("sx", "sy", "sz", "ex", "ey", "ez", "nx", "ny", "nz"))})


def choke(n_total: int = 10000, gap: float = 15.0, n_obstacles: int = 50, workload: int = 1,
          region: float = 40.0, offset: float = 10.0) -> ChokeScenario:
    """2-D crowd on the XZ plane (y = 0).  Group 0 spawns uniform in
    x in [-20, 20], z in [offset, offset + 40] and seeks the centre of group
    1's region; group 1 mirrors it at negative z.  Agents: radius 0.5,
    maxSpeed 3, maxForce 8, speed 0, random heading on the plane.  Lattice:
    2-D, cell size 5."""
    g = rng(workload)
    n0 = n_total // 2
    groups = []
    goals = []
    for side in (+1, -1):
        x = g.uniform(-region / 2, region / 2, size=n0)
        z = side * g.uniform(offset, offset + region, size=n0)
        pos = np.stack([x, np.zeros(n0), z], axis=1).astype(np.float32)
        ang = g.uniform(0.0, 2 * np.pi, size=n0)
        fwd = np.stack([np.cos(ang), np.zeros(n0), np.sin(ang)], axis=1)
        groups.append(make_agents(pos, fwd, np.zeros(n0, np.float32), 0.5, 3.0, 8.0))
        goals.append((0.0, 0.0, float(-side * (offset + region / 2))))
    walls = choke_walls(gap)
    # Obstacles uniform in the play area, outside the gap corridor and the
    # wall band.
    ox, oz, orad = [], [], []
    while len(ox) < n_obstacles:
        r = float(g.uniform(0.5, 2.0))
        x = float(g.uniform(-50.0, 50.0))
        z = float(g.uniform(-offset - region, offset + region))
        if abs(x) < gap / 2 + r + 1.0 or abs(z) < r + 1.0:
            continue
        ox.append(x)
        oz.append(z)
        orad.append(r)
    obstacles = (np.array(ox, np.float32), np.zeros(n_obstacles, np.float32), np.array(oz, np.float32),
                 np.array(orad, np.float32))
    extent = 200.0
    return ChokeScenario(groups=(groups[0], groups[1]), goals=(goals[0], goals[1]), walls=walls,
                         obstacles=obstacles, grid_origin=(-extent / 2, -5.0, -extent / 2),
                         grid_size=(extent, 10.0, extent), grid_cells=(int(extent / 5), 1, int(extent / 5)),
                         sep=FlockTerm(-1.0, 3.0, -0.707, 10.0, 4096))


# ---------------------------------------------------------------------------
# Ensembles of scenarios.


def boids_cells_for(n: int, radius: float = 50.0, cell_factor: float = 1.0, k: int = 7) -> int:
    """Cells per axis of a Boids lattice over the ball's 1.1 R margin cube for
    n agents: cell edge = cell_factor x the expected k-th neighbour distance
    (k / (n / V) = 4/3 pi r^3)."""
    r_k = (k * (4.0 / 3.0) * radius ** 3 / n) ** (1.0 / 3.0)
    return int(np.clip(np.floor(2.2 * radius / (cell_factor * r_k)), 1, 128))


def choke_cells_for(n_total: int, region: float = 40.0, extent: float = 200.0, cell_factor: float = 1.0,
                    k: int = 7) -> int:
    """Cells per axis (x and z) of a crowd lattice: cell edge = cell_factor x
    the expected k-th neighbour distance at the spawn density."""
    rho = (n_total / 2) / (region * region)
    r_k = np.sqrt(k / (np.pi * rho))
    return int(np.clip(np.floor(extent / (cell_factor * r_k)), 1, 2048))


def ensemble_mix(n_scenarios: int = 256, n_min: int = 1024, n_max: int = 65536, workload: int = 3,
                 cell_factor: float = 1.0, mode: int = 0):
    """Mixed ensemble: agent counts log-uniform over [n_min, n_max]; even
    scenarios are Boids flocks (the ball of radius 50 with the Boids plugin's
    forces) with the three weights jittered by U(0.7, 1.3); odd ones are
    choke-point crowds (two groups either side of a wall line with one gap)
    with gap widths cycling through 7.5, 10, ..., 32.5.  Lattices are sized to
    each scenario's density.  Returns a list of ("boids", state, params) and
    ("choke", ChokeScenario) specs."""
    g = rng(workload)
    gaps = [7.5 + 2.5 * i for i in range(13)]
    specs = []
    for i in range(n_scenarios):
        n = int(round(2.0 ** g.uniform(np.log2(n_min), np.log2(n_max))))
        if i % 2 == 0:
            st, p = boids_sphere(n=n, workload=100 + i)
            jit = g.uniform(0.7, 1.3, size=3)
            p.separation.weight *= float(jit[0])
            p.alignment.weight *= float(jit[1])
            p.cohesion.weight *= float(jit[2])
            c = boids_cells_for(n, cell_factor=cell_factor)
            p.grid_cells = (c, c, c)
            p.mode = mode
            specs.append(("boids", st, p))
        else:
            n = max(2, n - (n % 2))
            sc = choke(n_total=n, gap=gaps[(i // 2) % len(gaps)], workload=100 + i)
            c = choke_cells_for(n, cell_factor=cell_factor)
            sc.grid_cells = (c, 1, c)
            sc.mode = mode
            specs.append(("choke", sc))
    return specs


# ---------------------------------------------------------------------------
# Differentiable calibration (crowd scenarios, the adjoint path).

PARAM_NAMES = ("w_sep", "w_ali", "w_coh", "w_obs", "w_wall", "radius", "max_speed")


@dataclass
class CalibConfig:
    """Behaviour constants of the calibration model (weights, radius and max
    speed are the parameters).  GPU-kernel neighbourhoods (explicit minimum
    distance): separation 1.5-3 cone -0.707 masked by wall anti-penetration;
    alignment 1.5-4 cone 0.7; cohesion 1.5-5 cone -0.15; seek weight 1.5
    masked by walls / obstacles (ChokePoint.params mask 4480); walls 10 s and
    obstacles 5 s look-ahead; maxForce 8, mass 1, dt 0.05; penetration weight
    10; 100 steps, checkpoints every 10."""

    separation: FlockTerm = None
    alignment: FlockTerm = None
    cohesion: FlockTerm = None
    seek_weight: float = 1.5
    seek_mask: int = 4480
    walls_min_time: float = 10.0
    obstacles_min_time: float = 5.0
    max_force: float = 8.0
    mass: float = 1.0
    dt: float = 0.05
    penetration_weight: float = 10.0
    steps: int = 100
    checkpoint_every: int = 10

    def __post_init__(self):
        if self.separation is None:
            self.separation = FlockTerm(1.5, 3.0, -0.707, 0.0, 4096)
        if self.alignment is None:
            self.alignment = FlockTerm(1.5, 4.0, 0.7, 0.0, 0)
        if self.cohesion is None:
            self.cohesion = FlockTerm(1.5, 5.0, -0.15, 0.0, 0)


def calib_params(w_sep=(10.0, 10.0), w_ali=(2.0, 2.0), w_coh=(1.0, 1.0), w_obs=(1.0, 1.0), w_wall=(1.0, 1.0),
                 radius=(0.5, 0.5), max_speed=(3.0, 3.0)) -> np.ndarray:
    """The 14 parameters as a float64 vector in gs_calib_params order
    (per name: group 0, group 1)."""
    return np.array([*w_sep, *w_ali, *w_coh, *w_obs, *w_wall, *radius, *max_speed], dtype=np.float64)


def calib_scenarios(n_scenarios: int = 64, n_agents: int = 16384, gap: float = 15.0, workload: int = 4):
    """Choke-point crowds (two groups either side of a wall line with one gap
    of width 15, 50 obstacles) for calibration,
    lattices sized to density."""
    specs = []
    for i in range(n_scenarios):
        sc = choke(n_total=n_agents, gap=gap, workload=workload * 1000 + i)
        c = choke_cells_for(n_agents)
        sc.grid_cells = (c, 1, c)
        sc.mode = 1
        specs.append(("choke", sc))
    return specs


# ---------------------------------------------------------------------------
# The choke-point plugin of plugins/ChokePointNew.cpp driven by its data files.


@dataclass
class ChokeNewScenario:
    """One group of CtfGroup (plugins/ChokePointNew.cpp:203-491) set up from a
    parsed .params file and its split walls: agents, the group's path, the
    world lattice and the per-step constants of CtfGroup::update
    (:587-630)."""
    state: AgentState
    walls: WallState
    wall_mid: np.ndarray
    path_points: np.ndarray        # [n][3] as given (PolylinePathwayCUDA adds the closing point)
    path_radius: float
    path_cyclic: bool
    seek_target: Tuple[float, float, float]   # m_pPath->hvPoints()[1] (:622)
    world_size: Tuple[float, float, float]
    world_cells: Tuple[int, int, int]
    params: object                 # formats.SimParams
    dt: float
    # Wall KNN range: walls farther than this cannot touch a feeler, the
    # agent's radius or its step (the KNN to walls is capped there; the
    # forces equal the uncapped search's).
    wall_cap: float


def choke_point_new(params, walls: WallState, wall_mid: np.ndarray, n_agents: int = None, dt: float = 0.1,
                    start=None, max_start_radius: float = None, workload: int = 1) -> ChokeNewScenario:
    """Agents of the params' first group placed as CtfGroup::reset does
    (:445-491, randomizeStartingPositionAndHeading2D of
    plugins/PluginCommon.cpp:5-19): radius uniform in [min, max] start radius
    on a random XZ direction around the start position, random heading on
    the XZ plane, speed 0, radius 0.5, mass 1 (SimpleVehicle's defaults), no
    two agents closer than one radius (redrawn, as the reference does).  The
    draws are this module's own seeded stream, not the reference's rand().
    n_agents / start / max_start_radius override the file's values (smaller
    test crowds)."""
    gp = params.groups[0]
    n = gp.num_agents if n_agents is None else int(n_agents)
    c = np.array(gp.start_position if start is None else start, np.float64)
    r_lo = gp.min_start_radius
    r_hi = gp.max_start_radius if max_start_radius is None else max_start_radius
    g = rng(workload, 7)
    radius = 0.5
    cell = {}
    pos = np.zeros((n, 3), np.float32)
    filled = 0
    while filled < n:
        m = n - filled
        rr = g.uniform(r_lo, r_hi, size=m)
        th = g.uniform(0.0, 2.0 * np.pi, size=m)
        cand = np.stack([c[0] + rr * np.cos(th), np.full(m, c[1]), c[2] + rr * np.sin(th)], axis=1).astype(np.float32)
        for p in cand:
            key = (int(np.floor(p[0] / radius)), int(np.floor(p[2] / radius)))
            ok = True
            for dx in (-1, 0, 1):
                for dz in (-1, 0, 1):
                    for q in cell.get((key[0] + dx, key[1] + dz), ()):
                        if (p[0] - q[0]) ** 2 + (p[2] - q[2]) ** 2 < radius * radius:
                            ok = False
                            break
                    if not ok:
                        break
                if not ok:
                    break
            if not ok:
                continue
            cell.setdefault(key, []).append(p)
            pos[filled] = p
            filled += 1
            if filled == n:
                break
    hd = g.uniform(0.0, 2.0 * np.pi, size=n)
    fwd = np.stack([np.cos(hd), np.zeros(n), np.sin(hd)], axis=1)
    st = make_agents(pos, fwd, np.zeros(n, np.float32), radius, gp.max_speed, gp.max_force, 1.0)
    pts = np.asarray(gp.path, np.float32).reshape(-1, 3)
    a = walls.arrays
    half = 0.5 * np.sqrt((a["ex"] - a["sx"]) ** 2 + (a["ey"] - a["sy"]) ** 2 + (a["ez"] - a["sz"]) ** 2)
    reach = max(gp.max_speed * params.min_time_to_wall, 4.0 * np.sqrt(2.0) * radius, gp.max_speed * dt + radius)
    cap = float(reach + (half.max() if walls.n else 0.0) + 1.0)
    return ChokeNewScenario(state=st, walls=walls, wall_mid=np.asarray(wall_mid, np.float32), path_points=pts,
                            path_radius=float(gp.path_radius), path_cyclic=bool(gp.path_cyclic),
                            seek_target=tuple(float(v) for v in pts[1]), world_size=tuple(params.dimensions),
                            world_cells=tuple(params.cells), params=params, dt=dt, wall_cap=cap)


@dataclass
class MultiGroupScenario:
    """The groups of plugins/MultiGroup.cpp (Group, :180-230) set up from a
    parsed .params file: one agent state and path per group, the world
    lattice and the per-step constants of Group::update (:558-611)."""
    states: Tuple[AgentState, ...]
    paths: Tuple[Tuple[np.ndarray, float, bool], ...]   # (points, radius, cyclic) or points empty
    world_size: Tuple[float, float, float]
    world_cells: Tuple[int, int, int]
    params: object
    dt: float
    # KNN range of every group search: no agent farther than this can be a
    # collision threat within minTimeToCollision (closing speed at most twice
    # the largest max speed), nor a separation, flocking or overlap neighbour.
    knn_cap: float


def multi_group(params, n_per_group: int = None, dt: float = 0.1, max_start_radius: float = None,
                start_scale: float = 1.0, workload: int = 1) -> MultiGroupScenario:
    """Agents of every group placed as Group::reset does (:430-451): radius
    uniform in [min, max] start radius on a random XZ direction around the
    group's start, random XZ heading, speed = max speed, radius 0.5, mass 1, no
    overlap test.  The draws are this module's own seeded stream.
    n_per_group / max_start_radius override the file; start_scale moves the
    groups' starts toward the origin (smaller test crowds that meet sooner)."""
    g = rng(workload, 11)
    states, paths = [], []
    for gp in params.groups:
        n = gp.num_agents if n_per_group is None else int(n_per_group)
        hi = gp.max_start_radius if max_start_radius is None else max_start_radius
        rr = g.uniform(gp.min_start_radius, hi, size=n)
        th = g.uniform(0.0, 2.0 * np.pi, size=n)
        c = np.asarray(gp.start_position, np.float64) * start_scale
        pos = np.stack([c[0] + rr * np.cos(th), np.full(n, c[1]), c[2] + rr * np.sin(th)], axis=1).astype(np.float32)
        hd = g.uniform(0.0, 2.0 * np.pi, size=n)
        fwd = np.stack([np.cos(hd), np.zeros(n), np.sin(hd)], axis=1)
        states.append(make_agents(pos, fwd, np.full(n, gp.max_speed, np.float32), 0.5, gp.max_speed, gp.max_force))
        paths.append((np.asarray(gp.path, np.float32).reshape(-1, 3), float(gp.path_radius), bool(gp.path_cyclic)))
    vmax = max(gp.max_speed for gp in params.groups)
    reach = max(2.0 * vmax * params.min_time_to_collision + 2.0 * 0.5, params.max_flocking_distance,
                params.max_separation_distance, params.min_separation_distance + 2.0 * 0.5)
    return MultiGroupScenario(states=tuple(states), paths=tuple(paths), world_size=tuple(params.dimensions),
                              world_cells=tuple(params.cells), params=params, dt=dt, knn_cap=float(reach + 1.0))
