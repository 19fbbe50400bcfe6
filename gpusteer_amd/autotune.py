"""Autotuner: picks the bin lattice (cells along x, y, z -- z may be finer, so
the z-runs the KNN scans are trimmed more tightly) and the KNN fast path's
column radius (3 x 3 or 5 x 5 columns) for a flock, the launch shapes (block
size, agents per lane) of the k = 7 search and flocking kernels, and the
cell-size factor of an ensemble's per-scenario lattices, by timing short
runs of the real step on the device.  Every candidate computes the same
(exact) result; only the time differs."""
from __future__ import annotations

import copy
import time
from typing import List, Optional, Sequence, Tuple

from . import api
from ._capi import GS_MAX_CELLS

Cells = Tuple[int, int, int]
# The largest lattice gs_bins_create accepts.
MAX_CELLS = GS_MAX_CELLS
# Ties go to the earlier candidate (pick), so 48 x 48 x 144, the k = 16
# lattice, comes after the two that k = 7 picks between.
BOIDS_CANDIDATES: Tuple[Tuple[Cells, int], ...] = (
    ((48, 48, 48), 1), ((64, 64, 64), 1), ((48, 48, 96), 1), ((64, 64, 128), 1), ((64, 64, 192), 1),
    ((48, 48, 144), 1), ((80, 80, 160), 1), ((96, 96, 96), 2), ((128, 128, 128), 2))


# A candidate later in the list replaces an earlier one only when it is
# faster by more than this fraction, so timings within run-to-run noise
# give the same choice every time.
TIE_MARGIN = 0.02


def pick(times: Sequence[float], margin: float = TIE_MARGIN) -> int:
    """Index of the first candidate within `margin` of the fastest."""
    best = min(times)
    return next(i for i, t in enumerate(times) if t <= best * (1.0 + margin))


def scaled_candidates(n: int, base: Sequence[Tuple[Cells, int]] = BOIDS_CANDIDATES) -> List[Tuple[Cells, int]]:
    """The candidate lattices (set for 2^20 agents) scaled to n agents in the
    same volume: every axis by (n / 2^20)^(1/3), rounded to a multiple of 8,
    so the agents per cell stay what the base set was chosen for; lattices
    above the library's 2^28 cells are left out."""
    f = (n / float(1 << 20)) ** (1.0 / 3.0)
    out: List[Tuple[Cells, int]] = []
    for cells, rf in base:
        c = tuple(max(8, int(round(v * f / 8.0)) * 8) for v in cells)
        if c[0] * c[1] * c[2] <= MAX_CELLS and (c, rf) not in out:
            out.append((c, rf))
    return out


def tune_flock(ctx: api.Context, state, params, k: int = 7,
               candidates: Optional[Sequence[Tuple[Cells, int]]] = None, steps: int = 10, warmup: int = 3):
    """Times `steps` fused Boids steps (after `warmup`) of the given flock for
    each ((cells x, y, z), column radius) candidate (by default the base set
    scaled to the flock's size); returns the fastest candidate and the table
    of (cells, radius, ms per step)."""
    if candidates is None:
        candidates = scaled_candidates(state.n)
    table: List[Tuple[Cells, int, float]] = []
    for cells, rf in candidates:
        p = copy.deepcopy(params)
        p.grid_cells = tuple(cells)
        p.search_radius = rf
        A = api.AgentGroup(ctx, state)
        bins = api.KNNBinData(ctx, p.grid_cells, p.grid_size, origin=p.grid_origin)
        knn = api.KNNData(ctx, state.n, k)
        api.stepBoids(A, knn, bins, p, warmup)
        _, _, _, step_ms = api.stepBoidsTimed(A, knn, bins, p, steps)
        table.append((cells, rf, step_ms))
        knn.close()
        bins.close()
        A.close()
    best = table[pick([r[2] for r in table])]
    return (best[0], best[1]), table


# (threads per block, agents per lane) of the k = 7 search and flocking kernels.
# The library's default first: another shape is chosen only when clearly faster.
SHAPES: Tuple[Tuple[int, int], ...] = ((256, 1), (128, 1), (512, 1), (128, 2), (256, 2), (512, 2))


def tune_shapes(ctx: api.Context, state, params, k: int = 7, steps: int = 10, warmup: int = 3):
    """Times `steps` fused Boids steps of the flock (lattice and radius as in
    params) for each launch shape, set for both tuned kernels at once, and
    keeps for each kernel the shape of its lowest HIP-event time; sets them
    on ctx and returns ((knn shape, flock shape), table of (shape, knn ms,
    flocking ms, step ms)).  Only k = 7 has tunable shapes."""
    from ._capi import GS_SHAPE_FLOCK, GS_SHAPE_KNN

    if k != 7:
        return (ctx.shape(GS_SHAPE_KNN), ctx.shape(GS_SHAPE_FLOCK)), []
    table = []
    for shape in SHAPES:
        ctx.set_shape(GS_SHAPE_KNN, *shape)
        ctx.set_shape(GS_SHAPE_FLOCK, *shape)
        A = api.AgentGroup(ctx, state)
        bins = api.KNNBinData(ctx, params.grid_cells, params.grid_size, origin=params.grid_origin)
        knn = api.KNNData(ctx, state.n, k)
        api.stepBoids(A, knn, bins, params, warmup)
        _, knn_ms, flock_ms, step_ms = api.stepBoidsTimed(A, knn, bins, params, steps)
        table.append((shape, knn_ms, flock_ms, step_ms))
        knn.close()
        bins.close()
        A.close()
    best_knn = table[pick([r[1] for r in table])][0]
    best_flock = table[pick([r[2] for r in table])][0]
    ctx.set_shape(GS_SHAPE_KNN, *best_knn)
    ctx.set_shape(GS_SHAPE_FLOCK, *best_flock)
    return (best_knn, best_flock), table


def tune_ensemble_cells(ctx: api.Context, make_specs, factors: Sequence[float] = (0.8, 1.0, 1.25, 1.6, 2.0, 2.5, 3.2),
                        steps: int = 6):
    """Times an ensemble built by make_specs(cell_factor) for each factor
    (the first step after the upload and the graph capture excluded);
    returns the fastest factor and the table of (factor, ms per step)."""
    from .ensemble import Ensemble

    table: List[Tuple[float, float]] = []
    for f in factors:
        ens = Ensemble(ctx, make_specs(f))
        ens.step(3)
        ctx.synchronize()
        t0 = time.perf_counter()
        ens.step(steps)
        ctx.synchronize()
        table.append((f, (time.perf_counter() - t0) * 1e3 / steps))
        ens.close()
    best = table[pick([r[1] for r in table])]
    return best[0], table
