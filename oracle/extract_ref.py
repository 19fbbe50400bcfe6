"""Oracle (test infrastructure only; never imported by the product): a
numpy-float32 restatement of the per-cell extraction of a bin database and
its text writers, written from the reference's behaviour.

* cell_density: ComputeCellDensityKernel (include/OpenSteer/CUDA/ProcessData.cu:132-165):
  agents per cell (cellEnd - cellStart; 0 for an empty cell).
* cell_velocity: ComputeAvgCellVelocityKernel (:194-237): the sum of
  direction * speed over the cell's agents in the database's sorted order
  (ascending agent index, a stable sort by cell), one float32 addition at a
  time, divided by the count converted to float; (0, 0, 0) for an empty cell.
* Both in KNNBinDataV2's cell numbering ix + iz X + iy X Z
  (CUDA/KNNBinDataV2.cu:103-135).
* write_density / write_velocity: WriteCellDensity / WriteAvgCellVelocity
  (src/ExtractData.cpp:19-140): header, rows 0 .. Z-2 (the reference's loop
  bound, :63,132) of cells i X + j; floats as std::ostream prints them (%g).
"""
from __future__ import annotations

import numpy as np

F = np.float32


def _v2(cell_lq, cells):
    X, Y, Z = cells
    iz = cell_lq % Z
    iy = (cell_lq // Z) % Y
    ix = cell_lq // (Z * Y)
    return ix + iz * X + iy * X * Z


def cell_density(cell_lq, cells):
    """cell_lq[n]: each agent's cell in lq.c numbering (ix dy dz + iy dz + iz)."""
    C = cells[0] * cells[1] * cells[2]
    return np.bincount(_v2(np.asarray(cell_lq, np.int64), cells), minlength=C).astype(np.uint32)


def cell_velocity(cell_lq, cells, fwd, speed):
    """fwd[n][3], speed[n] (float32): per cell mean of fwd * speed."""
    C = cells[0] * cells[1] * cells[2]
    v2 = _v2(np.asarray(cell_lq, np.int64), cells)
    vel = (np.asarray(fwd, F) * np.asarray(speed, F)[:, None]).astype(F)
    s = np.zeros((C, 3), F)
    np.add.at(s, v2, vel)  # unbuffered, in agent index order: one float32 addition at a time
    cnt = np.bincount(v2, minlength=C)
    nz = cnt > 0
    s[nz] = (s[nz] / cnt[nz, None].astype(F)).astype(F)
    return s


def write_density(prefix, dens, cells, all_rows=False):
    X, Y, Z = cells
    out = [f"%\tCell densities\n%\tworldCells:\tX: {X}\tY: {Y}\tZ: {Z}\n%\tData format: \n%\tX\tY\tnumAgents\n"]
    for i in range(Z if all_rows else Z - 1):
        out.append("".join(f"{int(dens[i * X + j])}\t" for j in range(X)) + "\n")
    open(f"{prefix}.dat", "w").write("".join(out))


def write_velocity(prefix, vel, cells, all_rows=False):
    X, Y, Z = cells
    out = [f"%\tCell average velocities\n%\tworldCells:\tX: {X}\tY: {Y}\tZ: {Z}\n%\tData format: \n"
           "%\tX\tZ\tV.x\tV.z\n"]
    for i in range(Z if all_rows else Z - 1):
        for j in range(X):
            c = i * X + j
            out.append(f"{i}\t{j}\t{float(vel[c][0]):g}\t{float(vel[c][2]):g}\n")
    open(f"{prefix}.dat", "w").write("".join(out))
