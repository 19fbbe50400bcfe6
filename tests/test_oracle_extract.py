"""Pins the extraction oracle (oracle/extract_ref.py) by hand-computed cells:
counts and means of a few placed agents, the KNNBinDataV2 numbering of a
3-D lattice, and the writers' text for a 3 x 1 x 3 lattice."""
import numpy as np

import extract_ref as ex


def test_density_and_numbering():
    """Cells (ix, iy, iz) of a 3 x 2 x 3 lattice in lq.c numbering
    ix*6 + iy*3 + iz; KNNBinDataV2 numbers them ix + iz*3 + iy*9."""
    cells = (3, 2, 3)
    lq = [0 * 6 + 0 * 3 + 0, 2 * 6 + 0 * 3 + 1, 2 * 6 + 0 * 3 + 1, 1 * 6 + 1 * 3 + 2]
    d = ex.cell_density(lq, cells)
    assert d.sum() == 4 and d.shape == (18,)
    assert d[0] == 1 and d[2 + 1 * 3] == 2 and d[1 + 2 * 3 + 1 * 9] == 1


def test_velocity_means():
    cells = (3, 1, 3)
    lq = [4, 4, 0]  # ix = 1, iz = 1 twice; the origin cell once
    fwd = np.array([[1, 0, 0], [0, 0, 1], [0.6, 0, 0.8]], np.float32)
    speed = np.array([2, 4, 5], np.float32)
    v = ex.cell_velocity(lq, cells, fwd, speed)
    np.testing.assert_array_equal(v[1 + 1 * 3], [1.0, 0.0, 2.0])
    np.testing.assert_array_equal(v[0], np.array([0.6, 0, 0.8], np.float32) * np.float32(5))
    assert not v[np.setdiff1d(np.arange(9), [0, 4])].any()


def test_velocity_sum_order():
    """Three agents in one cell summed in index order then divided: 1e8 +
    1 - 1e8 in float32 is 0, not 1 (the order of the reference's loop)."""
    v = ex.cell_velocity([0, 0, 0], (1, 1, 1), np.array([[1, 0, 0]] * 3, np.float32),
                         np.array([1e8, 1, -1e8], np.float32))
    assert v[0][0] == 0.0


def test_writers_text(tmp_path):
    cells = (3, 1, 3)
    d = np.arange(9, dtype=np.uint32)
    ex.write_density(tmp_path / "d", d, cells)
    assert (tmp_path / "d.dat").read_text() == (
        "%\tCell densities\n%\tworldCells:\tX: 3\tY: 1\tZ: 3\n%\tData format: \n%\tX\tY\tnumAgents\n"
        "0\t1\t2\t\n3\t4\t5\t\n")
    v = np.zeros((9, 3), np.float32)
    v[4] = [0.5, 0.0, -1.25]
    ex.write_velocity(tmp_path / "v", v, cells)
    lines = (tmp_path / "v.dat").read_text().splitlines()
    assert lines[3] == "%\tX\tZ\tV.x\tV.z" and len(lines) == 4 + 6
    assert lines[4 + 4] == "1\t1\t0.5\t-1.25"
    ex.write_density(tmp_path / "a", d, cells, all_rows=True)
    assert (tmp_path / "a.dat").read_text().splitlines()[-1] == "6\t7\t8\t"
