"""GPU parity of the per-cell extraction (gs_group_cell_density,
gs_group_cell_velocity) and its text writers against the oracle
(oracle/extract_ref.py over the oracle's bin assignment), bit for bit and
byte for byte."""
import numpy as np
import pytest

from gpusteer_amd import scenarios

import extract_ref as ex

pytestmark = pytest.mark.gpu


def _flock(n, seed, dims):
    g = scenarios.rng(seed)
    pos = g.uniform(-45, 45, size=(n, 3)).astype(np.float32)
    if dims == 2:
        pos[:, 1] = 0.0
    fwd = scenarios.unit_vectors(g, n)
    st = scenarios.make_agents(pos, fwd, g.uniform(0, 3, n).astype(np.float32), 0.5, 3.0, 3.0)
    st.arrays["gx"][:] = g.normal(size=n)
    st.arrays["gz"][:] = g.normal(size=n)
    return st


@pytest.mark.parametrize("cells,dims", [((40, 1, 40), 2), ((12, 7, 9), 3), ((2, 1, 2), 2)],
                         ids=["2d", "3d", "10000-per-cell"])
def test_cell_density_and_velocity(gs, ctx, orc, cells, dims):
    """40,000 agents binned, then integrated once without a rebuild (the
    velocities are the current ones, the cells those of the build): density
    and mean velocity per cell equal the oracle's in KNNBinDataV2 numbering."""
    st = _flock(40000, 61, dims)
    size = (100.0, 100.0, 100.0)
    bins = gs.KNNBinData(ctx, cells, size)
    A = gs.AgentGroup(ctx, st)
    gs.updateKNNDatabase(A, bins)
    gs.updateGroup(A, 0.1, gs.MODE_GPUREF)
    ref = st.copy()
    orc.update(ref, 0.1, orc.MODE_GPUREF)
    key, _, _, _ = orc.bins_build(orc.grid((-50.0,) * 3, size, cells), np.stack(
        [st.arrays[c] for c in ("px", "py", "pz")], axis=1))
    C = cells[0] * cells[1] * cells[2]
    d = A.cell_density(C)
    v = A.cell_velocity(C)
    np.testing.assert_array_equal(d, ex.cell_density(key, cells))
    fwd = np.stack([ref.arrays[c] for c in ("fx", "fy", "fz")], axis=1)
    np.testing.assert_array_equal(v, ex.cell_velocity(key, cells, fwd, ref.arrays["speed"]))
    assert d.sum() == st.n and (d > 1).any()
    for h in (A, bins):
        h.close()


def test_cell_writers(gs, ctx, orc, tmp_path):
    """The writers' files equal the oracle writers' text; selected cells get
    one appended line per call."""
    cells = (30, 1, 30)
    st = _flock(20000, 62, 2)
    size = (100.0, 10.0, 100.0)
    bins = gs.KNNBinData(ctx, cells, size)
    A = gs.AgentGroup(ctx, st)
    gs.updateKNNDatabase(A, bins)
    key, _, _, _ = orc.bins_build(orc.grid((-50.0, -5.0, -50.0), size, cells), np.stack(
        [st.arrays[c] for c in ("px", "py", "pz")], axis=1))
    dens = ex.cell_density(key, cells)
    vel = ex.cell_velocity(key, cells, np.stack([st.arrays[c] for c in ("fx", "fy", "fz")], axis=1),
                           st.arrays["speed"])
    sel = [0, 465, 899]
    for all_rows in (False, True):
        for _ in range(2):
            A.write_cell_density(tmp_path / f"d{int(all_rows)}", sel, tmp_path / f"sd{int(all_rows)}", all_rows)
            A.write_cell_velocity(tmp_path / f"v{int(all_rows)}", sel, tmp_path / f"sv{int(all_rows)}", all_rows)
        ex.write_density(tmp_path / f"rd{int(all_rows)}", dens, cells, all_rows)
        ex.write_velocity(tmp_path / f"rv{int(all_rows)}", vel, cells, all_rows)
        for a, b in (("d", "rd"), ("v", "rv")):
            assert (tmp_path / f"{a}{int(all_rows)}.dat").read_text() == \
                (tmp_path / f"{b}{int(all_rows)}.dat").read_text()
        for c in sel:
            assert (tmp_path / f"sd{int(all_rows)}_{c}.dat").read_text() == f"{dens[c]}\n" * 2
            assert (tmp_path / f"sv{int(all_rows)}_{c}.dat").read_text() == \
                f"{float(vel[c][0]):g}\t{float(vel[c][2]):g}\n" * 2
    for h in (A, bins):
        h.close()
