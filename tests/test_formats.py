"""Scenario formats (host-side C-ABI, no GPU): wall maps, wall splitting at
bin-cell borders and simulation parameter files, on the reference's own data
files (gpusteer_amd/data, copied from data/), against the oracle's
restatement (oracle/formats_ref.py), hand-computed cases and properties the
splitting must have."""

import numpy as np
import pytest

from gpusteer_amd import formats
from gpusteer_amd._capi import GpuSteerError

import formats_ref as ref

FILES = formats.DATA_DIR
MAPS = sorted(FILES.glob("*.map"))
PARAMS = sorted(FILES.glob("*.params"))


def _as_rows(walls, mid):
    a = walls.arrays
    return [((a["sx"][i], a["sy"][i], a["sz"][i]), (a["ex"][i], a["ey"][i], a["ez"][i]),
             (a["nx"][i], a["ny"][i], a["nz"][i]), tuple(mid[i])) for i in range(walls.n)]


def _same(rows, ref_rows):
    assert len(rows) == len(ref_rows)
    for i, (r, q) in enumerate(zip(rows, ref_rows)):
        for part in range(4):
            np.testing.assert_array_equal(np.array(r[part], np.float32), np.array(q[part], np.float32),
                                          err_msg=f"wall {i} part {part}")


def test_fixture_files_present():
    assert len(MAPS) == 12 and len(PARAMS) == 7


@pytest.mark.parametrize("path", MAPS, ids=lambda p: p.name)
def test_load_map_equals_oracle(path):
    w = formats.WallSet.load_map(path)
    rows = _as_rows(*w.get())
    _same(rows, ref.load_map(path))
    # Every reference map holds the 6 walls of its choke point.
    assert len(rows) == 6


def test_load_map_drops_a_record_ending_the_file(tmp_path):
    """The reference keeps a record only while its stream is good: a last
    record whose final value ends the file (no line break) is dropped."""
    rec = "0 0 1  4 0 1  0 0 -1\n"
    p = tmp_path / "a.map"
    p.write_text(rec + rec.strip())
    assert len(formats.WallSet.load_map(p)) == 1 == len(ref.load_map(p))
    p.write_text(rec + rec)
    assert len(formats.WallSet.load_map(p)) == 2 == len(ref.load_map(p))
    p.write_text(rec + "1 2 3")  # incomplete record
    assert len(formats.WallSet.load_map(p)) == 1


LATTICES = [((635.0, 10.0, 635.0), (100, 1, 100)),     # data/MultiGroup.params
            ((200.0, 10.0, 200.0), (40, 1, 40)),
            ((300.0, 40.0, 300.0), (30, 4, 30))]


def test_split_of_the_choke_point_scenario():
    """data/ChokePoint.params: its map split by its own 512 x 1 x 512 lattice
    over 2048 x 10 x 2048 (1254 walls), against the oracle."""
    p = formats.load_params(FILES / "ChokePoint.params")
    path = FILES / p.map_filename
    w = formats.WallSet.load_map(path).split(p.dimensions, p.cells)
    rows = _as_rows(*w.get())
    assert len(rows) == 1254
    _same(rows, ref.split_walls(ref.load_map(path), ref.lattice_cells(p.dimensions, p.cells)))


@pytest.mark.parametrize("path", [MAPS[0], FILES / "ChokePoint15.map", FILES / "ChokePoint32.5.map"],
                         ids=lambda p: p.name)
@pytest.mark.parametrize("size,cells", LATTICES)
def test_split_equals_oracle(path, size, cells):
    w = formats.WallSet.load_map(path).split(size, cells)
    rows = _as_rows(*w.get())
    _same(rows, ref.split_walls(ref.load_map(path), ref.lattice_cells(size, cells)))
    # The explicit cell list (KNNBinData::hvCells) gives the same walls.
    lo, hi = formats.lattice_cells(size, cells)
    w2 = formats.WallSet.load_map(path).split_cells(lo, hi)
    _same(_as_rows(*w2.get()), rows)


@pytest.mark.parametrize("size,cells", LATTICES[0:2])
def test_split_properties(size, cells):
    """The pieces of each wall tile it end to end (shared endpoints, same
    normal), and no piece can be split again: every cell it enters, it
    enters within 0.001 of one of its ends."""
    path = FILES / "ChokePoint20.map"
    orig = ref.load_map(path)
    pieces = ref.split_walls(orig, ref.lattice_cells(size, cells))
    cl = ref.lattice_cells(size, cells)
    for s, e, n, _ in pieces:
        for lo, hi in cl:
            ip = ref._intersects(s, e, lo, hi)
            if ip is not None:
                assert not (ref._dist(s, ip) > ref.EPS and ref._dist(ip, e) > ref.EPS)
    for s0, e0, n0, _ in orig:
        # walk from the start through pieces sharing endpoints
        mine = [p for p in pieces if p[2] == n0 and _on_segment(p[0], s0, e0) and _on_segment(p[1], s0, e0)]
        cur, used = s0, 0
        while cur != e0:
            nxt = [p for p in mine if p[0] == cur]
            assert len(nxt) == 1, (cur, len(nxt))
            cur = nxt[0][1]
            used += 1
        assert used == len(mine)


def _on_segment(p, a, b, tol=1e-3):
    a, b, p = (np.array(x, np.float64) for x in (a, b, p))
    d = b - a
    t = np.dot(p - a, d) / np.dot(d, d)
    return -1e-6 <= t <= 1 + 1e-6 and np.linalg.norm(a + t * d - p) < tol


def test_split_hand_computed():
    """One wall along x at z = 0.5 from -3 to 3 through a lattice of 2-wide
    cells (x borders -2, 0, 2): it is cut at each border it enters, and the
    pieces come out in the order the list produces them -- [-3, -2] stays in
    front, [-2, 3] is cut at 0 and its tail [0, 3] at 2 -- with midpoints
    start + 0.5 * (end - start)."""
    from gpusteer_amd.state import WallState

    walls = WallState({"sx": np.array([-3.0], np.float32), "sy": np.zeros(1, np.float32),
                       "sz": np.array([0.5], np.float32), "ex": np.array([3.0], np.float32),
                       "ey": np.zeros(1, np.float32), "ez": np.array([0.5], np.float32),
                       "nx": np.zeros(1, np.float32), "ny": np.zeros(1, np.float32),
                       "nz": np.ones(1, np.float32)})
    w = formats.WallSet.from_walls(walls).split((8.0, 10.0, 8.0), (4, 1, 4))
    out, mid = w.get()
    np.testing.assert_array_equal(out.arrays["sx"], [-3, -2, 0, 2])
    np.testing.assert_array_equal(out.arrays["ex"], [-2, 0, 2, 3])
    np.testing.assert_array_equal(mid[:, 0], [-2.5, -1, 1, 2.5])
    np.testing.assert_array_equal(mid[:, 2], [0.5] * 4)


@pytest.mark.parametrize("path", PARAMS, ids=lambda p: p.name)
def test_params_equal_oracle(path):
    p = formats.load_params(path)
    q = ref.load_params(path)
    assert p.dimensions == tuple(np.float32(x) for x in q["world"]["dimensions"])
    assert p.cells == q["world"]["cells"]
    assert p.obstacle_count == q["world"]["obstacle_count"]
    assert p.map_filename == q["world"]["map_filename"]
    assert len(p.groups) == len(q["groups"])
    for g, h in zip(p.groups, q["groups"]):
        assert g.num_agents == h["num_agents"]
        for f in ("max_speed", "max_force", "min_start_radius", "max_start_radius", "path_radius"):
            assert np.float32(getattr(g, f)) == h[f], f
        assert g.start_position == tuple(np.float32(x) for x in h["start_position"])
        assert g.body_color == tuple(np.float32(x) for x in h["body_color"])
        assert [tuple(np.float32(v) for v in pt) for pt in g.path] == h["path"]
        assert g.path_cyclic == bool(h["path_cyclic"])
    for k, v in q.items():
        if k in ("world", "groups"):
            continue
        assert getattr(p, k) == v, k


def test_params_values_from_the_file_text():
    """Values read off data/ChokePoint.params by hand."""
    p = formats.load_params(FILES / "ChokePoint.params")
    assert (p.seed, p.knn, p.kno, p.knw, p.search_radius) == (12345, 6, 2, 3, 1)
    assert p.cos_max_flocking_angle == np.float32(-0.86602540378443864676372317075294)
    assert (p.mask_seek, p.mask_flee, p.mask_follow_path, p.mask_avoid_walls) == (4480, 4496, 16, 0)
    assert p.weight_alignment == np.float32(2.5) and p.weight_separation == 10.0
    assert p.dimensions == (2048.0, 10.0, 2048.0) and p.cells == (512, 1, 512)
    assert p.map_filename == "ChokePoint20.map"
    (g,) = p.groups
    assert (g.num_agents, g.max_speed, g.max_force, g.start_position) == (131072, 3.0, 3.0, (0.0, 0.0, 512.0))
    assert (g.min_start_radius, g.max_start_radius, g.path_radius, g.path_cyclic) == (0.0, 450.0, 20.0, False)
    assert g.path == [(0.0, 0.0, 13.5), (0.0, 0.0, -1650.0)]
    # commented-out key: MultiGroup.params has "#mapFilename ..."
    assert formats.load_params(FILES / "MultiGroup.params").map_filename == ""


def test_params_errors(tmp_path):
    with pytest.raises(GpuSteerError, match="unknown key 'worldDimensions"):
        formats.load_params(FILES / "10kAgentsChoke.param")
    with pytest.raises(GpuSteerError, match="cannot open"):
        formats.load_params(tmp_path / "missing.params")
    p = tmp_path / "open.params"
    p.write_text("seed = 1\nknn = 7\n")
    with pytest.raises(GpuSteerError, match="unexpected end of file"):
        formats.load_params(p)
    p.write_text("seed = x\n</simulation>\n")
    with pytest.raises(GpuSteerError, match="bad unsigned integer"):
        formats.load_params(p)
