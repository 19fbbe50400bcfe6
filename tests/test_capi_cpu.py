"""C-ABI checks that need no GPU: the library loads, exports every function
include/gpusteer.h declares, and rejects bad handles / arguments with status
codes instead of aborting."""
import ctypes as C
import re
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parents[1]


def declared_functions():
    text = (ROOT / "include" / "gpusteer.h").read_text()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"^\s*(?:const\s+)?\w+\*?\s+\*?(gs_\w+)\s*\(", text, flags=re.M)))


def test_library_exports_every_declared_symbol():
    from gpusteer_amd import _capi

    names = declared_functions()
    assert len(names) >= 35
    for n in names:
        assert hasattr(_capi.lib, n), f"{n} declared in gpusteer.h but not exported"
    # The ctypes table covers the whole header too.
    assert set(names) <= set(_capi.EXPORTED)


def test_enum_constants_match_header():
    """Every GS_* enum constant the header defines has the same value in
    the ctypes table."""
    from gpusteer_amd import _capi

    text = re.sub(r"/\*.*?\*/", "", (ROOT / "include" / "gpusteer.h").read_text(), flags=re.S)
    found = 0
    for body in re.findall(r"enum\s*\{(.*?)\}", text, flags=re.S):
        for name, expr in re.findall(r"(GS_\w+)\s*=\s*([^,]+)", body):
            m = re.fullmatch(r"(\d+)u?(?:\s*<<\s*(\d+))?", expr.strip())
            assert m, f"{name}: unexpected initializer {expr!r}"
            want = int(m.group(1)) << int(m.group(2) or 0)
            assert getattr(_capi, name) == want, name
            found += 1
    assert found >= 10


def test_status_strings():
    from gpusteer_amd import _capi

    for s in range(10):
        assert _capi.lib.gs_status_string(s) != b"unknown status"
    assert _capi.lib.gs_version() == 100


def test_null_handles_rejected():
    from gpusteer_amd import _capi as c

    L = c.lib
    assert L.gs_group_destroy(None) == c.GS_ERR_INVALID_HANDLE
    assert L.gs_update_knn_database(None, None) == c.GS_ERR_INVALID_HANDLE
    assert L.gs_find_knn(None, None, None, None, 1, 0) == c.GS_ERR_INVALID_HANDLE
    assert L.gs_steer_separation(None, None, None, 1, 2, 0, 1, 0, 0) == c.GS_ERR_INVALID_HANDLE
    assert L.gs_update_group(None, 0.1, 0) == c.GS_ERR_INVALID_HANDLE
    assert L.gs_knn_download(None, None, None) == c.GS_ERR_INVALID_HANDLE
    assert L.gs_update_group_anti_penetration(None, None, None, 0.1, 0) == c.GS_ERR_INVALID_HANDLE
    assert L.gs_anti_penetration_agents(None, None, None, 0) == c.GS_ERR_INVALID_HANDLE
    assert L.gs_steer_follow_path(None, None, 1.0, 1.0, 0) == c.GS_ERR_INVALID_HANDLE
    assert L.gs_steer_avoid_neighbors(None, None, None, 1.0, 0.5, 0, 1.0, 0, 1) == c.GS_ERR_INVALID_HANDLE
    assert L.gs_path_destroy(None) == c.GS_ERR_INVALID_HANDLE
    assert L.gs_group_cell_density(None, None) == c.GS_ERR_INVALID_HANDLE
    assert L.gs_write_cell_velocity(None, b"x", None, 0, None, 0) == c.GS_ERR_INVALID_HANDLE
    assert b"invalid" in L.gs_last_error()
    # A pointer that was never a handle is rejected, not dereferenced.
    bogus = C.c_void_p(0x1234)
    assert L.gs_context_synchronize(bogus) == c.GS_ERR_INVALID_HANDLE
    assert L.gs_step_boids(bogus, bogus, bogus, None, 1) == c.GS_ERR_INVALID_HANDLE
    # Frames: bad handles, and no recording to end.
    f = C.c_void_p()
    assert L.gs_frame_begin(None) == c.GS_ERR_INVALID_HANDLE
    assert L.gs_frame_end(None, C.byref(f)) == c.GS_ERR_INVALID_HANDLE
    assert L.gs_frame_launch(bogus, 1) == c.GS_ERR_INVALID_HANDLE
    assert L.gs_frame_destroy(bogus) == c.GS_ERR_INVALID_HANDLE
    assert L.gs_frame_destroy(None) == c.GS_OK
    assert L.gs_context_set_shape(None, c.GS_SHAPE_KNN, 256, 1) == c.GS_ERR_INVALID_HANDLE
    assert L.gs_context_get_shape(bogus, c.GS_SHAPE_FLOCK, None, None) == c.GS_ERR_INVALID_HANDLE


def test_context_without_gpu_reports_device_error():
    import torch

    if torch.cuda.is_available():
        pytest.skip("a GPU is present")
    from gpusteer_amd import _capi as c

    h = C.c_void_p()
    st = c.lib.gs_context_create(0, None, C.byref(h))
    assert st == c.GS_ERR_DEVICE
    assert not h.value


def test_package_import_requires_built_library(tmp_path):
    # The package has no CPU fallback: its loader raises when the .so is absent.
    import importlib.util

    src = (ROOT / "gpusteer_amd" / "_capi.py").read_text()
    assert "raise ImportError" in src
    spec = importlib.util.find_spec("gpusteer_amd")
    assert spec is not None


def test_ensemble_specs_and_struct():
    """Scenario generator and the gs_scenario struct (no GPU needed)."""
    from gpusteer_amd import _capi as capi
    from gpusteer_amd import scenarios
    from gpusteer_amd.ensemble import scenario_struct

    specs = scenarios.ensemble_mix(n_scenarios=8, n_min=1024, n_max=65536, workload=3)
    assert [s[0] for s in specs] == ["boids", "choke"] * 4
    for s in specs:
        st = scenario_struct(s)
        if s[0] == "boids":
            assert st.kind == capi.GS_SCENARIO_BOIDS and st.n[0] == s[1].n and st.wrap_shape == capi.GS_WRAP_SPHERE
            assert 1024 <= s[1].n <= 65536
        else:
            assert st.kind == capi.GS_SCENARIO_CHOKE and st.n[0] + st.n[1] == 2 * s[1].groups[0].n
            assert st.cells[1] == 1 and st.k_walls == 3 and st.k_obstacles == 2
    # Lattices follow density: more agents, more cells.
    assert scenarios.boids_cells_for(65536) > scenarios.boids_cells_for(1024)
    assert scenarios.choke_cells_for(65536) > scenarios.choke_cells_for(1024)


def test_lattice_limits_match_header():
    """The lattice limits the header defines are the ones the ctypes table
    and the autotuner use."""
    from gpusteer_amd import _capi, autotune

    text = (ROOT / "include" / "gpusteer.h").read_text()
    for name in ("GS_MAX_CELLS", "GS_MAX_SPLIT_CELLS"):
        m = re.search(rf"#define {name} \(1u << (\d+)\)", text)
        assert m, name
        assert getattr(_capi, name) == 1 << int(m.group(1))
    assert autotune.MAX_CELLS == _capi.GS_MAX_CELLS
