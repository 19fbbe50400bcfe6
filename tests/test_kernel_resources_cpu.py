"""Register and scratch budgets of the hot kernels, read from the code objects
the build wrote (gpusteer_amd/_build/*.o; no GPU): a change that pushes a
search or flocking kernel past its occupancy target or into spilling in its
hot loop shows up here before any timing does.  The budgets are the
measured figures of the tuned kernels with a little headroom."""
import re
import subprocess
import tempfile
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parents[1]
OBJ = ROOT / "gpusteer_amd" / "_build"
LLVM = Path("/opt/rocm/lib/llvm/bin")


def _kernels(obj: Path) -> dict:
    """name -> {vgpr_count, vgpr_spill_count, private_segment_fixed_size,
    group_segment_fixed_size} from the object's gfx950 code object."""
    with tempfile.TemporaryDirectory() as d:
        fb, co = Path(d) / "fb.bin", Path(d) / "k.co"
        r = subprocess.run([str(LLVM / "llvm-objcopy"), f"--dump-section=.hip_fatbin={fb}", str(obj),
                            str(Path(d) / "rest.o")], capture_output=True)
        if r.returncode:
            return {}  # no device code in this object
        subprocess.run([str(LLVM / "clang-offload-bundler"), "--unbundle", "--type=o", f"--input={fb}",
                        "--targets=hipv4-amdgcn-amd-amdhsa--gfx950", f"--output={co}"], check=True,
                       capture_output=True)
        notes = subprocess.run([str(LLVM / "llvm-readelf"), "--notes", str(co)], check=True, capture_output=True,
                               text=True).stdout
    out = {}
    # One YAML list item per kernel; its keys come in alphabetical order.
    for item in re.split(r"\n\s+- \.", notes)[1:]:
        m = re.search(r"\.name:\s+(\S+)", item)
        if not m:
            continue
        d = {k: int(v) for k, v in re.findall(
            r"\.(vgpr_count|vgpr_spill_count|private_segment_fixed_size|group_segment_fixed_size):\s+(\d+)", "." + item)}
        out[m.group(1)] = d
    return out


@pytest.fixture(scope="module")
def kernels():
    objs = sorted(OBJ.glob("*.o"))
    if not objs or not (LLVM / "clang-offload-bundler").exists():
        pytest.skip("library objects or LLVM tools not present (run the build first)")
    ks = {}
    for o in objs:
        ks.update(_kernels(o))
    assert ks, "no kernels found in the build objects"
    return ks


def _find(ks, pattern):
    hits = {n: d for n, d in ks.items() if re.search(pattern, n)}
    assert hits, f"no kernel matches {pattern}"
    return hits


# (pattern of the mangled name, max VGPRs, max spilled VGPRs)
BUDGETS = [
    (r"k_knn_binnedILi7ELi1ELi256ELi1E", 96, 64),   # k = 7 flat search: 5 wavefronts per SIMD
    (r"k_knn_binnedILi7ELi1ELi128ELi1E", 96, 64),
    (r"k_knn_binnedILi7ELi2E", 128, 0),             # 5 x 5 walk (choke-point plugin lattices)
    (r"k_knn_binnedILi8ELi1E", 96, 96),
    (r"k_knn_binnedILi8ELi2E", 136, 0),
    (r"k_knn_binnedILi16ELi1E", 128, 300),          # 4 wavefronts per SIMD (fewer spills were slower)
    (r"k_flock_updateILi7ELi\dELi\d+ELi1E", 96, 0),   # one agent per lane: 5 wavefronts per SIMD
    (r"k_flock_updateILi7ELi\dELi\d+ELi2E", 128, 0),
    (r"k_flock_updateILi16E", 112, 0),               # 16-entry lists summed 8 at a time: 4 wavefronts per SIMD
    (r"e_knnILi7ELi0E", 96, 128),                    # ensemble own-group search: 5 wavefronts per SIMD
    (r"e_knnILi8ELi0E", 96, 144),
    (r"e_knnILi7ELi1E", 152, 0),                     # ensemble crowd lists (other group, walls, obstacles)
    (r"e_knnILi8ELi1E", 152, 0),
    (r"k_knn_split_part", 400, 0),
    (r"k_knn_split_merge", 128, 0),
    (r"c4_backward", 128, 0),
    (r"c4_forward", 128, 0),
    (r"k_reorder", 64, 0),
]


@pytest.mark.parametrize("pattern,max_vgpr,max_spill", BUDGETS)
def test_kernel_register_budgets(kernels, pattern, max_vgpr, max_spill):
    for name, d in _find(kernels, pattern).items():
        assert d.get("vgpr_count", 0) <= max_vgpr, f"{name}: {d}"
        assert d.get("vgpr_spill_count", 0) <= max_spill, f"{name}: {d}"


def test_lds_fits_every_block(kernels):
    for name, d in kernels.items():
        assert d.get("group_segment_fixed_size", 0) <= 64 * 1024, f"{name}: {d}"
