"""Builds libgpusteer.so (HIP, gfx950) in-tree.

Every .hip file under csrc/ is compiled in parallel with hipcc and linked into
gpusteer_amd/lib/libgpusteer.so.  Floating-point flags are part of the
library's contract with its parity tests:

* -ffp-contract=off: no fused multiply-add; every fp32 operation rounds as
  written, like the CPU oracle built with the same flag.
* -fhip-fp32-correctly-rounded-divide-sqrt: IEEE division and square root.
* -fno-gpu-flush-denormals-to-zero: IEEE denormals.
* -Xclang -target-feature -Xclang -packed-fp32-ops: no packed fp32 ops.
"""
from __future__ import annotations

import concurrent.futures as cf
import os
import shutil
import subprocess
import sys
from pathlib import Path

PKG = Path(__file__).resolve().parent
CSRC = PKG / "csrc"
LIB_DIR = PKG / "lib"
LIB = LIB_DIR / "libgpusteer.so"
OBJ_DIR = PKG / "_build"
ARCH = os.environ.get("GPUSTEER_ARCH", "gfx950")

HIPCC_FLAGS = [
    f"--offload-arch={ARCH}",
    "-O3",
    "-std=c++17",
    "-fPIC",
    "-ffp-contract=off",
    "-fhip-fp32-correctly-rounded-divide-sqrt",
    "-fno-gpu-flush-denormals-to-zero",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall",
    "-Wno-unused-function",
    "-Wno-unused-command-line-argument",
]


def hipcc() -> str:
    exe = shutil.which("hipcc") or "/opt/rocm/bin/hipcc"
    if not Path(exe).exists():
        raise RuntimeError("hipcc not found: the HIP toolchain is required to build libgpusteer")
    return exe


def _compile(src: Path) -> Path:
    obj = OBJ_DIR / (src.stem + ".o")
    cmd = [hipcc(), *HIPCC_FLAGS, "-c", str(src), "-o", str(obj)]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"hipcc failed on {src.name}:\n{r.stderr[-6000:]}")
    return obj


def build(verbose: bool = False) -> Path:
    OBJ_DIR.mkdir(exist_ok=True)
    LIB_DIR.mkdir(exist_ok=True)
    sources = sorted(CSRC.glob("*.hip"))
    with cf.ThreadPoolExecutor(max_workers=min(8, len(sources))) as ex:
        objs = list(ex.map(_compile, sources))
    tmp = LIB.with_suffix(".so.tmp")
    cmd = [hipcc(), f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", str(tmp), *map(str, objs)]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"link failed:\n{r.stderr[-6000:]}")
    os.replace(tmp, LIB)
    build_cxx_examples()
    if verbose:
        print(f"built {LIB}")
    return LIB


ROOT = PKG.parent
CXX_EXAMPLES = {"boids_group_plugin": ROOT / "tests" / "cxx" / "boids_group_plugin.cpp",
                "choke_setup": ROOT / "tests" / "cxx" / "choke_setup.cpp",
                "choke_point_plugin": ROOT / "tests" / "cxx" / "choke_point_plugin.cpp",
                "multi_group_plugin": ROOT / "tests" / "cxx" / "multi_group_plugin.cpp",
                "ctf_seeker_enemies": ROOT / "tests" / "cxx" / "ctf_seeker_enemies.cpp"}


def build_cxx_examples() -> None:
    """Host-only C++ programs written against include/gpusteer_opensteer.hpp
    (the reference's group API over the C-ABI), linked to libgpusteer.so."""
    for name, src in CXX_EXAMPLES.items():
        if not src.exists():
            continue
        out = LIB_DIR / name
        cmd = [hipcc(), "--offload-arch=gfx950", "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops", "-O2",
               "-ffp-contract=off", "-std=c++17", f"-I{ROOT / 'include'}", str(src), f"-L{LIB_DIR}", "-lgpusteer",
               "-Wl,-rpath,$ORIGIN", "-o", str(out)]
        r = subprocess.run(cmd, capture_output=True, text=True)
        if r.returncode != 0:
            raise RuntimeError(f"hipcc failed on {src.name}:\n{r.stderr[-6000:]}")


if __name__ == "__main__":
    build(verbose=True)
    sys.exit(0)
