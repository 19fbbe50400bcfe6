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


# Host-side AddressSanitizer + UndefinedBehaviorSanitizer build of the
# sources without kernels (the C entry points and the scenario-file readers),
# linked with the regular objects of the kernel sources; -Xarch_host puts
# every sanitizer flag on the host pass only.
HOST_SANITIZE_FLAGS = [
    "--offload-host-only",
    "-O1",
    "-g",
    "-std=c++17",
    "-fPIC",
    "-fno-omit-frame-pointer",
    "-ffp-contract=off",
    "-Xarch_host", "-fsanitize=address",
    "-Xarch_host", "-fsanitize=undefined",
    "-Xarch_host", "-fno-sanitize-recover=undefined",
    "-Wno-unused-command-line-argument",
]
HOST_SANITIZED_SOURCES = ("gs_api.hip", "gs_formats.hip")


def build_host_sanitized(out_dir: Path) -> Path:
    """Builds out_dir/libgpusteer_hostsan.so -- the kernel-free sources
    (HOST_SANITIZED_SOURCES) under the sanitizers, the kernel sources from
    their regular objects -- and the exercise program
    tests/cxx/host_sanitize.cpp against it; returns the program's path."""
    out_dir = Path(out_dir)
    out_dir.mkdir(parents=True, exist_ok=True)
    OBJ_DIR.mkdir(exist_ok=True)
    objs = []
    for src in sorted(CSRC.glob("*.hip")):
        if src.name in HOST_SANITIZED_SOURCES:
            if "__global__" in src.read_text():
                raise RuntimeError(f"{src.name} defines kernels; it cannot be built host-only")
            obj = out_dir / (src.stem + ".o")
            cmd = [hipcc(), *HOST_SANITIZE_FLAGS, f"-I{ROOT / 'include'}", "-x", "hip", "-c", str(src), "-o",
                   str(obj)]
            r = subprocess.run(cmd, capture_output=True, text=True)
            if r.returncode != 0:
                raise RuntimeError(f"hipcc (host sanitizers) failed on {src.name}:\n{r.stderr[-6000:]}")
        else:
            obj = OBJ_DIR / (src.stem + ".o")
            if not obj.exists() or obj.stat().st_mtime < src.stat().st_mtime:
                obj = _compile(src)
        objs.append(obj)
    lib = out_dir / "libgpusteer_hostsan.so"
    link = [hipcc(), f"--offload-arch={ARCH}", "-shared", "-fPIC",
            "-Xarch_host", "-fsanitize=address",
            "-Xarch_host", "-fsanitize=undefined",
            "-o", str(lib), *map(str, objs)]
    r = subprocess.run(link, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"link (host sanitizers) failed:\n{r.stderr[-6000:]}")
    exe = out_dir / "host_sanitize"
    cmd = [hipcc(), "--offload-host-only", "-O1", "-g", "-std=c++17", "-fno-omit-frame-pointer",
           "-Xarch_host", "-fsanitize=address",
           "-Xarch_host", "-fsanitize=undefined",
           f"-I{ROOT / 'include'}", "-x", "c++", str(ROOT / "tests" / "cxx" / "host_sanitize.cpp"),
           "-x", "none", str(lib), f"-Wl,-rpath,{out_dir}", "-o", str(exe)]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"hipcc failed on host_sanitize.cpp:\n{r.stderr[-6000:]}")
    return exe


if __name__ == "__main__":
    build(verbose=True)
    sys.exit(0)
