"""AddressSanitizer + UndefinedBehaviorSanitizer on the CPU.

* The library's host code: gs_api.hip and gs_formats.hip (the C entry points,
  handle registry, frames, scenario-file readers) built host-only under the
  sanitizers and linked with the regular kernel objects; tests/cxx/host_sanitize.cpp
  drives the scenario readers over every data file and over truncated and
  byte-mutated copies, wall splitting, and the argument and handle checks,
  with leak detection on.
* The oracle: oracle.cpp built under the sanitizers (make -C oracle san) and
  loaded in place of liboracle.so (GS_ORACLE_LIB) by the oracle's own test
  files and by a few Boids and crowd steps, with libasan preloaded into the
  interpreter.

Any sanitizer report fails the test (halt on the first error)."""
import os
import subprocess
import sys
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parents[1]


def _asan_runtime() -> str:
    r = subprocess.run(["g++", "-print-file-name=libasan.so"], capture_output=True, text=True)
    p = r.stdout.strip()
    if r.returncode != 0 or not Path(p).exists():
        pytest.skip("g++'s libasan is not available")
    return p


def test_library_host_code_under_sanitizers(tmp_path):
    from gpusteer_amd import build

    exe = build.build_host_sanitized(tmp_path / "hostsan")
    work = tmp_path / "files"
    work.mkdir()
    env = dict(os.environ, ASAN_OPTIONS="detect_leaks=1:halt_on_error=1",
               UBSAN_OPTIONS="halt_on_error=1:print_stacktrace=1")
    r = subprocess.run([str(exe), str(ROOT / "gpusteer_amd" / "data"), str(work)], capture_output=True, text=True,
                       env=env, timeout=600)
    out = r.stdout + r.stderr
    assert r.returncode == 0, out[-6000:]
    assert "runtime error" not in out and "ERROR: AddressSanitizer" not in out, out[-6000:]
    assert "checks passed" in r.stdout


def test_oracle_under_sanitizers():
    asan = _asan_runtime()
    subprocess.run(["make", "-s", "-C", str(ROOT / "oracle"), "san"], check=True)
    lib = ROOT / "oracle" / "_san" / "liboracle_san.so"
    env = dict(os.environ, LD_PRELOAD=asan, GS_ORACLE_LIB=str(lib), ASAN_OPTIONS="detect_leaks=0:halt_on_error=1",
               UBSAN_OPTIONS="halt_on_error=1:print_stacktrace=1")
    tests = sorted(str(p) for p in (ROOT / "tests").glob("test_oracle_*.py"))
    r = subprocess.run([sys.executable, "-m", "pytest", "-x", "-q", "-p", "no:cacheprovider", *tests],
                       capture_output=True, text=True, env=env, cwd=ROOT, timeout=1200)
    out = r.stdout + r.stderr
    assert r.returncode == 0, out[-6000:]
    steps = (
        "import sys; sys.path[:0] = [%r, %r]\n"
        "import pyoracle as orc\n"
        "assert str(orc.LIB_PATH).endswith('liboracle_san.so')\n"
        "from gpusteer_amd import scenarios\n"
        "st, p = scenarios.boids_sphere()\n"
        "for _ in range(3): orc.boids_step(st, p)\n"
        "sc = scenarios.choke(n_total=2000)\n"
        "ref = [g.copy() for g in sc.groups]\n"
        "for _ in range(3): orc.choke_step(ref, sc)\n"
        "print('steps ok')\n" % (str(ROOT), str(ROOT / "oracle")))
    r = subprocess.run([sys.executable, "-c", steps], capture_output=True, text=True, env=env, cwd=ROOT, timeout=600)
    out = r.stdout + r.stderr
    assert r.returncode == 0 and "steps ok" in r.stdout, out[-6000:]
    assert "runtime error" not in out, out[-6000:]
