"""The host library's handler lifecycle under sanitizers, on the CPU.

libportlbm's sources (handler, C API, JSON, core, benchmark driver) are
compiled with AddressSanitizer + UndefinedBehaviorSanitizer, and separately
with ThreadSanitizer, against a CPU stand-in for the GPU library's entry
points (tests/cpp/lbm_cpu_stand_in.cpp), and driven through the C API by
tests/cpp/handler_lifecycle.cpp: the snapshot-restart sequence of the GPU
snapshot tests for every algorithm x layout x storage, pause/resume with a
concurrent reader, re-initialisation, destruction while running, debug mode
and CSV export.  A sanitizer report or a failed check fails the test.
"""
import os
import shutil
import subprocess
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parents[1]
HOST = ["core.cpp", "json.cpp", "hip_algorithm_handler.cpp", "c_api.cpp", "benchmark.cpp"]
SOURCES = [str(ROOT / "portlbm_amd" / "host" / f) for f in HOST] + [
    str(ROOT / "tests" / "cpp" / "lbm_cpu_stand_in.cpp"), str(ROOT / "tests" / "cpp" / "handler_lifecycle.cpp")]

pytestmark = pytest.mark.skipif(shutil.which("g++") is None, reason="needs g++")


def _build_and_run(tmp_path, flags, args, env_extra):
    exe = tmp_path / "lifecycle"
    subprocess.run(["g++", "-std=c++20", "-O1", "-g", "-fno-omit-frame-pointer", *flags, "-pthread",
                    f"-I{ROOT / 'include'}", *SOURCES, "-o", str(exe)], check=True)
    env = dict(os.environ, **env_extra)
    r = subprocess.run([str(exe), *args], stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True, env=env,
                       timeout=900)
    tail = "\n".join(l for l in r.stdout.splitlines() if "ERROR" in l or "WARNING" in l or "lifecycle:" in l
                     or "FAILED" in l or "runtime error" in l)
    assert r.returncode == 0, tail or r.stdout[-4000:]
    assert "lifecycle: " in r.stdout and ": ok" in r.stdout
    assert "runtime error" not in r.stdout and "WARNING: ThreadSanitizer" not in r.stdout


def test_handler_lifecycle_address_and_ub_sanitizer(tmp_path):
    _build_and_run(tmp_path, ["-fsanitize=address,undefined", "-fno-sanitize-recover=undefined"], ["1", "all"],
                   {"ASAN_OPTIONS": "detect_leaks=1", "UBSAN_OPTIONS": "print_stacktrace=1"})


def test_handler_lifecycle_thread_sanitizer(tmp_path):
    _build_and_run(tmp_path, ["-fsanitize=thread"], ["1", "collision"], {"TSAN_OPTIONS": "halt_on_error=1"})
