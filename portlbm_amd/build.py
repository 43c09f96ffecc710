"""Builds the native libraries of the package in place.

    python -m portlbm_amd.build        (or portlbm_amd.build.build())

Outputs (all under the repository, none installed elsewhere):
  portlbm_amd/lib/libportlbm_hip.so   HIP kernels + C ABI (include/lbm_hip.h), gfx950
  portlbm_amd/lib/libportlbm.so       C++ host library + handler C ABI (include/portlbm_c.h)
  portlbm_amd/bin/portlbm             command-line driver
  tests/cpp/portlbm_unit_tests        host-side unit tests (no GPU needed)
  oracle/liboracle.so                 the CPU oracle (test infrastructure, via oracle/Makefile)
"""
from __future__ import annotations

import concurrent.futures as cf
import os
import shutil
import subprocess
import sys
from pathlib import Path

PKG = Path(__file__).resolve().parent
ROOT = PKG.parent
INCLUDE = ROOT / "include"
LIB = PKG / "lib"
BIN = PKG / "bin"
OBJ = PKG / "_build"

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
CXX = os.environ.get("CXX", "g++")

# -ffp-contract=off: the FP64 path performs exactly the operations of the
# reference's source text (no fused multiply-adds), so that it can be compared
# with the CPU oracle bit for bit.  The packed-FP32 target feature is disabled
# for all device code.
HIP_FLAGS = [
    "--offload-arch=gfx950",
    "-O3",
    "-std=c++17",
    "-fPIC",
    "-ffp-contract=off",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall",
    f"-I{INCLUDE}",
]
CXX_FLAGS = ["-std=c++20", "-O2", "-fPIC", "-Wall", "-Wextra", "-pthread", f"-I{INCLUDE}"]


def _run(cmd: list[str], cwd: Path = ROOT) -> None:
    proc = subprocess.run(cmd, cwd=cwd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    # hipcc's host pass warns that it does not know the disabled feature
    out = "\n".join(l for l in proc.stdout.splitlines() if "packed-fp32-ops" not in l)
    if proc.returncode != 0:
        raise RuntimeError(f"command failed ({proc.returncode}): {' '.join(cmd)}\n{out}")
    if out.strip():
        print(out, file=sys.stderr)


def _hip_objects() -> list[Path]:
    OBJ.mkdir(exist_ok=True)
    sources = sorted((PKG / "csrc").glob("*.hip"))
    objs = [OBJ / (s.stem + ".o") for s in sources]
    with cf.ThreadPoolExecutor(max_workers=min(8, len(sources) or 1)) as pool:
        futs = [pool.submit(_run, [HIPCC, *HIP_FLAGS, "-c", str(s), "-o", str(o)]) for s, o in zip(sources, objs)]
        for f in futs:
            f.result()
    return objs


def _host_objects() -> list[Path]:
    OBJ.mkdir(exist_ok=True)
    sources = sorted(s for s in (PKG / "host").glob("*.cpp") if s.name != "main.cpp")
    objs = [OBJ / (s.stem + ".host.o") for s in sources]
    with cf.ThreadPoolExecutor(max_workers=min(8, len(sources) or 1)) as pool:
        futs = [pool.submit(_run, [CXX, *CXX_FLAGS, "-c", str(s), "-o", str(o)]) for s, o in zip(sources, objs)]
        for f in futs:
            f.result()
    return objs


def build(verbose: bool = True) -> None:
    LIB.mkdir(exist_ok=True)
    BIN.mkdir(exist_ok=True)
    with cf.ThreadPoolExecutor(max_workers=3) as pool:
        hip = pool.submit(_hip_objects)
        host = pool.submit(_host_objects)
        oracle = pool.submit(_run, ["make", "-s", "-C", str(ROOT / "oracle")])
        hip_objs, host_objs = hip.result(), host.result()
        oracle.result()
    hip_so = LIB / "libportlbm_hip.so"
    _run([HIPCC, "--offload-arch=gfx950", "-shared", "-fPIC", *map(str, hip_objs), "-o", str(hip_so)])
    host_so = LIB / "libportlbm.so"
    _run([CXX, "-shared", "-pthread", *map(str, host_objs), "-o", str(host_so), f"-L{LIB}", "-lportlbm_hip",
          "-Wl,-rpath,$ORIGIN"])
    _run([CXX, *CXX_FLAGS, str(PKG / "host" / "main.cpp"), "-o", str(BIN / "portlbm"), f"-L{LIB}", "-lportlbm",
          "-Wl,-rpath,$ORIGIN/../lib"])
    unit = ROOT / "tests" / "cpp" / "unit_tests.cpp"
    if unit.exists():
        _run([CXX, *CXX_FLAGS, str(unit), "-o", str(unit.parent / "portlbm_unit_tests"), f"-L{LIB}", "-lportlbm",
              f"-Wl,-rpath,{LIB}"])
    if verbose:
        print(f"built {hip_so.relative_to(ROOT)}, {host_so.relative_to(ROOT)}, "
              f"{(BIN / 'portlbm').relative_to(ROOT)}", file=sys.stderr)


def clean() -> None:
    shutil.rmtree(OBJ, ignore_errors=True)
    shutil.rmtree(LIB, ignore_errors=True)
    shutil.rmtree(BIN, ignore_errors=True)


if __name__ == "__main__":
    if "--clean" in sys.argv:
        clean()
    build()
